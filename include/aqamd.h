/*
 * aqamd.h -- C ABI of the MI355X (gfx950) CCSD amplitude engine.
 *
 * The tensor entry points replace the operator surface that Aquarius's
 * closed-shell CCSD issues through CTFTensor (src/tensor/ctf_tensor.hpp:16-285,
 * src/tensor/ctf_tensor.cxx:19-516).  Tensors are dense, all-NS (no packed
 * symmetry), double precision and column-major: the key of element
 * (i0,i1,...) is sum_k i_k * prod_{j<k} len_j, exactly CTF's key convention
 * (ctf_tensor.cxx:493-501).  Index strings use CTF/einsum label rules:
 *   - a label in both inputs and not in the output is contracted,
 *   - a label in one input only and not in the output is summed,
 *   - a label in the output only is broadcast,
 *   - a label repeated inside one tensor addresses its diagonal.
 *
 * Every function returns AQ_OK (0) or a positive error code; the message of
 * the last failure on the calling thread is returned by aq_last_error().
 * Nothing in the library aborts the process.
 *
 * Ordering: each context owns one HIP stream; operations are enqueued in call
 * order and run asynchronously, except those documented as synchronising
 * (aq_dot, aq_norm, aq_read*, aq_ctx_sync), which return host values.
 * Contexts share no mutable state, so separate contexts may be driven from
 * separate host threads.  One context must not be used from two threads at
 * once.
 */
#ifndef AQAMD_H
#define AQAMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct aq_ctx aq_ctx;
typedef struct aq_tensor aq_tensor;
typedef struct aq_rhfccsd aq_rhfccsd;

enum {
    AQ_OK = 0,
    AQ_ERR_INVALID = 1,      /* bad argument (ndim mismatch, bad index string, ...) */
    AQ_ERR_HIP = 2,          /* HIP runtime failure */
    AQ_ERR_OOM = 3,          /* device memory or the context's HBM budget exhausted */
    AQ_ERR_UNSUPPORTED = 4,
    AQ_ERR_NUMERIC = 5       /* e.g. singular DIIS system after the retry */
};

/* Contraction arithmetic of a context (see aq_ctx_set_precision). */
enum {
    AQ_PREC_FP64 = 0,        /* FP64 MFMA, FP64 accumulation */
    AQ_PREC_FP32 = 1,        /* operands rounded to FP32, FP32 MFMA, FP32 accumulation */
    AQ_PREC_BF16X3 = 2       /* operands split hi+lo BF16, three BF16 MFMA products, FP32 accumulation */
};

/* ---- errors / version ------------------------------------------------- */

/* Message of the last failing call on this thread ("" if none).
 * Reference: the TensorError family (tensor.hpp:834-898) and
 * Logger::error (task.cxx:103-116), which abort there. */
const char* aq_last_error(void);
int aq_version(void);

/* ---- context (Arena, src/util/distributed.hpp:28-74) ------------------- */

/* Create a context on HIP device `device` with a stream of default priority.
 * hbm_budget: bytes this context may hold in tensors and cached operand copies
 * (0 = 90% of the device's free memory). */
int aq_ctx_create(int device, size_t hbm_budget, aq_ctx** out);
int aq_ctx_destroy(aq_ctx* ctx);
/* Wait for all enqueued work of the context. */
int aq_ctx_sync(aq_ctx* ctx);
/* Arithmetic used by aq_contract.  Contractions with fewer than
 * `min_reduced_flops` multiply-add FLOPs (2*M*N*K) keep FP64 even when a
 * reduced precision is selected. */
int aq_ctx_set_precision(aq_ctx* ctx, int precision, double min_reduced_flops);
int aq_ctx_get_precision(const aq_ctx* ctx, int* precision);
/* Kernel variant of the large-tile GEMM for one arithmetic (AQ_PREC_*):
 * -1 = library default (or the autotuner's choice), FP64/FP32: 0 = BK 16,
 * 1 = BK 32, 2 = BK 16 with two waves per SIMD; BF16x3: 0 = 128x128 blocks,
 * 1 = 256x128, 2 = 128x256, 3 = 256x256 with eight waves (two per SIMD),
 * 4 = 256x256 with four waves.
 * A tuning knob: every variant computes identical results. */
int aq_ctx_set_gemm_variant(aq_ctx* ctx, int precision, int variant);
/* GEMM autotuning (no reference counterpart).  With enable != 0 the first
 * GEMM of every shape bucket (arithmetic, operand layouts, and M, N, K and
 * batch in half-octaves) times each kernel variant once on this context's
 * stream, into a scratch output, and the fastest serves that bucket from then
 * on for every context on the device.  Variants compute identical results
 * (the split-K partition does not depend on them), so tuning changes speed
 * only.  aq_ctx_set_gemm_variant (variant >= 0) overrides the tuner. */
int aq_ctx_set_autotune(aq_ctx* ctx, int enable);
typedef struct aq_tune_entry {
    int device, precision, trans_a, trans_b;
    int m_bucket, n_bucket, k_bucket, batch_bucket;   /* round(2*log2(x)) */
    int variant;                                      /* chosen */
    float ms[5];                                      /* trial time per variant, -1 = not tried */
} aq_tune_entry;
/* Copy up to `cap` entries of the tuning table; *n receives the total. */
int aq_autotune_table(aq_tune_entry* out, int cap, int* n);
/* Bytes currently held (tensors + caches + pooled scratch) and the budget. */
int aq_ctx_memory(const aq_ctx* ctx, size_t* in_use, size_t* budget);
/* High-water mark of the bytes held since creation or the last reset. */
int aq_ctx_peak_memory(aq_ctx* ctx, size_t* peak, int reset);
/* Drop cached operand copies and pooled scratch (keeps tensors), then return
 * the device pool's unused memory to the driver. */
int aq_ctx_trim(aq_ctx* ctx);
/* Per-kernel-class timing with HIP events on the context's stream.
 * When enabled every GEMM / elementwise launch is bracketed by events. */
int aq_ctx_enable_timing(aq_ctx* ctx, int enable);
/* Fill up to `cap` records; *n receives the number of classes.  Each record:
 * name (NUL-terminated, <= 63 chars), launches, total ms, algorithmic flops,
 * algorithmic bytes.  Synchronises. */
typedef struct aq_timing_record {
    char name[64];
    int64_t launches;
    double total_ms;
    double flops;
    double bytes;
} aq_timing_record;
int aq_ctx_timing(aq_ctx* ctx, aq_timing_record* rec, int cap, int* n);
int aq_ctx_timing_reset(aq_ctx* ctx);

/* ---- tensors ----------------------------------------------------------- */

/* CTFTensor(name, arena, ndim, len, sym, zero) (ctf_tensor.cxx:104-120):
 * dense all-NS tensor; zero != 0 clears it.  ndim may be 0 (a scalar). */
int aq_tensor_create(aq_ctx* ctx, int ndim, const int64_t* len, int zero, aq_tensor** out);
/* Deep copy (ctf_tensor.cxx:40-78). */
int aq_tensor_copy(aq_ctx* ctx, const aq_tensor* src, aq_tensor** out);
/* ~CTFTensor (ctf_tensor.cxx:123-127). */
int aq_tensor_destroy(aq_tensor* t);
/* CTFTensor::resize(ndim, len, sym, zero) (ctf_tensor.cxx:189-201): new
 * lengths (all NS), the old data freed and new storage allocated, cleared
 * when zero != 0 (otherwise undefined).  The handle stays valid; copies the
 * library cached from the old contents are dropped.  Not for a tensor that a
 * live solver borrows as an input. */
int aq_tensor_resize(aq_tensor* t, int ndim, const int64_t* len, int zero);
int aq_tensor_ndim(const aq_tensor* t, int* ndim);
int aq_tensor_lengths(const aq_tensor* t, int64_t* len);
int aq_tensor_size(const aq_tensor* t, int64_t* size);
/* Device address of the column-major data (for zero-copy interop).  Reading
 * through it is always safe once the context's queued work is done
 * (aq_ctx_sync).  After writing through it, call aq_tensor_touch before the
 * tensor is next used by the library. */
int aq_tensor_data(const aq_tensor* t, double** dptr);
/* Declare that X's data was changed outside the library (through the pointer
 * of aq_tensor_data): bumps X's write version, the key of every copy the
 * library derived from it (permuted, FP32 or BF16 operand copies in the
 * context's cache; a solver's denominators, spin-adapted blocks and packed
 * ladder operators when X is one of its inputs), so each is rebuilt before its
 * next use.  Without it those copies keep the old contents.  The counterpart
 * of writing a CTFTensor through its own write calls (ctf_tensor.hpp:139-247),
 * which the library's aq_write / aq_write_all already cover. */
int aq_tensor_touch(aq_tensor* X);

/* ---- operators (CTFTensor methods) ------------------------------------- */

/* C[idxC] = beta*C[idxC] + alpha * A[idxA]*B[idxB].
 * CTFTensor::mult (ctf_tensor.cxx:434-442).  C may alias A or B. */
int aq_contract(aq_ctx* ctx, double alpha, const aq_tensor* A, const char* idxA,
                const aq_tensor* B, const char* idxB, double beta, aq_tensor* C, const char* idxC);

/* B[idxB] = beta*B[idxB] + alpha*A[idxA].  CTFTensor::sum (ctf_tensor.cxx:454-458).
 * Correct when A aliases B (the permuted source is read before B is written). */
int aq_sum(aq_ctx* ctx, double alpha, const aq_tensor* A, const char* idxA,
           double beta, aq_tensor* B, const char* idxB);

/* X[idx] = alpha*X[idx].  CTFTensor::scale (ctf_tensor.cxx:461-464). */
int aq_scale(aq_ctx* ctx, aq_tensor* X, double alpha, const char* idx);

/* X = beta*X + alpha (scalar broadcast).  CTFTensor::sum(alpha, beta)
 * (ctf_tensor.cxx:445-451). */
int aq_fill(aq_ctx* ctx, aq_tensor* X, double alpha, double beta);

/* *out = sum A[idxA]*B[idxB] over all labels.  CTFTensor::dot
 * (ctf_tensor.cxx:467-478).  Synchronising. */
int aq_dot(aq_ctx* ctx, const aq_tensor* A, const char* idxA,
           const aq_tensor* B, const char* idxB, double* out);

/* X[i0..] /= (d_0[i0] + d_1[i1] + ... + shift), or 0 where that |sum| < 1e-4.
 * CTFTensor::weight (ctf_tensor.cxx:481-514).  d[k] points to len[k] host
 * doubles. */
int aq_weight(aq_ctx* ctx, aq_tensor* X, int ndim, const double* const* d, double shift);

/* aq_weight without any allocation: no reference counterpart; for callers
 * that cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, semantics and
 * results as aq_weight (bitwise), but the denominators are staged in one fixed
 * device array per GPU (AQ_NOALLOC_WEIGHT_MAX doubles, held until the process
 * ends) instead of pooled device memory, and no timing events are recorded.
 * The lengths of X must sum to at most AQ_NOALLOC_WEIGHT_MAX, otherwise
 * AQ_ERR_INVALID.  Calls on one GPU share that array: they take a per-device
 * lock and return only after the kernel has finished (synchronising), so calls
 * from several threads or contexts serialise instead of overwriting each
 * other's denominators. */
#define AQ_NOALLOC_WEIGHT_MAX 65536
int aq_weight_noalloc(aq_ctx* ctx, aq_tensor* X, int ndim, const double* const* d, double shift);

/* p = 0: max|x|, p = 1: sum|x|, p = 2: 2-norm.  CTFTensor::norm
 * (ctf_tensor.cxx:403-431).  Synchronising. */
int aq_norm(aq_ctx* ctx, const aq_tensor* X, int p, double* out);

/* Key/value I/O with column-major keys (ctf_tensor.hpp:139-247).
 * aq_read: getRemoteData; aq_write: writeRemoteData(alpha, beta, pairs)
 * (x[key] = beta*x[key] + alpha*val).  aq_read_all / aq_write_all move the
 * whole tensor in key order (getAllData).  aq_read* synchronise. */
int aq_read(aq_ctx* ctx, const aq_tensor* X, int64_t n, const int64_t* keys, double* vals);
int aq_write(aq_ctx* ctx, aq_tensor* X, int64_t n, const int64_t* keys, const double* vals,
             double alpha, double beta);
int aq_read_all(aq_ctx* ctx, const aq_tensor* X, double* host);
int aq_write_all(aq_ctx* ctx, aq_tensor* X, const double* host);

/* Elementwise C = beta*C + alpha*A/B over equal-shaped tensors; elements where
 * |B| <= DBL_MIN are left unchanged.  CTFTensor::div (ctf_tensor.cxx:271-342). */
int aq_div(aq_ctx* ctx, double alpha, const aq_tensor* A, const aq_tensor* B, double beta, aq_tensor* C);
/* Elementwise C = beta*C + alpha/A; elements where |A| <= DBL_MIN are left
 * unchanged.  CTFTensor::invert (ctf_tensor.cxx:344-371). */
int aq_invert(aq_ctx* ctx, double alpha, const aq_tensor* A, double beta, aq_tensor* C);

/* B[startB + r] = beta*B[startB + r] + alpha*A[startA + r] for r over the box
 * `extent` (CTFTensor::slice, ctf_tensor.cxx:219-256, NS only). */
int aq_slice(aq_ctx* ctx, double alpha, const aq_tensor* A, const int64_t* startA,
             double beta, aq_tensor* B, const int64_t* startB, const int64_t* extent);

/* ---- closed-shell CCSD engine (RHFCCSD, src/cc/rhfccsd.cxx) ------------ */

/* Task options.  Defaults (aq_rhfccsd_default_config) follow the task spec
 * (rhfccsd.cxx:332-353): convergence 1e-9, max_iterations 50, DIIS order 5,
 * start 1, damping 0.  The precision fields have no reference counterpart. */
typedef struct aq_rhfccsd_config {
    double convergence;
    int max_iterations;
    int diis_order;
    int diis_start;
    double diis_damping;
    /* contraction arithmetic while conv >= refine_below (AQ_PREC_*) */
    int precision;
    /* switch to FP64 contractions once conv < refine_below (0 = never); the
     * switch is final: later iterations stay FP64 even if conv rises again */
    double refine_below;
    /* contractions below this many FLOPs stay FP64 in reduced precision */
    double min_reduced_flops;
    /* 1: fuse weight + update + Tau + energy + conv into one kernel;
     * 0: issue them as separate operator calls, as the reference does */
    int fused_update;
    /* 1 (default): drop the DIIS history when the contractions switch to
     * FP64, so the refinement extrapolates only over FP64 residuals (residuals
     * of the reduced-precision phase carry errors near 1e-7 that slow the
     * tail).  No effect on an FP64-only solve. */
    int diis_reset_on_refine;
    /* > 0: with a BF16x3 bulk phase, run FP32 contractions while
     * fp32_stage_below <= conv < refine_below, and FP64 only below it
     * (0 = switch straight to FP64 at refine_below) */
    double fp32_stage_below;
    /* 1: run each iteration's contraction stage (rhfccsd.cxx:231-309) as a
     * HIP graph: captured on the second iteration at a contraction precision
     * (scratch pool and operand copies warm) and replayed on later ones;
     * dropped and captured again after a precision switch or whenever the
     * context frees device memory.  A capture that would allocate is
     * abandoned and that iteration runs eagerly (two such failures turn
     * graphs off for the solver).  Same kernels and arguments as the eager
     * path, so the results are bitwise identical.  Ignored while the context
     * records kernel timings.  0: every iteration eager. */
    int graph;
    /* 1: evaluate the particle-particle ladder (rhfccsd.cxx:303) on the
     * symmetric and antisymmetric combinations over packed pairs a<=b, e<=f,
     * i<=j (two GEMMs with a quarter of the FLOPs of the full contraction;
     * Scuseria, Janssen, Schaefer, J. Chem. Phys. 89, 7382 (1988)).  Needs
     * <ab|ef> = <ba|fe> and <ab|ij> = <ba|ji>, checked at setup to 1e-13
     * relative; otherwise, or when 0, the full contraction as the reference
     * issues it.  Costs (v(v+1)/2)^2 + (v(v-1)/2)^2 doubles of device memory. */
    int packed_ladder;
} aq_rhfccsd_config;

int aq_rhfccsd_default_config(aq_rhfccsd_config* cfg);

/* RHFCCSD::run setup (rhfccsd.cxx:45-98): takes the Fock blocks
 * fAI[v,o], fAB[v,v], fIJ[o,o] and the Dirac blocks VABCD[v,v,v,v],
 * VABCI[v,v,v,o], VABIJ[v,v,o,o], VAIBJ[v,o,v,o], VAIJB[v,o,o,v],
 * VAIJK[v,o,o,o], VIJKL[o,o,o,o] (layout of rhfaomoints.cxx:58-65), builds the
 * MP2 guess and the DIIS state.  The inputs are borrowed (not copied) and must
 * outlive the solver.  A caller that changes an input between iterations
 * (aq_write*, or a write through aq_tensor_data followed by aq_tensor_touch)
 * gets the next iteration computed from the new values: the solver rebuilds
 * the denominators, spin-adapted blocks and packed ladder operators it derived
 * from a changed input; the amplitudes and DIIS history are kept. */
int aq_rhfccsd_create(aq_ctx* ctx, const aq_tensor* const* fock3, const aq_tensor* const* v7,
                      const aq_rhfccsd_config* cfg, aq_rhfccsd** out);
int aq_rhfccsd_destroy(aq_rhfccsd* cc);
/* MP2 energy of the guess (rhfccsd.cxx:94). */
int aq_rhfccsd_mp2(const aq_rhfccsd* cc, double* mp2);
/* One RHFCCSD::iterate (rhfccsd.cxx:194-327): energy and conv as computed
 * before the DIIS extrapolation. */
int aq_rhfccsd_iterate(aq_rhfccsd* cc, double* energy, double* conv);
/* Iterative::run (iterative.hpp:126-166): iterate until conv < convergence or
 * max_iterations.  *converged = 0 reports non-convergence (the reference only
 * logs it).  `log_cb`, if non-null, is called after every iteration. */
typedef void (*aq_iter_callback)(int iter, double energy, double conv, double seconds, int precision, void* user);
int aq_rhfccsd_run(aq_rhfccsd* cc, aq_iter_callback log_cb, void* user,
                   double* energy, double* conv, int* iterations, int* converged);
/* Hbar products of the current amplitudes, formed after convergence
 * (rhfccsd.cxx:108-188): overwrites FAE, FMI, FME, WMNIJ, WMNEJ, WAMIJ,
 * WAMEI, WAMIE with the similarity-transformed Hamiltonian elements the EOM
 * consumers read (rhfeomeeccsd.cxx:55-80); with_wabej != 0 also forms
 * WABEJ [v,v,v,o] (the reference forms it when a consumer uses it).  FP64;
 * synchronising.  A later aq_rhfccsd_iterate overwrites the products again
 * with the iteration's intermediates. */
int aq_rhfccsd_hbar(aq_rhfccsd* cc, int with_wabej);
/* Amplitudes / state: which = 0: T1[v,o], 1: T2[v,v,o,o], 2: Z1, 3: Z2,
 * 4: Tau; 10..17: FAE, FMI, FME, WMNIJ, WMNEJ, WAMIJ, WAMEI, WAMIE;
 * 18: WABEJ (after aq_rhfccsd_hbar(cc, 1)). */
int aq_rhfccsd_tensor(const aq_rhfccsd* cc, int which, const aq_tensor** out);
/* Precision the next iteration's contractions will use. */
int aq_rhfccsd_current_precision(const aq_rhfccsd* cc, int* precision);
/* HIP-graph counters of the contraction stage (aq_rhfccsd_config.graph):
 * graphs captured, replays, and abandoned captures.  No reference counterpart
 * (TaskDAG::execute, src/task/task.cxx:495-588, runs every task eagerly). */
int aq_rhfccsd_graph_stats(const aq_rhfccsd* cc, int* captures, int* replays, int* failed);
/* 1 when the solver evaluates the ladder on packed pairs (see
 * aq_rhfccsd_config.packed_ladder), 0 when it runs the full contraction. */
int aq_rhfccsd_packed_ladder(const aq_rhfccsd* cc, int* packed);

/* ---- DIIS (DIIS<T>, src/convergence/diis.hpp:36-227) -------------------
 *
 * The Pulay extrapolator the CCSD drivers use, over device-resident history:
 * aq_diis_extrapolate(x, dx) stores copies of the nx tensors x and the ndx
 * residuals dx as the newest history slot, builds the bordered system from the
 * inner products of dx with every stored residual, solves it (a retry with
 * eps*B[n-1][n-1] added to the diagonal when a pivot is exactly zero, the
 * dsysv 'info > 0' path of diis.hpp:171-187; AQ_ERR_NUMERIC if that fails
 * too) and overwrites x and dx with the extrapolated combinations.  `start`
 * and `damping` as DIIS::start/damping (diis.hpp:151-164).  Any order >= 0
 * (history memory: order copies of every x and dx).  The tensors must keep
 * their sizes from call to call. */
typedef struct aq_diis aq_diis;
int aq_diis_create(aq_ctx* ctx, int order, int start, double damping, int nx, int ndx, aq_diis** out);
int aq_diis_destroy(aq_diis* d);
int aq_diis_extrapolate(aq_diis* d, aq_tensor* const* x, aq_tensor* const* dx);
/* The coefficients of the last solve (order + 1 values, the last one the
 * Lagrange multiplier; up to `cap` copied, *n receives the count) and whether
 * that solve took the eps retry. */
int aq_diis_coefficients(const aq_diis* d, double* c, int cap, int* n, int* used_eps_retry);

/* ---- synthetic closed-shell integrals (workload construction) ---------- */

/* Host: orbital energies (eps[o+v], occupied sorted then virtual sorted) and
 * symmetric three-index factors B[Q][p][q] (nQ*(o+v)^2 doubles) of the seeded
 * construction: eps_occ ~ U[-1.5,-0.3], eps_vir ~ U[0.2,5.0],
 * B^Q_pq = B^Q_qp ~ sigma*N(0,1), counter-based splitmix64 + Box-Muller.
 * sigma <= 0 selects the size-normalised default
 * aq_synth_default_sigma(v, nQ) = 0.05*sqrt(150/(v*sqrt(nQ))), which keeps the
 * effective CC coupling v*sigma^2*sqrt(nQ) of the v=18, nQ=72 system at every
 * size (sigma = 0.05 at any size makes CCSD diverge once v*sqrt(nQ) >> 150). */
int aq_synth_host(uint64_t seed, int o, int v, int nQ, double sigma, double* eps, double* B);
double aq_synth_default_sigma(int v, int nQ);
/* Device: build f blocks (fAI = 0, fAB/fIJ diagonal) and the seven Dirac
 * blocks <pq|rs> = (pr|qs), (pq|rs) = sum_Q B^Q_pq B^Q_rs, from host eps and B
 * as produced by aq_synth_host.  Creates 10 tensors: out[0..2] = fAI, fAB, fIJ;
 * out[3..9] = VABCD, VABCI, VABIJ, VAIBJ, VAIJB, VAIJK, VIJKL. */
int aq_synth_build(aq_ctx* ctx, int o, int v, int nQ, const double* eps, const double* B,
                   aq_tensor** out);

/* ---- integral front end for real molecules ------------------------------
 *
 * The reference builds the CCSD input with the tasks 1eints, 2eints,
 * localaoscf and aomoints (test/input.aq:12-17).  These entry points replace
 * them for closed shells: the molecule + basis, the AO integrals, RHF SCF
 * (J/K builds as device contractions) and the AO->MO step that yields the
 * ten tensors aq_rhfccsd_create takes. */
typedef struct aq_molecule aq_molecule;

/* Molecule with a GENBAS-format basis set: BasisSet::readBasisSet
 * (src/input/basis.cxx:15-75), Shell normalisation (src/integrals/shell.cxx:
 * 114-160), Molecule (src/input/molecule.cxx:180-182, 1237-1275).  Z[natoms],
 * xyz[3*natoms] in bohr (the reference converts angstrom with 1 A =
 * 1.88972612456506198632428439 bohr, molecule.cxx:1335-1336); `genbas` is the
 * text of a GENBAS file holding an entry for every element present;
 * spherical != 0 selects pure harmonics (the reference default,
 * molecule.cxx:1364-1367).  Each contraction of a generally contracted shell
 * becomes its own segmented shell. */
int aq_molecule_create(int natoms, const int* Z, const double* xyz, const char* genbas, int spherical,
                       int charge, aq_molecule** out);
int aq_molecule_destroy(aq_molecule* mol);
int aq_molecule_info(const aq_molecule* mol, int* nao, int* nelectrons, int* nshells,
                     double* nuclear_repulsion);
/* 1eints (src/integrals/1eints.cxx): overlap S, kinetic T and nuclear
 * attraction V, nao*nao each (symmetric), computed on the host. */
int aq_ints_1e(const aq_molecule* mol, double* S, double* T, double* V);
/* 2eints (src/integrals/2eints.cxx, os.cxx): eri[m + n*N + l*N^2 + s*N^3] =
 * (mn|ls) for all nao^4 elements, on the host with `nthreads` threads
 * (<= 0: every core).  aq_ints_2e_device uploads the same tensor to a new
 * [nao,nao,nao,nao] device tensor. */
int aq_ints_2e(const aq_molecule* mol, int nthreads, double* eri);
int aq_ints_2e_device(aq_ctx* ctx, const aq_molecule* mol, int nthreads, aq_tensor** out);

/* localaoscf (src/scf/aouhf.cxx, uhf.cxx:17-409) for closed shells.
 * Defaults (aouhf.cxx:266-282): convergence 1e-12 (energy change; the
 * orthonormal-basis commutator must also fall below 100x that), 150
 * iterations, DIIS order 6.  frozen_core is false there and here. */
typedef struct aq_scf_config {
    double convergence;
    int max_iterations;
    int diis_order;
} aq_scf_config;
typedef struct aq_scf_result {
    double energy;      /* total SCF energy including nuclear repulsion */
    double conv;        /* max |X^T (FDS - SDF) X| of the last iteration */
    int iterations;
    int converged;
    int nocc;           /* occupied (alpha) orbitals */
    int nocc_beta;      /* occupied beta orbitals (= nocc for RHF) */
} aq_scf_result;
int aq_scf_default_config(aq_scf_config* cfg);
/* eri: the device AO tensor of aq_ints_2e_device.  Outputs (optional):
 * eps[nao] orbital energies ascending, C[nao*nao] column-major MO
 * coefficients (C[m + p*nao]), F[nao*nao] the final AO Fock matrix. */
int aq_scf_rhf(aq_ctx* ctx, const aq_molecule* mol, const aq_tensor* eri, const aq_scf_config* cfg,
               aq_scf_result* res, double* eps, double* C, double* F);
/* aomoints for RHF (src/operator/rhfaomoints.cxx:38-180): the MO Fock blocks
 * f = C^T F C and the seven Dirac blocks <pq|rs> = (pr|qs) over the occupied
 * (first nocc) and virtual MOs.  Creates 10 tensors in the order of
 * aq_synth_build: fAI, fAB, fIJ, VABCD, VABCI, VABIJ, VAIBJ, VAIJB, VAIJK,
 * VIJKL. */
int aq_aomo_rhf(aq_ctx* ctx, const aq_tensor* eri, int nocc, const double* C, const double* F,
                aq_tensor** out);

/* ---- open shells: UHF and the spin-orbital CCSD path -------------------- */

/* localaoscf for any multiplicity (UHF, src/scf/uhf.cxx:17-409): n_alpha =
 * (N + mult - 1)/2, n_beta = (N - mult + 1)/2.  Outputs (optional): eps[2*nao]
 * alpha then beta orbital energies, C[2*nao*nao] alpha then beta column-major
 * coefficients, F[2*nao*nao] the final alpha and beta AO Fock matrices. */
int aq_scf_uhf(aq_ctx* ctx, const aq_molecule* mol, const aq_tensor* eri, int multiplicity,
               const aq_scf_config* cfg, aq_scf_result* res, double* eps, double* C, double* F);
/* Spin-orbital integrals of a UHF reference (the spin-orbital aomoints task,
 * src/operator/aomoints.cxx, as CCSD<U> reads them, src/cc/ccsd.cxx:92-105):
 * dense spin orbitals ordered occupied alpha, occupied beta, virtual alpha,
 * virtual beta (o = nalpha + nbeta, v = 2 nao - o).  C and F as written by
 * aq_scf_uhf.  Creates 13 tensors: fAI [v,o], fME [o,v], fAE [v,v], fMI [o,o],
 * and the antisymmetrised <pq||rs> blocks VABIJ [v,v,o,o], VMNEF [o,o,v,v],
 * VAMEF [v,o,v,v], VABEJ [v,v,v,o], VABEF [v,v,v,v], VMNIJ [o,o,o,o],
 * VMNEJ [o,o,v,o], VAMIJ [v,o,o,o], VAMEI [v,o,v,o]. */
int aq_aomo_uhf(aq_ctx* ctx, const aq_tensor* eri, int nalpha, int nbeta, const double* C, const double* F,
                aq_tensor** out);
/* Spin-orbital CCSD (CCSD<U>, src/cc/ccsd.cxx:30-193) on the 13 tensors of
 * aq_aomo_uhf, which it borrows.  The products into antisymmetric outputs are
 * antisymmetrised over index pairs taken from different operands, as the
 * reference's spin-orbital tensors do implicitly.  cfg as for aq_rhfccsd
 * (the task's defaults are the same); ccd != 0 holds T1 at zero (the CCD
 * equations).  aq_soccsd_tensor: 0 T1 [v,o], 1 T2 [v,v,o,o], 2 Z1, 3 Z2, 4 Tau. */
typedef struct aq_soccsd aq_soccsd;
int aq_soccsd_create(aq_ctx* ctx, const aq_tensor* const* in, const aq_rhfccsd_config* cfg, int ccd,
                     aq_soccsd** out);
int aq_soccsd_destroy(aq_soccsd* cc);
int aq_soccsd_mp2(const aq_soccsd* cc, double* mp2);
int aq_soccsd_iterate(aq_soccsd* cc, double* energy, double* conv);
int aq_soccsd_run(aq_soccsd* cc, aq_iter_callback log_cb, void* user,
                  double* energy, double* conv, int* iterations, int* converged);
int aq_soccsd_tensor(const aq_soccsd* cc, int which, const aq_tensor** out);

/* CCSD inputs from integrals already in an orthonormal MO basis (the role the
 * reference's FCIDUMP task plays, src/operator/fcidump.cxx:80-517): mo_eri is
 * a device tensor [n,n,n,n] with mo[p,q,r,s] = (pq|rs), h[p + q*n] the
 * one-electron integrals, the first nocc orbitals doubly occupied.  Builds the
 * closed-shell Fock matrix f = h + 2 J - K of those orbitals (J, K as device
 * contractions), returns the electronic SCF energy sum_i (h_ii + f_ii) (add
 * the core energy for the total) and, optionally, f[p + q*n], and creates the
 * ten tensors in the order of aq_synth_build. */
int aq_mo_rhf_blocks(aq_ctx* ctx, const aq_tensor* mo_eri, int nocc, const double* h, double* escf_elec,
                     double* f, aq_tensor** out);

/* FCIDUMP files (Knowles-Handy format), read as the reference's FCIDUMP task
 * reads them (src/operator/fcidump.cxx:43-140): header up to "/", "&END" or
 * "$END" with NORB, NELEC (and MS2); lines "value p q r s" with p = 0 the
 * core energy, q = 0 ignored, r = 0 the one-electron h_pq (mirrored to h_qp
 * unless full_1e), else (pq|rs) filling all eight symmetric positions (a line
 * whose ordered pair (p,q) precedes (r,s) is skipped, as there).  Outputs
 * h[p + q*norb], eri[p + q*n + r*n^2 + s*n^3] (norb^4 doubles) and the core
 * energy.  The writer emits the canonical order (i>=j, k>=l, ij>=kl) of the
 * integrals whose magnitude is at least `threshold`, then h (i>=j), then the
 * core energy. */
int aq_fcidump_header(const char* path, int* norb, int* nelec, int* ms2);
int aq_fcidump_read(const char* path, int full_1e, double* h, double* eri, double* ecore);
int aq_fcidump_write(const char* path, int norb, int nelec, int ms2, const double* h, const double* eri,
                     double ecore, double threshold);

/* ---- serving: continuous batching of CCSD jobs (no reference counterpart;
 * Aquarius runs one task DAG sequentially, src/task/task.cxx:495-588) -------
 *
 * A server owns `slots` worker threads on one device, each with its own
 * context (stream).  Jobs are admitted in arrival order against an HBM budget
 * (the sum of the running jobs' reservations); a job that does not fit waits
 * while later jobs that fit may start, unless the oldest waiting job has
 * waited longer than `max_bypass_s`, in which case admission pauses until it
 * fits.  Admitted jobs run concurrently, their kernels overlapping on the
 * device; each job keeps its integrals, amplitudes and DIIS history resident
 * until it converges.  The slot contexts autotune their GEMMs
 * (aq_ctx_set_autotune), sharing one tuning table.  Jobs share no mutable
 * state that affects arithmetic, so a job's result does not depend on what
 * else runs. */
typedef struct aq_server aq_server;

typedef struct aq_job_spec {
    uint64_t seed;          /* synthetic system (aq_synth_host construction) */
    int o, v, nQ;           /* nQ <= 0: 3*(o+v) */
    double sigma;           /* <= 0: aq_synth_default_sigma */
    aq_rhfccsd_config cfg;  /* solver options */
    /* Caller-supplied integrals instead of the synthetic system (at most one
     * of the two; both null = synthetic):
     *  - inputs: the ten tensors aq_rhfccsd_create takes, in the order of
     *    aq_synth_build (fAI, fAB, fIJ, VABCD, VABCI, VABIJ, VAIBJ, VAIJB,
     *    VAIJK, VIJKL), e.g. from aq_aomo_rhf (the reference's aomoints task
     *    feeding ccsd, rhfaomoints.cxx:38-180).  Tensors of any context on the
     *    server's device; submit waits for their context's queued work; they
     *    are borrowed, not counted against the server budget, and must stay
     *    unchanged until the job is done.  o and v are taken from fAI.
     *  - fcidump: path of an FCIDUMP file that the job's slot reads and turns
     *    into the ten tensors (aq_fcidump_read + aq_mo_rhf_blocks: the
     *    reference's fcidump task, fcidump.cxx:43-517); closed shell (MS2 = 0),
     *    o = NELEC/2, v = NORB - o.  The path is copied at submit. */
    aq_tensor* const* inputs;
    const char* fcidump;
} aq_job_spec;

enum { AQ_JOB_QUEUED = 0, AQ_JOB_RUNNING = 1, AQ_JOB_DONE = 2, AQ_JOB_FAILED = 3 };

typedef struct aq_job_result {
    int64_t id;
    int status;
    int slot;
    double t_submit, t_start, t_end;     /* seconds since the server started */
    double energy, conv, mp2;
    int iterations, converged;
    int reduced_iterations;              /* iterations run below FP64 */
    size_t bytes_reserved, bytes_peak;
    int error_code;
    char error[160];
} aq_job_result;

typedef struct aq_server_stats {
    int64_t submitted, completed, failed, queued, running;
    size_t budget, reserved;
    double elapsed_s;
} aq_server_stats;

/* Conservative device bytes one job needs (integrals, amplitudes, W/F
 * intermediates, DIIS history, operand copies and the integral build). */
size_t aq_job_bytes_estimate(int o, int v, int nQ, int diis_order, int precision);

int aq_server_create(int device, size_t hbm_budget, int slots, double max_bypass_s, aq_server** out);
/* Queue a job (rejected with AQ_ERR_OOM if it can never fit the budget). */
int aq_server_submit(aq_server* s, const aq_job_spec* spec, int64_t* job_id);
int aq_server_poll(aq_server* s, int64_t job_id, aq_job_result* out);
/* Block until the job is done or failed. */
int aq_server_wait(aq_server* s, int64_t job_id, aq_job_result* out);
/* Block until every submitted job is done or failed. */
int aq_server_drain(aq_server* s);
int aq_server_stats_get(aq_server* s, aq_server_stats* out);
/* Cross-job GEMM grouping (the serving loop's grouped launch; no reference
 * counterpart).  window_us > 0: the slots' eager GEMMs with 128x128 or 64x64
 * FP64/BF16x3 blocks are merged, across jobs, into one persistent grouped
 * kernel launch per (format, operand layout, block) key: a fixed grid of
 * resident blocks walks the tiles of every member GEMM.  A GEMM joins
 * same-key GEMMs that other slots have pending; with none pending, and
 * another slot inside a contraction stage, it waits up to window_us for one,
 * but only for keys whose earlier waits usually found a partner (a moving
 * average above 0.2).  A GEMM that finds no partner launches alone on its own
 * slot's stream, as without grouping.  Each element is computed exactly as in
 * a launch of its GEMM alone, so results do not change.  Stages replayed from
 * a HIP graph do not group (run jobs with cfg.graph = 0 to group every
 * stage).  window_us = 0 turns grouping off.  Only while no job is queued or
 * running. */
int aq_server_set_grouping(aq_server* s, int window_us);
/* Grouped launches so far, the GEMMs they carried, the largest group, and the
 * GEMMs offered for grouping that found no partner and ran alone. */
int aq_server_grouping_stats(aq_server* s, int64_t* launches, int64_t* gemms, int* max_group, int64_t* solo);
/* Drains, stops the workers and releases every context. */
int aq_server_destroy(aq_server* s);

#ifdef __cplusplus
}
#endif

#endif /* AQAMD_H */
