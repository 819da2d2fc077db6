// HBM-bound kernels: strided (permuting / reducing / broadcasting) sums and
// format conversions, CTF weight, deterministic reductions, key-value I/O,
// the split-K reduction and the fused CCSD amplitude update.
//
// Floating-point contraction is switched off in this file so that
// beta*y + alpha*x rounds exactly as the two-step CTF/numpy statement does.
#include "internal.hpp"

#include <algorithm>
#include <cmath>

#pragma clang fp contract(off)

namespace aqamd {

constexpr int TPB = 256;

static inline unsigned grid_for(int64_t n, int64_t per_block, unsigned cap = 1u << 20) {
    int64_t g = (n + per_block - 1) / per_block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return unsigned(g);
}

// ------------------------------------------------------------------ outputs

struct SumOut {
    double* dst;
    double alpha, beta;
    __device__ __forceinline__ void put(int64_t off, double v) const {
        const double a = alpha * v;
        dst[off] = beta == 0.0 ? a : beta * dst[off] + a;
    }
};

struct F32Out {
    float* dst;
    __device__ __forceinline__ void put(int64_t off, double v) const { dst[off] = float(v); }
};

__device__ __forceinline__ unsigned short bf16_rne(float f) {
    unsigned int u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ double bf16_to_double(unsigned short h) {
    return double(__uint_as_float((unsigned int)h << 16));
}

struct BF16X2Out {   // hi plane at dst[off], lo plane at dst[plane + off]
    unsigned short* dst;
    int64_t plane;
    __device__ __forceinline__ void put(int64_t off, double v) const {
        const unsigned short h = bf16_rne(float(v));
        const unsigned short l = bf16_rne(float(v - bf16_to_double(h)));
        dst[off] = h;
        dst[plane + off] = l;
    }
};

// ------------------------------------------------------ strided kernels

struct PlanDev {
    int nd, nr;
    int64_t len[MAXDIM], sdst[MAXDIM], ssrc[MAXDIM];
    int64_t rlen[MAXDIM], rsrc[MAXDIM];
    int64_t nout, nred;
};

static PlanDev to_dev(const StridedPlan& p) {
    PlanDev d;
    d.nd = p.nd;
    d.nr = p.nr;
    for (int k = 0; k < MAXDIM; ++k) {
        d.len[k] = k < p.nd ? p.len[k] : 1;
        d.sdst[k] = k < p.nd ? p.sdst[k] : 0;
        d.ssrc[k] = k < p.nd ? p.ssrc[k] : 0;
        d.rlen[k] = k < p.nr ? p.rlen[k] : 1;
        d.rsrc[k] = k < p.nr ? p.rsrc[k] : 0;
    }
    d.nout = p.nout;
    d.nred = p.nred;
    return d;
}

// one thread per output element; reduction over the summed source dims
// Index type I: the flat output index and its digits (uint32_t when every
// index and stride step stays below 2^32, which makes each digit a 32-bit
// division); element offsets are always 64-bit.
template <typename I, class Out>
__device__ __forceinline__ void strided_body(const PlanDev& p, const double* __restrict__ src, Out& out, bool read_src) {
    const I nout = I(p.nout), stride = I(gridDim.x) * I(TPB);
    for (I idx = I(blockIdx.x) * I(TPB) + I(threadIdx.x); idx < nout; idx += stride) {
        I r = idx;
        int64_t od = 0, os = 0;
        for (int k = 0; k < p.nd; ++k) {
            const I len = I(p.len[k]);
            const I q = r / len;
            const int64_t c = int64_t(r - q * len);
            r = q;
            od += c * p.sdst[k];
            os += c * p.ssrc[k];
        }
        double acc = 0.0;
        if (read_src) {
            if (p.nr == 0) {
                acc = src[os];
            } else {
                for (int64_t j = 0; j < p.nred; ++j) {
                    int64_t q = j, o2 = os;
                    for (int k = 0; k < p.nr; ++k) {
                        o2 += (q % p.rlen[k]) * p.rsrc[k];
                        q /= p.rlen[k];
                    }
                    acc += src[o2];
                }
            }
        }
        out.put(od, acc);
    }
}

template <class Out>
__global__ __launch_bounds__(TPB) void k_strided(PlanDev p, const double* __restrict__ src, Out out, bool read_src) {
    // idx + stride < 2^31 + 2^28 (grid <= 2^20 blocks of 256) never wraps 32 bits
    if (p.nout < (int64_t(1) << 31)) strided_body<uint32_t>(p, src, out, read_src);
    else strided_body<int64_t>(p, src, out, read_src);
}

// LDS-tiled transpose: dims d0 (unit stride in dst) and d1 (unit stride in
// src) are tiled T x T (T = 32); the other dims are walked per block.  Reads
// are coalesced along d1, writes along d0.  (A 64 x 64 tile for the BF16
// hi/lo output, 128-byte plane rows instead of 64, measured 7% slower on the
// o=48, v=288 splits: a quarter of the blocks per tile count, twice the LDS.)
template <class Out> struct TransposeTile { static constexpr int T = 32; };

template <class Out>
__global__ __launch_bounds__(TPB) void k_transpose(PlanDev p, int d0, int d1, int64_t tiles0, int64_t tiles1,
                                                    const double* __restrict__ src, Out out) {
    constexpr int T = TransposeTile<Out>::T;
    constexpr int R = TPB / T;   // rows per pass
    __shared__ double tile[T][T + 1];
    const int64_t b = blockIdx.x + int64_t(blockIdx.y) * gridDim.x;
    const int64_t nt = tiles0 * tiles1;
    const int64_t t0 = (b % nt) % tiles0, t1 = (b % nt) / tiles0;
    int64_t rest = b / nt;
    int64_t od = 0, os = 0;
    for (int k = 0; k < p.nd; ++k) {
        if (k == d0 || k == d1) continue;
        const int64_t c = rest % p.len[k];
        rest /= p.len[k];
        od += c * p.sdst[k];
        os += c * p.ssrc[k];
    }
    if (rest != 0) return;   // past the last slab
    const int tx = threadIdx.x % T, ty = threadIdx.x / T;
    const int64_t L0 = p.len[d0], L1 = p.len[d1];
    // read: fast along d1 (src unit stride)
#pragma unroll
    for (int r = 0; r < T; r += R) {
        const int64_t i1 = t1 * T + tx, i0 = t0 * T + ty + r;
        if (i0 < L0 && i1 < L1) tile[ty + r][tx] = src[os + i1 * p.ssrc[d1] + i0 * p.ssrc[d0]];
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < T; r += R) {
        const int64_t i0 = t0 * T + tx, i1 = t1 * T + ty + r;
        if (i0 < L0 && i1 < L1) out.put(od + i0 * p.sdst[d0] + i1 * p.sdst[d1], tile[tx][ty + r]);
    }
}

template <class Out>
static void run_strided(aq_ctx* ctx, const StridedPlan& sp, const double* src, const Out& out, bool read_src,
                        const char* name, double bytes) {
    if (sp.nout == 0) return;
    PlanDev p = to_dev(sp);
    int d0 = -1, d1 = -1;
    for (int k = 0; k < p.nd; ++k) {
        if (p.sdst[k] == 1 && d0 < 0) d0 = k;
        if (p.ssrc[k] == 1 && d1 < 0) d1 = k;
    }
    TimedLaunch tl(ctx, name, 0.0, bytes);
    if (read_src && p.nr == 0 && d0 >= 0 && d1 >= 0 && d0 != d1 && p.len[d0] >= 8 && p.len[d1] >= 8) {
        constexpr int T = TransposeTile<Out>::T;
        const int64_t tiles0 = (p.len[d0] + T - 1) / T, tiles1 = (p.len[d1] + T - 1) / T;
        const int64_t slabs = p.nout / (p.len[d0] * p.len[d1]);
        const int64_t nb = tiles0 * tiles1 * slabs;
        const int64_t gx = std::min<int64_t>(nb, 65535);
        const int64_t gy = (nb + gx - 1) / gx;
        AQ_CHECK(gy < 65536, AQ_ERR_UNSUPPORTED, "transpose grid too large");
        hipLaunchKernelGGL((k_transpose<Out>), dim3(unsigned(gx), unsigned(gy)), dim3(TPB), 0, ctx->stream, p, d0,
                           d1, tiles0, tiles1, src, out);
    } else {
        hipLaunchKernelGGL((k_strided<Out>), dim3(grid_for(p.nout, TPB * 4)), dim3(TPB), 0, ctx->stream, p, src, out,
                           read_src);
    }
    AQ_HIP(hipGetLastError());
}

void launch_strided_sum(aq_ctx* ctx, const StridedPlan& p, double alpha, const double* src, double beta,
                        double* dst) {
    SumOut o{dst, alpha, beta};
    const double bytes = 8.0 * (double(p.nout) * p.nred + p.nout * (beta != 0.0 ? 2 : 1));
    run_strided(ctx, p, src, o, alpha != 0.0, "strided_sum", bytes);
}

void launch_strided_convert(aq_ctx* ctx, const StridedPlan& p, const double* src, void* dst, Fmt fmt,
                            int64_t plane_stride) {
    AQ_CHECK(p.nr == 0, AQ_ERR_INVALID, "conversion cannot reduce");
    if (fmt == FMT_F64) {
        SumOut o{static_cast<double*>(dst), 1.0, 0.0};
        run_strided(ctx, p, src, o, true, "permute_f64", 16.0 * p.nout);
    } else if (fmt == FMT_F32) {
        F32Out o{static_cast<float*>(dst)};
        run_strided(ctx, p, src, o, true, "convert_f32", 12.0 * p.nout);
    } else {
        BF16X2Out o{static_cast<unsigned short*>(dst), plane_stride};
        run_strided(ctx, p, src, o, true, "split_bf16x2", 12.0 * p.nout);
    }
}

// -------------------------------------------------------------------- fill

__global__ __launch_bounds__(TPB) void k_fill(double* x, int64_t n, double alpha, double beta) {
    for (int64_t i = int64_t(blockIdx.x) * TPB + threadIdx.x; i < n; i += int64_t(gridDim.x) * TPB)
        x[i] = beta == 0.0 ? alpha : beta * x[i] + alpha;
}

void launch_fill(aq_ctx* ctx, double* x, int64_t n, double alpha, double beta) {
    if (n <= 0) return;
    TimedLaunch tl(ctx, "fill", 0.0, 8.0 * n * (beta != 0.0 ? 2 : 1));
    hipLaunchKernelGGL(k_fill, dim3(grid_for(n, TPB * 4)), dim3(TPB), 0, ctx->stream, x, n, alpha, beta);
    AQ_HIP(hipGetLastError());
}

// Zero `bytes` (a multiple of 4) with a kernel: the clear stays an ordinary
// kernel node when the stream is being captured into a graph.
__global__ __launch_bounds__(TPB) void k_zero_u32(uint32_t* x, int64_t n) {
    for (int64_t i = int64_t(blockIdx.x) * TPB + threadIdx.x; i < n; i += int64_t(gridDim.x) * TPB) x[i] = 0u;
}

void launch_zero_bytes(aq_ctx* ctx, void* p, size_t bytes) {
    AQ_CHECK(bytes % 4 == 0, AQ_ERR_INVALID, "zero fill needs a multiple of 4 bytes");
    const int64_t n = int64_t(bytes / 4);
    if (n <= 0) return;
    TimedLaunch tl(ctx, "zero", 0.0, double(bytes));
    hipLaunchKernelGGL(k_zero_u32, dim3(grid_for(n, TPB * 4)), dim3(TPB), 0, ctx->stream, static_cast<uint32_t*>(p), n);
    AQ_HIP(hipGetLastError());
}

// ------------------------------------------------------------- div / invert

// C = beta*C + alpha*A/B (A null: alpha/B); untouched where |B| <= DBL_MIN
__global__ __launch_bounds__(TPB) void k_div(const double* __restrict__ A, const double* __restrict__ B, double* C,
                                             int64_t n, double alpha, double beta) {
    for (int64_t i = int64_t(blockIdx.x) * TPB + threadIdx.x; i < n; i += int64_t(gridDim.x) * TPB) {
        const double b = B[i];
        if (fabs(b) > 2.2250738585072014e-308) {
            const double q = A ? alpha * A[i] / b : alpha / b;
            C[i] = beta * C[i] + q;
        }
    }
}

void launch_div(aq_ctx* ctx, const double* A, const double* B, double* C, int64_t n, double alpha, double beta) {
    if (n <= 0) return;
    TimedLaunch tl(ctx, "div", 0.0, 8.0 * n * (A ? 4 : 3));
    hipLaunchKernelGGL(k_div, dim3(grid_for(n, TPB * 4)), dim3(TPB), 0, ctx->stream, A, B, C, n, alpha, beta);
    AQ_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ weight

struct WeightArgs {
    int nd;
    int64_t len[MAXDIM];
    int64_t off[MAXDIM];
};

// x /= (d0[i0] + d1[i1] + ... + shift), zero where |.| < 1e-4 (CTFTensor::weight)
__global__ __launch_bounds__(TPB) void k_weight(double* x, int64_t n, WeightArgs w, const double* __restrict__ d,
                                                 double shift) {
    for (int64_t i = int64_t(blockIdx.x) * TPB + threadIdx.x; i < n; i += int64_t(gridDim.x) * TPB) {
        int64_t r = i;
        double den = 0.0;
        for (int k = 0; k < w.nd; ++k) {
            const int64_t c = r % w.len[k];
            r /= w.len[k];
            den += d[w.off[k] + c];
        }
        const double s = den + shift;
        x[i] = fabs(s) < 1e-4 ? 0.0 : x[i] / s;
    }
}

void launch_weight(aq_ctx* ctx, double* x, int ndim, const int64_t* len, const double* d_dev, const int64_t* d_off,
                   double shift, bool timed) {
    WeightArgs w;
    w.nd = ndim;
    int64_t n = 1;
    for (int k = 0; k < MAXDIM; ++k) {
        w.len[k] = k < ndim ? len[k] : 1;
        w.off[k] = k < ndim ? d_off[k] : 0;
        if (k < ndim) n *= len[k];
    }
    if (n == 0) return;
    if (!timed) {   // no event objects: the allocation-free entry point
        hipLaunchKernelGGL(k_weight, dim3(grid_for(n, TPB * 4)), dim3(TPB), 0, ctx->stream, x, n, w, d_dev, shift);
        AQ_HIP(hipGetLastError());
        return;
    }
    TimedLaunch tl(ctx, "weight", 0.0, 16.0 * n);
    hipLaunchKernelGGL(k_weight, dim3(grid_for(n, TPB * 4)), dim3(TPB), 0, ctx->stream, x, n, w, d_dev, shift);
    AQ_HIP(hipGetLastError());
}

// -------------------------------------------------------------- reductions
//
// Deterministic: a fixed number of blocks (a function of n only) each write
// one partial; one block combines the partials in a fixed tree order.

constexpr int RB = 1024;   // max partial blocks per slot

enum RedOp { RED_SUM = 0, RED_MAX = 1 };

__device__ __forceinline__ double red(double a, double b, int op) { return op == RED_MAX ? fmax(a, b) : a + b; }

__device__ double block_reduce(double v, int op) {
    __shared__ double sh[TPB];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = TPB / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] = red(sh[threadIdx.x], sh[threadIdx.x + s], op);
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

static unsigned red_blocks(int64_t n) { return grid_for(n, TPB * 8, RB); }

// mode: 0 dot, 1 max|x|, 2 sum|x|, 3 sum x^2
__global__ __launch_bounds__(TPB) void k_reduce_partial(const double* __restrict__ a, const double* __restrict__ b,
                                                         int64_t n, int mode, double* partial) {
    const int op = mode == 1 ? RED_MAX : RED_SUM;
    double acc = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * TPB + threadIdx.x; i < n; i += int64_t(gridDim.x) * TPB) {
        const double x = a[i];
        double v;
        if (mode == 0) v = x * b[i];
        else if (mode == 1 || mode == 2) v = fabs(x);
        else v = x * x;
        acc = red(acc, v, op);
    }
    const double r = block_reduce(acc, op);
    if (threadIdx.x == 0) partial[blockIdx.x] = r;
}

__global__ __launch_bounds__(TPB) void k_reduce_final(const double* partial, int np, int op, double* out) {
    double acc = 0.0;
    for (int i = threadIdx.x; i < np; i += TPB) acc = red(acc, partial[i], op);
    const double r = block_reduce(acc, op);
    if (threadIdx.x == 0) *out = r;
}

static double* partial_slot(aq_ctx* ctx, int slot) { return ctx->d_partials + int64_t(slot) * RB; }
static double* result_slot(aq_ctx* ctx, int slot) { return ctx->d_partials + int64_t(8) * RB + slot; }

static void finish(aq_ctx* ctx, int slot, unsigned np, int op) {
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(TPB), 0, ctx->stream, partial_slot(ctx, slot), int(np), op,
                       result_slot(ctx, slot));
    AQ_HIP(hipGetLastError());
    AQ_HIP(hipMemcpyAsync(ctx->h_result + slot, result_slot(ctx, slot), 8, hipMemcpyDeviceToHost, ctx->stream));
}

void launch_dot(aq_ctx* ctx, const double* a, const double* b, int64_t n, int slot) {
    const unsigned g = red_blocks(n);
    {
        TimedLaunch tl(ctx, "dot", 2.0 * n, 16.0 * n);
        hipLaunchKernelGGL(k_reduce_partial, dim3(g), dim3(TPB), 0, ctx->stream, a, b, n, 0, partial_slot(ctx, slot));
        AQ_HIP(hipGetLastError());
    }
    finish(ctx, slot, g, RED_SUM);
}

void launch_norm(aq_ctx* ctx, const double* x, int64_t n, int p, int slot) {
    const unsigned g = red_blocks(n);
    const int mode = p == 0 ? 1 : p == 1 ? 2 : 3;
    {
        TimedLaunch tl(ctx, "norm", 0.0, 8.0 * n);
        hipLaunchKernelGGL(k_reduce_partial, dim3(g), dim3(TPB), 0, ctx->stream, x, x, n, mode, partial_slot(ctx, slot));
        AQ_HIP(hipGetLastError());
    }
    finish(ctx, slot, g, p == 0 ? RED_MAX : RED_SUM);
    if (p == 2) {
        AQ_HIP(hipStreamSynchronize(ctx->stream));
        ctx->h_result[slot] = std::sqrt(ctx->h_result[slot]);
    }
}

double sync_result(aq_ctx* ctx, int slot) {
    AQ_HIP(hipStreamSynchronize(ctx->stream));
    resolve_timing(ctx);
    return ctx->h_result[slot];
}

// ---------------------------------------------------------------- key I/O

__global__ void k_gather(const double* x, int64_t n, const int64_t* keys, double* vals) {
    for (int64_t i = int64_t(blockIdx.x) * TPB + threadIdx.x; i < n; i += int64_t(gridDim.x) * TPB) vals[i] = x[keys[i]];
}

__global__ void k_scatter(double* x, int64_t n, const int64_t* keys, const double* vals, double alpha, double beta) {
    // duplicate keys: last writer wins per element order is not defined; the
    // host path serialises duplicates (see launch_scatter_keys)
    for (int64_t i = int64_t(blockIdx.x) * TPB + threadIdx.x; i < n; i += int64_t(gridDim.x) * TPB) {
        const double a = alpha * vals[i];
        x[keys[i]] = beta == 0.0 ? a : beta * x[keys[i]] + a;
    }
}

void launch_gather_keys(aq_ctx* ctx, const double* x, int64_t n, const int64_t* keys, double* vals) {
    hipLaunchKernelGGL(k_gather, dim3(grid_for(n, TPB)), dim3(TPB), 0, ctx->stream, x, n, keys, vals);
    AQ_HIP(hipGetLastError());
}

void launch_scatter_keys(aq_ctx* ctx, double* x, int64_t n, const int64_t* keys, const double* vals, double alpha,
                         double beta) {
    hipLaunchKernelGGL(k_scatter, dim3(grid_for(n, TPB)), dim3(TPB), 0, ctx->stream, x, n, keys, vals, alpha, beta);
    AQ_HIP(hipGetLastError());
}

// -------------------------------------------------------------- split-K

__global__ __launch_bounds__(TPB) void k_splitk(const double* __restrict__ P, int nsplit, int64_t m, int64_t n,
                                                 int64_t batch, double beta, double* C, int64_t ldc, int64_t strideC,
                                                 CMap map) {
    const int64_t mn = m * n, tot = mn * batch;
    for (int64_t i = int64_t(blockIdx.x) * TPB + threadIdx.x; i < tot; i += int64_t(gridDim.x) * TPB) {
        double s = 0.0;
        for (int k = 0; k < nsplit; ++k) s += P[int64_t(k) * tot + i];
        const int64_t b = i / mn, r = i % mn;
        const int64_t col = r / m, row = r % m;
        const int64_t off = map.mapped ? cmap_digits(map.nm, map.ml, map.ms, row) + cmap_digits(map.nn, map.nl, map.ns, col)
                                       : row + col * ldc;
        double* c = C + b * strideC + off;
        *c = beta == 0.0 ? s : beta * *c + s;
    }
}

void launch_reduce_splitk(aq_ctx* ctx, const double* P, int nsplit, int64_t m, int64_t n, int64_t batch, double beta,
                          double* C, int64_t ldc, int64_t strideC, const CMap& map) {
    const int64_t tot = m * n * batch;
    TimedLaunch tl(ctx, "splitk_reduce", double(tot) * nsplit, 8.0 * tot * (nsplit + 1));
    hipLaunchKernelGGL(k_splitk, dim3(grid_for(tot, TPB)), dim3(TPB), 0, ctx->stream, P, nsplit, m, n, batch, beta, C,
                       ldc, strideC, map);
    AQ_HIP(hipGetLastError());
}

// ---------------------------------------------------------- fused update
//
// rhfccsd.cxx:313-323 in two passes over the amplitudes:
//   pass 1 (v*o):      Z1 /= D1 (weight), T1 += Z1, partials of fAI.T1 and max|Z1|
//   pass 2 (v*v*o*o):  Z2 /= D2, T2 += Z2, Tau = T2 + T1(a,i) T1(b,j),
//                      partials of VABIJSA.Tau and max|Z2|
// The denominator is accumulated in the reference's order
// (((0 + dA[a]) + dA[b]) + dI[i]) + dI[j] and thresholded at 1e-4.

__global__ __launch_bounds__(TPB) void k_update1(UpdateArgs u, double* pe, double* pc) {
    const int64_t n = int64_t(u.v) * u.o;
    double e = 0.0, c = 0.0;
    for (int64_t idx = int64_t(blockIdx.x) * TPB + threadIdx.x; idx < n; idx += int64_t(gridDim.x) * TPB) {
        const int a = int(idx % u.v), i = int(idx / u.v);
        double den = 0.0;
        den += u.dA[a];
        den += u.dI[i];
        const double z = fabs(den) < 1e-4 ? 0.0 : u.Z1[idx] / den;
        u.Z1[idx] = z;
        const double t = u.T1[idx] + z;
        u.T1[idx] = t;
        e += u.fAI[idx] * t;
        c = fmax(c, fabs(z));
    }
    e = block_reduce(e, RED_SUM);
    c = block_reduce(c, RED_MAX);
    if (threadIdx.x == 0) {
        pe[blockIdx.x] = e;
        pc[blockIdx.x] = c;
    }
}

__global__ __launch_bounds__(TPB) void k_update2(UpdateArgs u, double* pe, double* pc) {
    const int64_t v = u.v, o = u.o;
    const int64_t n = v * v * o * o;
    double e = 0.0, c = 0.0;
    // (a, b, i, j) digits of the flat index in the mixed radix (v, v, o, o):
    // split once for the first element and for the grid stride, then advanced
    // by digit-wise addition with carries (no 64-bit division per element)
    const int64_t stride = int64_t(gridDim.x) * TPB;
    int64_t idx = int64_t(blockIdx.x) * TPB + threadIdx.x;
    auto digits = [&](int64_t x, int& d0, int& d1, int& d2, int64_t& d3) {
        d0 = int(x % v); x /= v;
        d1 = int(x % v); x /= v;
        d2 = int(x % o);
        d3 = x / o;
    };
    int a, b, i, sa, sb, si;
    int64_t j, sj;
    digits(idx, a, b, i, j);
    digits(stride, sa, sb, si, sj);
    // idx += stride, digit by digit (each digit sum < 2 * radix)
    auto step = [&]() {
        a += sa;
        int cy = a >= v;
        a -= cy ? int(v) : 0;
        b += sb + cy;
        cy = b >= v;
        b -= cy ? int(v) : 0;
        i += si + cy;
        cy = i >= o;
        i -= cy ? int(o) : 0;
        j += sj + cy;
    };
    // U consecutive elements of this thread per step: all their loads are
    // issued before any store (the arrays are distinct), then the elements
    // are finished in order, so e and c accumulate exactly as one at a time
    constexpr int U = 4;
    for (; idx < n; idx += U * stride) {
        int64_t p[U], dj[U];
        int da[U], db[U], di[U];
        double z2[U], t2[U], vs[U];
#pragma unroll
        for (int k = 0; k < U; ++k) {
            p[k] = idx + k * stride;
            da[k] = a;
            db[k] = b;
            di[k] = i;
            dj[k] = j;
            step();
        }
#pragma unroll
        for (int k = 0; k < U; ++k) {
            const bool ok = p[k] < n;
            z2[k] = ok ? u.Z2[p[k]] : 0.0;
            t2[k] = ok ? u.T2[p[k]] : 0.0;
            vs[k] = ok ? u.VABIJSA[p[k]] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < U; ++k) {
            if (p[k] >= n) break;
            double den = 0.0;
            den += u.dA[da[k]];
            den += u.dA[db[k]];
            den += u.dI[di[k]];
            den += u.dI[dj[k]];
            const double z = fabs(den) < 1e-4 ? 0.0 : z2[k] / den;
            u.Z2[p[k]] = z;
            const double t = t2[k] + z;
            u.T2[p[k]] = t;
            const double tau = t + u.T1[da[k] + di[k] * v] * u.T1[db[k] + dj[k] * v];
            u.Tau[p[k]] = tau;
            e += vs[k] * tau;
            c = fmax(c, fabs(z));
        }
    }
    e = block_reduce(e, RED_SUM);
    c = block_reduce(c, RED_MAX);
    if (threadIdx.x == 0) {
        pe[blockIdx.x] = e;
        pc[blockIdx.x] = c;
    }
}

void launch_ccsd_update(aq_ctx* ctx, const UpdateArgs& u, int slot_e1, int slot_e2, int slot_c1, int slot_c2) {
    const int64_t n1 = int64_t(u.v) * u.o;
    const int64_t n2 = n1 * n1;
    const unsigned g1 = red_blocks(n1), g2 = red_blocks(n2);
    {
        TimedLaunch tl(ctx, "ccsd_update_t1", 0.0, 8.0 * 5 * n1);
        hipLaunchKernelGGL(k_update1, dim3(g1), dim3(TPB), 0, ctx->stream, u, partial_slot(ctx, slot_e1),
                           partial_slot(ctx, slot_c1));
        AQ_HIP(hipGetLastError());
    }
    finish(ctx, slot_e1, g1, RED_SUM);
    finish(ctx, slot_c1, g1, RED_MAX);
    {
        // algorithmic bytes: read Z2, T2, VABIJSA; write Z2, T2, Tau
        TimedLaunch tl(ctx, "ccsd_update_t2", 0.0, 8.0 * 6 * n2);
        hipLaunchKernelGGL(k_update2, dim3(g2), dim3(TPB), 0, ctx->stream, u, partial_slot(ctx, slot_e2),
                           partial_slot(ctx, slot_c2));
        AQ_HIP(hipGetLastError());
    }
    finish(ctx, slot_e2, g2, RED_SUM);
    finish(ctx, slot_c2, g2, RED_MAX);
}

// ------------------------------------------------------- packed ladder
//
// The particle-particle ladder S(ab,ij) = sum_ef V(ab,ef) Tau(ef,ij) with
// V(ab,ef) = V(ba,fe) and Tau(ef,ij) = Tau(fe,ji) splits into a part
// symmetric and a part antisymmetric under e<->f (Scuseria, Janssen, Schaefer,
// J. Chem. Phys. 89, 7382 (1988)):
//   S+(ab,ij) = sum_{e<=f} V+(ab,ef) T+(ef,ij),  a<=b, i<=j
//   S-(ab,ij) = sum_{e<f}  V-(ab,ef) T-(ef,ij),  a<b,  i<j
//   V+(ab,ef) = V(ab,ef) + V(ab,fe) (V(ab,ee) on the diagonal), V- = V(ab,ef) - V(ab,fe)
//   T+-(ef,ij) = (Tau(ef,ij) +- Tau(fe,ij)) / 2
//   S(ab,ij) = S+(ab,ij) + S-(ab,ij), S+ symmetric and S- antisymmetric
//   under a<->b and under i<->j.
// Packed pair indices: tri_le(a,b) = a + b(b+1)/2 for a <= b,
// tri_lt(a,b) = a + b(b-1)/2 for a < b (column b holds the rows a).

__device__ __forceinline__ int64_t tri_le(int64_t a, int64_t b) { return a + b * (b + 1) / 2; }
__device__ __forceinline__ int64_t tri_lt(int64_t a, int64_t b) { return a + b * (b - 1) / 2; }

// inverse of tri_le: r -> (a, b), a <= b
__device__ __forceinline__ void untri_le(int64_t r, int64_t& a, int64_t& b) {
    int64_t bb = int64_t((sqrt(8.0 * double(r) + 1.0) - 1.0) * 0.5);
    while (bb * (bb + 1) / 2 > r) --bb;
    while ((bb + 1) * (bb + 2) / 2 <= r) ++bb;
    b = bb;
    a = r - bb * (bb + 1) / 2;
}

// V+ [P x P] and V- [Q x Q] (column-major) from V[v,v,v,v] (a fastest);
// grid.y walks the (e <= f) columns, threads the (a <= b) rows.
__global__ __launch_bounds__(TPB) void k_ladder_pack_v(const double* __restrict__ V, int64_t v, double* __restrict__ Vp,
                                                       double* __restrict__ Vm) {
    const int64_t P = v * (v + 1) / 2, Q = v * (v - 1) / 2;
    for (int64_t c = blockIdx.y; c < P; c += gridDim.y) {
        int64_t e, f;
        untri_le(c, e, f);
        const double* Vef = V + (e + f * v) * v * v;
        const double* Vfe = V + (f + e * v) * v * v;
        for (int64_t r = int64_t(blockIdx.x) * TPB + threadIdx.x; r < P; r += int64_t(gridDim.x) * TPB) {
            int64_t a, b;
            untri_le(r, a, b);
            const double x = Vef[a + b * v], y = Vfe[a + b * v];
            Vp[r + c * P] = e == f ? x : x + y;
            if (a < b && e < f) Vm[tri_lt(a, b) + tri_lt(e, f) * Q] = x - y;
        }
    }
}

// T+ [O+ x P] and T- [O- x Q] from Tau[v,v,o,o]: the occupied pair index is
// the contiguous one, so the ladder GEMM S = V T^T reads both operands
// row-contiguous.  Each block packs a tile of 64 (e <= f) pairs x 32 (i <= j)
// pairs: Tau is read along e (its contiguous index) and the tile is written
// through LDS along the pair (i,j).
constexpr int PK_R = 64, PK_C = 32;
__global__ __launch_bounds__(TPB) void k_ladder_pack_t(const double* __restrict__ T, int64_t v, int64_t o,
                                                       double* __restrict__ Tp, double* __restrict__ Tm) {
    static_assert(TPB == 256 && PK_R * PK_C == 8 * TPB, "tile of 8 elements per thread");
    __shared__ double sp[PK_R][PK_C + 1], sm[PK_R][PK_C + 1];
    const int64_t P = v * (v + 1) / 2, Op = o * (o + 1) / 2, Om = o * (o - 1) / 2;
    const int64_t r0 = int64_t(blockIdx.x) * PK_R, c0 = int64_t(blockIdx.y) * PK_C;
    const int tid = threadIdx.x;
    {
        // read: thread -> one (e,f) row, eight (i,j) columns
        const int rl = tid % PK_R;
        const int64_t r = r0 + rl;
        int64_t e = 0, f = 0;
        if (r < P) untri_le(r, e, f);
#pragma unroll
        for (int s = 0; s < PK_C / (TPB / PK_R); ++s) {
            const int cl = tid / PK_R + s * (TPB / PK_R);
            const int64_t c = c0 + cl;
            double x = 0.0, y = 0.0;
            if (r < P && c < Op) {
                int64_t i, j;
                untri_le(c, i, j);
                const double* Tij = T + (i + j * o) * v * v;
                x = Tij[e + f * v];
                y = Tij[f + e * v];
            }
            sp[rl][cl] = 0.5 * (x + y);
            sm[rl][cl] = 0.5 * (x - y);
        }
    }
    __syncthreads();
    {
        // write: thread -> one (i,j) column, eight (e,f) rows
        const int cl = tid % PK_C;
        const int64_t c = c0 + cl;
        int64_t i = 0, j = 0;
        if (c < Op) untri_le(c, i, j);
#pragma unroll
        for (int s = 0; s < PK_R / (TPB / PK_C); ++s) {
            const int rl = tid / PK_C + s * (TPB / PK_C);
            const int64_t r = r0 + rl;
            if (r >= P || c >= Op) continue;
            Tp[c + r * Op] = sp[rl][cl];
            if (i < j) {
                int64_t e, f;
                untri_le(r, e, f);
                // tri_lt(i,j) = tri_le(i,j) - j, tri_lt(e,f) = tri_le(e,f) - f
                if (e < f) Tm[(c - j) + (r - f) * Om] = sm[rl][cl];
            }
        }
    }
}

// Tau, TauSA and T2SA of one iteration in one pass (rhfccsd.cxx:231-238):
//   Tau(ab,ij)   = T2(ab,ij) + T1(a,i) T1(b,j)
//   TauSA(ab,ij) = 2 Tau(ab,ij) - Tau(ab,ji)
//   T2SA(ab,ij)  = 2 T2(ab,ij) - T2(ab,ji)
// Each block row handles the slab pair (i,j), (j,i) with i <= j, so T2 is
// read once and both slabs' outputs are written from registers.  Every value
// is rounded as the operator sequence rounds it (the product, then the sum;
// 2x exactly, then one rounded difference): contraction into FMAs is off.
__global__ __launch_bounds__(TPB) void k_tau_sa(const double* __restrict__ T1, const double* __restrict__ T2,
                                                int64_t v, int64_t o, double* __restrict__ Tau,
                                                double* __restrict__ TauSA, double* __restrict__ T2SA) {
#pragma clang fp contract(off)
    const int64_t vv = v * v, np = o * (o + 1) / 2;
    for (int64_t c = blockIdx.y; c < np; c += gridDim.y) {
        int64_t i, j;
        untri_le(c, i, j);
        const int64_t sij = (i + j * o) * vv, sji = (j + i * o) * vv;
        for (int64_t r = int64_t(blockIdx.x) * TPB + threadIdx.x; r < vv; r += int64_t(gridDim.x) * TPB) {
            const int64_t a = r % v, b = r / v;
            const double x = T2[sij + r], y = T2[sji + r];
            const double pij = T1[a + i * v] * T1[b + j * v];
            const double pji = T1[a + j * v] * T1[b + i * v];
            const double tij = x + pij, tji = y + pji;
            Tau[sij + r] = tij;
            TauSA[sij + r] = 2.0 * tij - tji;
            T2SA[sij + r] = 2.0 * x - y;
            if (i != j) {
                Tau[sji + r] = tji;
                TauSA[sji + r] = 2.0 * tji - tij;
                T2SA[sji + r] = 2.0 * y - x;
            }
        }
    }
}

// Z[v,v,o,o] += alpha * (S+ + S-) expanded to all (a,b,i,j):
//   Z(ab,ij) += S+ + S-,  Z(ba,ij) += S+ - S-,  Z(ab,ji) += S+ - S-,  Z(ba,ji) += S+ + S-
// for a < b, i < j (S- vanishes when a = b or i = j).  A block takes a 32 x 32
// tile of (a,b) pairs with a-tile <= b-tile and one pair i <= j: it reads the
// tile's packed values along a (contiguous in S+- for a <= b), keeps them in
// LDS and writes the (a,b) tile along a and its mirror (b,a) along b, for
// both slabs (i,j) and (j,i).  Every Z element is written by one block.  A
// diagonal tile (a-tile = b-tile) writes only its own (a,b) positions, reading
// the packed value of (min, max).
constexpr int UP_T = 32;
__global__ __launch_bounds__(TPB) void k_ladder_unpack(const double* __restrict__ Sp, const double* __restrict__ Sm,
                                                       int64_t v, int64_t o, double alpha, double* __restrict__ Z) {
    static_assert(TPB % UP_T == 0, "whole tile rows per pass");
    constexpr int R = TPB / UP_T;
    __shared__ double lp[UP_T][UP_T + 1], lm[UP_T][UP_T + 1];
    const int64_t P = v * (v + 1) / 2, Q = v * (v - 1) / 2;
    const int64_t vv = v * v, np = o * (o + 1) / 2;
    // upper tile pair (ta <= tb) of this block
    int64_t ta, tb;
    untri_le(blockIdx.x, ta, tb);
    const int64_t a0 = ta * UP_T, b0 = tb * UP_T;
    const bool diag = ta == tb;
    const int tx = threadIdx.x % UP_T, ty = threadIdx.x / UP_T;
    for (int64_t c = blockIdx.y; c < np; c += gridDim.y) {
        int64_t i, j;
        untri_le(c, i, j);
        const double* Spc = Sp + c * P;                                       // tri_le(i,j) = c
        const double* Smc = i < j ? Sm + (c - j) * Q : nullptr;               // tri_lt(i,j) = c - j
        __syncthreads();   // the previous pair's tile has been written out
        // load: lp/lm[al][bl] for the tile's (a,b); a fastest
#pragma unroll
        for (int r = 0; r < UP_T; r += R) {
            const int al = tx, bl = ty + r;
            const int64_t a = a0 + al, b = b0 + bl;
            double sp = 0.0, sm = 0.0;
            if (a < v && b < v && (!diag || al <= bl)) {
                sp = Spc[tri_le(a, b)];
                if (Smc && a < b) sm = Smc[tri_lt(a, b)];
            } else if (diag && a < v && b < v) {
                sp = Spc[tri_le(b, a)];
                if (Smc) sm = -Smc[tri_lt(b, a)];                             // S-(ab) = -S-(ba)
            }
            lp[al][bl] = sp;
            lm[al][bl] = sm;
        }
        __syncthreads();
        double* Zij = Z + (i + j * o) * vv;
        double* Zji = Z + (j + i * o) * vv;
        // (a,b) positions, a fastest: Z(ab,ij) += S+ + S-, Z(ab,ji) += S+ - S-
#pragma unroll
        for (int r = 0; r < UP_T; r += R) {
            const int al = tx, bl = ty + r;
            const int64_t a = a0 + al, b = b0 + bl;
            if (a < v && b < v) {
                const double sp = lp[al][bl], sm = lm[al][bl];
                Zij[a + b * v] += alpha * (sp + sm);
                if (i != j) Zji[a + b * v] += alpha * (sp - sm);
            }
        }
        if (!diag) {
            // mirror (b,a) positions, b fastest: Z(ba,ij) += S+ - S-, Z(ba,ji) += S+ + S-
#pragma unroll
            for (int r = 0; r < UP_T; r += R) {
                const int bl = tx, al = ty + r;
                const int64_t a = a0 + al, b = b0 + bl;
                if (a < v && b < v) {
                    const double sp = lp[al][bl], sm = lm[al][bl];
                    Zij[b + a * v] += alpha * (sp - sm);
                    if (i != j) Zji[b + a * v] += alpha * (sp + sm);
                }
            }
        }
    }
}

// max |V(a,b,e,f) - V(b,a,f,e)| into result slot `slot` (caller syncs)
__global__ __launch_bounds__(TPB) void k_pair_asym(const double* __restrict__ X, int64_t n1, int64_t n2,
                                                   double* partial) {
    // X[n1,n1,n2,n2]: deviation from X(p,q,r,s) = X(q,p,s,r)
    const int64_t n = n1 * n1 * n2 * n2;
    double acc = 0.0;
    for (int64_t k = int64_t(blockIdx.x) * TPB + threadIdx.x; k < n; k += int64_t(gridDim.x) * TPB) {
        int64_t t = k;
        const int64_t p = t % n1; t /= n1;
        const int64_t q = t % n1; t /= n1;
        const int64_t r = t % n2, s = t / n2;
        acc = fmax(acc, fabs(X[k] - X[q + n1 * (p + n1 * (s + n2 * r))]));
    }
    const double m = block_reduce(acc, RED_MAX);
    if (threadIdx.x == 0) partial[blockIdx.x] = m;
}

void launch_ladder_pack_v(aq_ctx* ctx, const double* V, int v, double* Vp, double* Vm) {
    const int64_t P = int64_t(v) * (v + 1) / 2;
    TimedLaunch tl(ctx, "ladder_pack_v", 0.0, 8.0 * (2.0 * P * P + 2.0 * P * P));
    const unsigned gx = unsigned(std::min<int64_t>((P + TPB - 1) / TPB, 64));
    const unsigned gy = unsigned(std::min<int64_t>(P, 65535));
    hipLaunchKernelGGL(k_ladder_pack_v, dim3(gx, gy), dim3(TPB), 0, ctx->stream, V, int64_t(v), Vp, Vm);
    AQ_HIP(hipGetLastError());
}

void launch_ladder_pack_t(aq_ctx* ctx, const double* T, int v, int o, double* Tp, double* Tm) {
    const int64_t P = int64_t(v) * (v + 1) / 2, Op = int64_t(o) * (o + 1) / 2;
    TimedLaunch tl(ctx, "ladder_pack_t", 0.0, 8.0 * (2.0 * P * Op + 2.0 * P * Op));
    const int64_t gx = (P + PK_R - 1) / PK_R, gy = (Op + PK_C - 1) / PK_C;
    AQ_CHECK(gx < (int64_t(1) << 31) && gy < 65536, AQ_ERR_UNSUPPORTED, "packed ladder too large");
    hipLaunchKernelGGL(k_ladder_pack_t, dim3(unsigned(gx), unsigned(gy)), dim3(TPB), 0, ctx->stream, T, int64_t(v),
                       int64_t(o), Tp, Tm);
    AQ_HIP(hipGetLastError());
}

void launch_tau_sa(aq_ctx* ctx, const double* T1, const double* T2, int v, int o, double* Tau, double* TauSA,
                   double* T2SA) {
    const int64_t vv = int64_t(v) * v, np = int64_t(o) * (o + 1) / 2;
    TimedLaunch tl(ctx, "tau_sa", 0.0, 8.0 * 4.0 * vv * o * o);
    const unsigned gx = unsigned(std::min<int64_t>((vv + TPB - 1) / TPB, 64));
    const unsigned gy = unsigned(std::min<int64_t>(np, 65535));
    hipLaunchKernelGGL(k_tau_sa, dim3(gx, gy), dim3(TPB), 0, ctx->stream, T1, T2, int64_t(v), int64_t(o), Tau, TauSA,
                       T2SA);
    AQ_HIP(hipGetLastError());
}

void launch_ladder_unpack(aq_ctx* ctx, const double* Sp, const double* Sm, int v, int o, double alpha, double* Z) {
    const int64_t vv = int64_t(v) * v, oo = int64_t(o) * o;
    TimedLaunch tl(ctx, "ladder_unpack", 0.0, 8.0 * 3.0 * vv * oo);
    const int64_t nt = (int64_t(v) + UP_T - 1) / UP_T;
    const int64_t gx = nt * (nt + 1) / 2, np = int64_t(o) * (o + 1) / 2;
    AQ_CHECK(gx < (int64_t(1) << 31), AQ_ERR_UNSUPPORTED, "ladder unpack grid too large");
    const unsigned gy = unsigned(std::min<int64_t>(np, 65535));
    hipLaunchKernelGGL(k_ladder_unpack, dim3(unsigned(gx), gy), dim3(TPB), 0, ctx->stream, Sp, Sm, int64_t(v),
                       int64_t(o), alpha, Z);
    AQ_HIP(hipGetLastError());
}

double pair_asymmetry(aq_ctx* ctx, const double* X, int n1, int n2, int slot) {
    AQ_CHECK(slot >= 0 && slot < 8, AQ_ERR_INVALID, "reduction slot out of range");
    const int64_t n = int64_t(n1) * n1 * n2 * n2;
    if (n == 0) return 0.0;
    const unsigned g = red_blocks(n);
    hipLaunchKernelGGL(k_pair_asym, dim3(g), dim3(TPB), 0, ctx->stream, X, int64_t(n1), int64_t(n2),
                       partial_slot(ctx, slot));
    AQ_HIP(hipGetLastError());
    finish(ctx, slot, g, RED_MAX);
    return sync_result(ctx, slot);
}

// ------------------------------------------------------------ AO integrals
//
// (pq|rs) [n,n,n,n] (p fastest) from the compact canonical shell-quartet
// storage (fe::CompactEri): the shell pair of (p,q) and of (r,s) is ordered
// larger shell first, the pair with the larger index is the bra, and the value
// is element [fA][fB][fC][fD] of block (bra, ket).  Indices at or past `total`
// read as zero (the layout never produces them; the guard keeps a corrupted
// table from reading out of bounds).
struct EriTables {
    const int* shell_of;
    const int* fn_of;
    const int* nf;
    const int64_t* pair_size;
    const int64_t* prefix;
    const int64_t* rowstart;
};

__global__ __launch_bounds__(TPB) void k_eri_expand(const double* __restrict__ compact, EriTables t, int64_t n,
                                                    int64_t total, double* __restrict__ out) {
    const int64_t n4 = n * n * n * n;
    for (int64_t k = int64_t(blockIdx.x) * TPB + threadIdx.x; k < n4; k += int64_t(gridDim.x) * TPB) {
        int64_t x = k;
        const int64_t p = x % n; x /= n;
        const int64_t q = x % n; x /= n;
        const int64_t r = x % n, s = x / n;
        int A = t.shell_of[p], fA = t.fn_of[p], B = t.shell_of[q], fB = t.fn_of[q];
        int Cs = t.shell_of[r], fC = t.fn_of[r], D = t.shell_of[s], fD = t.fn_of[s];
        if (A < B) { int u = A; A = B; B = u; u = fA; fA = fB; fB = u; }
        if (Cs < D) { int u = Cs; Cs = D; D = u; u = fC; fC = fD; fD = u; }
        // within a pair of one shell and between bra and ket of one pair, the
        // larger function index first: all eight permutations of (p,q,r,s)
        // read the same stored value, so the tensor is exactly symmetric
        if (A == B && fA < fB) { int u = fA; fA = fB; fB = u; }
        if (Cs == D && fC < fD) { int u = fC; fC = fD; fD = u; }
        int64_t P1 = int64_t(A) * (A + 1) / 2 + B, P2 = int64_t(Cs) * (Cs + 1) / 2 + D;
        if (P1 < P2 || (P1 == P2 && (fA < fC || (fA == fC && fB < fD)))) {
            int u = A; A = Cs; Cs = u; u = B; B = D; D = u;
            u = fA; fA = fC; fC = u; u = fB; fB = fD; fD = u;
            const int64_t w = P1; P1 = P2; P2 = w;
        }
        const int64_t idx = t.rowstart[P1] + t.pair_size[P1] * t.prefix[P2] +
                            ((int64_t(fA) * t.nf[B] + fB) * t.nf[Cs] + fC) * t.nf[D] + fD;
        out[k] = idx < total ? compact[idx] : 0.0;
    }
}

void launch_eri_expand(aq_ctx* ctx, const double* compact, const int* shell_of, const int* fn_of, const int* nf,
                       const int64_t* pair_size, const int64_t* prefix, const int64_t* rowstart, int64_t n,
                       int64_t total, double* out) {
    const int64_t n4 = n * n * n * n;
    if (n4 == 0) return;
    TimedLaunch tl(ctx, "eri_expand", 0.0, 16.0 * double(n4));
    EriTables t{shell_of, fn_of, nf, pair_size, prefix, rowstart};
    hipLaunchKernelGGL(k_eri_expand, dim3(grid_for(n4, TPB * 4)), dim3(TPB), 0, ctx->stream, compact, t, n, total, out);
    AQ_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ DIIS

constexpr int MAXK = 16;

struct PtrList {
    const double* p[MAXK];
};

// partial[slot_k][block] = sum x*y_k for k < nk (x read once)
__global__ __launch_bounds__(TPB) void k_multi_dot(const double* __restrict__ x, PtrList ys, int nk, int64_t n,
                                                    double* partials) {
    double acc[MAXK];
#pragma unroll
    for (int k = 0; k < MAXK; ++k) acc[k] = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * TPB + threadIdx.x; i < n; i += int64_t(gridDim.x) * TPB) {
        const double xv = x[i];
#pragma unroll
        for (int k = 0; k < MAXK; ++k)
            if (k < nk) acc[k] += xv * ys.p[k][i];
    }
    for (int k = 0; k < nk; ++k) {
        const double r = block_reduce(acc[k], RED_SUM);
        if (threadIdx.x == 0) partials[int64_t(k) * RB + blockIdx.x] = r;
    }
}

void launch_multi_dot(aq_ctx* ctx, const double* x, const double* const* ys, int k, int64_t n, int first_slot) {
    AQ_CHECK(k >= 1 && k <= MAXK && first_slot + k <= 8, AQ_ERR_INVALID, "multi-dot arity");
    PtrList pl{};
    for (int i = 0; i < k; ++i) pl.p[i] = ys[i];
    const unsigned g = red_blocks(n);
    {
        TimedLaunch tl(ctx, "diis_dot", 2.0 * n * k, 8.0 * n * (k + 1));
        hipLaunchKernelGGL(k_multi_dot, dim3(g), dim3(TPB), 0, ctx->stream, x, pl, k, n, partial_slot(ctx, first_slot));
        AQ_HIP(hipGetLastError());
    }
    for (int i = 0; i < k; ++i) finish(ctx, first_slot + i, g, RED_SUM);
}

// out = sum_k c_k * xs_k, accumulated in the reference's order
// (out = x0*c0; out += x_k*c_k, diis.hpp:205-226).  More than MAXK terms run
// as several launches; each later launch continues from `out` (cont = 1), so
// every element is summed left to right exactly as in one pass.
struct CoefList {
    double c[MAXK];
};

__global__ __launch_bounds__(TPB) void k_lincomb(double* out, PtrList xs, CoefList c, int nk, int64_t n, int cont) {
    for (int64_t i = int64_t(blockIdx.x) * TPB + threadIdx.x; i < n; i += int64_t(gridDim.x) * TPB) {
        double s = cont ? out[i] + xs.p[0][i] * c.c[0] : xs.p[0][i] * c.c[0];
        for (int k = 1; k < nk; ++k) s = s + xs.p[k][i] * c.c[k];
        out[i] = s;
    }
}

void launch_lincomb(aq_ctx* ctx, double* out, const double* const* xs, const double* c, int k, int64_t n) {
    AQ_CHECK(k >= 1, AQ_ERR_INVALID, "lincomb arity");
    for (int first = 0; first < k; first += MAXK) {
        const int kk = std::min(MAXK, k - first);
        PtrList pl{};
        CoefList cl{};
        for (int i = 0; i < kk; ++i) {
            pl.p[i] = xs[first + i];
            cl.c[i] = c[first + i];
        }
        TimedLaunch tl(ctx, "diis_lincomb", 2.0 * n * kk, 8.0 * n * (kk + 1 + (first > 0)));
        hipLaunchKernelGGL(k_lincomb, dim3(grid_for(n, TPB * 4)), dim3(TPB), 0, ctx->stream, out, pl, cl, kk, n,
                           int(first > 0));
        AQ_HIP(hipGetLastError());
    }
    AQ_HIP(hipGetLastError());
}

}  // namespace aqamd
