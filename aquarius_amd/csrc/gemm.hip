// MFMA GEMMs for gfx950 (CDNA4, 64-wide wavefronts).
//
//   C[m + n*ldc] = alpha * sum_k opA(m,k) * opB(k,n) + beta * C[m + n*ldc]
//   opA(m,k) = transA ? A[k + m*lda] : A[m + k*lda]
//   opB(k,n) = transB ? B[n + k*ldb] : B[k + n*ldb]
// with a batch dimension (grid z) and optional split-K (partials in FP64,
// reduced in a fixed order, so results do not depend on scheduling).
//
// FP64 path: v_mfma_f64_16x16x4f64.  FP32 path: v_mfma_f32_16x16x4f32.
// Both share one template: 256 threads = 4 waves, block tile BM x BN, BK = 16,
// operands staged through double-buffered LDS (one barrier per k-tile), each
// wave owning a (BM/WM) x (BN/WN) sub-tile of 16x16 MFMA blocks.
// The MFMA's 16x4 "src0" operand is fed from the B tile and its 4x16 "src1"
// from the A tile, so the 16x16 result lands with consecutive lanes on
// consecutive rows m of column-major C (coalesced epilogue).
//
// BF16x3 path: each FP64 operand is split into hi = bf16(x), lo = bf16(x-hi)
// planes; A*B ~ Ahi*Bhi + Ahi*Blo + Alo*Bhi with v_mfma_f32_16x16x32_bf16 and
// FP32 accumulation.  Its operands are always K-contiguous copies with a
// leading dimension padded to a multiple of 8 (16-byte vector loads).
//
// Tile order: the 1-D block id is remapped so that the blocks that land on
// one XCD (ids are dealt round-robin over the 8 XCDs) cover consecutive tiles,
// and consecutive tiles walk groups of 8 tile rows with m fastest, so one
// XCD's co-resident blocks cover a square of tiles and share row panels of A
// and column panels of B in that XCD's L2.
#include "internal.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>

namespace aqamd {

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef short i16x8 __attribute__((ext_vector_type(8)));

struct GemmArgs {
    int64_t M, N, K, kchunk;
    int nsplit;
    const void* A;
    int64_t lda, strideA, planeA;
    const void* B;
    int64_t ldb, strideB, planeB;
    double alpha, beta;
    double* C;
    int64_t ldc, strideC;
    double* P;   // split-K partials [split][batch][M*N], or null
    int batch;
    int transA, transB;
    int off32;   // every per-thread tile offset of A and B fits 32 bits (FP64/FP32 interior fast path)
    int raster;  // tile order: 0 = rows of tiles; G > 0 = groups of G tile rows (tile_of)
    CMap cmap;   // placement of the output in C (not used for split-K partials)
};

// C offsets of one lane's output columns n (FN blocks x 4 registers), as
// column-major C (n * ldc) or through the output map; computed once per tile.
template <int FN, class ColOf>
__device__ __forceinline__ void column_offsets(const GemmArgs& g, bool mapped, int64_t ldc, ColOf col,
                                               int64_t (&co)[FN][4]) {
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int64_t n = col(j, r);
            co[j][r] = mapped ? (n < g.N ? cmap_digits(g.cmap.nn, g.cmap.nl, g.cmap.ns, n) : 0) : n * ldc;
        }
}

// tile t -> (tile row, tile col).  G == 0: rows of tiles, n fastest.  G > 0:
// groups of G tile rows, m fastest within a group, so a run of consecutive
// tiles (one XCD's co-resident blocks after remap_tile) covers a G x (run/G)
// square of tiles and shares G row panels of A and run/G column panels of B.
__device__ __forceinline__ void tile_of(int64_t t, int64_t tilesM, int64_t tilesN, int G, int64_t& tm, int64_t& tn) {
    if (G <= 0) {
        tm = t / tilesN;
        tn = t % tilesN;
        return;
    }
    const int64_t per = int64_t(G) * tilesN;
    const int64_t first = (t / per) * G;
    const int64_t gsz = min(int64_t(G), tilesM - first);
    const int64_t r = t % per;
    tm = first + r % gsz;
    tn = r / gsz;
}

__device__ __forceinline__ int64_t remap_tile(int64_t bid, int64_t nb) {
    const int64_t full = (nb / 8) * 8;
    if (bid >= full) return bid;
    return (bid % 8) * (nb / 8) + bid / 8;
}

// Block (bx, bz) of a launch with grid (tiles, 1, slices) -> (tile, slice).
// The dispatcher deals blocks to the XCDs round-robin in linear order (x
// fastest, then z); remapping the linear id over all tiles x slices gives each
// XCD a contiguous run of tiles of one batch/split slice, whose operand
// panels it then shares in its L2 (remapping x alone scatters a slice over
// every XCD when the tile count is not a multiple of 8).
__device__ __forceinline__ void remap_block(int64_t tiles, int64_t& t, int& bz) {
    const int64_t r = remap_tile(t + tiles * bz, tiles * int64_t(gridDim.z));
    t = r % tiles;
    bz = int(r / tiles);
}

// Result layout of the 16x16 MFMAs (measured on gfx950): lane l, register r
// holds D[row][l % 16] with
//   f64 16x16x4:           row = l/16 + 4*r
//   f32 16x16x4, bf16 x32: row = 4*(l/16) + r
template <typename T> struct Mfma;
template <> struct Mfma<double> {
    typedef f64x4 acc_t;
    static __device__ __forceinline__ acc_t run(double x, double y, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int drow(int lane, int r) { return lane / 16 + 4 * r; }
};
template <> struct Mfma<float> {
    typedef f32x4 acc_t;
    static __device__ __forceinline__ acc_t run(float x, float y, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(x, y, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int drow(int lane, int r) { return 4 * (lane / 16) + r; }
};

// Element e (= tid + 256 r) of a ROWS x BK operand tile: its row and k index.
// KFAST: the operand is stored with k contiguous (rows are strided).
template <int ROWS, int BK, bool KFAST>
__device__ __forceinline__ void tile_coord(int e, int& row, int& kk) {
    if (KFAST) { kk = e % BK; row = e / BK; } else { row = e % ROWS; kk = e / ROWS; }
}

// Register-staged loader of one ROWS x BK operand tile (256 threads, every
// thread ROWS*BK/256 elements).  KC: the operand is stored k-contiguous
// (element (row,k) at row*ld + k), otherwise row-contiguous (k*ld + row).
// Each thread keeps one base pointer, a fixed coordinate and a validity mask;
// out-of-range elements read an in-bounds clamped address and are zeroed when
// the registers are written to LDS, so nothing consumes a loaded value before
// the MFMAs of the current tile (a select right after the load would make the
// wave wait for the whole memory latency there).
template <typename T, int ROWS, int BK, bool KC>
struct TileLoader {
    static_assert((ROWS * BK) % 256 == 0, "tile must divide over 256 threads");
    static constexpr int NE = ROWS * BK / 256;
    static constexpr int VSTEP = 256 / (KC ? BK : ROWS);   // step of the varying coordinate per element
    const T* base;      // element (row0 + fixed/var0 ..., kbeg + ...) of element 0
    const T* safe;      // any in-bounds element
    int64_t ld;
    int fixed, var0;
    unsigned rowmask;   // KC: bit r = row of element r in range; !KC: bit 0 = the fixed row in range
    unsigned okbits;    // bit r: element r of the loaded tile is inside the operand
    T v[NE];

    __device__ __forceinline__ void init(const T* P, int64_t ld_, int64_t row0, int64_t nrows, int64_t kbeg, int tid) {
        ld = ld_;
        safe = P;
        if (KC) {
            fixed = tid % BK;        // k within the tile
            var0 = tid / BK;         // row of element 0
            base = P + (row0 + var0) * ld + kbeg + fixed;
            rowmask = 0;
#pragma unroll
            for (int r = 0; r < NE; ++r)
                if (row0 + var0 + r * VSTEP < nrows) rowmask |= 1u << r;
        } else {
            fixed = tid % ROWS;      // row within the tile
            var0 = tid / ROWS;       // k of element 0
            base = P + (kbeg + var0) * ld + row0 + fixed;
            rowmask = (row0 + fixed < nrows) ? 1u : 0u;
        }
    }
    // load k-tile kt (k0 = kbeg + kt*BK) into registers
    __device__ __forceinline__ void load(int64_t kt, int64_t kbeg, int64_t kend) {
        const int64_t k0 = kbeg + kt * BK;
        unsigned bits = 0;
#pragma unroll
        for (int r = 0; r < NE; ++r) {
            bool ok;
            const T* p;
            if (KC) {
                ok = ((rowmask >> r) & 1u) && (k0 + fixed < kend);
                p = base + kt * BK + int64_t(r * VSTEP) * ld;
            } else {
                ok = rowmask && (k0 + var0 + r * VSTEP < kend);
                p = base + (kt * BK + r * VSTEP) * ld;
            }
            v[r] = *(ok ? p : safe);
            bits |= unsigned(ok) << r;
        }
        okbits = bits;
    }
    // registers -> LDS tile S[k][row], out-of-range elements as zeros
    template <int LD>
    __device__ __forceinline__ void store(T (*S)[LD]) const {
#pragma unroll
        for (int r = 0; r < NE; ++r) {
            const T x = ((okbits >> r) & 1u) ? v[r] : T(0);
            if (KC) S[fixed][var0 + r * VSTEP] = x;
            else S[var0 + r * VSTEP][fixed] = x;
        }
    }
};

// Loader of complete k-tiles (same element assignment as TileLoader).
// Element r of k-tile kt is at org + kt*kstep (the buffer resource, uniform)
// + r*rstep (scalar offset) + this thread's byte offset boff (vector offset):
// a load needs no address arithmetic of its own and no mask.  Every offset is
// below 2^31 bytes (GemmArgs::off32).  The caller owns the staging registers,
// so several k-tiles can be in flight.
// EDGE: the tile's rows may run past the operand (nrows).  Such rows read the
// operand's last row instead (every address stays inside the operand); they
// only feed accumulators of output rows or columns that are never stored, so
// the values do not matter.  Row-contiguous operands clamp the thread's one
// row; k-contiguous ones keep one clamped vector offset per element.
template <typename T, int ROWS, int BK, bool KC, bool EDGE = false>
struct FastLoader {
    static constexpr int NE = ROWS * BK / 256;
    static constexpr int VSTEP = 256 / (KC ? BK : ROWS);
    static constexpr int NOFF = (EDGE && KC) ? NE : 1;
    const T* org;
    int64_t kstep, rstep;
    uint32_t boff[NOFF];
    int fixed, var0;

    __device__ __forceinline__ void init(const T* P, int64_t ld, int64_t row0, int64_t nrows, int64_t kbeg, int tid) {
        fixed = KC ? tid % BK : tid % ROWS;
        var0 = KC ? tid / BK : tid / ROWS;
        org = KC ? P + row0 * ld + kbeg : P + kbeg * ld + row0;
        kstep = KC ? BK : BK * ld;
        rstep = int64_t(VSTEP) * ld;
        const int last = int(min<int64_t>(nrows - row0, ROWS)) - 1;   // last row of the tile inside the operand
        if (!EDGE) {
            boff[0] = uint32_t((int64_t(var0) * ld + fixed) * int64_t(sizeof(T)));
        } else if (!KC) {
            boff[0] = uint32_t((int64_t(var0) * ld + min(fixed, last)) * int64_t(sizeof(T)));
        } else {
#pragma unroll
            for (int r = 0; r < NOFF; ++r)
                boff[r] = uint32_t((int64_t(min(var0 + r * VSTEP, last)) * ld + fixed) * int64_t(sizeof(T)));
        }
    }
    __device__ __forceinline__ void load(int64_t kt, T (&v)[NE]) const {
        const __amdgpu_buffer_rsrc_t rs =
            __builtin_amdgcn_make_buffer_rsrc((void*)(org + kt * kstep), (short)0, 0x7fffffff, 0x00020000);
#pragma unroll
        for (int r = 0; r < NE; ++r) {
            // uniform by construction; readfirstlane keeps it in a scalar register
            const int so = (EDGE && KC) ? 0 : __builtin_amdgcn_readfirstlane(int(int64_t(r) * rstep * int64_t(sizeof(T))));
            const uint32_t vo = boff[(EDGE && KC) ? r : 0];
            if constexpr (sizeof(T) == 8)
                v[r] = __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(rs, vo, so, 0));
            else
                v[r] = __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(rs, vo, so, 0));
        }
    }
    template <int LD>
    __device__ __forceinline__ void store(T (*S)[LD], const T (&v)[NE]) const {
#pragma unroll
        for (int r = 0; r < NE; ++r) {
            if (KC) S[fixed][var0 + r * VSTEP] = v[r];
            else S[var0 + r * VSTEP][fixed] = v[r];
        }
    }
};

// One BM x BN output tile of batch/split slice bz: block bx of a launch of
// this GEMM alone (REMAP: XCD-aware order), or tile bx (tile_of order) of a
// grouped launch.
template <typename T, int BM, int BN, int WM, int WN, bool TA, bool TB, int BK, bool REMAP>
__device__ __forceinline__ void gemm_tile(const GemmArgs g, int64_t bx, int bz) {
    constexpr int PAD = 1;
    constexpr int WTM = BM / WM, WTN = BN / WN;   // wave tile
    constexpr int FM = WTM / 16, FN = WTN / 16;   // 16x16 blocks per wave
    static_assert(WM * WN == 4, "4 waves");
    typedef typename Mfma<T>::acc_t acc_t;

    __shared__ T As[2][BK][BM + PAD];
    __shared__ T Bs[2][BK][BN + PAD];

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave % WM, wn = wave / WM;

    const int64_t tilesN = (g.N + BN - 1) / BN;
    const int64_t tilesM = (g.M + BM - 1) / BM;
    int64_t t = bx;
    if (REMAP) remap_block(tilesM * tilesN, t, bz);
    int64_t tmi, tni;
    tile_of(t, tilesM, tilesN, g.raster, tmi, tni);
    const int64_t m0 = tmi * BM, n0 = tni * BN;
    const int b = bz / g.nsplit, split = bz % g.nsplit;
    const int64_t kbeg = int64_t(split) * g.kchunk;
    const int64_t kend = min(g.K, kbeg + g.kchunk);

    acc_t acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

    // Wave-uniform counts of this wave's 16-row / 16-column blocks that hold
    // any output inside C: on an edge tile the MFMAs of the other blocks are
    // skipped (their accumulators are never stored).
    const int ilim = __builtin_amdgcn_readfirstlane(
        int(max<int64_t>(0, min<int64_t>(FM, (g.M - (m0 + wm * WTM) + 15) / 16))));
    const int jlim = __builtin_amdgcn_readfirstlane(
        int(max<int64_t>(0, min<int64_t>(FN, (g.N - (n0 + wn * WTN) + 15) / 16))));

    // one k-tile of MFMAs from LDS buffer buf (LIM: skip blocks outside C)
    auto compute = [&](int buf, auto LIM) {
        constexpr bool lim = decltype(LIM)::value;
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            const int kr = kk + lane / 16;
            T x[FN], y[FM];
#pragma unroll
            for (int j = 0; j < FN; ++j) x[j] = Bs[buf][kr][wn * WTN + j * 16 + lane % 16];
#pragma unroll
            for (int i = 0; i < FM; ++i) y[i] = As[buf][kr][wm * WTM + i * 16 + lane % 16];
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j)
                    if (!lim || (i < ilim && j < jlim)) acc[i][j] = Mfma<T>::run(x[j], y[i], acc[i][j]);
        }
    };

    // opA(m,k) = A[k + m*lda] when TA (k contiguous); opB(k,n) = B[k + n*ldb] when !TB (k contiguous)
    const T* Ab = static_cast<const T*>(g.A) + int64_t(b) * g.strideA;
    const T* Bb = static_cast<const T*>(g.B) + int64_t(b) * g.strideB;
    const int64_t nk = kend > kbeg ? (kend - kbeg + BK - 1) / BK : 0;
    const int64_t nfull = kend > kbeg ? (kend - kbeg) / BK : 0;
    // Complete k-tiles through FastLoader, two k-tiles in flight (the tile
    // stored now was loaded two steps ago), so a load has two tiles of MFMA
    // work to arrive.  Register set s and LDS buffer s serve the k-tiles of
    // parity s.  EDGE: the tile runs past M or N (clamped rows, skipped blocks).
    auto fast = [&](auto EDGE) {
        constexpr bool edge = decltype(EDGE)::value;
        FastLoader<T, BM, BK, TA, edge> fa;
        FastLoader<T, BN, BK, !TB, edge> fb;
        fa.init(Ab, g.lda, m0, g.M, kbeg, tid);
        fb.init(Bb, g.ldb, n0, g.N, kbeg, tid);
        T va[2][FastLoader<T, BM, BK, TA>::NE], vb[2][FastLoader<T, BN, BK, !TB>::NE];
        if (nfull > 0) { fa.load(0, va[0]); fb.load(0, vb[0]); }
        if (nfull > 1) { fa.load(1, va[1]); fb.load(1, vb[1]); }
        auto step = [&](auto S, int64_t kt) {
            constexpr int s = decltype(S)::value;
            fa.store(As[s], va[s]);
            fb.store(Bs[s], vb[s]);
            __syncthreads();
            if (kt + 2 < nfull) { fa.load(kt + 2, va[s]); fb.load(kt + 2, vb[s]); }
            __builtin_amdgcn_sched_barrier(0);
            compute(s, EDGE);
        };
        for (int64_t kt = 0; kt < nfull; kt += 2) {
            step(std::integral_constant<int, 0>{}, kt);
            if (kt + 1 < nfull) step(std::integral_constant<int, 1>{}, kt + 1);
        }
        if (nk > nfull) {   // the partial last k-tile
            TileLoader<T, BM, BK, TA> la;
            TileLoader<T, BN, BK, !TB> lb;
            la.init(Ab, g.lda, m0, g.M, kbeg, tid);
            lb.init(Bb, g.ldb, n0, g.N, kbeg, tid);
            la.load(nfull, kbeg, kend);
            lb.load(nfull, kbeg, kend);
            const int buf = int(nfull & 1);
            la.store(As[buf]);
            lb.store(Bs[buf]);
            __syncthreads();
            compute(buf, EDGE);
        }
    };
    if (g.off32 && m0 + BM <= g.M && n0 + BN <= g.N) {
        fast(std::false_type{});
    } else if (g.off32) {
        fast(std::true_type{});
    } else {
        // offsets beyond 32 bits: predicated loads, one k-tile in flight
        TileLoader<T, BM, BK, TA> la;
        TileLoader<T, BN, BK, !TB> lb;
        la.init(Ab, g.lda, m0, g.M, kbeg, tid);
        lb.init(Bb, g.ldb, n0, g.N, kbeg, tid);
        if (nk > 0) { la.load(0, kbeg, kend); lb.load(0, kbeg, kend); }
        for (int64_t kt = 0; kt < nk; ++kt) {
            const int buf = int(kt & 1);
            la.store(As[buf]);
            lb.store(Bs[buf]);
            __syncthreads();
            // next tile's global loads go out before this tile's MFMAs, so their
            // latency hides behind the matrix work
            if (kt + 1 < nk) { la.load(kt + 1, kbeg, kend); lb.load(kt + 1, kbeg, kend); }
            __builtin_amdgcn_sched_barrier(0);
            compute(buf, std::true_type{});
        }
    }

    // epilogue: lane holds C[m = mb + lane%16][n = nb + drow(lane, r)]
    double* Cb = g.P ? g.P + (int64_t(split) * g.batch + b) * (g.M * g.N) : g.C + int64_t(b) * g.strideC;
    const int64_t ldc = g.P ? g.M : g.ldc;
    const bool mapped = !g.P && g.cmap.mapped;
    int64_t co[FN][4];
    column_offsets<FN>(g, mapped, ldc,
                       [&](int j, int r) { return n0 + wn * WTN + j * 16 + Mfma<T>::drow(lane, r); }, co);
    // per row of blocks, every C value is read before any is written: one
    // memory round trip per row instead of one per element (a load after a
    // store to C could alias it, so the compiler keeps that order)
    const bool addc = !g.P && g.beta != 0.0;
#pragma unroll
    for (int i = 0; i < FM; ++i) {
        const int64_t m = m0 + wm * WTM + i * 16 + lane % 16;
        const int64_t ro = mapped ? (m < g.M ? cmap_digits(g.cmap.nm, g.cmap.ml, g.cmap.ms, m) : 0) : m;
        double cv[FN][4];
        if (addc) {
#pragma unroll
            for (int j = 0; j < FN; ++j)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int64_t n = n0 + wn * WTN + j * 16 + Mfma<T>::drow(lane, r);
                    cv[j][r] = (m < g.M && n < g.N) ? Cb[ro + co[j][r]] : 0.0;
                }
        }
#pragma unroll
        for (int j = 0; j < FN; ++j) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t n = n0 + wn * WTN + j * 16 + Mfma<T>::drow(lane, r);
                if (m < g.M && n < g.N) {
                    double v = g.alpha * double(acc[i][j][r]);
                    if (addc) v += g.beta * cv[j][r];
                    Cb[ro + co[j][r]] = v;
                }
            }
        }
    }
}

template <typename T, int BM, int BN, int WM, int WN, bool TA, bool TB, int BK, int MINW>
__global__ __launch_bounds__(256, MINW) void k_gemm(GemmArgs g) {
    gemm_tile<T, BM, BN, WM, WN, TA, TB, BK, true>(g, blockIdx.x, blockIdx.z);
}

// ------------------------------------------------------------------ BF16x3

// NT = 64*WM*WN threads (4 waves: one per SIMD; 8 waves: two per SIMD, each
// holding half the accumulators, so one wave's LDS and barrier waits are
// covered by the other's MFMAs).  DIRECT: interior tiles load their operand
// tiles straight from global memory into LDS (global_load_lds_dwordx4), with
// no staging registers and no LDS store instructions; edge tiles, and every
// tile of a DIRECT = false instantiation, stage through registers.
template <int BM, int BN, int WM, int WN, bool REMAP, bool DIRECT>
__device__ __forceinline__ void gemm_tile_bf16x3(const GemmArgs g, int64_t bx, int bz) {
    constexpr int BK = 32;
    constexpr int NT = 64 * WM * WN;
    constexpr int NW = WM * WN;
    // LDS rows of BK bf16 (64 bytes, four 16-byte chunks), chunk c of row r
    // stored at chunk c ^ ((r >> 1) & 3).  Measured on gfx950 (bank-conflict
    // counter): this layout serves both the MFMA operand reads (lane -> row
    // lane%16, chunk lane/16) and the tile stores (lane -> row lane/4, chunk
    // lane%4) without bank conflicts; an 80-byte padded row costs 4 conflict
    // cycles per ds_read_b128 and 8 per ds_write_b128.
    constexpr int KP = BK;
    auto swz = [](int row, int k) { return (((k >> 3) ^ (row >> 1)) & 3) << 3; };
    constexpr int WTM = BM / WM, WTN = BN / WN;
    constexpr int FM = WTM / 16, FN = WTN / 16;
    constexpr int VA = BM * BK / 8, VB = BN * BK / 8;   // 8-element vectors per plane
    constexpr int NA = (VA + NT - 1) / NT, NB = (VB + NT - 1) / NT;
    static_assert(WM * WN == 4 || WM * WN == 8, "4 or 8 waves");
    static_assert(!DIRECT || ((BM / 16) % NW == 0 && (BN / 16) % NW == 0), "direct loads: whole 16-row groups per wave");

    // one LDS object per buffer, so that the compiler can tell the buffer a
    // direct load fills from the one the MFMAs read
    __shared__ i16x8 dummy_align;                   // keeps LDS 16-byte aligned
    __shared__ __attribute__((aligned(16))) unsigned short As0[2][BM][KP];   // [hi/lo][m][k]
    __shared__ __attribute__((aligned(16))) unsigned short As1[2][BM][KP];
    __shared__ __attribute__((aligned(16))) unsigned short Bs0[2][BN][KP];
    __shared__ __attribute__((aligned(16))) unsigned short Bs1[2][BN][KP];
    (void)dummy_align;
    auto As = [&](int buf) -> unsigned short (*)[BM][KP] { return buf ? As1 : As0; };
    auto Bs = [&](int buf) -> unsigned short (*)[BN][KP] { return buf ? Bs1 : Bs0; };

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave % WM, wn = wave / WM;

    const int64_t tilesN = (g.N + BN - 1) / BN;
    const int64_t tilesM = (g.M + BM - 1) / BM;
    int64_t t = bx;
    if (REMAP) remap_block(tilesM * tilesN, t, bz);
    int64_t tmi, tni;
    tile_of(t, tilesM, tilesN, g.raster, tmi, tni);
    const int64_t m0 = tmi * BM, n0 = tni * BN;
    const int b = bz / g.nsplit, split = bz % g.nsplit;
    const int64_t kbeg = int64_t(split) * g.kchunk;
    const int64_t kend = min(g.K, kbeg + g.kchunk);

    // operands: A[k + m*lda] (K-contiguous), B[k + n*ldb] (K-contiguous)
    const unsigned short* A = static_cast<const unsigned short*>(g.A) + int64_t(b) * g.strideA;
    const unsigned short* B = static_cast<const unsigned short*>(g.B) + int64_t(b) * g.strideB;

    // predicated loads: elements outside the operand are zeros
    auto load = [&](int64_t k0, i16x8 (&xa)[2][NA], i16x8 (&xb)[2][NB]) {
        const i16x8 z = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int r = 0; r < NA; ++r) {
            const int e = tid + NT * r;
            const int kk = (e % (BK / 8)) * 8, mm = e / (BK / 8);
            const int64_t m = m0 + mm, k = k0 + kk;
            i16x8 h = z, l = z;
            if (e < VA && m < g.M && k < kend) {
                h = *reinterpret_cast<const i16x8*>(A + k + m * g.lda);
                l = *reinterpret_cast<const i16x8*>(A + g.planeA + k + m * g.lda);
            }
            xa[0][r] = h;
            xa[1][r] = l;
        }
#pragma unroll
        for (int r = 0; r < NB; ++r) {
            const int e = tid + NT * r;
            const int kk = (e % (BK / 8)) * 8, nn = e / (BK / 8);
            const int64_t n = n0 + nn, k = k0 + kk;
            i16x8 h = z, l = z;
            if (e < VB && n < g.N && k < kend) {
                h = *reinterpret_cast<const i16x8*>(B + k + n * g.ldb);
                l = *reinterpret_cast<const i16x8*>(B + g.planeB + k + n * g.ldb);
            }
            xb[0][r] = h;
            xb[1][r] = l;
        }
    };
    auto store = [&](int buf, const i16x8 (&xa)[2][NA], const i16x8 (&xb)[2][NB]) {
        auto* as = As(buf);
        auto* bs = Bs(buf);
#pragma unroll
        for (int r = 0; r < NA; ++r) {
            const int e = tid + NT * r;
            if (e >= VA) continue;
            const int kk = (e % (BK / 8)) * 8, mm = e / (BK / 8);
            *reinterpret_cast<i16x8*>(&as[0][mm][swz(mm, kk)]) = xa[0][r];
            *reinterpret_cast<i16x8*>(&as[1][mm][swz(mm, kk)]) = xa[1][r];
        }
#pragma unroll
        for (int r = 0; r < NB; ++r) {
            const int e = tid + NT * r;
            if (e >= VB) continue;
            const int kk = (e % (BK / 8)) * 8, nn = e / (BK / 8);
            *reinterpret_cast<i16x8*>(&bs[0][nn][swz(nn, kk)]) = xb[0][r];
            *reinterpret_cast<i16x8*>(&bs[1][nn][swz(nn, kk)]) = xb[1][r];
        }
    };

    f32x4 acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    // MFMA operands: src0 (16 x 32) from the B tile, lane holds row lane%16,
    // k = 8*(lane/16) .. +7; src1 (32 x 16) from the A tile, lane holds column
    // lane%16, the same k.  One k-tile of MFMAs: the B fragments first, then
    // the A fragments of row i+1 are read while row i's MFMAs run (two
    // register pairs), so the MFMAs never wait for an LDS read they need next;
    // `mid` runs halfway.  The MFMA sequence of every accumulator (k-tiles in
    // order, three products per block in the order lo*hi, hi*lo, hi*hi) is the
    // same on every path and in every variant.
    const int kq = 8 * (lane / 16);
    auto mfma_tile = [&](const unsigned short (*as)[BM][KP], const unsigned short (*bs)[BN][KP], auto&& mid,
                         auto&& late) {
        bf16x8 xh[FN], xl[FN], yh[2], yl[2];
        auto read_b = [&](int j) {
            const int nn = wn * WTN + j * 16 + lane % 16;
            xh[j] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const i16x8*>(&bs[0][nn][swz(nn, kq)]));
            xl[j] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const i16x8*>(&bs[1][nn][swz(nn, kq)]));
        };
        auto read_a = [&](int i, int slot) {
            const int mm = wm * WTM + i * 16 + lane % 16;
            yh[slot] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const i16x8*>(&as[0][mm][swz(mm, kq)]));
            yl[slot] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const i16x8*>(&as[1][mm][swz(mm, kq)]));
        };
        // the first row of blocks reads each B fragment pair just ahead of its
        // MFMAs, so a k-tile starts after four reads rather than 2*FN + 2 (all
        // waves read at once after the barrier; measured +1-2% on 13824^3)
        read_b(0);
        read_a(0, 0);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int i = 0; i < FM; ++i) {
            if (i + 1 < FM && i > 0) read_a(i + 1, (i + 1) & 1);
            __builtin_amdgcn_sched_barrier(0);   // keep those reads ahead of row i's MFMAs
            if (i == FM / 2) mid();
            // every fragment of the last row is in registers by now: `late`
            // may let other waves overwrite this tile's buffers
            if (i == FM - 1 && FM > 1) late();
            if (i == 0) {
#pragma unroll
                for (int j = 0; j < FN; ++j) {
                    if (j + 1 < FN) read_b(j + 1);
                    if (j + 1 == FN && FM > 1) read_a(1, 1);
                    __builtin_amdgcn_sched_barrier(0);
                    acc[0][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xl[j], yh[0], acc[0][j], 0, 0, 0);
                    acc[0][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xh[j], yl[0], acc[0][j], 0, 0, 0);
                    acc[0][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xh[j], yh[0], acc[0][j], 0, 0, 0);
                }
                continue;
            }
#pragma unroll
            for (int j = 0; j < FN; ++j) {
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xl[j], yh[i & 1], acc[i][j], 0, 0, 0);
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xh[j], yl[i & 1], acc[i][j], 0, 0, 0);
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xh[j], yh[i & 1], acc[i][j], 0, 0, 0);
            }
        }
        if (FM == 1) late();
    };

    const int64_t nk = kend > kbeg ? (kend - kbeg + BK - 1) / BK : 0;
    if (DIRECT && m0 + BM <= g.M && n0 + BN <= g.N) {
        // Each wave instruction moves 64 lanes x 16 bytes into 1024 contiguous
        // LDS bytes: 16 rows of one plane.  Lane l fills row l/4 of the group
        // at physical chunk l%4, which holds the logical k-chunk
        // (l%4) ^ ((row >> 1) & 3) of the swizzled layout.  A chunk starting
        // at or beyond kend reads an in-bounds address and is zeroed in LDS
        // once its load has landed (only in the partial last k-tile).
        constexpr int GA = BM / 16 / NW, GB = BN / 16 / NW;   // row groups per wave and plane
        const int rl = lane >> 2, pc = lane & 3;
        auto issue = [&](auto S, int64_t k0) {
            constexpr int s = decltype(S)::value;
            auto* as = s ? As1 : As0;
            auto* bs = s ? Bs1 : Bs0;
#pragma unroll
            for (int q = 0; q < GA; ++q) {
                const int row = (wave + NW * q) * 16 + rl;
                const int64_t k = k0 + 8 * (pc ^ ((row >> 1) & 3));
                const unsigned short* src = A + (k < kend ? k : kbeg) + (m0 + row) * g.lda;
                __builtin_amdgcn_global_load_lds(src, (__attribute__((address_space(3))) void*)&as[0][(wave + NW * q) * 16][0], 16, 0, 0);
                __builtin_amdgcn_global_load_lds(src + g.planeA, (__attribute__((address_space(3))) void*)&as[1][(wave + NW * q) * 16][0], 16, 0, 0);
            }
#pragma unroll
            for (int q = 0; q < GB; ++q) {
                const int row = (wave + NW * q) * 16 + rl;
                const int64_t k = k0 + 8 * (pc ^ ((row >> 1) & 3));
                const unsigned short* src = B + (k < kend ? k : kbeg) + (n0 + row) * g.ldb;
                __builtin_amdgcn_global_load_lds(src, (__attribute__((address_space(3))) void*)&bs[0][(wave + NW * q) * 16][0], 16, 0, 0);
                __builtin_amdgcn_global_load_lds(src + g.planeB, (__attribute__((address_space(3))) void*)&bs[1][(wave + NW * q) * 16][0], 16, 0, 0);
            }
        };
        // wait for this wave's direct loads; in a partial k-tile, zero the
        // chunks past kend (each lane its own chunk, after its load landed)
        auto land = [&](auto S, int64_t k0) {
            constexpr int s = decltype(S)::value;
            __builtin_amdgcn_s_waitcnt(0);
            if (k0 + BK > kend) {
                auto* as = s ? As1 : As0;
                auto* bs = s ? Bs1 : Bs0;
                const i16x8 z = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
                for (int q = 0; q < GA; ++q) {
                    const int row = (wave + NW * q) * 16 + rl;
                    if (k0 + 8 * (pc ^ ((row >> 1) & 3)) >= kend) {
                        *reinterpret_cast<i16x8*>(&as[0][row][8 * pc]) = z;
                        *reinterpret_cast<i16x8*>(&as[1][row][8 * pc]) = z;
                    }
                }
#pragma unroll
                for (int q = 0; q < GB; ++q) {
                    const int row = (wave + NW * q) * 16 + rl;
                    if (k0 + 8 * (pc ^ ((row >> 1) & 3)) >= kend) {
                        *reinterpret_cast<i16x8*>(&bs[0][row][8 * pc]) = z;
                        *reinterpret_cast<i16x8*>(&bs[1][row][8 * pc]) = z;
                    }
                }
            }
        };
        // The barrier that publishes tile kt+1 comes before the last row of
        // tile kt's MFMAs (whose fragments are already in registers), so the
        // matrix core has that row's work across the barrier.
        auto step = [&](auto S, int64_t kt) {
            constexpr int s = decltype(S)::value;
            const int64_t k1 = kbeg + (kt + 1) * BK;
            if (kt + 1 < nk) issue(std::integral_constant<int, 1 - s>{}, k1);
            mfma_tile(s ? As1 : As0, s ? Bs1 : Bs0, [] {}, [&] {
                if (kt + 1 < nk) land(std::integral_constant<int, 1 - s>{}, k1);
                __syncthreads();
            });
        };
        if (nk > 0) {
            issue(std::integral_constant<int, 0>{}, kbeg);
            land(std::integral_constant<int, 0>{}, kbeg);
            __syncthreads();
        }
        for (int64_t kt = 0; kt < nk; kt += 2) {
            step(std::integral_constant<int, 0>{}, kt);
            if (kt + 1 < nk) step(std::integral_constant<int, 1>{}, kt + 1);
        }
    } else {
        // One barrier per k-tile.  Tile kt+1 (in registers) is stored to the
        // other LDS buffer halfway through tile kt's MFMAs, and the loads of
        // tile kt+2 are issued right after that store, so a global load has a
        // whole k-tile of MFMA work to arrive.
        i16x8 ra0[2][NA], rb0[2][NB];
        if (nk > 0) {
            load(kbeg, ra0, rb0);
            store(0, ra0, rb0);
            if (nk > 1) load(kbeg + BK, ra0, rb0);
            __syncthreads();
        }
        for (int64_t kt = 0; kt < nk; ++kt) {
            const int buf = int(kt & 1);
            mfma_tile(As(buf), Bs(buf), [&] {
                if (kt + 1 < nk) {
                    store(buf ^ 1, ra0, rb0);
                    if (kt + 2 < nk) load(kbeg + (kt + 2) * BK, ra0, rb0);
                }
            }, [] {});
            __syncthreads();
        }
    }

    double* Cb = g.P ? g.P + (int64_t(split) * g.batch + b) * (g.M * g.N) : g.C + int64_t(b) * g.strideC;
    const int64_t ldc = g.P ? g.M : g.ldc;
    const bool mapped = !g.P && g.cmap.mapped;
    int64_t co[FN][4];
    column_offsets<FN>(g, mapped, ldc, [&](int j, int r) { return n0 + wn * WTN + j * 16 + 4 * (lane / 16) + r; },
                       co);
    // per row of blocks, every C value is read before any is written (one
    // memory round trip per row instead of one per element)
    const bool addc = !g.P && g.beta != 0.0;
#pragma unroll
    for (int i = 0; i < FM; ++i) {
        const int64_t m = m0 + wm * WTM + i * 16 + lane % 16;
        const int64_t ro = mapped ? (m < g.M ? cmap_digits(g.cmap.nm, g.cmap.ml, g.cmap.ms, m) : 0) : m;
        double cv[FN][4];
        if (addc) {
#pragma unroll
            for (int j = 0; j < FN; ++j)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int64_t n = n0 + wn * WTN + j * 16 + 4 * (lane / 16) + r;
                    cv[j][r] = (m < g.M && n < g.N) ? Cb[ro + co[j][r]] : 0.0;
                }
        }
#pragma unroll
        for (int j = 0; j < FN; ++j) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t n = n0 + wn * WTN + j * 16 + 4 * (lane / 16) + r;
                if (m < g.M && n < g.N) {
                    double v = g.alpha * double(acc[i][j][r]);
                    if (addc) v += g.beta * cv[j][r];
                    Cb[ro + co[j][r]] = v;
                }
            }
        }
    }
}

template <int BM, int BN, int WM, int WN, bool DIRECT>
__global__ __launch_bounds__(64 * WM * WN, 1) void k_gemm_bf16x3(GemmArgs g) {
    gemm_tile_bf16x3<BM, BN, WM, WN, true, DIRECT>(g, blockIdx.x, blockIdx.z);
}

// ---------------------------------------------------------- grouped launch
//
// One launch over the output tiles of several independent GEMMs (problems
// from different contexts with the same format, operand layouts and block
// shape).  Problem p owns blocks [first[p], first[p+1]) of the 1-D grid,
// ordered as its own (tile, batch x split) grid; every block runs the same
// tile body as the single-problem kernels, so each element is computed with
// the same operations in the same order as in a launch of its GEMM alone.
constexpr int kMaxGroup = 8;
struct GemmGroup {
    GemmArgs p[kMaxGroup];
    int64_t first[kMaxGroup + 1];
    int n;
};

__device__ __forceinline__ int group_problem(const GemmGroup& G, int64_t bid) {
    int p = 0;
    while (p + 1 < G.n && bid >= G.first[p + 1]) ++p;
    return p;
}

// Persistent: the grid holds at most as many blocks as the device keeps
// resident, and each block walks the group's tiles with a stride of the grid
// size, so one launch of a fixed grid covers every problem of the group.
template <typename T, int BM, int BN, int WM, int WN, bool TA, bool TB, int BK, int MINW>
__global__ __launch_bounds__(256, MINW) void k_gemm_grouped(GemmGroup G) {
    const int64_t total = G.first[G.n];
    for (int64_t bid = blockIdx.x; bid < total; bid += gridDim.x) {
        const int p = group_problem(G, bid);
        const int64_t tiles = ((G.p[p].M + BM - 1) / BM) * ((G.p[p].N + BN - 1) / BN);
        const int64_t local = bid - G.first[p];
        gemm_tile<T, BM, BN, WM, WN, TA, TB, BK, false>(G.p[p], local % tiles, int(local / tiles));
        __syncthreads();   // the next tile reuses the LDS buffers
    }
}

template <int BM, int BN, int WM, int WN>
__global__ __launch_bounds__(256, 1) void k_gemm_bf16x3_grouped(GemmGroup G) {
    const int64_t total = G.first[G.n];
    for (int64_t bid = blockIdx.x; bid < total; bid += gridDim.x) {
        const int p = group_problem(G, bid);
        const int64_t tiles = ((G.p[p].M + BM - 1) / BM) * ((G.p[p].N + BN - 1) / BN);
        const int64_t local = bid - G.first[p];
        gemm_tile_bf16x3<BM, BN, WM, WN, false, false>(G.p[p], local % tiles, int(local / tiles));
        __syncthreads();   // the next tile reuses the LDS buffers
    }
}

// ---------------------------------------------------------------- dispatch

struct TileCfg {
    int BM, BN;
};

template <typename T, int BM, int BN, int WM, int WN>
static void run_cfg(aq_ctx* ctx, const GemmArgs& a, dim3 grid, int variant) {
#define AQ_LAUNCH(BK_, MW_)                                                                                       \
    do {                                                                                                          \
        if (a.transA && a.transB)                                                                                 \
            hipLaunchKernelGGL((k_gemm<T, BM, BN, WM, WN, true, true, BK_, MW_>), grid, dim3(256), 0, ctx->stream, a);   \
        else if (a.transA)                                                                                        \
            hipLaunchKernelGGL((k_gemm<T, BM, BN, WM, WN, true, false, BK_, MW_>), grid, dim3(256), 0, ctx->stream, a);  \
        else if (a.transB)                                                                                        \
            hipLaunchKernelGGL((k_gemm<T, BM, BN, WM, WN, false, true, BK_, MW_>), grid, dim3(256), 0, ctx->stream, a);  \
        else                                                                                                      \
            hipLaunchKernelGGL((k_gemm<T, BM, BN, WM, WN, false, false, BK_, MW_>), grid, dim3(256), 0, ctx->stream, a); \
    } while (0)
    // variants of the large tile: 0 = BK16 one wave/SIMD, 1 = BK32, 2 = BK16 register-lean (two waves/SIMD)
    // the one-short-side tiles (128 x 64, 64 x 128) always run two waves per SIMD
    if (BM == 128 && BN == 128 && variant == 1) AQ_LAUNCH(32, 1);
    else if (BM * BN >= 8192 && variant == 2) AQ_LAUNCH(16, 2);
    else AQ_LAUNCH(16, 1);
#undef AQ_LAUNCH
}

// Measured on MI355X (square 8192^3 / ladder / ring shapes): the register-lean
// BK-16 variant with two waves per SIMD hides the global-load latency that a
// single wave per SIMD exposes (FP64 35 -> 56 TF/s; FP32 unchanged at ~108).
int default_gemm_variant(Fmt fmt) { return fmt == FMT_BF16X2 ? 0 : 2; }

template <int BM, int BN, int WM, int WN, bool DIRECT = false>
static void run_cfg_bf16(aq_ctx* ctx, const GemmArgs& a, dim3 grid) {
    hipLaunchKernelGGL((k_gemm_bf16x3<BM, BN, WM, WN, DIRECT>), grid, dim3(64 * WM * WN), 0, ctx->stream, a);
}

// Kernel launch for one (format, block shape, variant); the caller checked the shape.
static void dispatch(aq_ctx* ctx, Fmt fmt, int BM, int BN, int variant, const GemmArgs& a, dim3 grid) {
    if (fmt == FMT_F64) {
        if (BM == 128 && BN == 128) run_cfg<double, 128, 128, 2, 2>(ctx, a, grid, variant);
        else if (BM == 128 && BN == 64) run_cfg<double, 128, 64, 2, 2>(ctx, a, grid, 2);
        else if (BM == 64 && BN == 128) run_cfg<double, 64, 128, 2, 2>(ctx, a, grid, 2);
        else if (BM == 64) run_cfg<double, 64, 64, 2, 2>(ctx, a, grid, 0);
        else if (BN == 16) run_cfg<double, 128, 16, 4, 1>(ctx, a, grid, 0);
        else run_cfg<double, 16, 128, 1, 4>(ctx, a, grid, 0);
    } else if (fmt == FMT_F32) {
        if (BM == 128 && BN == 128) run_cfg<float, 128, 128, 2, 2>(ctx, a, grid, variant);
        else if (BM == 128 && BN == 64) run_cfg<float, 128, 64, 2, 2>(ctx, a, grid, 2);
        else if (BM == 64 && BN == 128) run_cfg<float, 64, 128, 2, 2>(ctx, a, grid, 2);
        else if (BM == 64) run_cfg<float, 64, 64, 2, 2>(ctx, a, grid, 0);
        else if (BN == 16) run_cfg<float, 128, 16, 4, 1>(ctx, a, grid, 0);
        else run_cfg<float, 16, 128, 1, 4>(ctx, a, grid, 0);
    } else {
        if (BM == 256 && BN == 256 && variant == 3) run_cfg_bf16<256, 256, 2, 4, true>(ctx, a, grid);
        else if (BM == 256 && BN == 256) run_cfg_bf16<256, 256, 2, 2>(ctx, a, grid);
        else if (BM == 128 && BN == 128) run_cfg_bf16<128, 128, 2, 2>(ctx, a, grid);
        else if (BM == 256) run_cfg_bf16<256, 128, 2, 2>(ctx, a, grid);
        else if (BN == 256) run_cfg_bf16<128, 256, 2, 2>(ctx, a, grid);
        else if (BM == 64) run_cfg_bf16<64, 64, 2, 2>(ctx, a, grid);
        else if (BN == 16) run_cfg_bf16<128, 16, 4, 1>(ctx, a, grid);
        else run_cfg_bf16<16, 128, 1, 4>(ctx, a, grid);
    }
}

// Block shape of a variant: BF16x3 variants 1 and 2 enlarge the 128 x 128
// block to 256 x 128 / 128 x 256 (128 x 64 / 64 x 128 per wave), which reads a
// third less LDS per MFMA, variant 4 to 256 x 256 with four waves (128 x 128
// per wave, the accumulators in all 256 AGPRs, one wave per SIMD) and variant
// 3 to 256 x 256 with eight waves (128 x 64 per wave, two waves per SIMD); the
// FP64/FP32 variants change BK and occupancy only.
static void variant_shape(Fmt fmt, int variant, int64_t M, int64_t N, int& BM, int& BN) {
    if (fmt == FMT_BF16X2 && BM == 128 && BN == 128) {
        if (variant == 1 && M >= 256) BM = 256;
        else if (variant == 2 && N >= 256) BN = 256;
        else if ((variant == 3 || variant == 4) && M >= 256 && N >= 256) { BM = 256; BN = 256; }
    }
}

// ------------------------------------------------------------- autotuner
//
// Per device, one table from a shape bucket (format, operand layouts, M, N, K
// and batch in half-octaves) to the variant that ran fastest when the bucket
// was first met.  A context with autotune on times every candidate variant
// once on that first GEMM (HIP events on its stream, into a scratch output),
// records the winner and reuses it for every later GEMM of the bucket, from
// any context.  Variants never change results (the split-K partition does not
// depend on them), so tuning affects speed only.
namespace {

struct TuneKey {
    int device, fmt, ta, tb, m, n, k, batch;
    bool operator<(const TuneKey& o) const {
        return std::tie(device, fmt, ta, tb, m, n, k, batch) < std::tie(o.device, o.fmt, o.ta, o.tb, o.m, o.n, o.k, o.batch);
    }
};

struct TuneEntry {
    int variant = -1;
    float ms[5] = {-1.f, -1.f, -1.f, -1.f, -1.f};
};

std::mutex g_tune_mu;
std::map<TuneKey, TuneEntry> g_tune;

int half_octave(int64_t x) { return int(std::lround(2.0 * std::log2(double(std::max<int64_t>(x, 1))))); }

}  // namespace

static int tuned_variant(aq_ctx* ctx, const GemmDesc& d, int BM, int BN, const GemmArgs& a) {
    const int defv = default_gemm_variant(d.fmt);
    if (!(BM == 128 && BN == 128)) return defv;             // only the large tile has variants
    TuneKey key{ctx->device, int(d.fmt), d.transA, d.transB, half_octave(d.M), half_octave(d.N), half_octave(d.K),
                half_octave(d.batch)};
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune.find(key);
        if (it != g_tune.end()) return it->second.variant >= 0 ? it->second.variant : defv;
    }
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(ctx->stream, &cap) != hipSuccess || cap != hipStreamCaptureStatusNone) return defv;
    // trial output (and partials) so that C is untouched
    const size_t cbytes = size_t(a.M) * size_t(a.N) * size_t(a.batch) * 8;
    const size_t pbytes = a.nsplit > 1 ? size_t(a.nsplit) * cbytes : 0;
    if (ctx_in_use(ctx) + cbytes + pbytes > ctx->budget) return defv;
    TuneEntry e;
    {
        Scratch tc(ctx, cbytes), tp(ctx, pbytes);
        GemmArgs t = a;
        t.C = tc.as<double>();
        t.ldc = a.M;
        t.strideC = a.M * a.N;
        t.beta = 0.0;
        t.P = a.nsplit > 1 ? tp.as<double>() : nullptr;
        t.cmap = CMap{};   // the trial writes its scratch output column-major
        hipEvent_t e0, e1;
        AQ_HIP(hipEventCreate(&e0));
        AQ_HIP(hipEventCreate(&e1));
        float best = 0.f;
        const int nvar = d.fmt == FMT_BF16X2 ? 5 : 3;
        for (int v = 0; v < nvar; ++v) {
            int bm = BM, bn = BN;
            variant_shape(d.fmt, v, d.M, d.N, bm, bn);
            if (d.fmt == FMT_BF16X2 && v >= 1 && bm == BM && bn == BN) continue;   // = variant 0
            const int64_t tiles = ((d.M + bm - 1) / bm) * ((d.N + bn - 1) / bn);
            dim3 grid(unsigned(tiles), 1, unsigned(a.batch * a.nsplit));
            dispatch(ctx, d.fmt, bm, bn, v, t, grid);           // warm
            AQ_HIP(hipEventRecord(e0, ctx->stream));
            dispatch(ctx, d.fmt, bm, bn, v, t, grid);
            AQ_HIP(hipEventRecord(e1, ctx->stream));
            AQ_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            AQ_HIP(hipEventElapsedTime(&ms, e0, e1));
            e.ms[v] = ms;
            if (e.variant < 0 || ms < best) {
                best = ms;
                e.variant = v;
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        AQ_HIP(hipGetLastError());
    }
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune.emplace(key, e);
    return e.variant >= 0 ? e.variant : defv;
}

// ------------------------------------------------------- grouping (host)

namespace {

struct GroupKey {
    int fmt, ta, tb, bm, bn;
    bool operator==(const GroupKey& o) const {
        return fmt == o.fmt && ta == o.ta && tb == o.tb && bm == o.bm && bn == o.bn;
    }
};

// block shapes with a grouped kernel: the 128 x 128 and 64 x 64 FP64 and
// BF16x3 tiles (the FP32 comparison path and the skinny tiles launch alone)
bool group_shape(Fmt fmt, int BM, int BN) {
    return (fmt == FMT_F64 || fmt == FMT_BF16X2) && BM == BN && (BM == 128 || BM == 64);
}

// Resident blocks of the grouped kernels: 2 per CU (FP64 128 x 128 runs two
// waves per SIMD; the BF16x3 128 x 128 and the 64 x 64 tiles fit two blocks'
// LDS per CU), times the CU count.
void launch_grouped(const GroupKey& k, const GemmGroup& G, int64_t blocks, hipStream_t s, int cus) {
    const int64_t resident = int64_t(std::max(cus, 1)) * 2;
    const dim3 grid{unsigned(std::min(blocks, resident))}, block{256};
#define AQ_GROUPED(BM_, BK_, MW_)                                                                                      \
    do {                                                                                                           \
        if (k.ta && k.tb)                                                                                          \
            hipLaunchKernelGGL((k_gemm_grouped<double, BM_, BM_, 2, 2, true, true, BK_, MW_>), grid, block, 0, s, G);   \
        else if (k.ta)                                                                                             \
            hipLaunchKernelGGL((k_gemm_grouped<double, BM_, BM_, 2, 2, true, false, BK_, MW_>), grid, block, 0, s, G);  \
        else if (k.tb)                                                                                             \
            hipLaunchKernelGGL((k_gemm_grouped<double, BM_, BM_, 2, 2, false, true, BK_, MW_>), grid, block, 0, s, G);  \
        else                                                                                                       \
            hipLaunchKernelGGL((k_gemm_grouped<double, BM_, BM_, 2, 2, false, false, BK_, MW_>), grid, block, 0, s, G); \
    } while (0)
    if (k.fmt == FMT_F64) {
        if (k.bm == 128) AQ_GROUPED(128, 16, 2);   // the default FP64 variant's kernel
        else AQ_GROUPED(64, 16, 1);
    } else if (k.bm == 128) {
        hipLaunchKernelGGL((k_gemm_bf16x3_grouped<128, 128, 2, 2>), grid, block, 0, s, G);
    } else {
        hipLaunchKernelGGL((k_gemm_bf16x3_grouped<64, 64, 2, 2>), grid, block, 0, s, G);
    }
#undef AQ_GROUPED
}

}  // namespace

struct GemmBatcher {
    int device = 0;
    int cus = 0;                          // compute units of the device
    hipStream_t stream = nullptr;
    std::chrono::microseconds window{0};
    std::mutex mu;
    std::condition_variable cv;
    int in_stage = 0;                     // contexts inside a contraction stage
    struct Launch {
        hipEvent_t done = nullptr;        // recorded after the grouped kernel
        int refs = 0;                     // members that have not yet waited on it
    };
    struct Item {
        GroupKey key;
        GemmArgs a;
        int64_t blocks;
        hipEvent_t ready;                 // the member's operands are written
        Launch* launch = nullptr;
        bool failed = false;
    };
    std::vector<Item*> pending;           // in submission order
    std::vector<hipEvent_t> spare;
    // per key: moving average of whether a wait for a partner found one; a
    // key whose GEMMs rarely meet another slot's stops waiting (launches solo)
    std::map<std::tuple<int, int, int, int, int>, double> met;
    int64_t launches = 0, problems = 0, solo = 0;
    int max_group = 0;
};

static std::tuple<int, int, int, int, int> key_tuple(const GroupKey& k) {
    return std::make_tuple(k.fmt, k.ta, k.tb, k.bm, k.bn);
}

// Launch up to kMaxGroup pending items of `key` as one grouped kernel (mu held).
static void launch_group(GemmBatcher* b, const GroupKey& key) {
    std::vector<GemmBatcher::Item*> take;
    for (auto it = b->pending.begin(); it != b->pending.end() && int(take.size()) < kMaxGroup;) {
        if ((*it)->key == key) {
            take.push_back(*it);
            it = b->pending.erase(it);
        } else {
            ++it;
        }
    }
    if (take.empty()) return;
    auto* L = new GemmBatcher::Launch();
    try {
        GemmGroup G{};
        G.n = int(take.size());
        int64_t blocks = 0;
        for (int i = 0; i < G.n; ++i) {
            G.p[i] = take[size_t(i)]->a;
            G.first[i] = blocks;
            blocks += take[size_t(i)]->blocks;
            AQ_HIP(hipStreamWaitEvent(b->stream, take[size_t(i)]->ready, 0));
        }
        G.first[G.n] = blocks;
        AQ_CHECK(blocks > 0 && blocks < (int64_t(1) << 31), AQ_ERR_UNSUPPORTED, "grouped GEMM grid too large");
        if (!b->spare.empty()) {
            L->done = b->spare.back();
            b->spare.pop_back();
        } else {
            AQ_HIP(hipEventCreateWithFlags(&L->done, hipEventDisableTiming));
        }
        launch_grouped(key, G, blocks, b->stream, b->cus);
        AQ_HIP(hipGetLastError());
        AQ_HIP(hipEventRecord(L->done, b->stream));
    } catch (...) {
        if (L->done) b->spare.push_back(L->done);
        delete L;
        for (auto* it : take) it->failed = true;
        b->cv.notify_all();
        throw;
    }
    L->refs = int(take.size());
    for (auto* it : take) it->launch = L;
    b->launches += 1;
    b->problems += int64_t(take.size());
    b->max_group = std::max(b->max_group, int(take.size()));
    b->cv.notify_all();
}

// Offer one GEMM (arguments final, split-K partials allocated) for a grouped
// launch.  It joins same-key GEMMs that other slots have pending; when none
// is pending and another slot is inside a contraction stage, it waits up to
// the window for one (only for keys whose waits have usually found a
// partner).  Returns true once it ran in a grouped launch (ctx->stream is
// then ordered after it), false when it found no partner: the caller launches
// it alone on its own stream, exactly as without grouping.
static bool batcher_submit(aq_ctx* ctx, const GroupKey& key, const GemmArgs& a, int64_t blocks) {
    GemmBatcher* b = ctx->batcher;
    if (!ctx->group_ready) AQ_HIP(hipEventCreateWithFlags(&ctx->group_ready, hipEventDisableTiming));
    AQ_HIP(hipEventRecord(ctx->group_ready, ctx->stream));
    GemmBatcher::Item item{key, a, blocks, ctx->group_ready};
    std::unique_lock<std::mutex> lk(b->mu);
    b->pending.push_back(&item);
    struct Unlist {   // an item still pending when this frame unwinds is withdrawn (mu held)
        GemmBatcher* b;
        GemmBatcher::Item* it;
        ~Unlist() {
            auto pos = std::find(b->pending.begin(), b->pending.end(), it);
            if (pos != b->pending.end()) b->pending.erase(pos);
        }
    } unlist{b, &item};
    auto same = [&]() {
        int n = 0;
        for (const auto* x : b->pending) n += x->key == key ? 1 : 0;
        return n;
    };
    if (same() >= 2) {
        launch_group(b, key);
    } else {
        double& met = b->met.emplace(key_tuple(key), 1.0).first->second;
        if (b->in_stage >= 2 && met >= 0.2) {
            const auto deadline = std::chrono::steady_clock::now() + b->window;
            while (!item.launch && !item.failed && same() < 2 && std::chrono::steady_clock::now() < deadline)
                b->cv.wait_until(lk, deadline);
            const bool found = item.launch || same() >= 2;
            met = 0.9 * met + (found ? 0.1 : 0.0);
            if (!item.launch && !item.failed && same() >= 2) launch_group(b, key);
        }
    }
    if (!item.launch && !item.failed) {   // no partner: the caller launches it alone
        b->solo += 1;
        return false;
    }
    AQ_CHECK(!item.failed, AQ_ERR_HIP, "grouped GEMM launch failed");
    GemmBatcher::Launch* L = item.launch;
    const hipError_t e = hipStreamWaitEvent(ctx->stream, L->done, 0);
    if (--L->refs == 0) {
        b->spare.push_back(L->done);
        delete L;
    }
    AQ_HIP(e);
    return true;
}

GemmBatcher* batcher_create(int device, int window_us) {
    AQ_CHECK(window_us >= 0, AQ_ERR_INVALID, "grouping window must be >= 0 us");
    std::unique_ptr<GemmBatcher> b(new GemmBatcher());
    b->device = device;
    b->window = std::chrono::microseconds(window_us);
    AQ_HIP(hipSetDevice(device));
    AQ_HIP(hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, device));
    AQ_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    return b.release();
}

void batcher_destroy(GemmBatcher* b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    (void)hipStreamSynchronize(b->stream);
    for (auto e : b->spare) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(b->stream);
    delete b;
}

void batcher_stage(GemmBatcher* b, int delta) {
    std::lock_guard<std::mutex> lk(b->mu);
    b->in_stage += delta;
    b->cv.notify_all();
}

void batcher_stats(GemmBatcher* b, int64_t* launches, int64_t* problems, int* max_group, int64_t* solo) {
    std::lock_guard<std::mutex> lk(b->mu);
    if (launches) *launches = b->launches;
    if (problems) *problems = b->problems;
    if (max_group) *max_group = b->max_group;
    if (solo) *solo = b->solo;
}

void gemm(aq_ctx* ctx, const GemmDesc& d) {
    AQ_CHECK(d.M > 0 && d.N > 0 && d.K > 0 && d.batch > 0, AQ_ERR_INVALID, "empty GEMM");
    AQ_CHECK(d.batch < 65536, AQ_ERR_UNSUPPORTED, "GEMM batch too large");
    if (d.fmt == FMT_BF16X2)
        AQ_CHECK(d.transA && !d.transB && d.lda % 8 == 0 && d.ldb % 8 == 0 && d.planeA % 8 == 0 &&
                     d.planeB % 8 == 0 && d.strideA % 8 == 0 && d.strideB % 8 == 0,
                 AQ_ERR_INVALID, "BF16x3 GEMM needs K-contiguous operands with 8-aligned leading dimensions");

    // base block shape from the problem shape
    int BM = 128, BN = 128;
    if (d.N <= 16) { BM = 128; BN = 16; }
    else if (d.M <= 16) { BM = 16; BN = 128; }
    else if (((d.M + 127) / 128) * ((d.N + 127) / 128) * d.batch < 256) { BM = 64; BN = 64; }
    // FP64/FP32 with one short side (o = 48 against v^2 o = 4.0e6 rows): a
    // 64-wide block runs 64/48 of the useful MFMAs where 128 ran 128/48
    else if (d.fmt != FMT_BF16X2 && d.N <= 64) { BM = 128; BN = 64; }
    else if (d.fmt != FMT_BF16X2 && d.M <= 64) { BM = 64; BN = 128; }

    // Split-K when the output alone cannot fill the device.  The decision and
    // the k-partition depend only on the shape (base tile grid, 32-wide
    // k-steps that every variant's BK divides), so every kernel variant
    // computes each element with the same summation order.
    constexpr int64_t KSTEP = 32;
    const int64_t base_tiles = ((d.M + BM - 1) / BM) * ((d.N + BN - 1) / BN);
    AQ_CHECK(base_tiles < (int64_t(1) << 31), AQ_ERR_UNSUPPORTED, "GEMM grid too large");
    // The split count s (at most ~two rounds of W = 512 co-resident blocks,
    // 8 k-steps per split, 256 splits) minimises the rounds per unit of work,
    // ceil(T*s / W) / s with T tiles: e.g. 324 tiles split 3 ways fill two
    // rounds (972 blocks) where 4 ways need three (1296).  Ties take the
    // smaller s (less partial traffic).
    int nsplit = 1;
    const int64_t ksteps = (d.K + KSTEP - 1) / KSTEP;
    const int64_t T = base_tiles * d.batch, W = 512;
    // 64 x 64 blocks: four fit a CU, so a small grid is spread over up to
    // ~two rounds of 1024
    const int64_t Ws = (BM == 64 && BN == 64) ? 1024 : W;
    if (T < W && ksteps >= 16) {
        const int64_t W = Ws;
        const int64_t smax = std::max<int64_t>(1, std::min<int64_t>({(2 * W + T - 1) / T, ksteps / 8, 256}));
        int64_t best = 1;
        double best_cost = 1.0;   // s = 1: one round per unit of work
        for (int64_t sc = 2; sc <= smax; ++sc) {
            const double cost = double((T * sc + W - 1) / W) / double(sc);
            if (cost < best_cost - 1e-12) {
                best_cost = cost;
                best = sc;
            }
        }
        nsplit = int(best);
        while (nsplit > 1 && int64_t(nsplit) * d.batch >= 65536) nsplit /= 2;
    } else if (d.fmt == FMT_F64 && T > W && T <= 8 * W && ksteps >= 64) {
        // A grid of one to eight rounds with a ragged last round (the packed
        // ladder: 3260 and 2907 tiles = 6.4 and 5.7 rounds): split s ways when
        // that cuts the rounds per unit of work, ceil(T*s / W) / s, by at
        // least 4% (2907 tiles 7 ways: 4.8% on paper, 64.7 -> 59.6 ms
        // measured), with at least 8 k-steps per split and at most 4 GiB of
        // FP64 partials.
        // FP64 only: the BF16x3 phase keeps its operand copies in that memory.
        const double r1 = double((T + W - 1) / W);
        int64_t best = 1;
        double best_cost = r1;
        for (int64_t sc = 2; sc <= 8 && ksteps / sc >= 8; ++sc) {
            if (double(sc) * double(d.batch) * double(d.M) * double(d.N) * 8.0 > 4.0 * double(int64_t(1) << 30)) break;
            const double cost = double((T * sc + W - 1) / W) / double(sc);
            if (cost < best_cost - 1e-12) {
                best_cost = cost;
                best = sc;
            }
        }
        if (best_cost <= 0.96 * r1) nsplit = int(best);
        while (nsplit > 1 && int64_t(nsplit) * d.batch >= 65536) nsplit /= 2;
    } else if (d.fmt == FMT_BF16X2 && d.M >= 256 && d.N >= 256 && ksteps >= 64) {
        // BF16x3 GEMMs this large run on 256 x 256 blocks, one per CU: the
        // same ragged-round rule over the grid of 256 x 256 tiles and W2 = 256
        // (the packed ladder: 815 and 810 tiles = 3.2 rounds, split 5 ways
        // fill 16 rounds instead of 4 rounds of whole tiles), with the partials
        // capped at 2 GiB so that the operand cache keeps its copies.
        const int64_t W2 = 256, T2 = ((d.M + 255) / 256) * ((d.N + 255) / 256) * d.batch;
        if (T2 > W2 && T2 <= 8 * W2) {
            const double r1 = double((T2 + W2 - 1) / W2);
            int64_t best = 1;
            double best_cost = r1;
            for (int64_t sc = 2; sc <= 8 && ksteps / sc >= 8; ++sc) {
                if (double(sc) * double(d.batch) * double(d.M) * double(d.N) * 8.0 > 2.0 * double(int64_t(1) << 30))
                    break;
                const double cost = double((T2 * sc + W2 - 1) / W2) / double(sc);
                if (cost < best_cost - 1e-12) {
                    best_cost = cost;
                    best = sc;
                }
            }
            if (best_cost <= 0.96 * r1) nsplit = int(best);
            while (nsplit > 1 && int64_t(nsplit) * d.batch >= 65536) nsplit /= 2;
        }
    }
    int64_t kchunk = ((ksteps + nsplit - 1) / nsplit) * KSTEP;
    nsplit = int((d.K + kchunk - 1) / kchunk);

    GemmArgs a;
    a.M = d.M; a.N = d.N; a.K = d.K; a.kchunk = kchunk; a.nsplit = nsplit;
    a.A = d.A; a.lda = d.lda; a.strideA = d.strideA; a.planeA = d.planeA;
    a.B = d.B; a.ldb = d.ldb; a.strideB = d.strideB; a.planeB = d.planeB;
    a.alpha = d.alpha; a.beta = d.beta;
    a.C = d.C; a.ldc = d.ldc; a.strideC = d.strideC;
    a.batch = int(d.batch);
    a.transA = d.transA; a.transB = d.transB;
    a.P = nullptr;
    a.cmap = d.cmap;
    AQ_CHECK(!d.cmap.mapped || (d.M < (int64_t(1) << 31) && d.N < (int64_t(1) << 31)), AQ_ERR_INVALID,
             "mapped GEMM output needs M and N below 2^31");
    // an element of a tile lies below 144 * ld + 256 elements from the tile origin
    a.off32 = (144 * std::max(d.lda, d.ldb) + 256) * 8 < (int64_t(1) << 31) ? 1 : 0;
    // groups of 8 tile rows: measured on the 13824^3 ring GEMMs, BF16x3 279 -> 357 TFLOP/s
    // (FP64 unchanged in time, MFMA-bound, with less L2-to-fabric traffic)
    a.raster = 8;

    if (ctx->batcher && !ctx->capture && group_shape(d.fmt, BM, BN)) {
        std::unique_ptr<Scratch> part;
        if (nsplit > 1) {
            part.reset(new Scratch(ctx, size_t(nsplit) * d.batch * d.M * d.N * 8));
            a.P = part->as<double>();
        }
        if (batcher_submit(ctx, GroupKey{int(d.fmt), d.transA ? 1 : 0, d.transB ? 1 : 0, BM, BN}, a,
                           base_tiles * d.batch * nsplit)) {
            if (nsplit > 1)
                launch_reduce_splitk(ctx, a.P, nsplit, d.M, d.N, d.batch, d.beta, d.C, d.ldc, d.strideC, d.cmap);
            return;
        }
        a.P = nullptr;   // alone: the launch below allocates its own partials
    }

    int variant = ctx->gemm_variant[d.fmt];
    if (variant < 0) variant = ctx->autotune ? tuned_variant(ctx, d, BM, BN, a) : default_gemm_variant(d.fmt);
    variant_shape(d.fmt, variant, d.M, d.N, BM, BN);
    const int64_t tiles = ((d.M + BM - 1) / BM) * ((d.N + BN - 1) / BN);

    std::unique_ptr<Scratch> part;
    if (nsplit > 1) {
        part.reset(new Scratch(ctx, size_t(nsplit) * d.batch * d.M * d.N * 8));
        a.P = part->as<double>();
    }
    dim3 grid(unsigned(tiles), 1, unsigned(d.batch * nsplit));
    const double flops = 2.0 * double(d.M) * double(d.N) * double(d.K) * double(d.batch);
    const double bytes = double(fmt_bytes(d.fmt)) * (double(d.M) * d.K + double(d.N) * d.K) * d.batch +
                         8.0 * double(d.M) * d.N * d.batch * (d.beta != 0.0 ? 2 : 1);
    char name[64];
    // one timing class per kernel instantiation: tile and operand layouts
    // (N = opA rows / opB cols contiguous, T = k contiguous)
    snprintf(name, sizeof(name), "%s_%dx%d_%c%c", d.tag ? d.tag : "gemm", BM, BN, d.transA ? 'T' : 'N',
             d.transB ? 'T' : 'N');
    // AQ_GEMM_TRACE=1: one line per launch on stderr (shape, layout, tile, variant, split)
    static const bool trace = [] {
        const char* e = getenv("AQ_GEMM_TRACE");
        return e && e[0] == '1';
    }();
    if (trace)
        fprintf(stderr, "aq_gemm %s M=%lld N=%lld K=%lld batch=%lld beta=%g variant=%d nsplit=%d blocks=%lld\n", name,
                (long long)d.M, (long long)d.N, (long long)d.K, (long long)d.batch, d.beta, variant, nsplit,
                (long long)(tiles * d.batch * nsplit));
    {
        TimedLaunch tl(ctx, name, flops, bytes);
        dispatch(ctx, d.fmt, BM, BN, variant, a, grid);
        AQ_HIP(hipGetLastError());
    }
    if (nsplit > 1)
        launch_reduce_splitk(ctx, a.P, nsplit, d.M, d.N, d.batch, d.beta, d.C, d.ldc, d.strideC, d.cmap);
}

int autotune_entries(aq_tune_entry* out, int cap) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    int n = 0;
    for (const auto& kv : g_tune) {
        if (out && n < cap) {
            aq_tune_entry& e = out[n];
            e.device = kv.first.device;
            e.precision = kv.first.fmt == FMT_F64 ? AQ_PREC_FP64 : kv.first.fmt == FMT_F32 ? AQ_PREC_FP32 : AQ_PREC_BF16X3;
            e.trans_a = kv.first.ta;
            e.trans_b = kv.first.tb;
            e.m_bucket = kv.first.m;
            e.n_bucket = kv.first.n;
            e.k_bucket = kv.first.k;
            e.batch_bucket = kv.first.batch;
            e.variant = kv.second.variant;
            for (int v = 0; v < 5; ++v) e.ms[v] = kv.second.ms[v];
        }
        ++n;
    }
    return n;
}

}  // namespace aqamd
