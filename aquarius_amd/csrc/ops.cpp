// Tensor lifecycle and the CTFTensor operator surface (mult / sum / scale /
// dot / weight / norm / key-value I/O / slice) on dense column-major device
// tensors.  Reference semantics: src/tensor/ctf_tensor.cxx:403-514 and the
// expression front end src/tensor/indexable_tensor.hpp:288-346.
//
// A contraction is planned as at most: one strided pass per operand that is
// not already a matrix view (cached per tensor version, so constant integral
// blocks are permuted/converted once per solve), one MFMA GEMM (batched,
// split-K when M*N is small), and one strided pass into C when C's label
// order interleaves the two free groups.
#include "internal.hpp"

#include <algorithm>
#include <cstring>
#include <memory>
#include <numeric>

namespace aqamd {

// ------------------------------------------------------------------ labels

struct Labeled {
    const double* data = nullptr;
    const aq_tensor* t = nullptr;   // source tensor (for caching), may be null for temps
    std::vector<int64_t> len;       // per unique label
    std::vector<int64_t> stride;    // per unique label (sum over repeated positions)
    std::string lab;                // unique labels, first-occurrence order
    std::vector<int> dim;           // tensor dimension of each unique label (first occurrence)
    bool has_repeat = false;
    int64_t size() const {
        int64_t s = 1;
        for (auto l : len) s *= l;
        return s;
    }
    bool contiguous_canonical() const {   // strides are the packed col-major strides of `len`
        int64_t s = 1;
        for (size_t k = 0; k < len.size(); ++k) {
            if (len[k] != 1 && stride[k] != s) return false;
            s *= len[k];
        }
        return true;
    }
};

static Labeled label(const aq_tensor* t, const char* idx, const char* what) {
    AQ_CHECK(t, AQ_ERR_INVALID, std::string("null tensor ") + what);
    AQ_CHECK(idx, AQ_ERR_INVALID, std::string("null index string for ") + what);
    const size_t n = std::strlen(idx);
    AQ_CHECK((int)n == t->ndim, AQ_ERR_INVALID,
             std::string("invalid ndim: index '") + idx + "' for a " + std::to_string(t->ndim) + "-d tensor (" + what + ")");
    Labeled L;
    L.data = t->data;
    L.t = t;
    int64_t s = 1;
    for (size_t k = 0; k < n; ++k) {
        const char c = idx[k];
        const size_t pos = L.lab.find(c);
        if (pos == std::string::npos) {
            L.lab.push_back(c);
            L.len.push_back(t->len[k]);
            L.stride.push_back(s);
            L.dim.push_back((int)k);
        } else {
            AQ_CHECK(L.len[pos] == t->len[k], AQ_ERR_INVALID,
                     std::string("label '") + c + "' has inconsistent lengths in " + what);
            L.stride[pos] += s;
            L.has_repeat = true;
        }
        s *= t->len[k];
    }
    return L;
}

static int64_t label_len(const std::vector<const Labeled*>& Ls, char c) {
    int64_t l = -1;
    for (auto* L : Ls) {
        auto p = L->lab.find(c);
        if (p != std::string::npos) {
            if (l >= 0)
                AQ_CHECK(l == L->len[p], AQ_ERR_INVALID, std::string("label '") + c + "' has different lengths across tensors");
            l = L->len[p];
        }
    }
    return l;
}

// Plan dst[out labels] (+)= src[...] with reduction over src-only labels.
static StridedPlan make_plan(const Labeled& src, const Labeled& dst) {
    StridedPlan p;
    std::vector<int64_t> len, sd, ss, rl, rs;
    for (size_t k = 0; k < dst.lab.size(); ++k) {
        const char c = dst.lab[k];
        auto q = src.lab.find(c);
        if (q != std::string::npos)
            AQ_CHECK(src.len[q] == dst.len[k], AQ_ERR_INVALID, std::string("label '") + c + "' length mismatch");
        len.push_back(dst.len[k]);
        sd.push_back(dst.stride[k]);
        ss.push_back(q == std::string::npos ? 0 : src.stride[q]);
    }
    for (size_t k = 0; k < src.lab.size(); ++k) {
        if (dst.lab.find(src.lab[k]) == std::string::npos) {
            rl.push_back(src.len[k]);
            rs.push_back(src.stride[k]);
        }
    }
    // drop unit dims, then merge dims contiguous in both src and dst
    std::vector<int64_t> L2, D2, S2;
    for (size_t k = 0; k < len.size(); ++k) {
        if (len[k] == 1) continue;
        if (!L2.empty() && D2.back() * L2.back() == sd[k] && S2.back() * L2.back() == ss[k]) {
            L2.back() *= len[k];
            continue;
        }
        L2.push_back(len[k]);
        D2.push_back(sd[k]);
        S2.push_back(ss[k]);
    }
    std::vector<int64_t> RL2, RS2;
    for (size_t k = 0; k < rl.size(); ++k) {
        if (rl[k] == 1) continue;
        if (!RL2.empty() && RS2.back() * RL2.back() == rs[k]) {
            RL2.back() *= rl[k];
            continue;
        }
        RL2.push_back(rl[k]);
        RS2.push_back(rs[k]);
    }
    AQ_CHECK(L2.size() <= (size_t)MAXDIM && RL2.size() <= (size_t)MAXDIM, AQ_ERR_UNSUPPORTED,
             "more than 8 non-mergeable dimensions");
    p.nd = (int)L2.size();
    for (int k = 0; k < p.nd; ++k) {
        p.len[k] = L2[k];
        p.sdst[k] = D2[k];
        p.ssrc[k] = S2[k];
        p.nout *= L2[k];
    }
    p.nr = (int)RL2.size();
    for (int k = 0; k < p.nr; ++k) {
        p.rlen[k] = RL2[k];
        p.rsrc[k] = RS2[k];
        p.nred *= RL2[k];
    }
    for (size_t k = 0; k < len.size(); ++k)
        if (len[k] == 0) p.nout = 0;
    for (size_t k = 0; k < rl.size(); ++k)
        if (rl[k] == 0) p.nred = 0;
    return p;
}

// A packed column-major buffer described by unique labels.
static Labeled packed(const double* data, const std::string& lab, const std::vector<int64_t>& len) {
    Labeled L;
    L.data = data;
    L.lab = lab;
    L.len = len;
    int64_t s = 1;
    for (auto l : len) {
        L.stride.push_back(s);
        s *= l;
    }
    return L;
}

static std::vector<int64_t> lens_of(const std::string& labs, const std::vector<const Labeled*>& Ls) {
    std::vector<int64_t> out;
    for (char c : labs) out.push_back(label_len(Ls, c));
    return out;
}

static bool overlaps(const aq_tensor* a, const aq_tensor* b) { return a && b && a->data == b->data; }

// Cache key of a copy of tensor L.t laid out in label order `order`: the
// permutation is recorded in tensor dimensions, since the same tensor is
// addressed with different label names by different calls.
static std::string cache_key(const Labeled& L, const std::string& order, Fmt fmt) {
    std::string k = std::to_string(L.t->id) + "|";
    for (char c : order) k += std::to_string(L.dim[L.lab.find(c)]) + ",";
    return k + "|" + std::to_string(int(fmt));
}

// ------------------------------------------------------- operand cache

// Pins held for the duration of one operation.
struct PinSet {
    aq_ctx* ctx;
    std::vector<std::string> keys;
    explicit PinSet(aq_ctx* c) : ctx(c) {}
    ~PinSet() {
        for (auto& k : keys) {
            auto it = ctx->cache.find(k);
            if (it != ctx->cache.end()) it->second.pins -= 1;
        }
    }
};

struct Operand {
    const void* ptr = nullptr;
    int64_t ld = 0;       // leading dimension (elements)
    bool trans = false;
    int64_t bstride = 0;  // batch stride (elements)
    int64_t plane = 0;    // BF16X2 lo-plane offset (elements)
};

// Produce `src` laid out as packed column-major in label order `order`,
// in format `fmt`.  Cached by (tensor id, version, order, fmt) when `src`
// is a whole tensor; otherwise written to `tmp` (owned by the caller).
//
// `nlead` > 0: the first `nlead` labels of `order` form a leading group whose
// combined extent is padded to `lead_pad` elements (zero filled), as the BF16x3
// GEMM's 16-byte K-vector loads need.
static const void* materialize(aq_ctx* ctx, const Labeled& src, const std::string& order, Fmt fmt,
                               PinSet& pins, std::vector<std::unique_ptr<Scratch>>& tmp, int nlead = 0,
                               int64_t lead_pad = 0) {
    std::vector<int64_t> len;
    for (char c : order) len.push_back(src.len[src.lab.find(c)]);
    std::vector<int64_t> stride;
    int64_t s = 1, lead = 1;
    for (size_t k = 0; k < len.size(); ++k) {
        if (nlead > 0 && int(k) == nlead) s = lead_pad;
        stride.push_back(s);
        s *= len[k];
        if (int(k) < nlead) lead *= len[k];
    }
    const bool padded = nlead > 0 && lead_pad != lead;
    AQ_CHECK(nlead == 0 || lead_pad >= lead, AQ_ERR_INVALID, "bad operand padding");
    const int64_t n = nlead == int(len.size()) ? lead_pad : s;   // elements per plane
    const size_t bytes = std::max<size_t>(size_t(n) * fmt_bytes(fmt), 256);
    Labeled dst;
    dst.lab = order;
    dst.len = len;
    dst.stride = stride;
    StridedPlan p = make_plan(src, dst);
    auto clear = [&](void* ptr) {
        if (padded) launch_zero_bytes(ctx, ptr, bytes);
    };
    const bool cacheable = src.t != nullptr && !src.has_repeat;
    if (cacheable) {
        std::string key = cache_key(src, order, fmt) + (padded ? "|p" + std::to_string(lead_pad) : "");
        auto it = ctx->cache.find(key);
        if (it != ctx->cache.end() && it->second.bytes == bytes) {
            auto& e = it->second;
            ctx->cache_lru.splice(ctx->cache_lru.begin(), ctx->cache_lru, e.lru);
            e.pins += 1;
            pins.keys.push_back(key);
            if (!(e.valid && e.version == src.t->version)) {
                launch_strided_convert(ctx, p, src.data, e.ptr, fmt, n);
                e.version = src.t->version;
                e.valid = true;
                if (ctx->capture) ctx->capture->rewritten.push_back(e.ptr);
            } else if (ctx->capture) {
                const auto& rw = ctx->capture->rewritten;
                if (std::find(rw.begin(), rw.end(), e.ptr) == rw.end())
                    ctx->capture->fresh_reads.emplace_back(src.t, src.t->version);
            }
            return e.ptr;
        }
        AQ_CHECK(!ctx->capture, AQ_ERR_UNSUPPORTED, "operand-cache miss during graph capture");
        if (it != ctx->cache.end()) {   // size changed (cannot happen for a live tensor id)
            ctx->cache_lru.erase(it->second.lru);
            dfree(ctx, it->second.ptr);
            ctx->cache_bytes -= it->second.bytes;
            ctx->cache.erase(it);
        }
        if (ctx_in_use(ctx) + bytes > ctx->budget) evict_cache(ctx, bytes);
        if (ctx_in_use(ctx) + bytes <= ctx->budget) {
            void* ptr = dmalloc(ctx, bytes);
            if (ptr) {
                CacheEntry e;
                e.ptr = ptr;
                e.bytes = bytes;
                ctx->cache_lru.push_front(key);
                e.lru = ctx->cache_lru.begin();
                e.pins = 1;
                ctx->cache_bytes += bytes;
                note_cache_alloc(ctx);
                clear(ptr);
                launch_strided_convert(ctx, p, src.data, ptr, fmt, n);
                e.version = src.t->version;
                e.valid = true;
                ctx->cache[key] = e;
                pins.keys.push_back(key);
                return ptr;
            }
        }
        // fall through: uncached scratch copy
    }
    tmp.emplace_back(new Scratch(ctx, bytes));
    clear(tmp.back()->p);
    launch_strided_convert(ctx, p, src.data, tmp.back()->p, fmt, n);
    return tmp.back()->p;
}

// Reduce / collapse a labeled operand to the unique labels in `keep` (in
// first-occurrence order) when it has repeated labels or labels outside `keep`.
static Labeled simplify(aq_ctx* ctx, const Labeled& in, const std::string& keep,
                        std::vector<std::unique_ptr<Scratch>>& tmp) {
    std::string lab;
    for (char c : in.lab)
        if (keep.find(c) != std::string::npos) lab.push_back(c);
    if (lab == in.lab && !in.has_repeat) return in;
    std::vector<int64_t> len;
    for (char c : lab) len.push_back(in.len[in.lab.find(c)]);
    int64_t n = 1;
    for (auto l : len) n *= l;
    tmp.emplace_back(new Scratch(ctx, std::max<int64_t>(n, 1) * 8));
    Labeled out = packed(tmp.back()->as<double>(), lab, len);
    launch_strided_sum(ctx, make_plan(in, out), 1.0, in.data, 0.0, const_cast<double*>(out.data));
    return out;
}

static std::string filter(const std::string& s, const std::string& a, bool in_a, const std::string& b, bool in_b) {
    std::string r;
    for (char c : s)
        if ((a.find(c) != std::string::npos) == in_a && (b.find(c) != std::string::npos) == in_b) r.push_back(c);
    return r;
}

static int64_t prod_len(const std::string& labs, const Labeled& L) {
    int64_t p = 1;
    for (char c : labs) p *= L.len[L.lab.find(c)];
    return p;
}

// Can L be viewed as a column-major matrix [g1 ; g2] (g1 fastest) with the
// batch labels outermost?  Requires the packed layout of exactly that order.
static bool is_view(const Labeled& L, const std::string& order) {
    if (L.lab.size() != order.size()) return false;
    if (!L.contiguous_canonical()) return false;
    // unit-length labels may sit anywhere
    std::string a, b;
    for (char c : L.lab)
        if (L.len[L.lab.find(c)] != 1) a.push_back(c);
    for (char c : order)
        if (L.len[L.lab.find(c)] != 1) b.push_back(c);
    return a == b;
}

static bool cache_fresh(aq_ctx* ctx, const Labeled& L, const std::string& order, Fmt fmt) {
    if (!L.t || L.has_repeat) return false;
    auto it = ctx->cache.find(cache_key(L, order, fmt));
    return it != ctx->cache.end() && it->second.valid && it->second.version == L.t->version;
}

static void bump(aq_tensor* t) { t->version += 1; }

// ------------------------------------------------------------- contract

static void contract_impl(aq_ctx* ctx, double alpha, const aq_tensor* tA, const char* ia, const aq_tensor* tB,
                          const char* ib, double beta, aq_tensor* tC, const char* ic) {
    Labeled A0 = label(tA, ia, "A"), B0 = label(tB, ib, "B"), C0 = label(tC, ic, "C");
    for (char c : A0.lab + B0.lab + C0.lab) label_len({&A0, &B0, &C0}, c);

    std::vector<std::unique_ptr<Scratch>> tmp;
    PinSet pins(ctx);

    if (C0.size() == 0) return;
    // Step 1: operands with repeated labels or labels summed away on one side
    Labeled A = simplify(ctx, A0, B0.lab + C0.lab, tmp);
    Labeled B = simplify(ctx, B0, A.lab + C0.lab, tmp);
    A = simplify(ctx, A, B.lab + C0.lab, tmp);

    const std::string Mlab = filter(A.lab, B.lab, false, C0.lab, true);
    const std::string Klab = filter(A.lab, B.lab, true, C0.lab, false);
    const std::string Blab_c = filter(C0.lab, A.lab, true, B.lab, true);
    const std::string Nlab = filter(B.lab, A.lab, false, C0.lab, true);

    // C target: direct when C has unique labels covered by A/B and no aliasing
    std::string Cuse;   // labels of the GEMM output
    for (char c : C0.lab)
        if (A.lab.find(c) != std::string::npos || B.lab.find(c) != std::string::npos) Cuse.push_back(c);
    const bool c_alias = overlaps(tC, tA) || overlaps(tC, tB);
    const bool c_simple = !C0.has_repeat && Cuse == C0.lab && !c_alias;

    const int64_t M = prod_len(Mlab, A), N = prod_len(Nlab, B), K = prod_len(Klab, A),
                  NB = prod_len(Blab_c, A);
    const double flops = 2.0 * double(M) * double(N) * double(K) * double(NB);
    // reduced precision only where the contraction is MFMA work (K >= 16);
    // outer products and thin-K updates stay FP64
    Fmt fmt = FMT_F64;
    if (ctx->precision != AQ_PREC_FP64 && flops >= ctx->min_reduced_flops && K >= 16)
        fmt = ctx->precision == AQ_PREC_FP32 ? FMT_F32 : FMT_BF16X2;

    // K == 0: pure scaling of C
    if (K == 0 || M == 0 || N == 0 || NB == 0) {
        if (beta != 1.0) {
            Labeled Cp = C0;
            StridedPlan p = make_plan(Cp, Cp);
            launch_strided_sum(ctx, p, 0.0, tC->data, beta, tC->data);
            bump(tC);
        }
        return;
    }

    // Step 2: choose label orders / roles minimising strided passes
    auto order_in = [](const std::string& labs, const std::string& within) {
        std::string r;
        for (char c : within)
            if (labs.find(c) != std::string::npos) r.push_back(c);
        return r;
    };
    std::vector<std::string> Mc = {order_in(Mlab, C0.lab), order_in(Mlab, A.lab)};
    std::vector<std::string> Nc = {order_in(Nlab, C0.lab), order_in(Nlab, B.lab)};
    std::vector<std::string> Kc = {order_in(Klab, A.lab), order_in(Klab, B.lab)};
    const std::string Bo = order_in(Blab_c, C0.lab);

    struct Choice {
        double cost;
        std::string Mo, No, Ko;
        bool swap, cdirect, adirect, atrans, bdirect, btrans, cmapped;
    } best{1e300, "", "", "", false, false, false, false, false, false, false};

    // Can the GEMM output [rows cols Bo] be written straight into C through a
    // CMap?  At most three non-unit labels per side, at most one batch label,
    // and the first non-unit row label contiguous in C (a wavefront's 16
    // consecutive rows then write one contiguous run).
    auto non_unit = [&](const std::string& labs) {
        std::string r;
        for (char c : labs)
            if (C0.len[C0.lab.find(c)] != 1) r.push_back(c);
        return r;
    };
    auto mappable = [&](const std::string& rows, const std::string& cols) {
        const std::string r = non_unit(rows), c = non_unit(cols), b = non_unit(Bo);
        if (r.empty() || r.size() > 3 || c.size() > 3 || b.size() > 1) return false;
        if (C0.stride[C0.lab.find(r[0])] != 1) return false;
        return prod_len(rows, C0) < (int64_t(1) << 31) && prod_len(cols, C0) < (int64_t(1) << 31);
    };

    const double szA = double(A.size()), szB = double(B.size()), szC = double(C0.size());
    for (auto& Mo : Mc)
        for (auto& No : Nc)
            for (auto& Ko : Kc)
                for (int sw = 0; sw < 2; ++sw) {
                    Choice ch{0.0, Mo, No, Ko, sw == 1, false, false, false, false, false, false};
                    // GEMM output layout: [Mg Ng Bo]
                    const std::string outl = sw ? No + Mo + Bo : Mo + No + Bo;
                    ch.cdirect = c_simple && is_view(C0, outl);
                    if (!ch.cdirect) {
                        ch.cmapped = c_simple && (sw ? mappable(No, Mo) : mappable(Mo, No));
                        // mapped: the epilogue reads and writes C in place (a little index
                        // arithmetic); otherwise temp write + read + C read/write
                        ch.cost += (ch.cmapped ? 0.25 : 3.0) * szC;
                    }
                    // GEMM-A is A (rows Mo) or, swapped, B (rows No)
                    auto check = [&](const Labeled& L, const std::string& rows, bool& direct, bool& trans,
                                     double sz) {
                        direct = false;
                        trans = false;
                        if (fmt == FMT_F64) {
                            if (is_view(L, rows + Ko + Bo)) { direct = true; trans = false; return; }
                            if (is_view(L, Ko + rows + Bo)) { direct = true; trans = true; return; }
                        }
                        if (!cache_fresh(ctx, L, rows + Ko + Bo, fmt)) ch.cost += 2.0 * sz;
                    };
                    if (!sw) {
                        check(A, Mo, ch.adirect, ch.atrans, szA);
                        check(B, No, ch.bdirect, ch.btrans, szB);
                    } else {
                        check(B, No, ch.adirect, ch.atrans, szB);
                        check(A, Mo, ch.bdirect, ch.btrans, szA);
                    }
                    if (ch.cost < best.cost) best = ch;
                }

    const Choice& ch = best;
    const Labeled& GA = ch.swap ? B : A;
    const Labeled& GB = ch.swap ? A : B;
    const std::string Rg = ch.swap ? ch.No : ch.Mo;    // GEMM rows
    const std::string Cg = ch.swap ? ch.Mo : ch.No;    // GEMM cols
    const int64_t Mg = ch.swap ? N : M, Ng = ch.swap ? M : N;

    GemmDesc d{};
    d.M = Mg;
    d.N = Ng;
    d.K = K;
    d.batch = NB;
    d.fmt = fmt;
    d.alpha = alpha;
    // operand A of the GEMM: [Rg x K], batch outermost
    if (ch.adirect) {
        d.A = GA.data;
        d.transA = ch.atrans;
        d.lda = ch.atrans ? K : Mg;
    } else {
        d.A = materialize(ctx, GA, Rg + ch.Ko + Bo, fmt, pins, tmp);
        d.transA = false;
        d.lda = Mg;
    }
    d.strideA = Mg * K;
    d.planeA = Mg * K * NB;
    // operand B of the GEMM: [K x Cg]; GB holds labels Cg and K.
    // GB as [Cg ; K] (col-major) is the transposed view of the [K x Cg] operand.
    if (ch.bdirect) {
        d.B = GB.data;
        d.transB = !ch.btrans;            // check(): trans=false means [rows K] = [Cg K]
        d.ldb = ch.btrans ? K : Ng;
    } else {
        d.B = materialize(ctx, GB, Cg + ch.Ko + Bo, fmt, pins, tmp);
        d.transB = true;
        d.ldb = Ng;
    }
    d.strideB = Ng * K;
    d.planeB = Ng * K * NB;
    if (fmt == FMT_BF16X2) {
        // hi/lo BF16 planes, K-contiguous with K padded to a multiple of 8
        const int64_t Kp = (K + 7) / 8 * 8;
        const int nk = int(ch.Ko.size());
        d.A = materialize(ctx, GA, ch.Ko + Rg + Bo, fmt, pins, tmp, nk, Kp);
        d.transA = true;
        d.lda = Kp;
        d.strideA = Kp * Mg;
        d.planeA = Kp * Mg * NB;
        d.B = materialize(ctx, GB, ch.Ko + Cg + Bo, fmt, pins, tmp, nk, Kp);
        d.transB = false;
        d.ldb = Kp;
        d.strideB = Kp * Ng;
        d.planeB = Kp * Ng * NB;
    }
    d.tag = fmt == FMT_F64 ? "gemm_f64" : fmt == FMT_F32 ? "gemm_f32" : "gemm_bf16x3";

    if (ch.cdirect) {
        d.C = tC->data;
        d.ldc = Mg;
        d.strideC = Mg * Ng;
        d.beta = beta;
        gemm(ctx, d);
    } else if (ch.cmapped) {
        // the output [Rg Cg Bo] goes straight into the permuted C
        auto digits = [&](const std::string& labs, int& nd, int64_t* len, int64_t* st) {
            nd = 0;
            for (char c : labs) {
                const size_t k = C0.lab.find(c);
                if (C0.len[k] == 1) continue;
                len[nd] = C0.len[k];
                st[nd] = C0.stride[k];
                ++nd;
            }
        };
        d.cmap.mapped = 1;
        digits(Rg, d.cmap.nm, d.cmap.ml, d.cmap.ms);
        digits(Cg, d.cmap.nn, d.cmap.nl, d.cmap.ns);
        int nb = 0;
        int64_t blen[3] = {1, 1, 1}, bst[3] = {0, 0, 0};
        digits(Bo, nb, blen, bst);
        d.C = tC->data;
        d.ldc = Mg;
        d.strideC = nb ? bst[0] : 0;
        d.beta = beta;
        gemm(ctx, d);
    } else {
        // GEMM into a packed temp [Rg Cg Bo], then C = beta*C + temp (permuted,
        // broadcast over C-only labels, diagonal for repeated C labels)
        tmp.emplace_back(new Scratch(ctx, size_t(Mg) * Ng * NB * 8));
        d.C = tmp.back()->as<double>();
        d.ldc = Mg;
        d.strideC = Mg * Ng;
        d.beta = 0.0;
        gemm(ctx, d);
        const std::string tl = Rg + Cg + Bo;
        std::vector<int64_t> tlen;
        for (char c : tl) tlen.push_back(label_len({&A, &B}, c));
        Labeled T = packed(d.C, tl, tlen);
        launch_strided_sum(ctx, make_plan(T, C0), 1.0, d.C, beta, tC->data);
    }
    bump(tC);
}

static void sum_impl(aq_ctx* ctx, double alpha, const aq_tensor* tA, const char* ia, double beta, aq_tensor* tB,
                     const char* ib) {
    Labeled A = label(tA, ia, "A"), B = label(tB, ib, "B");
    for (char c : A.lab + B.lab) label_len({&A, &B}, c);
    if (B.size() == 0) return;
    StridedPlan p = make_plan(A, B);
    bool identity = p.nr == 0;
    for (int k = 0; k < p.nd && identity; ++k) identity = p.ssrc[k] == p.sdst[k];
    if (overlaps(tA, tB) && !identity) {
        // snapshot the source before writing the destination (aliased sum)
        Scratch snap(ctx, size_t(tA->size) * 8);
        AQ_HIP(hipMemcpyAsync(snap.p, tA->data, size_t(tA->size) * 8, hipMemcpyDeviceToDevice, ctx->stream));
        launch_strided_sum(ctx, p, alpha, snap.as<double>(), beta, tB->data);
    } else {
        launch_strided_sum(ctx, p, alpha, tA->data, beta, tB->data);
    }
    bump(tB);
}

static double dot_impl(aq_ctx* ctx, const aq_tensor* tA, const char* ia, const aq_tensor* tB, const char* ib) {
    Labeled A = label(tA, ia, "A"), B = label(tB, ib, "B");
    for (char c : A.lab + B.lab) label_len({&A, &B}, c);
    std::string sa = A.lab, sb = B.lab;
    std::sort(sa.begin(), sa.end());
    std::sort(sb.begin(), sb.end());
    if (sa == sb && !A.has_repeat && !B.has_repeat) {
        std::vector<std::unique_ptr<Scratch>> tmp;
        PinSet pins(ctx);
        const double* bptr = B.data;
        if (B.lab != A.lab) bptr = static_cast<const double*>(materialize(ctx, B, A.lab, FMT_F64, pins, tmp));
        launch_dot(ctx, A.data, bptr, A.size(), 0);
        return sync_result(ctx, 0);
    }
    // general label sets: contract into a scalar
    aq_tensor s;
    s.ctx = ctx;
    s.ndim = 0;
    s.size = 1;
    Scratch buf(ctx, 256);
    s.data = buf.as<double>();
    s.id = new_tensor_id();
    contract_impl(ctx, 1.0, tA, ia, tB, ib, 0.0, &s, "");
    double v = 0.0;
    AQ_HIP(hipMemcpyAsync(&v, s.data, 8, hipMemcpyDeviceToHost, ctx->stream));
    AQ_HIP(hipStreamSynchronize(ctx->stream));
    resolve_timing(ctx);
    return v;
}

}  // namespace aqamd

using namespace aqamd;

extern "C" {

int aq_tensor_create(aq_ctx* ctx, int ndim, const int64_t* len, int zero, aq_tensor** out) {
    return guarded([&] {
        AQ_CHECK(ctx && out, AQ_ERR_INVALID, "null argument");
        AQ_CHECK(ndim >= 0 && ndim <= 16, AQ_ERR_INVALID, "ndim out of range");
        AQ_CHECK(ndim == 0 || len, AQ_ERR_INVALID, "null lengths");
        auto* t = new aq_tensor();
        t->ctx = ctx;
        t->ndim = ndim;
        for (int k = 0; k < ndim; ++k) {
            if (len[k] < 0) {
                delete t;
                throw Error(AQ_ERR_INVALID, "negative length");
            }
            t->len.push_back(len[k]);
            t->size *= len[k];
        }
        try {
            t->data = tensor_alloc(ctx, size_t(t->size) * 8);
        } catch (...) {
            delete t;
            throw;
        }
        t->id = new_tensor_id();
        if (zero) AQ_HIP(hipMemsetAsync(t->data, 0, size_t(std::max<int64_t>(t->size, 1)) * 8, ctx->stream));
        *out = t;
    });
}

int aq_tensor_copy(aq_ctx* ctx, const aq_tensor* src, aq_tensor** out) {
    return guarded([&] {
        AQ_CHECK(ctx && src && out, AQ_ERR_INVALID, "null argument");
        aq_tensor* t = nullptr;
        int rc = aq_tensor_create(ctx, src->ndim, src->len.data(), 0, &t);
        if (rc) throw Error(rc, aq_last_error());
        AQ_HIP(hipMemcpyAsync(t->data, src->data, size_t(src->size) * 8, hipMemcpyDeviceToDevice, ctx->stream));
        *out = t;
    });
}

int aq_tensor_destroy(aq_tensor* t) {
    return guarded([&] {
        if (!t) return;
        aq_ctx* ctx = t->ctx;
        // drop cached copies of this tensor
        const std::string pre = std::to_string(t->id) + "|";
        for (auto it = ctx->cache.begin(); it != ctx->cache.end();) {
            if (it->first.compare(0, pre.size(), pre) == 0 && it->second.pins == 0) {
                dfree(ctx, it->second.ptr);
                ctx->cache_bytes -= it->second.bytes;
                ctx->cache_lru.erase(it->second.lru);
                it = ctx->cache.erase(it);
            } else {
                ++it;
            }
        }
        tensor_free(ctx, t->data, size_t(t->size) * 8);
        delete t;
    });
}

int aq_tensor_resize(aq_tensor* t, int ndim, const int64_t* len, int zero) {
    return guarded([&] {
        AQ_CHECK(t, AQ_ERR_INVALID, "null argument");
        AQ_CHECK(ndim >= 0 && ndim <= 16, AQ_ERR_INVALID, "ndim out of range");
        AQ_CHECK(ndim == 0 || len, AQ_ERR_INVALID, "null lengths");
        std::vector<int64_t> l;
        int64_t size = 1;
        for (int k = 0; k < ndim; ++k) {
            AQ_CHECK(len[k] >= 0, AQ_ERR_INVALID, "negative length");
            l.push_back(len[k]);
            size *= len[k];
        }
        aq_ctx* ctx = t->ctx;
        // drop cached copies of the old contents, then replace the storage
        const std::string pre = std::to_string(t->id) + "|";
        for (auto it = ctx->cache.begin(); it != ctx->cache.end();) {
            if (it->first.compare(0, pre.size(), pre) == 0 && it->second.pins == 0) {
                dfree(ctx, it->second.ptr);
                ctx->cache_bytes -= it->second.bytes;
                ctx->cache_lru.erase(it->second.lru);
                it = ctx->cache.erase(it);
            } else {
                ++it;
            }
        }
        double* fresh = tensor_alloc(ctx, size_t(size) * 8);
        tensor_free(ctx, t->data, size_t(t->size) * 8);
        t->data = fresh;
        t->ndim = ndim;
        t->len = l;
        t->size = size;
        t->id = new_tensor_id();   // a new tensor as far as any cache is concerned
        t->version = 0;
        if (zero) AQ_HIP(hipMemsetAsync(t->data, 0, size_t(std::max<int64_t>(size, 1)) * 8, ctx->stream));
    });
}

int aq_tensor_ndim(const aq_tensor* t, int* ndim) {
    return guarded([&] {
        AQ_CHECK(t && ndim, AQ_ERR_INVALID, "null argument");
        *ndim = t->ndim;
    });
}

int aq_tensor_lengths(const aq_tensor* t, int64_t* len) {
    return guarded([&] {
        AQ_CHECK(t && (len || t->ndim == 0), AQ_ERR_INVALID, "null argument");
        for (int k = 0; k < t->ndim; ++k) len[k] = t->len[k];
    });
}

int aq_tensor_size(const aq_tensor* t, int64_t* size) {
    return guarded([&] {
        AQ_CHECK(t && size, AQ_ERR_INVALID, "null argument");
        *size = t->size;
    });
}

int aq_tensor_data(const aq_tensor* t, double** dptr) {
    return guarded([&] {
        AQ_CHECK(t && dptr, AQ_ERR_INVALID, "null argument");
        *dptr = t->data;
    });
}

int aq_tensor_touch(aq_tensor* t) {
    return guarded([&] {
        AQ_CHECK(t, AQ_ERR_INVALID, "null argument");
        bump(t);
    });
}

int aq_contract(aq_ctx* ctx, double alpha, const aq_tensor* A, const char* idxA, const aq_tensor* B,
                const char* idxB, double beta, aq_tensor* C, const char* idxC) {
    return guarded([&] {
        AQ_CHECK(ctx, AQ_ERR_INVALID, "null context");
        contract_impl(ctx, alpha, A, idxA, B, idxB, beta, C, idxC);
    });
}

int aq_sum(aq_ctx* ctx, double alpha, const aq_tensor* A, const char* idxA, double beta, aq_tensor* B,
           const char* idxB) {
    return guarded([&] {
        AQ_CHECK(ctx, AQ_ERR_INVALID, "null context");
        sum_impl(ctx, alpha, A, idxA, beta, B, idxB);
    });
}

int aq_scale(aq_ctx* ctx, aq_tensor* X, double alpha, const char* idx) {
    return guarded([&] {
        AQ_CHECK(ctx && X, AQ_ERR_INVALID, "null argument");
        Labeled L = label(X, idx, "X");
        if (L.size() == 0) return;
        // elementwise in place over the addressed elements (diagonal if repeated)
        StridedPlan p = make_plan(L, L);
        launch_strided_sum(ctx, p, alpha, X->data, 0.0, X->data);
        bump(X);
    });
}

int aq_fill(aq_ctx* ctx, aq_tensor* X, double alpha, double beta) {
    return guarded([&] {
        AQ_CHECK(ctx && X, AQ_ERR_INVALID, "null argument");
        launch_fill(ctx, X->data, X->size, alpha, beta);
        bump(X);
    });
}

int aq_dot(aq_ctx* ctx, const aq_tensor* A, const char* idxA, const aq_tensor* B, const char* idxB, double* out) {
    return guarded([&] {
        AQ_CHECK(ctx && out, AQ_ERR_INVALID, "null argument");
        *out = dot_impl(ctx, A, idxA, B, idxB);
    });
}

int aq_weight(aq_ctx* ctx, aq_tensor* X, int ndim, const double* const* d, double shift) {
    return guarded([&] {
        AQ_CHECK(ctx && X, AQ_ERR_INVALID, "null argument");
        if (X->ndim == 0) return;   // ctf_tensor.cxx:484
        AQ_CHECK(ndim == X->ndim && d, AQ_ERR_INVALID, "weight needs one denominator vector per dimension");
        AQ_CHECK(X->ndim <= MAXDIM, AQ_ERR_UNSUPPORTED, "weight supports up to 8 dimensions");
        std::vector<double> host;
        std::vector<int64_t> off;
        for (int k = 0; k < ndim; ++k) {
            AQ_CHECK(d[k] || X->len[k] == 0, AQ_ERR_INVALID, "null denominator vector");
            off.push_back((int64_t)host.size());
            host.insert(host.end(), d[k], d[k] + X->len[k]);
        }
        if (X->size == 0) return;
        Scratch dd(ctx, std::max<size_t>(host.size(), 1) * 8);
        AQ_HIP(hipMemcpyAsync(dd.p, host.data(), host.size() * 8, hipMemcpyHostToDevice, ctx->stream));
        launch_weight(ctx, X->data, ndim, X->len.data(), dd.as<double>(), off.data(), shift);
        // the host vector must stay alive until the copy has been consumed
        AQ_HIP(hipStreamSynchronize(ctx->stream));
        bump(X);
    });
}

int aq_norm(aq_ctx* ctx, const aq_tensor* X, int p, double* out) {
    return guarded([&] {
        AQ_CHECK(ctx && X && out, AQ_ERR_INVALID, "null argument");
        AQ_CHECK(p == 0 || p == 1 || p == 2, AQ_ERR_INVALID, "norm order must be 0 (max), 1 or 2");
        launch_norm(ctx, X->data, X->size, p, 0);
        *out = sync_result(ctx, 0);
    });
}

int aq_read(aq_ctx* ctx, const aq_tensor* X, int64_t n, const int64_t* keys, double* vals) {
    return guarded([&] {
        AQ_CHECK(ctx && X && (n == 0 || (keys && vals)), AQ_ERR_INVALID, "null argument");
        for (int64_t i = 0; i < n; ++i)
            AQ_CHECK(keys[i] >= 0 && keys[i] < X->size, AQ_ERR_INVALID, "key out of range");
        if (n == 0) return;
        Scratch k(ctx, size_t(n) * 8), v(ctx, size_t(n) * 8);
        AQ_HIP(hipMemcpyAsync(k.p, keys, size_t(n) * 8, hipMemcpyHostToDevice, ctx->stream));
        launch_gather_keys(ctx, X->data, n, k.as<int64_t>(), v.as<double>());
        AQ_HIP(hipMemcpyAsync(vals, v.p, size_t(n) * 8, hipMemcpyDeviceToHost, ctx->stream));
        AQ_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int aq_write(aq_ctx* ctx, aq_tensor* X, int64_t n, const int64_t* keys, const double* vals, double alpha,
             double beta) {
    return guarded([&] {
        AQ_CHECK(ctx && X && (n == 0 || (keys && vals)), AQ_ERR_INVALID, "null argument");
        for (int64_t i = 0; i < n; ++i)
            AQ_CHECK(keys[i] >= 0 && keys[i] < X->size, AQ_ERR_INVALID, "key out of range");
        if (n == 0) return;
        Scratch k(ctx, size_t(n) * 8), v(ctx, size_t(n) * 8);
        AQ_HIP(hipMemcpyAsync(k.p, keys, size_t(n) * 8, hipMemcpyHostToDevice, ctx->stream));
        AQ_HIP(hipMemcpyAsync(v.p, vals, size_t(n) * 8, hipMemcpyHostToDevice, ctx->stream));
        launch_scatter_keys(ctx, X->data, n, k.as<int64_t>(), v.as<double>(), alpha, beta);
        AQ_HIP(hipStreamSynchronize(ctx->stream));
        bump(X);
    });
}

int aq_read_all(aq_ctx* ctx, const aq_tensor* X, double* host) {
    return guarded([&] {
        AQ_CHECK(ctx && X && (host || X->size == 0), AQ_ERR_INVALID, "null argument");
        if (X->size) AQ_HIP(hipMemcpyAsync(host, X->data, size_t(X->size) * 8, hipMemcpyDeviceToHost, ctx->stream));
        AQ_HIP(hipStreamSynchronize(ctx->stream));
        resolve_timing(ctx);
    });
}

int aq_write_all(aq_ctx* ctx, aq_tensor* X, const double* host) {
    return guarded([&] {
        AQ_CHECK(ctx && X && (host || X->size == 0), AQ_ERR_INVALID, "null argument");
        if (X->size) AQ_HIP(hipMemcpyAsync(X->data, host, size_t(X->size) * 8, hipMemcpyHostToDevice, ctx->stream));
        AQ_HIP(hipStreamSynchronize(ctx->stream));
        bump(X);
    });
}

int aq_div(aq_ctx* ctx, double alpha, const aq_tensor* A, const aq_tensor* B, double beta, aq_tensor* C) {
    return guarded([&] {
        AQ_CHECK(ctx && A && B && C, AQ_ERR_INVALID, "null argument");
        AQ_CHECK(A->len == C->len && B->len == C->len, AQ_ERR_INVALID, "div needs tensors of equal shape");
        launch_div(ctx, A->data, B->data, C->data, C->size, alpha, beta);
        bump(C);
    });
}

int aq_invert(aq_ctx* ctx, double alpha, const aq_tensor* A, double beta, aq_tensor* C) {
    return guarded([&] {
        AQ_CHECK(ctx && A && C, AQ_ERR_INVALID, "null argument");
        AQ_CHECK(A->len == C->len, AQ_ERR_INVALID, "invert needs tensors of equal shape");
        launch_div(ctx, nullptr, A->data, C->data, C->size, alpha, beta);
        bump(C);
    });
}

int aq_slice(aq_ctx* ctx, double alpha, const aq_tensor* A, const int64_t* startA, double beta, aq_tensor* B,
             const int64_t* startB, const int64_t* extent) {
    return guarded([&] {
        AQ_CHECK(ctx && A && B, AQ_ERR_INVALID, "null argument");
        AQ_CHECK(A->ndim == B->ndim, AQ_ERR_INVALID, "slice needs tensors of equal ndim");
        AQ_CHECK(A->ndim <= MAXDIM, AQ_ERR_UNSUPPORTED, "slice supports up to 8 dimensions");
        const int nd = A->ndim;
        int64_t sa = 1, sb = 1, offa = 0, offb = 0;
        StridedPlan p;
        p.nd = nd;
        for (int k = 0; k < nd; ++k) {
            AQ_CHECK(startA[k] >= 0 && extent[k] >= 0 && startA[k] + extent[k] <= A->len[k], AQ_ERR_INVALID,
                     "slice box outside A");
            AQ_CHECK(startB[k] >= 0 && startB[k] + extent[k] <= B->len[k], AQ_ERR_INVALID, "slice box outside B");
            p.len[k] = extent[k];
            p.ssrc[k] = sa;
            p.sdst[k] = sb;
            p.nout *= extent[k];
            offa += startA[k] * sa;
            offb += startB[k] * sb;
            sa *= A->len[k];
            sb *= B->len[k];
        }
        if (p.nout == 0) return;
        if (overlaps(A, B)) {
            Scratch snap(ctx, size_t(A->size) * 8);
            AQ_HIP(hipMemcpyAsync(snap.p, A->data, size_t(A->size) * 8, hipMemcpyDeviceToDevice, ctx->stream));
            launch_strided_sum(ctx, p, alpha, snap.as<double>() + offa, beta, B->data + offb);
        } else {
            launch_strided_sum(ctx, p, alpha, A->data + offa, beta, B->data + offb);
        }
        bump(B);
    });
}

}  // extern "C"
