// Closed-shell CCSD amplitude engine: RHFCCSD setup and iterate, DIIS and the
// fixed-point driver, issued through the tensor operator surface
// (include/aqamd.hpp) exactly as the reference issues them through CTFTensor.
//
//   setup      src/cc/rhfccsd.cxx:45-98
//   iterate    src/cc/rhfccsd.cxx:194-327
//   DIIS       src/convergence/diis.hpp:36-227
//   driver     src/util/iterative.hpp:126-188
//   D          src/operator/denominator.hpp:21-89
//
// Beyond the reference: the amplitude update + energy + convergence can run
// as one fused kernel pair, DIIS inner products and extrapolations are fused
// multi-operand kernels over device-resident history, and the contraction
// precision can be reduced until the residual drops below a threshold, after
// which the remaining iterations run in FP64 (iterative refinement).
#include "internal.hpp"

#include "../../include/aqamd.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <limits>
#include <memory>

using aqamd::tensor::Tensor;
using aqamd::tensor::scalar;

#include "diis.hpp"


using namespace aqamd;

struct aq_rhfccsd {
    aq_ctx* ctx = nullptr;
    aq_rhfccsd_config cfg{};
    int o = 0, v = 0;
    Tensor fAI, fAB, fIJ;
    Tensor VABCD, VABCI, VABIJ, VAIBJ, VAIJB, VAIJK, VIJKL;
    Tensor T1, T2, Z1, Z2, Tau, T2SA, TauSA, VABIJSA, VAIJBSA;
    Tensor FAE, FMI, FME, WMNIJ, WMNEJ, WAMIJ, WAMEI, WAMIE;
    Tensor WABEJ;                       // formed by hbar(true) only
    // packed ladder: V+- (constant), T+- and S+- per iteration
    bool packed = false;
    Tensor Vp, Vm, Tp, Tm, Sp, Sm;
    std::vector<double> dA, dI;
    Tensor d_dA, d_dI;
    std::unique_ptr<DIIS> diis;
    double mp2 = 0.0, energy = 0.0, conv = std::numeric_limits<double>::max();
    int iter = 0;
    int last_prec = AQ_PREC_FP64;

    // Stage of the precision policy reached so far: 0 = the bulk precision,
    // 1 = the optional FP32 stage, 2 = FP64 refinement.  Stages only advance:
    // once the residual has dropped below a threshold the remaining iterations
    // stay at that stage's precision even if DIIS lets conv rise above it again.
    int policy_stage = 0;

    int stage_for(double c) const {
        if (cfg.refine_below > 0.0 && c < cfg.refine_below) {
            // optional FP32 stage between the reduced phase and FP64 refinement
            if (cfg.fp32_stage_below > 0.0 && c >= cfg.fp32_stage_below && cfg.precision != AQ_PREC_FP32) return 1;
            return 2;
        }
        return 0;
    }

    int next_precision() const {
        if (cfg.precision == AQ_PREC_FP64) return AQ_PREC_FP64;
        const int st = std::max(policy_stage, stage_for(conv));
        return st == 2 ? AQ_PREC_FP64 : st == 1 ? AQ_PREC_FP32 : cfg.precision;
    }

    // Versions of the borrowed inputs that setup derived quantities from:
    // the denominators (fAB, fIJ), the spin-adapted blocks (VABIJ, VAIJB,
    // VAIBJ) and the packed ladder operators (VABCD, VABIJ).  A caller that
    // rewrites an input and calls aq_tensor_touch bumps its version; the next
    // iteration (or Hbar) rebuilds what was derived from it.  Inputs that the
    // contractions read directly need nothing more: cached operand copies are
    // keyed by the version.
    uint64_t v_fAB = 0, v_fIJ = 0, v_VABIJ = 0, v_VAIJB = 0, v_VAIBJ = 0, v_VABCD = 0;
    static uint64_t ver(const Tensor& t) { return t.handle()->version; }

    void note_input_versions() {
        v_fAB = ver(fAB);
        v_fIJ = ver(fIJ);
        v_VABIJ = ver(VABIJ);
        v_VAIJB = ver(VAIJB);
        v_VAIBJ = ver(VAIBJ);
        v_VABCD = ver(VABCD);
    }

    void refresh_inputs() {
        const bool f = ver(fAB) != v_fAB || ver(fIJ) != v_fIJ;
        const bool sa = ver(VABIJ) != v_VABIJ || ver(VAIJB) != v_VAIJB || ver(VAIBJ) != v_VAIBJ;
        const bool pk = ver(VABCD) != v_VABCD || ver(VABIJ) != v_VABIJ;
        if (!f && !sa && !pk) return;
        tensor::check(aq_ctx_set_precision(ctx, AQ_PREC_FP64, 0.0));
        if (f) derive_denominators();
        if (sa) derive_sa_blocks();
        if (pk) derive_packed();
        note_input_versions();
    }

    // Denominator: dA = -f_aa, dI = +f_ii (denominator.hpp:48-67)
    void derive_denominators() {
        std::vector<int64_t> ka(v), ki(o);
        for (int a = 0; a < v; ++a) ka[a] = int64_t(a) + int64_t(a) * v;
        for (int i = 0; i < o; ++i) ki[i] = int64_t(i) + int64_t(i) * o;
        dA.assign(v, 0.0);
        dI.assign(o, 0.0);
        tensor::check(aq_read(ctx, fAB.handle(), v, ka.data(), dA.data()));
        tensor::check(aq_read(ctx, fIJ.handle(), o, ki.data(), dI.data()));
        for (auto& x : dA) x = -x;
        if (!d_dA.handle()) {
            d_dA = Tensor(ctx, std::vector<int64_t>{v}, true);
            d_dI = Tensor(ctx, std::vector<int64_t>{o}, true);
        }
        d_dA.setAllData(dA);
        d_dI.setAllData(dI);
    }

    void derive_sa_blocks() {
        VABIJSA["abij"] = 2 * VABIJ["abij"];           // :88
        VABIJSA["abij"] -= VABIJ["abji"];              // :89

        VAIJBSA["aijb"] = 2 * VAIJB["aijb"];           // :91
        VAIJBSA["aijb"] -= VAIBJ["aibj"];              // :92
    }

    // The ladder can run on the (anti)symmetric packed forms when
    // <ab|ef> = <ba|fe> and <ab|ij> = <ba|ji> (so Tau(ef,ij) = Tau(fe,ji)).
    void derive_packed() {
        packed = false;
        if (cfg.packed_ladder && o >= 2 && v >= 2) {
            const double dV = pair_asymmetry(ctx, dptr(VABCD), v, v, 0);
            const double dI = pair_asymmetry(ctx, dptr(VABIJ), v, o, 0);
            const double nV = VABCD.norm(0), nI = VABIJ.norm(0);
            packed = dV <= 1e-13 * (1.0 + nV) && dI <= 1e-13 * (1.0 + nI);
        }
        if (!packed) {
            Vp = Tensor();
            Vm = Tensor();
            Tp = Tensor();
            Tm = Tensor();
            Sp = Tensor();
            Sm = Tensor();
            return;
        }
        const int64_t P = int64_t(v) * (v + 1) / 2, Q = int64_t(v) * (v - 1) / 2;
        const int64_t Op = int64_t(o) * (o + 1) / 2, Om = int64_t(o) * (o - 1) / 2;
        auto mk = [&](std::initializer_list<int64_t> l) { return Tensor(ctx, std::vector<int64_t>(l), true); };
        if (!Vp.handle()) {
            Vp = mk({P, P});
            Vm = mk({Q, Q});
            Tp = mk({Op, P});   // occupied pair fastest: the ladder GEMM reads T row-contiguous
            Tm = mk({Om, Q});
            Sp = mk({P, Op});
            Sm = mk({Q, Om});
        }
        launch_ladder_pack_v(ctx, dptr(VABCD), v, dptr(Vp), dptr(Vm));
        bump_t(Vp);
        bump_t(Vm);
    }

    // rhfccsd.cxx:45-98
    void setup() {
        const int64_t vo[2] = {v, o}, vvoo[4] = {v, v, o, o}, voov[4] = {v, o, o, v};
        auto mk = [&](std::initializer_list<int64_t> l) { return Tensor(ctx, std::vector<int64_t>(l), true); };
        (void)vo;
        (void)vvoo;
        (void)voov;
        T1 = mk({v, o});
        T2 = mk({v, v, o, o});
        Z1 = mk({v, o});
        Z2 = mk({v, v, o, o});
        Tau = mk({v, v, o, o});
        T2SA = mk({v, v, o, o});
        TauSA = mk({v, v, o, o});
        VABIJSA = mk({v, v, o, o});
        VAIJBSA = mk({v, o, o, v});
        FAE = mk({v, v});
        FMI = mk({o, o});
        FME = mk({o, v});
        WMNIJ = mk({o, o, o, o});
        WMNEJ = mk({o, o, v, o});
        WAMIJ = mk({v, o, o, o});
        WAMEI = mk({v, o, v, o});
        WAMIE = mk({v, o, o, v});

        derive_denominators();

        T1 = fAI;                                      // :79
        T1.weight({&dA, &dI});                         // :80
        T2 = VABIJ;                                    // :82
        T2.weight({&dA, &dA, &dI, &dI});               // :83

        Tau["abij"] = T2["abij"];                      // :85
        Tau["abij"] += T1["ai"] * T1["bj"];            // :86

        derive_sa_blocks();                            // :88-92

        mp2 = 2 * scalar(fAI["ai"] * T1["ai"]) + scalar(VABIJSA["abij"] * Tau["abij"]);   // :94

        derive_packed();
        note_input_versions();
        diis.reset(new DIIS(ctx, cfg.diis_order, cfg.diis_start, cfg.diis_damping, 2, 2));   // :98
    }

    // The contraction stage of one iteration (rhfccsd.cxx:231-309): every
    // operation is stream-ordered and none synchronises with the host.
    void contractions() {
        // :231-238 in one pass over T2, rounded as the six operator calls
        //   Tau["abij"] = T2["abij"];  Tau["abij"] += T1["ai"] * T1["bj"];
        //   TauSA["abij"] = 2 * Tau["abij"];  TauSA["abij"] -= Tau["abji"];
        //   T2SA["abij"] = 2 * T2["abij"];  T2SA["abij"] -= T2["abji"];
        launch_tau_sa(ctx, dptr(T1), dptr(T2), v, o, dptr(Tau), dptr(TauSA), dptr(T2SA));
        bump_t(Tau);
        bump_t(TauSA);
        bump_t(T2SA);

        FME["me"] = fAI["em"];                                     // :244
        FME["me"] += VABIJSA["efmn"] * T1["fn"];                   // :245

        FMI["mi"] = fIJ["mi"];                                     // :247
        FMI["mi"] += VABIJSA["efmn"] * T2["efin"];                 // :248
        FMI["mi"] += FME["me"] * T1["ei"];                         // :249
        FMI["mi"] += 2 * VAIJK["finm"] * T1["fn"];                 // :250
        FMI["mi"] -= VAIJK["fimn"] * T1["fn"];                     // :251

        FAE["ae"] = fAB["ae"];                                     // :253
        FAE["ae"] -= VABIJSA["efmn"] * T2["afmn"];                 // :254
        FAE["ae"] -= FME["me"] * T1["am"];                         // :255
        FAE["ae"] += 2 * VABCI["efam"] * T1["fm"];                 // :256
        FAE["ae"] -= VABCI["feam"] * T1["fm"];                     // :257

        WMNIJ["mnij"] = 0.5 * VIJKL["mnij"];                       // :259
        WMNIJ["mnij"] += 0.5 * VABIJ["efmn"] * Tau["efij"];        // :260
        WMNIJ["mnij"] += VAIJK["ejmn"] * T1["ei"];                 // :261
        WMNIJ["mnij"] += WMNIJ["nmji"];                            // :262 (aliased)

        WMNEJ["mnej"] = VAIJK["ejmn"];                             // :264
        WMNEJ["mnej"] += VABIJ["efmn"] * T1["fj"];                 // :265

        WAMIJ["amij"] = VAIJK["amij"];                             // :267
        WAMIJ["amij"] += VABCI["efam"] * Tau["efij"];              // :268
        WAMIJ["amij"] += VAIBJ["amej"] * T1["ei"];                 // :269
        WAMIJ["amij"] += VAIJB["amie"] * T1["ej"];                 // :270

        WAMIE["amie"] = VAIJBSA["amie"];                           // :272
        WAMIE["amie"] += 0.5 * VABIJSA["efmn"] * T2SA["afin"];     // :273
        WAMIE["amie"] += 2 * VABCI["feam"] * T1["fi"];             // :274
        WAMIE["amie"] -= VABCI["efam"] * T1["fi"];                 // :275
        WAMIE["amie"] -= 2 * WMNEJ["mnei"] * T1["an"];             // :276
        WAMIE["amie"] += WMNEJ["nmei"] * T1["an"];                 // :277

        WAMEI["amei"] = VAIBJ["amei"];                             // :279
        WAMEI["amei"] -= 0.5 * VABIJ["efnm"] * T2["afni"];         // :280
        WAMEI["amei"] += VABCI["efam"] * T1["fi"];                 // :281
        WAMEI["amei"] -= WMNEJ["nmei"] * T1["an"];                 // :282

        Z1["ai"] = fAI["ai"];                                      // :290
        Z1["ai"] += fAB["ae"] * T1["ei"];                          // :291
        Z1["ai"] -= FMI["mi"] * T1["am"];                          // :292
        Z1["ai"] += VAIJBSA["amie"] * T1["em"];                    // :293
        Z1["ai"] += FME["me"] * T2SA["aeim"];                      // :294
        Z1["ai"] += VABCI["efam"] * TauSA["efim"];                 // :295
        Z1["ai"] -= WMNEJ["mnei"] * T2SA["eamn"];                  // :296

        Z2["abij"] = 0.5 * VABIJ["abij"];                          // :298
        Z2["abij"] += VABCI["abej"] * T1["ei"];                    // :299
        Z2["abij"] -= WAMIJ["amij"] * T1["bm"];                    // :300
        Z2["abij"] += FAE["ae"] * T2["ebij"];                      // :301
        Z2["abij"] -= FMI["mi"] * T2["abmj"];                      // :302
        if (packed) {                                              // :303  particle-particle ladder,
            launch_ladder_pack_t(ctx, dptr(Tau), v, o, dptr(Tp), dptr(Tm));   // as S+ + S- on packed pairs
            bump_t(Tp);
            bump_t(Tm);
            Sp["pi"] = Vp["pq"] * Tp["iq"];
            Sm["pi"] = Vm["pq"] * Tm["iq"];
            launch_ladder_unpack(ctx, dptr(Sp), dptr(Sm), v, o, 0.5, dptr(Z2));
            bump_t(Z2);
        } else {
            Z2["abij"] += 0.5 * VABCD["abef"] * Tau["efij"];       // :303  particle-particle ladder
        }
        Z2["abij"] += 0.5 * WMNIJ["mnij"] * Tau["abmn"];           // :304
        Z2["abij"] += 0.5 * WAMIE["amie"] * T2SA["ebmj"];          // :305  ring
        Z2["abij"] -= 0.5 * WAMEI["amei"] * T2["ebjm"];            // :306  ring
        Z2["abij"] -= WAMEI["amej"] * T2["ebim"];                  // :307  ring

        Z2["abij"] += Z2["baji"];                                  // :309 (aliased)
    }

    // HIP graph of the contraction stage.  Captured when the previous stage
    // ran eagerly at the same precision (so every scratch buffer is in the
    // pool and every operand copy exists), replayed while the precision, the
    // context's memory generation and the versions of the copies it reads
    // without refreshing are unchanged.  Refreshing copies of amplitudes and
    // intermediates are kernels inside the graph, so a replay recomputes them
    // from the current data, exactly as the eager stage would.
    struct StageGraph {
        hipGraphExec_t exec = nullptr;
        int prec = -1;
        uint64_t gen = 0;
        std::vector<std::pair<const aq_tensor*, uint64_t>> fresh;
        int warm_prec = -1;     // precision of the last eager stage
        int captures = 0, replays = 0, failed = 0;
    } graph;

    void drop_graph() {
        if (!graph.exec) return;
        (void)hipStreamSynchronize(ctx->stream);   // the last replay may still run
        (void)hipGraphExecDestroy(graph.exec);
        graph.exec = nullptr;
        graph.fresh.clear();
        graph.warm_prec = -1;   // warm the pool and the copies eagerly before the next capture
    }

    bool graph_valid(int prec) const {
        if (!graph.exec || graph.prec != prec || graph.gen != ctx->mem_gen) return false;
        for (const auto& f : graph.fresh)
            if (f.first->version != f.second) return false;
        return true;
    }

    // Capture the stage and launch the graph; false (nothing executed, every
    // operand copy marked stale) when the capture had to allocate or failed.
    bool capture_stage(int prec) {
        CaptureLog log;
        const uint64_t gen0 = ctx->mem_gen;
        if (hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            (void)hipGetLastError();   // e.g. the caller is capturing this stream already
            graph.failed += 1;
            return false;
        }
        ctx->capture = &log;
        bool ok = true;
        try {
            contractions();
        } catch (...) {
            ok = false;
        }
        ctx->capture = nullptr;
        hipGraph_t g = nullptr;
        if (hipStreamEndCapture(ctx->stream, &g) != hipSuccess || !g) ok = false;
        if (ok && ctx->mem_gen != gen0) ok = false;
        for (const auto& f : log.fresh_reads)   // written after it was read unrefreshed
            if (f.first->version != f.second) ok = false;
        hipGraphExec_t exec = nullptr;
        if (ok && hipGraphInstantiate(&exec, g, nullptr, nullptr, 0) != hipSuccess) ok = false;
        if (g) (void)hipGraphDestroy(g);
        if (!ok) {
            (void)hipGetLastError();
            if (exec) (void)hipGraphExecDestroy(exec);
            // the capture marked copies fresh whose conversions never ran
            invalidate_cache(ctx);
            graph.failed += 1;
            return false;
        }
        AQ_HIP(hipGraphLaunch(exec, ctx->stream));
        graph.exec = exec;
        graph.prec = prec;
        graph.gen = ctx->mem_gen;
        graph.fresh = std::move(log.fresh_reads);
        graph.captures += 1;
        return true;
    }

    void stage(int prec) {
        const bool want = cfg.graph != 0 && !ctx->timing && graph.failed < 2;
        if (graph.exec && !(want && graph_valid(prec))) drop_graph();
        if (graph.exec) {
            AQ_HIP(hipGraphLaunch(graph.exec, ctx->stream));
            graph.replays += 1;
            return;
        }
        if (want && graph.warm_prec == prec && capture_stage(prec)) return;
        {
            StageMark mark(ctx->batcher);
            contractions();
        }
        graph.warm_prec = prec;
    }

    ~aq_rhfccsd() { drop_graph(); }

    // rhfccsd.cxx:194-327
    void iterate() {
        refresh_inputs();
        const int prec = next_precision();
        if (cfg.precision != AQ_PREC_FP64) policy_stage = std::max(policy_stage, stage_for(conv));
        if (cfg.diis_reset_on_refine && iter > 0 && prec == AQ_PREC_FP64 && last_prec != AQ_PREC_FP64)
            diis.reset(new DIIS(ctx, cfg.diis_order, cfg.diis_start, cfg.diis_damping, 2, 2));
        last_prec = prec;
        tensor::check(aq_ctx_set_precision(ctx, prec, cfg.min_reduced_flops));

        stage(prec);                                               // :231-309

        // the residual, energy and convergence are FP64 whatever the contraction precision
        tensor::check(aq_ctx_set_precision(ctx, AQ_PREC_FP64, 0.0));
        if (cfg.fused_update) {
            UpdateArgs u{o, v, dptr(Z1), dptr(Z2), dptr(T1), dptr(T2), dptr(Tau), dptr(VABIJSA), dptr(fAI),
                         dptr(d_dA), dptr(d_dI)};
            launch_ccsd_update(ctx, u, 0, 1, 2, 3);                // :313-323
            AQ_HIP(hipStreamSynchronize(ctx->stream));
            resolve_timing(ctx);
            for (const Tensor* t : {&Z1, &Z2, &T1, &T2, &Tau}) bump_t(*t);
            energy = 2 * ctx->h_result[0] + ctx->h_result[1];
            conv = std::max(ctx->h_result[2], ctx->h_result[3]);
        } else {
            Z1.weight({&dA, &dI});                                 // :313
            Z2.weight({&dA, &dA, &dI, &dI});                       // :314
            T1 += Z1;                                              // :316
            T2 += Z2;                                              // :317
            Tau["abij"] = T2["abij"];                              // :319
            Tau["abij"] += T1["ai"] * T1["bj"];                    // :320
            energy = 2 * scalar(fAI["ai"] * T1["ai"]) + scalar(VABIJSA["abij"] * Tau["abij"]);   // :322
            conv = std::max(Z1.norm(0), Z2.norm(0));               // :323
        }
        diis->extrapolate({&T1, &T2}, {&Z1, &Z2});                 // :325-326
        ++iter;
    }

    // Hbar products of the current amplitudes (rhfccsd.cxx:108-188), in FP64.
    void hbar(bool with_wabej) {
        refresh_inputs();
        tensor::check(aq_ctx_set_precision(ctx, AQ_PREC_FP64, 0.0));
        Tau["abij"] = T2["abij"];                                  // :119
        Tau["abij"] += T1["ai"] * T1["bj"];                        // :120
        T2SA["abij"] = 2 * T2["abij"];                             // :122
        T2SA["abij"] -= T2["abji"];                                // :123

        FME["me"] = fAI["em"];                                     // :125
        FME["me"] += VABIJSA["efmn"] * T1["fn"];                   // :126

        FMI["mi"] = fIJ["mi"];                                     // :128
        FMI["mi"] += VABIJSA["efmn"] * T2["efin"];                 // :129
        FMI["mi"] += FME["me"] * T1["ei"];                         // :130
        FMI["mi"] += 2 * VAIJK["finm"] * T1["fn"];                 // :131
        FMI["mi"] -= VAIJK["fimn"] * T1["fn"];                     // :132

        FAE["ae"] = fAB["ae"];                                     // :134
        FAE["ae"] -= VABIJSA["efmn"] * T2["afmn"];                 // :135
        FAE["ae"] -= FME["me"] * T1["am"];                         // :136
        FAE["ae"] += 2 * VABCI["efam"] * T1["fm"];                 // :137
        FAE["ae"] -= VABCI["feam"] * T1["fm"];                     // :138

        WMNIJ["mnij"] = 0.5 * VIJKL["mnij"];                       // :140
        WMNIJ["mnij"] += 0.5 * VABIJ["efmn"] * Tau["efij"];        // :141
        WMNIJ["mnij"] += VAIJK["ejmn"] * T1["ei"];                 // :142
        WMNIJ["mnij"] += WMNIJ["nmji"];                            // :143 (aliased)

        WMNEJ["mnej"] = VAIJK["ejmn"];                             // :145
        WMNEJ["mnej"] += VABIJ["efmn"] * T1["fj"];                 // :146

        WAMIJ["amij"] = VAIJK["amij"];                             // :148
        WAMIJ["amij"] += VABCI["efam"] * Tau["efij"];              // :149
        WAMIJ["amij"] += VAIBJ["amej"] * T1["ei"];                 // :150
        WAMIJ["amij"] += VAIJB["amie"] * T1["ej"];                 // :151
        WAMIJ["amij"] -= WMNIJ["nmij"] * T1["an"];                 // :152
        WAMIJ["amij"] += WMNEJ["nmej"] * T2SA["aein"];             // :153
        WAMIJ["amij"] -= WMNEJ["mnej"] * T2["aein"];               // :154
        WAMIJ["amij"] -= WMNEJ["mnei"] * T2["aenj"];               // :155
        WAMIJ["amij"] += FME["me"] * T2["aeij"];                   // :156

        WAMIE["amie"] = VAIJBSA["amie"];                           // :158
        WAMIE["amie"] += VABIJSA["efmn"] * T2SA["afin"];           // :159
        WAMIE["amie"] += 2 * VABCI["feam"] * T1["fi"];             // :160
        WAMIE["amie"] -= VABCI["efam"] * T1["fi"];                 // :161

        WAMEI["amei"] = VAIBJ["amei"];                             // :163
        WAMEI["amei"] -= VABIJ["efnm"] * T2["afni"];               // :164
        WAMEI["amei"] += VABCI["efam"] * T1["fi"];                 // :165

        if (with_wabej) {                                          // :167-181
            if (!WABEJ.handle()) WABEJ = Tensor(ctx, std::vector<int64_t>{v, v, v, o}, true);
            WABEJ["abej"] = VABCI["abej"];
            WABEJ["abej"] += WMNEJ["mnej"] * Tau["abmn"];
            WABEJ["abej"] -= 0.5 * WAMIE["bmje"] * T1["am"];
            WABEJ["abej"] -= 0.5 * WAMEI["bmej"] * T1["am"];
            WABEJ["abej"] -= WAMEI["amej"] * T1["bm"];
            WABEJ["abej"] += VABCD["abef"] * T1["fj"];
            WABEJ["abej"] += VABCI["efam"] * T2SA["fbmj"];
            WABEJ["abej"] -= VABCI["feam"] * T2["fbmj"];
            WABEJ["abej"] -= VABCI["febm"] * T2["afmj"];
            WABEJ["abej"] -= FME["me"] * T2["abmj"];
        }

        WAMIE["amie"] -= 2 * WMNEJ["mnei"] * T1["an"];             // :183
        WAMIE["amie"] += WMNEJ["nmei"] * T1["an"];                 // :184

        WAMEI["amei"] -= WMNEJ["nmei"] * T1["an"];                 // :186
    }
};

extern "C" {

int aq_rhfccsd_default_config(aq_rhfccsd_config* c) {
    return guarded([&] {
        AQ_CHECK(c, AQ_ERR_INVALID, "null config");
        c->convergence = 1e-9;
        c->max_iterations = 50;
        c->diis_order = 5;
        c->diis_start = 1;
        c->diis_damping = 0.0;
        c->precision = AQ_PREC_FP64;
        c->refine_below = 1e-5;
        c->min_reduced_flops = 1e8;
        c->fused_update = 1;
        c->diis_reset_on_refine = 1;
        c->fp32_stage_below = 0.0;
        c->graph = 0;
        c->packed_ladder = 1;
    });
}

int aq_rhfccsd_create(aq_ctx* ctx, const aq_tensor* const* f3, const aq_tensor* const* v7,
                      const aq_rhfccsd_config* cfg, aq_rhfccsd** out) {
    return guarded([&] {
        AQ_CHECK(ctx && f3 && v7 && out, AQ_ERR_INVALID, "null argument");
        for (int k = 0; k < 3; ++k) AQ_CHECK(f3[k], AQ_ERR_INVALID, "null Fock block");
        for (int k = 0; k < 7; ++k) AQ_CHECK(v7[k], AQ_ERR_INVALID, "null integral block");
        AQ_CHECK(f3[0]->ndim == 2, AQ_ERR_INVALID, "fAI must be [v,o]");
        const int64_t v = f3[0]->len[0], o = f3[0]->len[1];
        auto expect = [&](const aq_tensor* t, std::vector<int64_t> l, const char* nm) {
            AQ_CHECK(t->len == l, AQ_ERR_INVALID, std::string("invalid ndim/lengths of ") + nm);
        };
        expect(f3[1], {v, v}, "fAB");
        expect(f3[2], {o, o}, "fIJ");
        expect(v7[0], {v, v, v, v}, "VABCD");
        expect(v7[1], {v, v, v, o}, "VABCI");
        expect(v7[2], {v, v, o, o}, "VABIJ");
        expect(v7[3], {v, o, v, o}, "VAIBJ");
        expect(v7[4], {v, o, o, v}, "VAIJB");
        expect(v7[5], {v, o, o, o}, "VAIJK");
        expect(v7[6], {o, o, o, o}, "VIJKL");
        std::unique_ptr<aq_rhfccsd> cc(new aq_rhfccsd());
        cc->ctx = ctx;
        if (cfg) cc->cfg = *cfg;
        else aq_rhfccsd_default_config(&cc->cfg);
        AQ_CHECK(cc->cfg.diis_order >= 0, AQ_ERR_INVALID, "DIIS order must be >= 0");
        cc->o = int(o);
        cc->v = int(v);
        auto b = [&](const aq_tensor* t) { return Tensor::borrow(ctx, const_cast<aq_tensor*>(t)); };
        cc->fAI = b(f3[0]);
        cc->fAB = b(f3[1]);
        cc->fIJ = b(f3[2]);
        cc->VABCD = b(v7[0]);
        cc->VABCI = b(v7[1]);
        cc->VABIJ = b(v7[2]);
        cc->VAIBJ = b(v7[3]);
        cc->VAIJB = b(v7[4]);
        cc->VAIJK = b(v7[5]);
        cc->VIJKL = b(v7[6]);
        const int prev = ctx->precision;
        const double prevf = ctx->min_reduced_flops;
        ctx->precision = AQ_PREC_FP64;
        cc->setup();
        ctx->precision = prev;
        ctx->min_reduced_flops = prevf;
        *out = cc.release();
    });
}

int aq_rhfccsd_destroy(aq_rhfccsd* cc) {
    return guarded([&] { delete cc; });
}

int aq_rhfccsd_mp2(const aq_rhfccsd* cc, double* mp2) {
    return guarded([&] {
        AQ_CHECK(cc && mp2, AQ_ERR_INVALID, "null argument");
        *mp2 = cc->mp2;
    });
}

int aq_rhfccsd_iterate(aq_rhfccsd* cc, double* energy, double* conv) {
    return guarded([&] {
        AQ_CHECK(cc, AQ_ERR_INVALID, "null solver");
        cc->iterate();
        if (energy) *energy = cc->energy;
        if (conv) *conv = cc->conv;
    });
}

int aq_rhfccsd_run(aq_rhfccsd* cc, aq_iter_callback cb, void* user, double* energy, double* conv, int* iterations,
                   int* converged) {
    return guarded([&] {
        AQ_CHECK(cc, AQ_ERR_INVALID, "null solver");
        cc->conv = std::numeric_limits<double>::max();   // iterative.hpp:130
        int it = 1;
        for (; it <= cc->cfg.max_iterations && !(cc->conv < cc->cfg.convergence); ++it) {
            const int prec = cc->next_precision();
            auto t0 = std::chrono::steady_clock::now();
            cc->iterate();
            const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (cb) cb(it, cc->energy, cc->conv, dt, prec, user);
        }
        if (energy) *energy = cc->energy;
        if (conv) *conv = cc->conv;
        if (iterations) *iterations = it - 1;
        if (converged) *converged = cc->conv < cc->cfg.convergence ? 1 : 0;
    });
}

int aq_rhfccsd_hbar(aq_rhfccsd* cc, int with_wabej) {
    return guarded([&] {
        AQ_CHECK(cc, AQ_ERR_INVALID, "null solver");
        cc->hbar(with_wabej != 0);
        AQ_HIP(hipStreamSynchronize(cc->ctx->stream));
    });
}

int aq_rhfccsd_tensor(const aq_rhfccsd* cc, int which, const aq_tensor** out) {
    return guarded([&] {
        AQ_CHECK(cc && out, AQ_ERR_INVALID, "null argument");
        const Tensor* t = nullptr;
        switch (which) {
            case 0: t = &cc->T1; break;
            case 1: t = &cc->T2; break;
            case 2: t = &cc->Z1; break;
            case 3: t = &cc->Z2; break;
            case 4: t = &cc->Tau; break;
            case 10: t = &cc->FAE; break;
            case 11: t = &cc->FMI; break;
            case 12: t = &cc->FME; break;
            case 13: t = &cc->WMNIJ; break;
            case 14: t = &cc->WMNEJ; break;
            case 15: t = &cc->WAMIJ; break;
            case 16: t = &cc->WAMEI; break;
            case 17: t = &cc->WAMIE; break;
            case 18: t = &cc->WABEJ; break;
            default: throw Error(AQ_ERR_INVALID, "unknown tensor selector");
        }
        AQ_CHECK(t->handle(), AQ_ERR_INVALID, "WABEJ is formed by aq_rhfccsd_hbar(cc, 1)");
        *out = t->handle();
    });
}

int aq_rhfccsd_current_precision(const aq_rhfccsd* cc, int* precision) {
    return guarded([&] {
        AQ_CHECK(cc && precision, AQ_ERR_INVALID, "null argument");
        *precision = cc->next_precision();
    });
}

int aq_rhfccsd_packed_ladder(const aq_rhfccsd* cc, int* packed) {
    return guarded([&] {
        AQ_CHECK(cc && packed, AQ_ERR_INVALID, "null argument");
        *packed = cc->packed ? 1 : 0;
    });
}

int aq_rhfccsd_graph_stats(const aq_rhfccsd* cc, int* captures, int* replays, int* failed) {
    return guarded([&] {
        AQ_CHECK(cc, AQ_ERR_INVALID, "null solver");
        if (captures) *captures = cc->graph.captures;
        if (replays) *replays = cc->graph.replays;
        if (failed) *failed = cc->graph.failed;
    });
}

}  // extern "C"
