"""Benchmark: closed-shell CCSD iterations/s on one MI355X, mixed precision vs FP64.

Workload: the seeded synthetic closed-shell system o=48, v=288, nQ=1008
(eps_occ ~ U[-1.5,-0.3], eps_vir ~ U[0.2,5.0], (pq|rs) = sum_Q B^Q_pq B^Q_rs
with the size-normalised factor scale), integrals and amplitudes resident in
HBM (the full 55 GB FP64 vvvv block included).

A "step" is one CCSD iteration (RHFCCSD::iterate: intermediates, residual,
update, energy, DIIS).  After W untimed iterations in each arithmetic on a
throwaway solver (GEMM tuner trials, operand copies of the integrals), the
mixed solve (BF16x3 MFMA contractions with FP32 accumulation until the
residual drops below 1e-5, FP64 refinement afterwards) runs to convergence,
then the FP64 solve (the reference arithmetic) runs to convergence.  `value`
is the mixed solve's iterations/s over every iteration of that solve, FP64
refinement included, because the metric is defined at the converged energy;
`steps` is the number of iterations timed.  The BF16x3 phase's own rate is a
separate field.

The same run then solves the CPU leg's system (o=20, v=140) in FP64 on the
GPU, serves the 256-job stream (continuous batching; jobs/h, p50/p99
time-to-E and service time, one-at-a-time FP64 baseline, per-job |dE|) and
times the oracle on the host's cores on that o=20, v=140 system.

Prints one JSON line.  `--gpus N` with N > 1 exits: this engine runs on one
GPU (sharding over GPUs would split Z2's output pair index; it is not built).
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "CCSD iter/s mixed vs FP64 at |ΔE|≤1e-6 Eh; served jobs/h + p99 time-to-E"
PEAK = {"fp64": 78.6e12, "fp32": 157.3e12, "bf16x3": 2.5e15}

# HBM bytes per launch of a kernel class, from two rocprofv3 --pmc runs (FETCH_SIZE, then WRITE_SIZE)
# of `bench.py --only fp64 --steps 1 --warmup 1`, scaled by the same counters on a plain 1 GiB copy
# (fetch x1.9999, write x1.0); mean over the class's launches in the timed iteration.  Keyed by
# (timing class, o, v); the per-launch tables and the commands are in the profile summaries
# gemm_f64_128x128_NT_o48v288_hbm_traffic.txt and bf16x3_iteration_o48v288.txt
# (the BF16x3 class from `bench.py --only mixed --steps 1 --warmup 1`, the timed iteration, summarised
# in bf16x3_iteration_o48v288.txt).
TRAFFIC_PER_LAUNCH = {("gemm_f64_128x128_NT", 48, 288): 135.6e9,
                      ("gemm_bf16x3_256x256_TN", 48, 288): 17.39e9}


def flops_per_iter(o, v):
    """Algorithmic FLOPs of one iteration: the sum over the 40 contractions of
    rhfccsd.cxx:231-309 of 2 x the product of their label extents (outer
    products included), collected by power of o and v."""
    return (2 * o**2 * v**4 + 12 * o**3 * v**3 + 14 * o**2 * v**3 + 4 * o * v**3 + 4 * o**4 * v**2
            + 20 * o**3 * v**2 + 10 * o**2 * v**2 + 2 * o**4 * v + 4 * o**3 * v + 4 * (o**2 * v + o * v**2))


def flops_per_iter_executed(o, v, packed):
    """FLOPs the engine executes per iteration: the packed ladder replaces the
    2 o^2 v^4 term by two GEMMs over packed pairs (P = v(v+1)/2, Q = v(v-1)/2)."""
    F = flops_per_iter(o, v)
    if not packed:
        return F
    P, Q, Op, Om = v * (v + 1) // 2, v * (v - 1) // 2, o * (o + 1) // 2, o * (o - 1) // 2
    return F - 2 * o**2 * v**4 + 2 * P * P * Op + 2 * Q * Q * Om


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def roofline_of(timing, prefix, peak):
    """Dominant GEMM class with the given prefix: algorithmic FLOPs per launch
    over the mean launch time (HIP events on the context stream)."""
    cands = {k: t for k, t in timing.items() if k.startswith(prefix) and t["launches"] > 0}
    if not cands:
        return None
    name, t = max(cands.items(), key=lambda kv: kv[1]["total_ms"])
    per_launch_flops = t["flops"] / t["launches"]
    mean_s = t["total_ms"] / t["launches"] / 1e3
    ach = per_launch_flops / mean_s
    out = {"kernel": name, "bound": "mfma", "achieved": ach / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
           "frac": ach / peak, "traffic": None, "launches": t["launches"], "mean_ms": t["total_ms"] / t["launches"],
           "flops_per_launch": per_launch_flops}
    if prefix.startswith("gemm_bf16x3"):
        # the split product runs three BF16 MFMAs per algorithmic multiply-add
        out["executed_tflops"] = 3 * ach / 1e12
        out["executed_frac"] = 3 * ach / peak
    return out


def host_cpu():
    """Host cores this process may use and the CPU model name."""
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return cores, model


def cpu_baseline(seed, o, v, nq, iters, gpu_fp64_rate, o_big, v_big):
    """The oracle (numpy + the host BLAS) timed on this host's cores: `iters`
    iterations of the given synthetic system, each timed with a wall clock
    around RHFCCSD.iterate as Iterative::run does (iterative.hpp:134-141),
    after one untimed iteration (BLAS threads, page faults).  Setup (MP2
    guess) and the integral build are outside the timed region.  The larger
    workload's rate is a FLOP-ratio estimate, labelled as such."""
    import numpy as np
    from oracle import rhfccsd, synth
    try:
        from threadpoolctl import threadpool_info
        threads = max(int(i.get("num_threads", 1)) for i in threadpool_info()) or 1
    except Exception:
        threads = os.cpu_count() or 1
    cores, model = host_cpu()
    f, V, _, _ = synth.synthetic_system(seed, o, v, nq)
    cc = rhfccsd.RHFCCSD(f, V)
    cc.iterate()
    per = []
    for _ in range(iters):
        t0 = time.perf_counter()
        cc.iterate()
        per.append(time.perf_counter() - t0)
    s_it = float(np.median(per))
    out = {"value": 1.0 / s_it, "unit": "iter/s", "cores": threads, "kind": "port",
           "sample": f"{iters} timed iterations (after 1 untimed) of the oracle's RHF-CCSD on the synthetic system "
                     f"o={o}, v={v}, nQ={nq}, seed {seed}, FP64, numpy + host BLAS on {threads} threads; "
                     f"wall clock per iterate, median {s_it:.2f} s/iteration",
           "host": {"cpu_model": model, "cores_available": cores, "blas_threads": threads},
           "s_per_iter": per, "gpu_fp64_iter_per_s_same_system": gpu_fp64_rate,
           "gpu_fp64_vs_cpu": (gpu_fp64_rate * s_it) if gpu_fp64_rate else None}
    scale = flops_per_iter(o, v) / flops_per_iter(o_big, v_big)
    out["estimate_at_bench_workload"] = {
        "value": scale / s_it, "unit": "iter/s", "o": o_big, "v": v_big,
        "how": f"estimate, not measured: the rate above scaled by F({o},{v})/F({o_big},{v_big}) = {scale:.3e}"}
    return out


def job_stream(seed, n, o_min, o_max, rate):
    """The serving workload: o ~ U{o_min..o_max}, v = round(o * U[4,8]),
    Poisson arrivals at `rate` jobs/s, one synthetic system per job."""
    import numpy as np
    rng = np.random.default_rng(seed)
    o = rng.integers(o_min, o_max + 1, size=n)
    v = np.rint(o * rng.uniform(4.0, 8.0, size=n)).astype(int)
    t = np.cumsum(rng.exponential(1.0 / rate, size=n))
    t -= t[0]
    return [(int(1000 + i), int(o[i]), int(v[i]), float(t[i])) for i in range(n)]


def serve_stream(a, seed):
    """Continuous-batching serving run of the job stream; returns a dict.

    Time-to-E is submit -> converged energy (queueing included); service time
    is start -> converged energy on the device.  A sample of the same jobs is
    then solved one at a time in FP64 (no arrivals: back to back), which gives
    the one-at-a-time baseline and each sampled job's |dE| mixed vs FP64."""
    import numpy as np
    import aquarius_amd as aq
    jobs = job_stream(seed, a.jobs, a.o_min, a.o_max, a.rate)
    if a.serve_every > 1:   # a fixed subset of the stream (every k-th job)
        jobs = jobs[::a.serve_every]
    if a.burst:   # every job submitted at t = 0: the loop's throughput without arrival gaps
        jobs = [(s_, o_, v_, 0.0) for s_, o_, v_, _ in jobs]
    cfg = aq.default_config(packed_ladder=a.packed_ladder, graph=a.graph, convergence=a.convergence,
                            diis_order=a.diis, max_iterations=150, precision=a.mixed,
                            refine_below=a.refine_below, min_reduced_flops=1e8)
    # jobs with o <= graph_max_o replay their stage from a HIP graph (launch-bound sizes)
    cfg_small = aq.default_config(packed_ladder=a.packed_ladder, graph=1, convergence=a.convergence,
                                  diis_order=a.diis, max_iterations=150, precision=a.mixed,
                                  refine_below=a.refine_below, min_reduced_flops=1e8)
    res = []
    with aq.Server(slots=a.slots, hbm_budget=int(a.budget_gb * 2**30), max_bypass_s=a.max_bypass) as srv:
        if a.grouping_us > 0:
            srv.set_grouping(a.grouping_us)
        t0 = time.perf_counter()
        ids = []
        for jseed, o, v, t in jobs:
            dt = t - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            ids.append(srv.submit(jseed, o, v, config=cfg_small if o <= a.graph_max_o else cfg))
        for k, i in enumerate(ids):
            r = srv.wait(i)
            res.append(r)
            if k % 16 == 15 or k == len(ids) - 1:
                log(f"served {k + 1}/{len(ids)} (last: o={jobs[k][1]} v={jobs[k][2]} E={r['energy']:.8f} "
                    f"time-to-E {r['t_end'] - r['t_submit']:.2f}s)")
        wall = max(r["t_end"] for r in res) - min(r["t_submit"] for r in res)
        grouping = srv.grouping_stats()
    ok = [r for r in res if r["status"] == aq.JOB_DONE and r["converged"]]
    tte = np.array([r["t_end"] - r["t_submit"] for r in ok])
    svc = np.array([r["t_end"] - r["t_start"] for r in ok])

    def pct(x):
        return {"p50": float(np.percentile(x, 50)), "p99": float(np.percentile(x, 99)),
                "max": float(x.max()), "mean": float(x.mean())} if len(x) else None

    out = {"metric": "served jobs/h + p99 time-to-E", "value": 3600.0 * len(ok) / wall, "unit": "jobs/h",
           "n_gpus": 1, "higher_is_better": True, "dtype": f"{a.mixed}+fp64", "data": "synthetic",
           "config": {"workload": f"{a.jobs} synthetic CCSD jobs, o~U{{{a.o_min}..{a.o_max}}}, v=round(o*U[4,8]), "
                                  f"Poisson arrivals {a.rate}/s, HBM budget {a.budget_gb} GB, {a.slots} slots",
                      "seed": seed, "convergence": a.convergence, "diis_order": a.diis,
                      "graph": a.graph, "graph_max_o": a.graph_max_o, "grouping_us": a.grouping_us},
           "grouped_launches": grouping,
           "gemms_per_grouped_launch": (grouping["gemms"] / grouping["launches"]) if grouping["launches"] else None,
           "jobs": len(res), "converged": len(ok),
           "failed": sum(r["status"] == aq.JOB_FAILED for r in res),
           "first_error": next((r["error"] for r in res if r["status"] == aq.JOB_FAILED), ""),
           "wall_s": wall, "time_to_E_s": pct(tte), "service_time_s": pct(svc)}
    if a.fp64_sample:
        # Baselines on identical jobs: every k-th job of the stream (a fixed
        # subset) is (1) served again as a burst with the same server
        # settings, (2) solved one at a time in FP64, back to back, and, with
        # --mixed-one-at-a-time, (3) one at a time in the mixed policy.  The
        # served/FP64 ratio compares the two on the same work; per-job |dE|
        # compares each subset job's served energy with its FP64 solve.
        every = max(1, len(jobs) // a.fp64_sample)
        sub = [jobs[k] for k in range(0, len(jobs), every)][:a.fp64_sample]
        idx = [k for k in range(0, len(jobs), every)][:a.fp64_sample]
        c64 = aq.default_config(packed_ladder=a.packed_ladder, convergence=a.convergence, diis_order=a.diis,
                                max_iterations=150)
        with aq.Server(slots=a.slots, hbm_budget=int(a.budget_gb * 2**30), max_bypass_s=a.max_bypass) as srv:
            if a.grouping_us > 0:
                srv.set_grouping(a.grouping_us)
            ids = [srv.submit(js, o, v, config=cfg) for js, o, v, _ in sub]
            rs = [srv.wait(i) for i in ids]
            sub_wall = max(r["t_end"] for r in rs) - min(r["t_submit"] for r in rs)

        def one_at_a_time(config):
            ts, es = [], []
            with aq.Server(slots=1, hbm_budget=int(a.budget_gb * 2**30)) as srv:
                for js, o, v, _ in sub:
                    r = srv.wait(srv.submit(js, o, v, config=config))
                    ts.append(r["t_end"] - r["t_start"])
                    es.append(r["energy"])
            return ts, es

        t64, e64 = one_at_a_time(c64)
        out["subset"] = {"jobs": len(sub), "every": every,
                         "served_burst_jobs_per_h": 3600.0 * len(sub) / sub_wall,
                         "fp64_one_at_a_time_jobs_per_h": 3600.0 * len(sub) / float(np.sum(t64)),
                         "fp64_one_at_a_time_mean_solve_s": float(np.mean(t64))}
        if a.mixed_one_at_a_time:
            tmx, _ = one_at_a_time(cfg)
            out["subset"]["mixed_one_at_a_time_jobs_per_h"] = 3600.0 * len(sub) / float(np.sum(tmx))
        out["vs_fp64_one_at_a_time"] = (out["subset"]["served_burst_jobs_per_h"] /
                                        out["subset"]["fp64_one_at_a_time_jobs_per_h"])
        dE = [{"o": o, "v": v, "dE": res[k]["energy"] - e}
              for k, (_, o, v, _), e in zip(idx, sub, e64) if res[k]["status"] == aq.JOB_DONE]
        out["dE_mixed_vs_fp64_per_subset_job"] = dE
        out["max_abs_dE_mixed_vs_fp64"] = max((abs(d["dE"]) for d in dE), default=None)
    return out


def warm_tuner(a, ctx, ints):
    """One untimed iteration in each arithmetic, so that the GEMM autotuner's
    first-encounter trials and the operand caches' first fills of the
    integrals fall outside every timed solve.  The cached copies stay: with
    a trim here (cache dropped, pool memory handed back to the driver) the
    next FP64 iteration took 3.3 s instead of 0.66 s, measured."""
    import aquarius_amd as aq
    t0 = time.perf_counter()
    for prec in ("fp64", a.mixed):
        cc = aq.RHFCCSD(ctx, ints, packed_ladder=a.packed_ladder, convergence=a.convergence, diis_order=a.diis,
                        max_iterations=1000, precision=prec, refine_below=a.refine_below, min_reduced_flops=1e8)
        for _ in range(max(1, a.warmup)):
            cc.iterate()
        cc.close()
    log(f"GEMM tuner warmed in {time.perf_counter() - t0:.1f}s ({len(aq.autotune_table())} shape buckets)")


def copy_calibration(n=1 << 27, reps=3):
    """Plain copies B = A of a 2^27-element FP64 tensor (1 GiB read + 1 GiB
    written per launch): the known-size reference for the HBM counters."""
    import aquarius_amd as aq
    ctx = aq.Context(0)
    A = ctx.tensor((n,), zero=False)
    ctx.fill(A, 1.0, 0.0)
    B = ctx.tensor((n,), zero=True)
    ctx.enable_timing(True)
    ctx.timing_reset()
    for _ in range(reps):
        ctx.sum(1.0, A, "i", 0.0, B, "i")
    ctx.sync()
    tim = ctx.timing()
    A.free()
    B.free()
    ctx.close()
    print(json.dumps({"copy_calibration": {"elements": n, "bytes_read": 8 * n, "bytes_written": 8 * n,
                                           "launches": reps, "timing": tim}}))


def hbm_ceiling(v=288, o=48, reps=5):
    """Achievable HBM bandwidth next to the library's HBM-bound kernels on
    amplitude-sized FP64 tensors [v,v,o,o] (o=48, v=288: 1.53 GB each): a
    plain device copy through torch, then the permuted copies and adds,
    transposes, dot and the fused amplitude update as the solver issues them.
    Bytes are algorithmic (8 B per element read or written)."""
    import torch
    import aquarius_amd as aq
    out = []
    n = v * v * o * o
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(2):
        y.copy_(x)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        y.copy_(x)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    out.append({"kernel": "torch Tensor.copy_", "what": "y = x", "ms": ms, "GBps": 16.0 * n / ms / 1e6})
    del x, y
    torch.cuda.empty_cache()
    with aq.Context(0) as ctx:
        A = ctx.tensor((v, v, o, o), zero=False)
        A2 = ctx.tensor((v, o, v, o), zero=False)
        B = ctx.tensor((v, v, o, o), zero=False)
        for t in (A, A2, B):
            ctx.fill(t, 0.5, 0.0)
        cases = [(A, "abij", 0.0, "copy"), (A, "abij", 1.0, "add"), (A, "abji", 0.0, "copy, i<->j"),
                 (A, "abji", 1.0, "add, i<->j"), (A, "baji", 1.0, "add, a<->b and i<->j"),
                 (A2, "aibj", 0.0, "copy, LDS transpose"), (A2, "aibj", 1.0, "add, LDS transpose")]
        for src, ia, beta, what in cases:
            ctx.sum(1.0, src, ia, beta, B, "abij")
            ctx.sync()
            ctx.enable_timing(True)
            ctx.timing_reset()
            for _ in range(reps):
                ctx.sum(1.0, src, ia, beta, B, "abij")
            ctx.sync()
            for k, t in ctx.timing().items():
                if t["launches"]:
                    ms = t["total_ms"] / t["launches"]
                    out.append({"kernel": k, "what": f"B[abij] = {'B + ' if beta else ''}A[{ia}] ({what})",
                                "ms": ms, "GBps": t["bytes"] / t["launches"] / ms / 1e6})
            ctx.enable_timing(False)
        ctx.enable_timing(True)
        ctx.timing_reset()
        for _ in range(reps):
            ctx.dot(A, "abij", B, "abij")
        for k, t in ctx.timing().items():
            if t["launches"] and k == "dot":
                ms = t["total_ms"] / t["launches"]
                out.append({"kernel": k, "what": "sum A*B", "ms": ms, "GBps": t["bytes"] / t["launches"] / ms / 1e6})
        ctx.enable_timing(False)
    for r in out:
        print(json.dumps(r))


def trace_split(path, kernel="k_gemm<double, 128, 128, 2, 2, false, true, 16, 2>"):
    """Per-solve launch statistics of one kernel in a rocprofv3 kernel trace
    (run_kernel_trace.csv) of the default bench command: the trace is cut at
    the solver setups, each of which launches k_pair_asym twice (warm-up FP64,
    warm-up BF16x3, mixed solve, FP64 solve, then the later legs)."""
    import csv
    setups, launches = [], []
    with open(path) as fh:
        for r in csv.DictReader(fh):
            name = r["Kernel_Name"]
            if "k_pair_asym" in name:
                setups.append(int(r["Start_Timestamp"]))
            elif kernel in name:
                t0, t1 = int(r["Start_Timestamp"]), int(r["End_Timestamp"])
                launches.append((t0, (t1 - t0) / 1e6))
    setups.sort()
    launches.sort()
    cuts = setups[0:8:2] + [setups[8] if len(setups) > 8 else 1 << 62]
    out = {}
    for label, lo, hi in zip(("warm-up fp64", "warm-up bf16x3", "mixed solve", "fp64 solve"), cuts[:-1], cuts[1:]):
        d = [ms for t, ms in launches if lo <= t < hi]
        out[label] = {"launches": len(d), "mean_ms": sum(d) / len(d) if d else None}
    print(json.dumps({"kernel": kernel, "per_solve": out}))


def only(a):
    """W warmup + exactly K timed iterations of one solve (FP64, or the mixed
    policy from its first iteration) and nothing else: the command profiled
    for the per-kernel statistics and HBM counters."""
    import aquarius_amd as aq
    o, v = a.o, a.v
    nq = a.nq or 3 * (o + v)
    ctx = aq.Context(0)
    ctx.set_autotune(bool(a.autotune))
    ints = aq.Integrals.synthetic(ctx, a.seed, o, v, nq)
    ctx.sync()
    prec = "fp64" if a.only == "fp64" else a.mixed
    cc = aq.RHFCCSD(ctx, ints, packed_ladder=a.packed_ladder, convergence=a.convergence, diis_order=a.diis, max_iterations=1000, precision=prec,
                    refine_below=a.refine_below, min_reduced_flops=1e8)
    for _ in range(a.warmup):
        cc.iterate()
    ctx.enable_timing(True)
    ctx.timing_reset()
    ctx.sync()
    steps = max(a.steps, 1)
    t0 = time.perf_counter()
    for k in range(steps):
        e, c = cc.iterate()
        log(f"{a.only} it {k + 1} E={e:.12f} conv={c:.2e}")
    ctx.sync()
    dt = time.perf_counter() - t0
    tim = ctx.timing()
    peak = PEAK["fp64"] if a.only == "fp64" else PEAK[a.mixed]
    rf = roofline_of(tim, "gemm_f64" if a.only == "fp64" else ("gemm_bf16x3" if a.mixed == "bf16x3" else "gemm_f32"),
                     peak)
    cc.close()
    ints.free()
    ctx.close()
    print(json.dumps({"profiling_run": a.only, "o": o, "v": v, "steps": steps, "warmup": a.warmup,
                      "ms_per_iter": 1e3 * dt / steps, "iter_per_s": steps / dt, "roofline": rf,
                      "timing": {k: t for k, t in tim.items() if t["launches"] > 0}}))


KNOWN_ANSWERS = {   # test/input.aq:20-23 (H2O cc-pVDZ, bohr geometry)
    "h2o-bohr": {"scf": -74.550126456692, "mp2": -0.171348679568, "ccsd": -0.180145524753},
}


def molecular(a):
    """Real-molecule workload through the integral front end: host AO
    integrals, device SCF, device AO->MO, then the FP64 and the mixed CCSD
    solves to convergence.  Prints one JSON line."""
    import numpy as np
    import aquarius_amd as aq
    from aquarius_amd.molecules import H2O_ANGSTROM, H2O_REFERENCE_BOHR, water_cluster
    if a.molecule == "h2o":
        atoms, units, label = H2O_ANGSTROM, "angstrom", "H2O cc-pVDZ (angstrom geometry)"
    elif a.molecule == "h2o-bohr":
        atoms, units, label = H2O_REFERENCE_BOHR, "bohr", "H2O cc-pVDZ (bohr geometry of test/input.aq)"
    else:
        nw = int(a.molecule[len("water"):])
        atoms, units = water_cluster(a.seed, nw), "angstrom"
        label = f"(H2O){nw} cc-pVDZ, seed {a.seed}, 7 A cube, min O-O 2.7 A"
    ctx = aq.Context(0)
    ctx.set_autotune(bool(a.autotune))
    t = {}
    t0 = time.perf_counter()
    mol = aq.Molecule(atoms, a.basis, units=units)
    eri = mol.ints_2e_device(ctx)
    t["ao_integrals_s"] = time.perf_counter() - t0
    log(f"{label}: nao={mol.nao} AO integrals {t['ao_integrals_s']:.2f}s")
    t0 = time.perf_counter()
    scf = aq.SCF(ctx, mol, eri=eri)
    t["scf_s"] = time.perf_counter() - t0
    log(f"SCF E={scf.energy:.12f} in {scf.iterations} iterations, {t['scf_s']:.2f}s")
    t0 = time.perf_counter()
    ints = scf.integrals()
    ctx.sync()
    t["ao_to_mo_s"] = time.perf_counter() - t0
    eri.free()
    ctx.trim()
    o, v = scf.nocc, mol.nao - scf.nocc
    F = flops_per_iter(o, v)
    warm_tuner(a, ctx, ints)
    runs = {}
    for name, prec in (("fp64", "fp64"), ("mixed", a.mixed)):
        cc = aq.RHFCCSD(ctx, ints, packed_ladder=a.packed_ladder, convergence=a.convergence, diis_order=a.diis, max_iterations=150,
                        precision=prec, refine_below=a.refine_below, min_reduced_flops=1e8)
        mp2 = cc.mp2
        t0 = time.perf_counter()
        r = cc.run(callback=lambda h, n=name: log(f"{n} it {h['iter']:2d} E={h['energy']:.12f} "
                                                   f"conv={h['conv']:.2e} {h['seconds'] * 1e3:.1f} ms"))
        dt = time.perf_counter() - t0
        hist = cc.history
        red = [h["seconds"] for h in hist if h["precision"] != aq.AQ_PREC_FP64]
        full = [h["seconds"] for h in hist if h["precision"] == aq.AQ_PREC_FP64]
        if name == "fp64":
            t2_ref = cc.tensor("T2").copy()
            dt2 = 0.0
        else:
            d = cc.tensor("T2").copy()
            ctx.sum(-1.0, t2_ref, "abij", 1.0, d, "abij")
            dt2 = ctx.norm(d, 0)
            d.free()
            t2_ref.free()
        runs[name] = {"energy": r["energy"], "mp2": mp2, "iterations": r["iterations"], "converged": r["converged"],
                      "max_abs_dT2_vs_fp64": dt2,
                      "time_to_E_s": dt, "iter_per_s": len(hist) / sum(h["seconds"] for h in hist),
                      "reduced_iterations": len(red),
                      "reduced_iter_per_s": len(red) / sum(red) if red else None,
                      "fp64_iter_per_s": len(full) / sum(full) if full else None}
        cc.close()
    out = {"metric": METRIC, "value": runs["mixed"]["iter_per_s"], "unit": "iter/s", "n_gpus": 1,
           "higher_is_better": True, "dtype": f"{a.mixed}+fp64", "data": "molecular (integral front end)",
           "config": {"workload": label, "basis": a.basis, "nao": mol.nao, "o": o, "v": v,
                      "convergence": a.convergence, "diis_order": a.diis, "flops_per_iter": F,
                      "packed_ladder": bool(a.packed_ladder),
                      "flops_per_iter_executed": flops_per_iter_executed(o, v, a.packed_ladder)},
           "scf_energy": scf.energy, "scf_iterations": scf.iterations, "front_end_s": t,
           "fp64": runs["fp64"], "mixed": runs["mixed"],
           "dE_mixed_vs_fp64": runs["mixed"]["energy"] - runs["fp64"]["energy"],
           "speedup_iter_per_s": runs["mixed"]["iter_per_s"] / runs["fp64"]["iter_per_s"],
           "speedup_time_to_E": runs["fp64"]["time_to_E_s"] / runs["mixed"]["time_to_E_s"]}
    if a.molecule in KNOWN_ANSWERS:
        k = KNOWN_ANSWERS[a.molecule]
        out["vs_known_answers"] = {"scf": scf.energy - k["scf"], "mp2": runs["fp64"]["mp2"] - k["mp2"],
                                   "ccsd_fp64": runs["fp64"]["energy"] - k["ccsd"],
                                   "ccsd_mixed": runs["mixed"]["energy"] - k["ccsd"]}
    ints.free()
    ctx.close()
    print(json.dumps(out))


def timed_solve(aq, ctx, ints, a, prec, label):
    """One solve to convergence, every iteration timed with a wall clock
    around RHFCCSD::iterate (iterative.hpp:134-141; solver setup outside).
    Returns the solver (open) and the per-iteration history."""
    cc = aq.RHFCCSD(ctx, ints, packed_ladder=a.packed_ladder, convergence=a.convergence, diis_order=a.diis,
                    max_iterations=200, precision=prec, refine_below=a.refine_below, min_reduced_flops=1e8)
    hist = []
    while True:
        p = cc.current_precision()
        t1 = time.perf_counter()
        e, c = cc.iterate()
        dt = time.perf_counter() - t1
        hist.append({"energy": e, "conv": c, "precision": p, "seconds": dt})
        log(f"{label} it {len(hist):2d} [{['fp64', 'fp32', 'bf16x3'][p]}] E={e:.12f} conv={c:.2e} {dt * 1e3:.0f} ms")
        if c < a.convergence or len(hist) >= 200:
            break
    return cc, hist


def solve_summary(hist, convergence):
    n = len(hist)
    t = sum(h["seconds"] for h in hist)
    phases = {}
    for h in hist:
        ph = phases.setdefault(["fp64", "fp32", "bf16x3"][h["precision"]], {"iterations": 0, "seconds": 0.0})
        ph["iterations"] += 1
        ph["seconds"] += h["seconds"]
    for ph in phases.values():
        ph["iter_per_s"] = ph["iterations"] / ph["seconds"]
        ph["ms_per_iter"] = 1e3 * ph["seconds"] / ph["iterations"]
    return {"iterations": n, "converged": hist[-1]["conv"] < convergence, "energy": hist[-1]["energy"],
            "time_to_E_s": t, "iter_per_s": n / t, "ms_per_iter": 1e3 * t / n, "phases": phases}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--molecule", default=None, help="h2o | h2o-bohr | waterN: molecular workload via the "
                                                     "integral front end instead of the synthetic system")
    ap.add_argument("--basis", default="cc-pVDZ")
    ap.add_argument("--serving", action="store_true", help="run only the continuous-batching serving workload")
    ap.add_argument("--serving-in-line", type=int, default=1,
                    help="default run: also run the serving stream and report it as the line's `serving` object")
    ap.add_argument("--jobs", type=int, default=256)
    ap.add_argument("--o-min", type=int, default=4)
    ap.add_argument("--o-max", type=int, default=32)
    ap.add_argument("--rate", type=float, default=4.0)
    ap.add_argument("--budget-gb", type=float, default=240.0)
    ap.add_argument("--slots", type=int, default=4)
    ap.add_argument("--max-bypass", type=float, default=30.0)
    ap.add_argument("--fp64-sample", type=int, default=16,
                    help="serving: size of the job subset (every k-th job) for the one-at-a-time baselines")
    ap.add_argument("--burst", action="store_true", help="serving: submit every job at t = 0 (no arrival gaps)")
    ap.add_argument("--serve-every", type=int, default=1, help="serving: serve only every k-th job of the stream")
    ap.add_argument("--mixed-one-at-a-time", type=int, default=0,
                    help="serving: also solve the baseline subset one at a time in the mixed policy")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=0,
                    help="kept for the harness; the timed region is always the complete mixed solve (every "
                         "iteration to convergence, FP64 refinement included), because the metric is defined at "
                         "the converged energy")
    ap.add_argument("--warmup", type=int, default=1,
                    help="untimed iterations in each arithmetic on a throwaway solver before the timed solves "
                         "(GEMM tuner trials, operand copies of the integrals)")
    ap.add_argument("--o", type=int, default=48)
    ap.add_argument("--v", type=int, default=288)
    ap.add_argument("--nq", type=int, default=0)
    ap.add_argument("--seed", type=int, default=None, help="default: 3 (single solve), 4 (serving stream)")
    ap.add_argument("--convergence", type=float, default=1e-9)
    ap.add_argument("--diis", type=int, default=8)
    ap.add_argument("--mixed", default="bf16x3", choices=["bf16x3", "fp32"])
    ap.add_argument("--refine-below", type=float, default=1e-5)
    ap.add_argument("--cpu-sample", default="1,20,140,2",
                    help="seed,o,v,iterations of the timed CPU oracle leg ('' to skip)")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 reference solve")
    ap.add_argument("--only", choices=["fp64", "mixed"], default=None,
                    help="profiling run: W warmup + K timed iterations of one solve, then exit")
    ap.add_argument("--trace-split", default=None, metavar="RUN_KERNEL_TRACE_CSV",
                    help="per-solve launch statistics of the dominant kernel in a rocprofv3 trace of this bench")
    ap.add_argument("--hbm-ceiling", action="store_true",
                    help="time a plain device copy beside the permute/transpose/dot kernels (HBM bandwidth ceiling)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="profiling run: plain device copies of a 1 GiB FP64 tensor (HBM counter calibration)")
    ap.add_argument("--autotune", type=int, default=1, help="GEMM autotuner on (1) or off (0)")
    ap.add_argument("--packed-ladder", type=int, default=1,
                    help="particle-particle ladder on packed (anti)symmetric pairs (1) or as the full contraction (0)")
    ap.add_argument("--graph", type=int, default=0, help="serving: HIP-graph replay of each job's contraction stage "
                    "(stages replayed from a graph do not group)")
    ap.add_argument("--graph-max-o", type=int, default=0,
                    help="serving: jobs with o <= this use HIP-graph replay whatever --graph says")
    ap.add_argument("--grouping-us", type=int, default=50,
                    help="serving: cross-job grouped GEMM launches with this host window in us (0: off)")
    ap.add_argument("--compare-precisions", type=int, default=0,
                    help="also solve with the other reduced arithmetic (FP32 MFMA beside BF16x3) and report its "
                         "energy and T2 error against FP64")
    a = ap.parse_args()

    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(json.dumps({"metric": METRIC, "error": "this engine runs on one GPU; sharding over several GPUs is "
                          "not built", "n_gpus": a.gpus}))
        sys.exit(2)

    serving_seed = a.seed if (a.seed is not None and a.serving) else 4
    if a.seed is None:
        a.seed = 7 if a.molecule else (4 if a.serving else 3)
    if a.molecule:
        return molecular(a)
    if a.serving:
        print(json.dumps(serve_stream(a, serving_seed)))
        return
    if a.copy_calibration:
        return copy_calibration()
    if a.hbm_ceiling:
        return hbm_ceiling()
    if a.trace_split:
        return trace_split(a.trace_split)
    if a.only:
        return only(a)

    import aquarius_amd as aq

    o, v = a.o, a.v
    nq = a.nq or 3 * (o + v)
    F = flops_per_iter(o, v)
    ctx = aq.Context(0)
    ctx.set_autotune(bool(a.autotune))
    t0 = time.perf_counter()
    ints = aq.Integrals.synthetic(ctx, a.seed, o, v, nq)
    ctx.sync()
    log(f"integrals o={o} v={v} nQ={nq} built in {time.perf_counter() - t0:.1f}s; "
        f"HBM in use {ctx.memory()[0] / 2**30:.1f} GiB")
    warm_tuner(a, ctx, ints)
    t2_64 = None

    def max_dt2(cc):
        """max |T2 - T2(FP64)| of a converged solver."""
        if t2_64 is None:
            return None
        d = cc.tensor("T2").copy()
        ctx.sum(-1.0, t2_64, "abij", 1.0, d, "abij")
        r = ctx.norm(d, 0)
        d.free()
        return r

    # ---- mixed solve to convergence: the timed region of the headline value.
    # It runs first, right after the warm-up, so that each solve finds the
    # integrals' operand copies of its own arithmetic cached (the FP64 solve's
    # copies were made by the warm-up and its own refinement phase).
    ctx.enable_timing(True)
    ctx.timing_reset()
    ctx.sync()
    tw0 = time.perf_counter()
    cc, hm = timed_solve(aq, ctx, ints, a, a.mixed, "mixed")
    ctx.sync()
    t_window = time.perf_counter() - tw0
    timm = ctx.timing()
    sm = solve_summary(hm, a.convergence)
    t2_mixed = cc.tensor("T2").copy()
    cc.close()

    # ---- FP64 solve to convergence (the reference arithmetic)
    s64 = None
    dt2_mixed = None
    if not a.no_fp64:
        ctx.enable_timing(True)
        ctx.timing_reset()
        cc, h64 = timed_solve(aq, ctx, ints, a, "fp64", "fp64")
        tim64 = ctx.timing()
        s64 = solve_summary(h64, a.convergence)
        s64["roofline"] = roofline_of(tim64, "gemm_f64", PEAK["fp64"])
        t2_64 = cc.tensor("T2").copy()
        cc.close()
        d = t2_mixed.copy()
        ctx.sum(-1.0, t2_64, "abij", 1.0, d, "abij")
        dt2_mixed = ctx.norm(d, 0)
        d.free()
    t2_mixed.free()

    # ---- the other reduced arithmetic (optional): energy and T2 error
    other = None
    if a.compare_precisions and s64 is not None:
        prec2 = "fp32" if a.mixed == "bf16x3" else "bf16x3"
        c2, h2 = timed_solve(aq, ctx, ints, a, prec2, prec2)
        other = solve_summary(h2, a.convergence)
        other.update({"precision": prec2, "dE_vs_fp64": other["energy"] - s64["energy"],
                      "max_abs_dT2_vs_fp64": max_dt2(c2)})
        c2.close()
    if t2_64 is not None:
        t2_64.free()

    rf_red = roofline_of(timm, "gemm_bf16x3" if a.mixed == "bf16x3" else "gemm_f32", PEAK[a.mixed])
    rf_64 = roofline_of(timm, "gemm_f64", PEAK["fp64"])
    cands = [r for r in (rf_red, rf_64) if r]
    dominant = max(cands, key=lambda r: r["launches"] * r["mean_ms"]) if cands else None
    for r in (rf_red, rf_64, (s64 or {}).get("roofline")):
        if r:
            r["traffic"] = TRAFFIC_PER_LAUNCH.get((r["kernel"], o, v))
            r["traffic_unit"] = "bytes per launch (rocprofv3 --pmc, copy-calibrated)" if r["traffic"] else None

    steps = sm["iterations"]
    t_iters = sm["time_to_E_s"]
    res = {
        "metric": METRIC, "value": steps / t_iters, "unit": "iter/s", "n_gpus": 1, "steps": steps,
        "warmup": a.warmup, "ms_per_step": 1e3 * t_iters / steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": f"{a.mixed}+fp64", "data": "synthetic",
        "timed_region": "the complete mixed solve: every iteration from the first to convergence "
                        f"(conv < {a.convergence:g}), {a.mixed} phase and FP64 refinement, wall clock per iteration; "
                        f"--steps {a.steps} does not shorten it (the metric is defined at the converged energy)",
        "timed_window_wall_s": t_window,
        "config": {"workload": f"synthetic closed-shell RHF-CCSD o={o} v={v} nQ={nq} seed={a.seed}, full vvvv "
                               f"resident; {a.mixed} MFMA contractions until conv<{a.refine_below:g}, then FP64"
                               + ("; ladder on packed (anti)symmetric pairs" if a.packed_ladder else ""),
                   "o": o, "v": v, "nQ": nq, "seed": a.seed, "sigma": aq.synth_default_sigma(v, nq),
                   "convergence": a.convergence, "diis_order": a.diis, "flops_per_iter": F,
                   "packed_ladder": bool(a.packed_ladder),
                   "flops_per_iter_executed": flops_per_iter_executed(o, v, a.packed_ladder),
                   "gemm_autotune": bool(a.autotune)},
        "roofline": dominant,
        "mixed": dict(sm, roofline_reduced=rf_red, roofline_fp64=rf_64),
        "reduced_phase_iter_per_s": sm["phases"].get(a.mixed, {}).get("iter_per_s"),
        "fp64": s64,
    }
    if s64 is not None:
        res["dE_mixed_vs_fp64"] = sm["energy"] - s64["energy"]
        res["max_abs_dT2_mixed_vs_fp64"] = dt2_mixed
        res["other_reduced_precision"] = other
        res["fp64_iter_per_s"] = s64["iter_per_s"]
        res["speedup_iter_per_s"] = sm["iter_per_s"] / s64["iter_per_s"]
        res["speedup_time_to_E"] = s64["time_to_E_s"] / sm["time_to_E_s"]
        res["reduced_phase_speedup_iter_per_s"] = (res["reduced_phase_iter_per_s"] or 0) / s64["iter_per_s"]
    tab = aq.autotune_table()
    res["gemm_autotune"] = {"buckets": len(tab),
                            "variant_counts": {p: {str(vv): sum(1 for e in tab if e["precision"] == c and e["variant"] == vv)
                                                   for vv in (0, 1, 2, 3, 4)}
                                               for p, c in aq.PRECISIONS.items()}}
    ints.free()
    ctx.close()

    # ---- the CPU leg's system on the GPU (FP64 solve to convergence), in a
    # context of its own so that it starts as a solve of that size would
    cpu_sys = tuple(int(x) for x in a.cpu_sample.split(",")) if a.cpu_sample else None
    gpu_small_rate = None
    if cpu_sys:
        cs, co, cv, _ = cpu_sys
        ctx = aq.Context(0)
        ctx.set_autotune(bool(a.autotune))
        si = aq.Integrals.synthetic(ctx, cs, co, cv, 3 * (co + cv))
        c1, h1 = timed_solve(aq, ctx, si, a, "fp64", f"fp64 o={co} v={cv}")
        s1 = solve_summary(h1, a.convergence)
        gpu_small_rate = s1["iter_per_s"]
        res["fp64_cpu_leg_system"] = {"o": co, "v": cv, "seed": cs, **s1}
        c1.close()
        si.free()
        ctx.close()

    # ---- serving stream (continuous batching), same run
    if a.serving_in_line and a.jobs > 0:
        res["serving"] = serve_stream(a, 4)

    if cpu_sys:
        cs, co, cv, ci = cpu_sys
        log(f"cpu baseline: {ci} oracle iterations at o={co} v={cv}")
        res["cpu_baseline"] = cpu_baseline(cs, co, cv, 3 * (co + cv), ci, gpu_small_rate, o, v)
    else:
        res["cpu_baseline"] = None
    print(json.dumps(res))


if __name__ == "__main__":
    main()
