"""HIP-graph replay of the CCSD contraction stage (aq_rhfccsd_config.graph).

The graph replays the kernels and arguments the eager stage issued when it
was captured, so a solve with graphs must equal the eager solve bit for bit:
every iteration's energy and residual, the iteration count and the final
amplitudes, in FP64 and across the BF16x3 -> FP64 switch (which drops the
graph and captures a new one).  Each graph replay is followed by the eager
update and DIIS work and a host synchronisation, and every replay's result
is checked.  Also: a device free between iterations invalidates the graph
(it is captured again and the results do not change); a context that records
kernel timings never captures; the H2O/cc-pVDZ known answer through graphs."""

import numpy as np
import pytest

import aquarius_amd as aq

pytestmark = pytest.mark.gpu


def solve(seed, o, v, graph, steps=None, between=None, timing=False, **over):
    """Eager or graph solve in a fresh context; per-iteration (E, conv)."""
    with aq.Context() as ctx:
        if timing:
            ctx.enable_timing(True)
        ints = aq.Integrals.synthetic(ctx, seed, o, v)
        cc = aq.RHFCCSD(ctx, ints, graph=graph, **over)
        hist = []
        if steps is None:
            r = cc.run()
            hist = [(h["energy"], h["conv"]) for h in cc.history]
            assert r["converged"]
        else:
            for it in range(steps):
                hist.append(cc.iterate())
                if between is not None:
                    between(ctx, it)
        T2 = cc.tensor("T2").numpy()
        stats = cc.graph_stats()
        cc.close()
        ints.free()
    return hist, T2, stats


def test_graph_fp64_bitwise_equals_eager():
    print("case: FP64 o=6 v=26, 14 iterations, graph vs eager", flush=True)
    h0, T0, s0 = solve(41, 6, 26, graph=0, steps=14)
    h1, T1, s1 = solve(41, 6, 26, graph=1, steps=14)
    assert s0 == (0, 0, 0)
    captures, replays, failed = s1
    assert captures == 1 and failed == 0 and replays == 14 - 2, s1   # eager, capture, replays
    assert h0 == h1
    assert np.array_equal(T0, T1)


def test_graph_mixed_switch_bitwise_equals_eager():
    print("case: BF16x3 -> FP64 o=5 v=22 to 1e-9, graph vs eager", flush=True)
    over = dict(precision="bf16x3", refine_below=1e-5, min_reduced_flops=0.0, convergence=1e-9,
                diis_order=8, max_iterations=80)
    h0, T0, _ = solve(42, 5, 22, graph=0, **over)
    h1, T1, s1 = solve(42, 5, 22, graph=1, **over)
    captures, replays, failed = s1
    assert captures == 2 and failed == 0, s1        # one graph per contraction precision
    assert replays == len(h1) - 4, s1
    assert h0 == h1
    assert np.array_equal(T0, T1)


def test_graph_recaptured_after_a_device_free():
    print("case: FP64 o=4 v=18, a tensor freed after iteration 5", flush=True)

    def free_one(ctx, it):
        if it == 5:
            t = ctx.tensor((64,))
            t.free()

    h0, T0, _ = solve(43, 4, 18, graph=0, steps=12)
    h1, T1, s1 = solve(43, 4, 18, graph=1, steps=12, between=free_one)
    captures, replays, failed = s1
    # iterations 0 eager, 1 capture, 2-5 replays; 6 eager (graph dropped), 7 capture, 8-11 replays
    assert (captures, replays, failed) == (2, 8, 0), s1
    assert h0 == h1
    assert np.array_equal(T0, T1)


def test_no_graph_while_timing():
    print("case: FP64 o=4 v=16 with kernel timing on", flush=True)
    h0, T0, _ = solve(44, 4, 16, graph=0, steps=6)
    h1, T1, s1 = solve(44, 4, 16, graph=1, steps=6, timing=True)
    assert s1 == (0, 0, 0)
    assert h0 == h1 and np.array_equal(T0, T1)


def test_graph_h2o_ccpvdz_known_answer():
    print("case: H2O cc-pVDZ bohr through graphs", flush=True)
    from aquarius_amd.molecules import H2O_REFERENCE_BOHR
    with aq.Context() as ctx:
        mol = aq.Molecule(H2O_REFERENCE_BOHR, "cc-pVDZ", units="bohr")
        scf = aq.SCF(ctx, mol)
        ints = scf.integrals()
        cc = aq.RHFCCSD(ctx, ints, graph=1)
        r = cc.run()
        assert r["converged"]
        assert abs(r["energy"] - (-0.180145524753)) < 1e-9                 # test/input.aq:23
        captures, replays, failed = cc.graph_stats()
        assert captures == 1 and failed == 0 and replays == r["iterations"] - 2
        cc.close()
        ints.free()


def test_graph_serving_mix_equals_eager_alone():
    """Four slot contexts capture and replay concurrently (one graph per job
    and precision); every job equals its eager solve on a one-slot server."""
    print("case: serving 6 jobs, 4 slots, graphs on vs eager alone", flush=True)
    jobs = [(51, 4, 18), (52, 5, 24), (53, 3, 14), (54, 6, 30), (55, 5, 20), (56, 4, 26)]

    def cfg(graph):
        return aq.default_config(convergence=1e-9, diis_order=8, max_iterations=80, precision="bf16x3",
                                 refine_below=1e-5, min_reduced_flops=0.0, graph=graph)

    with aq.Server(slots=1) as s:
        alone = [s.wait(s.submit(seed, o, v, config=cfg(0))) for seed, o, v in jobs]
    with aq.Server(slots=4) as s:
        ids = [s.submit(seed, o, v, config=cfg(1)) for seed, o, v in jobs]
        mix = [s.wait(i) for i in ids]
    for a, m, job in zip(alone, mix, jobs):
        assert a["status"] == aq.JOB_DONE and m["status"] == aq.JOB_DONE, (job, a["error"], m["error"])
        assert a["energy"] == m["energy"] and a["iterations"] == m["iterations"], (job, a, m)


def test_hbar_after_graph_replays_equals_eager():
    """The Hbar products formed after a solve that replayed graphs read the
    operand cache the graph left behind; they must equal the eager solve's."""
    print("case: FP64 o=5 v=20, 10 iterations then hbar(with_wabej)", flush=True)
    out = []
    for graph in (0, 1):
        with aq.Context() as ctx:
            ints = aq.Integrals.synthetic(ctx, 45, 5, 20)
            cc = aq.RHFCCSD(ctx, ints, graph=graph)
            for _ in range(10):
                cc.iterate()
            cc.hbar(with_wabej=True)
            out.append({k: cc.tensor(k).numpy() for k in ("FAE", "FMI", "FME", "WMNIJ", "WMNEJ", "WAMIJ",
                                                          "WAMEI", "WAMIE", "WABEJ")})
            if graph:
                assert cc.graph_stats()[1] == 8
            cc.close()
            ints.free()
    for k in out[0]:
        assert np.array_equal(out[0][k], out[1][k]), k
