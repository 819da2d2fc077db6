"""Serving loop: continuous batching of CCSD jobs on one device.

Checks: a job's result is identical alone and in a concurrent mix (stated
tolerance 1e-12 Eh; the engine's reductions are deterministic, so equality is
expected bit for bit), equals the direct solver's result, admission respects
the HBM budget with and without backfill, oversize jobs are rejected at
submit, and a failing job does not disturb the others."""

import numpy as np
import pytest

import aquarius_amd as aq

pytestmark = pytest.mark.gpu

JOBS = [(11, 4, 18), (12, 5, 24), (13, 3, 14), (14, 6, 30), (15, 5, 20), (16, 4, 26)]


def cfg(prec="bf16x3"):
    return aq.default_config(convergence=1e-9, diis_order=8, max_iterations=80, precision=prec,
                             refine_below=1e-5, min_reduced_flops=0.0)


def run_alone(specs, prec):
    out = []
    with aq.Server(slots=1) as s:
        for seed, o, v in specs:
            out.append(s.wait(s.submit(seed, o, v, config=cfg(prec))))
    return out


@pytest.mark.parametrize("prec", ["fp64", "bf16x3"])
def test_alone_equals_in_the_mix(prec):
    alone = run_alone(JOBS, prec)
    with aq.Server(slots=4) as s:
        ids = [s.submit(seed, o, v, config=cfg(prec)) for seed, o, v in JOBS]
        mix = [s.wait(i) for i in ids]
    overlap = False
    for a, m, (seed, o, v) in zip(alone, mix, JOBS):
        assert a["status"] == aq.JOB_DONE and m["status"] == aq.JOB_DONE, (a["error"], m["error"])
        assert a["converged"] and m["converged"]
        assert abs(a["energy"] - m["energy"]) <= 1e-12, (seed, a["energy"], m["energy"])
        assert a["energy"] == m["energy"] and a["iterations"] == m["iterations"]
        print(f"job seed={seed} o={o} v={v}: E={m['energy']:.12f} it={m['iterations']} "
              f"(reduced {m['reduced_iterations']}) slot={m['slot']} peak={m['bytes_peak'] / 2**20:.1f} MiB "
              f"reserved={m['bytes_reserved'] / 2**20:.1f} MiB")
        assert m["bytes_peak"] <= m["bytes_reserved"]
    # several jobs were on the device at once
    spans = sorted((m["t_start"], m["t_end"]) for m in mix)
    for (s0, e0), (s1, e1) in zip(spans, spans[1:]):
        overlap |= s1 < e0
    assert overlap


@pytest.mark.parametrize("prec", ["fp64", "bf16x3"])
def test_grouped_launches_equal_alone(prec):
    """Cross-job grouped GEMM launches (aq_server_set_grouping): same-key
    GEMMs of the four slots run as one kernel, and every job's energy and
    iteration count are bit-identical to the job run alone."""
    alone = run_alone(JOBS, prec)
    with aq.Server(slots=4) as s:
        s.set_grouping(2000)
        ids = [s.submit(seed, o, v, config=cfg(prec)) for seed, o, v in JOBS]
        mix = [s.wait(i) for i in ids]
        st = s.grouping_stats()
        extra = s.submit(1, 3, 12, config=cfg(prec))
        with pytest.raises(aq.AqError):
            s.set_grouping(0)          # refused while a job is queued or running
        assert s.wait(extra)["status"] == aq.JOB_DONE
        s.set_grouping(0)
        assert s.grouping_stats() == {"launches": 0, "gemms": 0, "max_group": 0, "solo": 0}
    print(f"grouping {prec}: {st}")
    assert st["launches"] > 0 and st["gemms"] > st["launches"] and st["max_group"] >= 2
    for a, m, (seed, o, v) in zip(alone, mix, JOBS):
        assert a["status"] == aq.JOB_DONE and m["status"] == aq.JOB_DONE, (a["error"], m["error"])
        assert a["converged"] and m["converged"]
        assert a["energy"] == m["energy"] and a["iterations"] == m["iterations"], (seed, a["energy"], m["energy"])


def test_served_job_equals_direct_solver():
    seed, o, v = 21, 5, 22
    with aq.Server(slots=2) as s:
        r = s.wait(s.submit(seed, o, v, config=cfg("fp64")))
    with aq.Context() as ctx:
        ints = aq.Integrals.synthetic(ctx, seed, o, v)
        cc = aq.RHFCCSD(ctx, ints, config=cfg("fp64"))
        d = cc.run()
        assert r["energy"] == d["energy"] and r["iterations"] == d["iterations"]
        assert abs(r["mp2"] - cc.mp2) == 0.0
        cc.close()
        ints.free()


def test_admission_budget_and_bypass():
    big, small = (31, 8, 48), (32, 3, 12)
    need_big = aq.job_bytes_estimate(big[1], big[2], 0, 8, "fp64")
    need_small = aq.job_bytes_estimate(small[1], small[2], 0, 8, "fp64")
    budget = need_big + need_small + (1 << 20)      # one big + one small fit, never two big
    # backfill allowed: the small job overtakes the waiting second big job
    with aq.Server(slots=3, hbm_budget=budget, max_bypass_s=1e9) as s:
        b1 = s.submit(*big, config=cfg("fp64"))
        b2 = s.submit(*big, config=cfg("fp64"))
        sm = s.submit(*small, config=cfg("fp64"))
        r1, r2, rs = s.wait(b1), s.wait(b2), s.wait(sm)
    assert r2["t_start"] >= r1["t_end"] - 1e-6          # the two big jobs never overlap
    assert rs["t_start"] < r2["t_start"]                 # the small one backfilled
    # no bypass: the small job waits behind the blocked big one
    with aq.Server(slots=3, hbm_budget=budget, max_bypass_s=0.0) as s:
        b1 = s.submit(*big, config=cfg("fp64"))
        b2 = s.submit(*big, config=cfg("fp64"))
        sm = s.submit(*small, config=cfg("fp64"))
        r1, r2, rs = s.wait(b1), s.wait(b2), s.wait(sm)
    assert r2["t_start"] >= r1["t_end"] - 1e-6
    assert rs["t_start"] >= r2["t_start"] - 1e-6
    for r in (r1, r2, rs):
        assert r["status"] == aq.JOB_DONE


def test_oversize_rejected_and_failures_isolated():
    with aq.Server(slots=2, hbm_budget=aq.job_bytes_estimate(4, 16) + (1 << 20)) as s:
        with pytest.raises(aq.AqError) as e:
            s.submit(1, 20, 140, config=cfg("fp64"))
        assert e.value.code == 3
        bad = s.submit(2, 3, 12, config=aq.default_config(diis_order=-1))   # invalid DIIS order
        good = s.submit(3, 3, 12, config=cfg("fp64"))
        rb, rg = s.wait(bad), s.wait(good)
        st = s.stats()
    assert rb["status"] == aq.JOB_FAILED and rb["error_code"] == 1 and "DIIS order" in rb["error"]
    assert rg["status"] == aq.JOB_DONE and rg["converged"]
    assert st["completed"] == 1 and st["failed"] == 1 and st["reserved"] == 0
