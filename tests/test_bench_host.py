"""bench.py's host logic (no GPU): the FLOP model against the oracle's own
contractions, the packed-ladder FLOP count, the serving job stream, and the
refusal of more than one GPU, the per-solve split of a kernel trace."""

import json
import os
import subprocess
import sys

import numpy as np
import pytest

import bench
import oracle.rhfccsd as R
from oracle import synth

HERE = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("o,v", [(2, 3), (3, 5), (4, 7), (5, 6)])
def test_flop_model_equals_the_oracles_contractions(o, v, monkeypatch):
    """flops_per_iter(o, v) is exactly the FLOPs of the contractions one
    oracle iteration performs (2 x the product of every label's extent per
    call), so a wrong power or coefficient of any term fails at one of the
    sizes."""
    count = [0]
    mult = R.mult

    def counting(alpha, A, idxA, B, idxB, beta, C, idxC):
        ext = {}
        for arr, idx in ((A, idxA), (B, idxB), (C, idxC)):
            ext.update(zip(idx, arr.shape))
        count[0] += 2 * int(np.prod([ext[c] for c in set(idxA) | set(idxB)]))
        return mult(alpha, A, idxA, B, idxB, beta, C, idxC)

    monkeypatch.setattr(R, "mult", counting)
    f, V, _, _ = synth.synthetic_system(1, o, v)
    cc = R.RHFCCSD(f, V)
    count[0] = 0
    cc.iterate()
    assert count[0] == bench.flops_per_iter(o, v)


def test_packed_ladder_flops():
    o, v = 48, 288
    F = bench.flops_per_iter(o, v)
    assert bench.flops_per_iter_executed(o, v, False) == F
    P, Q = v * (v + 1) // 2, v * (v - 1) // 2
    Op, Om = o * (o + 1) // 2, o * (o - 1) // 2
    ex = bench.flops_per_iter_executed(o, v, True)
    assert F - ex == 2 * o**2 * v**4 - 2 * P * P * Op - 2 * Q * Q * Om
    # the packed ladder costs about a quarter of the full one
    assert 0.24 < (2 * P * P * Op + 2 * Q * Q * Om) / (2 * o**2 * v**4) < 0.26


def test_serving_job_stream():
    a = bench.job_stream(4, 256, 4, 32, 4.0)
    assert a == bench.job_stream(4, 256, 4, 32, 4.0) and a != bench.job_stream(5, 256, 4, 32, 4.0)
    o = np.array([x[1] for x in a])
    v = np.array([x[2] for x in a])
    t = np.array([x[3] for x in a])
    assert o.min() >= 4 and o.max() <= 32 and np.all(v >= 4 * o) and np.all(v <= 8 * o)
    assert t[0] == 0.0 and np.all(np.diff(t) >= 0)
    assert 40.0 < t[-1] < 90.0            # 255 gaps of mean 1/4 s
    assert len({x[0] for x in a}) == 256  # distinct seeds


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(HERE, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=HERE)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]


def test_trace_split_cuts_at_solver_setups(tmp_path, capsys):
    """bench.py --trace-split: the dominant kernel's launches are grouped by
    the solver setups (two k_pair_asym launches each) of a kernel trace."""
    nt = "void aqamd::k_gemm<double, 128, 128, 2, 2, false, true, 16, 2>(aqamd::GemmArgs)"
    rows, t = [], 0

    def add(name, ms):
        nonlocal t
        rows.append((name, t, t + int(ms * 1e6)))
        t += int(ms * 1e6) + 1000

    for seg, durations in enumerate(([1.0], [2.0, 2.0], [3.0, 5.0], [7.0])):
        add("aqamd::k_pair_asym(double const*, long, long, double*)", 0.1)
        add("aqamd::k_pair_asym(double const*, long, long, double*)", 0.1)
        for ms in durations:
            add(nt, ms)
            add("void aqamd::k_strided<aqamd::SumOut>(...)", 0.5)
    path = tmp_path / "run_kernel_trace.csv"
    with open(path, "w") as fh:
        fh.write('"Kind","Kernel_Name","Start_Timestamp","End_Timestamp"\n')
        for name, a, b in rows:
            fh.write(f'"KERNEL_DISPATCH","{name}",{a},{b}\n')
    bench.trace_split(str(path))
    out = json.loads(capsys.readouterr().out.strip().splitlines()[-1])["per_solve"]
    assert out["warm-up fp64"] == {"launches": 1, "mean_ms": pytest.approx(1.0, rel=1e-6)}
    assert out["warm-up bf16x3"]["launches"] == 2
    assert out["mixed solve"] == {"launches": 2, "mean_ms": pytest.approx(4.0, rel=1e-6)}
    assert out["fp64 solve"] == {"launches": 1, "mean_ms": pytest.approx(7.0, rel=1e-6)}
