"""Write the committed golden vectors of the oracle's RHF-CCSD on the seeded
synthetic construction (no reference code is run: the oracle is the numpy
restatement in oracle/).

    python tests/golden/make_fixtures.py [--only NAME]

Fixtures (tests/golden/<name>.npz):
  synth_o20_v140  o=20, v=140, nQ=480, seed 1: the FP64 parity workload
  synth_o6_v18    o=6,  v=18,  nQ=72,  seed 5: small regression case
Each stores the per-iteration energy/convergence trajectory, the MP2 energy,
the converged T1, and T2 at 20000 fixed column-major keys plus its norms.
"""

import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import rhfccsd, synth  # noqa: E402

CASES = {
    "synth_o20_v140": dict(seed=1, o=20, v=140, nQ=480, convergence=1e-9, order=5, maxiter=50),
    "synth_o6_v18": dict(seed=5, o=6, v=18, nQ=72, convergence=1e-10, order=8, maxiter=60),
}

NKEYS = 20000


def sample_keys(size, seed=12345):
    rng = np.random.default_rng(seed)
    k = rng.choice(size, size=min(NKEYS, size), replace=False)
    return np.sort(k).astype(np.int64)


def make(name, c):
    t0 = time.time()
    f, V, B, _ = synth.synthetic_system(c["seed"], c["o"], c["v"], c["nQ"])
    cfg = rhfccsd.default_config()
    cfg["convergence"] = c["convergence"]
    cfg["max_iterations"] = c["maxiter"]
    cfg["diis"]["order"] = c["order"]
    cc = rhfccsd.RHFCCSD(f, V, cfg)

    def log(s):
        e, cv = s.history[-1]
        print(f"{name} iter {len(s.history)} E={e:.14f} conv={cv:.3e} t={time.time() - t0:.1f}s", flush=True)

    ok = cc.run(callback=log)
    T2 = cc.T2.ravel(order="F")
    keys = sample_keys(T2.size)
    out = os.path.join(HERE, name + ".npz")
    np.savez_compressed(
        out, seed=c["seed"], o=c["o"], v=c["v"], nQ=c["nQ"], sigma=synth.default_sigma(c["v"], c["nQ"]),
        convergence=c["convergence"], order=c["order"],
        maxiter=c["maxiter"], converged=ok, mp2=cc.mp2, energy=np.array([h[0] for h in cc.history]),
        conv=np.array([h[1] for h in cc.history]), T1=cc.T1, t2_keys=keys, t2_vals=T2[keys],
        t2_norm2=np.sqrt(np.sum(T2 * T2)), t2_norm1=np.sum(np.abs(T2)), t2_max=np.max(np.abs(T2)))
    print(f"wrote {out} ({time.time() - t0:.1f}s)")


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", default=None)
    a = ap.parse_args()
    for n, c in CASES.items():
        if a.only is None or a.only == n:
            make(n, c)
