"""Seeded synthetic closed-shell integrals (oracle side; test infrastructure only).

Construction (the synthetic CCSD workloads o=20/v=140, o=48/v=288 and the
serving stream):
  * occupied orbital energies ~ U[-1.5,-0.3] Eh, virtual ~ U[0.2,5.0] Eh,
    each sorted ascending; Fock f = diag(eps), so f_ai = 0;
  * three-index factors B^Q_pq = B^Q_qp ~ N(0,1)*sigma for p<=q over the o+v
    orbitals (occupied first), Q = 0..nQ-1; sigma defaults to
    0.05*sqrt(150/(v*sqrt(nQ))), which keeps the effective coupling
    v*sigma^2*sqrt(nQ) of the v=18/nQ=72 system (sigma ~ 0.05) at every size;
    a fixed 0.05 makes CCSD diverge for v*sqrt(nQ) >> 150;
  * Mulliken (pq|rs) = sum_Q B^Q_pq B^Q_rs (8-fold symmetric, PSD).

The random stream is counter based (splitmix64 finaliser + Box-Muller), so the
engine's host generator and this restatement draw identical uniforms; the
normals agree to the last ulp or two (libm vs numpy transcendental rounding).

``blocks_from_B`` maps (pq|rs) onto the seven Dirac blocks in the layout of
rhfaomoints.cxx:58-65 (transcription <pq|rs> = (pr|qs), aomoints.cxx:806).
"""

import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
STREAM = np.uint64(0xD1B54A32D192ED03)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)

STREAM_EPS_OCC = 1
STREAM_EPS_VIR = 2
STREAM_B = 3


def mix64(z):
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * M1
        z = (z ^ (z >> np.uint64(27))) * M2
    return z ^ (z >> np.uint64(31))


def u01(seed, stream, counters):
    """Uniform doubles in [0,1) with 53 random bits, one per counter."""
    counters = np.asarray(counters, dtype=np.uint64)
    with np.errstate(over="ignore"):
        key = mix64(np.uint64(seed) ^ (np.uint64(stream) * STREAM))
        x = mix64(key + (counters + np.uint64(1)) * GOLDEN)
    return (x >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def normals(seed, stream, counters):
    """Box-Muller: normal k uses the uniform pair (2*(k//2), 2*(k//2)+1)."""
    counters = np.asarray(counters, dtype=np.uint64)
    j = counters >> np.uint64(1)
    u1 = u01(seed, stream, j * np.uint64(2))
    u2 = u01(seed, stream, j * np.uint64(2) + np.uint64(1))
    r = np.sqrt(-2.0 * np.log1p(-u1))
    th = 2.0 * np.pi * u2
    odd = (counters & np.uint64(1)).astype(bool)
    return np.where(odd, r * np.sin(th), r * np.cos(th))


def orbital_energies(seed, o, v):
    eo = np.sort(-1.5 + 1.2 * u01(seed, STREAM_EPS_OCC, np.arange(o)))
    ev = np.sort(0.2 + 4.8 * u01(seed, STREAM_EPS_VIR, np.arange(v)))
    return eo, ev


def default_sigma(v, nQ):
    return 0.05 * np.sqrt(150.0 / (v * np.sqrt(nQ)))


def three_index(seed, o, v, nQ, sigma=0.0):
    """B[Q,p,q] (symmetric in p,q).  Counter of (Q, p<=q) is
    Q*npair + p*n - p*(p-1)/2 + (q-p)."""
    sg = sigma if sigma > 0 else default_sigma(v, nQ)
    n = o + v
    npair = n * (n + 1) // 2
    p, q = np.triu_indices(n)
    t = p * n - p * (p - 1) // 2 + (q - p)
    B = np.empty((nQ, n, n))
    for Q in range(nQ):
        z = sg * normals(seed, STREAM_B, np.uint64(Q * npair) + t.astype(np.uint64))
        B[Q, p, q] = z
        B[Q, q, p] = z
    return B


def fock(eo, ev):
    o, v = len(eo), len(ev)
    return {"fAI": np.zeros((v, o)), "fAB": np.diag(ev), "fIJ": np.diag(eo)}


def blocks_from_B(B, o):
    """Seven Dirac blocks from B (Mulliken (pq|rs) = sum_Q B_pq B_rs)."""
    Bvv = B[:, o:, o:]
    Bvo = B[:, o:, :o]
    Boo = B[:, :o, :o]
    es = lambda s, X, Y: np.einsum(s, X, Y, optimize=True)
    return {
        "VABCD": es("Qac,Qbd->abcd", Bvv, Bvv),   # (ac|bd)
        "VABCI": es("Qac,Qbi->abci", Bvv, Bvo),   # (ac|bi)
        "VABIJ": es("Qai,Qbj->abij", Bvo, Bvo),   # (ai|bj)
        "VAIBJ": es("Qab,Qij->aibj", Bvv, Boo),   # (ab|ij)
        "VAIJB": es("Qaj,Qbi->aijb", Bvo, Bvo),   # (aj|ib)
        "VAIJK": es("Qaj,Qik->aijk", Bvo, Boo),   # (aj|ik)
        "VIJKL": es("Qik,Qjl->ijkl", Boo, Boo),   # (ik|jl)
    }


def blocks_from_eri(eri, o):
    """Seven Dirac blocks from a full Mulliken tensor eri[p,q,r,s] = (pq|rs)."""
    O = slice(0, o)
    Vs = slice(o, eri.shape[0])
    # <pq|rs> = (pr|qs)
    g = lambda s1, s2, s3, s4: np.ascontiguousarray(eri[s1, s3, s2, s4].transpose(0, 2, 1, 3))
    return {
        "VABCD": g(Vs, Vs, Vs, Vs),
        "VABCI": g(Vs, Vs, Vs, O),
        "VABIJ": g(Vs, Vs, O, O),
        "VAIBJ": g(Vs, O, Vs, O),
        "VAIJB": g(Vs, O, O, Vs),
        "VAIJK": g(Vs, O, O, O),
        "VIJKL": g(O, O, O, O),
    }


def synthetic_system(seed, o, v, nQ=None, sigma=0.0):
    """(f, V, B, eps) of the synthetic construction; nQ defaults to 3(o+v)."""
    nQ = 3 * (o + v) if nQ is None else nQ
    eo, ev = orbital_energies(seed, o, v)
    B = three_index(seed, o, v, nQ, sigma)
    return fock(eo, ev), blocks_from_B(B, o), B, (eo, ev)
