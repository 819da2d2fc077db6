"""CPU check of the packed particle-particle ladder (rhfccsd.cxx:303 as S+ + S-).

For V(ab,ef) = V(ba,fe) and Tau(ef,ij) = Tau(fe,ji):
  sum_ef V(ab,ef) Tau(ef,ij) = S+(ab,ij) + S-(ab,ij)
  S+ = sum_{e<=f} V+(ab,ef) T+(ef,ij),  V+ = V(ab,ef) + V(ab,fe) (V(ab,ee) on e=f),
  S- = sum_{e<f}  V-(ab,ef) T-(ef,ij),  V- = V(ab,ef) - V(ab,fe),
  T+- = (Tau(ef,ij) +- Tau(fe,ij)) / 2.
The packing below follows the device kernels' index conventions (column-major,
tri_le(a,b) = a + b(b+1)/2 for a <= b, tri_lt(a,b) = a + b(b-1)/2 for a < b,
the same inverse of tri_le) and records every index it writes or reads, so
the test also shows that the kernels' accesses stay inside their buffers.
The identity itself is checked against the plain einsum on random tensors
with the two symmetries (an exact rewrite: agreement to rounding)."""

import math

import numpy as np
import pytest


def tri_le(a, b):
    return a + b * (b + 1) // 2


def tri_lt(a, b):
    return a + b * (b - 1) // 2


def untri_le(r):
    # the kernels' inverse: float estimate, then integer correction
    b = int((math.sqrt(8.0 * r + 1.0) - 1.0) * 0.5)
    while b * (b + 1) // 2 > r:
        b -= 1
    while (b + 1) * (b + 2) // 2 <= r:
        b += 1
    return r - b * (b + 1) // 2, b


def pack_v(V, v):
    P, Q = v * (v + 1) // 2, v * (v - 1) // 2
    Vf = V.reshape(-1, order="F")
    Vp, Vm = np.full(P * P, np.nan), np.full(max(Q * Q, 1), np.nan)
    for c in range(P):
        e, f = untri_le(c)
        for r in range(P):
            a, b = untri_le(r)
            ix, iy = a + b * v + (e + f * v) * v * v, a + b * v + (f + e * v) * v * v
            assert 0 <= ix < v ** 4 and 0 <= iy < v ** 4
            x, y = Vf[ix], Vf[iy]
            k = r + c * P
            assert 0 <= k < P * P
            Vp[k] = x if e == f else x + y
            if a < b and e < f:
                k = tri_lt(a, b) + tri_lt(e, f) * Q
                assert 0 <= k < Q * Q
                Vm[k] = x - y
    return Vp.reshape(P, P, order="F"), Vm[:Q * Q].reshape(Q, Q, order="F")


def pack_t(T, v, o):
    """T+ [O+ x P] and T- [O- x Q]: the occupied pair index fastest."""
    P, Q, Op, Om = v * (v + 1) // 2, v * (v - 1) // 2, o * (o + 1) // 2, o * (o - 1) // 2
    Tf = T.reshape(-1, order="F")
    Tp, Tm = np.full(P * Op, np.nan), np.full(max(Q * Om, 1), np.nan)
    for c in range(Op):
        i, j = untri_le(c)
        base = (i + j * o) * v * v
        for r in range(P):
            e, f = untri_le(r)
            x, y = Tf[base + e + f * v], Tf[base + f + e * v]
            assert base + max(e + f * v, f + e * v) < v * v * o * o
            Tp[c + r * Op] = 0.5 * (x + y)
            if e < f and i < j:
                # the kernel's form of tri_lt(i,j) + tri_lt(e,f) * Om
                assert c - j == tri_lt(i, j) and r - f == tri_lt(e, f)
                k = (c - j) + (r - f) * Om
                assert 0 <= k < Q * Om
                Tm[k] = 0.5 * (x - y)
    return Tp.reshape(Op, P, order="F"), Tm[:Q * Om].reshape(Om, Q, order="F")


def unpack(Sp, Sm, v, o, alpha, Z):
    P, Q = v * (v + 1) // 2, v * (v - 1) // 2
    Spf, Smf = Sp.reshape(-1, order="F"), Sm.reshape(-1, order="F")
    Zf = Z.reshape(-1, order="F").copy()
    vv = v * v
    for c in range(o * o):
        i, j = c % o, c // o
        ilo, ihi = min(i, j), max(i, j)
        sij = 1.0 if i < j else -1.0
        for r in range(vv):
            a, b = r % v, r // v
            alo, ahi = min(a, b), max(a, b)
            k = tri_le(alo, ahi) + tri_le(ilo, ihi) * P
            assert 0 <= k < Spf.size
            s = Spf[k]
            if i != j and a != b:
                k = tri_lt(alo, ahi) + tri_lt(ilo, ihi) * Q
                assert 0 <= k < Smf.size
                s += (sij if a < b else -sij) * Smf[k]
            Zf[r + c * vv] += alpha * s
    return Zf.reshape(Z.shape, order="F")


def unpack_tiled(Sp, Sm, v, o, alpha, Z, T=4):
    """The device kernel's schedule: upper (a-tile <= b-tile) tiles of T x T
    (a,b) pairs and one pair i <= j per block; the (a,b) tile and its mirror
    (b,a), slabs (i,j) and (j,i).  Records that every Z element is updated
    exactly once and every read stays inside S+ / S-."""
    P, Q = v * (v + 1) // 2, v * (v - 1) // 2
    Spf, Smf = Sp.reshape(-1, order="F"), Sm.reshape(-1, order="F")
    Zf = Z.reshape(-1, order="F").copy()
    hits = np.zeros(Zf.size, dtype=int)
    vv, nt = v * v, (v + T - 1) // T
    for bx in range(nt * (nt + 1) // 2):
        ta, tb = untri_le(bx)
        diag = ta == tb
        for c in range(o * (o + 1) // 2):
            i, j = untri_le(c)
            lp, lm = np.zeros((T, T)), np.zeros((T, T))
            for al in range(T):
                for bl in range(T):
                    a, b = ta * T + al, tb * T + bl
                    if a >= v or b >= v:
                        continue
                    lo, hi = (a, b) if (not diag or al <= bl) else (b, a)
                    k = tri_le(lo, hi) + c * P
                    assert 0 <= k < Spf.size
                    lp[al, bl] = Spf[k]
                    if i < j and lo < hi:
                        k = tri_lt(lo, hi) + (c - j) * Q
                        assert c - j == tri_lt(i, j) and 0 <= k < Smf.size
                        lm[al, bl] = Smf[k] if lo == a else -Smf[k]
            for al in range(T):
                for bl in range(T):
                    a, b = ta * T + al, tb * T + bl
                    if a >= v or b >= v:
                        continue
                    sp, sm = lp[al, bl], lm[al, bl]
                    writes = [(a + b * v + (i + j * o) * vv, sp + sm)]
                    if i != j:
                        writes.append((a + b * v + (j + i * o) * vv, sp - sm))
                    if not diag:
                        writes.append((b + a * v + (i + j * o) * vv, sp - sm))
                        if i != j:
                            writes.append((b + a * v + (j + i * o) * vv, sp + sm))
                    for k, s in writes:
                        Zf[k] += alpha * s
                        hits[k] += 1
    assert (hits == 1).all()
    return Zf.reshape(Z.shape, order="F")


def symmetric_inputs(rng, v, o):
    V = rng.standard_normal((v, v, v, v))
    V = V + V.transpose(1, 0, 3, 2)                  # V(ab,ef) = V(ba,fe)
    Tau = rng.standard_normal((v, v, o, o))
    Tau = Tau + Tau.transpose(1, 0, 3, 2)            # Tau(ef,ij) = Tau(fe,ji)
    return V, Tau


@pytest.mark.parametrize("v,o", [(2, 2), (3, 2), (5, 3), (7, 4)])
def test_packed_ladder_equals_full_contraction(v, o):
    rng = np.random.default_rng(100 * v + o)
    V, Tau = symmetric_inputs(rng, v, o)
    Vp, Vm = pack_v(V, v)
    Tp, Tm = pack_t(Tau, v, o)
    assert not np.isnan(Vp).any() and not np.isnan(Vm).any()
    assert not np.isnan(Tp).any() and not np.isnan(Tm).any()
    Z0 = rng.standard_normal((v, v, o, o))
    Z = unpack(Vp @ Tp.T, Vm @ Tm.T, v, o, 0.5, Z0)
    ref = Z0 + 0.5 * np.einsum("abef,efij->abij", V, Tau)
    np.testing.assert_allclose(Z, ref, rtol=1e-12, atol=1e-12)
    # the tiled schedule of the device kernel (ragged last tile when T does not divide v)
    for T in (2, 3, 4):
        np.testing.assert_array_equal(unpack_tiled(Vp @ Tp.T, Vm @ Tm.T, v, o, 0.5, Z0, T), Z)


def test_packed_ladder_flop_ratio():
    # two packed GEMMs against the full v^4 o^2 contraction
    v, o = 288, 48
    P, Q, Op, Om = v * (v + 1) // 2, v * (v - 1) // 2, o * (o + 1) // 2, o * (o - 1) // 2
    full = 2.0 * v ** 4 * o ** 2
    packed = 2.0 * P * P * Op + 2.0 * Q * Q * Om
    assert 0.24 < packed / full < 0.26


def test_asymmetric_input_breaks_the_identity():
    """The rewrite needs V(ab,ef) = V(ba,fe): without it the two sides differ,
    which is why the solver checks the symmetry before packing."""
    rng = np.random.default_rng(7)
    v, o = 4, 3
    V = rng.standard_normal((v, v, v, v))
    _, Tau = symmetric_inputs(rng, v, o)
    Vp, Vm = pack_v(V, v)
    Tp, Tm = pack_t(Tau, v, o)
    Z = unpack(Vp @ Tp.T, Vm @ Tm.T, v, o, 1.0, np.zeros((v, v, o, o)))
    ref = np.einsum("abef,efij->abij", V, Tau)
    assert np.abs(Z - ref).max() > 1e-3


@pytest.mark.parametrize("v", [2, 3, 17, 288, 1000])
def test_untri_le_inverts_tri_le(v):
    P = v * (v + 1) // 2
    rs = range(P) if P < 5000 else list(range(2000)) + list(range(P - 2000, P))
    for r in rs:
        a, b = untri_le(r)
        assert 0 <= a <= b < v and tri_le(a, b) == r
