"""Pins the oracle's RHF-CCSD restatement (rhfccsd.cxx:45-327), DIIS
(diis.hpp:71-227) and driver (iterative.hpp:126-188) against checks that do
not reuse its formulas: an independent spin-orbital CCSD, textbook closed-shell
MP2, exact symmetries, size consistency and the bordered Pulay system."""

import numpy as np
import pytest

from oracle import rhfccsd, synth
from oracle.spinorbital_ccsd import SpinOrbitalCCSD, spin_orbital_system


def small_system(seed, o, v, nQ=None, offdiag=0.0):
    """Synthetic integrals; optional symmetric off-diagonal Fock perturbation
    (non-canonical orbitals, f_ai != 0) to exercise every f term."""
    f, V, B, (eo, ev) = synth.synthetic_system(seed, o, v, nQ)
    n = o + v
    F = np.diag(np.concatenate([eo, ev]))
    if offdiag:
        rng = np.random.default_rng(seed)
        P = rng.standard_normal((n, n)) * offdiag
        P = P + P.T
        np.fill_diagonal(P, 0.0)
        F = F + P
    f = {"fAI": F[o:, :o].copy(), "fAB": F[o:, o:].copy(), "fIJ": F[:o, :o].copy()}
    eri = np.einsum("Qpq,Qrs->pqrs", B, B)
    return f, V, F, eri


def tight_config(order=8):
    cfg = rhfccsd.default_config()
    cfg["convergence"] = 1e-12
    cfg["max_iterations"] = 200
    cfg["diis"]["order"] = order
    return cfg


@pytest.mark.parametrize("offdiag", [0.0, 0.03])
def test_rhf_matches_independent_spin_orbital_ccsd(offdiag):
    o, v = 3, 6
    f, V, F, eri = small_system(11, o, v, nQ=14, offdiag=offdiag)
    cc = rhfccsd.RHFCCSD(f, V, tight_config())
    assert cc.run()
    Fso, Wso = spin_orbital_system(F, eri, o)
    so = SpinOrbitalCCSD(Fso, Wso, 2 * o)
    e_so, _ = so.solve(tol=1e-12)
    assert abs(cc.energy - e_so) < 1e-10
    # amplitudes: t_ai = t[i alpha, a alpha]; T2[a,b,i,j] = t[i alpha, j beta, a alpha, b beta]
    ia, ib = 2 * np.arange(o), 2 * np.arange(o) + 1
    aa, ab = 2 * np.arange(v), 2 * np.arange(v) + 1
    np.testing.assert_allclose(cc.T1, so.t1[np.ix_(ia, aa)].T, atol=1e-9)
    t2so = so.t2[np.ix_(ia, ib, aa, ab)].transpose(2, 3, 0, 1)
    np.testing.assert_allclose(cc.T2, t2so, atol=1e-9)


def test_spin_orbital_reference_detects_a_planted_sign_error():
    """The cross-check has teeth: flipping the sign of one RHF term (the ring
    term at rhfccsd.cxx:307) on the oracle copy moves the energy far beyond
    the tolerance used above."""
    o, v = 3, 6
    f, V, F, eri = small_system(11, o, v, nQ=14)

    class Broken(rhfccsd.RHFCCSD):
        def iterate(self):
            orig = rhfccsd.mult

            def patched(alpha, A, ia, B, ib, beta, C, ic):
                if (ia, ib, ic) == ("amej", "ebim", "abij"):
                    alpha = -alpha
                return orig(alpha, A, ia, B, ib, beta, C, ic)
            rhfccsd.mult = patched
            try:
                super().iterate()
            finally:
                rhfccsd.mult = orig

    good = rhfccsd.RHFCCSD(f, V, tight_config())
    good.run()
    bad = Broken(f, V, tight_config())
    bad.run()
    assert abs(good.energy - bad.energy) > 1e-6


def test_mp2_closed_form():
    """E_MP2 (rhfccsd.cxx:94) equals the textbook closed-shell expression
    built directly from Mulliken integrals, with a non-canonical f (t_ai != 0)."""
    o, v = 3, 5
    f, V, F, eri = small_system(5, o, v, nQ=10, offdiag=0.02)
    cc = rhfccsd.RHFCCSD(f, V)
    e = 0.0
    fo, fv = np.diag(F)[:o], np.diag(F)[o:]
    t1 = np.zeros((v, o))
    for a in range(v):
        for i in range(o):
            t1[a, i] = F[o + a, i] / (fo[i] - fv[a])
            e += 2 * F[o + a, i] * t1[a, i]
    for a in range(v):
        for b in range(v):
            for i in range(o):
                for j in range(o):
                    K = eri[o + a, i, o + b, j]          # <ab|ij> = (ai|bj)
                    Kx = eri[o + a, j, o + b, i]         # <ab|ji> = (aj|bi)
                    D = fo[i] + fo[j] - fv[a] - fv[b]
                    e += (2 * K - Kx) * (K / D + t1[a, i] * t1[b, j])
    assert abs(cc.mp2 - e) < 1e-13 * max(1.0, abs(e))


def test_blocks_layout_matches_mulliken():
    o, v = 2, 3
    _, V, B, _ = synth.synthetic_system(3, o, v, 7)
    eri = np.einsum("Qpq,Qrs->pqrs", B, B)
    a, b, c, d, i, j, k, l = 1, 2, 0, 1, 0, 1, 1, 0
    A = lambda x: o + x
    assert np.isclose(V["VABCD"][a, b, c, d], eri[A(a), A(c), A(b), A(d)])
    assert np.isclose(V["VABCI"][a, b, c, i], eri[A(a), A(c), A(b), i])
    assert np.isclose(V["VABIJ"][a, b, i, j], eri[A(a), i, A(b), j])
    assert np.isclose(V["VAIBJ"][a, i, b, j], eri[A(a), A(b), i, j])
    assert np.isclose(V["VAIJB"][a, i, j, b], eri[A(a), j, i, A(b)])
    assert np.isclose(V["VAIJK"][a, i, j, k], eri[A(a), j, i, k])
    assert np.isclose(V["VIJKL"][i, j, k, l], eri[i, k, j, l])
    Vx = synth.blocks_from_eri(eri, o)
    for name in Vx:
        np.testing.assert_allclose(Vx[name], V[name], atol=1e-15)


def test_t2_and_z2_symmetric():
    """Z2 leaves the aliased sum (rhfccsd.cxx:309) exactly symmetric under
    (ai)<->(bj); the weight's denominator sum ((dA+dA)+dI)+dI is not
    associative, so after the update symmetry holds to an ulp."""
    o, v = 3, 5
    f, V, _, _ = small_system(2, o, v, nQ=12, offdiag=0.02)
    cc = rhfccsd.RHFCCSD(f, V)
    for _ in range(4):
        cc.iterate()
        np.testing.assert_allclose(cc.Z2, cc.Z2.transpose(1, 0, 3, 2), rtol=1e-14, atol=1e-18)
        np.testing.assert_allclose(cc.T2, cc.T2.transpose(1, 0, 3, 2), rtol=1e-14, atol=1e-18)


def test_size_consistency():
    """Two non-interacting copies: E(AB) = E(A) + E(B)."""
    o1, v1, o2, v2 = 2, 4, 1, 3
    f1, V1, F1, eri1 = small_system(21, o1, v1, nQ=9)
    f2, V2, F2, eri2 = small_system(22, o2, v2, nQ=6)
    n1, n2 = o1 + v1, o2 + v2
    occ = list(range(o1)) + list(range(n1, n1 + o2))
    vir = list(range(o1, n1)) + list(range(n1 + o2, n1 + n2))
    perm = occ + vir
    N = n1 + n2
    F = np.zeros((N, N))
    F[:n1, :n1] = F1
    F[n1:, n1:] = F2
    eri = np.zeros((N, N, N, N))
    eri[:n1, :n1, :n1, :n1] = eri1
    eri[n1:, n1:, n1:, n1:] = eri2
    F = F[np.ix_(perm, perm)]
    eri = eri[np.ix_(perm, perm, perm, perm)]
    o = o1 + o2
    f = {"fAI": F[o:, :o], "fAB": F[o:, o:], "fIJ": F[:o, :o]}
    V = synth.blocks_from_eri(eri, o)
    cfg = tight_config()
    e = []
    for ff, VV in ((f1, V1), (f2, V2), (f, V)):
        cc = rhfccsd.RHFCCSD(ff, VV, tight_config())
        assert cc.run()
        e.append(cc.energy)
    assert abs(e[2] - (e[0] + e[1])) < 1e-10
    del cfg


def test_diis_coefficients_solve_bordered_pulay_system():
    rng = np.random.default_rng(0)
    d = rhfccsd.DIIS({"order": 4, "start": 1, "damping": 0.0}, 1, 1)
    xs, dxs = [], []
    for it in range(7):
        x = [rng.standard_normal(5)]
        dx = [rng.standard_normal(5) * 0.5 ** it]
        xs.append(x[0].copy())
        dxs.append(dx[0].copy())
        xin, dxin = [x[0].copy()], [dx[0].copy()]
        d.extrapolate(xin, dxin)
        k = min(it + 1, 4)
        if k == 1:
            np.testing.assert_array_equal(xin[0], x[0])
            continue
        hist_dx = dxs[-k:][::-1]   # slot 0 = newest
        hist_x = xs[-k:][::-1]
        Bm = np.zeros((k + 1, k + 1))
        for a in range(k):
            for b in range(k):
                Bm[a, b] = hist_dx[a] @ hist_dx[b]
        Bm[:k, k] = Bm[k, :k] = -1.0
        rhs = np.zeros(k + 1)
        rhs[k] = -1.0
        c = np.linalg.solve(Bm, rhs)
        np.testing.assert_allclose(d.c[:k + 1], c, rtol=1e-9, atol=1e-12)
        np.testing.assert_allclose(xin[0], sum(ci * xi for ci, xi in zip(c[:k], hist_x)), rtol=1e-9, atol=1e-12)
        np.testing.assert_allclose(dxin[0], sum(ci * xi for ci, xi in zip(c[:k], hist_dx)), rtol=1e-9, atol=1e-12)


def test_diis_singular_b_takes_epsilon_retry():
    """Identical residuals make B exactly singular; the reference retries with
    eps*B[n-1][n-1] on the diagonal (diis.hpp:177-185)."""
    d = rhfccsd.DIIS({"order": 3, "start": 1, "damping": 0.0}, 1, 1)
    r = np.array([1.0, 2.0, 3.0])
    d.extrapolate([np.array([1.0, 1.0, 1.0])], [r.copy()])
    x, dx = [np.array([3.0, 3.0, 3.0])], [r.copy()]
    d.extrapolate(x, dx)
    assert d.used_eps_retry
    # the retried system is perturbed by ~1 ulp of B, so the split between
    # the two coefficients is set by roundoff; the unity constraint and the
    # extrapolation itself are exact properties
    assert abs(d.c[0] + d.c[1] - 1.0) < 1e-12
    np.testing.assert_allclose(x[0], d.c[0] * 3.0 + d.c[1] * 1.0, rtol=1e-14)
    np.testing.assert_allclose(dx[0], r, rtol=1e-12)


def test_diis_start_delays_and_damps():
    d = rhfccsd.DIIS({"order": 3, "start": 3, "damping": 0.25}, 1, 1)
    x1 = [np.array([1.0, 0.0])]
    d.extrapolate(x1, [np.array([1.0, 0.0])])
    x2 = [np.array([0.0, 2.0])]
    d.extrapolate(x2, [np.array([0.0, 1.0])])
    np.testing.assert_array_equal(x2[0], [0.0, 2.0])          # skipped (start 3 -> 2 > 1)
    # the damping rewrites the stored history slot only (diis.hpp:155-161)
    np.testing.assert_allclose(d.old_x[0][0], (0.25 - 1) * np.array([0.0, 2.0]) + 0.25 * np.array([1.0, 0.0]))
    x3 = [np.array([1.0, 1.0])]
    d.extrapolate(x3, [np.array([0.5, 0.5])])                 # start 2 -> 1: extrapolates
    assert not np.array_equal(x3[0], [1.0, 1.0])


def test_iterative_stop_rule_and_nonconvergence():
    o, v = 2, 4
    f, V, _, _ = small_system(4, o, v, nQ=8)
    cfg = rhfccsd.default_config()
    cfg["max_iterations"] = 3
    cc = rhfccsd.RHFCCSD(f, V, cfg)
    assert cc.run() is False             # logged, state kept (iterative.hpp:160-165)
    assert len(cc.history) == 3
    cfg2 = rhfccsd.default_config()
    cc2 = rhfccsd.RHFCCSD(f, V, cfg2)
    assert cc2.run()
    assert cc2.history[-1][1] < cfg2["convergence"]
    assert all(h[1] >= cfg2["convergence"] for h in cc2.history[:-1])


# ------------------------------------------------------------ Hbar products

HBAR = ("FAE", "FMI", "FME", "WMNIJ", "WMNEJ", "WAMIJ", "WAMEI", "WAMIE", "WABEJ")


def test_hbar_reduces_to_bare_integrals_at_t_zero():
    """rhfccsd.cxx:108-188 with T1 = T2 = 0: every product is its bare block."""
    f, V, _, _ = small_system(4, 3, 7, offdiag=0.05)
    cc = rhfccsd.RHFCCSD(f, V)
    cc.T1[...] = 0.0
    cc.T2[...] = 0.0
    cc.hbar(with_wabej=True)
    np.testing.assert_array_equal(cc.FAE, f["fAB"])
    np.testing.assert_array_equal(cc.FMI, f["fIJ"])
    np.testing.assert_array_equal(cc.FME, f["fAI"].T)
    np.testing.assert_allclose(cc.WMNIJ, V["VIJKL"], atol=1e-15)          # (<mn|ij> + <nm|ji>)/2
    np.testing.assert_array_equal(cc.WMNEJ, V["VAIJK"].transpose(2, 3, 0, 1))
    np.testing.assert_array_equal(cc.WAMIJ, V["VAIJK"])
    np.testing.assert_allclose(cc.WAMIE, 2 * V["VAIJB"] - V["VAIBJ"].transpose(0, 1, 3, 2), atol=1e-15)
    np.testing.assert_array_equal(cc.WAMEI, V["VAIBJ"])
    np.testing.assert_array_equal(cc.WABEJ, V["VABCI"])


def test_hbar_fme_is_half_the_energy_gradient():
    """E = 2 f.T1 + VABIJSA.(T2 + T1 T1) is quadratic in T1 and
    dE/dT1[e,m] = 2 FME[m,e] exactly; central differences check it."""
    f, V, _, _ = small_system(5, 3, 6, offdiag=0.05)
    cc = rhfccsd.RHFCCSD(f, V)
    rng = np.random.default_rng(1)
    cc.T1[...] = rng.standard_normal(cc.T1.shape) * 0.05
    cc.T2[...] = rng.standard_normal(cc.T2.shape) * 0.02
    cc.hbar()

    def energy(T1):
        tau = cc.T2 + np.einsum("ai,bj->abij", T1, T1)
        return 2 * np.sum(f["fAI"] * T1) + np.sum(cc.VABIJSA * tau)

    h = 1e-3
    grad = np.zeros_like(cc.T1)
    for e in range(cc.nv):
        for m in range(cc.no):
            d = np.zeros_like(cc.T1)
            d[e, m] = h
            grad[e, m] = (energy(cc.T1 + d) - energy(cc.T1 - d)) / (2 * h)
    np.testing.assert_allclose(grad, 2 * cc.FME.T, rtol=0, atol=1e-11)


def test_hbar_shapes_and_wabej_order():
    """WABEJ reads WAMIE/WAMEI before their final T1 terms (rhfccsd.cxx:167-186):
    forming it must not change the other products."""
    f, V, _, _ = small_system(6, 3, 5, offdiag=0.05)
    a = rhfccsd.RHFCCSD(f, V)
    a.run(max_iterations=5)
    b = rhfccsd.RHFCCSD(f, V)
    b.run(max_iterations=5)
    a.hbar(False)
    b.hbar(True)
    for k in HBAR[:-1]:
        np.testing.assert_array_equal(getattr(a, k), getattr(b, k))
    assert b.WABEJ.shape == (5, 5, 5, 3)


def test_hbar_products_pinned_by_their_eom_consumer():
    """At converged amplitudes the EOM matrix minus E_cc is the Jacobian of the
    CC residual: for any R, the singlet sigma vector that the reference's
    EOM-EE-CCSD builds from the Hbar products (rhfeomeeccsd.cxx:381-424) must
    equal the derivative of the residual Z(T) (rhfccsd.cxx:231-309) along R.
    Both are polynomials of degree <= 4 in the amplitudes, so the five-point
    central difference is exact up to rounding.  This pins every product the
    consumer reads: FAE, FMI, FME, WMNIJ, WMNEJ, WAMIJ, WAMEI, WAMIE, WABEJ."""
    f, V, _, _ = synth.synthetic_system(17, 3, 9)
    cfg = rhfccsd.default_config()
    cfg["convergence"] = 1e-13
    cfg["max_iterations"] = 200
    cfg["diis"]["order"] = 8
    cc = rhfccsd.RHFCCSD(f, V, cfg)
    assert cc.run()
    cc.hbar(with_wabej=True)
    rng = np.random.default_rng(4)
    v, o = cc.T1.shape
    for trial in range(3):
        R1 = rng.standard_normal((v, o))
        R2 = rng.standard_normal((v, v, o, o))
        R2 = R2 + R2.transpose(1, 0, 3, 2)                  # the symmetry of T2
        s1, s2 = rhfccsd.eom_singlet_sigma(cc, R1, R2)
        h = 1e-3
        Z = {k: cc.residual_at(cc.T1 + k * h * R1, cc.T2 + k * h * R2) for k in (-2, -1, 1, 2)}
        d1 = (-Z[2][0] + 8 * Z[1][0] - 8 * Z[-1][0] + Z[-2][0]) / (12 * h)
        d2 = (-Z[2][1] + 8 * Z[1][1] - 8 * Z[-1][1] + Z[-2][1]) / (12 * h)
        scale = max(np.abs(s1).max(), np.abs(s2).max())
        # measured: ~2e-15 relative (rounding of the difference quotient)
        assert np.abs(s1 - d1).max() <= 1e-12 * scale, (trial, np.abs(s1 - d1).max(), scale)
        assert np.abs(s2 - d2).max() <= 1e-12 * scale, (trial, np.abs(s2 - d2).max(), scale)
    # every product matters: a relative error of 1e-6 in any one of them is detected
    for name in ("FAE", "FMI", "FME", "WMNIJ", "WMNEJ", "WAMIJ", "WAMEI", "WAMIE", "WABEJ"):
        x = getattr(cc, name)
        keep = x.copy()
        x *= 1.0 + 1e-6
        t1, t2 = rhfccsd.eom_singlet_sigma(cc, R1, R2)
        x[...] = keep
        assert max(np.abs(t1 - d1).max(), np.abs(t2 - d2).max()) > 1e-12 * scale, name
