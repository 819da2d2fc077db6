"""Pins the oracle's integral front end (oracle/integrals.py).

* The reference's known answers for H2O at the bohr geometry of
  test/input.aq:3-11, each to the reference's own tolerance 1e-9 Eh:
  cc-pVDZ SCF -74.550126456692, MP2 -0.171348679568, CCD -0.179753103625,
  CCSD -0.180145524753 (test/input.aq:20-23); DZ SCF -74.491837303542 and
  MP2 -0.089766142431 (test/input.aq:42-43).
* The Boys function against direct quadrature.
* s-type integrals against their closed forms.
* Rotational invariance of the generalised (H, S) spectrum, which fails for a
  wrong pure-harmonic transform of p or d shells.
"""

import os

import numpy as np
import pytest
from scipy import integrate
from scipy.linalg import eigh

from aquarius_amd.molecules import H2O_REFERENCE_BOHR
from oracle import integrals as I
from oracle import rhfccsd as R

BASIS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "aquarius_amd", "basis")


def genbas(name):
    with open(os.path.join(BASIS, name + ".gbs")) as f:
        return f.read()


@pytest.mark.parametrize("T", [0.0, 1e-9, 0.3, 2.0, 9.7, 24.99, 25.01, 60.0, 800.0])
def test_boys_against_quadrature(T):
    F = I.boys(8, np.array([T]))[:, 0]
    for m in range(9):
        q = integrate.quad(lambda t: t ** (2 * m) * np.exp(-T * t * t), 0.0, 1.0, epsabs=0, epsrel=2e-14, limit=400)[0]
        assert F[m] == pytest.approx(q, rel=1e-12, abs=1e-300)


def one_s_basis(a, b, A, B):
    shells = {"H": [(0, np.array([a]), np.array([[1.0]]))], "HE": [(0, np.array([b]), np.array([[1.0]]))]}
    return I.Basis([(1, A), (2, B)], shells)


def test_s_integrals_closed_forms():
    a, b = 0.8, 1.7
    A, B = np.array([0.1, -0.2, 0.3]), np.array([0.9, 0.4, -0.5])
    bas = one_s_basis(a, b, A, B)
    S, T, V = I.one_electron(bas)
    na, nb = (2 * a / np.pi) ** 0.75, (2 * b / np.pi) ** 0.75
    p, mu = a + b, a * b / (a + b)
    R2 = np.sum((A - B) ** 2)
    Sab = na * nb * (np.pi / p) ** 1.5 * np.exp(-mu * R2)
    assert S[0, 0] == pytest.approx(1.0, abs=1e-14) and S[1, 1] == pytest.approx(1.0, abs=1e-14)
    assert S[0, 1] == pytest.approx(Sab, rel=1e-13)
    assert T[0, 1] == pytest.approx(mu * (3 - 2 * mu * R2) * Sab, rel=1e-13)
    assert T[0, 0] == pytest.approx(1.5 * a, rel=1e-13)
    P = (a * A + b * B) / p
    F0 = lambda x: I.boys(0, np.array([x]))[0, 0]
    Vab = -sum(Z * 2 * np.pi / p * na * nb * np.exp(-mu * R2) * F0(p * np.sum((P - C) ** 2))
               for Z, C in ((1, A), (2, B)))
    assert V[0, 1] == pytest.approx(Vab, rel=1e-13)
    eri = I.two_electron(bas)
    # (ab|ab) closed form
    q = p
    g = 2 * np.pi ** 2.5 / (p * q * np.sqrt(p + q)) * (na * nb * np.exp(-mu * R2)) ** 2 * F0(0.0)
    assert eri[0, 1, 0, 1] == pytest.approx(g, rel=1e-13)
    # (aa|bb): two separated charge distributions
    pa, pb = 2 * a, 2 * b
    al = pa * pb / (pa + pb)
    g2 = 2 * np.pi ** 2.5 / (pa * pb * np.sqrt(pa + pb)) * na ** 2 * nb ** 2 * F0(al * R2)
    assert eri[0, 0, 1, 1] == pytest.approx(g2, rel=1e-13)
    for perm in [(1, 0, 2, 3), (0, 1, 3, 2), (2, 3, 0, 1)]:
        np.testing.assert_allclose(eri, eri.transpose(perm), rtol=0, atol=1e-15)


def test_pure_harmonics():
    for L in range(5):
        M = I.solid_harmonics(L)
        S = I.cart_overlap_axial(L)
        np.testing.assert_allclose(M @ S @ M.T, np.eye(2 * L + 1), atol=1e-13)
    # the d set is orthogonal to the contaminant r^2 = x^2 + y^2 + z^2
    P = I.cart_powers(2)
    r2 = np.zeros(6)
    for k, pw in enumerate(P):
        if sorted(pw) == [0, 0, 2]:
            r2[k] = 1.0
    np.testing.assert_allclose(I.solid_harmonics(2) @ I.cart_overlap_axial(2) @ r2, 0.0, atol=1e-14)


def test_rotational_invariance_of_one_electron_spectrum():
    text = genbas("cc-pVDZ")
    rng = np.random.default_rng(5)
    q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    geom = np.array([a[1:] for a in H2O_REFERENCE_BOHR]) + np.array([0.3, -0.1, 0.2])
    rot = geom @ q.T
    spec = []
    for g in (geom, rot):
        b = I.molecule([(a[0], *x) for a, x in zip(H2O_REFERENCE_BOHR, g)], text)
        S, T, V = I.one_electron(b)
        spec.append(eigh(T + V, S, eigvals_only=True))
    np.testing.assert_allclose(spec[0], spec[1], rtol=0, atol=1e-10)


def test_genbas_parser():
    shells = I.parse_genbas(genbas("cc-pVDZ"))
    L, exps, coef = shells["O"][0]
    assert L == 0 and exps.shape == (9,) and coef.shape == (9, 3)
    assert [s[0] for s in shells["O"]] == [0, 1, 2] and [s[0] for s in shells["H"]] == [0, 1]
    with pytest.raises(ValueError):
        I.parse_genbas("no header here\n")


def scf_and_cc(name):
    b = I.molecule(H2O_REFERENCE_BOHR, genbas(name))
    S, T, V = I.one_electron(b)
    eri = I.two_electron(b)
    nocc = b.n_electrons() // 2
    E, eps, C, F = I.rhf(S, T + V, eri, nocc, b.nuclear_repulsion())
    f, Vb = I.mo_blocks(eri, C, F, nocc)
    return E, f, Vb


@pytest.fixture(scope="module")
def h2o_pvdz():
    return scf_and_cc("cc-pVDZ")


def test_reference_known_answers_ccpvdz(h2o_pvdz):
    E, f, Vb = h2o_pvdz
    assert abs(E - (-74.550126456692)) < 1e-9                       # test/input.aq:20
    cfg = R.default_config()
    cfg["convergence"] = 1e-11
    cfg["max_iterations"] = 100
    cc = R.RHFCCSD(f, Vb, cfg)
    assert abs(cc.mp2 - (-0.171348679568)) < 1e-9                   # test/input.aq:21
    assert cc.run()
    assert abs(cc.energy - (-0.180145524753)) < 1e-9                # test/input.aq:23
    cfg["freeze_t1"] = True
    cd = R.RHFCCSD(f, Vb, cfg)
    assert cd.run()
    assert abs(cd.energy - (-0.179753103625)) < 1e-9                # test/input.aq:22


def test_reference_known_answers_dz():
    E, f, Vb = scf_and_cc("DZ")
    assert abs(E - (-74.491837303542)) < 1e-9                       # test/input.aq:42
    cc = R.RHFCCSD(f, Vb)
    assert abs(cc.mp2 - (-0.089766142431)) < 1e-9                   # test/input.aq:43
