"""Pins the oracle's open-shell path: UHF (oracle.integrals.uhf) and the
spin-orbital CCSD restatement of src/cc/ccsd.cxx:30-193 (oracle.socc).

* The reference's CH2 triplet known answers at the bohr geometry of
  test/input.aq:47-72, to its tolerance 1e-9: UHF SCF -37.090409355231,
  MP2 -0.085465008793, CCD -0.098998464453, CCSD -0.099500248606.
* On a closed shell (H2O/DZ) the spin-orbital path equals the spin-adapted
  restatement of rhfccsd.cxx and the independent Stanton-Gauss code.
* The implicit antisymmetriser leaves the amplitudes antisymmetric."""

import numpy as np
import pytest

import aquarius_amd as aq
from aquarius_amd.molecules import H2O_REFERENCE_BOHR
from oracle import integrals as I
from oracle import rhfccsd as R
from oracle import socc
from oracle import spinorbital_ccsd as SG

CH2_BOHR = [("C", 0.0, 0.0, 0.11726921), ("H", 0.75698224, 0.0, -0.46907685), ("H", -0.75698224, 0.0, -0.46907685)]


def tight():
    cfg = R.default_config()
    cfg["convergence"] = 1e-11
    cfg["max_iterations"] = 200
    cfg["diis"]["order"] = 8
    return cfg


@pytest.fixture(scope="module")
def ch2():
    b = I.molecule(CH2_BOHR, aq.basis_text("cc-pVDZ"))
    S, T, V = I.one_electron(b)
    eri = I.two_electron(b)
    E, eps, Cs, Fs = I.uhf(S, T + V, eri, 5, 3, b.nuclear_repulsion())
    f, g, spin = socc.spin_orbitals(Cs[0], Cs[1], Fs[0], Fs[1], eri, 5, 3)
    return E, f, g


def test_ch2_triplet_known_answers(ch2):
    E, f, g = ch2
    assert abs(E - (-37.090409355231)) < 1e-9                     # test/input.aq:67
    cc = socc.SpinOrbitalCCSD(f, g, 8, tight())
    assert abs(cc.mp2 - (-0.085465008793)) < 1e-9                 # test/input.aq:68
    assert cc.run()
    assert abs(cc.energy - (-0.099500248606)) < 1e-9              # test/input.aq:70
    cd = socc.SpinOrbitalCCSD(f, g, 8, tight(), freeze_t1=True)
    assert cd.run()
    assert abs(cd.energy - (-0.098998464453)) < 1e-9              # test/input.aq:69
    # amplitudes stay antisymmetric under the implicit antisymmetriser
    np.testing.assert_allclose(cc.T2, -cc.T2.transpose(1, 0, 2, 3), rtol=0, atol=1e-12)
    np.testing.assert_allclose(cc.T2, -cc.T2.transpose(0, 1, 3, 2), rtol=0, atol=1e-12)


def test_closed_shell_equals_rhf_and_stanton_gauss():
    b = I.molecule(H2O_REFERENCE_BOHR, aq.basis_text("DZ"))
    S, T, V = I.one_electron(b)
    eri = I.two_electron(b)
    E_r, eps, C, F = I.rhf(S, T + V, eri, 5, b.nuclear_repulsion())
    E_u, eps_u, Cs, Fs = I.uhf(S, T + V, eri, 5, 5, b.nuclear_repulsion())
    assert abs(E_u - E_r) < 1e-10
    f, g, spin = socc.spin_orbitals(Cs[0], Cs[1], Fs[0], Fs[1], eri, 5, 5)
    so = socc.SpinOrbitalCCSD(f, g, 10, tight())
    assert so.run()
    fr, Vr = I.mo_blocks(eri, C, F, 5)
    rc = R.RHFCCSD(fr, Vr, tight())
    assert rc.run()
    assert abs(so.energy - rc.energy) < 1e-10
    assert abs(so.mp2 - rc.mp2) < 1e-10
    # independent Stanton-Gauss code on the same spin-orbital integrals
    o = 10
    sg = SG.SpinOrbitalCCSD(f, g, o)
    e_sg, _ = sg.solve(tol=1e-11)
    assert abs(e_sg - so.energy) < 1e-10
