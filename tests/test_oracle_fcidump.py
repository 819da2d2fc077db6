"""FCIDUMP interchange: the oracle's restatement of the reference's reader
(src/operator/fcidump.cxx:43-517) and the library's host reader and writer
(no GPU).  Pins: a file written from converged H2O/DZ SCF orbitals (bohr
geometry of test/input.aq) reproduces the reference's SCF energy and MP2
(test/input.aq:42-43); a hand-written file exercises every line kind."""

import numpy as np
import pytest

import aquarius_amd as aq
from aquarius_amd.molecules import H2O_REFERENCE_BOHR
from oracle import fcidump as FD
from oracle import integrals as I
from oracle import rhfccsd as R


@pytest.fixture(scope="module")
def h2o_dz_mo():
    b = I.molecule(H2O_REFERENCE_BOHR, aq.basis_text("DZ"))
    S, T, V = I.one_electron(b)
    eri = I.two_electron(b)
    E, eps, C, F = I.rhf(S, T + V, eri, 5, b.nuclear_repulsion())
    h = C.T @ (T + V) @ C
    mo = np.einsum("abcd,ap,bq,cr,ds->pqrs", eri, C, C, C, C, optimize=True)
    return h, mo, b.nuclear_repulsion(), E


def test_write_read_round_trip_and_oracle_reader_agree(tmp_path, h2o_dz_mo):
    h, mo, ecore, _ = h2o_dz_mo
    path = tmp_path / "FCIDUMP"
    aq.fcidump_write(path, h, mo, ecore, nelec=10, threshold=0.0)
    h2, mo2, ec2, n, ne, ms2 = aq.fcidump_read(path)
    assert (n, ne, ms2) == (14, 10, 0)
    # %.16E carries 17 significant digits: doubles survive the round trip exactly
    np.testing.assert_array_equal(h2, np.tril(h) + np.tril(h, -1).T)
    # the file holds one representative per eight-fold class; the MO tensor is
    # symmetric only to rounding
    np.testing.assert_allclose(mo2, mo, rtol=0, atol=5e-13)
    assert ec2 == ecore
    # exactness: writing the read-back integrals again returns them bit for bit
    path2 = tmp_path / "FCIDUMP2"
    aq.fcidump_write(path2, h2, mo2, ec2, nelec=10, threshold=0.0)
    h3, mo3, ec3, *_ = aq.fcidump_read(path2)
    np.testing.assert_array_equal(mo3, mo2)
    np.testing.assert_array_equal(h3, h2)
    ho, moo, eco, no, neo = FD.read(path)
    np.testing.assert_array_equal(ho, h2)
    np.testing.assert_array_equal(moo, mo2)
    assert eco == ec2


def test_fcidump_reproduces_reference_known_answers(tmp_path, h2o_dz_mo):
    h, mo, ecore, E = h2o_dz_mo
    path = tmp_path / "FCIDUMP"
    aq.fcidump_write(path, h, mo, ecore, nelec=10)
    hh, ee, ec, n, ne = FD.read(path)
    fock, V, e = FD.blocks(hh, ee, ne // 2)
    assert abs(ec + e - (-74.491837303542)) < 1e-9                      # test/input.aq:42
    assert np.abs(fock["fAI"]).max() < 1e-8                               # canonical orbitals
    cc = R.RHFCCSD(fock, V)
    assert abs(cc.mp2 - (-0.089766142431)) < 1e-9                         # test/input.aq:43


def test_hand_written_file(tmp_path):
    text = """ &FCI NORB=3,NELEC=2,MS2=0,
  ORBSYM=1,1,1,
  ISYM=1,
 &END
  0.5D+00   1   1   1   1
  0.25      2   1   1   1
  0.125     1   1   2   1
  0.1       3   2   2   1
 -1.5       1   1   0   0
  0.2       2   1   0   0
  3.0       1   0   0   0
  7.0       0   0   0   0
"""
    path = tmp_path / "FCIDUMP"
    path.write_text(text)
    h, eri, ec, n, ne, ms2 = aq.fcidump_read(path)
    assert (n, ne, ec) == (3, 2, 7.0)
    assert h[0, 0] == -1.5 and h[1, 0] == 0.2 and h[0, 1] == 0.2 and h[2, 2] == 0.0
    assert eri[0, 0, 0, 0] == 0.5
    # (21|11) is canonical; (11|21) precedes it and is skipped
    for x in [(1, 0, 0, 0), (0, 1, 0, 0), (0, 0, 1, 0), (0, 0, 0, 1)]:
        assert eri[x] == 0.25
    for x in [(2, 1, 1, 0), (1, 2, 1, 0), (2, 1, 0, 1), (1, 2, 0, 1), (1, 0, 2, 1), (0, 1, 2, 1), (1, 0, 1, 2),
              (0, 1, 1, 2)]:
        assert eri[x] == 0.1
    assert np.count_nonzero(eri) == 1 + 4 + 8
    ho, eo, eco, no, neo = FD.read(path)
    np.testing.assert_array_equal(ho, h)
    np.testing.assert_array_equal(eo, eri)
    # full_1e: no mirroring
    h2 = aq.fcidump_read(path, full_1e=True)[0]
    assert h2[1, 0] == 0.2 and h2[0, 1] == 0.0


def test_fcidump_errors(tmp_path):
    p = tmp_path / "bad"
    p.write_text(" &FCI NORB=2,NELEC=2,\n")
    with pytest.raises(aq.AqError) as e:
        aq.fcidump_read(p)
    assert "&END" in str(e.value)
    p.write_text(" &FCI NELEC=2, &END\n")
    with pytest.raises(aq.AqError):
        aq.fcidump_read(p)
    p.write_text(" &FCI NORB=2,NELEC=2, &END\n 1.0 3 1 1 1\n")
    with pytest.raises(aq.AqError) as e:
        aq.fcidump_read(p)
    assert "out of range" in str(e.value)
