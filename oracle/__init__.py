"""CPU oracle for the closed-shell CCSD amplitude path.

TEST INFRASTRUCTURE ONLY.  This package is a plain numpy restatement of the
reference's algorithm (Aquarius ``RHFCCSD::iterate`` and the CTF operator
semantics under it).  It exists to check the HIP engine; it is never the thing
measured and never a fallback for it.  Only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` may import it.

Modules
-------
ctf_ops
    Operator semantics of ``CTFTensor`` (contract / sum / scale / dot / weight /
    norm) on numpy arrays whose axis order equals the label order.
rhfccsd
    ``RHFCCSD`` setup + iterate, ``DIIS``, ``Denominator``, the ``Iterative``
    fixed-point driver, the Hbar products and the EOM-EE singlet sigma vector
    that consumes them.
spinorbital_ccsd
    An independent textbook spin-orbital CCSD (Stanton-Gauss form) used only to
    cross-check the spin-adapted restatement on tiny closed-shell systems.
synth
    Seeded synthetic closed-shell integrals (counter-based RNG, Cholesky-like
    three-index factors) and the Dirac-block layout the engine consumes.
integrals
    The integral front end: GENBAS basis sets, the Boys function,
    McMurchie-Davidson one- and two-electron integrals, RHF and UHF SCF, and
    the AO->MO Dirac blocks.
fcidump
    FCIDUMP reading (the reference's semantics) and the closed-shell Fock
    matrix and blocks from MO-basis integrals.
socc
    The spin-orbital CCSD / CCD task on a UHF reference.

Pinning status (what checks each function; see tests/test_oracle_*.py)
---------------------------------------------------------------------
* ctf_ops.*: brute-force loops on tiny shapes, closed forms for weight/norm,
  aliasing snapshot test.
* rhfccsd.RHFCCSD.setup (MP2 guess): closed-form textbook RHF-MP2 and the
  reference's MP2 known answer (test/input.aq:21).
* rhfccsd.RHFCCSD.iterate: converged energy and amplitudes equal the
  independent spin-orbital CCSD; Z2 exact (ai)<->(bj) symmetry; size
  consistency on two non-interacting copies; the reference's CCSD and CCD
  known answers for H2O/cc-pVDZ at the bohr geometry (test/input.aq:22-23),
  through the integral front end below.
* rhfccsd.DIIS: coefficients against numpy.linalg.solve of the bordered
  Pulay system; singular-B retry path; start/rotation bookkeeping against a
  brute-force re-derivation from the stored history.
* rhfccsd.RHFCCSD.hbar: the T = 0 limit, FME = dE/dT1 / 2, and the EOM
  Jacobian identity (the sigma vector from the products equals the exact
  derivative of the residual at converged T).
* integrals.*: Boys function against quadrature; closed forms for s-type
  integrals; 8-fold ERI symmetry; rotational invariance; the reference's SCF
  and MP2 known answers for H2O/cc-pVDZ and H2O/DZ (test/input.aq:20-21,
  42-43) and the CH2 triplet UHF known answer (test/input.aq:67).
* fcidump.*: a file written from converged H2O/DZ orbitals reproduces the
  reference's SCF and MP2 known answers; a hand-written file with every line
  kind.
* socc.*: the CH2 triplet MP2, CCD and CCSD known answers
  (test/input.aq:68-70); on a closed shell equal to the spin-adapted
  restatement.
"""
