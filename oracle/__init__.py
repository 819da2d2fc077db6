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
    ``RHFCCSD`` setup + iterate, ``DIIS``, ``Denominator`` and the
    ``Iterative`` fixed-point driver.
spinorbital_ccsd
    An independent textbook spin-orbital CCSD (Stanton-Gauss form) used only to
    cross-check the spin-adapted restatement on tiny closed-shell systems.
synth
    Seeded synthetic closed-shell integrals (counter-based RNG, Cholesky-like
    three-index factors) and the Dirac-block layout the engine consumes.

Pinning status (what checks each function; see tests/test_oracle_*.py)
---------------------------------------------------------------------
* ctf_ops.*: brute-force loops on tiny shapes, closed forms for weight/norm,
  aliasing snapshot test.
* rhfccsd.RHFCCSD.setup (MP2 guess): closed-form textbook RHF-MP2.
* rhfccsd.RHFCCSD.iterate: converged energy and amplitudes equal the
  independent spin-orbital CCSD; Z2 exact (ai)<->(bj) symmetry; size
  consistency on two non-interacting copies.
* rhfccsd.DIIS: coefficients against numpy.linalg.solve of the bordered
  Pulay system; singular-B retry path; start/rotation bookkeeping against a
  brute-force re-derivation from the stored history.
* The whole solve on the reference's known-answer molecule (H2O/cc-pVDZ,
  test/input.aq:20-23) needs an AO integral front end that this oracle does
  not have: that anchor is *not* reproduced here (parity of the whole solve on
  real molecules is pinned only through the checks above).
"""
