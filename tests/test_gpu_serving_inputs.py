"""Served jobs with real integrals (aq_job_spec.inputs / .fcidump): H2O in
cc-pVDZ at the bohr geometry of test/input.aq:3-11, its integrals built by
the front end in the caller's context (the reference's aomoints task feeding
ccsd) or written to an FCIDUMP file (the reference's fcidump task), served
next to synthetic jobs: the CCSD energy reaches the reference's known answer
(test/input.aq:23) to 1e-9 Eh, and the mixed-precision serving path reaches
the FP64 one to 1e-6 Eh."""

import numpy as np
import pytest

import aquarius_amd as aq
from aquarius_amd.molecules import H2O_REFERENCE_BOHR
from oracle import integrals as I

pytestmark = pytest.mark.gpu

E_CCSD = -0.180145524753   # test/input.aq:23


def cfg(prec="fp64"):
    return aq.default_config(convergence=1e-11, diis_order=8, max_iterations=100, precision=prec,
                             refine_below=1e-5, min_reduced_flops=0.0)


def test_served_molecular_and_fcidump_jobs(tmp_path):
    print("case: H2O cc-pVDZ bohr served from front-end tensors and from an FCIDUMP file", flush=True)
    b = I.molecule(H2O_REFERENCE_BOHR, aq.basis_text("cc-pVDZ"))
    S, T, V = I.one_electron(b)
    eri = I.two_electron(b)
    E, eps, C, F = I.rhf(S, T + V, eri, 5, b.nuclear_repulsion())
    h = C.T @ (T + V) @ C
    mo = np.einsum("abcd,ap,bq,cr,ds->pqrs", eri, C, C, C, C, optimize=True)
    path = tmp_path / "FCIDUMP"
    aq.fcidump_write(path, h, mo, b.nuclear_repulsion(), nelec=10)
    with aq.Context(0) as ctx:
        mol = aq.Molecule(H2O_REFERENCE_BOHR, "cc-pVDZ", units="bohr")
        scf = aq.SCF(ctx, mol)
        ints = scf.integrals()
        with aq.Server(slots=3) as s:
            ids = {
                "tensors": s.submit(0, 0, 0, config=cfg(), inputs=ints),
                "tensors_mixed": s.submit(0, 0, 0, config=cfg("bf16x3"), inputs=ints),
                "fcidump": s.submit(0, 0, 0, config=cfg(), fcidump=str(path)),
                "synthetic": s.submit(21, 6, 30, config=cfg()),
            }
            res = {k: s.wait(i) for k, i in ids.items()}
        for k, r in res.items():
            assert r["status"] == aq.JOB_DONE and r["converged"], (k, r["error"])
            print(f"{k}: E = {r['energy']:.12f} ({r['iterations']} it, {r['reduced_iterations']} reduced)")
        assert abs(res["tensors"]["energy"] - E_CCSD) < 1e-9
        assert abs(res["fcidump"]["energy"] - E_CCSD) < 1e-9
        assert abs(res["tensors_mixed"]["energy"] - res["tensors"]["energy"]) < 1e-6
        # the borrowed tensors were not counted against the server budget and are untouched
        assert res["tensors"]["bytes_reserved"] < aq.job_bytes_estimate(5, 19)
        ints.free()


def test_bad_inputs_are_refused():
    with aq.Context(0) as ctx:
        ints = aq.Integrals.synthetic(ctx, 3, 4, 10, 20)
        with aq.Server(slots=1) as s:
            with pytest.raises(aq.AqError):
                s.submit(0, 0, 0, config=cfg(), inputs=ints, fcidump="/nonexistent")
            with pytest.raises(aq.AqError):
                s.submit(0, 0, 0, config=cfg(), fcidump="/nonexistent/FCIDUMP")
            swapped = aq.Integrals(ctx, dict(ints.t, VABIJ=ints.t["VAIBJ"], VAIBJ=ints.t["VABIJ"]))
            with pytest.raises(aq.AqError) as e:
                s.submit(0, 0, 0, config=cfg(), inputs=swapped)
            assert "wrong shape" in str(e.value)
        ints.free()
