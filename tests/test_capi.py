"""CPU-side checks of the C ABI: the library loads, exports every function
include/aqamd.h declares, reports errors without a device, and its host-side
synthetic generator follows the stated construction (no GPU compute)."""

import os
import re

import numpy as np
import pytest

import aquarius_amd as aq
from oracle import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "aqamd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(aq_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    L = aq.lib()
    names = declared_functions()
    assert len(names) >= 35
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the Python binding covers the same surface
    assert set(names) <= set(aq._SIGS), set(names) - set(aq._SIGS)


def test_version_and_error_channel():
    L = aq.lib()
    assert L.aq_version() >= 100
    # a null-context call fails with a message and no abort
    with pytest.raises(aq.AqError) as ei:
        aq._ck(L.aq_ctx_sync(None))
    assert "null context" in str(ei.value)


def test_default_config_matches_reference_spec():
    """rhfccsd.cxx:332-353: convergence 1e-9, max_iterations 50, DIIS order 5, start 1, damping 0."""
    cfg = aq.default_config()
    assert cfg.convergence == 1e-9 and cfg.max_iterations == 50
    assert cfg.diis_order == 5 and cfg.diis_start == 1 and cfg.diis_damping == 0.0
    assert cfg.precision == aq.AQ_PREC_FP64


@pytest.mark.parametrize("seed,o,v,nQ", [(1, 2, 3, 4), (2024, 5, 9, 6)])
def test_host_generator_follows_construction(seed, o, v, nQ):
    eps, B = aq.synth_host(seed, o, v, nQ)
    eo, ev = synth.orbital_energies(seed, o, v)
    np.testing.assert_array_equal(eps, np.concatenate([eo, ev]))     # identical uniforms
    np.testing.assert_allclose(B, synth.three_index(seed, o, v, nQ), rtol=0, atol=1e-15)
    assert np.array_equal(B, B.transpose(0, 2, 1))
    assert np.all(np.diff(eps[:o]) >= 0) and np.all(np.diff(eps[o:]) >= 0)
    assert eps[:o].min() >= -1.5 and eps[:o].max() < -0.3
    assert eps[o:].min() >= 0.2 and eps[o:].max() < 5.0


def test_host_generator_statistics():
    n = 24
    iu = np.triu_indices(n)
    for sigma in (0.05, 0.0):
        eps, B = aq.synth_host(99, 4, 20, 50, sigma)
        s = sigma if sigma > 0 else 0.05 * np.sqrt(150.0 / (20 * np.sqrt(50)))
        assert s == pytest.approx(sigma if sigma > 0 else aq.synth_default_sigma(20, 50), rel=1e-15)
        z = B[:, iu[0], iu[1]].ravel() / s
        assert abs(z.mean()) < 0.02 and abs(z.std() - 1.0) < 0.02


def test_job_bytes_estimate():
    e = aq.job_bytes_estimate
    # monotone in o and v, reduced precision needs its extra operand copies
    assert e(4, 16) < e(5, 16) < e(5, 20)
    assert e(20, 140, 0, 8, "bf16x3") > e(20, 140, 0, 8, "fp64")
    assert e(20, 140, 0, 8) > e(20, 140, 0, 5)
    # covers at least the resident FP64 integrals + DIIS history
    o, v = 32, 256
    floor = 8 * (v**4 + v**3 * o + 2 * 8 * o * o * v * v)
    assert e(o, v, 0, 8) > floor
    assert e(0, 10) == 0


def test_no_device_is_a_clean_error_here():
    if os.path.exists("/dev/kfd") and os.environ.get("AQ_EXPECT_GPU"):
        pytest.skip("device present")
    try:
        ctx = aq.Context()
    except aq.AqError as e:
        assert e.code == 2
    else:
        ctx.close()
