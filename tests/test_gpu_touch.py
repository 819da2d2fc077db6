"""Writes through aq_tensor_data followed by aq_tensor_touch: cached operand
copies (BF16 hi/lo planes, permuted FP64 copies) and a solver's packed ladder
operators are rebuilt from the new data, so results equal those computed
from scratch on the new values; the solver's next iteration after a
mid-solve change of VABCD equals the oracle's with the same change."""

import ctypes as C

import numpy as np
import pytest

import aquarius_amd as aq
from oracle import rhfccsd, synth

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipMemcpy.restype = C.c_int
    return _hip


def raw_write(ctx, t, arr):
    """Overwrite t's device data with arr (column-major) without telling the library."""
    ctx.sync()
    a = np.asfortranarray(np.asarray(arr, dtype=np.float64))
    rc = hip().hipMemcpy(C.c_void_p(t.data_ptr()), a.ctypes.data_as(C.c_void_p), a.nbytes, 1)  # host to device
    assert rc == 0


@pytest.mark.parametrize("prec,idx", [("bf16x3", ("km", "kn", "mn")), ("fp64", ("mk", "nk", "nm"))])
def test_touch_refreshes_cached_operand_copies(prec, idx):
    rng = np.random.default_rng(5)
    K, M, N = 96, 72, 40
    ia, ib, ic = idx
    dims = {"k": K, "m": M, "n": N}
    a0 = rng.standard_normal([dims[c] for c in ia])
    a1 = rng.standard_normal([dims[c] for c in ia])
    b = rng.standard_normal([dims[c] for c in ib])
    with aq.Context(0) as ctx:
        ctx.set_precision(prec, 0.0)
        A, B = ctx.from_numpy(a0), ctx.from_numpy(b)
        C1 = ctx.tensor([dims[c] for c in ic])
        ctx.contract(1.0, A, ia, B, ib, 0.0, C1, ic)      # builds the cached copy of A
        first = C1.numpy()
        raw_write(ctx, A, a1)
        A.touch()
        ctx.contract(1.0, A, ia, B, ib, 0.0, C1, ic)
        A2 = ctx.from_numpy(a1)                           # a fresh tensor: no cached copy
        C2 = ctx.tensor([dims[c] for c in ic])
        ctx.contract(1.0, A2, ia, B, ib, 0.0, C2, ic)
        assert np.array_equal(C1.numpy(), C2.numpy())
        assert not np.array_equal(C1.numpy(), first)
        for t in (A, B, C1, A2, C2):
            t.free()


def test_solver_follows_a_touched_vabcd():
    """Two iterations, then VABCD scaled by 1.01 through the device pointer
    (touch), then two more: per-iteration energies equal the oracle's with the
    same change (the packed V+- are rebuilt), and equal a second solver whose
    VABCD was changed through aq_write_all."""
    o, v, nQ, seed = 4, 12, 22, 6
    f, V, B, _ = synth.synthetic_system(seed, o, v, nQ)
    cfg = rhfccsd.default_config()
    cfg["diis"]["order"] = 5
    ref = rhfccsd.RHFCCSD(f, V, cfg)
    with aq.Context(0) as ctx:
        eps, Bh = aq.synth_host(seed, o, v, nQ)
        runs = []
        for how in ("touch", "write_all"):
            ints = aq.Integrals.synthetic(ctx, seed, o, v, nQ, host=(eps, Bh))
            cc = aq.RHFCCSD(ctx, ints, diis_order=5)
            assert cc.packed_ladder
            es = [cc.iterate()[0] for _ in range(2)]
            vabcd = ints.t["VABCD"].numpy() * 1.01
            if how == "touch":
                raw_write(ctx, ints.t["VABCD"], vabcd)
                ints.t["VABCD"].touch()
            else:
                ints.t["VABCD"].write_all(vabcd)
            es += [cc.iterate()[0] for _ in range(2)]
            assert cc.packed_ladder
            runs.append(es)
            cc.close()
            ints.free()
        assert runs[0] == runs[1]
        eref = []
        for k in range(4):
            if k == 2:
                ref.V["VABCD"] *= 1.01
            ref.iterate()
            eref.append(ref.energy)
        assert max(abs(a - b) for a, b in zip(runs[0], eref)) < 1e-10, (runs[0], eref)
        # the change matters: an untouched solve differs after the change
        ints = aq.Integrals.synthetic(ctx, seed, o, v, nQ, host=(eps, Bh))
        cc = aq.RHFCCSD(ctx, ints, diis_order=5)
        plain = [cc.iterate()[0] for _ in range(4)]
        cc.close()
        ints.free()
        assert abs(plain[3] - runs[0][3]) > 1e-8
