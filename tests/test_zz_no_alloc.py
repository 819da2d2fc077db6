"""The allocation-free weight (aq_weight_noalloc) through the C ABI: bitwise
equal to aq_weight on the same inputs, on two calls in a row with the second
smaller, with denominators that hit the |den| < 1e-4 cut; lengths beyond the
fixed device table are refused and leave the tensor unchanged."""

import ctypes as C

import numpy as np
import pytest

import aquarius_amd as aq
from oracle import ctf_ops as ops

pytestmark = pytest.mark.gpu


def call_noalloc(ctx, X, ds, shift=0.0):
    arrs = [np.ascontiguousarray(d, dtype=np.float64) for d in ds]
    ptrs = (aq._DP * len(arrs))(*[aq._dp(a) for a in arrs])
    return aq.lib().aq_weight_noalloc(ctx.h, X.h, len(arrs), ptrs, shift)


def denominators(rng, o, v):
    dI = rng.uniform(-1.5, -0.3, o)
    dA = -rng.uniform(0.2, 5.0, v)
    # planted cuts: dA[0] + dA[1] + dI[0] + dI[1] = 0 and a sum of 5e-5 for [0, 2, 0, 1]
    dA[1] = -(dA[0] + dI[0] + dI[1])
    dA[2] = 5e-5 - (dA[0] + dI[0] + dI[1])
    return dA, dI


@pytest.mark.parametrize("shift", [0.0, 0.125])
def test_noalloc_weight_equals_weight(shift):
    rng = np.random.default_rng(11)
    with aq.Context(0) as ctx:
        for (o, v) in ((12, 40), (3, 7)):          # the second call is the smaller input
            dA, dI = denominators(rng, o, v)
            x = rng.standard_normal((v, v, o, o))
            X1 = ctx.from_numpy(x)
            X2 = ctx.from_numpy(x)
            ctx.weight(X1, [dA, dA, dI, dI], shift)
            rc = call_noalloc(ctx, X2, [dA, dA, dI, dI], shift)
            assert rc == 0, aq.lib().aq_last_error()
            a, b = X1.numpy(), X2.numpy()
            assert np.array_equal(a, b)
            ref = x.copy()
            ops.weight(ref, [dA, dA, dI, dI], shift)
            assert np.array_equal(b, ref)
            if shift == 0.0:
                assert b[0, 1, 0, 1] == 0.0 and b[0, 2, 0, 1] == 0.0
            X1.free()
            X2.free()
        # a two-dimensional tensor after the four-dimensional ones: the table's
        # earlier contents must not leak into the result
        dA, dI = denominators(rng, 5, 9)
        x = rng.standard_normal((9, 5))
        X1, X2 = ctx.from_numpy(x), ctx.from_numpy(x)
        ctx.weight(X1, [dA, dI], shift)
        assert call_noalloc(ctx, X2, [dA, dI], shift) == 0
        assert np.array_equal(X1.numpy(), X2.numpy())
        X1.free()
        X2.free()


def test_noalloc_weight_refuses_lengths_beyond_the_table():
    limit = 65536   # AQ_NOALLOC_WEIGHT_MAX
    with aq.Context(0) as ctx:
        x = np.arange(limit + 1, dtype=np.float64)
        X = ctx.from_numpy(x)
        rc = call_noalloc(ctx, X, [np.full(limit + 1, 2.0)])
        assert rc == 1
        assert "fixed device table" in aq.lib().aq_last_error().decode()
        assert np.array_equal(X.numpy(), x)
        X.free()
        # exactly at the limit is accepted and equals aq_weight
        x = np.linspace(-1.0, 1.0, limit)
        d = np.linspace(0.5, 3.0, limit)
        X1, X2 = ctx.from_numpy(x), ctx.from_numpy(x)
        ctx.weight(X1, [d])
        assert call_noalloc(ctx, X2, [d]) == 0
        assert np.array_equal(X1.numpy(), X2.numpy())
        X1.free()
        X2.free()
