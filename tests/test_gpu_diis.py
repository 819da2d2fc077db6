"""The device DIIS (aq_diis_*) against the oracle's restatement of
diis.hpp:36-227, whose solve is LAPACK dsysv: extrapolated amplitudes and
residuals and the coefficients after every call, at an order above the
multi-dot and lincomb kernels' per-launch arity (so the chunked paths run);
the eps retry on an exactly singular B (identical residuals) and the
AQ_ERR_NUMERIC failure when the retry cannot help (zero residuals)."""

import numpy as np
import pytest

import aquarius_amd as aq
from oracle import rhfccsd

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("order,start,damping", [(18, 1, 0.0), (5, 3, 0.25)])
def test_device_diis_equals_oracle(order, start, damping):
    rng = np.random.default_rng(order)
    n1, n2 = 700, 1300
    ref = rhfccsd.DIIS({"order": order, "start": start, "damping": damping}, 2, 2)
    with aq.Context(0) as ctx:
        d = aq.DIIS(ctx, order, start, damping, 2, 2)
        for it in range(order + 4):
            x = [rng.standard_normal(n1), rng.standard_normal(n2)]
            # residuals that shrink, as in a converging solve
            dx = [rng.standard_normal(n1) * 0.7 ** it, rng.standard_normal(n2) * 0.7 ** it]
            tx = [ctx.from_numpy(a) for a in x]
            tdx = [ctx.from_numpy(a) for a in dx]
            d.extrapolate(tx, tdx)
            ref.extrapolate(x, dx)
            for t, a in zip(tx + tdx, x + dx):
                got = t.numpy()
                np.testing.assert_allclose(got, a, rtol=1e-10, atol=1e-12 * np.abs(a).max(), err_msg=f"it {it}")
                t.free()
            c, retry = d.coefficients()
            assert not retry
            if it >= 1 and it + 1 >= start:
                m = min(it + 1, order) + 1
                np.testing.assert_allclose(c[:m], ref.c[:m], rtol=1e-9, atol=1e-12)
        d.close()


def test_identical_residuals_take_the_eps_retry():
    """B is exactly singular; both the oracle (dsysv info > 0) and the device
    (exactly zero pivot) retry with eps*B[n-1][n-1] on the diagonal.  The
    retried system is within one ulp of singular, so the split between the two
    coefficients is roundoff; their sum and the extrapolation are exact."""
    r = np.array([1.0, 2.0, 3.0, -0.5])
    ref = rhfccsd.DIIS({"order": 3, "start": 1, "damping": 0.0}, 1, 1)
    with aq.Context(0) as ctx:
        d = aq.DIIS(ctx, 3, 1, 0.0, 1, 1)
        for xv in (1.0, 3.0):
            x, dx = [np.full(4, xv)], [r.copy()]
            tx, tdx = [ctx.from_numpy(x[0])], [ctx.from_numpy(dx[0])]
            d.extrapolate(tx, tdx)
            ref.extrapolate(x, dx)
        c, retry = d.coefficients()
        assert retry and ref.used_eps_retry
        assert abs(c[0] + c[1] - 1.0) < 1e-12
        np.testing.assert_allclose(tx[0].numpy(), c[0] * 3.0 + c[1] * 1.0, rtol=1e-14)
        np.testing.assert_allclose(tdx[0].numpy(), r, rtol=1e-12)
        d.close()


def test_zero_residuals_fail_as_the_oracle_does():
    """All-zero residuals: eps*B[n-1][n-1] = 0, the retried system is still
    singular; the oracle raises (dsysv info > 0 twice), the device returns
    AQ_ERR_NUMERIC."""
    ref = rhfccsd.DIIS({"order": 3, "start": 1, "damping": 0.0}, 1, 1)
    with aq.Context(0) as ctx:
        d = aq.DIIS(ctx, 3, 1, 0.0, 1, 1)
        z = np.zeros(5)
        d.extrapolate([ctx.from_numpy(np.ones(5))], [ctx.from_numpy(z)])
        ref.extrapolate([np.ones(5)], [z.copy()])
        with pytest.raises(aq.AqError) as ei:
            d.extrapolate([ctx.from_numpy(np.full(5, 2.0))], [ctx.from_numpy(z)])
        assert ei.value.code == 5
        with pytest.raises(RuntimeError):
            ref.extrapolate([np.full(5, 2.0)], [z.copy()])
        d.close()
