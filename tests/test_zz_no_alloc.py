"""ihg_dp853_finish_noalloc (no context, no allocation, the partials in the
library's fixed device array) against ihg_dp853_finish on the same inputs:
y_new bit for bit and the two error sums bit for bit (same kernels, same
work-group count, same reduction order).  Two calls in a row, the second
with a smaller state, so the second reads what the first left in the shared
array only where it writes first.  The first (64^3, 4160 work-groups) sums
its block partials in two levels, the second (8 x 16 x 8) in one.  The sums
are also checked against numpy's restatement of error()
(src/dopri853.c:469-486) to rounding."""
import ctypes

import numpy as np
import pytest

from tests.states import BOX_CONFIG, make_state

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ih():
    import instaheat_amd
    instaheat_amd.lib()
    return instaheat_amd


def _numpy_sums(y, ks, dt, rtol, atol, nlive, tab):
    b, bhh, er = tab["b"], tab["bhh"], tab["er"]
    idx = np.r_[0:nlive, y.size - 1]
    K = [k[idx] for k in ks]
    yv = y[idx]
    bs = sum(b[j] * K[j] for j in range(12) if b[j] != 0.0)
    yn = yv + dt * bs
    yerr = bs - bhh[0] * K[0] - bhh[1] * K[8] - bhh[2] * K[11]
    yerr2 = sum(er[j] * K[j] for j in range(12) if er[j] != 0.0)
    sk = atol + rtol * np.maximum(np.abs(yv), np.abs(yn))
    return np.array([np.sum((yerr2 / sk) ** 2), np.sum((yerr / sk) ** 2)]), yn


def _case(ih, shape, gw, seed, dt):
    rng = np.random.default_rng(seed)
    g = ih.Gpu(shape, bounds=BOX_CONFIG, gw=gw, filt=True, bins=8)
    y = make_state(shape, seed=seed, gw=True)  # h blocks filled in both cases
    ks = [y * 1e-2 + 1e-3 * rng.standard_normal(y.size) for _ in range(12)]
    d_y = g.to_device(y)
    d_k = [g.to_device(k) for k in ks]
    garbage = rng.standard_normal(y.size)
    d_yn_ctx = g.to_device(garbage)
    d_yn_free = g.to_device(garbage)
    rtol, atol = 1e-10, 1e-14
    print(f"no_alloc case {shape} gw={gw} seed={seed}: ihg_dp853_finish",
          flush=True)
    s_ctx = g.dp853_finish(dt, d_y, d_k, d_yn_ctx, rtol, atol)
    g.sync()  # the context's stream is done before the null stream starts
    nlive = g.Ntot - 1 if gw else 4 * g.N
    arr = (ctypes.c_void_p * 12)(*[k.ptr for k in d_k])
    s_free = np.zeros(2)
    print(f"no_alloc case {shape} gw={gw}: ihg_dp853_finish_noalloc "
          f"nlive={nlive} ntot={g.Ntot}", flush=True)
    rc = ih.lib().ihg_dp853_finish_noalloc(
        None, nlive, g.Ntot, dt, d_y.ptr, arr, d_yn_free.ptr, rtol, atol,
        s_free.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert rc == 0, ih.lib().ihg_last_error()
    yn_ctx = d_yn_ctx.download()
    yn_free = d_yn_free.download()
    np.testing.assert_array_equal(yn_free, yn_ctx)
    np.testing.assert_array_equal(s_free, s_ctx)
    ref, yn_ref = _numpy_sums(y, ks, dt, rtol, atol, nlive, ih.tableau())
    np.testing.assert_allclose(s_free, ref, rtol=1e-12)
    idx = np.r_[0:nlive, y.size - 1]
    np.testing.assert_allclose(yn_free[idx], yn_ref, rtol=1e-14, atol=1e-15)
    # outside the live part y_new is not written
    if not gw:
        np.testing.assert_array_equal(yn_free[nlive:-1], garbage[nlive:-1])
    g.close()
    return s_free


def test_noalloc_finish_equals_context_finish_twice(ih):
    big = _case(ih, (64, 64, 64), True, 3, 1e-2)
    small = _case(ih, (8, 16, 8), False, 4, 3e-3)
    assert np.all(big > 0) and np.all(small > 0)


def test_noalloc_finish_rejects_bad_sizes(ih):
    s = np.zeros(2)
    sp = s.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    arr = (ctypes.c_void_p * 12)()
    L = ih.lib()
    for nlive, ntot in ((3, 10), (0, 10), (10, 10),
                        (2 ** 31, 2 ** 31 + 1)):
        rc = L.ihg_dp853_finish_noalloc(None, nlive, ntot, 1e-2, None, arr,
                                        None, 1e-10, 1e-14, sp)
        assert rc == -22, (nlive, ntot)  # IHG_EINVAL
        assert b"nlive" in L.ihg_last_error()
