"""CPU checks of the log-spaced spectrum bins of the oracle (log_bins = 1,
no counterpart in the reference, which bins linearly): a single Fourier
mode lands in the bin whose edges k_lo (k_max / k_lo)^(b / bins) enclose
its |k|, and binning moves power between bins but never changes the total.
No GPU."""
import numpy as np
import pytest

from oracle import oracle as orc

BOX = ((0.0, 2 * np.pi / 0.1), (0.0, 2 * np.pi / 0.2), (0.0, 2 * np.pi / 0.15))


def _phi_ps(shape, bins, log_bins, phi):
    o = orc.Oracle(*shape, bounds=BOX, mass=1.0, gw=False, filt=False,
                   bins=bins, log_bins=log_bins, threads=1)
    N = int(np.prod(shape))
    y = np.zeros(o.Ntot)
    y[:N] = phi.ravel()
    y[N:2 * N] = -0.7
    y[-1] = 1.0
    o.mk_rhs(y, output=1, field=y)
    return orc.unpack_obs(o.obs(), bins)["phi_ps"]


@pytest.mark.parametrize("mode", [(1, 0, 0), (0, 3, 2), (5, 7, 1), (8, 8, 8),
                                  (2, 15, 0)])
def test_single_mode_lands_between_its_log_edges(mode):
    shape, bins = (16, 32, 16), 24
    ix, iy, iz = mode
    k1 = [2 * np.pi / (b - a) for a, b in BOX]
    x = [np.arange(n) * (b - a) / n for n, (a, b) in zip(shape, BOX)]
    X, Y, Z = np.meshgrid(*x, indexing="ij")
    phi = 1.0 + 1e-3 * np.cos(ix * k1[0] * X + iy * k1[1] * Y + iz * k1[2] * Z)
    ps = _phi_ps(shape, bins, True, phi)
    b = int(np.argmax(ps))
    assert np.count_nonzero(ps > 1e-12 * ps.max()) == 1
    k = np.sqrt((ix * k1[0]) ** 2 + (iy * k1[1]) ** 2 + (iz * k1[2]) ** 2)
    k_lo = min(k1)
    k_max = np.sqrt(sum((n // 2 * kk) ** 2 for n, kk in zip(shape, k1)))
    edges = k_lo * (k_max / k_lo) ** (np.arange(bins + 1) / bins)
    assert edges[b] * (1 - 1e-12) <= k <= edges[b + 1] * (1 + 1e-12)


def test_log_and_linear_bins_hold_the_same_power():
    shape = (16, 16, 32)
    rng = np.random.default_rng(4)
    phi = 1.0 + 1e-3 * rng.standard_normal(shape)
    lin = _phi_ps(shape, 20, False, phi)
    log = _phi_ps(shape, 20, True, phi)
    assert lin.sum() == pytest.approx(log.sum(), rel=1e-13)
    assert not np.allclose(lin, log)
