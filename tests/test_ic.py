"""CPU checks of the oracle's seeded Gaussian initial data (oracle/ic.py):
splitmix64 against its published outputs, and the spectrum against the
distribution the recipe must produce.  No GPU."""
import numpy as np
import pytest
from scipy import stats

from oracle import ic

BOX = ((0.0, 2 * np.pi / 0.1),) * 3


def test_splitmix64_published_outputs():
    # the first outputs of splitmix64 seeded with 1234567 (Vigna's
    # splitmix64.c; the state advances by the golden gamma before mixing)
    want = [6457827717110365317, 3203168211198807973, 9817491932198370423,
            4593380528125082431, 16408922859458223821]
    states = np.uint64(1234567) + np.arange(5, dtype=np.uint64) * ic.GAMMA
    assert [int(v) for v in ic.splitmix64(states)] == want


def test_unit_interval_ends():
    assert ic.unit_open0(np.uint64(0)) == 2.0 ** -53
    assert ic.unit_open0(np.uint64(2 ** 64 - 1)) == 1.0


@pytest.mark.parametrize("seed", [1, 7])
def test_spectrum_distribution(seed):
    """|phi_k|^2 / sigma_k^2 is Exp(1) (Box-Muller radius), the phase is
    uniform, modes outside 0 < |k| <= k_cut are zero."""
    shape = (48, 32, 64)
    spec, live, sigma = ic.gaussian_spectrum(shape, BOX, 1.0, seed, 1e-4, 1.3,
                                             k_cut=2.0)
    assert not live[0, 0, 0] and spec[0, 0, 0] == 0
    assert np.all(spec[~live] == 0)
    kx, ky, kz = ic.wavenumbers(shape, BOX)
    k2 = kx[:, None, None] ** 2 + ky[None, :, None] ** 2 + kz[None, None, :] ** 2
    assert np.all(k2[live] <= 4.0 + 1e-12) and np.all((k2[~live] > 4.0) | (k2[~live] == 0))
    z = (np.abs(spec[live]) / sigma[live]) ** 2
    n = z.size
    assert n > 3000
    assert abs(z.mean() - 1.0) < 4.0 / np.sqrt(n)
    assert stats.kstest(z, "expon").pvalue > 1e-3
    ph = np.angle(spec[live])
    assert stats.kstest((ph + np.pi) / (2 * np.pi), "uniform").pvalue > 1e-3
    # sigma_k^2 = A^2 / (2 a0^3 omega_k), omega_k = sqrt(k^2 + m^2)
    k = np.sqrt(k2[live])
    assert np.allclose(sigma[live] ** 2, 1e-8 / (2 * 1.3 ** 3 * np.sqrt(k ** 2 + 1.0)),
                       rtol=1e-14)


def test_seeds_and_modes_independent():
    shape = (16, 16, 16)
    a, live, _ = ic.gaussian_spectrum(shape, BOX, 1.0, 3, 1e-5, 1.0, k_cut=0.8)
    b, _, _ = ic.gaussian_spectrum(shape, BOX, 1.0, 4, 1e-5, 1.0, k_cut=0.8)
    c, _, _ = ic.gaussian_spectrum(shape, BOX, 1.0, 3, 1e-5, 1.0, k_cut=0.8)
    assert np.array_equal(a, c)
    assert np.abs(np.corrcoef(np.abs(a[live]), np.abs(b[live]))[0, 1]) < 0.2
    # a mode's value depends on (seed, index) only: a larger cut keeps it
    d, live2, _ = ic.gaussian_spectrum(shape, BOX, 1.0, 3, 1e-5, 1.0, k_cut=1.2)
    assert live2.sum() > live.sum()
    assert np.array_equal(a[live], d[live])


def test_state_layout():
    shape = (8, 16, 32)
    N = 8 * 16 * 32
    ntot = 4 * N + 4 * 8 * 16 * 17 * 2 + 1
    y = ic.gaussian_state(shape, BOX, 1.0, 5, 1e-3, 1.0093, -0.7137, 1.2, 1.0, ntot)
    assert y[-1] == 1.2
    assert np.all(y[N:2 * N] == -0.7137)
    assert np.all(y[2 * N:-1] == 0)
    d = y[:N] - 1.0093
    assert abs(d.mean()) < 1e-15  # no k = 0 mode
    assert d.std() > 0
