"""GPU parity of the 3D r2c / c2r transforms (fft / ifft of
src/toolbox.c:806-827) against numpy's pocketfft and the oracle FFT."""
import numpy as np
import pytest

from tests.states import rel_l2

pytestmark = pytest.mark.gpu

SHAPES = [(4, 4, 4), (8, 8, 8), (16, 16, 16), (32, 32, 32), (64, 64, 64),
          (8, 16, 32), (32, 16, 8), (4, 64, 8), (128, 32, 16), (16, 8, 512),
          (512, 8, 4), (256, 16, 8), (128, 128, 128)]

# FFTW-class accuracy: relative L2 error of an FP64 FFT grows like eps*log(N)
TOL = 1e-14 * 8


@pytest.fixture(scope="module")
def ih():
    import instaheat_amd
    instaheat_amd.lib()
    return instaheat_amd


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_r2c_matches_numpy(ih, shape):
    rng = np.random.default_rng(sum(shape))
    x = rng.standard_normal(shape)
    g = ih.Gpu(shape)
    X = g.rfftn(x)
    assert X.shape == (shape[0], shape[1], shape[2] // 2 + 1)
    assert rel_l2(X, np.fft.rfftn(x)) < TOL


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_c2r_matches_numpy(ih, shape):
    rng = np.random.default_rng(1 + sum(shape))
    cs = (shape[0], shape[1], shape[2] // 2 + 1)
    # deliberately non-Hermitian input: c2r must take the real part of the
    # inverse of the Hermitian extension (FFTW / pocketfft semantics)
    X = rng.standard_normal(cs) + 1j * rng.standard_normal(cs)
    g = ih.Gpu(shape)
    y = g.irfftn(X)
    ref = np.fft.irfftn(X, s=shape, axes=(0, 1, 2)) * np.prod(shape)
    assert rel_l2(y, ref) < TOL


@pytest.mark.parametrize("shape", [(16, 16, 16), (32, 16, 64)],
                         ids=lambda s: "x".join(map(str, s)))
def test_fft_matches_oracle(ih, oracle_mod, shape):
    rng = np.random.default_rng(7)
    x = rng.standard_normal(shape)
    g = ih.Gpu(shape)
    o = oracle_mod.Oracle(*shape)
    assert rel_l2(g.rfftn(x), o.rfftn(x)) < TOL
    X = o.rfftn(x) * (1 + 0.5j)
    assert rel_l2(g.irfftn(X), o.irfftn(X)) < TOL


def test_round_trip_and_plane_wave(ih):
    shape = (32, 16, 8)
    g = ih.Gpu(shape)
    n = np.indices(shape)
    kvec = (3, 5, 2)
    x = np.cos(2 * np.pi * sum(kvec[i] * n[i] / shape[i] for i in range(3)))
    X = g.rfftn(x)
    # a plane wave is a single spike of height N/2 at k and its mirror
    N = np.prod(shape)
    assert abs(X[3, 5, 2] - N / 2) < 1e-10
    assert abs(X[-3, -5, 2]) < 1e-10
    X2 = X.copy()
    X2[3, 5, 2] = 0
    assert np.abs(X2).max() < 1e-10
    # c2r(r2c(x)) = N x
    assert rel_l2(g.irfftn(X) / N, x) < TOL


def test_rejects_unsupported_sizes(ih):
    with pytest.raises(ih.IhgError):
        ih.Gpu((12, 8, 8))
    with pytest.raises(ih.IhgError):
        ih.Gpu((1024, 8, 8))
