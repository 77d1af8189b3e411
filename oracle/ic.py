"""CPU restatement of the seeded Gaussian initial data -- TEST
INFRASTRUCTURE ONLY (tests/ and nothing else use it).

The recipe of instaheat_amd/csrc/ihg_ic.hip in numpy: splitmix64 counters
per mode of the r2c half spectrum, Box-Muller with a uniform phase, variance
amplitude^2 / (2 a0^3 omega_k) for 0 < |k| <= k_cut, the unnormalised c2r
(FFTW semantics), phi0 added, dphi0 set, a0 last.  The reference generates
its initial data on the host with GSL (src/setup.c:771-790, absent here);
this recipe is the build's own, pinned in tests/test_ic.py by splitmix64's
published outputs and by the distribution the spectrum must have.
"""
import numpy as np

GAMMA = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)
TWOPI = 6.283185307179586476925286766559


def splitmix64(x):
    """splitmix64 output for state x (x advanced by the golden gamma first),
    element-wise on uint64 arrays, wrapping mod 2^64."""
    with np.errstate(over="ignore"):
        x = np.asarray(x, dtype=np.uint64) + GAMMA
        x = (x ^ (x >> np.uint64(30))) * M1
        x = (x ^ (x >> np.uint64(27))) * M2
        return x ^ (x >> np.uint64(31))


def unit_open0(r):
    return ((r >> np.uint64(11)).astype(np.float64) + 1.0) * (1.0 / 9007199254740992.0)


def wavenumbers(shape, bounds):
    Nx, Ny, Nz = shape
    Mz = Nz // 2 + 1
    k1 = [TWOPI / (b - a) for a, b in bounds]
    ix = np.arange(Nx)
    iy = np.arange(Ny)
    kx = k1[0] * np.where(ix > Nx // 2, ix - Nx, ix).astype(np.float64)
    ky = k1[1] * np.where(iy > Ny // 2, iy - Ny, iy).astype(np.float64)
    kz = k1[2] * np.arange(Mz, dtype=np.float64)
    return kx, ky, kz


def gaussian_spectrum(shape, bounds, mass, seed, amplitude, a0, k_cut):
    """the half spectrum (Nx, Ny, Nz/2 + 1) of delta phi"""
    Nx, Ny, Nz = shape
    Mz = Nz // 2 + 1
    kx, ky, kz = wavenumbers(shape, bounds)
    k2 = (kx[:, None, None] * kx[:, None, None] + ky[None, :, None] * ky[None, :, None]) \
        + kz[None, None, :] * kz[None, None, :]
    live = (k2 != 0.0) & ~(k2 > k_cut * k_cut)
    idx = np.arange(Nx * Ny * Mz, dtype=np.uint64).reshape(Nx, Ny, Mz)
    key = splitmix64(np.uint64(seed) ^ np.uint64(0x5DEECE66D))
    with np.errstate(over="ignore"):
        base = key + np.uint64(2) * idx
        u1 = unit_open0(splitmix64(base))
        u2 = unit_open0(splitmix64(base + np.uint64(1))) - 2.0 ** -53
    omega = np.sqrt(k2 + mass * mass)
    sigma = amplitude * np.sqrt(1.0 / (2.0 * a0 * a0 * a0 * omega))
    with np.errstate(divide="ignore", invalid="ignore"):
        r = np.sqrt(-2.0 * np.log(u1)) * sigma / np.sqrt(2.0)
    ph = TWOPI * u2
    spec = np.where(live, r * np.cos(ph) + 1j * r * np.sin(ph), 0.0)
    return spec, live, sigma


def gaussian_state(shape, bounds, mass, seed, amplitude, phi0, dphi0, a0,
                   k_cut, ntot):
    """the state vector: phi0 + delta phi, dphi0, zeros, a0"""
    spec, _, _ = gaussian_spectrum(shape, bounds, mass, seed, amplitude, a0, k_cut)
    N = int(np.prod(shape))
    dphi = np.fft.irfftn(spec, s=shape, axes=(0, 1, 2)) * N  # unnormalised c2r
    y = np.zeros(ntot)
    y[:N] = dphi.ravel() + phi0
    y[N:2 * N] = dphi0
    y[-1] = a0
    return y
