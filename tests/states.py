"""Seeded synthetic states for the parity tests (layout of include/ihg.h)."""
import numpy as np

BOX_CONFIG = ((0.0, 2 * np.pi / 0.1),) * 3  # L = 2 pi / k_min, k_min = 0.1


def sizes(shape):
    nx, ny, nz = shape
    N = nx * ny * nz
    M = nx * ny * (nz // 2 + 1)
    Next = 2 * M
    return N, M, Next, 4 * N + 4 * Next + 1


def kgrids(shape, bounds):
    """Full wave numbers k_i (FFT frequency order, half axis along z)."""
    nx, ny, nz = shape
    L = [b[1] - b[0] for b in bounds]
    kx = 2 * np.pi / L[0] * np.fft.fftfreq(nx, 1.0 / nx)
    ky = 2 * np.pi / L[1] * np.fft.fftfreq(ny, 1.0 / ny)
    kz = 2 * np.pi / L[2] * np.arange(nz // 2 + 1)
    return np.meshgrid(kx, ky, kz, indexing="ij")


def smooth_field(rng, shape, bounds, rms, kscale):
    """Band-limited Gaussian random field with spectrum exp(-k^2/kscale^2)."""
    KX, KY, KZ = kgrids(shape, bounds)
    k2 = KX ** 2 + KY ** 2 + KZ ** 2
    X = (rng.standard_normal(k2.shape) + 1j * rng.standard_normal(k2.shape))
    X *= np.exp(-k2 / kscale ** 2)
    x = np.fft.irfftn(X, s=shape, axes=(0, 1, 2))
    s = x.std()
    return x * (rms / s) if s > 0 else x


def make_state(shape, bounds=BOX_CONFIG, seed=0, gw=False, a=1.0,
               phi0=1.0093, dphi0=-0.7137, amp=1e-3, psi_amp=1e-5,
               h_amp=1e-7, kscale=None):
    """A state vector y (Ntot doubles) with smooth random fluctuations."""
    rng = np.random.default_rng(seed)
    N, M, Next, Ntot = sizes(shape)
    if kscale is None:
        kscale = 0.35 * min(shape) * 2 * np.pi / max(b[1] - b[0] for b in bounds) / 2
    y = np.zeros(Ntot)
    y[0:N] = phi0 + smooth_field(rng, shape, bounds, amp, kscale).ravel()
    y[N:2 * N] = dphi0 + smooth_field(rng, shape, bounds, amp, kscale).ravel()
    y[2 * N:3 * N] = smooth_field(rng, shape, bounds, psi_amp, kscale).ravel()
    y[3 * N:4 * N] = smooth_field(rng, shape, bounds, psi_amp, kscale).ravel()
    if gw:
        KX, KY, KZ = kgrids(shape, bounds)
        dec = np.exp(-(KX ** 2 + KY ** 2 + KZ ** 2) / kscale ** 2).ravel()
        for b in range(4):
            z = (rng.standard_normal(M) + 1j * rng.standard_normal(M)) * dec * h_amp
            y[4 * N + b * Next:4 * N + (b + 1) * Next] = z.view(np.float64)
    y[-1] = a
    return y


def blocks(shape):
    """Named slices of the state vector."""
    N, M, Next, Ntot = sizes(shape)
    return {
        "phi": slice(0, N), "dphi": slice(N, 2 * N), "psi": slice(2 * N, 3 * N),
        "dpsi": slice(3 * N, 4 * N),
        "h1": slice(4 * N, 4 * N + Next), "h2": slice(4 * N + Next, 4 * N + 2 * Next),
        "dh1": slice(4 * N + 2 * Next, 4 * N + 3 * Next),
        "dh2": slice(4 * N + 3 * Next, 4 * N + 4 * Next),
        "a": slice(Ntot - 1, Ntot),
    }


def rel_l2(a, b):
    """||a - b|| / ||b|| (real or complex)."""
    a = np.asarray(a)
    b = np.asarray(b)
    nb = np.linalg.norm(b)
    d = np.linalg.norm(a - b)
    return d / nb if nb > 0 else d
