"""GPU checks of the device-side initial data and the .dat format: the
seeded Gaussian state against the oracle's restatement, and .dat export /
import (IC_FROM_DAT_FILE_WITH_PSI and _WITHOUT_PSI) round trips."""
import os

import numpy as np
import pytest

from oracle import ic

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ih():
    import instaheat_amd
    instaheat_amd.host_lib()
    return instaheat_amd


def _sim(ih, shape, L=2 * np.pi / 0.1, gw=True, a0=1.0):
    return ih.Sim(shape, bounds=((0.0, L),) * 3, mass=1.0, gw=gw, bins=16,
                  method=ih.IH_RK4, tf=1.0, dt=0.01, a_initial=a0)


@pytest.mark.parametrize("shape,gw,a0,amp,kcut", [((32, 32, 32), True, 1.0, 1e-5, 0.8),
                                                  ((16, 64, 8), False, 1.3, 1e-4, 1.5),
                                                  ((64, 64, 64), True, 1.0, 1e-4, 1e9)])
def test_device_ic_matches_oracle(ih, shape, gw, a0, amp, kcut):
    L = 2 * np.pi / 0.1
    sim = _sim(ih, shape, gw=gw, a0=a0)
    sim.init_gaussian(seed=9, amplitude=amp, phi0=1.0093, dphi0=-0.7137, k_cut=kcut)
    y = sim.field()
    want = ic.gaussian_state(shape, ((0.0, L),) * 3, 1.0, 9, amp, 1.0093, -0.7137,
                             a0, kcut, sim.Ntot)
    N = sim.N
    d, dw = y[:N] - 1.0093, want[:N] - 1.0093
    err = np.linalg.norm(d - dw) / np.linalg.norm(dw)
    print(shape, "delta phi rel L2 vs oracle", err)
    assert err < 1e-13
    assert np.array_equal(y[N:], want[N:])
    sim.close()


def test_dat_round_trip(ih, tmp_path):
    shape = (16, 8, 32)
    a = _sim(ih, shape)
    a.init_gaussian(seed=2, amplitude=1e-3, k_cut=2.0)
    a.mk_initial_psi()
    ya = a.field()
    p = tmp_path / "ic.dat"
    a.write_dat(p)
    lines = open(p).read().splitlines()
    assert len(lines) == a.N
    i, j, k = (int(v) for v in lines[8 * 32 + 3].split()[:3])
    assert (i, j, k) == (1, 0, 3)
    b = _sim(ih, shape)
    b.read_dat(p, with_psi=True)
    yb = b.field()
    N = a.N
    assert np.array_equal(yb[:4 * N], ya[:4 * N])
    assert yb[-1] == 1.0 and np.all(yb[4 * N:-1] == 0)
    # WITHOUT_PSI: psi, dpsi come from mk_initial_psi of the file's phi, dphi
    c = _sim(ih, shape)
    c.read_dat(p, with_psi=False)
    yc = c.field()
    assert np.array_equal(yc[:2 * N], ya[:2 * N])
    assert np.allclose(yc[2 * N:4 * N], ya[2 * N:4 * N], rtol=0, atol=1e-15 * np.abs(ya[2 * N:4 * N]).max())
    # a truncated file is an error and names nothing else
    with open(p) as fh:
        head = fh.readlines()[:N // 2]
    q = tmp_path / "short.dat"
    q.write_text("".join(head))
    with pytest.raises(RuntimeError):
        b.read_dat(q)
    with pytest.raises(RuntimeError):
        b.read_dat(tmp_path / "missing.dat")
    for s in (a, b, c):
        s.close()
