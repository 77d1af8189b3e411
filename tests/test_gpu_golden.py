"""The HIP path against the committed golden fixtures (tests/golden/):
FFTs against numpy's vectors, and the RHS with observables, RK4 and forced
DOPRI853 end states and the device initial data against the oracle's
vectors on a 16 x 8 x 8 grid, GW on and off."""
import os

import numpy as np
import pytest

from tests.states import BOX_CONFIG, rel_l2

pytestmark = pytest.mark.gpu

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SHAPE = (16, 8, 8)


@pytest.fixture(scope="module")
def ih():
    import instaheat_amd
    instaheat_amd.host_lib()
    return instaheat_amd


@pytest.fixture(scope="module")
def grid():
    return np.load(os.path.join(G, "grid_16x8x8.npz"))


@pytest.mark.parametrize("shape", [(16, 8, 4), (8, 8, 8), (4, 16, 32)])
def test_fft_matches_numpy_vectors(ih, shape):
    d = np.load(os.path.join(G, "fft_numpy.npz"))
    tag = "x".join(str(v) for v in shape)
    x, X = d[f"fft_{tag}_in"], d[f"fft_{tag}_out"]
    g = ih.Gpu(shape)
    assert rel_l2(g.rfftn(x), X) < 1e-13
    assert rel_l2(g.irfftn(X), x * x.size) < 1e-13


@pytest.mark.parametrize("gw", [True, False])
def test_rhs_and_steppers_match_grid_vectors(ih, grid, gw):
    tag = "gw" if gw else "nogw"
    y = grid[f"grid_{tag}_rhs_in"]
    g = ih.Gpu(SHAPE, bounds=BOX_CONFIG, mass=4.0, gw=gw, bins=8)
    f = g.to_device(y)
    r = g.alloc(g.Ntot)
    g.set_field(f)
    g.mk_rhs(f, r, flags=3)
    assert rel_l2(f.download(), grid[f"grid_{tag}_rhs_filtered"]) < 1e-12
    assert rel_l2(r.download(), grid[f"grid_{tag}_rhs_out"]) < 1e-12
    assert rel_l2(g.observables(), grid[f"grid_{tag}_obs"]) < 1e-9
    kw = dict(bounds=BOX_CONFIG, mass=4.0, gw=gw, bins=8, skip=10 ** 9)
    s = ih.Sim(SHAPE, method=ih.IH_RK4, ti=0.0, tf=0.2, dt=0.01, **kw)
    s.set_field(y)
    assert s.run_rk4() == grid[f"grid_{tag}_rk4_steps"][0]
    assert rel_l2(s.field(), grid[f"grid_{tag}_rk4_end"]) < 1e-10
    s.close()
    s = ih.Sim(SHAPE, method=ih.IH_DOPRI853, ti=0.0, tf=1.0, dt=1e-3, **kw)
    s.set_field(y)
    s.run_dopri853(forced=grid[f"grid_{tag}_dp_dts"])
    assert rel_l2(s.field(), grid[f"grid_{tag}_dp_end"]) < 1e-10
    s.close()


def test_device_ic_matches_vector(ih, grid):
    seed, amp, phi0, dphi0, a0, kcut = grid["ic_params"]
    want = grid["ic_state"]
    s = ih.Sim(SHAPE, bounds=BOX_CONFIG, mass=4.0, gw=True, bins=8,
               method=ih.IH_RK4, tf=1.0, dt=0.01, a_initial=a0)
    s.init_gaussian(seed=int(seed), amplitude=amp, phi0=phi0, dphi0=dphi0,
                    k_cut=kcut)
    y = s.field()
    N = int(np.prod(SHAPE))
    assert rel_l2(y[:N] - phi0, want[:N] - phi0) < 1e-13
    assert np.array_equal(y[N:], want[N:])
    s.close()
