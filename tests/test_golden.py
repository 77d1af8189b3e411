"""The oracle against the committed golden fixtures (tests/golden/, written
by tests/golden/make_golden.py): its FFT against numpy's vectors, and its
RHS, observables, RK4 and forced DOPRI853 end states and initial data
against the vectors it wrote (a change to the oracle shows here).  No GPU."""
import os

import numpy as np
import pytest

from oracle import ic
from oracle import oracle as orc
from tests.states import BOX_CONFIG, rel_l2

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SHAPE = (16, 8, 8)


@pytest.fixture(scope="module")
def fft():
    return np.load(os.path.join(G, "fft_numpy.npz"))


@pytest.fixture(scope="module")
def grid():
    return np.load(os.path.join(G, "grid_16x8x8.npz"))


@pytest.mark.parametrize("shape", [(16, 8, 4), (8, 8, 8), (4, 16, 32)])
def test_oracle_fft_matches_numpy_vectors(fft, shape):
    tag = "x".join(str(v) for v in shape)
    x, X = fft[f"fft_{tag}_in"], fft[f"fft_{tag}_out"]
    o = orc.Oracle(*shape)
    got = o.rfftn(x)
    assert rel_l2(got, X) < 1e-14
    back = o.irfftn(X)
    assert rel_l2(back, x * x.size) < 1e-14


@pytest.mark.parametrize("gw", [True, False])
def test_oracle_reproduces_grid_vectors(grid, gw):
    tag = "gw" if gw else "nogw"
    o = orc.Oracle(*SHAPE, bounds=BOX_CONFIG, mass=4.0, gw=gw, bins=8)
    y = grid[f"grid_{tag}_rhs_in"]
    f = y.copy()
    r = o.mk_rhs(f, filter=1, output=1, field=f)
    assert rel_l2(f, grid[f"grid_{tag}_rhs_filtered"]) < 1e-14
    assert rel_l2(r, grid[f"grid_{tag}_rhs_out"]) < 1e-13
    assert rel_l2(o.obs(), grid[f"grid_{tag}_obs"]) < 1e-12
    f = y.copy()
    assert o.run_rk4(f, 0.0, 0.2, 0.01, skip=10 ** 9) == grid[f"grid_{tag}_rk4_steps"][0]
    assert rel_l2(f, grid[f"grid_{tag}_rk4_end"]) < 1e-13
    f = y.copy()
    o.run_dopri853(f, np.zeros_like(f), 0.0, 1.0, 1e-3,
                   forced=grid[f"grid_{tag}_dp_dts"], skip=10 ** 9)
    assert rel_l2(f, grid[f"grid_{tag}_dp_end"]) < 1e-13


def test_oracle_ic_reproduces_vector(grid):
    seed, amp, phi0, dphi0, a0, kcut = grid["ic_params"]
    want = grid["ic_state"]
    got = ic.gaussian_state(SHAPE, BOX_CONFIG, 4.0, int(seed), amp, phi0, dphi0,
                            a0, kcut, want.size)
    assert np.array_equal(got, want)
