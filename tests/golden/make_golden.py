"""Writes the golden fixtures of tests/golden/ (run from the repository root:
python tests/golden/make_golden.py).

The reference holds no numeric fixtures for its hot path, so the vectors
come from two sources built here:
  fft_*     numpy's pocketfft (independent of this project's code): the
            unnormalised r2c of seeded real arrays, FFTW's half layout;
  grid_*    the oracle (oracle/, the CPU restatement of the reference) on a
            16 x 8 x 8 grid, box [0, 2 pi / 0.1)^3, m = 4: one filtered
            output mk_rhs with its observables, GW on and off; an RK4 end
            state (20 steps, dt 0.01); a forced-step DOPRI853 end state
            (5 steps); and the seeded Gaussian initial data of oracle/ic.py.
Inputs are stored beside the outputs, so a fixture does not depend on the
generators in tests/states.py staying the same.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import ic  # noqa: E402
from oracle import oracle as orc  # noqa: E402
from tests.states import BOX_CONFIG, make_state  # noqa: E402

SHAPE = (16, 8, 8)
FFT_SHAPES = ((16, 8, 4), (8, 8, 8), (4, 16, 32))


def fft_fixtures():
    out = {}
    rng = np.random.default_rng(2024)
    for s in FFT_SHAPES:
        x = rng.standard_normal(s)
        tag = "x".join(str(v) for v in s)
        out[f"fft_{tag}_in"] = x
        out[f"fft_{tag}_out"] = np.fft.rfftn(x, axes=(0, 1, 2))
    return out


def grid_fixtures():
    out = {}
    for gw in (True, False):
        tag = "gw" if gw else "nogw"
        o = orc.Oracle(*SHAPE, bounds=BOX_CONFIG, mass=4.0, gw=gw, bins=8,
                       threads=1)
        y = make_state(SHAPE, seed=77, gw=gw, a=1.25)
        out[f"grid_{tag}_rhs_in"] = y.copy()
        f = y.copy()
        r = o.mk_rhs(f, filter=1, output=1, field=f)
        out[f"grid_{tag}_rhs_filtered"] = f
        out[f"grid_{tag}_rhs_out"] = r
        out[f"grid_{tag}_obs"] = o.obs()
        # RK4, 20 steps of 0.01 from the same input
        f = y.copy()
        n = o.run_rk4(f, 0.0, 0.2, 0.01, skip=10 ** 9)
        out[f"grid_{tag}_rk4_end"] = f
        out[f"grid_{tag}_rk4_steps"] = np.array([n])
        # DOPRI853 along a forced step sequence
        f = y.copy()
        df = np.zeros_like(f)
        dts = np.array([1e-3, 4e-3, 0.01, 0.02, 0.02])
        o.run_dopri853(f, df, 0.0, 1.0, 1e-3, forced=dts, skip=10 ** 9)
        out[f"grid_{tag}_dp_dts"] = dts
        out[f"grid_{tag}_dp_end"] = f
    N = int(np.prod(SHAPE))
    M = SHAPE[0] * SHAPE[1] * (SHAPE[2] // 2 + 1)
    ntot = 4 * N + 8 * M + 1
    out["ic_state"] = ic.gaussian_state(SHAPE, BOX_CONFIG, 4.0, 11, 1e-4, 1.0093,
                                        -0.7137, 1.1, 3.0, ntot)
    out["ic_params"] = np.array([11, 1e-4, 1.0093, -0.7137, 1.1, 3.0])
    return out


def main():
    np.savez_compressed(os.path.join(HERE, "fft_numpy.npz"), **fft_fixtures())
    np.savez_compressed(os.path.join(HERE, "grid_16x8x8.npz"), **grid_fixtures())
    for f in ("fft_numpy.npz", "grid_16x8x8.npz"):
        print(f, os.path.getsize(os.path.join(HERE, f)), "bytes")


if __name__ == "__main__":
    main()
