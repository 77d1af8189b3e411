"""GPU checks of the precision form (ihg_params.gw_fp32): the GW source
(products, FFTs, TT projection) and the GW spectrum in FP32, everything else
FP64.  A single RHS against the FP64 oracle, and the GW spectrum per k bin
of a DOPRI853 run against the FP64-only run over the same forced steps."""
import numpy as np
import pytest

from tests.states import BOX_CONFIG, blocks, make_state, rel_l2

pytestmark = pytest.mark.gpu

NON_H = ("phi", "dphi", "psi", "dpsi", "a")


@pytest.fixture(scope="module")
def ih():
    import instaheat_amd
    instaheat_amd.host_lib()
    return instaheat_amd


@pytest.mark.parametrize("shape,flags", [((16, 16, 16), 3), ((32, 16, 8), 1),
                                         ((64, 64, 64), 0)])
def test_fp32_gw_rhs(ih, oracle_mod, shape, flags):
    y = make_state(shape, seed=5, gw=True, a=1.3)
    out = {}
    for fp32 in (False, True):
        g = ih.Gpu(shape, bounds=BOX_CONFIG, mass=4.0, gw=True, bins=20,
                   gw_fp32=fp32)
        f = g.to_device(y)
        r = g.alloc(g.Ntot)
        g.mk_rhs(f, r, flags=flags)
        out[fp32] = r.download()
    o = oracle_mod.Oracle(*shape, bounds=BOX_CONFIG, mass=4.0, gw=True, bins=20)
    fo = y.copy()
    ro = o.mk_rhs(fo, filter=flags & 1, output=(flags >> 1) & 1, field=fo)
    B = blocks(shape)
    err = {k: rel_l2(out[True][B[k]], ro[B[k]]) for k in ("dh1", "dh2")}
    print("FP32 GW source, dh errors against the FP64 oracle", err)
    for k in NON_H:
        assert np.array_equal(out[True][B[k]], out[False][B[k]]), k
    for k in ("dh1", "dh2"):
        assert err[k] < 2e-6
        assert rel_l2(out[False][B[k]], ro[B[k]]) < 1e-12


def test_fp32_gw_spectrum_per_bin(ih):
    """32^3, GW on, amplitude 1e-4, DOPRI853 rtol 1e-10 to t = 10: the GW
    spectrum of the precision form per k bin against the FP64-only run over
    the same step sequence."""
    shape, bins = (32, 32, 32), 16
    L = 2 * np.pi / 0.1
    runs = {}
    forced = None
    for fp32 in (False, True):
        sim = ih.Sim(shape, bounds=((0.0, L),) * 3, mass=1.0, gw=True,
                     bins=bins, method=ih.IH_DOPRI853, tf=10.0, dt=1e-7,
                     rtol=1e-10, atol=1e-14, hubble_frac=1e-2, skip=10**9,
                     gw_fp32=fp32)
        sim.init_gaussian(seed=3, amplitude=1e-4, k_cut=32 * 0.1 / 4)
        sim.mk_initial_psi()
        saves = []
        r = sim.run_dopri853(forced=forced, saves=saves, log_cap=10000)
        if forced is None:
            acc = r["log"][:, 7] > 0.5
            forced = r["log"][acc, 1]
        runs[fp32] = (sim.t, ih.unpack_obs(saves[-1][2], bins)["gw"])
    assert runs[True][0] == pytest.approx(runs[False][0], rel=1e-14)
    g64, g32 = runs[False][1], runs[True][1]
    live = g64 > 1e-6 * g64.max()
    rel = np.abs(g32[live] - g64[live]) / g64[live]
    print("GW spectrum relative error per k bin (FP32 source vs FP64):",
          np.array2string(rel, precision=2))
    assert live.sum() >= bins // 2
    assert rel.max() < 1e-4
