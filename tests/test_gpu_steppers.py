"""GPU parity of the steppers: RK4 end states (src/rk4.c:27-132) and
DOPRI853 with a forced step sequence and free running
(src/dopri853.c:112-540), driven by the C99 host against the CPU oracle."""
import numpy as np
import pytest

from tests.states import BOX_CONFIG, blocks, make_state, rel_l2

pytestmark = pytest.mark.gpu

TOL_RK4 = 1e-10
TOL_DP = 1e-10


@pytest.fixture(scope="module")
def ih():
    import instaheat_amd
    instaheat_amd.host_lib()
    return instaheat_amd


def _live(shape, gw):
    B = blocks(shape)
    keys = ["phi", "dphi", "psi", "dpsi", "a"] + (["h1", "h2", "dh1", "dh2"] if gw else [])
    return B, keys


def test_rk4_config1(ih, oracle_mod):
    """The small plain case: 32^3, k_min = 0.1, m = 1, seeded Gaussian
    fluctuations of amplitude 1e-5, psi from the constraint, GW off, RK4 with
    dt = 0.01 for 500 steps."""
    shape = (32, 32, 32)
    bins = 32
    sim = ih.Sim(shape, bounds=BOX_CONFIG, mass=1.0, gw=False, filt=True,
                 bins=bins, method=ih.IH_RK4, ti=0.0, tf=5.0, dt=0.01,
                 skip=50)
    sim.init_gaussian(seed=1, amplitude=1e-5, phi0=1.0093, dphi0=-0.7137,
                      k_cut=32 * 0.1 / 4)
    sim.mk_initial_psi()
    y0 = sim.field()
    saves_g = []
    n = sim.run_rk4(saves=saves_g)
    yg = sim.field()
    o = oracle_mod.Oracle(*shape, bounds=BOX_CONFIG, mass=1.0, gw=False,
                          filt=True, bins=bins)
    yo = y0.copy()
    saves_o = []
    no = o.run_rk4(yo, 0.0, 5.0, 0.01, skip=50, saves=saves_o)
    assert n == no == 500
    B, keys = _live(shape, False)
    mask = np.zeros(o.Ntot, bool)
    for k in keys:
        mask[B[k]] = True
    e = rel_l2(yg[mask], yo[mask])
    print("rk4 end state rel L2", e, {k: rel_l2(yg[B[k]], yo[B[k]]) for k in keys})
    assert e < TOL_RK4
    assert len(saves_g) == len(saves_o) == 11
    for (tg, ag, obg), (to, ao, obo) in zip(saves_g, saves_o):
        assert tg == pytest.approx(to, rel=1e-15, abs=1e-15)
        assert ag == pytest.approx(ao, rel=1e-12)
        assert rel_l2(obg[:24], obo[:24]) < 1e-9
    # the fluctuations grew and the background rolled: a real evolution
    assert yg[-1] > 1.5 and abs(yg[0] - y0[0]) > 0.1


def _dp_case(ih, oracle_mod, shape, gw, tf, forced=None, skip=2):
    bins = 16
    y = make_state(shape, seed=21, gw=gw, amp=1e-3)
    kw = dict(bounds=BOX_CONFIG, mass=1.0, gw=gw, filt=True, bins=bins)
    o = oracle_mod.Oracle(*shape, **kw)
    yo = y.copy()
    dyo = np.zeros_like(y)
    saves_o = []
    ro = o.run_dopri853(yo, dyo, 0.0, tf, 1e-3, rtol=1e-10, atol=1e-14,
                        skip=skip, forced=forced, saves=saves_o)
    sim = ih.Sim(shape, method=ih.IH_DOPRI853, ti=0.0, tf=tf, dt=1e-3,
                 rtol=1e-10, atol=1e-14, skip=skip, **kw)
    sim.set_field(y)
    saves_g = []
    rg = sim.run_dopri853(forced=forced, saves=saves_g, log_cap=10000)
    return o, ro, yo, rg, sim.field(), saves_o, saves_g


@pytest.mark.parametrize("gw", [False, True])
def test_dopri853_forced_sequence(ih, oracle_mod, gw):
    shape = (16, 16, 16)
    # a step sequence of the adaptive method itself, then forced on both
    o0 = oracle_mod.Oracle(*shape, bounds=BOX_CONFIG, gw=gw, bins=16)
    y = make_state(shape, seed=21, gw=gw, amp=1e-3)
    r0 = o0.run_dopri853(y.copy(), np.zeros_like(y), 0.0, 0.5, 1e-3)
    forced = r0["dt"][r0["accepted"]][:12]
    assert len(forced) >= 6
    o, ro, yo, rg, yg, so, sg = _dp_case(ih, oracle_mod, shape, gw, 10.0,
                                         forced=forced)
    assert rg["n_stp"] == ro["n_stp"] == len(forced)
    B, keys = _live(shape, gw)
    errs = {k: rel_l2(yg[B[k]], yo[B[k]]) for k in keys}
    print("forced DOPRI853 end-state errors", errs)
    mask = np.zeros(o.Ntot, bool)
    for k in keys:
        mask[B[k]] = True
    assert rel_l2(yg[mask], yo[mask]) < TOL_DP
    for k in keys:
        assert errs[k] < 1e-8, k
    # The embedded error norms of these short steps are rounding noise of
    # the near-zero psi and h components (sk ~ atol there), so they agree
    # only roughly; what the controller does with them must agree: below
    # (safe/maxscale)^8 the step growth is clamped to maxscale on both sides.
    eg = np.array([r[2] for r in rg["log"]])
    clamp = (0.9 / 3.0) ** 8
    assert np.array_equal(eg < clamp, ro["err"] < clamp)
    assert np.allclose(eg, ro["err"], rtol=0.5)
    assert len(sg) == len(so)


def test_dopri853_free_running(ih, oracle_mod):
    shape = (16, 16, 16)
    o, ro, yo, rg, yg, so, sg = _dp_case(ih, oracle_mod, shape, True, 0.05,
                                         skip=3)
    print("oracle", ro["n_stp"], ro["n_ok"], ro["n_bad"], "gpu", rg["n_stp"],
          rg["n_ok"], rg["n_bad"])
    assert (rg["n_stp"], rg["n_ok"], rg["n_bad"]) == (ro["n_stp"], ro["n_ok"],
                                                       ro["n_bad"])
    dts_g = np.array([r[1] for r in rg["log"]])
    assert np.allclose(dts_g, ro["dt"], rtol=1e-8)
    assert rg["t"] == pytest.approx(ro["t"], rel=1e-14)
    B, keys = _live(shape, True)
    mask = np.zeros(o.Ntot, bool)
    for k in keys:
        mask[B[k]] = True
    assert rel_l2(yg[mask], yo[mask]) < TOL_DP
    assert len(sg) == len(so)
    for (tg, ag, obg), (to, ao, obo) in zip(sg, so):
        assert tg == pytest.approx(to, rel=1e-13)
        assert rel_l2(obg, obo) < 1e-8


def test_dopri853_free_running_32_to_t3(ih, oracle_mod):
    """32^3, GW on, the adaptive method left to itself from t = 0 to 3
    (about 80 steps, growing from 1e-3 by the controller's scaling): the
    same accepted and rejected steps with the same sizes as the oracle's
    run, and end states within 1e-9."""
    shape = (32, 32, 32)
    o, ro, yo, rg, yg, so, sg = _dp_case(ih, oracle_mod, shape, True, 3.0,
                                         skip=10)
    print("oracle", ro["n_stp"], ro["n_ok"], ro["n_bad"], "gpu", rg["n_stp"],
          rg["n_ok"], rg["n_bad"])
    assert rg["n_ok"] >= 50
    assert (rg["n_stp"], rg["n_ok"], rg["n_bad"]) == (ro["n_stp"], ro["n_ok"],
                                                       ro["n_bad"])
    dts_g = np.array([r[1] for r in rg["log"]])
    assert np.allclose(dts_g, ro["dt"], rtol=1e-8)
    acc_g = np.array([r[7] for r in rg["log"]]) > 0.5
    assert np.array_equal(acc_g, ro["accepted"])
    assert rg["t"] == pytest.approx(ro["t"], rel=1e-14)
    B, keys = _live(shape, True)
    mask = np.zeros(o.Ntot, bool)
    for k in keys:
        mask[B[k]] = True
    err = rel_l2(yg[mask], yo[mask])
    print(f"32^3 free-running DOPRI853 to t = 3: end state {err:.3e}")
    assert err < 1e-9
    assert len(sg) == len(so)
    for (tg, ag, obg), (to, ao, obo) in zip(sg, so):
        assert tg == pytest.approx(to, rel=1e-13)
        assert rel_l2(obg, obo) < 1e-8


def test_stage_kernels_match_numpy(ih):
    """ihg_dp853_stage / _finish and the RK4 kernels against numpy on random
    arrays with the tableau the library reports."""
    shape = (8, 8, 8)
    g = ih.Gpu(shape, gw=True)
    rng = np.random.default_rng(3)
    n = g.Ntot
    y = rng.standard_normal(n)
    ks = [rng.standard_normal(n) for _ in range(12)]
    dy = g.to_device(y)
    dks = [g.to_device(k) for k in ks]
    out = g.alloc(n)
    tab = ih.tableau()
    dt = 0.37
    for s in range(2, 13):
        g.dp853_stage(s, dt, dy, dks, out)
        acc = sum(tab["a"][s - 1][j] * ks[j] for j in range(s - 1))
        assert rel_l2(out.download(), y + dt * acc) < 1e-14
    ynew = g.alloc(n)
    rtol, atol = 1e-6, 1e-9
    sums = g.dp853_finish(dt, dy, dks, ynew, rtol, atol)
    bs = sum(tab["b"][j] * ks[j] for j in range(12))
    yn = y + dt * bs
    assert rel_l2(ynew.download(), yn) < 1e-14
    yerr = bs - tab["bhh"][0] * ks[0] - tab["bhh"][1] * ks[8] - tab["bhh"][2] * ks[11]
    yerr2 = sum(tab["er"][j] * ks[j] for j in range(12))
    sk = atol + rtol * np.maximum(np.abs(y), np.abs(yn))
    assert sums[0] == pytest.approx(np.sum((yerr2 / sk) ** 2), rel=1e-12)
    assert sums[1] == pytest.approx(np.sum((yerr / sk) ** 2), rel=1e-12)
    g.rk4_stage(out, dy, dks[0], dt, True)
    assert rel_l2(out.download(), y + dt * ks[0] / 2) < 1e-15
    g.rk4_stage(out, dy, dks[1], dt, False)
    assert rel_l2(out.download(), y + dt * ks[1]) < 1e-15
    g.rk4_final(dy, dks[0], dks[1], dks[2], dks[3], dt)
    ref = y + dt * (ks[0] + 2 * ks[1] + 2 * ks[2] + ks[3]) / 6
    assert rel_l2(dy.download(), ref) < 1e-15


def test_snapshot_restore_and_stream_copy(ih):
    """Sim.snapshot / restore (ihg_stream_copy, odd lengths included) give
    back the state and time exactly."""
    sim = ih.Sim((16, 16, 16), bounds=BOX_CONFIG, gw=True, method=ih.IH_DOPRI853,
                 tf=1e9, dt=1e-3)
    y = make_state((16, 16, 16), seed=9, gw=True)
    sim.set_field(y)
    snap = sim.snapshot()
    t0 = sim.t
    sim.set_max_steps(1)
    sim.run_dopri853()
    assert sim.t > t0
    sim.restore(snap)
    assert sim.t == t0
    assert np.array_equal(sim.field(), y)
    sim.drop(snap)
    g = sim.gpu if hasattr(sim, "gpu") else None
    L = ih.lib()
    for n in (1, 2, 7, 1001):
        a = ih.Gpu((4, 4, 4))
        src = a.to_device(np.arange(n, dtype=np.float64) + 0.5)
        dst = a.alloc(n)
        ih._check(L.ihg_stream_copy(a.ctx, dst.ptr, src.ptr, n))
        assert np.array_equal(dst.download(), np.arange(n) + 0.5)
