"""GPU checks of the host surface of run_dopri853 and the toolbox names:

* MAX_RUNTIME (src/dopri853.c:145-149): a run stops after its first step
  once the limit has passed, then wraps up;
* the stiffness check (src/dopri853.c:295-299, 550-580): dtlamb of a forced
  run equals the oracle's;
* the wrap-up (src/dopri853.c:223-231): with an HDF5 file, the final slice
  only when the file's buffer holds slices; the followup state written when
  ENABLE_FOLLOWUP is set and not otherwise;
* prepare_and_save_timeslice, mean and variance (src/toolbox.c:96-104,
  726-754) against the oracle and numpy;
* the HDF5 restarts set INITIAL_TIME, so that the GSL driver's output times
  i (tf - ti) / n (src/gsl_stepper.c:60) are the reference's.
"""
import ctypes

import numpy as np
import pytest

from tests.states import BOX_CONFIG, make_state, rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ih():
    import instaheat_amd
    instaheat_amd.host_lib()
    return instaheat_amd


def _sim(ih, **kw):
    args = dict(bounds=BOX_CONFIG, mass=4.0, gw=True, bins=16,
                method=ih.IH_DOPRI853, ti=0.0, tf=10.0, dt=1e-3, skip=10 ** 9)
    args.update(kw)
    return ih.Sim((16, 16, 16), **args)


def test_max_runtime_stops_after_the_first_step(ih):
    y0 = make_state((16, 16, 16), seed=2, gw=True)
    sim = _sim(ih, max_runtime=1e-9)
    sim.set_field(y0)
    saves = []
    r = sim.run_dopri853(saves=saves)
    assert r["stop"] == ih.IH_STOP_MAX_RUNTIME
    assert r["n_stp"] == 0 and r["n_ok"] + r["n_bad"] == 1
    # the initial save and the wrap-up's final slice
    assert len(saves) == 2 and saves[-1][0] == sim.t > 0.0
    sim.close()
    sim = _sim(ih, max_runtime=0.0)  # no limit
    sim.set_field(y0)
    sim.set_max_steps(2)
    r = sim.run_dopri853()
    assert r["stop"] == ih.IH_STOP_MAX_STEPS and r["n_ok"] + r["n_bad"] == 3
    sim.close()


@pytest.mark.parametrize("gw", [True, False])
def test_stiffness_check_dtlamb_matches_oracle(ih, oracle_mod, gw):
    shape = (16, 16, 16)
    y0 = make_state(shape, seed=6, gw=gw, a=1.2)
    dts = [0.01, 0.012, 0.008, 0.011]
    sim = _sim(ih, gw=gw, stiffness_check=True)
    sim.set_field(y0)
    r = sim.run_dopri853(forced=dts)
    assert r["stop"] == ih.IH_STOP_FORCED
    o = oracle_mod.Oracle(*shape, bounds=BOX_CONFIG, mass=4.0, gw=gw, bins=16)
    yo = y0.copy()
    ro = o.run_dopri853(yo, np.zeros_like(yo), 0.0, 10.0, dts[0], forced=dts,
                        skip=10 ** 9, stiffness_check=True)
    print("dtlamb GPU", r["dtlamb"], "oracle", ro["dtlamb"])
    assert ro["dtlamb"] > 0.0 and not ro["stiff"]
    # n_ok % 10 == 0 only on the first step: dtlamb of step 1
    assert r["dtlamb"] == pytest.approx(ro["dtlamb"], rel=1e-6)
    assert rel_l2(sim.field(), yo) < 1e-10
    sim.close()


def test_wrapup_follows_the_file_buffer_and_writes_the_followup(ih, tmp_path):
    from instaheat_amd import h5
    y0 = make_state((16, 16, 16), seed=8, gw=True)
    dts = [0.01, 0.01, 0.01, 0.01]
    # skip 3: saves at t0 and after step 3; the 4th step leaves the last
    # save older than t.  Buffer 2: flushed after the second save, so the
    # reference's wrap-up adds no slice; buffer 3: one slice pending, the
    # final slice is saved.
    for buf, slices, followup in ((2, 2, True), (3, 3, False)):
        sim = _sim(ih, skip=3, followup=followup)
        sim.set_field(y0)
        path = tmp_path / f"buf{buf}.h5"
        with h5.H5Output(str(path), sim.cfg, buf_size=buf, sim=sim) as out:
            r = sim.run_dopri853(forced=dts, h5=out)
            assert out.slices == slices, (buf, out.slices)
        t = h5.read(str(path), "time")
        assert len(t) == slices
        y = sim.field()
        if followup:
            f = h5.read(str(path), "followup").ravel()
            assert np.array_equal(f, y)
        else:
            with pytest.raises(Exception):
                h5.read(str(path), "followup")
        assert r["stop"] == ih.IH_STOP_FORCED
        sim.close()


def test_prepare_and_save_mean_variance(ih, oracle_mod):
    shape = (16, 16, 16)
    y0 = make_state(shape, seed=9, gw=True, a=1.1)
    sim = _sim(ih)
    sim.set_field(y0)
    t, a, obs = sim.prepare_and_save_timeslice()
    o = oracle_mod.Oracle(*shape, bounds=BOX_CONFIG, mass=4.0, gw=True, bins=16)
    yo = y0.copy()
    o.mk_rhs(yo, filter=1, output=1, field=yo)
    assert t == 0.0 and a == y0[-1]
    assert rel_l2(obs, o.obs()) < 1e-9
    assert rel_l2(sim.field(), yo) < 1e-12  # filtered in place
    # mean / variance of device arrays, and the compat names
    L, H = ih.lib(), ih.host_lib()
    N = int(np.prod(shape))
    x = np.random.default_rng(1).standard_normal(N) * 3 + 0.5
    g = ih.Gpu(shape)
    dx = g.to_device(x)
    m = ctypes.c_double()
    ih._check(L.ihg_mean(g.ctx, dx.ptr, N, ctypes.byref(m)))
    v = ctypes.c_double()
    ih._check(L.ihg_variance(g.ctx, m.value, dx.ptr, N, ctypes.byref(v)))
    assert m.value == pytest.approx(x.mean(), rel=1e-14)
    assert v.value == pytest.approx(x.var(ddof=1), rel=1e-13)
    H.ih_compat_bind(ctypes.c_void_p(sim.sim))
    H.mean.restype = ctypes.c_double
    H.mean.argtypes = [ctypes.c_void_p, ctypes.c_size_t]
    H.variance.restype = ctypes.c_double
    H.variance.argtypes = [ctypes.c_double, ctypes.c_void_p, ctypes.c_size_t]
    phi = H.ih_field(sim.sim)
    mp = H.mean(phi, N)
    vp = H.variance(mp, phi, N)
    f = sim.field()[:N]
    assert mp == pytest.approx(f.mean(), rel=1e-14)
    assert vp == pytest.approx(f.var(ddof=1), rel=1e-11)
    H.ih_compat_bind(None)
    g.close()
    sim.close()


def test_followup_restart_sets_initial_time(ih, tmp_path):
    """After h5_read_followup the GSL driver's output times are those of the
    reference: i (tf - ti) / n with ti = the file's last time.  From
    t = 0.25 with tf = 0.5 and n = 2 the first interval ends at 0.125,
    before t: GSL_EINVAL (4), as in the reference.  With INITIAL_TIME reset
    to 0 the intervals end at 0.25 and 0.5."""
    from instaheat_amd import h5
    y0 = make_state((16, 16, 16), seed=10, gw=True)
    sim = _sim(ih, tf=0.25, dt=1e-2)
    sim.set_field(y0)
    path = tmp_path / "run.h5"
    with h5.H5Output(str(path), sim.cfg, sim=sim) as out:
        sim.run_dopri853(forced=[0.0625] * 4, h5=out)
        out.write_followup()
    for reset, status in ((False, 4), (True, 0)):
        g = _sim(ih, method=ih.IH_DOPRI89, tf=0.5, dt=0.01)
        h5.read_followup(g, str(path))
        cfg = ctypes.cast(ih.host_lib().ih_cfg(g.sim),
                          ctypes.POINTER(ih.IhConfig)).contents
        assert g.t == 0.25 and cfg.ti == 0.25
        if reset:
            g.set_ti(0.0)
        st = g.run_gsl(n_output=2)
        assert st["status"] == status, (reset, st)
        if reset:
            assert g.t == pytest.approx(0.5)
        g.close()
    sim.close()
