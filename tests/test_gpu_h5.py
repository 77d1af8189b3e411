"""GPU checks of the HDF5 output file written from the host's save() hook
(include/instaheat_h5.h, src/io.c): every slice of a DOPRI853 run with
step-based and dense outputs holds exactly the time, a, observables, the
strided full fields of the saved state and rho of its output call; the
meta data and the followup snapshot; IC_FROM_H5_FILE restarts (followup,
time slice) continue a run as if it had not stopped."""
import ctypes

import numpy as np
import pytest

from tests.states import BOX_CONFIG, make_state

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ih():
    import instaheat_amd
    instaheat_amd.host_lib()
    return instaheat_amd


@pytest.fixture(scope="module")
def h5():
    from instaheat_amd import h5 as mod
    mod.h5_lib()
    return mod


def _strided(y, shape, stride):
    return y.reshape(shape)[::stride[0], ::stride[1], ::stride[2]].ravel()


def test_dopri853_run_file_matches_saves(ih, h5, tmp_path):
    shape, bins, stride = (16, 16, 16), 12, (2, 1, 4)
    N = int(np.prod(shape))
    sim = ih.Sim(shape, bounds=BOX_CONFIG, mass=4.0, gw=True, bins=bins,
                 method=ih.IH_DOPRI853, ti=0.0, tf=10.0, skip=2)
    y0 = make_state(shape, seed=31, gw=True, a=1.1)
    sim.set_field(y0)
    sim.t = 0.0
    p = tmp_path / "run.h5"
    w = h5.H5Output(p, sim.cfg, outputs=h5.OUT_DEFAULT | h5.OUT_FIELDS, buf_size=3,
                    stride=stride, seed=31, sim=sim)
    saves, states, rhos, dense = [], [], [], []
    L, H = ih.lib(), ih.host_lib()

    def grab(t):
        states.append(sim._dl(H.ih_save_state(sim.sim), sim.Ntot))
        r = np.empty(N)
        assert L.ihg_aux(sim.ctx, 3, r.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) == 0
        rhos.append(r)
        dense.append(t in t_out)
    dts = [0.01, 0.02, 0.02, 0.03, 0.03, 0.03, 0.04]
    t_out = [0.015, 0.07, 0.1]
    st = sim.run_dopri853(forced=dts, saves=saves, t_out=t_out, on_save=grab, h5=w)
    assert w.status == 0
    assert w.slices == len(saves) > 6
    w.write_dp_stats(st)
    w.write_followup()
    final = sim.field()
    w.close()
    times = np.array([s[0] for s in saves])
    assert np.array_equal(h5.read(p, "time"), times)
    assert np.array_equal(h5.read(p, "a"), np.array([s[1] for s in saves]))
    # the a of each slice is the saved state's (dense outputs included)
    assert np.array_equal(h5.read(p, "a"), np.array([y[-1] for y in states]))
    obs = np.array([s[2] for s in saves])
    assert np.array_equal(h5.read(p, "phi_summary"), obs[:, 0:4])
    assert np.array_equal(h5.read(p, "pressure_summary"), obs[:, 20:24])
    assert np.array_equal(h5.read(p, "constraints"), obs[:, 24:28])
    assert np.array_equal(h5.read(p, "gravitational_wave_spectrum"),
                          obs[:, 28 + 3 * bins:28 + 4 * bins])
    for j, n in enumerate(["phi", "dphi", "psi", "dpsi"]):
        want = np.array([_strided(y[j * N:(j + 1) * N], shape, stride) for y in states])
        assert np.array_equal(h5.read(p, n), want), n
    want = np.array([_strided(r, shape, stride) for r in rhos])
    assert np.array_equal(h5.read(p, "rho"), want)
    # the phi summary of the first, last and dense-output slices describes
    # that slice's state; those of the step-based slices in between read
    # the state before the step, as the reference's mk_summary reads the
    # global `field` during try_step (src/toolbox.c:634-647)
    assert sum(dense) == len(t_out)
    for i, y in enumerate(states):
        if dense[i] or i in (0, len(states) - 1):
            assert abs(obs[i, 0] - y[:N].mean()) <= 1e-12 * abs(y[:N]).max(), i
    assert h5.read(p, "steps_total")[0] == st["n_stp"]
    assert h5.read(p, "steps_ok")[0] == st["n_ok"]
    assert h5.read(p, "calls_rhs")[0] == st["calls_rhs"]
    assert np.array_equal(h5.read(p, "followup"), final)
    sim.close()


def _rk4(ih, shape, tf, filt=False, skip=1, dt=0.01):
    return ih.Sim(shape, bounds=BOX_CONFIG, mass=4.0, gw=True, bins=8,
                  method=ih.IH_RK4, ti=0.0, tf=tf, dt=dt, skip=skip, filt=filt)


def test_followup_restart_continues_the_run(ih, h5, tmp_path):
    """IC_FROM_H5_FILE (src/setup.c:780-781): a run stopped after 10 steps
    with its followup snapshot, read into a new simulation and continued
    for 10 more, ends bit-identical to the run of 20 steps that did not stop
    (RK4 without filter: each step depends only on the state and t; dt a
    power of two, so the times add up exactly)"""
    shape = (16, 16, 16)
    dt = 2.0 ** -7
    y0 = make_state(shape, seed=8, gw=True, a=1.0)
    a = _rk4(ih, shape, 20 * dt, dt=dt)
    a.set_field(y0)
    a.run_rk4()
    want = a.field()
    b = _rk4(ih, shape, 10 * dt, dt=dt)
    b.set_field(y0)
    p = tmp_path / "first.h5"
    with h5.H5Output(p, b.cfg, sim=b) as w:
        b.run_rk4(h5=w)
        w.write_followup()
    t_stop = b.t
    assert t_stop == 10 * dt
    # an RK4 run's file ends with main's calls_rhs and runtime_total
    assert h5.read(p, "calls_rhs")[0] == ih.lib().ihg_calls_rhs(b.ctx)
    assert h5.read(p, "runtime_total")[0] > 0.0
    # closing the writer after its simulation is destroyed is safe
    e = _rk4(ih, shape, 2 * dt, dt=dt)
    e.set_field(y0)
    q = tmp_path / "closed_late.h5"
    w2 = h5.H5Output(q, e.cfg, sim=e)
    e.run_rk4(h5=w2)
    calls = ih.lib().ihg_calls_rhs(e.ctx)
    e.close()
    w2.close()
    assert h5.read(q, "calls_rhs")[0] == calls
    assert h5.read(p, "time")[-1] == t_stop
    c = _rk4(ih, shape, 20 * dt, dt=dt)
    h5.read_followup(c, p)
    assert c.t == t_stop
    assert np.array_equal(c.field(), b.field())
    c.run_rk4()
    assert np.array_equal(c.field(), want)
    # a simulation of another size refuses the snapshot
    d = _rk4(ih, (16, 16, 32), 20 * dt, dt=dt)
    with pytest.raises(ih.IhgError, match="specifications"):
        h5.read_followup(d, p)
    for s in (a, b, c, d):
        s.close()


def test_timeslice_initial_data(ih, h5, tmp_path):
    """h5_read_timeslice (src/io.c:640-711): the first slice at or after t
    (DBL_EPSILON slack), else the last; phi, dphi, psi, dpsi and a from it,
    the metric perturbations zero"""
    shape = (16, 8, 16)
    N = int(np.prod(shape))
    s = _rk4(ih, shape, 0.05)
    s.set_field(make_state(shape, seed=4, gw=True, a=1.2))
    p = tmp_path / "slices.h5"
    states = []
    with h5.H5Output(p, s.cfg, outputs=h5.OUT_DEFAULT | (h5.OUT_FIELDS & ~h5.OUT_RHO),
                     sim=s) as w:
        s.run_rk4(h5=w, on_save=lambda t: states.append(
            s._dl(ih.host_lib().ih_save_state(s.sim), s.Ntot)))
    times = h5.read(p, "time")
    assert len(times) == len(states) == 6
    assert np.allclose(times, [0.0, 0.01, 0.02, 0.03, 0.04, 0.05], rtol=0, atol=1e-15)
    for t, want_i in [(0.025, 3), (times[2], 2), (0.0, 0), (9.0, 5),
                      (times[4] + 1e-17, 4)]:
        r = _rk4(ih, shape, 1.0)
        got_t = h5.read_timeslice(r, p, t)
        assert got_t == times[want_i] and r.t == times[want_i]
        y = r.field()
        assert np.array_equal(y[:4 * N], states[want_i][:4 * N])
        assert y[-1] == states[want_i][-1]
        assert not np.any(y[4 * N:-1])
        r.close()
    # a file without full fields at stride 1 cannot seed a simulation
    q = tmp_path / "smry.h5"
    with h5.H5Output(q, s.cfg) as w:
        w.append(0.0, 1.0, np.zeros(28 + 4 * 8))
    r = _rk4(ih, shape, 1.0)
    with pytest.raises(ih.IhgError, match="no dataset phi"):
        h5.read_timeslice(r, q, 0.0)
    r.close()
    s.close()


def test_download_strided_matches_numpy(ih):
    """ihg_download_strided (the device gather under full-field output)
    against numpy slicing, rho included; bad strides are errors"""
    shape = (16, 8, 32)
    N = int(np.prod(shape))
    g = ih.Gpu(shape, bounds=BOX_CONFIG, mass=4.0, gw=True, bins=8)
    y = make_state(shape, seed=3, gw=True, a=1.0)
    d, res = g.alloc(g.Ntot), g.alloc(g.Ntot)
    d.upload(y)
    g.mk_rhs(d, res, t=0.0)
    rho = g.aux(3)
    L = ih.lib()
    dp = ctypes.POINTER(ctypes.c_double)
    for stride in ((1, 1, 1), (2, 1, 4), (3, 5, 7), (16, 8, 32), (1, 8, 1)):
        st = (ctypes.c_int * 3)(*stride)
        want = y[N:2 * N].reshape(shape)[::stride[0], ::stride[1], ::stride[2]].ravel()
        out = np.full(want.size, np.nan)
        assert L.ihg_download_strided(g.ctx, d.ptr + 8 * N, st, out.ctypes.data_as(dp)) == 0
        assert np.array_equal(out, want), stride
        want = rho.reshape(shape)[::stride[0], ::stride[1], ::stride[2]].ravel()
        out = np.full(want.size, np.nan)
        assert L.ihg_download_strided(g.ctx, None, st, out.ctypes.data_as(dp)) == 0
        assert np.array_equal(out, want), ("rho", stride)
    out = np.zeros(N)
    for bad in ((0, 1, 1), (1, 9, 1), (1, 1, -2)):
        st = (ctypes.c_int * 3)(*bad)
        assert L.ihg_download_strided(g.ctx, d.ptr, st, out.ctypes.data_as(dp)) != 0
    g.close()


def test_c_host_writes_the_file(h5, tmp_path):
    """the C host of the integration guide (tests/test_h5.py) run on the
    GPU: DOPRI853 at 16^3 with ih_h5_save as its save() hook"""
    import subprocess
    from tests.test_h5 import build_c_host
    rc, err, exe = build_c_host(tmp_path)
    assert rc == 0, err
    p = tmp_path / "c_run.h5"
    r = subprocess.run([str(exe), str(p)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    t = h5.read(p, "time")
    assert t[0] == 0.0 and t[-1] == pytest.approx(0.5)
    n = h5.read(p, "steps_total")[0]
    assert n > 0 and len(t) >= n / 5
    assert h5.read(p, "phi_power_spectrum").shape == (len(t), 50)
    assert h5.dims(p, "followup")[0] == 4 * 16 ** 3 + 4 * 2 * 16 * 16 * 9 + 1
    assert h5.read(p, "calls_rhs")[0] > 12 * n
    assert h5.read(p, "runtime_total")[0] > 0.0
