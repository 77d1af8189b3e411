"""GPU checks of DOPRI853 dense output (time-based output): the library's
interpolated states inside forced steps against the oracle's DOP853 step,
extra stages and interpolant (oracle/dense.py with scipy's independent
table), and the observables of a dense output against the oracle's output
call on the oracle's interpolated state."""
import numpy as np
import pytest
from scipy.integrate._ivp import dop853_coefficients as SC

from oracle import dense
from tests.states import BOX_CONFIG, make_state, rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ih():
    import instaheat_amd
    instaheat_amd.host_lib()
    return instaheat_amd


@pytest.mark.parametrize("shape,gw", [((16, 16, 16), True), ((8, 16, 32), False)])
def test_dense_output_matches_oracle(ih, oracle_mod, shape, gw):
    bins = 16
    y0 = make_state(shape, seed=21, gw=gw, a=1.2)
    dts = [0.04, 0.05]
    t0 = 0.3
    t1 = t0 + dts[0]
    # inside step 1, its end, inside step 2, its end (the same sums the
    # steps make, so the ends are exact)
    t_out = [t0 + 0.011, t1, t1 + 0.031, t1 + dts[1]]
    sim = ih.Sim(shape, bounds=BOX_CONFIG, mass=4.0, gw=gw, bins=bins,
                 method=ih.IH_DOPRI853, ti=t0, tf=10.0, skip=10 ** 9)
    sim.set_field(y0)
    sim.t = t0
    got, obs_gpu = {}, {}
    saves = []

    def grab(t):
        # the first save at an output time is the dense one (the final
        # step-based save may fall on the same time)
        if t in t_out and t not in got:
            got[t] = sim.dense_state()
            obs_gpu[t] = saves[-1][2]
    sim.run_dopri853(forced=dts, saves=saves, t_out=t_out, on_save=grab)
    assert sorted(got) == t_out

    o = oracle_mod.Oracle(*shape, bounds=BOX_CONFIG, mass=4.0, gw=gw, bins=bins)

    def rhs(t, y):
        f = y.copy()
        return o.mk_rhs(f, t=t, filter=0, output=0, field=f)
    y = y0.copy()
    k1 = o.mk_rhs(y, t=t0, filter=1, output=1, field=y)
    t = t0
    for h in dts:
        K, y_new = dense.step(rhs, t, y, k1, h, SC.A, SC.B, SC.C)
        K.append(o.mk_rhs(y_new, t=t + h, filter=1, output=0, field=y_new))
        for i in range(3):
            st = 13 + i
            yi = dense.extra_stage_input(y, K, h, SC.A[st][:st])
            K.append(rhs(t + SC.C[st] * h, yi))
        for tt in t_out:
            if t < tt <= t + h:
                want = dense.interpolate(y, y_new, K, h, (tt - t) / h, SC.D)
                err = rel_l2(got[tt], want)
                print(shape, "dense output at t =", tt, "rel L2", err)
                assert err < 1e-10
                # observables of the interpolated state (summaries read it)
                w = want.copy()
                o.mk_rhs(w, t=tt, filter=0, output=1, field=w)
                ob = o.obs()
                assert rel_l2(obs_gpu[tt], ob) < 1e-9
        t, y, k1 = t + h, y_new, K[12]
    sim.close()


def test_dense_end_points(ih):
    """theta = 0 gives the step's start, theta = 1 its end; the FSAL count
    of RHS evaluations grows by 3 per step that holds an output time"""
    shape = (16, 16, 16)
    y0 = make_state(shape, seed=3, gw=True, a=1.0)
    sim = ih.Sim(shape, bounds=BOX_CONFIG, mass=4.0, gw=True, bins=16,
                 method=ih.IH_DOPRI853, ti=0.0, tf=10.0, skip=10 ** 9, filt=False)
    sim.set_field(y0)
    sim.t = 0.0
    ends = {}
    sim.run_dopri853(forced=[0.03, 0.03], t_out=[0.03],
                     on_save=lambda t: ends.setdefault(t, sim.dense_state())
                     if t == 0.03 else None)
    calls_dense = ih.lib().ihg_calls_rhs(sim.ctx)
    y1 = None
    sim2 = ih.Sim(shape, bounds=BOX_CONFIG, mass=4.0, gw=True, bins=16,
                  method=ih.IH_DOPRI853, ti=0.0, tf=10.0, skip=10 ** 9, filt=False)
    sim2.set_field(y0)
    sim2.t = 0.0
    sim2.run_dopri853(forced=[0.03])
    y1 = sim2.field()
    calls_plain = ih.lib().ihg_calls_rhs(sim2.ctx)
    # one step more (12 RHS) plus 3 extra stages and 1 output call
    assert calls_dense - calls_plain == 12 + 3 + 1
    assert rel_l2(ends[0.03], y1) < 1e-15
    sim.close()
    sim2.close()


def test_time_based_log_binned_outputs(ih):
    """The 64^3 GW-on DOPRI853 configuration with outputs every 0.5 time
    units (dense output) and 64 log-spaced spectrum bins, to t = 2."""
    n, bins = 64, 64
    L = 2 * np.pi / 0.1
    sim = ih.Sim((n, n, n), bounds=((0.0, L),) * 3, mass=1.0, gw=True,
                 bins=bins, method=ih.IH_DOPRI853, ti=0.0, tf=2.0, dt=1e-7,
                 rtol=1e-10, atol=1e-14, hubble_frac=1e-2, skip=10 ** 9,
                 log_bins=True)
    sim.init_gaussian(seed=2, amplitude=1e-4, k_cut=n * 0.1 / 4)
    sim.mk_initial_psi()
    saves = []
    t_out = [0.5, 1.0, 1.5, 2.0]
    sim.run_dopri853(saves=saves, t_out=t_out)
    times = [t for t, _, _ in saves]
    for t in t_out:
        assert t in times
    for t, a, obs in saves:
        o = ih.unpack_obs(obs, bins)
        assert np.all(np.isfinite(o["phi_ps"])) and np.all(np.isfinite(o["gw"]))
        assert o["phi_ps"].sum() > 0
    assert sim.t == pytest.approx(2.0)
    sim.close()


def _rho(ih, sim, N):
    """rho of the simulation's last mk_rhs (ihg_download_strided, stride 1)"""
    import ctypes
    out = np.zeros(N)
    st = (ctypes.c_int * 3)(1, 1, 1)
    rc = ih.lib().ihg_download_strided(sim.ctx, None, st,
                                       out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert rc == 0, ih.lib().ihg_last_error()
    return out


@pytest.mark.parametrize("gw", [True, False])
def test_step_saves_unchanged_by_dense_saves(ih, gw):
    """Steps with a step-based save (skip = 1) and dense saves inside them:
    the step-based saves carry the observables and rho of the step's own
    output call, those of the same forced run without time-based outputs
    (bit for bit but for the spectra's atomic bin sums), and a t_out equal
    to a step's end is saved once."""
    shape = (16, 16, 16)
    N = 16 ** 3
    y0 = make_state(shape, seed=5, gw=gw, a=1.1)
    dts = [0.02, 0.03, 0.025]
    t0 = 0.2
    ends = list(np.cumsum([t0] + dts)[1:])
    # interior points of steps 1 and 3, and the end of step 2 (exactly the
    # sum the stepper forms)
    t_e2 = (t0 + dts[0]) + dts[1]
    t_out = [t0 + 0.007, t_e2, t_e2 + 0.01, t_e2 + 0.02]

    def run(with_dense):
        sim = ih.Sim(shape, bounds=BOX_CONFIG, mass=4.0, gw=gw, bins=16,
                     method=ih.IH_DOPRI853, ti=t0, tf=10.0, skip=1)
        sim.set_field(y0)
        sim.t = t0
        saves, rhos, dense_flag = [], [], []

        def grab(t):
            rhos.append(_rho(ih, sim, N))
            dense_flag.append(sim.dense_state_active())
        sim.run_dopri853(forced=dts, saves=saves,
                         t_out=t_out if with_dense else None, on_save=grab)
        y = sim.field()
        sim.close()
        return saves, rhos, dense_flag, y

    s_plain, r_plain, _, y_plain = run(False)
    s_dense, r_dense, flag, y_dense = run(True)
    np.testing.assert_array_equal(y_dense, y_plain)
    step_saves = [(s, r) for s, r, f in zip(s_dense, r_dense, flag) if not f]
    assert len(step_saves) == len(s_plain) == len(r_plain)
    for (sd, rd), sp, rp in zip(step_saves, s_plain, r_plain):
        assert sd[0] == sp[0] and sd[1] == sp[1]
        # summaries and constraints (fixed-order reductions) bit for bit; the
        # spectra to the rounding of their atomic bin sums
        np.testing.assert_array_equal(sd[2][:28], sp[2][:28])
        np.testing.assert_allclose(sd[2][28:], sp[2][28:], rtol=1e-13, atol=0)
        np.testing.assert_array_equal(rd, rp)
    times = [s[0] for s in s_dense]
    assert times == sorted(times)
    assert times.count(t_e2) == 1
    dense_times = [s[0] for s, f in zip(s_dense, flag) if f]
    assert dense_times == [t for t in t_out if t != t_e2]
    assert [s[0] for s in s_plain][1:] == pytest.approx(ends, abs=0)
