"""GPU parity of the GSL stepper path (src/gsl_stepper.c:41-76: RKF45,
RKCK, DOPRI89 = RK8PD under GSL's driver and standard control) against the
oracle's restatement (oracle/gsl.py) on identical seeded states: the same
attempts and rejections, every interval ending on its output time, the
saved observables, the suggested next step and the end state."""
import numpy as np
import pytest

from oracle import gsl
from tests.states import BOX_CONFIG, make_state, rel_l2

pytestmark = pytest.mark.gpu

# fixed steps (the same h on both sides): the state gate of a forced step
# sequence
TOL_STATE = 1e-10
TOL_OBS = 1e-9
# The standard control's ratio max_i |yerr_i| / D_i divides an embedded
# difference that cancels the stage sums down to about rtol of their size,
# so the ~1e-15 rounding differences of the RHS (FFT order of operations)
# reach it at about 1e-15 / rtol relative: the free-running runs must make
# the same accept / reject decisions, with the suggested step equal to
# TOL_H and the end states, whose step sizes then differ at that level,
# within a few rtol.
TOL_H = 1e-3


@pytest.fixture(scope="module")
def ih():
    import instaheat_amd
    instaheat_amd.host_lib()
    return instaheat_amd


# Free-running cases whose decisions do not hinge on rounding, which
# tests/test_gsl.py checks on the oracle (other reduction orders, the RHS
# perturbed by up to 1e-13 relative: the same attempts and rejections).
# RK8PD at tight tolerances is not such a case: its error estimate sinks to
# the rounding level of the stage sums, and the control then follows the
# noise (32^3, rtol 1e-7, 5 intervals took 14, 15 or 16 attempts with the
# oracle's reductions at different thread counts).
CASES = [
    # method, shape, gw, tf, n_output, h0, rtol, atol
    (gsl.RKF45, (16, 16, 16), True, 0.4, 4, 0.1, 1e-9, 1e-12),
    (gsl.RKCK, (16, 16, 16), True, 0.4, 4, 0.1, 1e-9, 1e-12),
    (gsl.DOPRI89, (16, 16, 16), True, 0.4, 4, 0.1, 1e-6, 1e-9),
    (gsl.RKCK, (8, 16, 32), False, 0.3, 3, 0.05, 1e-9, 1e-12),
    (gsl.DOPRI89, (32, 32, 32), True, 1.0, 4, 0.05, 1e-6, 1e-9),
]


@pytest.mark.parametrize("method,shape,gw,tf,nout,h0,rtol,atol", CASES)
def test_gsl_run_matches_oracle(ih, oracle_mod, method, shape, gw, tf, nout, h0,
                                rtol, atol):
    bins = 16
    y0 = make_state(shape, seed=12, gw=gw, a=1.1)
    sim = ih.Sim(shape, bounds=BOX_CONFIG, mass=4.0, gw=gw, bins=bins,
                 method=method, ti=0.0, tf=tf, dt=h0, dt_min=1e-9, rtol=rtol,
                 atol=atol)
    sim.set_field(y0)
    sim.t = 0.0
    saves = []
    st = sim.run_gsl(nout, saves=saves)
    y = sim.field()
    o = oracle_mod.Oracle(*shape, bounds=BOX_CONFIG, mass=4.0, gw=gw, bins=bins)
    status, t, want, d, osaves = gsl.run(o, y0, method, 0.0, tf, h0, atol, rtol,
                                         1e-9, nout, obs=[])
    print(method, shape, "attempts", st["attempts"], d.count, "failed",
          st["failed_steps"], d.failed, "h", st["h"], d.h)
    assert st["status"] == status == gsl.GSL_SUCCESS
    assert st["attempts"] == d.count
    assert st["failed_steps"] == d.failed
    assert st["steps"] == d.count - d.failed
    assert st["intervals"] == nout
    assert st["t"] == t == sim.t
    assert st["h"] == pytest.approx(d.h, rel=TOL_H)
    err = rel_l2(y, want)
    print("end state rel L2", err)
    assert err < max(TOL_STATE, 10 * rtol)
    assert [s[0] for s in saves] == [s[0] for s in osaves]
    for (tg, _, og), (to, oo) in zip(saves, osaves):
        assert rel_l2(og, oo) < max(TOL_OBS, 100 * rtol), tg
    # per RHS call: one k1 per accepted step and stages - 1 per attempt,
    # plus the output calls
    stages = gsl.tableau(method)["stages"]
    assert st["calls_rhs"] == st["steps"] + (stages - 1) * st["attempts"] + nout + 1
    sim.close()


@pytest.mark.parametrize("method,shape,gw", [(gsl.RKF45, (16, 16, 16), True),
                                              (gsl.RKCK, (16, 16, 16), True),
                                              (gsl.DOPRI89, (16, 16, 16), True),
                                              (gsl.RKF45, (8, 16, 32), False),
                                              (gsl.DOPRI89, (32, 8, 16), False)])
def test_gsl_fixed_steps_match_oracle(ih, oracle_mod, method, shape, gw):
    """four steps of fixed sizes through the C ABI (mk_rhs, ihg_gsl_stage,
    ihg_gsl_finish) against the oracle's step: every state to the forced
    step gate, every error ratio to TOL_H (with GW off the kernels skip the
    h blocks, whose entries stay zero and count in the maximum as 0)"""
    g = ih.Gpu(shape, bounds=BOX_CONFIG, mass=4.0, gw=gw, bins=8)
    o = oracle_mod.Oracle(*shape, bounds=BOX_CONFIG, mass=4.0, gw=gw, bins=8)
    tab = gsl.tableau(method)
    S = tab["stages"]
    y0 = make_state(shape, seed=19, gw=gw, a=1.05)
    n = g.Ntot
    dy, dn, tmp = g.alloc(n), g.alloc(n), g.alloc(n)
    ks = [g.alloc(n) for _ in range(S)]
    dy.upload(y0)

    def f(t, yy):
        w = yy.copy()
        return o.mk_rhs(w, t=t, filter=0, output=0, field=w)
    y, t = y0.copy(), 0.2
    atol, rtol = 1e-12, 1e-9
    # step sizes whose error ratios lie well above the rounding level
    hs = (0.05, 0.1, 0.08, 0.15) if method == gsl.DOPRI89 else (0.01, 0.02, 0.015, 0.03)
    for h in hs:
        g.mk_rhs(dy, ks[0], t=t)
        for s in range(2, S + 1):
            g.gsl_stage(method, s, h, dy, ks, tmp)
            g.mk_rhs(tmp, ks[s - 1], t=t + tab["c"][s - 1] * h)
        r = g.gsl_finish(method, h, dy, ks, dn, atol, rtol)
        y_new, yerr, _ = gsl.step(f, t, y, f(t, y), h, tab)
        r_want = gsl.rmax(y_new, yerr, atol, rtol)
        err = rel_l2(dn.download(), y_new)
        print(method, "h", h, "state", err, "ratio", r, r_want)
        assert r_want > 1.0
        assert err < TOL_STATE
        assert r == pytest.approx(r_want, rel=TOL_H)
        dy, dn = dn, dy
        y, t = y_new, t + h
    g.close()


def test_gsl_enoprog_matches_oracle(ih, oracle_mod):
    """a tolerance that needs steps below hmin: both stop with GSL_ENOPROG
    at the same time, after the same attempts, with no save for the
    unfinished interval"""
    shape = (16, 16, 16)
    y0 = make_state(shape, seed=5, gw=True, a=1.0)
    sim = ih.Sim(shape, bounds=BOX_CONFIG, mass=4.0, gw=True, bins=8,
                 method=ih.IH_RKF45, ti=0.0, tf=1.0, dt=0.2, dt_min=0.01,
                 rtol=1e-13, atol=1e-16)
    sim.set_field(y0)
    sim.t = 0.0
    saves = []
    st = sim.run_gsl(2, saves=saves)
    o = oracle_mod.Oracle(*shape, bounds=BOX_CONFIG, mass=4.0, gw=True, bins=8)
    status, t, want, d, osaves = gsl.run(o, y0, gsl.RKF45, 0.0, 1.0, 0.2, 1e-16,
                                         1e-13, 0.01, 2, obs=[])
    assert status == st["status"] == gsl.GSL_ENOPROG
    assert st["attempts"] == d.count and st["failed_steps"] == d.failed
    # the accepted step comes from the control's ratio (TOL_H)
    assert st["t"] == pytest.approx(t, rel=TOL_H)
    assert len(saves) == len(osaves) == 1
    assert rel_l2(sim.field(), want) < 1e-6
    sim.close()


def test_gsl_methods_are_rejected_by_the_other_steppers(ih):
    sim = ih.Sim((8, 8, 8), bounds=BOX_CONFIG, bins=8, method=ih.IH_RKCK, tf=0.1, dt=0.01)
    with pytest.raises(ih.IhgError):
        sim.run_gsl(0)
    sim.close()
    sim = ih.Sim((8, 8, 8), bounds=BOX_CONFIG, bins=8, method=ih.IH_DOPRI853, tf=0.1, dt=0.01)
    with pytest.raises(ih.IhgError):
        sim.run_gsl(2)
    sim.close()


def test_gsl_initial_time_quirk_matches_oracle(ih, oracle_mod):
    """with INITIAL_TIME > 0 the reference's first output time
    (tf - ti) / n lies before ti: GSL's driver returns GSL_EINVAL and the
    run ends after the initial save, on both sides"""
    shape = (8, 8, 8)
    y0 = make_state(shape, seed=2, gw=False, a=1.0)
    sim = ih.Sim(shape, bounds=BOX_CONFIG, mass=4.0, gw=False, bins=8,
                 method=ih.IH_RKCK, ti=0.5, tf=1.0, dt=0.01)
    sim.set_field(y0)
    sim.t = 0.5
    saves = []
    st = sim.run_gsl(2, saves=saves)
    o = oracle_mod.Oracle(*shape, bounds=BOX_CONFIG, mass=4.0, gw=False, bins=8)
    status, t, want, d, osaves = gsl.run(o, y0, gsl.RKCK, 0.5, 1.0, 0.01, 1e-14,
                                         1e-10, 1e-9, 2, obs=[])
    assert st["status"] == status == gsl.GSL_EINVAL
    assert st["steps"] == 0 == d.count
    assert len(saves) == len(osaves) == 1
    assert st["t"] == t == 0.5
    sim.close()
