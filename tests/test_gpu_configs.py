"""GPU parity at the full workloads of the three added forms and of the
64^3 DOPRI853 configuration:

* precision: 256^3, k_min 0.05, the GW source and spectrum in FP32 -- one
  filtered output RHS against the FP64 oracle, and the GW spectrum per k bin
  after 40 DOPRI853 steps against the FP64 GPU run over the same steps;
* batched: the mixed ensemble 48 x 32^3 + 16 x 64^3 + 4 x 128^3 (seeds
  1..68, amplitudes log-uniform in [1e-6, 1e-4], RK4 dt 0.005, GW and
  filter on) for 20 steps in HIP graphs, per size and grouped (one launch
  per pass for all sizes), the two equal bit for bit -- one realization per
  size equal to run_rk4 of that realization alone, bit for bit, and one
  realization's RHS per size against the FP64 oracle;
* learned: 256^3 FP64 DOPRI853 rtol 1e-10 with the step proposer from a
  state in the error-controlled phase (t > 20) -- its end state after four
  accepted steps against the oracle taking the same steps;
* the 64^3 DOPRI853 configuration with GW on, outputs every 0.5 and 64
  log-spaced bins, t = 0 to 100 -- every output saved, and 6 steps from
  t = 50 against the oracle taking the same steps.

Each case prints its name and shape before it touches the GPU."""
import numpy as np
import pytest

from tests.states import blocks, make_state, rel_l2

pytestmark = pytest.mark.gpu

K_MIN = 0.05
BOX_256 = ((0.0, 2 * np.pi / K_MIN),) * 3
NON_H = ("phi", "dphi", "psi", "dpsi", "a")


@pytest.fixture(scope="module")
def ih():
    import instaheat_amd
    instaheat_amd.host_lib()
    return instaheat_amd


@pytest.mark.timeout(600)
def test_precision_form_rhs_256(ih, oracle_mod):
    """256^3 filtered output RHS: with the FP32 GW source the dh blocks are
    within 2e-6 of the FP64 oracle and every other block (and the filtered
    state) equals the FP64 path bit for bit; the FP64 path is within 1e-12
    of the oracle.  The GW spectrum of the FP32 output call per k bin is
    within 1e-4 of the oracle's."""
    shape = (256, 256, 256)
    bins = 64
    y = make_state(shape, bounds=BOX_256, seed=9, gw=True, a=1.4)
    out, filt, obs = {}, {}, {}
    for fp32 in (False, True):
        print(f"precision 256^3 RHS, gw_fp32={fp32}", flush=True)
        g = ih.Gpu(shape, bounds=BOX_256, mass=1.0, gw=True, bins=bins,
                   gw_fp32=fp32)
        f = g.to_device(y)
        r = g.alloc(g.Ntot)
        g.mk_rhs(f, r, flags=ih.IHG_FILTER | ih.IHG_OUTPUT)
        out[fp32], filt[fp32], obs[fp32] = r.download(), f.download(), g.observables()
        del f, r
        g.close()
    print("precision 256^3: oracle RHS", flush=True)
    o = oracle_mod.Oracle(*shape, bounds=BOX_256, mass=1.0, gw=True, bins=bins)
    fo = y.copy()
    ro = o.mk_rhs(fo, filter=1, output=1, field=fo)
    oo = o.obs()
    B = blocks(shape)
    err32 = {k: rel_l2(out[True][B[k]], ro[B[k]]) for k in ("dh1", "dh2")}
    err64 = {k: rel_l2(out[False][B[k]], ro[B[k]]) for k in B}
    print("256^3 FP32 GW source: dh vs FP64 oracle", err32,
          "FP64 path max block", max(err64.values()))
    for k in NON_H:
        assert np.array_equal(out[True][B[k]], out[False][B[k]]), k
    assert np.array_equal(filt[True], filt[False])
    for k in ("dh1", "dh2"):
        assert err32[k] < 2e-6
    assert rel_l2(out[False][:-1], ro[:-1]) < 1e-12
    assert max(err64.values()) < 1e-9
    gw32 = ih.unpack_obs(obs[True], bins)["gw"]
    gwo = ih.unpack_obs(oo, bins)["gw"]
    live = gwo > 1e-6 * gwo.max()
    rel = np.abs(gw32[live] - gwo[live]) / gwo[live]
    print("256^3 GW spectrum FP32 vs FP64 oracle per bin: max", rel.max(),
          "bins", int(live.sum()))
    assert live.sum() >= bins // 2
    assert rel.max() < 1e-4
    assert rel_l2(obs[False], oo) < 1e-9


@pytest.mark.timeout(600)
def test_precision_form_spectrum_256_steps(ih):
    """256^3 k_min 0.05, amplitude 1e-5, DOPRI853 rtol 1e-10: 40 steps of the
    FP64 form, then the precision form over the same steps; the GW spectrum
    per k bin of the two end states agrees to 1e-4, and every FP64 block of
    the end states to 1e-12 (the FP32 source reaches them only through dh)."""
    n, bins = 256, 64
    runs = {}
    forced = None
    for fp32 in (False, True):
        print(f"precision 256^3 DOPRI853 40 steps, gw_fp32={fp32}", flush=True)
        sim = ih.Sim((n, n, n), bounds=BOX_256, mass=1.0, gw=True, bins=bins,
                     method=ih.IH_DOPRI853, tf=1e9, dt=1e-7, rtol=1e-10,
                     atol=1e-14, hubble_frac=1e-2, skip=10 ** 9,
                     gw_fp32=fp32)
        sim.init_gaussian(seed=1, amplitude=1e-5, k_cut=n * K_MIN / 4)
        sim.mk_initial_psi()
        saves = []
        if forced is None:
            sim.set_max_steps(39)
            r = sim.run_dopri853(saves=saves, log_cap=1000)
            forced = r["log"][r["log"][:, 7] > 0.5, 1]
        else:
            sim.run_dopri853(forced=forced, saves=saves)
        runs[fp32] = (sim.t, ih.unpack_obs(saves[-1][2], bins)["gw"],
                      sim.field())
        sim.close()
    assert len(forced) == 40
    assert runs[True][0] == runs[False][0]
    g64, g32 = runs[False][1], runs[True][1]
    live = g64 > 1e-6 * g64.max()
    rel = np.abs(g32[live] - g64[live]) / g64[live]
    print(f"256^3 GW spectrum at t = {runs[False][0]:.4f}, FP32 vs FP64 per "
          f"bin: max {rel.max():.3e} median {np.median(rel):.3e} over "
          f"{int(live.sum())} bins")
    assert live.sum() >= bins // 2
    assert rel.max() < 1e-4
    B = blocks((n, n, n))
    for k in NON_H:
        assert rel_l2(runs[True][2][B[k]], runs[False][2][B[k]]) < 1e-12, k


MIXED = ((32, 48), (64, 16), (128, 4))


@pytest.mark.timeout(600)
def test_mixed_ensemble_full_size(ih, oracle_mod):
    """The mixed ensemble at its full size for 20 graph-replayed RK4 steps,
    per size and grouped (every pass of all sizes one launch), the two equal
    bit for bit; the last realization of each size equals run_rk4 of that realization
    alone (its final prepare_and_save filter applied to the ensemble's state
    too), bit for bit; the RHS of that realization's end state matches the
    oracle to 1e-12; the ensemble mean and variance of the GW spectrum are
    those of the per-realization spectra."""
    from instaheat_amd import ensemble as E
    steps, dt, kmin, bins = 20, 0.005, 0.1, 32
    rng = np.random.default_rng(1)
    ens, seeds, amps = [], {}, {}
    first = 1
    for n, nreal in MIXED:
        print(f"mixed ensemble: {nreal} x {n}^3", flush=True)
        e = E.Ensemble(n, nreal, kmin=kmin, bins=bins)
        a = 10.0 ** rng.uniform(-6.0, -4.0, nreal)
        s = list(range(first, first + nreal))
        e.init(s, a)
        seeds[n], amps[n] = s, a
        first += nreal
        ens.append(e)
    print("mixed ensemble: 20 RK4 steps in graphs", flush=True)
    E.run(ens, steps, dt, use_graph=True)
    # the grouped form (every pass of the three sizes one launch, one graph
    # per step) from the same initial data: the same states bit for bit
    grp = []
    for n, nreal in MIXED:
        e = E.Ensemble(n, nreal, kmin=kmin, bins=bins)
        e.init(seeds[n], amps[n])
        grp.append(e)
    print("mixed ensemble: 20 grouped RK4 steps in graphs", flush=True)
    E.run(grp, steps, dt, use_graph=True, grouped=True)
    for a, b in zip(ens, grp):
        assert np.array_equal(a.state(), b.state()), a.n
        b.close()
    L = 2 * np.pi / kmin
    for e in ens:
        n, r = e.n, e.nreal - 1
        assert e.t == pytest.approx(steps * dt, rel=1e-14)
        y_ens = e.state()[r]
        print(f"mixed ensemble: run_rk4 of {n}^3 realization {r}", flush=True)
        sim = ih.Sim((n, n, n), bounds=((0.0, L),) * 3, mass=1.0, gw=True,
                     filt=True, bins=bins, method=ih.IH_RK4, tf=1.0,
                     max_steps=steps, dt=dt, skip=10 ** 9)
        sim.init_gaussian(seed=seeds[n][r], amplitude=amps[n][r],
                          k_cut=n * kmin / 4)
        sim.mk_initial_psi()
        sim.run_rk4()
        y_sim = sim.field()
        sim.close()
        # run_rk4 ends with prepare_and_save, whose filtered output call
        # filters the state once more: the same call on the ensemble's state
        g = ih.Gpu((n, n, n), bounds=((0.0, L),) * 3, mass=1.0, gw=True,
                   filt=True, bins=bins)
        f = g.to_device(y_ens)
        res = g.alloc(g.Ntot)
        g.set_field(f)
        g.mk_rhs(f, res, flags=ih.IHG_FILTER | ih.IHG_OUTPUT)
        assert np.array_equal(f.download(), y_sim), n
        # the RHS of the end state (unfiltered call) against the oracle
        f2 = g.to_device(y_ens)
        g.mk_rhs(f2, res, flags=0)
        rg = res.download()
        g.close()
        o = oracle_mod.Oracle(n, n, n, bounds=((0.0, L),) * 3, mass=1.0,
                              gw=True, filt=True, bins=bins)
        fo = y_ens.copy()
        ro = o.mk_rhs(fo, filter=0, output=0, field=fo)
        B = blocks((n, n, n))
        err = {k: rel_l2(rg[B[k]], ro[B[k]]) for k in B}
        print(f"{n}^3 realization {r}: RHS vs oracle", max(err.values()))
        assert rel_l2(rg[:-1], ro[:-1]) < 1e-12
        assert max(err.values()) < 1e-9
        gw = e.gw_spectra()
        assert gw.shape == (e.nreal, bins) and np.all(np.isfinite(gw))
        assert np.all(gw.mean(0) >= 0) and np.all(gw.var(0, ddof=1) >= 0)
    for e in ens:
        e.close()


@pytest.mark.timeout(900)
def test_tuned_form_256_matches_oracle_on_its_steps(ih, oracle_mod):
    """256^3 FP64 DOPRI853 rtol 1e-10, GW on: the classic controller to
    t > 20 (the error-controlled phase), 8 steps of the tuned form to fill
    its history, then 4 accepted tuned steps; the oracle takes the same 4
    steps from the same state, and the end states agree within 1e-9."""
    from instaheat_amd.proposer import MlpProposer
    n, bins = 256, 64
    print("tuned form 256^3: pre-roll to t = 20", flush=True)
    sim = ih.Sim((n, n, n), bounds=BOX_256, mass=1.0, gw=True, bins=bins,
                 method=ih.IH_DOPRI853, tf=20.0, dt=1e-7, rtol=1e-10,
                 atol=1e-14, hubble_frac=1e-2, skip=10 ** 9)
    sim.init_gaussian(seed=1, amplitude=1e-5, k_cut=n * K_MIN / 4)
    sim.mk_initial_psi()
    w = sim.run_dopri853()
    assert sim.t == pytest.approx(20.0, rel=1e-14)
    prop = MlpProposer.load()
    sim.set_tf(1e9)
    sim.set_dt(w["dt_next"])
    sim.set_max_steps(7)
    h = sim.run_dopri853(proposer=prop)
    print("tuned form 256^3: 4 steps from t =", sim.t, flush=True)
    y0, t0 = sim.field(), sim.t
    sim.set_dt(h["dt_next"])
    sim.set_max_steps(3)
    r = sim.run_dopri853(proposer=prop, log_cap=100)
    log = r["log"]
    dts = log[log[:, 7] > 0.5, 1]
    assert len(dts) == 4
    # the controller's phase: steps below the Hubble cap 1e-2 sqrt(3/rho)
    caps = 1e-2 * np.sqrt(3.0 / log[log[:, 7] > 0.5, 3])
    print(f"tuned steps {dts}, Hubble caps {caps}, rejected {r['n_bad']}, "
          f"proposals {prop.calls}")
    assert np.any(dts < 0.99 * caps)
    assert prop.calls > 0
    yg = sim.field()
    sim.close()
    print("tuned form 256^3: oracle replays the 4 steps", flush=True)
    o = oracle_mod.Oracle(n, n, n, bounds=BOX_256, mass=1.0, gw=True,
                          filt=True, bins=bins)
    yo = y0.copy()
    ro = o.run_dopri853(yo, np.zeros_like(yo), t0, 1e9, dts[0], rtol=1e-10,
                        atol=1e-14, hubble_frac=1e-2, skip=10 ** 9, forced=dts)
    assert ro["t"] == pytest.approx(t0 + dts.sum(), rel=1e-14)
    err = rel_l2(yg, yo)
    print(f"256^3 tuned form end state vs the oracle on its 4 steps: {err:.3e}")
    assert err < 1e-9


@pytest.mark.timeout(900)
def test_config2_dopri853_to_t100(ih, oracle_mod):
    """The 64^3 DOPRI853 configuration at its workload: GW and filter on,
    k_min 0.1, amplitude 1e-4, rtol 1e-10, atol 1e-14, outputs every 0.5
    time units (dense output) with 64 log-spaced spectrum bins, t = 0 to
    100.  At t = 50 the oracle takes the GPU's next 6 accepted steps from
    the same state, and the end states agree within 1e-9; every one of the
    200 output times is saved, with finite observables and positive
    spectra; the scale factor grows at every save."""
    n, bins, kmin = 64, 64, 0.1
    bounds = ((0.0, 2 * np.pi / kmin),) * 3
    kw = dict(bounds=bounds, mass=1.0, gw=True, filt=True, bins=bins)
    t_out = np.arange(1, 201) * 0.5
    sim = ih.Sim((n, n, n), method=ih.IH_DOPRI853, ti=0.0, tf=50.0, dt=1e-7,
                 rtol=1e-10, atol=1e-14, hubble_frac=1e-2, skip=10 ** 9,
                 log_bins=True, **kw)
    sim.init_gaussian(seed=2, amplitude=1e-4, k_cut=n * kmin / 4)
    sim.mk_initial_psi()
    saves = []
    print("config 2, 64^3 DOPRI853: t = 0 to 50", flush=True)
    r1 = sim.run_dopri853(saves=saves, t_out=t_out[t_out <= 50.0])
    assert sim.t == pytest.approx(50.0, rel=1e-14)
    # the window the oracle replays: 6 accepted steps from t = 50
    y50, t50 = sim.field(), sim.t
    sim.set_tf(100.0)
    sim.set_dt(r1["dt_next"])
    sim.set_max_steps(5)
    rw = sim.run_dopri853(saves=saves, t_out=t_out[t_out > t50], log_cap=100)
    log = rw["log"]
    dts = log[log[:, 7] > 0.5, 1]
    assert len(dts) == 6
    yw, tw = sim.field(), sim.t
    print(f"config 2: window of 6 steps from t = {t50} to {tw}; to t = 100",
          flush=True)
    sim.set_max_steps(10 ** 15)
    sim.set_dt(rw["dt_next"])
    r2 = sim.run_dopri853(saves=saves, t_out=t_out[t_out > tw])
    assert sim.t == pytest.approx(100.0, rel=1e-14)
    n_ok = r1["n_ok"] + rw["n_ok"] + r2["n_ok"]
    n_bad = r1["n_bad"] + rw["n_bad"] + r2["n_bad"]
    print(f"config 2: {n_ok} + {n_bad} rejected steps to t = 100, "
          f"{ih.lib().ihg_calls_rhs(sim.ctx)} RHS, a = {sim.field()[-1]:.4f}",
          flush=True)
    assert n_bad < 0.1 * n_ok
    times = np.array([t for t, _, _ in saves])
    for t in t_out:
        assert np.any(np.abs(times - t) < 1e-9), t
    a_prev = 0.0
    for t, a, obs in sorted(saves, key=lambda s: s[0]):
        o = ih.unpack_obs(obs, bins)
        assert np.all(np.isfinite(obs)), t
        assert o["phi_ps"].sum() > 0 and np.all(o["gw"] >= 0), t
        if t >= 1.0:  # h grows from zero
            assert o["gw"].sum() > 0, t
        assert a >= a_prev, t
        a_prev = a
    assert a_prev > 1.5
    sim.close()
    print("config 2: the oracle replays the 6 steps from t = 50", flush=True)
    o = oracle_mod.Oracle(n, n, n, **kw)
    yo = y50.copy()
    ro = o.run_dopri853(yo, np.zeros_like(yo), t50, 100.0, dts[0], rtol=1e-10,
                        atol=1e-14, hubble_frac=1e-2, skip=10 ** 9, forced=dts)
    assert ro["t"] == pytest.approx(tw, rel=1e-14)
    err = rel_l2(yw, yo)
    print(f"64^3 end state vs the oracle on the GPU's 6 steps: {err:.3e}")
    assert err < 1e-9
