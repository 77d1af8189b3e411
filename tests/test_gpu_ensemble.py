"""GPU checks of the batched form: an ensemble context's one-launch RHS of
every realization equals the single-realization RHS of each, its RK4 steps
equal single-realization RK4 (run_rk4), and the HIP-graph replay of a step
equals the eager step after every replay."""
import ctypes

import numpy as np
import pytest

from tests.states import BOX_CONFIG, make_state, sizes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ih():
    import instaheat_amd
    instaheat_amd.host_lib()
    return instaheat_amd


@pytest.fixture(scope="module")
def E():
    from instaheat_amd import ensemble
    return ensemble


@pytest.mark.parametrize("shape,gw,flags,fp32", [((16, 16, 16), True, 0, False),
                                                 ((16, 8, 32), True, 1, False),
                                                 ((32, 32, 32), False, 1, False),
                                                 ((16, 16, 16), True, 1, True)])
def test_ensemble_rhs_equals_single_realizations(ih, shape, gw, flags, fp32):
    nreal = 3
    p = ih.make_params(shape, BOX_CONFIG, 4.0, gw, True, 20, fp32)
    h = ctypes.c_void_p()
    ih._check(ih.lib().ihg_create_ensemble(ctypes.byref(p), nreal, ctypes.byref(h)))
    ctx = h.value
    ens = (ctypes.c_long * 2)()
    ih._check(ih.lib().ihg_ensemble(ctx, ens))
    stride = ens[1]
    g = ih.Gpu(shape, bounds=BOX_CONFIG, mass=4.0, gw=gw, bins=20, gw_fp32=fp32)
    ys = [make_state(shape, seed=40 + r, gw=gw, a=1.1 + 0.2 * r) for r in range(nreal)]
    full = np.zeros((nreal, stride))
    for r, y in enumerate(ys):
        full[r, :y.size] = y
    L = ih.lib()
    n = nreal * stride
    fd = ctypes.c_void_p()
    rd = ctypes.c_void_p()
    ih._check(L.ihg_malloc(ctx, ctypes.byref(fd), n))
    ih._check(L.ihg_malloc(ctx, ctypes.byref(rd), n))
    ih._check(L.ihg_upload(ctx, fd, full.ctypes.data_as(ctypes.c_void_p), n))
    ih._check(L.ihg_mk_rhs_ensemble(ctx, 0.0, fd, rd, flags))
    got = np.empty(n)
    ih._check(L.ihg_download(ctx, got.ctypes.data_as(ctypes.c_void_p), rd, n))
    filtered = np.empty(n)
    ih._check(L.ihg_download(ctx, filtered.ctypes.data_as(ctypes.c_void_p), fd, n))
    got = got.reshape(nreal, stride)
    filtered = filtered.reshape(nreal, stride)
    for r, y in enumerate(ys):
        f = g.to_device(y)
        res = g.alloc(g.Ntot)
        g.mk_rhs(f, res, flags=flags)
        assert np.array_equal(got[r, :y.size], res.download()), r
        assert np.array_equal(filtered[r, :y.size], f.download()), r
    L.ihg_free(ctx, fd)
    L.ihg_free(ctx, rd)
    L.ihg_destroy(ctx)


@pytest.mark.parametrize("gw", [True, False])
def test_ensemble_rk4_equals_run_rk4(ih, E, gw):
    """Realization r of an ensemble stepped by RK4 (the stage sums folded
    into the RHS kernels) equals run_rk4 of one simulation from the same
    initial data (separate stage sweeps), bit for bit.  (Filter off on both
    sides: run_rk4's final prepare_and_save filters the state once more.)"""
    n, nreal, steps, dt = 16, 3, 6, 0.01
    e = E.Ensemble(n, nreal, kmin=0.1, bins=16, filt=False, gw=gw)
    seeds, amps = [7, 8, 9], [1e-5, 3e-5, 1e-4]
    e.init(seeds, amps)
    E.run([e], steps, dt, use_graph=False)
    ens_end = e.state()
    L = 2 * np.pi / 0.1
    for r in range(nreal):
        # max_steps rather than tf, so that no last step is shortened
        sim = ih.Sim((n, n, n), bounds=((0.0, L),) * 3, mass=1.0, gw=gw,
                     filt=False, bins=16, method=ih.IH_RK4, tf=1.0,
                     max_steps=steps, dt=dt, skip=10**9)
        sim.init_gaussian(seed=seeds[r], amplitude=amps[r], k_cut=n * 0.1 / 4)
        sim.mk_initial_psi()
        sim.run_rk4()
        assert np.array_equal(sim.field(), ens_end[r]), r
    e.close()


def test_graph_replay_equals_eager_steps(ih, E):
    """Two ensembles (16^3 x 3, 32^3 x 2) in one joined graph: after every
    replay, with eager work and synchronisation in between, the states equal
    the eager steps exactly."""
    def make():
        a = E.Ensemble(16, 3, bins=16)
        a.init([1, 2, 3], [1e-5, 2e-5, 4e-5])
        b = E.Ensemble(32, 2, bins=16)
        b.init([4, 5], [1e-4, 1e-6])
        return a, b
    ea, eb = make()
    ga, gb = make()
    dt = 0.005
    H, L = ih.host_lib(), ih.lib()
    pa, pb = ctypes.c_void_p(), ctypes.c_void_p()
    ih._check(H.ih_ens_capture_step(ga.h, dt, ctypes.byref(pa)))
    ih._check(H.ih_ens_capture_step(gb.h, dt, ctypes.byref(pb)))
    parts = (ctypes.c_void_p * 2)(pa.value, pb.value)
    joined = ctypes.c_void_p()
    ih._check(L.ihg_graph_join(parts, 2, ctypes.byref(joined)))
    for step in range(4):
        E.run([ea, eb], 1, dt, use_graph=False)
        ih._check(L.ihg_graph_launch(joined, ga.ctx))
        ih._check(L.ihg_sync(ga.ctx))
        # eager work of the program itself between the replays
        _ = ea.observables(0)
        assert np.array_equal(ga.state(), ea.state()), step
        assert np.array_equal(gb.state(), eb.state()), step
    L.ihg_graph_destroy(joined)
    for p in parts:
        L.ihg_graph_destroy(p)
    for e in (ea, eb, ga, gb):
        e.close()


def test_mixed_ensemble_short(E):
    """The mixed-size ensemble machinery at reduced size: graph and eager
    stepping agree (the states exactly; the GW spectra to the rounding of
    their atomic bin sums), and the GW spectrum statistics are finite."""
    sizes = ((16, 4), (32, 2))
    a = E.run_mixed(nsteps=5, dt=0.005, use_graph=True, sizes=sizes, bins=16,
                    keep_states=True)
    b = E.run_mixed(nsteps=5, dt=0.005, use_graph=False, sizes=sizes, bins=16,
                    keep_states=True)
    for n, _ in sizes:
        assert np.array_equal(a["sizes"][n]["states"], b["sizes"][n]["states"])
        assert np.allclose(a["sizes"][n]["gw_mean"], b["sizes"][n]["gw_mean"],
                           rtol=1e-12, atol=0)
        assert np.all(np.isfinite(a["sizes"][n]["gw_var"]))
        assert a["sizes"][n]["t"] == pytest.approx(0.025)
    assert a["rhs_evals"] == 4 * 5 * 6


def _candidates(ih, ctx):
    cnt, cur = (ctypes.c_int * 2)(), (ctypes.c_int * 2)()
    ih._check(ih.lib().ihg_fft_candidates(ctx, cnt, cur))
    return list(cnt), list(cur)


@pytest.mark.parametrize("shape", [(16, 16, 16), (64, 32, 16), (8, 128, 32),
                                   (256, 16, 4)])
def test_every_fft_candidate_matches_oracle(ih, oracle_mod, shape):
    """Each (plan, work-group) candidate of the strided y and x passes, set by
    hand: the filtered RHS against the oracle, every block within 1e-12."""
    from tests.states import blocks, rel_l2
    y = make_state(shape, seed=11, gw=True, a=1.2)
    o = oracle_mod.Oracle(*shape, bounds=BOX_CONFIG, mass=4.0, gw=True, bins=20)
    fo = y.copy()
    ro = o.mk_rhs(fo, filter=1, output=0, field=fo)
    B = blocks(shape)
    g = ih.Gpu(shape, bounds=BOX_CONFIG, mass=4.0, gw=True, bins=20)
    (ny, nx), cur = _candidates(ih, g.ctx)
    assert cur[0] == 0 and 0 <= cur[1] < nx  # per-length defaults
    assert ny >= 1 and nx >= 1
    for cy in range(ny):
        for cx in range(nx):
            ih._check(ih.lib().ihg_set_fft_candidates(g.ctx, cy, cx))
            f = g.to_device(y)
            r = g.alloc(g.Ntot)
            g.mk_rhs(f, r, flags=1)
            got = r.download()
            err = {k: rel_l2(got[B[k]], ro[B[k]]) for k in B}
            assert max(err.values()) < 1e-12, (cy, cx, err)
            assert rel_l2(f.download(), fo) < 1e-12, (cy, cx)
    with pytest.raises(RuntimeError):
        ih._check(ih.lib().ihg_set_fft_candidates(g.ctx, ny, 0))
    with pytest.raises(RuntimeError):
        ih._check(ih.lib().ihg_set_fft_candidates(g.ctx, 0, -1))


def test_autotuned_ensemble_rhs_matches_oracle(ih, E, oracle_mod):
    """ihg_autotune on a 32^3 x 4 ensemble picks valid candidates, reports a
    time for each, and the ensemble RHS with its choice still matches the
    oracle for every realization."""
    from tests.states import blocks, rel_l2
    n, nreal = 32, 4
    e = E.Ensemble(n, nreal, bins=16)
    e.init([1, 2, 3, 4], [1e-5, 2e-5, 3e-5, 1e-4])
    counts, _ = _candidates(ih, e.ctx)
    (cy, cx), ms = e.autotune(reps=3)
    assert 0 <= cy < counts[0] and 0 <= cx < counts[1]
    assert _candidates(ih, e.ctx)[1] == [cy, cx]
    nt = counts[0] + counts[1]
    assert np.all(ms[:nt] > 0) and np.all(ms[nt:] == 0)
    states = e.state()
    L = ih.lib()
    stride = e.stride
    nn = nreal * stride
    fd, rd = ctypes.c_void_p(), ctypes.c_void_p()
    ih._check(L.ihg_malloc(e.ctx, ctypes.byref(fd), nn))
    ih._check(L.ihg_malloc(e.ctx, ctypes.byref(rd), nn))
    full = np.zeros((nreal, stride))
    full[:, :e.Ntot] = states
    ih._check(L.ihg_upload(e.ctx, fd, full.ctypes.data_as(ctypes.c_void_p), nn))
    ih._check(L.ihg_mk_rhs_ensemble(e.ctx, 0.0, fd, rd, 1))
    got = np.empty(nn)
    ih._check(L.ihg_download(e.ctx, got.ctypes.data_as(ctypes.c_void_p), rd, nn))
    got = got.reshape(nreal, stride)
    shape = (n, n, n)
    Lb = 2 * np.pi / 0.1
    o = oracle_mod.Oracle(*shape, bounds=((0.0, Lb),) * 3, mass=1.0, gw=True,
                          filt=True, bins=16)
    B = blocks(shape)
    for r in range(nreal):
        fo = states[r].copy()
        ro = o.mk_rhs(fo, filter=1, output=0, field=fo)
        for k in B:
            assert rel_l2(got[r, :e.Ntot][B[k]], ro[B[k]]) < 1e-12, (r, k)
    L.ihg_free(e.ctx, fd)
    L.ihg_free(e.ctx, rd)
    e.close()


@pytest.mark.parametrize("shape,nreal,gw,fp32,flags", [
    ((16, 16, 16), 3, True, False, 1), ((32, 32, 32), 2, False, False, 0),
    ((128, 128, 128), 1, True, False, 0), ((16, 16, 16), 2, True, True, 1),
    ((16, 256, 16), 2, True, False, 0)])
def test_mk_rhs_rk4_equals_rhs_then_stage(ih, shape, nreal, gw, fp32, flags):
    """ihg_mk_rhs_rk4 (the stage or the final RK4 sum written by the kernels
    that write the slope) equals ihg_mk_rhs_ensemble followed by
    ihg_rk4_stage / ihg_rk4_final, bit for bit: 16^3 and 128^3 with GW
    (fused TT pass), 32^3 without GW, and the FP32 GW source and 256-point
    y lines (no fused TT pass: the separate calls run)."""
    L = ih.lib()
    p = ih.make_params(shape, BOX_CONFIG, 4.0, gw, True, 16, fp32)
    h = ctypes.c_void_p()
    ih._check(L.ihg_create_ensemble(ctypes.byref(p), nreal, ctypes.byref(h)))
    ctx = h.value
    ens = (ctypes.c_long * 2)()
    ih._check(L.ihg_ensemble(ctx, ens))
    n = nreal * ens[1]
    rng = np.random.default_rng(5)

    def dev(x=None):
        q = ctypes.c_void_p()
        ih._check(L.ihg_malloc(ctx, ctypes.byref(q), n))
        if x is not None:
            ih._check(L.ihg_upload(ctx, q, x.ctypes.data_as(ctypes.c_void_p), n))
        return q

    def host(q):
        out = np.empty(n)
        ih._check(L.ihg_download(ctx, out.ctypes.data_as(ctypes.c_void_p), q, n))
        return out

    ys = np.zeros((nreal, ens[1]))
    for r in range(nreal):
        y0 = make_state(shape, seed=60 + r, gw=gw, a=1.2 + 0.1 * r)
        ys[r, :y0.size] = y0
    ys = ys.ravel()
    # the input of the stage RHS: y plus a perturbation (as a stage input is)
    fin = ys * (1.0 + 1e-3 * rng.standard_normal(n))
    # slopes on the state's elements as an RHS leaves them: zero h blocks
    # with GW off, zero padding between realizations
    N = int(np.prod(shape))
    ntot = sizes(shape)[3]
    live = np.zeros((nreal, ens[1]), dtype=bool)
    live[:, :4 * N] = True
    if gw:
        live[:, 4 * N:ntot - 1] = True
    live[:, ntot - 1] = True
    live = live.ravel()
    ks = [np.where(live, rng.standard_normal(n) * 1e-2, 0.0) for _ in range(3)]
    dt = 0.01
    for mode, half in ((ih.IHG_RK4_STAGE, 1), (ih.IHG_RK4_STAGE, 0),
                       (ih.IHG_RK4_FINAL, 0)):
        print(f"ihg_mk_rhs_rk4 {shape} x {nreal}, mode {mode}, half {half}", flush=True)
        # fused
        y1, f1, r1, o1 = dev(ys), dev(fin), dev(), dev()
        k1, k2, k3 = dev(ks[0]), dev(ks[1]), dev(ks[2])
        ep = ih.IhgRk4Ep(mode, half, dt, y1.value, o1.value, k1.value, k2.value,
                         k3.value)
        ih._check(L.ihg_mk_rhs_rk4(ctx, 0.0, f1, r1, flags, ctypes.byref(ep)))
        # separate
        y2, f2, r2, o2 = dev(ys), dev(fin), dev(), dev()
        ih._check(L.ihg_mk_rhs_ensemble(ctx, 0.0, f2, r2, flags))
        if mode == ih.IHG_RK4_STAGE:
            ih._check(L.ihg_rk4_stage(ctx, o2, y2, r2, dt, half))
            assert np.array_equal(host(o1)[live], host(o2)[live])
            assert np.array_equal(host(r1)[live], host(r2)[live])
        else:
            ih._check(L.ihg_rk4_final(ctx, y2, k1, k2, k3, r2, dt))
            assert np.array_equal(host(y1)[live], host(y2)[live])
        assert np.array_equal(host(f1), host(f2))  # the filtered input
        for q in (y1, f1, r1, o1, k1, k2, k3, y2, f2, r2, o2):
            L.ihg_free(ctx, q)
    L.ihg_destroy(ctx)
