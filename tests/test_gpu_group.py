"""GPU checks of the grouped form: ensembles of several grid sizes whose
passes run as one launch each (ihg_group_*, ih_ens_run_rk4_grouped).

* one grouped mk_rhs equals ihg_mk_rhs_ensemble on every context, bit for
  bit (the result and the filtered input), with and without the filter and
  with GW on and off in one group;
* grouped RK4 steps, eager and as one graph per step, equal the
  per-ensemble steps bit for bit, after every graph replay with the
  program's own eager work and synchronisation in between;
* each grouped call is one launch per pass, whatever the number of
  contexts;
* the group rejects what its launches do not serve (non-cubic grids, other
  lengths than 16 to 128, the FP32 GW source, more than 4 or repeated
  contexts).
"""
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
def E():
    from instaheat_amd import ensemble
    return ensemble


class Ctx:
    """An ensemble context with nreal seeded states on the device."""

    def __init__(self, ih, shape, nreal, gw=True, filt=True, seed=0):
        self.ih, self.L = ih, ih.lib()
        p = ih.make_params(shape, BOX_CONFIG, 4.0, gw, filt, 16, False)
        h = ctypes.c_void_p()
        ih._check(self.L.ihg_create_ensemble(ctypes.byref(p), nreal, ctypes.byref(h)))
        self.ctx = h.value
        ens = (ctypes.c_long * 2)()
        ih._check(self.L.ihg_ensemble(self.ctx, ens))
        self.nreal, self.stride = nreal, ens[1]
        self.n = nreal * self.stride
        ys = [make_state(shape, seed=seed + r, gw=gw, a=1.1 + 0.1 * r)
              for r in range(nreal)]
        self.host = np.zeros((nreal, self.stride))
        for r, y in enumerate(ys):
            self.host[r, :y.size] = y
        self.f = self.alloc()
        self.r = self.alloc()
        self.upload(self.f, self.host)

    def alloc(self):
        p = ctypes.c_void_p()
        self.ih._check(self.L.ihg_malloc(self.ctx, ctypes.byref(p), self.n))
        return p.value

    def upload(self, p, x):
        x = np.ascontiguousarray(x, dtype=np.float64).ravel()
        self.ih._check(self.L.ihg_upload(self.ctx, p, x.ctypes.data_as(ctypes.c_void_p),
                                         self.n))

    def download(self, p):
        out = np.empty(self.n)
        self.ih._check(self.L.ihg_download(self.ctx, out.ctypes.data_as(ctypes.c_void_p),
                                           p, self.n))
        return out

    def close(self):
        self.L.ihg_destroy(self.ctx)


def _ptrs(xs):
    return (ctypes.c_void_p * len(xs))(*xs)


def _group(ih, ctxs):
    g = ctypes.c_void_p()
    ih._check(ih.lib().ihg_group_create(_ptrs([c.ctx for c in ctxs]), len(ctxs),
                                        ctypes.byref(g)))
    return g.value


@pytest.mark.parametrize("flags", [0, 1])
def test_group_rhs_equals_per_context(ih, flags):
    L = ih.lib()
    # 128 and 64 with GW, 32 without GW, 16 without the filter
    ctxs = [Ctx(ih, (16, 16, 16), 3, gw=True, filt=False, seed=10),
            Ctx(ih, (64, 64, 64), 1, gw=True, seed=20),
            Ctx(ih, (32, 32, 32), 2, gw=False, seed=30),
            Ctx(ih, (128, 128, 128), 1, gw=True, seed=40)]
    g = _group(ih, ctxs)
    print(f"grouped mk_rhs, flags {flags}", flush=True)
    n0 = L.ihg_group_launches(g)
    ih._check(L.ihg_group_mk_rhs(g, 0.0, _ptrs([c.f for c in ctxs]),
                                 _ptrs([c.r for c in ctxs]), flags))
    # filter (5 passes) when any context filters, phi_hat for the others
    # (2), fan-out with and without the kept spectrum, y, z c2r + rho, means,
    # update, and the 3 GW passes
    nl = L.ihg_group_launches(g) - n0
    assert nl == (5 if flags else 0) + 2 + (2 if flags else 1) + 4 + 3, nl
    ih._check(L.ihg_sync(ctxs[0].ctx))
    got = [(c.download(c.r), c.download(c.f)) for c in ctxs]
    for c, (res, filt) in zip(ctxs, got):
        print(f"per-context mk_rhs, {c.nreal} realizations", flush=True)
        f2, r2 = c.alloc(), c.alloc()
        c.upload(f2, c.host)
        ih._check(L.ihg_mk_rhs_ensemble(c.ctx, 0.0, f2, r2, flags))
        assert np.array_equal(res, c.download(r2))
        assert np.array_equal(filt, c.download(f2))
        L.ihg_free(c.ctx, f2)
        L.ihg_free(c.ctx, r2)
    L.ihg_group_destroy(g)
    for c in ctxs:
        c.close()


def _mixed(E, sizes, seed0=1):
    ens, s = [], seed0
    for n, nreal in sizes:
        e = E.Ensemble(n, nreal, bins=16)
        e.init(list(range(s, s + nreal)), list(10.0 ** np.linspace(-6, -4, nreal)))
        s += nreal
        ens.append(e)
    return ens


SIZES = ((16, 4), (32, 3), (64, 2), (128, 1))


def test_grouped_rk4_equals_per_ensemble_steps(ih, E):
    ref = _mixed(E, SIZES)
    grp = _mixed(E, SIZES)
    gg = _mixed(E, SIZES)
    steps, dt = 3, 0.005
    print("per-ensemble RK4 steps", flush=True)
    E.run(ref, steps, dt, use_graph=False)
    print("grouped RK4 steps, eager", flush=True)
    E.run(grp, steps, dt, use_graph=False, grouped=True)
    print("grouped RK4 steps, one graph per step", flush=True)
    E.run(gg, steps, dt, use_graph=True, grouped=True)
    for a, b, c in zip(ref, grp, gg):
        assert np.array_equal(a.state(), b.state()), a.n
        assert np.array_equal(a.state(), c.state()), a.n
        assert b.t == pytest.approx(steps * dt) and c.t == b.t
        assert b.calls_rhs == a.calls_rhs == c.calls_rhs
    for e in ref + grp + gg:
        e.close()


def test_grouped_graph_replays_with_eager_work_between(ih, E):
    """One captured grouped step replayed 4 times; between the replays the
    program runs its own eager work (an observables call) and synchronises;
    after every replay the states equal the eager grouped steps."""
    H, L = ih.host_lib(), ih.lib()
    sizes = ((16, 3), (32, 2), (64, 1))
    ea = _mixed(E, sizes, seed0=5)
    ga = _mixed(E, sizes, seed0=5)
    dt = 0.005
    graph = ctypes.c_void_p()
    arr = _ptrs([e.h for e in ga])
    print("capture one grouped RK4 step", flush=True)
    ih._check(H.ih_ens_capture_step_grouped(arr, len(ga), dt, ctypes.byref(graph)))
    for step in range(4):
        print(f"grouped graph replay {step}", flush=True)
        E.run(ea, 1, dt, use_graph=False, grouped=True)
        ih._check(L.ihg_graph_launch(graph, ga[0].ctx))
        ih._check(L.ihg_sync(ga[0].ctx))
        _ = ga[1].observables(0)  # eager work of the program between replays
        for a, b in zip(ea, ga):
            assert np.array_equal(a.state(), b.state()), (step, a.n)
    L.ihg_graph_destroy(graph)
    for e in ea + ga:
        e.close()


def test_group_rejects_what_it_does_not_serve(ih):
    L = ih.lib()
    ok = Ctx(ih, (16, 16, 16), 2)
    bad = [Ctx(ih, (16, 8, 32), 1), Ctx(ih, (8, 8, 8), 1)]
    g = ctypes.c_void_p()
    for c in bad:
        rc = L.ihg_group_create(_ptrs([ok.ctx, c.ctx]), 2, ctypes.byref(g))
        assert rc != 0 and g.value is None
        assert b"grouped launches serve cubic grids" in L.ihg_last_error()
    rc = L.ihg_group_create(_ptrs([ok.ctx, ok.ctx]), 2, ctypes.byref(g))
    assert rc != 0 and b"repeated" in L.ihg_last_error()
    rc = L.ihg_group_create(_ptrs([ok.ctx] * 5), 5, ctypes.byref(g))
    assert rc != 0
    ih._check(L.ihg_set_gw_fp32(ok.ctx, 1))
    rc = L.ihg_group_create(_ptrs([ok.ctx]), 1, ctypes.byref(g))
    assert rc != 0 and b"FP32" in L.ihg_last_error()
    for c in [ok] + bad:
        c.close()
