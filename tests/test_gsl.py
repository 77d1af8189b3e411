"""CPU checks of the GSL stepper path (src/gsl_stepper.c): the Butcher
tables of RKF45, RKCK and RK8PD (Prince-Dormand 8(7)) against their order
conditions over every rooted tree, and against the convergence rates they
must show on problems with closed-form solutions; the device library's
copy of the tables against the oracle's; the standard control's step
adjustment (C host and oracle) against the rules GSL documents; the
oracle's driver on a closed-form problem (intervals end exactly, the error
stays near the tolerance, a step below hmin stops it with GSL_ENOPROG)."""
import ctypes
import math
from fractions import Fraction as F

import numpy as np
import pytest

from oracle import gsl

METHODS = [gsl.RKF45, gsl.RKCK, gsl.DOPRI89]


# --- rooted trees and their order conditions -------------------------------

_TREES = {1: [()]}


def trees(n):
    """rooted trees with n nodes, each a sorted tuple of its subtrees"""
    if n in _TREES:
        return _TREES[n]
    out = set()

    def parts(rem, maxkey, acc):
        if rem == 0:
            out.add(tuple(sorted(acc)))
            return
        for m in range(1, rem + 1):
            for t in trees(m):
                key = (m, t)
                if maxkey is None or key <= maxkey:
                    parts(rem - m, key, acc + [t])
    parts(n - 1, None, [])
    _TREES[n] = sorted(out)
    return _TREES[n]


def _order(t):
    return 1 + sum(_order(c) for c in t)


def _gamma(t):
    g = _order(t)
    for c in t:
        g *= _gamma(c)
    return g


def _phi(t, A):
    s = len(A)
    v = [F(1)] * s
    for c in t:
        pc = _phi(c, A)
        w = [sum(A[i][j] * pc[j] for j in range(i)) for i in range(s)]
        v = [v[i] * w[i] for i in range(s)]
    return v


def test_tree_counts():
    # OEIS A000081: 1, 1, 2, 4, 9, 20, 48, 115, 286
    assert [len(trees(n)) for n in range(1, 10)] == [1, 1, 2, 4, 9, 20, 48, 115, 286]


def _residuals(A, w, n):
    s = len(A)
    return [abs(float(sum(w[i] * ph[i] for i in range(s)) - F(1, _gamma(t))))
            for t in trees(n) for ph in [_phi(t, A)]]


@pytest.mark.parametrize("method", METHODS)
def test_order_conditions(method):
    """b satisfies every condition up to the method's order q and not all
    of q + 1; the embedded weights up to q - 1 and not all of q.  The
    RK8PD tables are rational approximations (about 1e-17)."""
    T = gsl.TABLEAUS[method]
    A, c, b, s, q = T["a"], T["c"], T["b"], T["stages"], T["order"]
    emb = gsl.EMBEDDED[method]
    bh = [F(emb.get(j, 0)) for j in range(1, s + 1)]
    tol = 0 if method != gsl.DOPRI89 else 1e-16
    for i in range(s):
        assert abs(float(sum(A[i]) - c[i])) <= tol, ("row sum", i)
        assert all(A[i][j] == 0 for j in range(i, s)), "explicit"
    for n in range(1, q + 1):
        assert max(_residuals(A, b, n)) <= tol, ("b", n)
    assert max(_residuals(A, b, q + 1)) > 1e-7
    for n in range(1, q):
        assert max(_residuals(A, bh, n)) <= tol, ("bhat", n)
    assert max(_residuals(A, bh, q)) > 1e-7
    # the error weights are the difference of the two solutions
    if T["err_mode"] == 0:
        assert T["w"] == [b[j] - bh[j] for j in range(s)]
    else:
        assert T["w"] == bh


def _lib():
    import instaheat_amd as ih
    return ih.lib(), ih.host_lib()


@pytest.mark.parametrize("method", METHODS)
def test_device_tableau_equals_oracle(method):
    L, _ = _lib()
    info = (ctypes.c_int * 3)()
    c, a, b, w = np.zeros(13), np.zeros(13 * 12), np.zeros(13), np.zeros(13)
    p = ctypes.POINTER(ctypes.c_double)
    assert L.ihg_gsl_tableau(method, info, c.ctypes.data_as(p), a.ctypes.data_as(p),
                             b.ctypes.data_as(p), w.ctypes.data_as(p)) == 0
    t = gsl.tableau(method)
    s = t["stages"]
    assert list(info) == [s, t["order"], t["err_mode"]]
    assert np.array_equal(c[:s], t["c"])
    assert np.array_equal(a.reshape(13, 12)[:s, :s - 1], t["a"][:, :s - 1])
    assert np.array_equal(b[:s], t["b"])
    assert np.array_equal(w[:s], t["w"])


def test_device_tableau_rejects_unknown_methods():
    L, _ = _lib()
    info = (ctypes.c_int * 3)()
    z = np.zeros(13 * 12)
    p = z.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    for m in (0, 4, -1):
        assert L.ihg_gsl_tableau(m, info, p, p, p, p) != 0


# --- convergence on closed-form problems ------------------------------------

def _f(t, y):
    # y0' = y0 cos t (y0 = exp(sin t)); (y1, y2)' = (y2, -y1) (cos, -sin)
    return np.array([y[0] * math.cos(t), y[2], -y[1]])


def _exact(t):
    return np.array([math.exp(math.sin(t)), math.cos(t), -math.sin(t)])


@pytest.mark.parametrize("method", METHODS)
def test_convergence_rates(method):
    """fixed steps to t = 1: the global error falls as h^q; the embedded
    estimate of one step tracks the lower-order solution's local error,
    which falls as h^q"""
    tab = gsl.tableau(method)
    q = tab["order"]

    def run(n):
        h = 1.0 / n
        y, t = _exact(0.0), 0.0
        for _ in range(n):
            y, _, _ = gsl.step(_f, t, y, _f(t, y), h, tab)
            t += h
        return np.abs(y - _exact(1.0)).max()
    ns = (4, 8) if q == 8 else (8, 16)
    e1, e2 = run(ns[0]), run(ns[1])
    rate = math.log2(e1 / e2)
    print(method, "global errors", e1, e2, "rate", rate)
    assert q - 0.6 < rate < q + 1.2
    # one step: the error estimate falls as h^q (the lower order's local
    # error), and the step's own error is a small fraction of it
    est, own = [], []
    for h in ((0.2, 0.1) if q == 8 else (0.1, 0.05)):
        y0 = _exact(0.3)
        y_new, yerr, _ = gsl.step(_f, 0.3, y0, _f(0.3, y0), h, tab)
        est.append(np.abs(yerr).max())
        own.append(np.abs(y_new - _exact(0.3 + h)).max())
    rate = math.log2(est[0] / est[1])
    print(method, "estimate", est, "rate", rate, "own", own)
    assert q - 0.7 < rate < q + 0.7
    assert own[1] < 0.2 * est[1]


# --- the standard control ---------------------------------------------------

def test_hadjust_rules_and_host_match_oracle():
    _, H = _lib()
    hh = ctypes.c_double()
    cases = [
        # r, q, verdict, factor (GSL's documented rules, S = 0.9)
        (2.0, 5, -1, 0.9 * 2.0 ** (-1 / 5)),
        (1e12, 5, -1, 0.2),                      # at most 5 times smaller
        (1.1000001, 8, -1, 0.9 * 1.1000001 ** (-1 / 8)),
        (1.1, 5, 0, 1.0),                        # 0.5 <= r <= 1.1: keep
        (0.5, 8, 0, 1.0),
        (0.3, 5, 1, 0.9 * 0.3 ** (-1 / 6)),
        (0.49, 8, 1, 1.0),                       # S r^(-1/9) < 1: no decrease
        (1e-30, 8, 1, 5.0),                      # at most 5 times larger
        (np.finfo(float).tiny, 5, 1, 5.0),
    ]
    for r, q, verdict, factor in cases:
        v, h = gsl.hadjust(r, 0.125, q)
        assert v == verdict, (r, q)
        assert h == pytest.approx(0.125 * factor, rel=1e-15), (r, q)
        hh.value = 0.125
        assert H.ih_gsl_hadjust(r, ctypes.byref(hh), q) == verdict
        assert hh.value == h
    rng = np.random.default_rng(3)
    for r in 10.0 ** rng.uniform(-20, 6, 400):
        for q in (5, 8):
            v, h = gsl.hadjust(float(r), 0.37, q)
            hh.value = 0.37
            assert H.ih_gsl_hadjust(float(r), ctypes.byref(hh), q) == v
            assert hh.value == h


def test_rmax_skips_nan_and_starts_at_dbl_min():
    y = np.array([1.0, 0.0, 2.0])
    e = np.array([1e-12, 0.0, np.nan])
    # the zero entry with eps_abs = 0 is 0/0 = NaN: skipped
    assert gsl.rmax(y, e, 0.0, 1e-10) == pytest.approx(1e-2)
    assert gsl.rmax(np.zeros(3), np.zeros(3), 0.0, 1e-10) == np.finfo(float).tiny


# --- the driver -------------------------------------------------------------

@pytest.mark.parametrize("method", METHODS)
def test_driver_on_closed_form(method):
    d = gsl.Driver(_f, method, 1e-3, 1e-12, 1e-10, hmin=1e-14)
    y, t = _exact(0.0), 0.0
    for i in range(1, 11):
        t1 = i * 0.7
        status, t, y = d.apply(t, t1, y)
        assert status == gsl.GSL_SUCCESS
        assert t == t1  # the last step of an interval lands on it exactly
        err = np.abs(y - _exact(t1)).max()
        assert err < 1e-7, (t1, err)
    accepted = d.count - d.failed
    cut = [tr for tr in d.trace if tr[2] != gsl.HADJ_DEC]
    assert accepted == len(cut)
    print(method, "attempts", d.count, "failed", d.failed, "rhs", d.rhs)
    # the first step from 1e-3 grows by at most 5x per step
    hs = [h for _, h, v in d.trace[:4]]
    assert all(b <= 5.0 * a * (1 + 1e-15) for a, b in zip(hs, hs[1:]))


def test_driver_enoprog_below_hmin():
    # a stiff decay with a loose hmin: the control asks for steps below it
    def f(t, y):
        return -1e6 * (y - np.cos(t))
    d = gsl.Driver(f, gsl.RKF45, 1e-2, 1e-14, 1e-12, hmin=1e-5)
    status, t, y = d.apply(0.0, 1.0, np.array([2.0]))
    assert status == gsl.GSL_ENOPROG
    assert 0.0 < t < 1.0
    assert abs(d.h) < 1e-5


def _conditioning(case):
    """decisions and final suggested step of one free-running case of the
    GPU test on the oracle, with its reductions at 1 and 3 threads and its
    RHS perturbed by +-2e-15 and +-1e-13 relative: [(perturbation,
    (status, attempts, rejections), h / h_unperturbed - 1)]"""
    from oracle import oracle as orc
    from tests.states import BOX_CONFIG, make_state
    method, shape, gw, tf, nout, h0, rtol, atol = case
    out = []
    for threads, noise in ((1, 0.0), (3, 0.0), (1, 2e-15), (1, -2e-15),
                           (1, 1e-13), (1, -1e-13)):
        o = orc.Oracle(*shape, bounds=BOX_CONFIG, mass=4.0, gw=gw, bins=16,
                       threads=threads)
        rng = np.random.default_rng(1)
        plain = o.mk_rhs

        def mk_rhs(f, _plain=plain, _noise=noise, _rng=rng, **kw):
            r = _plain(f, **kw)
            if _noise and not kw.get("output", 0):
                r = r * (1.0 + _noise * _rng.standard_normal(r.size))
            return r
        o.mk_rhs = mk_rhs
        y0 = make_state(shape, seed=12, gw=gw, a=1.1)
        status, t, y, d, _ = gsl.run(o, y0, method, 0.0, tf, h0, atol, rtol, 1e-9,
                                     nout)
        out.append(((threads, noise), (status, d.count, d.failed), d.h))
    h_ref = out[0][2]
    return [(p, dec, h / h_ref - 1.0) for p, dec, h in out]


def test_driver_einval_for_an_interval_behind_t():
    """gsl_odeiv2_driver_apply refuses t1 < t; run_gsl_stepper's output
    times i (tf - ti) / n hit that whenever ti > 0 (src/gsl_stepper.c:60-64)"""
    d = gsl.Driver(_f, gsl.RKCK, 1e-2, 1e-12, 1e-10)
    y0 = _exact(0.5)
    status, t, y = d.apply(0.5, 0.25, y0)
    assert status == gsl.GSL_EINVAL and t == 0.5 and y is y0 and d.count == 0
    status, t, y = d.apply(0.5, 0.5, y0)
    assert status == gsl.GSL_SUCCESS and d.count == 0


def test_gpu_cases_are_well_conditioned():
    """The free-running GPU parity cases make the same decisions whatever
    the rounding: with the oracle's reductions at 1 or 3 threads and its
    RHS perturbed by 2e-15 (about the device's FFT rounding differences)
    or 1e-13 relative; the final suggested step moves by less than TOL_H / 5
    under rounding-sized changes.  So the GPU test compares decisions, not
    rounding noise.  (The oracle's thread count is process-wide: this test
    leaves it at 1.)"""
    from tests.test_gpu_gsl import CASES, TOL_H
    for case in CASES:
        res = _conditioning(case)
        print(case[:2], res)
        for p, dec, dh in res:
            assert dec == res[0][1], (case[:2], p)
            if abs(p[1]) <= 2e-15:
                assert abs(dh) < 0.2 * TOL_H, (case[:2], p)
