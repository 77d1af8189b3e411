"""CPU checks of DOPRI853 dense output: the oracle's restatement
(oracle/dense.py) against scipy's DOP853 dense output, the library's
continuous-extension table against scipy's independent copy, and the
interpolant's end points and order.  No GPU."""
import ctypes

import numpy as np
import pytest
from scipy.integrate import solve_ivp
from scipy.integrate._ivp import dop853_coefficients as SC
from scipy.integrate._ivp.rk import DOP853

from oracle import dense


def _fun(t, y):
    # a nonlinear oscillator pair with time dependence
    return np.array([y[1], -np.sin(y[0]) - 0.1 * y[1] + 0.3 * np.cos(t),
                     y[3], -4.0 * y[2] * (1 + 0.2 * y[2] ** 2)])


def test_oracle_interpolant_equals_scipy():
    s = DOP853(_fun, 0.0, np.array([1.0, 0.0, 0.5, -0.2]), 20.0, rtol=1e-8,
               atol=1e-10)
    nsteps = 0
    while s.status == "running" and nsteps < 40:
        s.step()
        nsteps += 1
        sol = s.dense_output()
        K = s.K_extended
        h = s.t - s.t_old
        # the extra stages: inputs from the oracle, derivatives through fun
        for i in range(3):
            st = 13 + i
            yi = dense.extra_stage_input(s.y_old, K, h, SC.A[st][:st])
            assert np.allclose(_fun(s.t_old + SC.C[st] * h, yi), K[st],
                               rtol=1e-14, atol=1e-14)
        for th in (0.0, 0.13, 0.5, 0.91, 1.0):
            want = sol(s.t_old + th * h)
            got = dense.interpolate(s.y_old, s.y, K, h, th, SC.D)
            assert np.allclose(got, want, rtol=1e-13, atol=1e-14), (nsteps, th)
    assert nsteps >= 10


def test_oracle_step_equals_scipy_step():
    y0 = np.array([1.0, 0.0, 0.5, -0.2])
    s = DOP853(_fun, 0.0, y0, 20.0, rtol=1e-8, atol=1e-10, first_step=0.05)
    s.step()
    h = s.t - s.t_old
    K, y_new = dense.step(_fun, 0.0, y0, _fun(0.0, y0), h, SC.A, SC.B, SC.C)
    assert np.allclose(y_new, s.y, rtol=1e-15, atol=1e-15)
    assert np.allclose(np.array(K), s.K[:12], rtol=1e-14, atol=1e-14)


def test_interpolant_order():
    """the dense output's error at mid-step shrinks like h^8 (order 7
    interpolant of an 8th-order method), against a tight solve_ivp"""
    y0 = np.array([1.0, 0.0, 0.5, -0.2])
    ref = solve_ivp(_fun, (0, 1.0), y0, method="DOP853", rtol=1e-13,
                    atol=1e-15, dense_output=True)
    errs = []
    for h in (0.2, 0.1):
        t, y, k1 = 0.0, y0.copy(), _fun(0.0, y0)
        e = 0.0
        while t < 1.0 - 1e-12:
            K, y_new = dense.step(_fun, t, y, k1, h, SC.A, SC.B, SC.C)
            K.append(_fun(t + h, y_new))
            for i in range(3):
                st = 13 + i
                yi = dense.extra_stage_input(y, K, h, SC.A[st][:st])
                K.append(_fun(t + SC.C[st] * h, yi))
            mid = dense.interpolate(y, y_new, K, h, 0.5, SC.D)
            e = max(e, np.max(np.abs(mid - ref.sol(t + 0.5 * h))))
            t, y, k1 = t + h, y_new, K[12]
        errs.append(e)
    rate = np.log2(errs[0] / errs[1])
    print("dense-output errors", errs, "rate", rate)
    assert rate > 6.5


def test_library_dense_table_equals_scipy():
    import instaheat_amd as ih
    L = ih.lib()
    c = np.zeros(3)
    a = np.zeros(48)
    d = np.zeros(64)
    ih._check(L.ihg_dp853_dense_tableau(c.ctypes.data_as(ih._dptr),
                                        a.ctypes.data_as(ih._dptr),
                                        d.ctypes.data_as(ih._dptr)))
    assert np.allclose(c, SC.C[13:16], rtol=1e-15, atol=0)
    assert np.allclose(a.reshape(3, 16), SC.A[13:16], rtol=1e-14, atol=1e-17)
    assert np.allclose(d.reshape(4, 16), SC.D, rtol=1e-14, atol=1e-17)
