"""CPU checks of the learned step proposer's host side: the C feature
extraction equals the Python one used to build training data, the training
targets follow their closed form, the oracle's MLP gradient is pinned by
finite differences, and the committed model is well formed."""
import ctypes
import json
import os

import numpy as np
import pytest

import instaheat_amd as ih
from instaheat_amd import proposer as P

pytestmark = pytest.mark.filterwarnings("ignore::RuntimeWarning")


def _records(rng, n):
    recs = (ih.IhDpRecord * n)()
    arr = np.zeros((n, len(ih.RECORD_FIELDS)))
    for i in range(n):
        v = dict(t=rng.random(), dt=10 ** rng.uniform(-4, 0),
                 err=10 ** rng.uniform(-14, 0.5), rho_mean=10 ** rng.uniform(-4, 0),
                 a=1 + 10 * rng.random(), var_phi=10 ** rng.uniform(-12, -3),
                 mean_psi2=10 ** rng.uniform(-14, -6),
                 accepted=int(rng.random() > 0.2))
        for k, x in v.items():
            setattr(recs[i], k, x)
        arr[i] = [v[f] for f in ih.RECORD_FIELDS]
    return recs, arr


@pytest.mark.parametrize("n", [1, 3, 8, 13])
def test_c_features_equal_python_features(n):
    rng = np.random.default_rng(n)
    recs, arr = _records(rng, n)
    xc = np.zeros(ih.MLP_IN)
    ih.host_lib().ih_proposer_features(recs, n, xc.ctypes.data_as(ih._dptr))
    xp = P.raw_features(arr, n - 1)
    assert np.allclose(xc, xp, rtol=0, atol=1e-14)
    # history padding: with one record every history slot is that record
    if n == 1:
        assert np.all(xc[:8] == xc[0]) and np.all(xc[8:16] == xc[8])


def test_targets_closed_form():
    """constant error rate err = C dt^8: dt* = dt (e_t / (C dt^8))^(1/8)"""
    C = 3.0e4
    dts = np.array([0.1, 0.11, 0.12, 0.125, 0.13, 0.12])
    log = np.zeros((len(dts), len(ih.RECORD_FIELDS)))
    F = P.FIELDS
    log[:, F["dt"]] = dts
    log[:, F["err"]] = C * dts ** 8
    log[:, F["accepted"]] = 1
    log[:, F["a"]] = 1.0
    log[:, F["rho_mean"]] = 0.5
    log[:, F["var_phi"]] = 1e-6
    log[:, F["mean_psi2"]] = 1e-8
    X, T = P.samples(log, err_target=0.5)
    dt_star = (0.5 / C) ** 0.125
    assert len(T) == len(dts) - 1
    assert np.allclose(T, np.log(dt_star / dts[:-1]), rtol=1e-12)
    # a rejected attempt drops its own sample and the one after it
    log[2, F["accepted"]] = 0
    X2, T2 = P.samples(log, err_target=0.5)
    assert len(T2) == len(dts) - 3


def test_oracle_mlp_gradient_matches_finite_differences():
    from oracle import mlp as M
    rng = np.random.default_rng(0)
    p = rng.normal(0, 0.3, M.NPARAM)
    x = rng.normal(size=(37, M.NI))
    t = rng.normal(size=37)
    _, g = M.loss_grad(p, x, t)
    idx = np.concatenate([rng.choice(M.NPARAM, 300, replace=False),
                          np.arange(M.OFF_W3, M.NPARAM)])
    h = 1e-6
    for q in idx:
        e = np.zeros(M.NPARAM)
        e[q] = h
        fd = (M.loss(p + e, x, t) - M.loss(p - e, x, t)) / (2 * h)
        assert abs(fd - g[q]) <= 1e-7 * max(1.0, abs(g[q])), q


def test_oracle_mlp_forward_closed_form():
    """zero weights in the first two layers: y = W3 . tanh(b2) + b3"""
    from oracle import mlp as M
    rng = np.random.default_rng(1)
    p = np.zeros(M.NPARAM)
    p[M.OFF_B2:M.OFF_W3] = rng.normal(size=M.NH)
    p[M.OFF_W3:M.OFF_B3] = rng.normal(size=M.NH)
    p[M.OFF_B3] = 0.25
    y = M.forward(p, rng.normal(size=(5, M.NI)))
    ref = np.tanh(p[M.OFF_B2:M.OFF_W3]) @ p[M.OFF_W3:M.OFF_B3] + 0.25
    assert np.allclose(y, ref, rtol=1e-14)


def test_committed_model_is_well_formed():
    with open(P.DEFAULT_MODEL) as fh:
        m = json.load(fh)
    assert len(m["params"]) == ih.lib().ihg_mlp_nparams() == 5569
    assert len(m["in_mean"]) == len(m["in_std"]) == ih.MLP_IN
    assert all(s > 0 for s in m["in_std"]) and m["out_std"] > 0
    assert np.all(np.isfinite(m["params"]))
    # the proposer may not rescale more than success() does
    assert m["min_ratio"] >= 0.333 - 1e-12 and m["max_ratio"] <= 3.0 + 1e-12
