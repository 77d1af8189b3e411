"""Pins of the oracle's metric terms against what the mathematics fixes.

* rho, p, phi_tt, psi_tt (update_phi_psi and mk_rho_and_p, src/toolbox.c:
  225-243, 282-299) against the stress tensor, the Klein-Gordon equation and
  the trace of the Einstein equations of the metric of README.md:91, derived
  by sympy in tests/einstein.py, on a state with large psi terms;
* update_h (src/toolbox.c:311-346) against the linearised Einstein tensor of
  a transverse-traceless wave, G_ij = TT part of d_i phi d_j phi, with the
  source projected by the textbook projector, on every mode and branch;
* free tensor waves (phi = dphi = 0, so rho = p = H = 0 and a is constant):
  every mode of h follows h0 cos(w t) + (dh0 / w) sin(w t), w = |k| / a,
  with the full Nyquist wave number, under the oracle's RK4;
* constraint propagation, as in the reference's own evaluation
  (evaluation/evalConstraints.m: errors that scale with the square of the
  perturbation): from mk_initial_psi data the Hamiltonian and momentum
  constraints stay of second order in the fluctuation amplitude;
* every one of these pins fails on an oracle with a planted mistake (a
  changed coefficient or sign in the terms above), compiled from a copy of
  oracle/oracle.c on the CPU.
"""
import json
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

from tests import einstein as E
from tests.states import BOX_CONFIG, kgrids, make_state

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MASS = 1.3
# the grids here are tiny: one oracle thread each (many OpenMP threads on
# loops of a few hundred points only wait on each other)


def _derivs(f, shape, bounds):
    """first derivatives (Nyquist entry zeroed) and second derivatives (full
    wave number) of a real field, spectrally"""
    F = np.fft.rfftn(f.reshape(shape))
    KX, KY, KZ = kgrids(shape, bounds)
    out = {}
    for name, K, n, ax in (("x", KX, shape[0], 0), ("y", KY, shape[1], 1),
                           ("z", KZ, shape[2], 2)):
        Kd = K.copy()
        idx = [slice(None)] * 3
        idx[ax] = n // 2
        Kd[tuple(idx)] = 0.0
        out[name] = np.fft.irfftn(1j * Kd * F, s=shape, axes=(0, 1, 2)).ravel()
        out[name + name] = np.fft.irfftn(-K ** 2 * F, s=shape,
                                         axes=(0, 1, 2)).ravel()
    return out


def _scalar_state(shape, seed):
    # large phi fluctuations and psi ~ 1e-2, so that every psi term is far
    # above rounding
    return make_state(shape, BOX_CONFIG, seed=seed, gw=True, a=1.3,
                      amp=5e-2, psi_amp=1e-2, h_amp=1e-2)


def _evaluate(fn_names, env):
    f, names = fn_names
    return f(*[env[n] for n in names])


def scalar_residuals(y, out, shape):
    """max relative deviations of the oracle's rho, p, a_t, phi_tt, psi_tt
    (out: rhs, rho, p) from the metric's equations"""
    N = int(np.prod(shape))
    rhs, rho_o, p_o = out["rhs"], out["rho"], out["p"]
    a = y[-1]
    phi, dphi, psi, dpsi = (y[i * N:(i + 1) * N] for i in range(4))
    dp = _derivs(phi, shape, BOX_CONFIG)
    ds = _derivs(psi, shape, BOX_CONFIG)
    eq = E.lambdified(E.scalar_equations())
    env = {"m": MASS, "a": a, "phi": phi, "phi_t": dphi, "psi": psi,
           "psi_t": dpsi}
    for k in ("x", "y", "z", "xx", "yy", "zz"):
        env["phi_" + k] = dp[k]
        env["psi_" + k] = ds[k]
    rho = _evaluate(eq["rho"], env)
    p = _evaluate(eq["p"], env)
    H = np.sqrt(rho.mean() / 3.0)
    env["a_t"] = a * H
    env["p_mean"] = p.mean()
    ddphi = _evaluate(eq["phi_tt"], env)
    ddpsi = _evaluate(eq["psi_tt"], env)
    lap = dp["xx"] + dp["yy"] + dp["zz"]
    sc_phi = np.abs(lap).max() / a ** 2 + np.abs(MASS ** 2 * phi).max() + \
        np.abs(3 * H * dphi).max()
    return {
        "rho": np.abs(rho_o - rho).max() / np.abs(rho).max(),
        "p": np.abs(p_o - p).max() / np.abs(p).max(),
        "a_t": abs(rhs[-1] - a * H) / (a * H),
        "dphi": np.abs(rhs[:N] - dphi).max() / np.abs(dphi).max(),
        "dpsi": np.abs(rhs[2 * N:3 * N] - dpsi).max() / np.abs(dpsi).max(),
        "phi_tt": np.abs(rhs[N:2 * N] - ddphi).max() / sc_phi,
        "psi_tt": np.abs(rhs[3 * N:4 * N] - ddpsi).max() / np.abs(p).max(),
    }


def _full_k(shape):
    """full wave numbers with the Nyquist entry of x and y taken as +N/2
    (kvec.xf / yf, src/setup.c:609-666; the TT projector is odd in a single
    component, so the sign at the Nyquist planes is a convention)"""
    KX, KY, KZ = kgrids(shape, BOX_CONFIG)
    KX[shape[0] // 2] = np.abs(KX[shape[0] // 2])
    KY[:, shape[1] // 2] = np.abs(KY[:, shape[1] // 2])
    return KX, KY, KZ


def _tt_source(phi, shape):
    """the two components the reference keeps of the TT part of the
    transform of d_i phi d_j phi, by the textbook projector
    Lambda_ij,lm = P_il P_jm - P_ij P_lm / 2, P = 1 - k k / k^2 (full k)"""
    d = _derivs(phi, shape, BOX_CONFIG)
    g = [d["x"], d["y"], d["z"]]
    S = np.zeros((3, 3) + (shape[0], shape[1], shape[2] // 2 + 1), complex)
    for i in range(3):
        for j in range(i, 3):
            S[i, j] = S[j, i] = np.fft.rfftn((g[i] * g[j]).reshape(shape))
    KX, KY, KZ = _full_k(shape)
    K = np.stack([KX, KY, KZ])
    k2 = (K ** 2).sum(0)
    k2s = np.where(k2 > 0, k2, 1.0)
    P = np.eye(3)[:, :, None, None, None] - K[:, None] * K[None, :] / k2s
    PSP = np.einsum("il...,lm...,mj...->ij...", P, S, P)
    tr = np.einsum("lm...,lm...->...", P, S)
    TT = PSP - 0.5 * P * tr
    TT[..., 0, 0, 0] = 0.0
    eps = np.finfo(float).eps
    c0 = np.where(np.abs(KZ) > eps, TT[0, 0],
                  np.where(np.abs(KY) > eps, TT[0, 0], TT[1, 1]))
    c1 = np.where(np.abs(KZ) > eps, TT[0, 1],
                  np.where(np.abs(KY) > eps, TT[0, 2], TT[1, 2]))
    return c0.ravel(), c1.ravel(), k2.ravel()


def tensor_residual(y, out, shape):
    """max deviation of the oracle's dh (the h slots of the RHS) from the
    linearised Einstein equation G_ij = TT(d_i phi d_j phi), relative to its
    largest term"""
    N = int(np.prod(shape))
    M = shape[0] * shape[1] * (shape[2] // 2 + 1)
    rhs = out["rhs"]
    a = y[-1]
    rho, p = out["rho"], out["p"]
    H = np.sqrt(rho.mean() / 3.0)
    c0, c1, k2 = _tt_source(y[:N], shape)
    eq = E.lambdified(E.tensor_equations())
    worst = 0.0
    for slot, (name, src) in enumerate((("G11", c0), ("G12", c1))):
        h = y[4 * N + slot * 2 * M:4 * N + (slot + 1) * 2 * M].view(complex)
        dh = y[4 * N + (2 + slot) * 2 * M:4 * N + (3 + slot) * 2 * M].view(complex)
        # d h / dt is the dh slot
        got_dh = rhs[4 * N + slot * 2 * M:4 * N + (slot + 1) * 2 * M].view(complex)
        worst = max(worst, np.abs(got_dh - dh).max() / np.abs(dh).max())
        ddh = rhs[4 * N + (2 + slot) * 2 * M:4 * N + (3 + slot) * 2 * M].view(complex)
        sym = "hp" if name == "G11" else "hx"
        env = {"a": a, "a_t": a * H, "p_mean": p.mean(), sym: h,
               sym + "_t": dh, sym + "_tt": ddh, sym + "_zz": -k2 * h}
        G = _evaluate(eq[name], env)
        scale = max(np.abs(a ** 2 * ddh / 2).max(), np.abs(src).max(),
                    np.abs(k2 * h / 2).max())
        worst = max(worst, np.abs(G - src).max() / scale)
    return worst


def oracle_outputs(orc, y, shape, gw=True):
    o = orc.Oracle(*shape, bounds=BOX_CONFIG, mass=MASS, gw=gw, filt=False,
                   bins=8, threads=1)
    f = y.copy()
    rhs = o.mk_rhs(f, filter=0, output=0, field=f)
    return {"rhs": rhs, "rho": o.aux(5), "p": o.aux(6)}


SHAPES = [(16, 16, 16), (8, 16, 32)]


@pytest.mark.parametrize("shape", SHAPES)
def test_scalar_terms_follow_the_metric(oracle_mod, shape):
    y = _scalar_state(shape, 3)
    r = scalar_residuals(y, oracle_outputs(oracle_mod, y, shape), shape)
    print(shape, r)
    for k, v in r.items():
        assert v < 1e-12, (k, v)


@pytest.mark.parametrize("shape", SHAPES)
def test_tensor_equation_is_the_linearised_einstein_tensor(oracle_mod, shape):
    y = _scalar_state(shape, 4)
    r = tensor_residual(y, oracle_outputs(oracle_mod, y, shape), shape)
    print(shape, "tensor residual", r)
    assert r < 1e-12


def free_wave_error(orc, shape, steps=800):
    """max relative error of every h mode against the closed form after
    `steps` RK4 steps of 1/200 of the fastest mode's period"""
    N = int(np.prod(shape))
    M = shape[0] * shape[1] * (shape[2] // 2 + 1)
    rng = np.random.default_rng(8)
    a = 1.7
    y = np.zeros(4 * N + 8 * M + 1)
    blocks = [(rng.standard_normal(M) + 1j * rng.standard_normal(M))
              for _ in range(4)]
    for b in range(4):
        y[4 * N + b * 2 * M:4 * N + (b + 1) * 2 * M] = blocks[b].view(float)
    y[-1] = a
    KX, KY, KZ = kgrids(shape, BOX_CONFIG)
    w = np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2).ravel() / a
    dt = 0.005 * (2 * np.pi / w.max())
    T = steps * dt
    o = orc.Oracle(*shape, bounds=BOX_CONFIG, mass=MASS, gw=True, filt=False,
                   bins=8, threads=1)
    o.run_rk4(y, 0.0, T * 1.5, dt, max_steps=steps, skip=10 ** 9)
    err = 0.0
    for slot in range(2):
        h0, dh0 = blocks[slot], blocks[2 + slot]
        ws = np.where(w > 0, w, 1.0)
        h = np.where(w > 0, h0 * np.cos(w * T) + dh0 / ws * np.sin(w * T),
                     h0 + dh0 * T)
        dh = np.where(w > 0, -h0 * w * np.sin(w * T) + dh0 * np.cos(w * T), dh0)
        got_h = y[4 * N + slot * 2 * M:4 * N + (slot + 1) * 2 * M].view(complex)
        got_dh = y[4 * N + (2 + slot) * 2 * M:4 * N + (3 + slot) * 2 * M].view(complex)
        err = max(err, np.abs(got_h - h).max() / np.abs(h).max(),
                  np.abs(got_dh - dh).max() / np.abs(dh).max())
    assert y[-1] == a
    return err


@pytest.mark.parametrize("shape", [(8, 8, 8), (4, 8, 16)])
def test_free_tensor_waves_closed_form(oracle_mod, shape):
    """Every mode incl. the Nyquist planes; RK4's error at w dt = 2 pi / 200
    over 4 periods of the fastest mode is about 1e-7."""
    err = free_wave_error(oracle_mod, shape)
    print(shape, "free waves", err)
    assert err < 1e-6


def constraint_ratios(orc, steps=100, dt=0.01):
    """Hamiltonian and momentum constraint RMS at t = steps dt from
    mk_initial_psi data of fluctuation amplitude A and 2 A (16^3, k_min 0.1,
    m = 1, RK4): their ratios, 4 for violations of second order in A, 2 for
    first order"""
    from oracle import ic
    shape, n = (16, 16, 16), 16
    rms = []
    for amp in (1e-4, 2e-4):
        o = orc.Oracle(*shape, bounds=BOX_CONFIG, mass=1.0, gw=False,
                       filt=True, bins=8, threads=1)
        y = ic.gaussian_state(shape, BOX_CONFIG, 1.0, 1, amp, 1.0093, -0.7137,
                              1.0, n * 0.1 / 4, o.Ntot)
        o.mk_initial_psi(y, np.zeros_like(y))
        saves = []
        o.run_rk4(y, 0.0, 1e9, dt, max_steps=steps, skip=10 ** 9, saves=saves)
        c = saves[-1][2][24:28]
        rms.append(np.sqrt(c[[0, 2]]))
    return rms[1] / rms[0]


def test_constraints_stay_second_order(oracle_mod):
    """evaluation/evalConstraints.m's check: the constraint errors scale with
    the square of the perturbation (here at t = 1 after 100 RK4 steps)."""
    r = constraint_ratios(oracle_mod)
    print("constraint RMS ratio at doubled amplitude (ham, mom):", r)
    assert np.all(np.abs(r - 4.0) < 0.1)


# planted mistakes: (text in oracle/oracle.c, replacement); each changes one
# coefficient, sign or slot of the terms the pins above cover
MISTAKES = [
    ("(1.0 + 4.0 * ps) * c->lap[i]", "(1.0 + 2.0 * ps) * c->lap[i]"),
    ("(h3 - 4.0 * dps) * dph", "(h3 + 4.0 * dps) * dph"),
    ("(1.0 + 2.0 * ps) * potential_prime", "(1.0 - 2.0 * ps) * potential_prime"),
    ("(ps - 0.5) * pm", "(ps + 0.5) * pm"),
    ("4.0 * hubble * dps;", "3.0 * hubble * dps;"),
    ("0.5 * c->pressure[i]", "c->pressure[i]"),
    ("double t1 = (0.5 - ps) * dph * dph;", "double t1 = (0.5 + ps) * dph * dph;"),
    ("double t2 = (0.5 + ps) * c->grad[i] / a2;",
     "double t2 = (0.5 - ps) * c->grad[i] / a2;"),
    ("t1 - t2 / 3.0", "t1 - t2 / 2.0"),
    ("-(2.0 * pm + c->ksq[i] / a2)", "-(pm + c->ksq[i] / a2)"),
    ("-(2.0 * pm + c->ksq[i] / a2)", "-(2.0 * pm - c->ksq[i] / a2)"),
    ("2.0 * creal(stt[0][i]) / a2", "creal(stt[0][i]) / a2"),
    ("result[Nh1 + i] = f[Ndh1 + i];", "result[Nh1 + i] = f[Nh1 + i];"),
    ("tmp * f[Nh1 + i1] - h3 * f[Ndh1 + i1]", "tmp * f[Nh1 + i1] - hubble * f[Ndh1 + i1]"),
    ("result[Ndh2 + i1] = tmp * f[Nh2 + i1]", "result[Ndh2 + i1] = tmp * f[Nh1 + i1]"),
    ("s[1][i] = s[1][i] - fx * k2 - fy * k1 + 0.5 * fx * ky * s1;",
     "s[1][i] = s[1][i] - fx * k2 - fy * k1 - 0.5 * fx * ky * s1;"),
    ("result[c->Ntot - 1] = f[c->Ntot - 1] * sqrt(c->rho_mean / 3.0);",
     "result[c->Ntot - 1] = f[c->Ntot - 1] * sqrt(c->rho_mean / 2.0);"),
]


def pin_report(orc):
    """every pin's figure on the oracle in use"""
    rep = {}
    for shape in SHAPES[:1]:
        y = _scalar_state(shape, 3)
        out = oracle_outputs(orc, y, shape)
        rep["scalar"] = max(scalar_residuals(y, out, shape).values())
        rep["tensor"] = tensor_residual(y, out, shape)
    rep["free_waves"] = free_wave_error(orc, (8, 8, 8), steps=200)
    rep["constraints"] = float(np.abs(constraint_ratios(orc, steps=50) - 4.0).max())
    return rep


def _caught(rep):
    return [k for k, v, lim in (("scalar", rep["scalar"], 1e-12),
                                ("tensor", rep["tensor"], 1e-12),
                                ("free_waves", rep["free_waves"], 1e-6),
                                ("constraints", rep["constraints"], 0.1))
            if not v < lim]


@pytest.mark.parametrize("mistake", MISTAKES, ids=[m[1][:40] for m in MISTAKES])
def test_planted_mistake_is_caught(oracle_mod, tmp_path, mistake):
    src = open(os.path.join(ROOT, "oracle", "oracle.c")).read()
    old, new = mistake
    assert src.count(old) == 1, old
    d = tmp_path / "orc"
    d.mkdir()
    (d / "oracle.c").write_text(src.replace(old, new))
    shutil.copy(os.path.join(ROOT, "oracle", "oracle.h"), d / "oracle.h")
    so = d / "libmutant.so"
    subprocess.run(["gcc", "-std=c99", "-O1", "-fopenmp", "-fPIC", "-shared",
                    str(d / "oracle.c"), "-o", str(so), "-lm"], check=True)
    with oracle_mod.library(str(so)):
        rep = pin_report(oracle_mod)
    caught = _caught(rep)
    print(new, "->", rep, "caught by", caught)
    assert caught, rep


def test_unchanged_oracle_passes_every_pin(oracle_mod):
    rep = pin_report(oracle_mod)
    print(rep)
    assert not _caught(rep), rep
