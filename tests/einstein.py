"""Field equations derived from the metric, for pinning the oracle.

The reference evolves a scalar field in the perturbed FLRW metric

    ds^2 = -(1 + 2 psi) dt^2 + a^2 (1 - 2 psi) delta_ij dx^i dx^j

(README.md:91), with the tensor modes h_ij in the transverse-traceless gauge
on top of the background, g_ij = a^2 (delta_ij + h_ij), units 8 pi G = 1
(H^2 = rho_mean / 3, src/toolbox.c:87).  This module derives, with sympy,
from the metric alone (Christoffel symbols, Ricci and Einstein tensors, the
covariant d'Alembertian, the scalar field's stress tensor):

* rho = -T^0_0 and p = T^i_i / 3 to first order in psi;
* phi_tt from the Klein-Gordon equation  box phi = V'(phi),  V = m^2 phi^2 / 2,
  to first order in psi;
* psi_tt from the trace of the spatial Einstein equations
  G^i_i = T^i_i to first order in psi, the background part closed with the
  spatial means (2 a_tt / a + H^2 = -p_mean, as the reference closes its
  background);
* the linearised G_11 and G_12 (lower indices) of a transverse-traceless
  plane wave h_+ , h_x along z, whose equation G_ij = pi_ij (the
  transverse-traceless part of d_i phi d_j phi) fixes the reference's tensor
  equation.

The results are plain functions of numbers (numpy arrays of the fields and
their spatial derivatives), so the oracle's right-hand side can be checked
against them pointwise.  Nothing here restates the oracle's formulas: they
come out of the metric.
"""
import functools

import sympy as sp

t, x, y, z = sp.symbols("t x y z", real=True)
X = (t, x, y, z)
eps = sp.Symbol("epsilon")
m = sp.Symbol("m", positive=True)


def _christoffel(g, ginv):
    n = 4
    G = [[[0] * n for _ in range(n)] for _ in range(n)]
    dg = [[[sp.diff(g[i, j], X[k]) for k in range(n)] for j in range(n)]
          for i in range(n)]
    for a in range(n):
        for b in range(n):
            for c in range(b, n):
                s = 0
                for d in range(n):
                    if ginv[a, d] != 0:
                        s += ginv[a, d] * (dg[d][b][c] + dg[d][c][b] - dg[b][c][d])
                G[a][b][c] = G[a][c][b] = s / 2
    return G


def _ricci(G):
    n = 4
    R = sp.zeros(n, n)
    for b in range(n):
        for c in range(b, n):
            s = 0
            for a in range(n):
                s += sp.diff(G[a][b][c], X[a]) - sp.diff(G[a][b][a], X[c])
                for d in range(n):
                    s += G[a][a][d] * G[d][b][c] - G[a][c][d] * G[d][b][a]
            R[b, c] = R[c, b] = s
    return R


def _lin(e):
    """e to first order in eps (eps kept)"""
    return e.subs(eps, 0) + eps * sp.diff(e, eps).subs(eps, 0)


def _first_order(e):
    """e to first order in eps, eps then set to 1"""
    return _lin(e).subs(eps, 1)


def _symbols_of(expr, funcs):
    """Replace a(t), psi(...), phi(...) and their derivatives by symbols."""
    reps = {}
    for f, name in funcs:
        for d in sorted(expr.atoms(sp.Derivative), key=lambda d: -len(d.variables)):
            if d.expr == f:
                key = name + "_" + "".join(str(v) for v in d.variables)
                reps[d] = sp.Symbol(key, real=True)
    e = expr.xreplace(reps)
    for f, name in funcs:
        e = e.xreplace({f: sp.Symbol(name, real=True)})
    return e


@functools.lru_cache(maxsize=None)
def scalar_equations():
    """rho, p, phi_tt and psi_tt (first order in psi) as sympy expressions in
    the symbols phi, phi_t, phi_x.., phi_xx.., psi, psi_t, psi_x.., a, a_t,
    and (for psi_tt) p_mean."""
    a = sp.Function("a")(t)
    psi = sp.Function("psi")(*X)
    phi = sp.Function("phi")(*X)
    g = sp.diag(-(1 + 2 * eps * psi), *([a ** 2 * (1 - 2 * eps * psi)] * 3))
    ginv = sp.diag(*[1 / g[i, i] for i in range(4)])
    sqrtg = sp.sqrt(-g.det())
    V = m ** 2 * phi ** 2 / 2
    dphi = [sp.diff(phi, v) for v in X]
    kin = sum(ginv[i, i] * dphi[i] ** 2 for i in range(4))
    # T^mu_nu = g^mu al d_al phi d_nu phi - delta^mu_nu (kin / 2 + V)
    T00 = ginv[0, 0] * dphi[0] ** 2 - (kin / 2 + V)
    Tii = sum(ginv[i, i] * dphi[i] ** 2 for i in (1, 2, 3)) - 3 * (kin / 2 + V)
    rho = _first_order(-T00)
    p = _first_order(Tii / 3)
    # Klein-Gordon: (1/sqrt(-g)) d_mu (sqrt(-g) g^mu nu d_nu phi) = V'(phi)
    box = sum(sp.diff(sqrtg * ginv[i, i] * dphi[i], X[i]) for i in range(4)) / sqrtg
    kg = _lin(sp.expand(box - sp.diff(V, phi)))
    phi_tt = sp.Symbol("phi_tt")
    kg = kg.xreplace({sp.diff(phi, t, 2): phi_tt})
    sol = sp.solve(sp.Eq(kg, 0), phi_tt)[0]
    ddphi = _first_order(sol)
    # Einstein: trace of the mixed spatial components
    G = _christoffel(g, ginv)
    R = _ricci(G)
    Rs = sum(ginv[i, i] * R[i, i] for i in range(4))
    Gii = sum(ginv[i, i] * R[i, i] for i in (1, 2, 3)) - sp.Rational(3, 2) * Rs
    Gii = _lin(sp.expand(Gii))
    psi_tt = sp.Symbol("psi_tt")
    a_tt = sp.Symbol("a_tt")
    pm = sp.Symbol("p_mean")
    Gii = Gii.xreplace({sp.diff(psi, t, 2): psi_tt}).xreplace(
        {sp.diff(a, t, 2): a_tt})
    H = sp.diff(a, t) / a
    # background closed with the means: 2 a_tt / a + H^2 = -p_mean
    Gii = Gii.subs(a_tt, a * (-pm - H ** 2) / 2)
    # the truncated equation is linear in psi_tt with a constant coefficient
    sol = sp.solve(sp.Eq(Gii.subs(eps, 1), 3 * p), psi_tt)[0]
    ddpsi = sol
    funcs = [(phi, "phi"), (psi, "psi"), (a, "a")]
    out = {}
    for k, e in (("rho", rho), ("p", p), ("phi_tt", ddphi), ("psi_tt", ddpsi)):
        out[k] = sp.expand(_symbols_of(sp.expand(e), funcs))
    return out


@functools.lru_cache(maxsize=None)
def tensor_equations():
    """First-order G_11 and G_12 (lower indices) of the plane wave
    g_11 = a^2 (1 + hp), g_22 = a^2 (1 - hp), g_12 = a^2 hx, g_33 = a^2,
    hp, hx functions of (t, z); background closed with the means
    (2 a_tt / a + H^2 = -p_mean); in symbols hp, hp_t, hp_tt, hp_zz, ...,
    a, a_t, p_mean."""
    a = sp.Function("a")(t)
    hp = sp.Function("hp")(t, z)
    hx = sp.Function("hx")(t, z)
    g = sp.Matrix([[-1, 0, 0, 0],
                   [0, a ** 2 * (1 + eps * hp), a ** 2 * eps * hx, 0],
                   [0, a ** 2 * eps * hx, a ** 2 * (1 - eps * hp), 0],
                   [0, 0, 0, a ** 2]])
    ginv = g.inv()
    ginv = ginv.applyfunc(_lin)
    G = _christoffel(g, ginv)
    R = _ricci(G)
    Rs = sum(ginv[i, j] * R[i, j] for i in range(4) for j in range(4))
    out = {}
    a_tt = sp.Symbol("a_tt")
    pm = sp.Symbol("p_mean")
    H = sp.diff(a, t) / a
    for name, (i, j) in (("G11", (1, 1)), ("G12", (1, 2))):
        Gij = R[i, j] - Rs * g[i, j] / 2
        # the first-order part only (the background part is -a^2 (2 a_tt/a
        # + H^2) delta_ij, which the background equation removes)
        e1 = sp.diff(Gij, eps).subs(eps, 0)
        e1 = sp.expand(e1).xreplace({sp.diff(a, t, 2): a_tt})
        e1 = e1.subs(a_tt, a * (-pm - H ** 2) / 2)
        out[name] = sp.expand(_symbols_of(sp.expand(e1),
                                          [(hp, "hp"), (hx, "hx"), (a, "a")]))
    return out


def lambdified(exprs):
    """{name: (function of keyword arrays, the symbol names it takes)}"""
    res = {}
    for k, e in exprs.items():
        syms = sorted(e.free_symbols, key=str)
        res[k] = (sp.lambdify(syms, e, "numpy"), [str(s) for s in syms])
    return res
