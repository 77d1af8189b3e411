"""CPU restatement of the reference's GSL stepper path -- TEST
INFRASTRUCTURE ONLY (tests/ and nothing else use it).

src/gsl_stepper.c:41-76 runs GSL's ODE driver (GSL 2.x, gsl_odeiv2; GSL is
not installed here and its source is not vendored in the reference) with
one of three embedded Runge-Kutta steppers:

  RKF45    gsl_odeiv2_step_rkf45   Fehlberg 4(5), advances with the 5th
           order, error = h sum_j e_j k_j                   (order 5)
  RKCK     gsl_odeiv2_step_rkck    Cash-Karp 4(5), likewise (order 5)
  DOPRI89  gsl_odeiv2_step_rk8pd   Prince-Dormand 8(7) 13M, advances with
           the 8th order, error = h (sum bhat k - sum b k)  (order 8)

driven by gsl_odeiv2_driver_alloc_y_new(sys, T, DELTA_T,
ABSOLUTE_TOLERANCE, RELATIVE_TOLERANCE) with hmin = MINIMAL_DELTA_T, over
GSL_OUTPUT_NUMBER intervals [t, i dt], dt = (tf - ti) / GSL_OUTPUT_NUMBER,
with prepare_and_save_timeslice() after each.  This module restates GSL's
published algorithm for that path:

* driver_apply: GSL_EINVAL for an interval ending before t; step until t
  reaches t1; after every step the suggested h is capped by hmax and
  checked against hmin (GSL_ENOPROG below it).  The reference's output
  times i (tf - ti) / n are the intended ones only for ti = 0; with ti > 0
  the first intervals end before ti and the run stops there;
* evolve_apply: dydt_in = f(t0, y) once per call; a step longer than the
  rest of the interval is cut to it and lands exactly on t1; the standard
  control's verdict DEC (with an actual decrease that moves t by at least
  one ulp) undoes the step and retries with the new h;
* the standard control y_new(eps_abs, eps_rel): D_i = eps_rel |y_i| +
  eps_abs on the new y, r = max_i |yerr_i| / D_i; r > 1.1: h *= max(0.2,
  0.9 r^(-1/q)); r < 0.5: h *= min(5, max(1, 0.9 r^(-1/(q+1)))), q the
  stepper's order (5, 5, 8).

The Butcher tables are the published ones (Fehlberg 1969; Cash and Karp,
ACM TOMS 16 (1990) 201; Prince and Dormand, J. Comp. Appl. Math. 7 (1981)
67), pinned in tests/test_gsl.py by their order conditions over all rooted
trees and by convergence rates on closed-form problems; the driver and
control by the rules GSL documents.  Parity unpinned against GSL itself
(absent here): which order q the control uses (the stepper's order, 5 / 5 /
8) and that RK8PD's error is h (sum bhat k - sum b k) follow GSL's source as
published, not a run of GSL.  GSL also evaluates the RHS at every
attempt's new point (dydt_out), which this control never reads; it is not
evaluated here.
"""
from fractions import Fraction as F

import numpy as np

RKF45, RKCK, DOPRI89 = 1, 2, 3  # INTEGRATION_METHOD, src/main_template.h:52-54
GSL_SUCCESS, GSL_EINVAL, GSL_EMAXITER, GSL_ENOPROG = 0, 4, 11, 27


def _tab(c, a, b, w, order, err_mode):
    s = len(c)
    A = [[F(0)] * s for _ in range(s)]
    for i, row in a.items():
        for j, v in row.items():
            A[i - 1][j - 1] = F(v)
    B = [F(0)] * s
    W = [F(0)] * s
    for j, v in b.items():
        B[j - 1] = F(v)
    for j, v in w.items():
        W[j - 1] = F(v)
    return {"stages": s, "order": order, "err_mode": err_mode,
            "c": [F(v) for v in c], "a": A, "b": B, "w": W}


def _rkf45():
    c = [0, F(1, 4), F(3, 8), F(12, 13), 1, F(1, 2)]
    a = {2: {1: F(1, 4)},
         3: {1: F(3, 32), 2: F(9, 32)},
         4: {1: F(1932, 2197), 2: F(-7200, 2197), 3: F(7296, 2197)},
         5: {1: F(439, 216), 2: -8, 3: F(3680, 513), 4: F(-845, 4104)},
         6: {1: F(-8, 27), 2: 2, 3: F(-3544, 2565), 4: F(1859, 4104),
             5: F(-11, 40)}}
    b5 = {1: F(16, 135), 3: F(6656, 12825), 4: F(28561, 56430), 5: F(-9, 50),
          6: F(2, 55)}
    b4 = {1: F(25, 216), 3: F(1408, 2565), 4: F(2197, 4104), 5: F(-1, 5)}
    e = {j: b5.get(j, 0) - b4.get(j, 0) for j in range(1, 7)}
    return _tab(c, a, b5, e, 5, 0), b4


def _rkck():
    c = [0, F(1, 5), F(3, 10), F(3, 5), 1, F(7, 8)]
    a = {2: {1: F(1, 5)},
         3: {1: F(3, 40), 2: F(9, 40)},
         4: {1: F(3, 10), 2: F(-9, 10), 3: F(6, 5)},
         5: {1: F(-11, 54), 2: F(5, 2), 3: F(-70, 27), 4: F(35, 27)},
         6: {1: F(1631, 55296), 2: F(175, 512), 3: F(575, 13824),
             4: F(44275, 110592), 5: F(253, 4096)}}
    b5 = {1: F(37, 378), 3: F(250, 621), 4: F(125, 594), 6: F(512, 1771)}
    b4 = {1: F(2825, 27648), 3: F(18575, 48384), 4: F(13525, 55296),
          5: F(277, 14336), 6: F(1, 4)}
    e = {j: b5.get(j, 0) - b4.get(j, 0) for j in range(1, 7)}
    return _tab(c, a, b5, e, 5, 0), b4


def _rk8pd():
    c = [0, F(1, 18), F(1, 12), F(1, 8), F(5, 16), F(3, 8), F(59, 400),
         F(93, 200), F(5490023248, 9719169821), F(13, 20),
         F(1201146811, 1299019798), 1, 1]
    a = {2: {1: F(1, 18)},
         3: {1: F(1, 48), 2: F(1, 16)},
         4: {1: F(1, 32), 3: F(3, 32)},
         5: {1: F(5, 16), 3: F(-75, 64), 4: F(75, 64)},
         6: {1: F(3, 80), 4: F(3, 16), 5: F(3, 20)},
         7: {1: F(29443841, 614563906), 4: F(77736538, 692538347),
             5: F(-28693883, 1125000000), 6: F(23124283, 1800000000)},
         8: {1: F(16016141, 946692911), 4: F(61564180, 158732637),
             5: F(22789713, 633445777), 6: F(545815736, 2771057229),
             7: F(-180193667, 1043307555)},
         9: {1: F(39632708, 573591083), 4: F(-433636366, 683701615),
             5: F(-421739975, 2616292301), 6: F(100302831, 723423059),
             7: F(790204164, 839813087), 8: F(800635310, 3783071287)},
         10: {1: F(246121993, 1340847787), 4: F(-37695042795, 15268766246),
              5: F(-309121744, 1061227803), 6: F(-12992083, 490766935),
              7: F(6005943493, 2108947869), 8: F(393006217, 1396673457),
              9: F(123872331, 1001029789)},
         11: {1: F(-1028468189, 846180014), 4: F(8478235783, 508512852),
              5: F(1311729495, 1432422823), 6: F(-10304129995, 1701304382),
              7: F(-48777925059, 3047939560), 8: F(15336726248, 1032824649),
              9: F(-45442868181, 3398467696), 10: F(3065993473, 597172653)},
         12: {1: F(185892177, 718116043), 4: F(-3185094517, 667107341),
              5: F(-477755414, 1098053517), 6: F(-703635378, 230739211),
              7: F(5731566787, 1027545527), 8: F(5232866602, 850066563),
              9: F(-4093664535, 808688257), 10: F(3962137247, 1805957418),
              11: F(65686358, 487910083)},
         13: {1: F(403863854, 491063109), 4: F(-5068492393, 434740067),
              5: F(-411421997, 543043805), 6: F(652783627, 914296604),
              7: F(11173962825, 925320556), 8: F(-13158990841, 6184727034),
              9: F(3936647629, 1978049680), 10: F(-160528059, 685178525),
              11: F(248638103, 1413531060)}}
    b8 = {1: F(14005451, 335480064), 6: F(-59238493, 1068277825),
          7: F(181606767, 758867731), 8: F(561292985, 797845732),
          9: F(-1041891430, 1371343529), 10: F(760417239, 1151165299),
          11: F(118820643, 751138087), 12: F(-528747749, 2220607170),
          13: F(1, 4)}
    b7 = {1: F(13451932, 455176623), 6: F(-808719846, 976000145),
          7: F(1757004468, 5645159321), 8: F(656045339, 265891186),
          9: F(-3867574721, 1518517206), 10: F(465885868, 322736535),
          11: F(53011238, 667516719), 12: F(2, 45)}
    return _tab(c, a, b8, b7, 8, 1), b7


TABLEAUS = {}
EMBEDDED = {}
for _m, _f in ((RKF45, _rkf45), (RKCK, _rkck), (DOPRI89, _rk8pd)):
    TABLEAUS[_m], EMBEDDED[_m] = _f()


def tableau(method):
    """float64 arrays: c (s), a (s x s), b (s), w (s), order, err_mode"""
    t = TABLEAUS[method]
    s = t["stages"]
    return {"stages": s, "order": t["order"], "err_mode": t["err_mode"],
            "c": np.array([float(v) for v in t["c"]]),
            "a": np.array([[float(v) for v in row] for row in t["a"]]),
            "b": np.array([float(v) for v in t["b"]]),
            "w": np.array([float(v) for v in t["w"]])}


def embedded_weights(method):
    """the lower-order weights (4th, 4th, 7th) as float64"""
    e = EMBEDDED[method]
    s = TABLEAUS[method]["stages"]
    return np.array([float(e.get(j, 0)) for j in range(1, s + 1)])


# ---------------------------------------------------------------------------
# one step and the standard control
# ---------------------------------------------------------------------------

def step(f, t, y, k1, h, tab):
    """one embedded RK step from (t, y) with k1 = f(t, y): (y_new, yerr,
    the stage derivatives).  Stage inputs y + h sum_j a_sj k_j, y_new =
    y + h sum_j b_j k_j, yerr = h sum_j e_j k_j (err_mode 0) or
    h (sum_j bhat_j k_j - sum_j b_j k_j) (err_mode 1)."""
    s, A, c = tab["stages"], tab["a"], tab["c"]
    K = [k1]
    for i in range(1, s):
        acc = A[i][0] * K[0]
        for j in range(1, i):
            if A[i][j] != 0.0:
                acc = acc + A[i][j] * K[j]
        K.append(f(t + c[i] * h, y + h * acc))
    b, w = tab["b"], tab["w"]
    sb = sum(b[j] * K[j] for j in range(s) if b[j] != 0.0)
    sw = sum(w[j] * K[j] for j in range(s) if w[j] != 0.0)
    y_new = y + h * sb
    yerr = h * sw if tab["err_mode"] == 0 else h * (sw - sb)
    return y_new, yerr, K


HADJ_DEC, HADJ_NIL, HADJ_INC = -1, 0, 1


def rmax(y_new, yerr, eps_abs, eps_rel):
    """max_i |yerr_i| / (eps_rel |y_i| + eps_abs), NaN ratios skipped (the
    control keeps the larger of r and the running max by `>`), starting
    from DBL_MIN"""
    with np.errstate(divide="ignore", invalid="ignore"):
        r = np.abs(yerr) / np.abs(eps_rel * np.abs(y_new) + eps_abs)
    r = r[~np.isnan(r)]
    return max(float(r.max()) if r.size else 0.0, np.finfo(float).tiny)


def hadjust(r, h_old, order):
    """the standard control's verdict and new h for the error ratio r"""
    S = 0.9
    if r > 1.1:
        q = S / r ** (1.0 / order)
        if q < 0.2:
            q = 0.2
        return HADJ_DEC, q * h_old
    if r < 0.5:
        q = S / r ** (1.0 / (order + 1.0))
        if q > 5.0:
            q = 5.0
        if q < 1.0:
            q = 1.0
        return HADJ_INC, q * h_old
    return HADJ_NIL, h_old


class Driver:
    """gsl_odeiv2_driver over an RHS f(t, y) (a copy of y is passed)."""

    def __init__(self, f, method, h, eps_abs, eps_rel, hmin=0.0, nmax=0,
                 hmax=np.finfo(float).max):
        self.f = f
        self.tab = tableau(method)
        self.h = h
        self.eps_abs, self.eps_rel = eps_abs, eps_rel
        self.hmin, self.hmax, self.nmax = hmin, hmax, nmax
        self.count = 0          # attempts
        self.failed = 0         # retried after DEC
        self.rhs = 0            # RHS evaluations made here
        self.trace = []         # (t0, h, verdict) of every attempt

    def _f(self, t, y):
        self.rhs += 1
        return self.f(t, y)

    def evolve_apply(self, t, t1, y):
        t0 = t
        h0 = self.h
        dt = t1 - t0
        k1 = self._f(t0, y.copy())
        while True:
            final = h0 > dt
            if final:
                h0 = dt
            y_new, yerr, _ = step(self._f, t0, y, k1, h0, self.tab)
            self.count += 1
            t_new = t1 if final else t0 + h0
            h_old = h0
            verdict, h0 = hadjust(rmax(y_new, yerr, self.eps_abs, self.eps_rel),
                                  h_old, self.tab["order"])
            self.trace.append((t0, h_old, verdict))
            if verdict == HADJ_DEC:
                if abs(h0) < abs(h_old) and t_new + h0 != t_new:
                    self.failed += 1
                    continue
                h0 = h_old
            self.h = h0
            return t_new, y_new

    def apply(self, t, t1, y):
        """evolve y from t to t1: (status, t, y); an interval that ends
        before t (for a positive step) is GSL_EINVAL"""
        if t1 - t < 0.0:
            return GSL_EINVAL, t, y
        n = 0
        while t1 - t > 0.0:
            t, y = self.evolve_apply(t, t1, y)
            if self.nmax > 0 and n == self.nmax:
                return GSL_EMAXITER, t, y
            if abs(self.h) > self.hmax:
                self.h = self.hmax
            if abs(self.h) < self.hmin:
                return GSL_ENOPROG, t, y
            n += 1
        return GSL_SUCCESS, t, y


def run(orc, y, method, ti, tf, h, eps_abs, eps_rel, hmin, n_output,
        filter=True, obs=None):
    """run_gsl_stepper (src/gsl_stepper.c:41-76) on the oracle `orc`
    (oracle.Oracle): prepare_and_save_timeslice at ti, then n_output
    intervals ending at i (tf - ti) / n_output, each followed by
    prepare_and_save_timeslice.  Returns (status, t, y, driver, saves) with
    saves = [(t, obs)] when obs is a list."""
    saves = []

    def save(t, y):
        # filter on, output on (src/toolbox.c:96-104); y is filtered in place
        orc.mk_rhs(y, t=t, filter=1 if filter else 0, output=1, field=y)
        if obs is not None:
            saves.append((t, orc.obs()))

    def f(t, yy):
        w = yy.copy()
        return orc.mk_rhs(w, t=t, filter=0, output=0, field=w)
    y = y.copy()
    t = ti
    save(t, y)
    d = Driver(f, method, h, eps_abs, eps_rel, hmin=hmin)
    dt = (tf - ti) / n_output
    status = GSL_SUCCESS
    for i in range(1, n_output + 1):
        t1 = i * dt
        status, t, y = d.apply(t, t1, y)
        if status != GSL_SUCCESS:
            break
        save(t, y)
    return status, t, y, d, saves
