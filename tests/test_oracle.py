"""CPU checks that pin the oracle (oracle/oracle.c) to independent answers:
numpy's FFT and a brute-force DFT, the closed forms of the reference's own
MATLAB checks (evaluation/testLaplacian.m, evaluation/mkPowerSpectrum.m), the
textbook TT projector, the homogeneous FLRW equations (the three ODEs of
helpers/flrw_homogeneous.py, restated) integrated by scipy, scipy's
independent DOP853 table and the Runge-Kutta order conditions."""
import numpy as np
import pytest

from tests.states import BOX_CONFIG, blocks, kgrids, make_state, rel_l2


# --------------------------------------------------------------------------
# FFT (fft / ifft of src/toolbox.c:806-827)
# --------------------------------------------------------------------------

@pytest.mark.parametrize("shape", [(4, 6, 8), (8, 8, 8), (6, 5, 7), (32, 16, 8),
                                   (16, 32, 64)])
def test_fft_matches_numpy(oracle_mod, shape):
    o = oracle_mod.Oracle(*shape)
    rng = np.random.default_rng(0)
    x = rng.standard_normal(shape)
    assert rel_l2(o.rfftn(x), np.fft.rfftn(x)) < 1e-14
    X = rng.standard_normal(o.cshape) + 1j * rng.standard_normal(o.cshape)
    ref = np.fft.irfftn(X, s=shape, axes=(0, 1, 2)) * np.prod(shape)
    assert rel_l2(o.irfftn(X), ref) < 1e-14


def test_fft_brute_force_dft(oracle_mod):
    shape = (4, 6, 8)
    o = oracle_mod.Oracle(*shape)
    rng = np.random.default_rng(1)
    x = rng.standard_normal(shape)
    n = np.indices(shape).reshape(3, -1)
    X = np.zeros((4, 6, 5), complex)
    for kx in range(4):
        for ky in range(6):
            for kz in range(5):
                ph = kx * n[0] / 4 + ky * n[1] / 6 + kz * n[2] / 8
                X[kx, ky, kz] = np.sum(x.ravel() * np.exp(-2j * np.pi * ph))
    assert rel_l2(o.rfftn(x), X) < 1e-14


# --------------------------------------------------------------------------
# k grid and spectral derivatives (mk_k_grid, src/setup.c:609-666;
# mk_gradient_squared_and_laplacian, src/toolbox.c:129-149)
# --------------------------------------------------------------------------

def test_kgrid_conventions(oracle_mod):
    shape = (8, 6, 8)
    bounds = ((0.0, 4.0), (-1.0, 2.0), (0.0, 2 * np.pi))
    o = oracle_mod.Oracle(*shape, bounds=bounds)
    k = o.kgrid()
    L = [b[1] - b[0] for b in bounds]
    # testLaplacian.m: derivative k = [0:N/2-1 0 -N/2+1:-1], k^2 from
    # [0:N/2 -N/2+1:-1]
    def deriv(n, l):
        return 2 * np.pi / l * np.array(list(range(n // 2)) + [0] +
                                        list(range(-n // 2 + 1, 0)))

    def full(n, l):
        return 2 * np.pi / l * np.array(list(range(n // 2 + 1)) +
                                        list(range(-n // 2 + 1, 0)))
    kxd = deriv(8, L[0])[:, None, None]
    kyd = deriv(6, L[1])[None, :, None]
    kzd = deriv(8, L[2])[:5][None, None, :]
    kzd = np.where(np.arange(5) == 4, 0.0, 2 * np.pi / L[2] * np.arange(5))[None, None, :]
    assert np.allclose(k["x"], np.broadcast_to(kxd, o.cshape), rtol=0, atol=1e-15)
    assert np.allclose(k["y"], np.broadcast_to(kyd, o.cshape), rtol=0, atol=1e-15)
    assert np.allclose(k["z"], np.broadcast_to(kzd, o.cshape), rtol=0, atol=1e-15)
    kxf, kyf = full(8, L[0]), full(6, L[1])
    kzf = 2 * np.pi / L[2] * np.arange(5)
    KX, KY, KZ = np.meshgrid(kxf, kyf, kzf, indexing="ij")
    assert np.allclose(k["xf"], KX, rtol=1e-15)
    assert np.allclose(k["sq"], KX ** 2 + KY ** 2 + KZ ** 2, rtol=1e-14)
    kmax = np.sqrt(sum((n // 2 * 2 * np.pi / l) ** 2 for n, l in zip(shape, L)))
    assert k["k_max"] == pytest.approx(kmax, rel=1e-15)


def _flat_state(o, shape, phi, dphi=0.0, a=1.0):
    y = np.zeros(o.Ntot)
    N = o.N
    y[:N] = np.broadcast_to(phi, shape).ravel()
    y[N:2 * N] = dphi
    y[-1] = a
    return y


def test_laplacian_and_gradient_closed_form(oracle_mod):
    """evaluation/testLaplacian.m: phi = sin x sin y sin z on [-pi, pi)^3."""
    n = 16
    shape = (n, n, n)
    o = oracle_mod.Oracle(*shape, bounds=((-np.pi, np.pi),) * 3, mass=0.0)
    x = -np.pi + 2 * np.pi * np.arange(n) / n
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    phi = np.sin(X) * np.sin(Y) * np.sin(Z)
    y = _flat_state(o, shape, phi)
    r = o.mk_rhs(y)
    lap = o.aux(3).reshape(shape)
    grad = o.aux(4).reshape(shape)
    assert np.abs(lap + 3 * phi).max() < 1e-13
    g = (np.cos(X) ** 2 * np.sin(Y) ** 2 * np.sin(Z) ** 2 +
         np.sin(X) ** 2 * np.cos(Y) ** 2 * np.sin(Z) ** 2 +
         np.sin(X) ** 2 * np.sin(Y) ** 2 * np.cos(Z) ** 2)
    assert np.abs(grad - g).max() < 1e-13
    # with dphi = psi = 0, a = 1, m = 0: ddphi = lap phi
    assert np.abs(r[o.N:2 * o.N] - lap.ravel()).max() < 1e-13


def test_nyquist_mode_zeroed_for_derivative_kept_for_laplacian(oracle_mod):
    n = 16
    shape = (n, n, n)
    o = oracle_mod.Oracle(*shape, bounds=((-np.pi, np.pi),) * 3, mass=0.0)
    x = -np.pi + 2 * np.pi * np.arange(n) / n
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    phi = np.cos(8 * X) * np.sin(Y) + np.cos(8 * Z)
    o.mk_rhs(_flat_state(o, shape, phi))
    gx, gz = o.aux(0), o.aux(2)
    # d/dx of the Nyquist cosine is dropped (kvec.x zeroed at N/2)
    assert np.abs(gx).max() == 0.0 or np.abs(gx).max() < 1e-13
    assert np.abs(gz).max() < 1e-13
    lap = o.aux(3).reshape(shape)
    ref = -65 * np.cos(8 * X) * np.sin(Y) - 64 * np.cos(8 * Z)
    assert np.abs(lap - ref).max() < 1e-11


def test_filter_mask(oracle_mod):
    """filter_window exp(-36 x^36), x = max_i |k_i| / k_i,max
    (src/setup.c:710-754) with the full (unzeroed) Nyquist wave numbers."""
    shape = (8, 16, 8)
    o = oracle_mod.Oracle(*shape, bounds=BOX_CONFIG)
    F = o.filter_mask()
    assert F[0, 0, 0] == 1.0
    # the Nyquist corner of every axis reaches x = 1
    assert F[4, 8, 4] == pytest.approx(np.exp(-36.0), rel=1e-15)
    assert F[4, 0, 0] == pytest.approx(np.exp(-36.0), rel=1e-15)
    assert np.all(F <= 1.0) and np.all(F > 0.0)
    # half the maximal wave number: exp(-36 / 2^36), i.e. barely filtered
    assert F[2, 0, 0] == pytest.approx(np.exp(-36.0 * 0.5 ** 36), rel=1e-15)


# --------------------------------------------------------------------------
# homogeneous limit: the three FLRW equations
# --------------------------------------------------------------------------

def _flrw(t, y, m):
    phi, dphi, a = y
    H = np.sqrt((0.5 * dphi ** 2 + 0.5 * m ** 2 * phi ** 2) / 3.0)
    return [dphi, -3.0 * H * dphi - m ** 2 * phi, H * a]


def test_homogeneous_rhs_is_flrw(oracle_mod):
    shape = (8, 8, 8)
    o = oracle_mod.Oracle(*shape, bounds=BOX_CONFIG, mass=1.3, gw=True)
    y = _flat_state(o, shape, 1.1, dphi=-0.4, a=1.7)
    r = o.mk_rhs(y)
    d = _flrw(0, [1.1, -0.4, 1.7], 1.3)
    N = o.N
    assert np.allclose(r[:N], d[0], rtol=1e-15)
    assert np.allclose(r[N:2 * N], d[1], rtol=1e-14)
    # psi'' = p / 2 + (psi - 1/2) <p> - ...: p - <p> over O(1) values, so
    # zero to the rounding of the mean (its order depends on the threads)
    assert np.abs(r[2 * N:4 * N]).max() < 1e-14
    assert r[-1] == pytest.approx(d[2], rel=1e-15)
    assert np.abs(r[4 * N:-1]).max() == 0.0


def _flrw_ref(t, m=1.0, y0=(1.0093, -0.7137, 1.0)):
    from scipy.integrate import solve_ivp
    s = solve_ivp(_flrw, (0.0, t), list(y0), method="DOP853", rtol=1e-13,
                  atol=1e-15, args=(m,))
    return s.y[:, -1]


def test_rk4_homogeneous_converges_fourth_order(oracle_mod):
    shape = (4, 4, 4)
    ref = _flrw_ref(2.0)
    errs = []
    for dt in (0.04, 0.02):
        o = oracle_mod.Oracle(*shape, bounds=BOX_CONFIG, mass=1.0)
        y = _flat_state(o, shape, 1.0093, dphi=-0.7137, a=1.0)
        n = o.run_rk4(y, 0.0, 2.0, dt, skip=10 ** 6)
        assert n == round(2.0 / dt)
        errs.append(abs(y[0] - ref[0]) + abs(y[o.N] - ref[1]) + abs(y[-1] - ref[2]))
    assert errs[1] < 1e-8
    assert 13 < errs[0] / errs[1] < 19  # 2^4


def test_dopri853_homogeneous_converges_eighth_order(oracle_mod):
    shape = (4, 4, 4)
    ref = _flrw_ref(2.0)
    errs = []
    for dt in (0.4, 0.2):
        o = oracle_mod.Oracle(*shape, bounds=BOX_CONFIG, mass=1.0)
        y = _flat_state(o, shape, 1.0093, dphi=-0.7137, a=1.0)
        r = o.run_dopri853(y, np.zeros_like(y), 0.0, 2.0, dt,
                           forced=[dt] * round(2.0 / dt))
        assert r["t"] == pytest.approx(2.0)
        errs.append(abs(y[0] - ref[0]) + abs(y[o.N] - ref[1]) + abs(y[-1] - ref[2]))
    assert 150 < errs[0] / errs[1] < 400  # 2^8 = 256


def test_dopri853_adaptive_meets_tolerance(oracle_mod):
    shape = (4, 4, 4)
    o = oracle_mod.Oracle(*shape, bounds=BOX_CONFIG, mass=1.0)
    y = _flat_state(o, shape, 1.0093, dphi=-0.7137, a=1.0)
    r = o.run_dopri853(y, np.zeros_like(y), 0.0, 2.0, 1e-3, rtol=1e-10,
                       atol=1e-14, hubble_frac=0.0)
    ref = _flrw_ref(2.0)
    assert r["t"] == 2.0
    assert abs(y[0] - ref[0]) < 1e-8 and abs(y[-1] - ref[2]) < 1e-8
    # the loop counter is not advanced on the final break (trap of
    # src/dopri853.c:118-140): n_stp + 1 steps were taken
    assert r["n_ok"] + r["n_bad"] == r["n_stp"] + 1


# --------------------------------------------------------------------------
# DOP853 tableau (src/dopri853_constants.c) and controller
# --------------------------------------------------------------------------

def test_tableau_matches_scipy(oracle_mod):
    import scipy.integrate._ivp.dop853_coefficients as d
    t = oracle_mod.tableau()
    assert np.allclose(t["a"][:12, :12], d.A[:12, :12], rtol=2e-16, atol=0)
    assert np.allclose(t["c"][:12], d.C[:12], rtol=2e-16, atol=0)
    assert np.allclose(t["b"], d.B, rtol=2e-16, atol=0)
    bhh = np.array([d.B[0] - d.E3[0], d.B[8] - d.E3[8], d.B[11] - d.E3[11]])
    assert np.allclose(t["bhh"], bhh, rtol=4e-16, atol=0)
    assert np.allclose(t["er"], d.E5[:12], rtol=2e-16, atol=0)


def test_tableau_order_conditions(oracle_mod):
    t = oracle_mod.tableau()
    A, c, b = t["a"][:12, :12], t["c"][:12], t["b"]
    assert np.allclose(A.sum(axis=1), c, atol=1e-14)
    for q in range(1, 9):
        assert np.dot(b, c ** (q - 1)) == pytest.approx(1.0 / q, rel=1e-13)
    assert t["bhh"].sum() == pytest.approx(1.0, rel=1e-15)
    assert abs(t["er"].sum()) < 1e-14


def test_success_closed_forms(oracle_mod):
    s = oracle_mod.success
    # accept: scale = safe * err^(-1/8), next dt = dt * scale
    ok, dt, nxt, eo, rej = s(0.5, 0.01, rho_mean=1e-12)
    assert ok and dt == 0.01 and rej == 0
    assert nxt == pytest.approx(0.01 * 0.9 * 0.5 ** -0.125, rel=1e-15)
    assert eo == 0.5
    # Hubble cap: dt_next <= fraction * sqrt(3 / rho_mean)
    ok, dt, nxt, eo, rej = s(0.01, 1.0, rho_mean=0.75)
    assert nxt == pytest.approx(0.01 * 2.0, rel=1e-15)
    # err == 0: maxscale;  tiny err: clamped to maxscale; err_old >= 1e-4
    assert s(0.0, 0.01, rho_mean=1e-12)[2] == pytest.approx(0.03)
    assert s(1e-12, 0.01, rho_mean=1e-12)[2:4] == pytest.approx((0.03, 1e-4))
    # after a rejection the step may not grow
    assert s(0.01, 0.01, reject=1, rho_mean=1e-12)[2] == pytest.approx(0.01)
    # reject: dt *= max(safe * err^(-1/8), minscale)
    ok, dt, nxt, eo, rej = s(2.0, 0.01)
    assert not ok and rej == 1
    assert dt == pytest.approx(0.01 * 0.9 * 2.0 ** -0.125, rel=1e-15)
    assert s(1e9, 0.01)[1] == pytest.approx(0.01 * 0.333, rel=1e-15)


# --------------------------------------------------------------------------
# observables
# --------------------------------------------------------------------------

def _matlab_power_spectrum(f, bins, L):
    """Restatement of evaluation/mkPowerSpectrum.m (full FFT, 3D)."""
    N = f.size
    fk2 = np.abs(np.fft.fftn(f)) ** 2
    ks1 = [np.array(list(range(n // 2 + 1)) + list(range(-n // 2 + 1, 0)))
           for n in f.shape]
    X, Y, Z = np.meshgrid(*ks1, indexing="ij")
    ks = np.sqrt(X ** 2 + Y ** 2 + Z ** 2) * 2 * np.pi / L
    kmax = ks.max()
    ps = np.zeros(bins)
    for i in range(1, N):
        idx = int(np.fix(bins * (ks.flat[i] / kmax) - 1e-14))
        ps[idx] += fk2.flat[i] / N
    return ps


def test_power_spectrum_matches_matlab_restatement(oracle_mod):
    shape = (16, 16, 16)
    bins = 12
    L = 2 * np.pi / 0.1
    o = oracle_mod.Oracle(*shape, bounds=((0, L),) * 3, bins=bins, filt=False)
    y = make_state(shape, seed=9)
    o.mk_rhs(y, output=1, field=y)
    obs = oracle_mod.unpack_obs(o.obs(), bins)
    ref = _matlab_power_spectrum(y[:o.N].reshape(shape), bins, L)
    assert rel_l2(obs["phi_ps"], ref) < 1e-12


def test_summary_closed_form(oracle_mod):
    shape = (8, 8, 8)
    o = oracle_mod.Oracle(*shape, bounds=BOX_CONFIG, bins=4)
    y = make_state(shape, seed=2)
    o.mk_rhs(y.copy(), output=1, field=y)
    obs = oracle_mod.unpack_obs(o.obs(), 4)
    N = o.N
    for key, blk in [("phi_smry", y[:N]), ("dpsi_smry", y[3 * N:4 * N])]:
        s = obs[key]
        assert s[0] == pytest.approx(blk.mean(), rel=1e-14)
        assert s[1] == pytest.approx(blk.var(ddof=1), rel=1e-10)
        assert s[2] == blk.min() and s[3] == blk.max()


def _tt_projector(K, Sm, i, j):
    k2 = K[0] ** 2 + K[1] ** 2 + K[2] ** 2
    k2 = np.where(k2 == 0, 1.0, k2)
    P = [[(1.0 if a == b else 0.0) - K[a] * K[b] / k2 for b in range(3)]
         for a in range(3)]
    out = 0
    for l in range(3):
        for m in range(3):
            out = out + (P[i][l] * P[j][m] - 0.5 * P[i][j] * P[l][m]) * Sm(l, m)
    return out


def test_tt_projection_is_the_textbook_projector(oracle_mod):
    """The two kept components of mk_gw_sources_tt (src/toolbox.c:363-397)
    equal Lambda_ij,lm S_lm, Lambda = P P - P P / 2, P = 1 - k k / k^2."""
    shape = (16, 8, 8)
    bounds = ((0, 10.0), (-3, 5.0), (1, 4.0))
    o = oracle_mod.Oracle(*shape, bounds=bounds, gw=True, bins=10)
    y = make_state(shape, bounds, seed=3, gw=True)
    o.mk_rhs(y.copy())
    g = [o.aux(i).reshape(shape) for i in range(3)]
    Sh = {(a, b): np.fft.rfftn(g[a] * g[b]) for a in range(3) for b in range(a, 3)}

    def Sm(a, b):
        return Sh[(min(a, b), max(a, b))]
    s0 = o.aux(8).view(np.complex128).reshape(o.cshape)
    s1 = o.aux(9).view(np.complex128).reshape(o.cshape)
    nx, ny, nz = shape
    L = [b[1] - b[0] for b in bounds]
    kx = 2 * np.pi / L[0] * np.array([i if i <= nx // 2 else i - nx for i in range(nx)])
    ky = 2 * np.pi / L[1] * np.array([j if j <= ny // 2 else j - ny for j in range(ny)])
    kz = 2 * np.pi / L[2] * np.arange(nz // 2 + 1)
    K = np.meshgrid(kx, ky, kz, indexing="ij")
    eps = np.finfo(float).eps
    b1 = np.abs(K[2]) > eps
    b2 = ~b1 & (np.abs(K[1]) > eps)
    b3 = ~b1 & ~b2
    for m in (b1, b2, b3):
        m[0, 0, 0] = False
    for mask, (c0, c1) in [(b1, ((0, 0), (0, 1))), (b2, ((0, 0), (0, 2))),
                           (b3, ((1, 1), (1, 2)))]:
        assert mask.sum() > 0
        L0 = _tt_projector(K, Sm, *c0)
        L1 = _tt_projector(K, Sm, *c1)
        assert rel_l2(s0[mask], L0[mask]) < 1e-13
        assert rel_l2(s1[mask], L1[mask]) < 1e-13
    assert s0[0, 0, 0] == 0 and s1[0, 0, 0] == 0


def _full_tt(k, comps, vals):
    idx = {(0, 0): 0, (0, 1): 1, (0, 2): 2, (1, 1): 3, (1, 2): 4, (2, 2): 5}

    def I(a, b):
        return idx[(min(a, b), max(a, b))]
    A, rhs = [], []
    for a in range(3):
        row = np.zeros(6, complex)
        for b in range(3):
            row[I(a, b)] += k[b]
        A.append(row)
        rhs.append(0)
    row = np.zeros(6, complex)
    row[[0, 3, 5]] = 1
    A.append(row)
    rhs.append(0)
    for c, v in zip(comps, vals):
        row = np.zeros(6, complex)
        row[I(*c)] = 1
        A.append(row)
        rhs.append(v)
    h = np.linalg.lstsq(np.array(A), np.array(rhs, complex), rcond=None)[0]
    return np.array([[h[I(a, b)] for b in range(3)] for a in range(3)])


def test_gw_spectrum_is_full_tt_norm(oracle_mod):
    """Per mode, the GW power of mk_gw_spectrum (src/toolbox.c:435-484) is
    sum_ij |dh_ij|^2 of the transverse traceless tensor whose two kept
    components are dh1, dh2; checked on single-mode states."""
    shape = (8, 8, 8)
    bounds = ((0, 6.0), (0, 7.0), (0, 5.0))
    bins = 6
    o = oracle_mod.Oracle(*shape, bounds=bounds, gw=True, bins=bins)
    N, M = o.N, o.M
    L = [b[1] - b[0] for b in bounds]
    kmax = np.sqrt(sum((4 * 2 * np.pi / l) ** 2 for l in L))
    rng = np.random.default_rng(4)
    for (ix, iy, iz), comps in [((1, 2, 3), ((0, 0), (0, 1))),
                                ((6, 3, 0), ((0, 0), (0, 2))),
                                ((3, 0, 0), ((1, 1), (1, 2)))]:
        y = make_state(shape, bounds, seed=1, gw=False)
        Next = 2 * M
        d1 = complex(*rng.standard_normal(2))
        d2 = complex(*rng.standard_normal(2))
        lin = (ix * 8 + iy) * 5 + iz
        base1 = 4 * N + 2 * Next
        y[base1 + 2 * lin:base1 + 2 * lin + 2] = [d1.real, d1.imag]
        y[base1 + Next + 2 * lin:base1 + Next + 2 * lin + 2] = [d2.real, d2.imag]
        o.mk_rhs(y.copy(), output=1, field=y)
        gw = oracle_mod.unpack_obs(o.obs(), bins)["gw"]
        k = np.array([2 * np.pi / L[0] * (ix if ix <= 4 else ix - 8),
                      2 * np.pi / L[1] * (iy if iy <= 4 else iy - 8),
                      2 * np.pi / L[2] * iz])
        pw = np.sum(np.abs(_full_tt(k, comps, (d1, d2))) ** 2)
        if iz != 0:
            pw *= 2.0
        rho_mean = o.means()[0]
        H = np.sqrt(rho_mean / 3)
        fac = np.pi * 0.01 ** (1 / 3) / (3 * H * H * np.prod(L))
        kk = np.linalg.norm(k)
        expect = np.zeros(bins)
        expect[int(np.trunc(bins * kk / kmax - 1e-14))] = fac * kk ** 3 * pw / N
        assert rel_l2(gw, expect) < 1e-13


# --------------------------------------------------------------------------
# psi initialisation (mk_initial_psi, src/setup.c:825-903)
# --------------------------------------------------------------------------

def test_initial_psi_satisfies_constraints(oracle_mod):
    shape = (16, 16, 16)
    bins = 8
    o = oracle_mod.Oracle(*shape, bounds=BOX_CONFIG, bins=bins, filt=False)
    y = make_state(shape, seed=5, psi_amp=0.0, amp=1e-5)
    dy = np.zeros_like(y)
    o.mk_initial_psi(y, dy)
    N = o.N
    assert np.abs(y[2 * N:3 * N]).max() > 0
    o.mk_rhs(y.copy(), output=1, field=y)
    c = oracle_mod.unpack_obs(o.obs(), bins)["constraints"]
    rho = o.aux(5)
    drho2 = np.mean((rho - rho.mean()) ** 2)
    # the momentum constraint holds to rounding by construction; the
    # Hamiltonian one to first order in the fluctuations
    assert c[2] < 1e-20 * max(drho2, 1e-30) + 1e-28
    assert c[0] < 1e-6 * drho2
