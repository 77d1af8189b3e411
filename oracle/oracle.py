"""ctypes binding of the CPU oracle (oracle/oracle.c).

Test infrastructure only: imported by tests/, ``__graft_entry__.smoke()`` and
the CPU-baseline leg of ``bench.py``, never by the shipped package.
"""
import contextlib
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dptr = ctypes.POINTER(ctypes.c_double)


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("xa", ctypes.c_double), ("xb", ctypes.c_double),
        ("ya", ctypes.c_double), ("yb", ctypes.c_double),
        ("za", ctypes.c_double), ("zb", ctypes.c_double),
        ("mass", ctypes.c_double),
        ("enable_gw", ctypes.c_int), ("enable_filter", ctypes.c_int),
        ("bins", ctypes.c_int), ("threads", ctypes.c_int),
        ("log_bins", ctypes.c_int),
    ]


class DpOpts(ctypes.Structure):
    _fields_ = [
        ("ti", ctypes.c_double), ("tf", ctypes.c_double), ("dt", ctypes.c_double),
        ("dt_min", ctypes.c_double), ("max_steps", ctypes.c_long),
        ("beta", ctypes.c_double), ("safe", ctypes.c_double),
        ("minscale", ctypes.c_double), ("maxscale", ctypes.c_double),
        ("rtol", ctypes.c_double), ("atol", ctypes.c_double),
        ("max_dt_hubble_fraction", ctypes.c_double),
        ("skip", ctypes.c_long),
        ("forced_dt", _dptr), ("n_forced", ctypes.c_long),
        ("enable_stiffness_check", ctypes.c_int),
    ]


class DpStats(ctypes.Structure):
    _fields_ = [("n_stp", ctypes.c_long), ("n_ok", ctypes.c_long),
                ("n_bad", ctypes.c_long), ("t", ctypes.c_double),
                ("last_err", ctypes.c_double), ("dtlamb", ctypes.c_double),
                ("stiff", ctypes.c_int)]


class Stiffness(ctypes.Structure):
    _fields_ = [("dtlamb", ctypes.c_double), ("stiff", ctypes.c_int),
                ("nonstiff", ctypes.c_int)]


class Ctrl(ctypes.Structure):
    _fields_ = [("err_old", ctypes.c_double), ("dt_next", ctypes.c_double),
                ("reject", ctypes.c_int)]


SAVE_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_double,
                           ctypes.c_double, _dptr, ctypes.c_size_t)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        _lib = _bind(ctypes.CDLL(_LIB_PATH))
    return _lib


@contextlib.contextmanager
def library(path):
    """Use the oracle built at `path` (a variant compiled by a test, e.g.
    with a planted mistake) for every call inside the block."""
    global _lib
    saved = _lib
    _lib = _bind(ctypes.CDLL(os.path.abspath(path)))
    try:
        yield _lib
    finally:
        _lib = saved


def _bind(L):
    """argument and result types of the oracle's C functions"""
    vp = ctypes.c_void_p
    L.orc_create.restype = vp
    L.orc_create.argtypes = [ctypes.POINTER(OrcParams)]
    L.orc_destroy.argtypes = [vp]
    L.orc_sizes.argtypes = [vp, ctypes.POINTER(ctypes.c_size_t)]
    L.orc_threads.argtypes = [vp]
    L.orc_threads.restype = ctypes.c_int
    L.orc_fft_r2c.argtypes = [vp, _dptr, _dptr]
    L.orc_fft_c2r.argtypes = [vp, _dptr, _dptr]
    L.orc_kgrid.argtypes = [vp] + [_dptr] * 8
    L.orc_filter_mask.argtypes = [vp, _dptr]
    L.orc_set_field.argtypes = [vp, _dptr]
    L.orc_mk_rhs.argtypes = [vp, ctypes.c_double, _dptr, _dptr,
                             ctypes.c_int, ctypes.c_int]
    L.orc_means.argtypes = [vp, _dptr]
    L.orc_obs.argtypes = [vp, _dptr]
    L.orc_calls_rhs.argtypes = [vp]
    L.orc_calls_rhs.restype = ctypes.c_ulong
    L.orc_aux.argtypes = [vp, ctypes.c_int, _dptr]
    L.orc_mk_initial_psi.argtypes = [vp, ctypes.c_double, _dptr, _dptr]
    L.orc_mean.argtypes = [_dptr, ctypes.c_size_t]
    L.orc_mean.restype = ctypes.c_double
    L.orc_variance.argtypes = [ctypes.c_double, _dptr, ctypes.c_size_t]
    L.orc_variance.restype = ctypes.c_double
    L.orc_run_rk4.argtypes = [vp, _dptr, ctypes.c_double, ctypes.c_double,
                              ctypes.c_double, ctypes.c_long, ctypes.c_long,
                              SAVE_FN, vp]
    L.orc_run_rk4.restype = ctypes.c_long
    L.orc_run_dopri853.argtypes = [vp, _dptr, _dptr, ctypes.POINTER(DpOpts),
                                   ctypes.POINTER(DpStats), SAVE_FN, vp,
                                   _dptr, _dptr,
                                   ctypes.POINTER(ctypes.c_int), ctypes.c_long]
    L.orc_run_dopri853.restype = ctypes.c_long
    L.orc_success.argtypes = [ctypes.POINTER(Ctrl), ctypes.POINTER(DpOpts),
                              ctypes.c_double, _dptr, ctypes.c_double]
    L.orc_success.restype = ctypes.c_int
    L.orc_dp_tableau.argtypes = [_dptr] * 5
    L.orc_stiffness_update.argtypes = [ctypes.POINTER(Stiffness),
                                       ctypes.c_double, ctypes.c_double,
                                       ctypes.c_double]
    L.orc_stiffness_update.restype = ctypes.c_int
    return L


def _p(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dptr)


def obs_len(bins):
    return 24 + 4 + 4 * bins


def unpack_obs(obs, bins):
    """Split a packed observable vector into named arrays."""
    obs = np.asarray(obs)
    smry = obs[:24].reshape(6, 4)
    return {
        "phi_smry": smry[0], "dphi_smry": smry[1], "psi_smry": smry[2],
        "dpsi_smry": smry[3], "rho_smry": smry[4], "pressure_smry": smry[5],
        "constraints": obs[24:28],
        "phi_ps": obs[28:28 + bins], "psi_ps": obs[28 + bins:28 + 2 * bins],
        "rho_ps": obs[28 + 2 * bins:28 + 3 * bins],
        "gw": obs[28 + 3 * bins:28 + 4 * bins],
    }


class Oracle:
    """One oracle context (the reference's globals for one parameter set)."""

    def __init__(self, nx, ny, nz, bounds=((0.0, 1.0),) * 3, mass=1.0,
                 gw=False, filt=True, bins=50, threads=0, log_bins=False):
        p = OrcParams(nx, ny, nz, bounds[0][0], bounds[0][1], bounds[1][0],
                      bounds[1][1], bounds[2][0], bounds[2][1], mass, int(gw),
                      int(filt), bins, threads, int(log_bins))
        self.params = p
        self.bins = bins
        self._L = lib()  # the library that owns ctx (see library())
        self.ctx = self._L.orc_create(ctypes.byref(p))
        if not self.ctx:
            raise ValueError("orc_create rejected the parameters")
        s = (ctypes.c_size_t * 4)()
        lib().orc_sizes(self.ctx, s)
        self.N, self.M, self.Next, self.Ntot = [int(v) for v in s]
        self.shape = (nx, ny, nz)
        self.cshape = (nx, ny, nz // 2 + 1)
        self._field = None

    def __del__(self):
        if getattr(self, "ctx", None):
            self._L.orc_destroy(self.ctx)
            self.ctx = None

    @property
    def threads(self):
        return lib().orc_threads(self.ctx)

    def rfftn(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        out = np.zeros(2 * self.M)
        lib().orc_fft_r2c(self.ctx, _p(x), _p(out))
        return out.view(np.complex128).reshape(self.cshape)

    def irfftn(self, X):
        X = np.ascontiguousarray(X, dtype=np.complex128).reshape(-1)
        Xd = X.view(np.float64)
        out = np.zeros(self.N)
        lib().orc_fft_c2r(self.ctx, _p(Xd), _p(out))
        return out.reshape(self.shape)

    def kgrid(self):
        arrs = [np.zeros(self.M) for _ in range(7)]
        kmax = np.zeros(1)
        lib().orc_kgrid(self.ctx, *[_p(a) for a in arrs], _p(kmax))
        names = ["x", "y", "z", "xf", "yf", "zf", "sq"]
        d = {n: a.reshape(self.cshape) for n, a in zip(names, arrs)}
        d["k_max"] = float(kmax[0])
        return d

    def filter_mask(self):
        out = np.zeros(self.M)
        lib().orc_filter_mask(self.ctx, _p(out))
        return out.reshape(self.cshape)

    def set_field(self, field):
        self._field = field
        lib().orc_set_field(self.ctx, _p(field))

    def mk_rhs(self, f, t=0.0, filter=0, output=0, field=None):
        """mk_rhs(t, f, result); f is mutated when filter=1."""
        if field is not None:
            self.set_field(field)
        elif self._field is None:
            self.set_field(f)
        res = np.zeros(self.Ntot)
        lib().orc_mk_rhs(self.ctx, t, _p(f), _p(res), int(filter), int(output))
        return res

    def means(self):
        out = np.zeros(2)
        lib().orc_means(self.ctx, _p(out))
        return out

    def obs(self):
        out = np.zeros(obs_len(self.bins))
        lib().orc_obs(self.ctx, _p(out))
        return out

    def aux(self, which):
        n = 2 * self.M if which >= 8 else self.N
        out = np.zeros(n)
        lib().orc_aux(self.ctx, which, _p(out))
        return out

    @property
    def calls_rhs(self):
        return int(lib().orc_calls_rhs(self.ctx))

    def mk_initial_psi(self, field, dfield, ti=0.0):
        self.set_field(field)
        lib().orc_mk_initial_psi(self.ctx, ti, _p(field), _p(dfield))

    def run_rk4(self, field, ti, tf, dt, max_steps=10**15, skip=1, saves=None):
        self.set_field(field)

        def cb(user, t, a, obs, n):
            if saves is not None:
                saves.append((t, a, np.ctypeslib.as_array(obs, (n,)).copy()))
        fn = SAVE_FN(cb)
        return lib().orc_run_rk4(self.ctx, _p(field), ti, tf, dt,
                                 min(max_steps, 2**62), skip, fn, None)

    def run_dopri853(self, field, dfield, ti, tf, dt, rtol=1e-10, atol=1e-14,
                     dt_min=1e-9, max_steps=10**15, beta=0.0, safe=0.9,
                     minscale=0.333, maxscale=3.0, hubble_frac=1e-2, skip=1,
                     forced=None, saves=None, log_cap=100000,
                     stiffness_check=False):
        self.set_field(field)
        fz = None
        if forced is not None:
            fz = np.ascontiguousarray(forced, dtype=np.float64)
        o = DpOpts(ti, tf, dt, dt_min, min(max_steps, 2**62), beta, safe,
                   minscale, maxscale, rtol, atol, hubble_frac, skip,
                   _p(fz) if fz is not None else None,
                   len(fz) if fz is not None else 0, int(bool(stiffness_check)))
        st = DpStats()
        dts = np.zeros(log_cap)
        errs = np.zeros(log_cap)
        acc = np.zeros(log_cap, dtype=np.int32)

        def cb(user, t, a, obs, n):
            if saves is not None:
                saves.append((t, a, np.ctypeslib.as_array(obs, (n,)).copy()))
        fn = SAVE_FN(cb)
        n = lib().orc_run_dopri853(
            self.ctx, _p(field), _p(dfield), ctypes.byref(o), ctypes.byref(st),
            fn, None, _p(dts), _p(errs),
            acc.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), log_cap)
        m = min(n, log_cap)
        return {"attempts": n, "dt": dts[:m], "err": errs[:m],
                "accepted": acc[:m].astype(bool), "n_stp": st.n_stp,
                "n_ok": st.n_ok, "n_bad": st.n_bad, "t": st.t,
                "last_err": st.last_err, "dtlamb": st.dtlamb,
                "stiff": bool(st.stiff)}


def tableau():
    a = np.zeros(13 * 12)
    c = np.zeros(13)
    b = np.zeros(12)
    bhh = np.zeros(3)
    er = np.zeros(12)
    lib().orc_dp_tableau(_p(a), _p(c), _p(b), _p(bhh), _p(er))
    return {"a": a.reshape(13, 12), "c": c, "b": b, "bhh": bhh, "er": er}


def success(err, dt, err_old=1e-4, reject=0, rho_mean=1.0, beta=0.0,
            safe=0.9, minscale=0.333, maxscale=3.0, hubble_frac=1e-2):
    """success() on one controller state; returns (ok, dt, dt_next, err_old,
    reject)."""
    s = Ctrl(err_old, 0.0, reject)
    o = DpOpts(0.0, 1.0, dt, 0.0, 1, beta, safe, minscale, maxscale, 0.0, 0.0,
               hubble_frac, 1, None, 0)
    d = ctypes.c_double(dt)
    ok = lib().orc_success(ctypes.byref(s), ctypes.byref(o), err,
                           ctypes.byref(d), rho_mean)
    return bool(ok), d.value, s.dt_next, s.err_old, s.reject
