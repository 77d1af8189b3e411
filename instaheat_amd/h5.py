"""Python mirror of libinstaheat_h5.so (include/instaheat_h5.h): the
reference's HDF5 output file (src/io.c) written from the host's save() hook,
and its readers (IC_FROM_H5_FILE).

h5py is not needed: the file is written and read through the HDF5 C
library that libinstaheat_h5.so links.  The library is optional; calls
raise IhgError when it was not built (no HDF5 found at build time).
"""
import ctypes
import os

import numpy as np

from . import LIBDIR, IhConfig, IhDpStats, IhgError, host_lib

H5_LIB_PATH = os.path.join(LIBDIR, "libinstaheat_h5.so")

OUT_PHI = 1 << 0
OUT_DPHI = 1 << 1
OUT_PSI = 1 << 2
OUT_DPSI = 1 << 3
OUT_RHO = 1 << 4
OUT_PHI_PS = 1 << 5
OUT_PSI_PS = 1 << 6
OUT_RHO_PS = 1 << 7
OUT_PHI_SMRY = 1 << 8
OUT_DPHI_SMRY = 1 << 9
OUT_PSI_SMRY = 1 << 10
OUT_DPSI_SMRY = 1 << 11
OUT_RHO_SMRY = 1 << 12
OUT_PRESSURE_SMRY = 1 << 13
OUT_CONSTRAINTS = 1 << 14
OUT_FIELDS = OUT_PHI | OUT_DPHI | OUT_PSI | OUT_DPSI | OUT_RHO
OUT_DEFAULT = (OUT_PHI_PS | OUT_PSI_PS | OUT_RHO_PS | OUT_PHI_SMRY |
               OUT_DPHI_SMRY | OUT_PSI_SMRY | OUT_DPSI_SMRY | OUT_RHO_SMRY |
               OUT_PRESSURE_SMRY | OUT_CONSTRAINTS)

# every function include/instaheat_h5.h declares
H5_SYMBOLS = [
    "ih_h5_opts_defaults", "ih_h5_create", "ih_h5_append", "ih_h5_bind",
    "ih_h5_save", "ih_h5_status", "ih_h5_slices", "ih_h5_write_simple",
    "ih_h5_write_dp_stats", "ih_h5_write_followup", "ih_h5_close",
    "ih_h5_last_error", "ih_h5_read_followup", "ih_h5_read_timeslice",
    "ih_h5_read", "ih_h5_read_string", "ih_h5_buffered", "ih_h5_followup",
]

_dptr = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p
_szp = ctypes.POINTER(ctypes.c_size_t)


class IhH5Opts(ctypes.Structure):
    _fields_ = [("outputs", ctypes.c_uint), ("buf_size", ctypes.c_long),
                ("stride", ctypes.c_long * 3), ("seed", ctypes.c_double),
                ("inflaton_mass", ctypes.c_double),
                ("bunch_davies_cutoff", ctypes.c_double),
                ("threads", ctypes.c_double), ("fftw_threads", ctypes.c_double),
                ("commit_hash", ctypes.c_char_p)]


_h5 = None


def h5_lib():
    """Load libinstaheat_h5.so; raise if it is not built."""
    global _h5
    if _h5 is not None:
        return _h5
    host_lib()
    if not os.path.exists(H5_LIB_PATH):
        raise IhgError(f"{H5_LIB_PATH} is missing: build it with "
                       "`python -m instaheat_amd.build` (needs HDF5)")
    L = ctypes.CDLL(H5_LIB_PATH)
    i, d, sz = ctypes.c_int, ctypes.c_double, ctypes.c_size_t
    sig = {
        "ih_h5_opts_defaults": (None, [ctypes.POINTER(IhH5Opts)]),
        "ih_h5_create": (i, [ctypes.c_char_p, ctypes.POINTER(IhConfig),
                             ctypes.POINTER(IhH5Opts), ctypes.POINTER(_vp)]),
        "ih_h5_append": (i, [_vp, d, d, _dptr, sz, _dptr, _dptr]),
        "ih_h5_bind": (i, [_vp, _vp]),
        "ih_h5_save": (None, [_vp, d, d, _dptr, sz]),
        "ih_h5_status": (i, [_vp]),
        "ih_h5_slices": (ctypes.c_long, [_vp]),
        "ih_h5_write_simple": (i, [_vp, ctypes.c_char_p, _dptr, sz]),
        "ih_h5_write_dp_stats": (i, [_vp, ctypes.POINTER(IhDpStats)]),
        "ih_h5_write_followup": (i, [_vp]),
        "ih_h5_buffered": (ctypes.c_long, [_vp]),
        "ih_h5_followup": (i, [_vp]),
        "ih_h5_close": (i, [_vp]),
        "ih_h5_last_error": (ctypes.c_char_p, []),
        "ih_h5_read_followup": (i, [_vp, ctypes.c_char_p]),
        "ih_h5_read_timeslice": (i, [_vp, ctypes.c_char_p, d, _dptr]),
        "ih_h5_read": (i, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_long,
                           _dptr, sz, _szp]),
        "ih_h5_read_string": (i, [ctypes.c_char_p, ctypes.c_char_p,
                                  ctypes.c_char_p, sz]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _h5 = L
    return L


def _check(rc):
    if rc != 0:
        msg = h5_lib().ih_h5_last_error().decode(errors="replace")
        raise IhgError(f"ih_h5 error {rc}: {msg}")
    return rc


def default_opts():
    o = IhH5Opts()
    h5_lib().ih_h5_opts_defaults(ctypes.byref(o))
    return o


class H5Output:
    """One output file (h5_create_empty_by_path ... h5_close).  `cfg` is an
    IhConfig (a Sim's .cfg); `sim` binds a Sim for full-field output and
    the followup snapshot."""

    def __init__(self, path, cfg, outputs=OUT_DEFAULT, buf_size=5,
                 stride=(1, 1, 1), seed=0.0, inflaton_mass=-1.0,
                 bunch_davies_cutoff=0.0, commit_hash=None, sim=None):
        o = default_opts()
        o.outputs = outputs
        o.buf_size = buf_size
        for d in range(3):
            o.stride[d] = stride[d]
        o.seed, o.inflaton_mass = seed, inflaton_mass
        o.bunch_davies_cutoff = bunch_davies_cutoff
        self._hash = commit_hash.encode() if commit_hash else None
        o.commit_hash = self._hash
        self.h = None
        h = _vp()
        _check(h5_lib().ih_h5_create(os.fsencode(path), ctypes.byref(cfg),
                                     ctypes.byref(o), ctypes.byref(h)))
        self.h = h.value
        self.path = path
        if sim is not None:
            self.bind(sim)

    def bind(self, sim):
        _check(h5_lib().ih_h5_bind(self.h, sim.sim))
        self._sim = sim  # the simulation outlives the writer's use of it

    def append(self, t, a, obs, fields=None, rho=None):
        obs = np.ascontiguousarray(obs, dtype=np.float64)
        f = None if fields is None else np.ascontiguousarray(fields, dtype=np.float64)
        r = None if rho is None else np.ascontiguousarray(rho, dtype=np.float64)
        _check(h5_lib().ih_h5_append(
            self.h, t, a, obs.ctypes.data_as(_dptr), obs.size,
            None if f is None else f.ctypes.data_as(_dptr),
            None if r is None else r.ctypes.data_as(_dptr)))

    def save(self, t, a, obs, n):
        """the save() hook; obs is the pointer the host passed"""
        h5_lib().ih_h5_save(self.h, t, a, obs, n)

    @property
    def status(self):
        return h5_lib().ih_h5_status(self.h)

    @property
    def slices(self):
        return h5_lib().ih_h5_slices(self.h)

    def write_simple(self, name, values):
        v = np.ascontiguousarray(np.atleast_1d(values), dtype=np.float64)
        _check(h5_lib().ih_h5_write_simple(self.h, name.encode(),
                                           v.ctypes.data_as(_dptr), v.size))

    def write_dp_stats(self, stats):
        """stats: the dict Sim.run_dopri853 returns"""
        st = IhDpStats()
        for k, _ in IhDpStats._fields_:
            if k in stats:
                setattr(st, k, stats[k])
        _check(h5_lib().ih_h5_write_dp_stats(self.h, ctypes.byref(st)))

    def wrapup_hooks(self):
        """(buffered, followup, file) for ih_dp_opts: run_dopri853's wrap-up
        saves the final slice only when this file's buffer holds slices, and
        writes the followup state here when ENABLE_FOLLOWUP is set"""
        from instaheat_amd import BUFFERED_FN, FOLLOWUP_FN
        L = h5_lib()
        return (ctypes.cast(L.ih_h5_buffered, BUFFERED_FN),
                ctypes.cast(L.ih_h5_followup, FOLLOWUP_FN), self.h)

    def write_followup(self):
        sim = getattr(self, "_sim", None)
        if sim is not None and not sim.sim:
            raise IhgError("write_followup: the bound simulation is closed")
        _check(h5_lib().ih_h5_write_followup(self.h))

    def close(self):
        if self.h:
            h, self.h = self.h, None
            _check(h5_lib().ih_h5_close(h))

    def __del__(self):
        # an unclosed file still gets its buffered slices; errors here
        # have nowhere to go
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def dims(path, name):
    d = (ctypes.c_size_t * 2)()
    _check(h5_lib().ih_h5_read(os.fsencode(path), name.encode(), -1, None, 0, d))
    return int(d[0]), int(d[1])


def read(path, name, row=None):
    """A double dataset: 2-D (slices, dim), 1-D for rank 1; one slice with
    `row`."""
    n0, n1 = dims(path, name)
    if row is None:
        out = np.empty(n0 * n1)
        _check(h5_lib().ih_h5_read(os.fsencode(path), name.encode(), -1,
                                   out.ctypes.data_as(_dptr), out.size, None))
        return out.reshape(n0, n1) if n1 != 1 else out
    out = np.empty(n1)
    _check(h5_lib().ih_h5_read(os.fsencode(path), name.encode(), int(row),
                               out.ctypes.data_as(_dptr), out.size, None))
    return out


def read_string(path, name):
    buf = ctypes.create_string_buffer(256)
    _check(h5_lib().ih_h5_read_string(os.fsencode(path), name.encode(), buf, 256))
    return buf.value.decode()


def read_followup(sim, path):
    """IC_FROM_H5_FILE: state and time from a file's followup snapshot"""
    _check(h5_lib().ih_h5_read_followup(sim.sim, os.fsencode(path)))


def read_timeslice(sim, path, t):
    """phi, dphi, psi, dpsi, a of the first slice at or after t; returns
    that slice's time"""
    out = ctypes.c_double()
    _check(h5_lib().ih_h5_read_timeslice(sim.sim, os.fsencode(path), t,
                                         ctypes.byref(out)))
    return out.value
