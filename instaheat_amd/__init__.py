"""instaheat_amd -- InstaHeat's RHS + time-stepping hot path on MI355X.

Python mirror of the two native libraries built in this directory:

* ``libihg.so`` (HIP, gfx950): the C ABI of ``include/ihg.h`` -- fft / ifft,
  mk_rhs, the observables, mk_initial_psi and the RK4 / DOPRI853 stage
  arithmetic, on arrays resident in HBM.
* ``libinstaheat.so`` (C99): the host of ``include/instaheat.h`` -- the
  reference's run_rk4 / run_dopri853 control flow over libihg.

There is no CPU fallback: importing works without a GPU, but every call that
needs the device fails loudly when the libraries are missing or no GPU is
present.
"""
import ctypes
import os

import numpy as np

__all__ = ["Gpu", "DeviceArray", "Sim", "IhgError", "lib", "host_lib",
           "LIB_PATH", "HOST_LIB_PATH", "IHG_FILTER", "IHG_OUTPUT",
           "IH_RK4", "IH_RKF45", "IH_RKCK", "IH_DOPRI89", "IH_DOPRI853", "obs_len", "unpack_obs", "tableau",
           "default_config", "IhCtrl", "IhConfig"]

HERE = os.path.dirname(os.path.abspath(__file__))
# INSTAHEAT_AMD_LIBDIR points at another build of the same two libraries
LIBDIR = os.environ.get("INSTAHEAT_AMD_LIBDIR", HERE)
LIB_PATH = os.path.join(LIBDIR, "libihg.so")
HOST_LIB_PATH = os.path.join(LIBDIR, "libinstaheat.so")

IHG_FILTER = 1
IHG_OUTPUT = 2
IH_RK4 = 0
IH_RKF45 = 1
IH_RKCK = 2
IH_DOPRI89 = 3
IH_DOPRI853 = 4

_dptr = ctypes.POINTER(ctypes.c_double)
_fptr = ctypes.POINTER(ctypes.c_float)
_vp = ctypes.c_void_p


class IhgError(RuntimeError):
    pass


class IhgParams(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("xa", ctypes.c_double), ("xb", ctypes.c_double),
        ("ya", ctypes.c_double), ("yb", ctypes.c_double),
        ("za", ctypes.c_double), ("zb", ctypes.c_double),
        ("mass", ctypes.c_double),
        ("enable_gw", ctypes.c_int), ("enable_filter", ctypes.c_int),
        ("bins", ctypes.c_int), ("gw_fp32", ctypes.c_int),
        ("log_bins", ctypes.c_int),
    ]


class IhgRk4Ep(ctypes.Structure):
    """ihg_rk4_ep (include/ihg.h): the RK4 epilogue of ihg_mk_rhs_rk4."""
    _fields_ = [("mode", ctypes.c_int), ("half", ctypes.c_int),
                ("dt", ctypes.c_double), ("y", ctypes.c_void_p),
                ("out", ctypes.c_void_p), ("k1", ctypes.c_void_p),
                ("k2", ctypes.c_void_p), ("k3", ctypes.c_void_p)]


IHG_RK4_STAGE = 1
IHG_RK4_FINAL = 2


class IhConfig(ctypes.Structure):
    _fields_ = [
        ("grid", IhgParams),
        ("integration_method", ctypes.c_int),
        ("ti", ctypes.c_double), ("tf", ctypes.c_double), ("dt", ctypes.c_double),
        ("dt_min", ctypes.c_double), ("max_steps", ctypes.c_long),
        ("beta", ctypes.c_double), ("safe", ctypes.c_double),
        ("minscale", ctypes.c_double), ("maxscale", ctypes.c_double),
        ("rtol", ctypes.c_double), ("atol", ctypes.c_double),
        ("max_dt_hubble_fraction", ctypes.c_double),
        ("skip", ctypes.c_long), ("a_initial", ctypes.c_double),
        ("max_runtime", ctypes.c_double), ("enable_followup", ctypes.c_int),
        ("enable_stiffness_check", ctypes.c_int),
    ]


class IhDpRecord(ctypes.Structure):
    _fields_ = [("t", ctypes.c_double), ("dt", ctypes.c_double),
                ("err", ctypes.c_double), ("rho_mean", ctypes.c_double),
                ("a", ctypes.c_double), ("var_phi", ctypes.c_double),
                ("mean_psi2", ctypes.c_double), ("accepted", ctypes.c_int)]


RECORD_FIELDS = ("t", "dt", "err", "rho_mean", "a", "var_phi", "mean_psi2",
                 "accepted")
MLP_IN = 20
MLP_HIDDEN = 64
IH_HISTORY = 8


class IhDpStats(ctypes.Structure):
    _fields_ = [("n_stp", ctypes.c_long), ("n_ok", ctypes.c_long),
                ("n_bad", ctypes.c_long), ("attempts", ctypes.c_long),
                ("t", ctypes.c_double), ("dt_next", ctypes.c_double),
                ("err_old", ctypes.c_double),
                ("integration_seconds", ctypes.c_double),
                ("calls_rhs", ctypes.c_ulong), ("stop", ctypes.c_int),
                ("dtlamb", ctypes.c_double)]


# why run_dopri853 stopped (ih_dp_stats.stop)
IH_STOP_FINAL_TIME, IH_STOP_MAX_STEPS, IH_STOP_MAX_RUNTIME, IH_STOP_DT_MIN, \
    IH_STOP_UNDERFLOW, IH_STOP_FORCED, IH_STOP_STIFF = range(1, 8)
IH_ESTIFF = -2001


class IhStiffness(ctypes.Structure):
    _fields_ = [("dtlamb", ctypes.c_double), ("stiff", ctypes.c_int),
                ("nonstiff", ctypes.c_int)]


BUFFERED_FN = ctypes.CFUNCTYPE(ctypes.c_long, _vp)
FOLLOWUP_FN = ctypes.CFUNCTYPE(ctypes.c_int, _vp)


PROPOSER_FN = ctypes.CFUNCTYPE(ctypes.c_double, _vp,
                               ctypes.POINTER(IhDpRecord), ctypes.c_long,
                               ctypes.c_double)


class IhGslStats(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_long), ("attempts", ctypes.c_long),
                ("failed_steps", ctypes.c_long), ("intervals", ctypes.c_long),
                ("status", ctypes.c_int), ("t", ctypes.c_double),
                ("h", ctypes.c_double), ("integration_seconds", ctypes.c_double),
                ("calls_rhs", ctypes.c_ulong)]


class IhDpOpts(ctypes.Structure):
    _fields_ = [("forced_dt", _dptr), ("n_forced", ctypes.c_long),
                ("log", ctypes.POINTER(IhDpRecord)), ("log_cap", ctypes.c_long),
                ("proposer", PROPOSER_FN), ("proposer_user", _vp),
                ("t_out", _dptr), ("n_out", ctypes.c_long),
                ("buffered", BUFFERED_FN), ("followup", FOLLOWUP_FN),
                ("file", _vp)]


class IhMlpProposer(ctypes.Structure):
    _fields_ = [("mlp", _vp),
                ("in_mean", ctypes.c_float * MLP_IN),
                ("in_std", ctypes.c_float * MLP_IN),
                ("out_mean", ctypes.c_float), ("out_std", ctypes.c_float),
                ("min_ratio", ctypes.c_double), ("max_ratio", ctypes.c_double),
                ("calls", ctypes.c_long)]


class IhCtrl(ctypes.Structure):
    _fields_ = [("err_old", ctypes.c_double), ("dt_next", ctypes.c_double),
                ("reject", ctypes.c_int)]


SAVE_FN = ctypes.CFUNCTYPE(None, _vp, ctypes.c_double, ctypes.c_double, _dptr,
                           ctypes.c_size_t)

# every symbol declared in include/ihg.h and include/instaheat.h
IHG_SYMBOLS = [
    "ihg_create", "ihg_destroy", "ihg_sizes", "ihg_last_error", "ihg_malloc",
    "ihg_free", "ihg_upload", "ihg_download", "ihg_copy", "ihg_zero",
    "ihg_stream_copy",
    "ihg_sync", "ihg_fft_r2c", "ihg_fft_c2r", "ihg_set_field", "ihg_mk_rhs",
    "ihg_means", "ihg_observables", "ihg_calls_rhs", "ihg_mk_initial_psi",
    "ihg_rk4_stage", "ihg_rk4_final", "ihg_dp853_stage", "ihg_dp853_finish",
    "ihg_dp853_finish_noalloc", "ihg_mean", "ihg_variance", "ihg_diff_norms",
    "ihg_dp853_tableau", "ihg_dp853_dense_stage", "ihg_dp853_dense",
    "ihg_dp853_dense_tableau", "ihg_gsl_stage", "ihg_gsl_finish",
    "ihg_gsl_tableau", "ihg_download_strided", "ihg_aux", "ihg_prof_enable", "ihg_prof_reset",
    "ihg_prof_read", "ihg_prof_nfam", "ihg_prof_family", "ihg_features",
    "ihg_set_gw_fp32", "ihg_init_gaussian", "ihg_autotune", "ihg_fft_candidates",
    "ihg_set_fft_candidates", "ihg_create_ensemble", "ihg_ensemble",
    "ihg_mk_rhs_ensemble", "ihg_graph_begin", "ihg_graph_end",
    "ihg_graph_join", "ihg_graph_launch", "ihg_graph_destroy",
    "ihg_group_create", "ihg_group_destroy", "ihg_group_launches",
    "ihg_group_mk_rhs", "ihg_group_rk4_stage", "ihg_group_rk4_final",
    "ihg_group_mk_rhs_rk4", "ihg_mk_rhs_rk4",
    "ihg_mlp_nparams", "ihg_mlp_create", "ihg_mlp_destroy",
    "ihg_mlp_get_params", "ihg_mlp_set_params", "ihg_mlp_forward",
    "ihg_mlp_grad", "ihg_mlp_train",
]
HOST_SYMBOLS = [
    "ih_config_defaults", "ih_create", "ih_destroy", "ih_gpu", "ih_cfg",
    "ih_field", "ih_dfield", "ih_flags", "ih_time", "ih_set_time", "ih_set_dt",
    "ih_set_max_steps", "ih_set_tf", "ih_set_ti",
    "ih_prepare_and_save_timeslice", "ih_stiffness_update",
    "ih_mk_rhs", "ih_init_gaussian", "ih_mk_initial_psi", "ih_read_dat",
    "ih_write_dat", "ih_run_rk4",
    "ih_run_dopri853", "ih_dense_state", "ih_save_state", "ih_success", "ih_error_norm", "ih_proposer_features",
    "ih_mlp_propose", "ih_ens_create", "ih_ens_destroy", "ih_ens_gpu",
    "ih_ens_state", "ih_ens_stride", "ih_ens_nreal", "ih_ens_time",
    "ih_ens_calls_rhs", "ih_ens_init", "ih_ens_capture_step",
    "ih_ens_run_rk4", "ih_ens_run_rk4_grouped", "ih_ens_capture_step_grouped",
    "ih_ens_observables",
    "ih_run_gsl", "ih_gsl_hadjust",
]

_lib = None
_host = None


def lib():
    """Load libihg.so (the HIP library); raise if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise IhgError(f"{LIB_PATH} is missing: build it with "
                       "`python -m instaheat_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    i, d, sz = ctypes.c_int, ctypes.c_double, ctypes.c_size_t
    sig = {
        "ihg_create": (i, [ctypes.POINTER(IhgParams), ctypes.POINTER(_vp)]),
        "ihg_destroy": (i, [_vp]),
        "ihg_sizes": (i, [_vp, ctypes.POINTER(sz)]),
        "ihg_last_error": (ctypes.c_char_p, []),
        "ihg_malloc": (i, [_vp, ctypes.POINTER(_vp), sz]),
        "ihg_free": (i, [_vp, _vp]),
        "ihg_upload": (i, [_vp, _vp, _vp, sz]),
        "ihg_download": (i, [_vp, _vp, _vp, sz]),
        "ihg_copy": (i, [_vp, _vp, _vp, sz]),
        "ihg_zero": (i, [_vp, _vp, sz]),
        "ihg_stream_copy": (i, [_vp, _vp, _vp, sz]),
        "ihg_sync": (i, [_vp]),
        "ihg_fft_r2c": (i, [_vp, _vp, _vp]),
        "ihg_fft_c2r": (i, [_vp, _vp, _vp]),
        "ihg_set_field": (i, [_vp, _vp]),
        "ihg_mk_rhs": (i, [_vp, d, _vp, _vp, ctypes.c_uint]),
        "ihg_means": (i, [_vp, _dptr]),
        "ihg_observables": (i, [_vp, _dptr, sz]),
        "ihg_calls_rhs": (ctypes.c_ulong, [_vp]),
        "ihg_mk_initial_psi": (i, [_vp, d, _vp, _vp]),
        "ihg_rk4_stage": (i, [_vp, _vp, _vp, _vp, d, i]),
        "ihg_rk4_final": (i, [_vp, _vp, _vp, _vp, _vp, _vp, d]),
        "ihg_dp853_stage": (i, [_vp, i, d, _vp, ctypes.POINTER(_vp), _vp]),
        "ihg_dp853_finish": (i, [_vp, d, _vp, ctypes.POINTER(_vp), _vp, d, d,
                                 _dptr]),
        "ihg_dp853_finish_noalloc": (i, [_vp, ctypes.c_long, ctypes.c_long, d,
                                         _vp, ctypes.POINTER(_vp), _vp, d, d,
                                         _dptr]),
        "ihg_mean": (i, [_vp, _vp, sz, _dptr]),
        "ihg_variance": (i, [_vp, d, _vp, sz, _dptr]),
        "ihg_diff_norms": (i, [_vp, _vp, _vp, _vp, _vp, sz, _dptr]),
        "ihg_dp853_tableau": (i, [_dptr] * 5),
        "ihg_dp853_dense_stage": (i, [_vp, i, d, _vp, _vp, _vp]),
        "ihg_dp853_dense": (i, [_vp, d, d, _vp, _vp, _vp, _vp]),
        "ihg_dp853_dense_tableau": (i, [_dptr] * 3),
        "ihg_aux": (i, [_vp, i, _dptr]),
        "ihg_download_strided": (i, [_vp, _vp, ctypes.POINTER(ctypes.c_int), _dptr]),
        "ihg_gsl_stage": (i, [_vp, i, i, ctypes.c_double, _vp,
                              ctypes.POINTER(_vp), _vp]),
        "ihg_gsl_finish": (i, [_vp, i, ctypes.c_double, _vp, ctypes.POINTER(_vp),
                               _vp, ctypes.c_double, ctypes.c_double, _dptr]),
        "ihg_gsl_tableau": (i, [i, ctypes.POINTER(ctypes.c_int), _dptr, _dptr,
                                _dptr, _dptr]),
        "ihg_prof_enable": (i, [_vp, i]),
        "ihg_prof_reset": (i, [_vp]),
        "ihg_prof_read": (i, [_vp, _dptr, _dptr, ctypes.POINTER(ctypes.c_long), i]),
        "ihg_prof_nfam": (i, []),
        "ihg_prof_family": (ctypes.c_char_p, [i]),
        "ihg_features": (i, [_vp, _vp, _dptr]),
        "ihg_set_gw_fp32": (i, [_vp, i]),
        "ihg_autotune": (i, [_vp, i, ctypes.POINTER(i), _fptr, i]),
        "ihg_init_gaussian": (i, [_vp, _vp, ctypes.c_uint64, d, d, d, d, d]),
        "ihg_fft_candidates": (i, [_vp, ctypes.POINTER(i), ctypes.POINTER(i)]),
        "ihg_set_fft_candidates": (i, [_vp, i, i]),
        "ihg_create_ensemble": (i, [ctypes.POINTER(IhgParams), i,
                                    ctypes.POINTER(_vp)]),
        "ihg_ensemble": (i, [_vp, ctypes.POINTER(ctypes.c_long)]),
        "ihg_mk_rhs_ensemble": (i, [_vp, d, _vp, _vp, ctypes.c_uint]),
        "ihg_graph_begin": (i, [_vp]),
        "ihg_graph_end": (i, [_vp, ctypes.POINTER(_vp)]),
        "ihg_graph_join": (i, [ctypes.POINTER(_vp), i, ctypes.POINTER(_vp)]),
        "ihg_graph_launch": (i, [_vp, _vp]),
        "ihg_group_create": (i, [ctypes.POINTER(_vp), i, ctypes.POINTER(_vp)]),
        "ihg_group_destroy": (i, [_vp]),
        "ihg_group_launches": (ctypes.c_long, [_vp]),
        "ihg_group_mk_rhs": (i, [_vp, d, ctypes.POINTER(_vp),
                                 ctypes.POINTER(_vp), ctypes.c_uint]),
        "ihg_group_rk4_stage": (i, [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                    ctypes.POINTER(_vp), d, i]),
        "ihg_group_rk4_final": (i, [_vp] + [ctypes.POINTER(_vp)] * 5 + [d]),
        "ihg_group_mk_rhs_rk4": (i, [_vp, d, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                     ctypes.c_uint, ctypes.POINTER(IhgRk4Ep)]),
        "ihg_mk_rhs_rk4": (i, [_vp, d, _vp, _vp, ctypes.c_uint,
                               ctypes.POINTER(IhgRk4Ep)]),
        "ihg_graph_destroy": (i, [_vp]),
        "ihg_mlp_nparams": (i, []),
        "ihg_mlp_create": (i, [ctypes.POINTER(_vp), _fptr, ctypes.c_ulong]),
        "ihg_mlp_destroy": (i, [_vp]),
        "ihg_mlp_get_params": (i, [_vp, _fptr]),
        "ihg_mlp_set_params": (i, [_vp, _fptr]),
        "ihg_mlp_forward": (i, [_vp, _fptr, _fptr, i]),
        "ihg_mlp_grad": (i, [_vp, _fptr, _fptr, i, _fptr, _fptr]),
        "ihg_mlp_train": (i, [_vp, _fptr, _fptr, i, i, i, ctypes.c_float,
                              ctypes.c_ulong, _fptr]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def host_lib():
    """Load libinstaheat.so (the C99 host); raise if it is not built."""
    global _host
    if _host is not None:
        return _host
    lib()
    if not os.path.exists(HOST_LIB_PATH):
        raise IhgError(f"{HOST_LIB_PATH} is missing: build it with "
                       "`python -m instaheat_amd.build`")
    H = ctypes.CDLL(HOST_LIB_PATH)
    i, d = ctypes.c_int, ctypes.c_double
    sig = {
        "ih_config_defaults": (None, [ctypes.POINTER(IhConfig)]),
        "ih_create": (i, [ctypes.POINTER(IhConfig), ctypes.POINTER(_vp)]),
        "ih_destroy": (None, [_vp]),
        "ih_gpu": (_vp, [_vp]),
        "ih_dense_state": (_vp, [_vp]),
        "ih_save_state": (_vp, [_vp]),
        "ih_field": (_vp, [_vp]),
        "ih_dfield": (_vp, [_vp]),
        "ih_time": (d, [_vp]),
        "ih_set_time": (None, [_vp, d]),
        "ih_set_dt": (None, [_vp, d]),
        "ih_set_max_steps": (None, [_vp, ctypes.c_long]),
        "ih_set_tf": (None, [_vp, ctypes.c_double]),
        "ih_set_ti": (None, [_vp, ctypes.c_double]),
        "ih_prepare_and_save_timeslice": (i, [_vp, SAVE_FN, _vp]),
        "ih_stiffness_update": (i, [ctypes.POINTER(IhStiffness), d, d, d]),
        "ih_cfg": (_vp, [_vp]),
        "ih_flags": (_vp, [_vp]),
        "ih_mk_rhs": (i, [_vp, d, _vp, _vp]),
        "ih_init_gaussian": (i, [_vp, ctypes.c_uint64, d, d, d, d]),
        "ih_read_dat": (i, [_vp, ctypes.c_char_p, i]),
        "ih_write_dat": (i, [_vp, ctypes.c_char_p]),
        "ih_mk_initial_psi": (i, [_vp]),
        "ih_run_rk4": (ctypes.c_long, [_vp, SAVE_FN, _vp]),
        "ih_run_dopri853": (ctypes.c_long, [_vp, ctypes.POINTER(IhDpOpts),
                                            SAVE_FN, _vp,
                                            ctypes.POINTER(IhDpStats)]),
        "ih_success": (i, [ctypes.POINTER(IhCtrl), ctypes.POINTER(IhConfig), d,
                           _dptr, d]),
        "ih_ens_create": (i, [ctypes.POINTER(IhConfig), i, ctypes.POINTER(_vp)]),
        "ih_ens_destroy": (None, [_vp]),
        "ih_ens_gpu": (_vp, [_vp]),
        "ih_ens_state": (_vp, [_vp]),
        "ih_ens_stride": (ctypes.c_long, [_vp]),
        "ih_ens_nreal": (i, [_vp]),
        "ih_ens_time": (d, [_vp]),
        "ih_ens_calls_rhs": (ctypes.c_ulong, [_vp]),
        "ih_ens_init": (i, [_vp, ctypes.POINTER(ctypes.c_uint64), _dptr, d, d, d]),
        "ih_ens_capture_step": (i, [_vp, d, ctypes.POINTER(_vp)]),
        "ih_ens_run_rk4": (i, [ctypes.POINTER(_vp), i, ctypes.c_long, d, i]),
        "ih_ens_run_rk4_grouped": (i, [ctypes.POINTER(_vp), i, ctypes.c_long, d,
                                       i]),
        "ih_ens_capture_step_grouped": (i, [ctypes.POINTER(_vp), i, d,
                                            ctypes.POINTER(_vp)]),
        "ih_ens_observables": (i, [_vp, i, _dptr, ctypes.c_size_t]),
        "ih_proposer_features": (None, [ctypes.POINTER(IhDpRecord),
                                        ctypes.c_long, _dptr]),
        "ih_mlp_propose": (d, [_vp, ctypes.POINTER(IhDpRecord), ctypes.c_long,
                               d]),
        "ih_error_norm": (d, [d, d, d, ctypes.c_size_t]),
        "ih_run_gsl": (ctypes.c_long, [_vp, ctypes.c_long, SAVE_FN, _vp,
                                       ctypes.POINTER(IhGslStats)]),
        "ih_gsl_hadjust": (i, [d, _dptr, i]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(H, name)
        fn.restype = res
        fn.argtypes = args
    _host = H
    return H


def _check(rc):
    if rc != 0:
        msg = lib().ihg_last_error().decode(errors="replace")
        raise IhgError(f"ihg error {rc}: {msg}")
    return rc


def obs_len(bins):
    return 28 + 4 * bins


def unpack_obs(obs, bins):
    """Name the parts of a packed observable vector (include/ihg.h)."""
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


def make_params(shape, bounds=((0.0, 1.0),) * 3, mass=1.0, gw=False,
                filt=True, bins=50, gw_fp32=False, log_bins=False):
    nx, ny, nz = shape
    return IhgParams(nx, ny, nz, bounds[0][0], bounds[0][1], bounds[1][0],
                     bounds[1][1], bounds[2][0], bounds[2][1], mass, int(gw),
                     int(filt), bins, int(gw_fp32), int(log_bins))


class DeviceArray:
    """A block of doubles in HBM owned by a Gpu context."""

    def __init__(self, gpu, n, ptr=None, owned=True):
        self.gpu = gpu
        self.n = int(n)
        self.owned = owned
        if ptr is None:
            p = _vp()
            _check(lib().ihg_malloc(gpu.ctx, ctypes.byref(p), self.n))
            ptr = p.value
        self.ptr = ptr

    def __del__(self):
        if self.owned and getattr(self, "ptr", None) and self.gpu.ctx:
            lib().ihg_free(self.gpu.ctx, self.ptr)
            self.ptr = None

    def offset(self, k):
        return DeviceArray(self.gpu, self.n - k, self.ptr + 8 * k, owned=False)

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        assert a.size <= self.n
        _check(lib().ihg_upload(self.gpu.ctx, self.ptr,
                                a.ctypes.data_as(_vp), a.size))

    def download(self, n=None):
        n = self.n if n is None else n
        out = np.empty(n)
        _check(lib().ihg_download(self.gpu.ctx, out.ctypes.data_as(_vp),
                                  self.ptr, n))
        return out


class Gpu:
    """One libihg context: the reference's globals for one parameter set."""

    def __init__(self, shape, bounds=((0.0, 1.0),) * 3, mass=1.0, gw=False,
                 filt=True, bins=50, gw_fp32=False, log_bins=False):
        self.params = make_params(shape, bounds, mass, gw, filt, bins, gw_fp32,
                                  log_bins)
        self.bins = bins
        c = _vp()
        _check(lib().ihg_create(ctypes.byref(self.params), ctypes.byref(c)))
        self.ctx = c.value
        s = (ctypes.c_size_t * 4)()
        _check(lib().ihg_sizes(self.ctx, s))
        self.N, self.M, self.Next, self.Ntot = [int(v) for v in s]
        self.shape = tuple(shape)
        self.cshape = (shape[0], shape[1], shape[2] // 2 + 1)

    def close(self):
        if getattr(self, "ctx", None):
            lib().ihg_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        self.close()

    def alloc(self, n):
        return DeviceArray(self, n)

    def to_device(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        d = DeviceArray(self, a.size)
        d.upload(a)
        return d

    def sync(self):
        _check(lib().ihg_sync(self.ctx))

    def rfftn(self, x):
        d_in = self.to_device(x)
        d_out = self.alloc(2 * self.M)
        _check(lib().ihg_fft_r2c(self.ctx, d_in.ptr, d_out.ptr))
        return d_out.download().view(np.complex128).reshape(self.cshape)

    def irfftn(self, X):
        Xd = np.ascontiguousarray(X, dtype=np.complex128).reshape(-1).view(np.float64)
        d_in = self.to_device(Xd)
        d_out = self.alloc(self.N)
        _check(lib().ihg_fft_c2r(self.ctx, d_in.ptr, d_out.ptr))
        return d_out.download().reshape(self.shape)

    def set_field(self, d):
        _check(lib().ihg_set_field(self.ctx, d.ptr if d is not None else None))

    def mk_rhs(self, d_f, d_res, t=0.0, flags=0):
        _check(lib().ihg_mk_rhs(self.ctx, t, d_f.ptr, d_res.ptr, flags))

    def means(self):
        out = np.zeros(2)
        _check(lib().ihg_means(self.ctx, out.ctypes.data_as(_dptr)))
        return out

    def observables(self):
        out = np.zeros(obs_len(self.bins))
        _check(lib().ihg_observables(self.ctx, out.ctypes.data_as(_dptr),
                                     out.size))
        return out

    @property
    def calls_rhs(self):
        return int(lib().ihg_calls_rhs(self.ctx))

    def mk_initial_psi(self, d_field, d_dfield, ti=0.0):
        _check(lib().ihg_mk_initial_psi(self.ctx, ti, d_field.ptr, d_dfield.ptr))

    def rk4_stage(self, out, y, k, dt, half):
        _check(lib().ihg_rk4_stage(self.ctx, out.ptr, y.ptr, k.ptr, dt, int(half)))

    def rk4_final(self, y, k1, k2, k3, k4, dt):
        _check(lib().ihg_rk4_final(self.ctx, y.ptr, k1.ptr, k2.ptr, k3.ptr,
                                   k4.ptr, dt))

    def dp853_stage(self, s, dt, y, ks, out):
        arr = (_vp * 12)(*[k.ptr if k is not None else None for k in ks])
        _check(lib().ihg_dp853_stage(self.ctx, s, dt, y.ptr, arr, out.ptr))

    def dp853_finish(self, dt, y, ks, ynew, rtol, atol):
        arr = (_vp * 12)(*[k.ptr if k is not None else None for k in ks])
        sums = np.zeros(2)
        _check(lib().ihg_dp853_finish(self.ctx, dt, y.ptr, arr, ynew.ptr, rtol,
                                      atol, sums.ctypes.data_as(_dptr)))
        return sums

    def gsl_stage(self, method, s, dt, y, ks, out):
        arr = (_vp * 13)(*[k.ptr if k is not None else None for k in ks])
        _check(lib().ihg_gsl_stage(self.ctx, method, s, dt, y.ptr, arr, out.ptr))

    def gsl_finish(self, method, dt, y, ks, ynew, eps_abs, eps_rel):
        """y_new and the standard control's error ratio"""
        arr = (_vp * 13)(*[k.ptr if k is not None else None for k in ks])
        r = ctypes.c_double()
        _check(lib().ihg_gsl_finish(self.ctx, method, dt, y.ptr, arr, ynew.ptr,
                                    eps_abs, eps_rel, ctypes.byref(r)))
        return r.value

    def aux(self, which):
        out = np.zeros(self.N)
        _check(lib().ihg_aux(self.ctx, which, out.ctypes.data_as(_dptr)))
        return out

    def features(self, f):
        """{a, variance of phi, mean of psi^2} of the state f (ihg_features)"""
        out = np.zeros(3)
        _check(lib().ihg_features(self.ctx, f.ptr, out.ctypes.data_as(_dptr)))
        return out


def prof_enable(ctx, on=True):
    _check(lib().ihg_prof_enable(ctx, int(on)))


def prof_reset(ctx):
    _check(lib().ihg_prof_reset(ctx))


def prof_read(ctx):
    """{family: (milliseconds, algorithmic bytes, launches)} since reset."""
    n = lib().ihg_prof_nfam()
    ms = np.zeros(n)
    by = np.zeros(n)
    cnt = np.zeros(n, dtype=np.int64)
    _check(lib().ihg_prof_read(ctx, ms.ctypes.data_as(_dptr),
                               by.ctypes.data_as(_dptr),
                               cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_long)),
                               n))
    return {lib().ihg_prof_family(f).decode(): (float(ms[f]), float(by[f]),
                                                 int(cnt[f]))
            for f in range(n)}


def tableau():
    a = np.zeros(13 * 12)
    c = np.zeros(13)
    b = np.zeros(12)
    bhh = np.zeros(3)
    er = np.zeros(12)
    _check(lib().ihg_dp853_tableau(*[x.ctypes.data_as(_dptr)
                                     for x in (a, c, b, bhh, er)]))
    return {"a": a.reshape(13, 12), "c": c, "b": b, "bhh": bhh, "er": er}


def default_config():
    cfg = IhConfig()
    host_lib().ih_config_defaults(ctypes.byref(cfg))
    return cfg


class Sim:
    """The C99 host (include/instaheat.h): one simulation on the GPU."""

    def __init__(self, shape, bounds=((0.0, 1.0),) * 3, mass=1.0, gw=False,
                 filt=True, bins=50, method=IH_DOPRI853, ti=0.0, tf=1.0,
                 dt=1e-7, dt_min=1e-9, max_steps=10**15, rtol=1e-10,
                 atol=1e-14, hubble_frac=1e-2, skip=1, a_initial=1.0,
                 beta=0.0, safe=0.9, minscale=0.333, maxscale=3.0,
                 gw_fp32=False, log_bins=False, max_runtime=None,
                 followup=False, stiffness_check=False):
        cfg = default_config()
        cfg.grid = make_params(shape, bounds, mass, gw, filt, bins, gw_fp32,
                               log_bins)
        cfg.integration_method = method
        cfg.ti, cfg.tf, cfg.dt, cfg.dt_min = ti, tf, dt, dt_min
        cfg.max_steps = min(int(max_steps), 2**62)
        cfg.beta, cfg.safe, cfg.minscale, cfg.maxscale = beta, safe, minscale, maxscale
        cfg.rtol, cfg.atol, cfg.max_dt_hubble_fraction = rtol, atol, hubble_frac
        cfg.skip, cfg.a_initial = skip, a_initial
        if max_runtime is not None:
            cfg.max_runtime = max_runtime
        cfg.enable_followup = int(bool(followup))
        cfg.enable_stiffness_check = int(bool(stiffness_check))
        self.cfg = cfg
        self.bins = bins
        s = _vp()
        _check(host_lib().ih_create(ctypes.byref(cfg), ctypes.byref(s)))
        self.sim = s.value
        self.ctx = host_lib().ih_gpu(self.sim)
        sz = (ctypes.c_size_t * 4)()
        _check(lib().ihg_sizes(self.ctx, sz))
        self.N, self.M, self.Next, self.Ntot = [int(v) for v in sz]
        self.shape = tuple(shape)

    def close(self):
        if getattr(self, "sim", None):
            host_lib().ih_destroy(self.sim)
            self.sim = None

    def __del__(self):
        self.close()

    @property
    def t(self):
        return host_lib().ih_time(self.sim)

    @t.setter
    def t(self, v):
        host_lib().ih_set_time(self.sim, v)

    def set_ti(self, ti):
        """INITIAL_TIME of the next run (the GSL driver's output times)"""
        host_lib().ih_set_ti(self.sim, ti)

    def prepare_and_save_timeslice(self, on_save=None):
        """prepare_and_save_timeslice: a filtered output RHS on the state and
        a save; returns (t, a, observables)"""
        got = []

        def cb(user, t, a, obs, n):
            got.append((t, a, np.ctypeslib.as_array(obs, (n,)).copy()))
            if on_save is not None:
                on_save(t)
        fn = SAVE_FN(cb)
        _check(host_lib().ih_prepare_and_save_timeslice(self.sim, fn, None))
        return got[0]

    def set_dt(self, dt):
        host_lib().ih_set_dt(self.sim, dt)
        self.cfg.dt = dt

    def set_max_steps(self, n):
        host_lib().ih_set_max_steps(self.sim, int(n))
        self.cfg.max_steps = int(n)

    def set_tf(self, tf):
        host_lib().ih_set_tf(self.sim, float(tf))
        self.cfg.tf = float(tf)

    def _dl(self, ptr, n):
        out = np.empty(n)
        _check(lib().ihg_download(self.ctx, out.ctypes.data_as(_vp), ptr, n))
        return out

    def field(self):
        return self._dl(host_lib().ih_field(self.sim), self.Ntot)

    def dfield(self):
        return self._dl(host_lib().ih_dfield(self.sim), self.Ntot)

    def snapshot(self):
        """A device copy of the state and its time (for restore)."""
        buf = _vp()
        _check(lib().ihg_malloc(self.ctx, ctypes.byref(buf), self.Ntot))
        _check(lib().ihg_stream_copy(self.ctx, buf,
                                     host_lib().ih_field(self.sim), self.Ntot))
        return {"buf": buf, "t": self.t}

    def restore(self, snap):
        """Back to a snapshot's state and time (clears the step history)."""
        _check(lib().ihg_stream_copy(self.ctx, host_lib().ih_field(self.sim),
                                     snap["buf"], self.Ntot))
        self.t = snap["t"]

    def drop(self, snap):
        if snap.get("buf"):
            _check(lib().ihg_free(self.ctx, snap["buf"]))
            snap["buf"] = None

    def set_field(self, y):
        y = np.ascontiguousarray(y, dtype=np.float64)
        assert y.size == self.Ntot
        _check(lib().ihg_upload(self.ctx, host_lib().ih_field(self.sim),
                                y.ctypes.data_as(_vp), self.Ntot))

    def init_gaussian(self, seed=1, amplitude=1e-5, phi0=1.0093, dphi0=-0.7137,
                      k_cut=None):
        if k_cut is None:
            k_cut = self.shape[0] * (2 * np.pi / (self.cfg.grid.xb - self.cfg.grid.xa)) / 4
        _check(host_lib().ih_init_gaussian(self.sim, seed, amplitude, phi0,
                                           dphi0, k_cut))

    def mk_initial_psi(self):
        _check(host_lib().ih_mk_initial_psi(self.sim))

    def dense_state(self):
        """the latest dense-output state (host copy)"""
        out = np.empty(self.Ntot)
        _check(lib().ihg_download(self.ctx, out.ctypes.data_as(ctypes.c_void_p),
                                  host_lib().ih_dense_state(self.sim), self.Ntot))
        return out

    def dense_state_active(self):
        """inside a save() call: whether it saves a dense-output state"""
        return (host_lib().ih_save_state(self.sim) ==
                host_lib().ih_dense_state(self.sim))

    def read_dat(self, path, with_psi=True):
        """initial data from a .dat file (IC_FROM_DAT_FILE_WITH_PSI, or
        WITHOUT_PSI with with_psi=False)"""
        _check(host_lib().ih_read_dat(self.sim, os.fsencode(path), int(with_psi)))

    def write_dat(self, path):
        _check(host_lib().ih_write_dat(self.sim, os.fsencode(path)))

    def run_rk4(self, saves=None, h5=None, on_save=None):
        """run_rk4; `on_save(t)` runs inside every save() call; `h5`: an
        instaheat_amd.h5.H5Output bound to this Sim that receives every
        save() as well"""
        def cb(user, t, a, obs, n):
            if saves is not None:
                saves.append((t, a, np.ctypeslib.as_array(obs, (n,)).copy()))
            if on_save is not None:
                on_save(t)
            if h5 is not None:
                h5.save(t, a, obs, n)
        fn = SAVE_FN(cb)
        n = host_lib().ih_run_rk4(self.sim, fn, None)
        if n < 0:
            _check(n)
        return n

    def run_gsl(self, n_output, saves=None, on_save=None, h5=None):
        """run_gsl_stepper with the stepper of `method` (IH_RKF45, IH_RKCK,
        IH_DOPRI89) over n_output (GSL_OUTPUT_NUMBER) intervals; returns
        the IhGslStats fields as a dict"""
        def cb(user, t, a, obs, n):
            if saves is not None:
                saves.append((t, a, np.ctypeslib.as_array(obs, (n,)).copy()))
            if on_save is not None:
                on_save(t)
            if h5 is not None:
                h5.save(t, a, obs, n)
        fn = SAVE_FN(cb)
        st = IhGslStats()
        n = host_lib().ih_run_gsl(self.sim, int(n_output), fn, None, ctypes.byref(st))
        if n < 0:
            _check(n)
        return {k: getattr(st, k) for k, _ in IhGslStats._fields_}

    def run_dopri853(self, forced=None, saves=None, log_cap=0, proposer=None,
                     t_out=None, on_save=None, h5=None):
        """run_dopri853; `proposer` is a Python callable
        (history records, n, dt_classic) -> dt, or an object with a
        ``c_proposer()`` method returning (C function, user pointer), such as
        instaheat_amd.proposer.MlpProposer.  `t_out`: increasing times of
        dense (time-based) outputs; `on_save(t)` runs inside every save()
        call (where dense_state() is the state of a time-based output);
        `h5`: an instaheat_amd.h5.H5Output bound to this Sim that receives
        every save() as well."""
        fz = None
        if forced is not None:
            fz = np.ascontiguousarray(forced, dtype=np.float64)
        to = None
        if t_out is not None:
            to = np.ascontiguousarray(t_out, dtype=np.float64)
        log = (IhDpRecord * log_cap)() if log_cap else None
        user = None
        if proposer is None:
            prop = PROPOSER_FN()
        elif hasattr(proposer, "c_proposer"):
            prop, user = proposer.c_proposer()
        else:
            prop = PROPOSER_FN(lambda u, h, n, dtc: proposer(h, n, dtc))
        buffered, followup, file = BUFFERED_FN(), FOLLOWUP_FN(), None
        if h5 is not None:
            buffered, followup, file = h5.wrapup_hooks()
        opts = IhDpOpts(fz.ctypes.data_as(_dptr) if fz is not None else None,
                        len(fz) if fz is not None else 0,
                        ctypes.cast(log, ctypes.POINTER(IhDpRecord)) if log else None,
                        log_cap, prop, user,
                        to.ctypes.data_as(_dptr) if to is not None else None,
                        len(to) if to is not None else 0,
                        buffered, followup, file)
        st = IhDpStats()

        def cb(user, t, a, obs, n):
            if saves is not None:
                saves.append((t, a, np.ctypeslib.as_array(obs, (n,)).copy()))
            if on_save is not None:
                on_save(t)
            if h5 is not None:
                h5.save(t, a, obs, n)
        fn = SAVE_FN(cb)
        n = host_lib().ih_run_dopri853(self.sim, ctypes.byref(opts), fn, None,
                                       ctypes.byref(st))
        if n == IH_ESTIFF:
            raise IhgError(f"potential stiffness detected at t = {st.t} "
                           f"(dtlamb {st.dtlamb})")
        if n < 0:
            _check(n)
        out = {k: getattr(st, k) for k, _ in IhDpStats._fields_}
        if log:
            m = min(n, log_cap)
            out["log"] = np.array([tuple(getattr(log[i], f) for f in RECORD_FIELDS)
                                   for i in range(m)], dtype=np.float64).reshape(m, len(RECORD_FIELDS))
        return out
