"""The batched form: ensembles of realizations stepped together.

``Ensemble`` wraps ih_ens (include/instaheat.h): nreal realizations of one
grid whose RHS kernels run as one launch each (the realization is the third
grid dimension).  ``run`` steps several ensembles (one per grid size) by
fixed-step RK4, by default as one HIP graph launch per step and ensemble,
each on its ensemble's own stream; with ``grouped`` every pass of all
ensembles is one launch (ih_ens_run_rk4_grouped), optionally captured as one
graph per step.  ``run_mixed`` is the mixed-size ensemble of 48 x 32^3 + 16 x 64^3 +
4 x 128^3 realizations (independent seeds, amplitudes log-uniform in
[1e-6, 1e-4], RK4 dt 0.005) with the ensemble mean and variance of the GW
spectrum per size.  The reference runs one realization per process and has
no counterpart of this form.
"""
import ctypes
import time

import numpy as np

import instaheat_amd as ih

MIXED = ((32, 48), (64, 16), (128, 4))


class Ensemble:
    def __init__(self, n, nreal, kmin=0.1, gw=True, filt=True, bins=32,
                 mass=1.0, gw_fp32=False):
        L = 2 * np.pi / kmin
        cfg = ih.default_config()
        cfg.grid = ih.make_params((n, n, n), ((0.0, L),) * 3, mass, gw, filt,
                                  bins, gw_fp32)
        cfg.integration_method = ih.IH_RK4
        self.cfg, self.n, self.nreal, self.bins, self.kmin = cfg, n, nreal, bins, kmin
        h = ctypes.c_void_p()
        ih._check(ih.host_lib().ih_ens_create(ctypes.byref(cfg), nreal,
                                              ctypes.byref(h)))
        self.h = h.value
        H = ih.host_lib()
        self.ctx = H.ih_ens_gpu(self.h)
        self.stride = H.ih_ens_stride(self.h)
        sz = (ctypes.c_size_t * 4)()
        ih._check(ih.lib().ihg_sizes(self.ctx, sz))
        self.N, self.M, self.Next, self.Ntot = [int(v) for v in sz]
        self.obs_len = 28 + 4 * bins

    def close(self):
        if getattr(self, "h", None):
            ih.host_lib().ih_ens_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    @property
    def t(self):
        return ih.host_lib().ih_ens_time(self.h)

    @property
    def calls_rhs(self):
        return ih.host_lib().ih_ens_calls_rhs(self.h)

    def init(self, seeds, amplitudes, phi0=1.0093, dphi0=-0.7137, k_cut=None):
        seeds = (ctypes.c_uint64 * self.nreal)(*[int(v) for v in seeds])
        amps = np.ascontiguousarray(amplitudes, dtype=np.float64)
        if k_cut is None:
            k_cut = self.n * self.kmin / 4
        ih._check(ih.host_lib().ih_ens_init(self.h, seeds,
                                            amps.ctypes.data_as(ih._dptr),
                                            phi0, dphi0, k_cut))

    def state(self):
        """all realizations' states, shape (nreal, Ntot)"""
        n = self.nreal * self.stride
        out = np.empty(n)
        ih._check(ih.lib().ihg_download(self.ctx, out.ctypes.data_as(ctypes.c_void_p),
                                        ih.host_lib().ih_ens_state(self.h), n))
        return out.reshape(self.nreal, self.stride)[:, :self.Ntot].copy()

    def set_state(self, y):
        y = np.asarray(y, dtype=np.float64).reshape(self.nreal, self.Ntot)
        full = np.zeros((self.nreal, self.stride))
        full[:, :self.Ntot] = y
        ih._check(ih.lib().ihg_upload(self.ctx, ih.host_lib().ih_ens_state(self.h),
                                      full.ctypes.data_as(ctypes.c_void_p),
                                      full.size))

    def autotune(self, reps=5):
        """ihg_autotune on this ensemble: the strided-pass candidates in use
        afterwards {y, x} and every candidate's time (ms)."""
        choice = (ctypes.c_int * 2)()
        ms = np.zeros(16, dtype=np.float32)
        ih._check(ih.lib().ihg_autotune(self.ctx, reps, choice,
                                        ms.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                        16))
        self.tuned = (choice[0], choice[1])
        return self.tuned, ms

    def observables(self, r):
        obs = np.zeros(self.obs_len)
        ih._check(ih.host_lib().ih_ens_observables(self.h, r,
                                                   obs.ctypes.data_as(ih._dptr),
                                                   self.obs_len))
        return obs

    def gw_spectra(self):
        return np.array([ih.unpack_obs(self.observables(r), self.bins)["gw"]
                         for r in range(self.nreal)])


def run(ensembles, nsteps, dt, use_graph=True, grouped=False):
    """RK4 steps of several ensembles: per ensemble (each on its own stream,
    eager or one graph per ensemble and step), or grouped (every pass of all
    ensembles in one launch, eager or one graph per step)."""
    arr = (ctypes.c_void_p * len(ensembles))(*[e.h for e in ensembles])
    H = ih.host_lib()
    f = H.ih_ens_run_rk4_grouped if grouped else H.ih_ens_run_rk4
    ih._check(f(arr, len(ensembles), int(nsteps), dt, int(bool(use_graph))))


def run_mixed(nsteps=1000, dt=0.005, use_graph=True, seed=1, sizes=MIXED,
              bins=32, keep_states=False, autotune=False, grouped=False):
    """The mixed-size ensemble; returns per size the GW spectrum mean and
    variance over its realizations and the timings."""
    rng = np.random.default_rng(seed)
    ens, seeds = [], 1
    for n, nreal in sizes:
        e = Ensemble(n, nreal, bins=bins)
        amps = 10.0 ** rng.uniform(-6.0, -4.0, nreal)
        e.init(range(seeds, seeds + nreal), amps)
        e.amplitudes = amps
        seeds += nreal
        ens.append(e)
    tuned = {}
    if autotune:
        for e in ens:
            tuned[e.n], _ = e.autotune()
    for e in ens:
        ih.lib().ihg_sync(e.ctx)
    calls0 = sum(e.calls_rhs for e in ens)
    t0 = time.perf_counter()
    run(ens, nsteps, dt, use_graph, grouped)
    wall = time.perf_counter() - t0
    out = {"wall_s": wall, "steps": nsteps, "dt": dt, "graph": bool(use_graph),
           "grouped": bool(grouped),
           "rhs_evals": int(sum(e.calls_rhs for e in ens) - calls0),
           "realization_steps_per_s": sum(e.nreal for e in ens) * nsteps / wall,
           "autotuned": {str(k): v for k, v in tuned.items()}, "sizes": {}}
    for e in ens:
        gw = e.gw_spectra()
        out["sizes"][e.n] = {"nreal": e.nreal, "t": e.t,
                             "gw_mean": gw.mean(0).tolist(),
                             "gw_var": gw.var(0, ddof=1).tolist() if e.nreal > 1 else None}
        if keep_states:
            out["sizes"][e.n]["states"] = e.state()
    for e in ens:
        e.close()
    return out
