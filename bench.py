"""Benchmark of the InstaHeat hot path on one MI355X: tuned vs plain GPU.

Workload: 256^3 FP64, DOPRI853 (rtol 1e-10, atol 1e-14, Hubble cap 1e-2),
GW tensor source on, spectral filter on, k_min = 0.05 (L = 125.66), m = 1,
phi0 = 1.0093, dphi0 = -0.7137, a0 = 1, seeded Gaussian fluctuations of
amplitude 1e-5 (|dphi_k|^2 = 1/(2 a^3 omega_k), cut at N k_min / 4), psi from
the Einstein constraint, observables every TIME_STEP_SKIPS = 10 steps.

A "step" is one accepted DOPRI853 step (perform_step).  An untimed pre-roll
of the classic controller takes the run from DELTA_T at t = 0 to t = 23,
through the Hubble-capped phase (where every controller takes the capped
step) into the phase where the error estimate sets the step; W warm-up
steps follow.  From that state (a device snapshot) each form takes 8
untimed steps, filling the step history, then K timed steps:
  plain  success() alone (the reference's controller);
  tuned  the learned step proposer (instaheat_amd/proposer_model.json)
         before success(); the embedded error still judges every step.
A third form, precision (the GW source and spectrum in FP32), replays the
plain form's steps and reports its GW spectrum error per k bin.
value = the tuned form's simulated time advanced / wall seconds of its K
timed steps (the mon.integration span), every array resident in HBM; the
plain form's figures are reported beside it.  The per-kernel figures of the
roofline come from one more replay of the plain form's steps under the
live profiler.  Both forms also run the whole span t = 0 to 200 from the
initial data (RHS evaluations, rejected steps, integration seconds), which
is reported beside the timed window (--no-long-run skips it).

Prints ONE JSON line.  --gpus > 1 is refused: this build runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "sim-time/wall-s at 256³ FP64 DOPRI853 rtol 1e-10: tuned vs plain GPU"
FFT_FAMILIES = ("z_r2c", "z_c2r", "y_pass", "x_pass", "x_fanout", "z_c2r_rho",
                "z_r2c_sources", "y_tt_update_h", "x_roundtrip")
# the passes that do nothing but transform (the others fuse derivatives,
# rho and p, the GW products, the TT projection or the filter into a pass)
FFT_PLAIN = ("z_r2c", "z_c2r", "y_pass", "x_pass")
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--skip", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-long-run", action="store_true",
                    help="leave out the t = 0 to 200 comparison of both forms")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--ensemble", action="store_true",
                    help="only run the mixed-size ensemble (48 x 32^3 + 16 x "
                         "64^3 + 4 x 128^3, RK4 dt 0.005) graph vs eager")
    ap.add_argument("--long-run", type=float, default=0.0, metavar="TF",
                    help="only run both forms from t = 0 to TF at the grid "
                         "(RHS evaluations, rejections, wall time and the "
                         "end-state difference tuned vs plain)")
    ap.add_argument("--copy-probe", type=float, default=0.0, metavar="GB",
                    help="only time a plain copy kernel of GB gigabytes each "
                         "way (the HBM ceiling and the counter calibration)")
    ap.add_argument("--ensemble-forms", default="eager,graph,tuned,grouped,grouped_graph",
                    help="with --ensemble: the forms to run, comma separated "
                         "(eager, graph, tuned, grouped, grouped_graph)")
    ap.add_argument("--mlp", action="store_true",
                    help="only time the step proposer's FP32 kernels: "
                         "training (Adam epochs over a training-set-sized "
                         "sample), the batched gradient and inference")
    ap.add_argument("--next-rows", action="store_true",
                    help="only measure the follow-on paths at the workload's "
                         "grid: device initial data, dense output, HDF5 "
                         "saves, the GSL RK8PD stepper")
    return ap.parse_args()


def copy_probe(gb, reps=10):
    """ihg_stream_copy of gb GB -> gb GB, timed by the live profiler."""
    import ctypes
    import instaheat_amd as ih
    g = ih.Gpu((8, 8, 8))
    n = int(gb * 1e9 / 8) // 2 * 2
    a, b = g.alloc(n), g.alloc(n)
    L = ih.lib()
    for _ in range(2):
        ih._check(L.ihg_stream_copy(g.ctx, b.ptr, a.ptr, n))
    ih.prof_enable(g.ctx, True)
    ih.prof_reset(g.ctx)
    for _ in range(reps):
        ih._check(L.ihg_stream_copy(g.ctx, b.ptr, a.ptr, n))
    ms, by, cnt = ih.prof_read(g.ctx)["stream_copy"]
    return {"probe": "stream_copy", "bytes_each_way": 8 * n, "launches": cnt,
            "avg_us": round(1e3 * ms / cnt, 2),
            "GB_per_s": round(by / (ms / 1e3) / 1e9, 1)}


HISTORY_STEPS = 8
# simulated time of the untimed pre-roll: past the Hubble-capped phase of
# the 256^3 workload, so that the error estimate sets the timed steps
PRE_ROLL_T = 23.0
LONG_RUN_T = 200.0


def workload(n):
    kmin = 0.05
    L = 2 * np.pi / kmin
    return dict(shape=(n, n, n), bounds=((0.0, L),) * 3, mass=1.0, gw=True,
                filt=True, bins=64, kmin=kmin, kcut=n * kmin / 4,
                amplitude=1e-5, phi0=1.0093, dphi0=-0.7137, seed=1)


def timed_form(ih, sim, snap, dt0, prop, steps, profile, forced=None, tag=""):
    """HISTORY_STEPS untimed then `steps` timed accepted steps from snap;
    forced = (history dts, timed dts) replays another form's steps."""
    sim.restore(snap)
    sim.set_dt(dt0)
    sim.set_max_steps(HISTORY_STEPS - 1)
    h = sim.run_dopri853(proposer=prop, log_cap=1000,
                         forced=None if forced is None else forced[0])
    sim.set_dt(h["dt_next"])
    sim.set_max_steps(steps - 1)
    if profile:
        ih.prof_enable(sim.ctx, True)
        ih.prof_reset(sim.ctx)
    t0 = sim.t
    calls0 = ih.lib().ihg_calls_rhs(sim.ctx)
    saves = []
    wall0 = time.perf_counter()
    r = sim.run_dopri853(proposer=prop, log_cap=1000, saves=saves,
                         forced=None if forced is None else forced[1])
    wall1 = time.perf_counter()
    prof = None
    if profile:
        prof = ih.prof_read(sim.ctx)
        ih.prof_enable(sim.ctx, False)
    calls = ih.lib().ihg_calls_rhs(sim.ctx) - calls0
    sim_time = sim.t - t0
    print(f"{tag}: history to t = {t0:.4f} (dt_next {h['dt_next']:.4g}), timed "
          f"{r['n_ok']} + {r['n_bad']} steps to t = {sim.t:.4f} in "
          f"{r['integration_seconds']:.3f} s, {calls} RHS",
          file=sys.stderr, flush=True)
    if not sim_time > 0.0:
        raise RuntimeError("the timed run advanced no simulated time")
    n = r["n_ok"] + r["n_bad"]
    secs = r["integration_seconds"]

    def accepted(log):
        return log[log[:, 7] > 0.5, 1]
    return dict(value=sim_time / secs, sim_time=sim_time, secs=secs,
                wall=wall1 - wall0, steps=n, n_ok=r["n_ok"], n_bad=r["n_bad"],
                attempts=r["attempts"], calls_rhs=int(calls), prof=prof,
                t_start=t0, t_end=sim.t, dt=sim_time / max(n, 1),
                ms_per_step=1e3 * secs / max(n, 1),
                rhs_per_sim_time=calls / sim_time,
                dts=(accepted(h["log"]), accepted(r["log"])),
                last_obs=saves[-1][2] if saves else None)


def make_sim(ih, wl, skip, tf=1e9):
    sim = ih.Sim(wl["shape"], bounds=wl["bounds"], mass=wl["mass"], gw=wl["gw"],
                 filt=wl["filt"], bins=wl["bins"], method=ih.IH_DOPRI853,
                 ti=0.0, tf=tf, dt=1e-7, rtol=1e-10, atol=1e-14,
                 hubble_frac=1e-2, skip=skip)
    sim.init_gaussian(seed=wl["seed"], amplitude=wl["amplitude"],
                      phi0=wl["phi0"], dphi0=wl["dphi0"], k_cut=wl["kcut"])
    sim.mk_initial_psi()
    return sim


def long_run(args, wl, tf):
    """Both forms from the same initial state to tf: RHS evaluations,
    accepted / rejected steps, integration wall time, and how far the tuned
    form's end state and GW spectrum are from the plain form's."""
    import instaheat_amd as ih
    from instaheat_amd.proposer import MlpProposer
    prop = MlpProposer.load()
    forms, ends, gws = {}, {}, {}
    n = wl["shape"][0]
    live = slice(0, 4 * n ** 3)
    for form in ("plain", "tuned"):
        sim = make_sim(ih, wl, args.skip, tf)
        saves = []
        r = sim.run_dopri853(proposer=prop if form == "tuned" else None,
                             saves=saves)
        calls = int(ih.lib().ihg_calls_rhs(sim.ctx))
        forms[form] = {"t_end": sim.t, "steps_ok": r["n_ok"], "steps_bad": r["n_bad"],
                       "rhs_evals": calls,
                       "integration_s": round(r["integration_seconds"], 3),
                       "sim_time_per_wall_s": round(sim.t / r["integration_seconds"], 4)}
        print(f"{form}: {forms[form]}", file=sys.stderr, flush=True)
        y = sim.field()
        ends[form] = y
        gws[form] = ih.unpack_obs(saves[-1][2], wl["bins"])["gw"]
        forms[form]["a_end"] = float(y[-1])
        sim.close()
    p, t = forms["plain"], forms["tuned"]
    yp, yt = ends["plain"][live], ends["tuned"][live]
    g0, g1 = gws["plain"], gws["tuned"]
    ok = g0 > 1e-6 * g0.max()
    return {"long_run": f"{n}^3 FP64 DOPRI853 rtol 1e-10 atol 1e-14, t = 0 to {tf:g}",
            "plain": p, "tuned": t,
            "rhs_evals_tuned_over_plain": round(t["rhs_evals"] / p["rhs_evals"], 4),
            "speed_tuned_over_plain": round(t["sim_time_per_wall_s"] /
                                            p["sim_time_per_wall_s"], 4),
            "end_state_rel_l2_tuned_vs_plain": float(np.linalg.norm(yt - yp) /
                                                     np.linalg.norm(yp)),
            "gw_spectrum_rel_err_max_tuned_vs_plain": float(
                np.max(np.abs(g1[ok] - g0[ok]) / g0[ok]))}


def run_gpu(args, wl):
    import instaheat_amd as ih
    from instaheat_amd.proposer import MlpProposer
    sim = make_sim(ih, wl, args.skip, tf=PRE_ROLL_T)
    # pre-roll to t = PRE_ROLL_T, then W warm-up steps (the loop takes
    # max_steps + 1 steps)
    pre = sim.run_dopri853()
    print(f"pre-roll: {pre['n_ok']} + {pre['n_bad']} rejected steps to "
          f"t = {pre['t']:.4f}, dt_next {pre['dt_next']:.4g}", file=sys.stderr,
          flush=True)
    sim.set_tf(1e9)
    sim.set_dt(pre["dt_next"])
    w = {"dt_next": pre["dt_next"], "t": sim.t}
    if args.warmup > 0:
        sim.set_max_steps(args.warmup - 1)
        w = sim.run_dopri853()
    print(f"warm-up: {args.warmup} steps to t = {w['t']:.4f}, "
          f"dt_next {w['dt_next']:.4g}", file=sys.stderr, flush=True)
    snap = sim.snapshot()
    forms = {}
    prop = MlpProposer.load()
    for form in ("plain", "tuned"):
        forms[form] = timed_form(ih, sim, snap, w["dt_next"],
                                 prop if form == "tuned" else None,
                                 args.steps, profile=False, tag=form)
    forms["tuned"]["proposals"] = int(prop.calls)
    # per-kernel figures: the plain form's steps once more under the live
    # profiler (which keeps every launch on one stream, so each kernel is
    # timed alone)
    prof = timed_form(ih, sim, snap, w["dt_next"], None, args.steps,
                      profile=True, forced=forms["plain"]["dts"],
                      tag="profile")["prof"]
    # the copy ceiling at the stage kernels' footprint: a plain copy kernel
    # of one state array (Ntot doubles each way; restore copies the
    # snapshot into the state)
    ih.prof_enable(sim.ctx, True)
    ih.prof_reset(sim.ctx)
    for _ in range(10):
        sim.restore(snap)
    ms, by, cnt = ih.prof_read(sim.ctx)["stream_copy"]
    ih.prof_enable(sim.ctx, False)
    copy_gbs = by / (ms / 1e3) / 1e9
    # the precision form (FP32 GW source and spectrum) over the plain form's
    # steps, its GW spectrum per k bin against the plain form's
    ih._check(ih.lib().ihg_set_gw_fp32(sim.ctx, 1))
    forms["precision"] = timed_form(ih, sim, snap, w["dt_next"], None,
                                    args.steps, profile=False,
                                    forced=forms["plain"]["dts"],
                                    tag="precision")
    ih._check(ih.lib().ihg_set_gw_fp32(sim.ctx, 0))
    g64 = ih.unpack_obs(forms["plain"]["last_obs"], wl["bins"])["gw"]
    g32 = ih.unpack_obs(forms["precision"]["last_obs"], wl["bins"])["gw"]
    live = g64 > 1e-6 * g64.max()
    forms["precision"]["gw_rel_err"] = (np.abs(g32 - g64)[live] / g64[live]).tolist()
    check = None
    if not args.no_cpu_baseline:
        # ORACLE_STEPS forced steps of the plain form's first step size from
        # the warm-up state, for the oracle to take on the host cores
        sim.restore(snap)
        y_snap, t_snap = sim.field(), sim.t
        dts = np.full(ORACLE_STEPS, forms["plain"]["dts"][0][0])
        sim.run_dopri853(forced=dts)
        check = dict(y=y_snap, t=t_snap, dts=dts, y_end=sim.field())
    sim.drop(snap)
    sim.close()
    return dict(forms=forms, prof=prof, check=check, warm_t=w["t"],
                copy_gbs=copy_gbs, pre_steps=pre["n_ok"] + pre["n_bad"])


ENSEMBLE_FORMS = {"eager": (False, False, False), "graph": (True, False, False),
                  "tuned": (True, True, False), "grouped": (False, False, True),
                  "grouped_graph": (True, False, True)}


def ensemble_bench(steps, forms=tuple(ENSEMBLE_FORMS)):
    """The batched form: realization-steps per second of the mixed ensemble,
    per size on its own stream (eager launches, one graph per size and step,
    graphs with the autotuned strided passes) and grouped (every pass of all
    sizes in one launch, eager and as one graph per step), with the ensemble
    mean and variance of the GW spectrum per size (the last form run).
    `forms` picks a subset (a profiler run takes the eager ones)."""
    from instaheat_amd import ensemble
    for f in forms:
        if f not in ENSEMBLE_FORMS:
            raise SystemExit(f"unknown ensemble form {f!r}")
    out = {}
    for name in forms:
        graph, tune, grouped = ENSEMBLE_FORMS[name]
        out[name] = ensemble.run_mixed(nsteps=steps, dt=0.005, use_graph=graph,
                                       autotune=tune, grouped=grouped,
                                       keep_states=True)
        print(f"ensemble {name}: {out[name]['realization_steps_per_s']:.1f} "
              f"realization-steps/s", file=sys.stderr, flush=True)
    base = out.get("eager", out[forms[0]])

    def rate(name):
        r = out[name]
        return {"value": round(r["realization_steps_per_s"], 2),
                "wall_s": round(r["wall_s"], 3),
                "ms_per_step": round(1e3 * r["wall_s"] / steps, 3),
                "over_eager": round(r["realization_steps_per_s"] /
                                    base["realization_steps_per_s"], 3)}
    head = out.get("grouped_graph", out[forms[-1]])
    res = {"metric": "realization-steps/s of the mixed ensemble "
                     "(48 x 32^3 + 16 x 64^3 + 4 x 128^3, FP64 RK4 dt 0.005, GW on)",
           "steps": steps,
           "value": round(head["realization_steps_per_s"], 2),
           "form": ("grouped_graph (one launch per pass for all sizes, one "
                    "graph per step)") if "grouped_graph" in out else forms[-1],
           "rhs_evals": head["rhs_evals"]}
    for name in out:
        res[name] = rate(name)
    if "tuned" in out:
        res["tuned"]["candidates_yx"] = out["tuned"]["autotuned"]
    # every form's states equal the first form's bit for bit (the tuned
    # form's strided passes differ in rounding); the GW spectra agree to the
    # rounding of their atomic bin sums
    ref = out[forms[0]]
    res["states_equal"] = {
        name: all(np.array_equal(r["sizes"][n]["states"], ref["sizes"][n]["states"])
                  for n in ref["sizes"]) for name, r in out.items() if name != "tuned"}
    res["spectra_match"] = all(
        np.allclose(r["sizes"][n]["gw_mean"], ref["sizes"][n]["gw_mean"],
                    rtol=1e-12, atol=0)
        for name, r in out.items() if name != "tuned" for n in ref["sizes"])
    if "tuned" in out:
        t = out["tuned"]
        res["tuned"]["gw_mean_rel_diff_max"] = max(
            float(np.max(np.abs(np.array(t["sizes"][n]["gw_mean"]) -
                                np.array(ref["sizes"][n]["gw_mean"])) /
                         np.maximum(np.abs(ref["sizes"][n]["gw_mean"]), 1e-300)))
            for n in ref["sizes"])
    res["gw_spectrum"] = {str(n): {"nreal": v["nreal"], "t": v["t"],
                                   "mean": v["gw_mean"], "var": v["gw_var"]}
                          for n, v in head["sizes"].items()}
    return res


def mlp_bench(n=419611, epochs=3, batch=1024):
    """The learned proposer's kernels (ihg_mlp_*): Adam training epochs over
    n synthetic samples of the training set's size and shape, the gradient
    of the whole set, and inference of one sample (the per-step proposal)
    and of the whole set; FLOPs counted as 2 per multiply-add of the three
    layers (forward) and 3x that for a training sample (forward and
    backward)."""
    from instaheat_amd.proposer import Mlp
    rng = np.random.default_rng(1)
    X = rng.standard_normal((n, 20)).astype(np.float32)
    T = (0.1 * X[:, 0] - 0.05 * X[:, 3]).astype(np.float32)
    m = Mlp(seed=1)
    macs = 20 * 64 + 64 * 64 + 64
    m.train(X[:4096], T[:4096], 1, batch=batch)  # warm-up
    t0 = time.perf_counter()
    hist = m.train(X, T, epochs, batch=batch)
    train_s = time.perf_counter() - t0
    m.grad(X[:4096], T[:4096])
    t0 = time.perf_counter()
    for _ in range(3):
        m.grad(X, T)
    grad_s = (time.perf_counter() - t0) / 3
    one = X[:1]
    for _ in range(10):
        m.forward(one)
    t0 = time.perf_counter()
    for _ in range(200):
        m.forward(one)
    fwd1_us = (time.perf_counter() - t0) / 200 * 1e6
    t0 = time.perf_counter()
    for _ in range(3):
        m.forward(X)
    fwd_s = (time.perf_counter() - t0) / 3
    m.close()
    sps = n * epochs / train_s
    return {"metric": "step-proposer kernels (20-64-64-1 tanh MLP, FP32)",
            "samples": n, "epochs": epochs, "batch": batch,
            "train_samples_per_s": round(sps, 1),
            "train_s_per_epoch": round(train_s / epochs, 4),
            "train_GFLOP_per_s": round(sps * 6 * macs / 1e9, 2),
            "grad_full_set_ms": round(grad_s * 1e3, 3),
            "forward_one_sample_us": round(fwd1_us, 2),
            "forward_full_set_ms": round(fwd_s * 1e3, 3),
            "epoch_loss": [float(v) for v in hist],
            "note": "includes the host-device copies of each call (host "
                    "arrays at the C ABI)"}


def measured_traffic(kernel, grid):
    """HBM bytes per launch of the dominant kernel from the committed counter
    summary (two rocprofv3 --pmc passes scaled by a copy kernel), or None."""
    path = os.path.join(HERE, "profiles", "hbm_traffic_256cube_pmc.json")
    if grid != 256 or not os.path.exists(path):
        return None
    with open(path) as fh:
        t = json.load(fh)
    return t["traffic_bytes_per_launch"] if t.get("kernel") == "k_" + kernel else None


def roofline(prof, grid=256, copy_gbs=None):
    fams = {k: v for k, v in prof.items() if v[2] > 0}
    dom = max(fams, key=lambda k: fams[k][0])
    ms, by, cnt = fams[dom]
    avg_s = ms / cnt / 1e3
    per_launch = by / cnt
    ach = per_launch / avg_s / 1e9
    fft_ms = sum(prof[k][0] for k in FFT_FAMILIES if k in prof)
    fft_by = sum(prof[k][1] for k in FFT_FAMILIES if k in prof)
    pl_ms = sum(prof[k][0] for k in FFT_PLAIN if k in prof)
    pl_by = sum(prof[k][1] for k in FFT_PLAIN if k in prof)
    kernels = {k: {"ms": round(v[0], 3), "launches": v[2],
                   "avg_us": round(1e3 * v[0] / v[2], 2),
                   "GB_per_s": round(v[1] / (v[0] / 1e3) / 1e9, 1)}
               for k, v in fams.items()}
    return {"bound": "hbm", "kernel": dom, "achieved": round(ach, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(dom, grid),
            "bytes_per_launch": per_launch, "avg_launch_us": round(avg_s * 1e6, 2),
            # the same kernel against a plain copy kernel of one state array
            # on this GPU (the measured ceiling beside the 8 TB/s spec)
            "copy_ceiling_GB_per_s": None if copy_gbs is None else round(copy_gbs, 1),
            "frac_of_copy_ceiling": None if copy_gbs is None else round(ach / copy_gbs, 4),
            "fft_all_passes_GB_per_s": round(fft_by / (fft_ms / 1e3) / 1e9, 1),
            "fft_all_passes_frac": round(fft_by / (fft_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
            "fft_plain_passes_GB_per_s": round(pl_by / (pl_ms / 1e3) / 1e9, 1),
            "fft_plain_passes_frac": round(pl_by / (pl_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
            "kernels": kernels}


def _fam(prof, name):
    """(avg microseconds per launch, GB/s, fraction of the HBM peak) of one
    live-profiler family"""
    ms, by, cnt = prof[name]
    if cnt == 0:
        return None
    gbs = by / (ms / 1e3) / 1e9
    return {"launches": int(cnt), "avg_us": round(1e3 * ms / cnt, 2),
            "GB_per_s": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}


def next_rows(args, wl):
    """The follow-on paths at the workload's grid, each timed on the GPU
    from the warm-up state of the headline run:
      initial_data  ih_init_gaussian + mk_initial_psi (device) against
                    numpy's restatement of the same recipe on the host;
      dense_output  DOPRI853 with one time-based output inside every step
                    against the same forced steps without: the extra cost
                    per output, and k_dp_dense's HBM rate;
      hdf5          save() into libinstaheat_h5 (default outputs, then full
                    fields at stride 2) at the dense outputs of the same
                    steps, against those steps without the file;
      gsl_rk8pd     run_gsl_stepper with RK8PD (rtol 1e-10, atol 1e-14)
                    against DOPRI853 over the same span from the same
                    state, and the GSL kernels' HBM rates."""
    import shutil
    import tempfile
    import instaheat_amd as ih
    from instaheat_amd import h5 as ih5
    from oracle import ic
    n = wl["shape"][0]
    L = ih.lib()
    out = {"metric": f"follow-on paths at {n}^3 FP64 (GW on, filter on)"}
    sim = make_sim(ih, wl, 10 ** 9)
    # --- device initial data vs the host restatement ----------------------
    reps = 5
    L.ihg_sync(sim.ctx)
    t0 = time.perf_counter()
    for _ in range(reps):
        sim.init_gaussian(seed=wl["seed"], amplitude=wl["amplitude"],
                          phi0=wl["phi0"], dphi0=wl["dphi0"], k_cut=wl["kcut"])
        sim.mk_initial_psi()
    ih._check(L.ihg_sync(sim.ctx))
    gpu_s = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    ic.gaussian_state(wl["shape"], wl["bounds"], wl["mass"], wl["seed"],
                      wl["amplitude"], wl["phi0"], wl["dphi0"], 1.0, wl["kcut"],
                      sim.Ntot)
    host_s = time.perf_counter() - t0
    out["initial_data"] = {"gpu_ms": round(1e3 * gpu_s, 3),
                           "host_numpy_ms_phi_dphi_only": round(1e3 * host_s, 1),
                           "note": "GPU: Gaussian phi, dphi and psi, dpsi from "
                                   "the constraints; host: phi, dphi only"}
    print(f"initial data: {out['initial_data']}", file=sys.stderr, flush=True)
    # --- the headline run's pre-roll to t = PRE_ROLL_T and its W warm-up
    # steps: the error-controlled phase ------------------------------------
    sim.set_tf(PRE_ROLL_T)
    w = sim.run_dopri853()
    sim.set_tf(1e9)
    sim.set_dt(w["dt_next"])
    if args.warmup > 0:
        sim.set_max_steps(args.warmup - 1)
        w = sim.run_dopri853()
    print(f"next rows: pre-roll and warm-up to t = {w['t']:.4f}",
          file=sys.stderr, flush=True)
    snap = sim.snapshot()
    nsteps = 8
    sim.set_dt(w["dt_next"])
    sim.set_max_steps(nsteps - 1)
    r = sim.run_dopri853(log_cap=100)
    dts = r["log"][r["log"][:, 7] > 0.5, 1][:nsteps]
    t_w = snap["t"]

    def forced_run(**kw):
        sim.restore(snap)
        L.ihg_sync(sim.ctx)
        t0 = time.perf_counter()
        sim.run_dopri853(forced=dts, **kw)
        ih._check(L.ihg_sync(sim.ctx))
        return time.perf_counter() - t0
    # --- dense output: one output time inside every step --------------------
    ends = t_w + np.cumsum(dts)
    t_out = list(ends - 0.5 * dts)
    base = min(forced_run() for _ in range(2))
    ih.prof_enable(sim.ctx, True)
    ih.prof_reset(sim.ctx)
    dense = forced_run(t_out=t_out)
    prof = ih.prof_read(sim.ctx)
    ih.prof_enable(sim.ctx, False)
    dense = min(dense, forced_run(t_out=t_out))
    out["dense_output"] = {"steps": int(nsteps), "outputs": len(t_out),
                           "ms_per_step_without": round(1e3 * base / nsteps, 2),
                           "extra_ms_per_output": round(1e3 * (dense - base) / len(t_out), 2),
                           "k_dp_dense": _fam(prof, "dp_dense"),
                           "extra_stage_inputs": _fam(prof, "dp_dense_stage")}
    print(f"dense output: {out['dense_output']}", file=sys.stderr, flush=True)
    # --- HDF5 saves ---------------------------------------------------------
    tmp = tempfile.mkdtemp(prefix="ih_h5_")
    try:
        hd = {}
        for tag, outputs, stride in (("default_outputs", ih5.OUT_DEFAULT, (1, 1, 1)),
                                     ("full_fields_stride2", ih5.OUT_DEFAULT | ih5.OUT_FIELDS,
                                      (2, 2, 2))):
            path = os.path.join(tmp, tag + ".h5")
            with ih5.H5Output(path, sim.cfg, outputs=outputs, buf_size=1,
                              stride=stride, sim=sim) as wr:
                secs = forced_run(t_out=t_out, h5=wr)
                slices = wr.slices
            plain = forced_run(t_out=t_out, saves=[])
            size = os.path.getsize(path)
            os.remove(path)
            hd[tag] = {"saves": int(slices), "file_MB": round(size / 1e6, 1),
                       "ms_per_save": round(1e3 * (secs - plain) / slices, 2)}
        out["hdf5"] = hd
        print(f"hdf5: {hd}", file=sys.stderr, flush=True)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    # --- GSL RK8PD against DOPRI853 over the same span ---------------------
    span = float(np.sum(dts))
    sim.restore(snap)
    y_w = sim.field()
    g = ih.Sim(wl["shape"], bounds=wl["bounds"], mass=wl["mass"], gw=wl["gw"],
               filt=wl["filt"], bins=wl["bins"], method=ih.IH_DOPRI89, ti=0.0,
               tf=t_w + span, dt=float(dts[0]), dt_min=1e-9, rtol=1e-10,
               atol=1e-14)
    g.set_field(y_w)
    g.t = t_w
    ih.prof_enable(g.ctx, True)
    ih.prof_reset(g.ctx)
    st = g.run_gsl(1)
    gprof = ih.prof_read(g.ctx)
    ih.prof_enable(g.ctx, False)
    g.set_field(y_w)
    g.t = t_w
    st = g.run_gsl(1)
    g.close()
    dp_s = forced_run()
    out["gsl_rk8pd"] = {
        "span": round(span, 4), "t_start": round(t_w, 4),
        "steps": st["steps"], "attempts": st["attempts"], "rejected": st["failed_steps"],
        "status": st["status"],
        "sim_time_per_wall_s": round(span / st["integration_seconds"], 4),
        "ms_per_step": round(1e3 * st["integration_seconds"] / max(st["steps"], 1), 2),
        "dopri853_same_span_sim_time_per_wall_s": round(span / dp_s, 4),
        "dopri853_steps": int(nsteps),
        "k_gsl_stage": _fam(gprof, "gsl_stage"),
        "k_gsl_finish": _fam(gprof, "gsl_finish")}
    print(f"gsl: {out['gsl_rk8pd']}", file=sys.stderr, flush=True)
    sim.drop(snap)
    sim.close()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = next_rows_cpu(args)
    return out


def next_rows_cpu(args, n=128):
    """The oracle beside the GPU for two follow-on rows, on an n^3 sample
    of the workload (its Python restatements stream full state vectors
    with numpy, too slowly at 256^3 for a bounded sample), the GPU timed
    on the same sample from the same state and step size:
      dense_output  the extra cost of one output inside a DOPRI853 step:
                    three extra stages, the interpolant, the output call;
      gsl_rk8pd     one RK8PD step: 13 RHS, the stage sums, the error
                    ratio."""
    from scipy.integrate._ivp import dop853_coefficients as SC
    import instaheat_amd as ih
    from oracle import dense, gsl
    from oracle import oracle as orc
    wl = workload(n)
    L = ih.lib()
    threads = args.cpu_threads or len(os.sched_getaffinity(0))
    sim = make_sim(ih, wl, 10 ** 9)
    sim.set_max_steps(99)
    w = sim.run_dopri853()
    h = float(w["dt_next"])
    snap = sim.snapshot()
    y, t = sim.field(), sim.t

    def gpu_dense(t_out):
        sim.restore(snap)
        L.ihg_sync(sim.ctx)
        t0 = time.perf_counter()
        sim.run_dopri853(forced=[h], t_out=t_out)
        ih._check(L.ihg_sync(sim.ctx))
        return time.perf_counter() - t0
    gd = min(gpu_dense([t + 0.5 * h]) - gpu_dense(None) for _ in range(3))
    g = ih.Sim(wl["shape"], bounds=wl["bounds"], mass=wl["mass"], gw=wl["gw"],
               filt=wl["filt"], bins=wl["bins"], method=ih.IH_DOPRI89, ti=0.0,
               tf=t + h, dt=h, dt_min=1e-9, rtol=1e-10, atol=1e-14)
    gg = []
    for _ in range(3):
        g.set_field(y)
        g.t = t
        st = g.run_gsl(1)
        gg.append(st["integration_seconds"] / max(st["attempts"], 1))
    g.close()
    sim.drop(snap)
    sim.close()
    o = orc.Oracle(n, n, n, bounds=wl["bounds"], mass=wl["mass"], gw=wl["gw"],
                   filt=wl["filt"], bins=wl["bins"], threads=threads)

    def rhs(tt, yy):
        f = yy.copy()
        return o.mk_rhs(f, t=tt, filter=0, output=0, field=f)
    K, y_new = dense.step(rhs, t, y, rhs(t, y), h, SC.A, SC.B, SC.C)
    K.append(rhs(t + h, y_new))
    t0 = time.perf_counter()
    for i in range(3):
        st_ = 13 + i
        K.append(rhs(t + SC.C[st_] * h, dense.extra_stage_input(y, K, h, SC.A[st_][:st_])))
    yd = dense.interpolate(y, y_new, K, h, 0.5, SC.D)
    o.mk_rhs(yd, t=t + 0.5 * h, filter=0, output=1, field=yd)
    o.obs()
    cd = time.perf_counter() - t0
    del K
    tab = gsl.tableau(gsl.DOPRI89)
    t0 = time.perf_counter()
    y8, e8, _ = gsl.step(rhs, t, y, rhs(t, y), h, tab)
    gsl.rmax(y8, e8, 1e-14, 1e-10)
    cg = time.perf_counter() - t0
    res = {"sample": (f"{n}^3 state after 100 DOPRI853 steps (t = {t:.3f}), step "
                      f"{h:.4g}; oracle (C99+OpenMP RHS, numpy stage sums) on "
                      f"{threads} host threads"),
           "cores": threads, "kind": "port",
           "dense_output_extra_s": {"cpu": round(cd, 3), "gpu": round(gd, 5),
                                    "gpu_over_cpu": round(cd / gd, 1)},
           "gsl_rk8pd_step_s": {"cpu": round(cg, 3), "gpu": round(min(gg), 5),
                                "gpu_over_cpu": round(cg / min(gg), 1)}}
    print(f"cpu: {res}", file=sys.stderr, flush=True)
    return res


ORACLE_STEPS = 2


def cpu_baseline(args, wl, check, dt_mean):
    """The oracle (C99 + OpenMP restatement of the reference) on the host
    cores: ORACLE_STEPS forced DOPRI853 steps from the warm-up state (its
    initial filtered output call, then 11 stages + the FSAL call per step,
    stage arithmetic included), the same steps the GPU took.  Its seconds per
    step give simulated time per wall second at the plain form's mean
    accepted step; its end state is the GPU's parity check at full size."""
    from oracle import oracle as orc
    threads = args.cpu_threads or len(os.sched_getaffinity(0))
    n = wl["shape"][0]
    o = orc.Oracle(n, n, n, bounds=wl["bounds"], mass=wl["mass"], gw=wl["gw"],
                   filt=wl["filt"], bins=wl["bins"], threads=threads)
    y = check["y"].copy()
    t0 = time.perf_counter()
    o.run_dopri853(y, np.zeros_like(y), check["t"], 1e9, check["dts"][0],
                   rtol=1e-10, atol=1e-14, hubble_frac=1e-2, skip=10**9,
                   forced=check["dts"])
    secs = time.perf_counter() - t0
    step_s = secs / len(check["dts"])
    err = float(np.linalg.norm(check["y_end"] - y) / np.linalg.norm(y))
    return {"value": round(dt_mean / step_s, 6), "unit": "sim-time/wall-s",
            "cores": threads, "kind": "port",
            "sample": (f"oracle (C99+OpenMP) on {threads} host threads: "
                       f"{len(check['dts'])} forced DOPRI853 steps of dt "
                       f"{check['dts'][0]:.4g} from the {n}^3 warm-up state at "
                       f"t = {check['t']:.3f} ({secs:.1f} s, incl. the initial "
                       f"output call), scaled by the GPU plain form's mean dt "
                       f"{dt_mean:.4g}; this is the project's own restatement "
                       f"(radix-2 FFTs), not the reference's FFTW build, so "
                       f"the GPU/CPU ratio overstates the gain over the "
                       f"reference"),
            "seconds_per_step": round(step_s, 3),
            "gpu_end_state_rel_err_vs_oracle": err}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU path is not built", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe > 0:
        print(json.dumps(copy_probe(args.copy_probe)))
        return
    if args.ensemble:
        print(json.dumps(ensemble_bench(args.steps,
                                        args.ensemble_forms.split(","))))
        return
    if args.mlp:
        print(json.dumps(mlp_bench()))
        return
    wl = workload(args.grid)
    if args.next_rows:
        print(json.dumps(next_rows(args, wl)))
        return
    if args.long_run > 0:
        print(json.dumps(long_run(args, wl, args.long_run)))
        return
    gpu = run_gpu(args, wl)
    tuned, plain = gpu["forms"]["tuned"], gpu["forms"]["plain"]

    def summary(f):
        return {"value": round(f["value"], 6), "ms_per_step": round(f["ms_per_step"], 3),
                "steps_ok": f["n_ok"], "steps_bad": f["n_bad"],
                "rhs_evals": f["calls_rhs"],
                "rhs_evals_per_s": round(f["calls_rhs"] / f["secs"], 2),
                "rhs_evals_per_sim_time": round(f["rhs_per_sim_time"], 3),
                "mean_dt": f["dt"], "t_start": f["t_start"], "t_end": f["t_end"],
                "host_wall_s": round(f["wall"], 4)}
    rec = {
        "metric": METRIC,
        "value": round(tuned["value"], 6),
        "unit": "sim-time/wall-s",
        "n_gpus": 1,
        "steps": tuned["steps"],
        "warmup": args.warmup,
        "ms_per_step": round(tuned["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded Gaussian inflaton fluctuations (no dataset)",
        "config": {
            "workload": (f"{args.grid}^3 FP64 DOPRI853 rtol 1e-10 atol 1e-14, GW on, "
                         f"filter on, k_min 0.05, m 1, amplitude 1e-5, "
                         f"TIME_STEP_SKIPS {args.skip}; untimed pre-roll of "
                         f"{gpu['pre_steps']} steps from t = 0 to "
                         f"{PRE_ROLL_T:g} (error-controlled phase), "
                         f"{args.warmup} warm-up steps to t = "
                         f"{gpu['warm_t']:.2f}, then {HISTORY_STEPS} untimed + "
                         f"{args.steps} timed steps per form"),
            "grid": args.grid, "form": "tuned (learned step proposer + success())",
            "tuned": dict(summary(tuned), proposals=tuned["proposals"]),
            "plain": summary(plain),
            "tuned_over_plain": round(tuned["value"] / plain["value"], 4),
            "precision": dict(
                summary(gpu["forms"]["precision"]),
                form="FP32 GW source + spectrum over the plain form's steps",
                gw_spectrum_rel_err_max=max(gpu["forms"]["precision"]["gw_rel_err"]),
                gw_spectrum_rel_err_median=float(np.median(gpu["forms"]["precision"]["gw_rel_err"])),
                gw_bins_compared=len(gpu["forms"]["precision"]["gw_rel_err"])),
        },
        "roofline": roofline(gpu["prof"], args.grid, gpu["copy_gbs"]),
    }
    if not args.no_long_run:
        rec["config"]["long_run"] = long_run(args, wl, LONG_RUN_T)
    if not args.no_cpu_baseline:
        rec["cpu_baseline"] = cpu_baseline(args, wl, gpu["check"], plain["dt"])
    else:
        rec["cpu_baseline"] = None
    print(json.dumps(rec))


if __name__ == "__main__":
    main()
