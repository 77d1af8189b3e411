"""Benchmark of the InstaHeat hot path on one MI355X.

Workload: 256^3 FP64, DOPRI853 (rtol 1e-10, atol 1e-14, Hubble cap 1e-2),
GW tensor source on, spectral filter on, k_min = 0.05 (L = 125.66), m = 1,
phi0 = 1.0093, dphi0 = -0.7137, a0 = 1, seeded Gaussian fluctuations of
amplitude 1e-5 (|dphi_k|^2 = 1/(2 a^3 omega_k), cut at N k_min / 4), psi from
the Einstein constraint, observables every TIME_STEP_SKIPS = 10 steps.

A "step" is one accepted DOPRI853 step (perform_step).  W warm-up steps bring
the controller from DELTA_T to its Hubble-capped step size; the K timed
steps run from there.  value = simulated time advanced / wall seconds of the
K timed steps (the mon.integration span), with every array resident in HBM.

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
                "z_r2c_sources", "x_roundtrip")
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=14)
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--skip", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    return ap.parse_args()


def workload(n):
    kmin = 0.05
    L = 2 * np.pi / kmin
    return dict(shape=(n, n, n), bounds=((0.0, L),) * 3, mass=1.0, gw=True,
                filt=True, bins=64, kmin=kmin, kcut=n * kmin / 4,
                amplitude=1e-5, phi0=1.0093, dphi0=-0.7137, seed=1)


def run_gpu(args, wl):
    import instaheat_amd as ih
    sim = ih.Sim(wl["shape"], bounds=wl["bounds"], mass=wl["mass"], gw=wl["gw"],
                 filt=wl["filt"], bins=wl["bins"], method=ih.IH_DOPRI853,
                 ti=0.0, tf=1e9, dt=1e-7, rtol=1e-10, atol=1e-14,
                 hubble_frac=1e-2, skip=args.skip)
    sim.init_gaussian(seed=wl["seed"], amplitude=wl["amplitude"],
                      phi0=wl["phi0"], dphi0=wl["dphi0"], k_cut=wl["kcut"])
    sim.mk_initial_psi()
    y0 = sim.field() if not args.no_cpu_baseline else None
    # warm-up: W accepted steps (the loop takes max_steps + 1 steps)
    sim.set_max_steps(max(args.warmup - 1, 0))
    w = sim.run_dopri853()
    sim.set_dt(w["dt_next"])
    sim.set_max_steps(args.steps - 1)
    ih.prof_enable(sim.ctx, True)
    ih.prof_reset(sim.ctx)
    t0 = sim.t
    calls0 = ih.lib().ihg_calls_rhs(sim.ctx)
    wall0 = time.perf_counter()
    r = sim.run_dopri853()
    wall1 = time.perf_counter()
    prof = ih.prof_read(sim.ctx)
    ih.prof_enable(sim.ctx, False)
    calls = ih.lib().ihg_calls_rhs(sim.ctx) - calls0
    sim_time = sim.t - t0
    steps = r["n_ok"] + r["n_bad"]
    secs = r["integration_seconds"]
    out = dict(sim_time=sim_time, secs=secs, wall=wall1 - wall0, steps=steps,
               n_ok=r["n_ok"], n_bad=r["n_bad"], attempts=r["attempts"],
               calls_rhs=int(calls), prof=prof, t_end=sim.t,
               warm_t=t0, dt=sim_time / max(steps, 1), y0=y0)
    sim.close()
    return out


def roofline(prof):
    fams = {k: v for k, v in prof.items() if v[2] > 0}
    dom = max(fams, key=lambda k: fams[k][0])
    ms, by, cnt = fams[dom]
    avg_s = ms / cnt / 1e3
    per_launch = by / cnt
    ach = per_launch / avg_s / 1e9
    fft_ms = sum(prof[k][0] for k in FFT_FAMILIES if k in prof)
    fft_by = sum(prof[k][1] for k in FFT_FAMILIES if k in prof)
    kernels = {k: {"ms": round(v[0], 3), "launches": v[2],
                   "avg_us": round(1e3 * v[0] / v[2], 2),
                   "GB_per_s": round(v[1] / (v[0] / 1e3) / 1e9, 1)}
               for k, v in fams.items()}
    return {"bound": "hbm", "kernel": dom, "achieved": round(ach, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
            "bytes_per_launch": per_launch, "avg_launch_us": round(avg_s * 1e6, 2),
            "fft_all_passes_GB_per_s": round(fft_by / (fft_ms / 1e3) / 1e9, 1),
            "fft_all_passes_frac": round(fft_by / (fft_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
            "kernels": kernels}


def cpu_baseline(args, wl, gpu):
    """The oracle (C99 + OpenMP restatement of the reference) on the host
    cores: one plain mk_rhs and one filtered + output mk_rhs at the same
    256^3 state, scaled to a DOPRI853 step (11 plain RHS + 1 filtered FSAL
    call; stage arithmetic not counted), then to simulated time per wall
    second with the GPU run's mean accepted step size."""
    from oracle import oracle as orc
    threads = args.cpu_threads or len(os.sched_getaffinity(0))
    n = wl["shape"][0]
    y = gpu["y0"]
    o = orc.Oracle(n, n, n, bounds=wl["bounds"], mass=wl["mass"], gw=wl["gw"],
                   filt=wl["filt"], bins=wl["bins"], threads=threads)
    f = y.copy()
    t0 = time.perf_counter()
    o.mk_rhs(f, field=f)
    t1 = time.perf_counter()
    o.mk_rhs(f, filter=1, output=1, field=f)
    t2 = time.perf_counter()
    step_s = 11 * (t1 - t0) + (t2 - t1)
    return {"value": round(gpu["dt"] / step_s, 6), "unit": "sim-time/wall-s",
            "cores": threads, "kind": "port",
            "sample": (f"oracle (C99+OpenMP) on {threads} host threads: 1 plain "
                       f"+ 1 filter/output mk_rhs at {n}^3 GW on "
                       f"({t1 - t0:.2f} s, {t2 - t1:.2f} s), scaled to 11+1 RHS "
                       f"per step and the GPU run's mean dt {gpu['dt']:.4g}"),
            "seconds_per_step": round(step_s, 3)}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU path is not built", file=sys.stderr)
        sys.exit(2)
    wl = workload(args.grid)
    gpu = run_gpu(args, wl)
    value = gpu["sim_time"] / gpu["secs"]
    rec = {
        "metric": METRIC,
        "value": round(value, 6),
        "unit": "sim-time/wall-s",
        "n_gpus": 1,
        "steps": gpu["steps"],
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * gpu["secs"] / gpu["steps"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded Gaussian inflaton fluctuations (no dataset)",
        "config": {
            "workload": (f"{args.grid}^3 FP64 DOPRI853 rtol 1e-10 atol 1e-14, GW on, "
                         f"filter on, k_min 0.05, m 1, amplitude 1e-5, "
                         f"TIME_STEP_SKIPS {args.skip}, plain GPU form"),
            "grid": args.grid, "form": "plain",
            "rhs_evals": gpu["calls_rhs"],
            "rhs_evals_per_s": round(gpu["calls_rhs"] / gpu["secs"], 2),
            "steps_ok": gpu["n_ok"], "steps_bad": gpu["n_bad"],
            "attempts": gpu["attempts"], "mean_dt": gpu["dt"],
            "t_start": gpu["warm_t"], "t_end": gpu["t_end"],
            "host_wall_s": round(gpu["wall"], 4),
        },
        "roofline": roofline(gpu["prof"]),
    }
    if not args.no_cpu_baseline:
        rec["cpu_baseline"] = cpu_baseline(args, wl, gpu)
    else:
        rec["cpu_baseline"] = None
    print(json.dumps(rec))


if __name__ == "__main__":
    main()
