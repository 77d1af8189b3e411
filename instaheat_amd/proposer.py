"""The learned DOPRI853 step proposer.

* Training data: the stepper's accept/reject logs (``Sim.run_dopri853(
  log_cap=...)``), one sample per accepted step that did not follow a
  rejection and has a successor attempt.  Inputs are the 20 raw features of
  ``ih_proposer_features`` (include/instaheat.h); the target is the
  hindsight step ratio log(dt* / dt_i), where dt* = dt_{i+1} (e_t /
  err_{i+1})^(1/8) is the step that would have met the error target e_t
  under the 8th-order error model err ~ dt^8.
* Training: Adam on the GPU through ihg_mlp (include/ihg.h), the network's
  own FP32 forward / gradient / update kernels.
* Use: ``MlpProposer`` hands ih_mlp_propose and its state to
  ``Sim.run_dopri853(proposer=...)``; success() still accepts or rejects
  every step and handles every rejection.

The network has no counterpart in the reference, whose controller is
success() (src/dopri853.c:505-540).

Command line::

    python -m instaheat_amd.proposer generate OUT.npz [--quick]
    python -m instaheat_amd.proposer train LOGS.npz [LOGS.npz ...] --out MODEL.json
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

import instaheat_amd as ih

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_MODEL = os.path.join(HERE, "proposer_model.json")
ERR_TARGET = 0.6
FIELDS = {name: i for i, name in enumerate(ih.RECORD_FIELDS)}


def raw_features(log, i):
    """ih_proposer_features of records log[max(0, i-7) .. i] (rows of the
    array Sim.run_dopri853 returns): the same formula, for building data."""
    x = np.empty(ih.MLP_IN)
    for j in range(ih.IH_HISTORY):
        k = max(i - ih.IH_HISTORY + 1 + j, 0)
        x[j] = np.log10(max(log[k, FIELDS["err"]], 1e-12))
        x[ih.IH_HISTORY + j] = np.log10(log[k, FIELDS["dt"]])
    r = log[i]
    x[16] = np.log10(r[FIELDS["a"]])
    x[17] = np.log10(np.sqrt(r[FIELDS["rho_mean"]] / 3.0))
    x[18] = np.log10(r[FIELDS["var_phi"]] + 1e-300)
    x[19] = np.log10(r[FIELDS["mean_psi2"]] + 1e-300)
    return x


def samples(log, err_target=ERR_TARGET):
    """(X, T) from one run's log."""
    acc = log[:, FIELDS["accepted"]] > 0.5
    X, T = [], []
    for i in range(len(log) - 1):
        if not acc[i] or (i > 0 and not acc[i - 1]):
            continue
        dt_i = log[i, FIELDS["dt"]]
        dt_n = log[i + 1, FIELDS["dt"]]
        err_n = max(log[i + 1, FIELDS["err"]], 1e-12)
        dt_star = dt_n * (err_target / err_n) ** 0.125
        X.append(raw_features(log, i))
        T.append(np.log(dt_star / dt_i))
    return np.array(X).reshape(-1, ih.MLP_IN), np.array(T)


class Mlp:
    """ihg_mlp: the network on the GPU."""

    def __init__(self, params=None, seed=0):
        L = ih.lib()
        self.n = L.ihg_mlp_nparams()
        h = ctypes.c_void_p()
        p = None
        if params is not None:
            self._p = np.ascontiguousarray(params, dtype=np.float32)
            assert self._p.size == self.n
            p = self._p.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        ih._check(L.ihg_mlp_create(ctypes.byref(h), p, ctypes.c_ulong(seed)))
        self.h = h

    def close(self):
        if self.h:
            ih.lib().ihg_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _f(a):
        return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))

    def params(self):
        out = np.empty(self.n, dtype=np.float32)
        ih._check(ih.lib().ihg_mlp_get_params(self.h, self._f(out)))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        ih._check(ih.lib().ihg_mlp_set_params(self.h, self._f(p)))

    def forward(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        y = np.empty(X.shape[0], dtype=np.float32)
        ih._check(ih.lib().ihg_mlp_forward(self.h, self._f(X), self._f(y),
                                           X.shape[0]))
        return y

    def grad(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        g = np.empty(self.n, dtype=np.float32)
        loss = ctypes.c_float()
        ih._check(ih.lib().ihg_mlp_grad(self.h, self._f(X), self._f(T),
                                        X.shape[0], self._f(g),
                                        ctypes.byref(loss)))
        return g, loss.value

    def train(self, X, T, epochs, batch=1024, lr=1e-3, seed=1):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        hist = np.zeros(max(epochs, 1), dtype=np.float32)
        ih._check(ih.lib().ihg_mlp_train(self.h, self._f(X), self._f(T),
                                         X.shape[0], epochs, batch,
                                         ctypes.c_float(lr), ctypes.c_ulong(seed),
                                         self._f(hist)))
        return hist[:epochs]


class MlpProposer:
    """A trained network plus its input / output scaling, usable as
    Sim.run_dopri853(proposer=...)."""

    def __init__(self, model):
        self.model = model
        self.mlp = Mlp(np.array(model["params"], dtype=np.float32))
        st = ih.IhMlpProposer()
        st.mlp = self.mlp.h
        for i in range(ih.MLP_IN):
            st.in_mean[i] = model["in_mean"][i]
            st.in_std[i] = model["in_std"][i]
        st.out_mean = model["out_mean"]
        st.out_std = model["out_std"]
        st.min_ratio = model["min_ratio"]
        st.max_ratio = model["max_ratio"]
        st.calls = 0
        self.state = st

    @property
    def calls(self):
        return self.state.calls

    def c_proposer(self):
        fn = ctypes.cast(ih.host_lib().ih_mlp_propose, ih.PROPOSER_FN)
        return fn, ctypes.cast(ctypes.byref(self.state), ctypes.c_void_p)

    @classmethod
    def load(cls, path=DEFAULT_MODEL):
        with open(path) as fh:
            return cls(json.load(fh))


def fit(X, T, epochs=150, batch=1024, lr=1e-3, seed=1, min_ratio=0.333,
        max_ratio=3.0):
    """Standardise, train on the GPU, return (model dict, epoch losses)."""
    in_mean = X.mean(0)
    in_std = X.std(0)
    in_std[in_std < 1e-6] = 1.0
    out_mean = float(T.mean())
    out_std = float(T.std()) if T.std() > 1e-9 else 1.0
    Xs = (X - in_mean) / in_std
    Ts = (T - out_mean) / out_std
    net = Mlp(seed=seed)
    hist = net.train(Xs, Ts, epochs, batch=batch, lr=lr, seed=seed)
    model = dict(params=[float(v) for v in net.params()],
                 in_mean=[float(v) for v in in_mean],
                 in_std=[float(v) for v in in_std],
                 out_mean=out_mean, out_std=out_std,
                 min_ratio=min_ratio, max_ratio=max_ratio,
                 err_target=ERR_TARGET, n_samples=int(X.shape[0]),
                 epochs=int(epochs), final_loss=float(hist[-1]) if len(hist) else None)
    net.close()
    return model, hist


# --------------------------------------------------------------------------
# training runs: the plain method at the small configurations (32^3 without
# and 64^3 with the GW source, k_min 0.1), several seeds, amplitudes and
# tolerances
# --------------------------------------------------------------------------

def run_specs(quick=False):
    specs = []
    seeds = [1] if quick else [1, 2, 3, 4]
    for n in (32, 64):
        for gw in (False, True):
            for kmin in (0.1, 0.05):
                for rtol in (1e-9, 1e-10, 1e-11):
                    for amp in (1e-6, 1e-5, 1e-4):
                        for seed in seeds:
                            specs.append(dict(n=n, gw=gw, rtol=rtol, amp=amp,
                                              seed=seed, kmin=kmin,
                                              tf=40.0 if quick else 150.0))
    return specs


def make_sim(n, gw, rtol, amp, seed, tf, kmin=0.1, max_steps=10**9):
    """The training / test configuration: the physics of the 32^3 and 64^3
    configurations (m = 1, phi0 = 1.0093, dphi0 = -0.7137, Hubble cap 1e-2)
    at the given grid, tolerance, amplitude and seed."""
    L = 2 * np.pi / kmin
    sim = ih.Sim((n, n, n), bounds=((0.0, L),) * 3, mass=1.0, gw=gw, filt=True,
                 bins=16, method=ih.IH_DOPRI853, ti=0.0, tf=tf, dt=1e-7,
                 rtol=rtol, atol=1e-14, hubble_frac=1e-2, skip=10**9,
                 max_steps=max_steps)
    sim.init_gaussian(seed=seed, amplitude=amp, phi0=1.0093, dphi0=-0.7137,
                      k_cut=n * kmin / 4)
    sim.mk_initial_psi()
    return sim


class JitterProposer:
    """The classic step times exp(U(lo, hi)): training runs whose step
    histories are not tied to one controller."""

    def __init__(self, lo=-0.25, hi=0.15, seed=0):
        self.rng = np.random.default_rng(seed)
        self.lo, self.hi = lo, hi

    def __call__(self, history, n, dt_classic):
        return dt_classic * float(np.exp(self.rng.uniform(self.lo, self.hi)))


def generate(out, quick=False, log_cap=200000, jitter=None, model=None,
             seeds=None):
    """Accept/reject logs of the training runs; each run's steps come from
    the classic controller, the classic one jittered (jitter = (lo, hi)),
    or a trained model (on-policy data)."""
    logs = {}
    specs = run_specs(quick)
    if seeds is not None:
        specs = [sp for sp in specs if sp["seed"] in seeds]
    prop = MlpProposer(model) if model is not None else None
    for k, sp in enumerate(specs):
        t0 = time.perf_counter()
        sim = make_sim(**sp)
        if jitter is not None:
            prop = JitterProposer(jitter[0], jitter[1], seed=k)
        r = sim.run_dopri853(log_cap=log_cap, proposer=prop)
        logs["run%03d" % k] = r["log"]
        logs["spec%03d" % k] = np.array([sp["n"], sp["gw"], sp["rtol"],
                                         sp["amp"], sp["seed"], sp["tf"],
                                         sp["kmin"]])
        print("run %d %s: %d attempts, %d rejected, t %.2f, %.1f s" %
              (k, sp, len(r["log"]), r["n_bad"], r["t"],
               time.perf_counter() - t0), flush=True)
        del sim
    np.savez_compressed(out, **logs)


def compare(model, n=32, gw=False, rtol=1e-10, amp=1e-5, seed=11, kmin=0.1,
            t_start=0.0, tf=150.0):
    """The plain controller and the learned proposer over the same interval
    from the same state: RHS evaluations, rejections, simulated time and the
    two end states."""
    out = {}
    prop = MlpProposer(model) if isinstance(model, dict) else model
    for form in ("plain", "tuned"):
        sim = make_sim(n, gw, rtol, amp, seed, tf, kmin=kmin)
        if t_start > 0.0:
            sim.set_tf(t_start)
            sim.run_dopri853()
            sim.set_tf(tf)
        calls0 = ih.lib().ihg_calls_rhs(sim.ctx)
        t0 = sim.t
        r = sim.run_dopri853(proposer=prop if form == "tuned" else None)
        out[form] = dict(calls=int(ih.lib().ihg_calls_rhs(sim.ctx) - calls0),
                         n_ok=int(r["n_ok"]), n_bad=int(r["n_bad"]),
                         sim_time=float(sim.t - t0),
                         seconds=float(r["integration_seconds"]),
                         state=sim.field())
    return out


def load_logs(paths):
    runs = []
    for p in paths:
        with np.load(p, allow_pickle=False) as z:
            for k in sorted(z.files):
                if k.startswith("run"):
                    runs.append(z[k])
    return runs


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m instaheat_amd.proposer")
    sub = ap.add_subparsers(dest="cmd", required=True)
    g = sub.add_parser("generate")
    g.add_argument("out")
    g.add_argument("--quick", action="store_true")
    g.add_argument("--jitter", type=float, nargs=2, default=None)
    g.add_argument("--model", default=None, help="on-policy runs of a model")
    g.add_argument("--seeds", type=int, nargs="*", default=None)
    t = sub.add_parser("train")
    t.add_argument("logs", nargs="+")
    t.add_argument("--out", default=DEFAULT_MODEL)
    t.add_argument("--epochs", type=int, default=150)
    t.add_argument("--lr", type=float, default=1e-3)
    t.add_argument("--margin", type=float, default=0.0,
                   help="subtracted from the predicted log step ratio")
    a = ap.parse_args(argv)
    if a.cmd == "generate":
        model = None
        if a.model:
            with open(a.model) as fh:
                model = json.load(fh)
        generate(a.out, quick=a.quick, jitter=a.jitter, model=model,
                 seeds=a.seeds)
    else:
        runs = load_logs(a.logs)
        XT = [samples(r) for r in runs]
        X = np.concatenate([x for x, _ in XT])
        T = np.concatenate([t for _, t in XT])
        print("samples", X.shape[0], flush=True)
        model, hist = fit(X, T, epochs=a.epochs, lr=a.lr)
        model["out_mean"] -= a.margin
        model["margin"] = a.margin
        print("loss first / last epoch %.4g / %.4g" % (hist[0], hist[-1]))
        with open(a.out, "w") as fh:
            json.dump(model, fh)
        print("wrote", a.out)


if __name__ == "__main__":
    sys.exit(main())
