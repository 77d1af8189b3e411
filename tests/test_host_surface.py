"""CPU checks of the host surface around the steppers:

* ih_config_from_macros (include/instaheat_params.h): a C file that defines
  the macros configure writes into main.h from src/parameters.sh
  (src/main_template.h:116-162, switches as NAME / NOT_NAME) gets every
  ih_config field it names, field by field;
* the stiffness check's counters (src/dopri853.c:550-580) of the host equal
  the oracle's on random verdict sequences and stop on the 15th stiff
  verdict in a row;
* the compat layer exports prepare_and_save_timeslice, mean and variance
  (src/toolbox.h:29, 34-35).
No GPU call is made: the counters are host arithmetic, and the C file only
fills a struct.
"""
import ctypes
import json
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# main.h as configure writes it from src/parameters.sh's defaults: values
# substituted as they stand, switches "1" -> NAME, "0" -> NOT_NAME
PARAMETERS_SH = {
    "GRIDPOINTS_X": "(1024)", "GRIDPOINTS_Y": "(1)", "GRIDPOINTS_Z": "(1)",
    "SPATIAL_LOWER_BOUND_X": "(-5.0)", "SPATIAL_UPPER_BOUND_X": "(5.0)",
    "SPATIAL_LOWER_BOUND_Y": "(-5.0)", "SPATIAL_UPPER_BOUND_Y": "(5.0)",
    "SPATIAL_LOWER_BOUND_Z": "(-5.0)", "SPATIAL_UPPER_BOUND_Z": "(5.0)",
    "MASS": "(4.0)", "A_INITIAL": "(1.0)", "FINAL_TIME": "(1.0e0)",
    "POWER_SPECTRUM_BINS": "(50)", "TIME_STEP_SKIPS": "(1)",
    "RELATIVE_TOLERANCE": "(1.0e-10)", "ABSOLUTE_TOLERANCE": "(1.0e-14)",
    "MAX_DT_HUBBLE_FRACTION": "(1.0e-2)", "INTEGRATION_METHOD": "(DOPRI853)",
    "INITIAL_TIME": "(0.0)", "DELTA_T": "(1.0e-7)",
    "MINIMAL_DELTA_T": "(1.0e-9)", "MAX_STEPS": "(1e15)",
    "SMALLEST_SCALING": "(0.333)", "LARGEST_SCALING": "(3.0)",
    "BETA": "(0.0)", "SAFE": "(0.9)", "MAX_RUNTIME": "(172500)",
    "WRITE_OUT_BUFFER_NUMBER": "(5)", "SEED": "(93)",
}
SWITCHES = {"ENABLE_FFT_FILTER": 1, "ENABLE_GW": 1, "ENABLE_FOLLOWUP": 0,
            "ENABLE_STIFFNESSCHECK": 0}

FIELDS = [("nx", "grid.nx", "%d"), ("ny", "grid.ny", "%d"), ("nz", "grid.nz", "%d"),
          ("xa", "grid.xa", "%.17g"), ("xb", "grid.xb", "%.17g"),
          ("ya", "grid.ya", "%.17g"), ("yb", "grid.yb", "%.17g"),
          ("za", "grid.za", "%.17g"), ("zb", "grid.zb", "%.17g"),
          ("mass", "grid.mass", "%.17g"), ("enable_gw", "grid.enable_gw", "%d"),
          ("enable_filter", "grid.enable_filter", "%d"),
          ("bins", "grid.bins", "%d"),
          ("method", "integration_method", "%d"), ("ti", "ti", "%.17g"),
          ("tf", "tf", "%.17g"), ("dt", "dt", "%.17g"), ("dt_min", "dt_min", "%.17g"),
          ("max_steps", "max_steps", "%ld"), ("beta", "beta", "%.17g"),
          ("safe", "safe", "%.17g"), ("minscale", "minscale", "%.17g"),
          ("maxscale", "maxscale", "%.17g"), ("rtol", "rtol", "%.17g"),
          ("atol", "atol", "%.17g"),
          ("hubble", "max_dt_hubble_fraction", "%.17g"), ("skip", "skip", "%ld"),
          ("a_initial", "a_initial", "%.17g"),
          ("max_runtime", "max_runtime", "%.17g"),
          ("followup", "enable_followup", "%d"),
          ("stiffness", "enable_stiffness_check", "%d")]


@pytest.fixture(scope="module")
def ih():
    import instaheat_amd
    instaheat_amd.host_lib()
    return instaheat_amd


def _config_from(tmp_path, ih, macros, switches, with_main_names=True):
    lines = []
    if with_main_names:
        # src/main_template.h:51-55
        lines += ["#define RK4 (0)", "#define RKF45 (1)", "#define RKCK (2)",
                  "#define DOPRI89 (3)", "#define DOPRI853 (4)"]
    lines += [f"#define {k} {v}" for k, v in macros.items()]
    lines += [f"#define {k}" if v else f"#define NOT_{k}"
              for k, v in switches.items()]
    body = ", ".join(f'\\"{n}\\": {fmt}' for n, _, fmt in FIELDS)
    args = ", ".join(f"cfg.{f}" for _, f, _ in FIELDS)
    src = tmp_path / "params.c"
    src.write_text("\n".join(lines) + f"""
#include <stdio.h>
#include "instaheat_params.h"
int main(void) {{
    ih_config cfg;
    ih_config_from_macros(&cfg);
    printf("{{{body}}}\\n", {args});
    return 0;
}}
""")
    libdir = os.path.dirname(ih.HOST_LIB_PATH)
    exe = tmp_path / "params"
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I",
                        os.path.join(ROOT, "include"), str(src), "-L", libdir,
                        "-linstaheat", "-lihg", "-Wl,-rpath," + libdir,
                        "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout)


def _expected(macros, switches):
    v = {k: float(eval(s.strip("()").replace("DOPRI853", "4").replace("RK4", "0")))
         for k, s in macros.items()}
    return {"nx": v["GRIDPOINTS_X"], "ny": v["GRIDPOINTS_Y"], "nz": v["GRIDPOINTS_Z"],
            "xa": v["SPATIAL_LOWER_BOUND_X"], "xb": v["SPATIAL_UPPER_BOUND_X"],
            "ya": v["SPATIAL_LOWER_BOUND_Y"], "yb": v["SPATIAL_UPPER_BOUND_Y"],
            "za": v["SPATIAL_LOWER_BOUND_Z"], "zb": v["SPATIAL_UPPER_BOUND_Z"],
            "mass": v["MASS"], "enable_gw": switches["ENABLE_GW"],
            "enable_filter": switches["ENABLE_FFT_FILTER"],
            "bins": v["POWER_SPECTRUM_BINS"], "method": v["INTEGRATION_METHOD"],
            "ti": v["INITIAL_TIME"], "tf": v["FINAL_TIME"], "dt": v["DELTA_T"],
            "dt_min": v["MINIMAL_DELTA_T"], "max_steps": v["MAX_STEPS"],
            "beta": v["BETA"], "safe": v["SAFE"], "minscale": v["SMALLEST_SCALING"],
            "maxscale": v["LARGEST_SCALING"], "rtol": v["RELATIVE_TOLERANCE"],
            "atol": v["ABSOLUTE_TOLERANCE"], "hubble": v["MAX_DT_HUBBLE_FRACTION"],
            "skip": v["TIME_STEP_SKIPS"], "a_initial": v["A_INITIAL"],
            "max_runtime": v["MAX_RUNTIME"],
            "followup": switches["ENABLE_FOLLOWUP"],
            "stiffness": switches["ENABLE_STIFFNESSCHECK"]}


def test_config_from_parameters_sh_defaults(tmp_path, ih):
    got = _config_from(tmp_path, ih, PARAMETERS_SH, SWITCHES)
    want = _expected(PARAMETERS_SH, SWITCHES)
    for k, v in want.items():
        assert got[k] == v, (k, got[k], v)


def test_config_from_another_parameter_set(tmp_path, ih):
    m = dict(PARAMETERS_SH)
    m.update({"GRIDPOINTS_X": "(64)", "GRIDPOINTS_Y": "(32)", "GRIDPOINTS_Z": "(16)",
              "SPATIAL_UPPER_BOUND_Y": "(62.83185307179586)", "MASS": "(1.0)",
              "INTEGRATION_METHOD": "(RK4)", "DELTA_T": "(0.01)",
              "TIME_STEP_SKIPS": "(50)", "MAX_RUNTIME": "(3600)",
              "FINAL_TIME": "(5.0)", "POWER_SPECTRUM_BINS": "(64)"})
    sw = {"ENABLE_FFT_FILTER": 0, "ENABLE_GW": 0, "ENABLE_FOLLOWUP": 1,
          "ENABLE_STIFFNESSCHECK": 1}
    got = _config_from(tmp_path, ih, m, sw, with_main_names=False)
    want = _expected(m, sw)
    for k, v in want.items():
        assert got[k] == v, (k, got[k], v)


def test_config_without_macros_is_the_defaults(tmp_path, ih):
    got = _config_from(tmp_path, ih, {}, {})
    cfg = ih.default_config()
    for name, field, _ in FIELDS:
        obj = cfg
        for part in field.split("."):
            obj = getattr(obj, part)
        assert got[name] == obj, name


def test_stiffness_counters_match_oracle(ih, oracle_mod):
    rng = np.random.default_rng(3)
    H = ih.host_lib()
    for trial in range(50):
        a, b = ih.IhStiffness(), oracle_mod.Stiffness()
        for step in range(200):
            dt = 10 ** rng.uniform(-3, 0)
            # verdicts around the threshold dt sqrt(num / den) = 6.1, and
            # den = 0 (dtlamb kept)
            den = 0.0 if rng.random() < 0.05 else 10 ** rng.uniform(-6, 0)
            num = den * (6.1 * 10 ** rng.normal(0, 0.3) / dt) ** 2
            r1 = H.ih_stiffness_update(ctypes.byref(a), dt, num, den)
            r2 = oracle_mod.lib().orc_stiffness_update(ctypes.byref(b), dt, num, den)
            assert (r1, a.stiff, a.nonstiff, a.dtlamb) == \
                (r2, b.stiff, b.nonstiff, b.dtlamb), (trial, step)
            if r1:
                break


def test_stiffness_stops_on_the_fifteenth_stiff_verdict(ih):
    H = ih.host_lib()
    st = ih.IhStiffness()
    # 14 stiff, 5 non-stiff (no reset), then the 15th stiff stops
    seq = [7.0] * 14 + [1.0] * 5 + [7.0]
    out = [H.ih_stiffness_update(ctypes.byref(st), 1.0, v * v, 1.0) for v in seq]
    assert out == [0] * 19 + [1]
    # 6 non-stiff verdicts in a row reset the count
    st = ih.IhStiffness()
    seq = [7.0] * 14 + [1.0] * 6 + [7.0] * 14
    out = [H.ih_stiffness_update(ctypes.byref(st), 1.0, v * v, 1.0) for v in seq]
    assert sum(out) == 0 and st.stiff == 14
    # den = 0 keeps the last dtlamb
    st = ih.IhStiffness()
    H.ih_stiffness_update(ctypes.byref(st), 2.0, 9.0, 1.0)
    assert st.dtlamb == 6.0
    H.ih_stiffness_update(ctypes.byref(st), 2.0, 0.0, 0.0)
    assert st.dtlamb == 6.0 and st.nonstiff == 2


def test_compat_layer_exports_the_remaining_toolbox_names(ih):
    H = ctypes.CDLL(ih.HOST_LIB_PATH)
    for name in ["prepare_and_save_timeslice", "mean", "variance",
                 "ih_compat_file_hooks"]:
        assert hasattr(H, name), name
