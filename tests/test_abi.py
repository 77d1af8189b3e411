"""CPU checks of the C ABI: both libraries load without a GPU, export every
function their headers declare, and reject bad arguments before touching the
device; the host controller success() / error() of the C99 host match the
oracle's on random inputs."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"\b(ihg_\w+|ih_\w+)\s*\(", txt)
    return sorted(set(n for n in names if not n.endswith("_fn")))


@pytest.fixture(scope="module")
def ih():
    import instaheat_amd
    instaheat_amd.host_lib()
    return instaheat_amd


def test_headers_declare_what_the_wrapper_binds(ih):
    assert _declared("ihg.h") == sorted(ih.IHG_SYMBOLS)
    host = [n for n in _declared("instaheat.h") if n not in ("ih_dt_proposer",)]
    assert host == sorted(ih.HOST_SYMBOLS)


def test_libraries_export_every_declared_symbol(ih):
    L = ctypes.CDLL(ih.LIB_PATH)
    H = ctypes.CDLL(ih.HOST_LIB_PATH)
    for name in _declared("ihg.h"):
        assert hasattr(L, name), name
    for name in ih.HOST_SYMBOLS:
        assert hasattr(H, name), name


@pytest.mark.parametrize("shape", [(12, 8, 8), (1024, 8, 8), (8, 8, 2), (8, 3, 8)])
def test_create_rejects_unsupported_grids(ih, shape):
    with pytest.raises(ih.IhgError, match="power of two"):
        ih.Gpu(shape)


def test_create_rejects_bad_bounds_and_bins(ih):
    with pytest.raises(ih.IhgError, match="bounds"):
        ih.Gpu((8, 8, 8), bounds=((1.0, 0.0), (0.0, 1.0), (0.0, 1.0)))
    with pytest.raises(ih.IhgError, match="BINS"):
        ih.Gpu((8, 8, 8), bins=0)


def test_null_context_is_an_error_not_a_crash(ih):
    L = ih.lib()
    assert L.ihg_mk_rhs(None, 0.0, None, None, 0) != 0
    assert L.ihg_sync(None) != 0
    assert L.ihg_dp853_stage(None, 2, 0.1, None, None, None) != 0
    assert b"NULL" in L.ihg_last_error()


def test_tableau_of_the_library_equals_the_oracle(ih, oracle_mod):
    t1 = ih.tableau()
    t2 = oracle_mod.tableau()
    for k in t1:
        assert np.array_equal(t1[k], t2[k]), k


def test_host_success_matches_oracle(ih, oracle_mod):
    H = ih.host_lib()
    rng = np.random.default_rng(0)
    cfg = ih.default_config()
    for _ in range(500):
        err = float(10 ** rng.uniform(-6, 1.5)) * (rng.random() > 0.05)
        dt = float(10 ** rng.uniform(-4, 0))
        err_old = float(10 ** rng.uniform(-4, 0))
        reject = int(rng.random() < 0.3)
        rho = float(10 ** rng.uniform(-3, 1))
        cfg.beta = float(rng.choice([0.0, 0.04]))
        cfg.max_dt_hubble_fraction = float(rng.choice([0.0, 1e-2]))
        ctl = ih.IhCtrl(err_old, 0.0, reject)
        d = ctypes.c_double(dt)
        ok = H.ih_success(ctypes.byref(ctl), ctypes.byref(cfg), err,
                          ctypes.byref(d), rho)
        ref = oracle_mod.success(err, dt, err_old=err_old, reject=reject,
                                 rho_mean=rho, beta=cfg.beta,
                                 hubble_frac=cfg.max_dt_hubble_fraction)
        assert (bool(ok), d.value, ctl.dt_next if ok else ref[2],
                ctl.err_old if ok else ref[3], ctl.reject) == \
            (ref[0], ref[1], ref[2], ref[3], ref[4])


def test_host_error_norm_formula(ih):
    """error() of src/dopri853.c:469-486 from the two sums."""
    H = ih.host_lib()
    e5, e3, dt, n = 2.5, 7.0, 0.01, 1000
    deno = e5 + 0.01 * e3
    assert H.ih_error_norm(e5, e3, dt, n) == pytest.approx(
        dt * e5 * np.sqrt(1.0 / (n * deno)), rel=1e-15)
    # deno <= 0 falls back to 1
    assert H.ih_error_norm(0.0, 0.0, dt, n) == 0.0


def test_compat_layer_exports_the_reference_names(ih):
    H = ctypes.CDLL(ih.HOST_LIB_PATH)
    for name in ["mk_rhs", "fft", "ifft", "run_rk4", "run_dopri853",
                 "run_gsl_stepper", "ih_compat_gsl_output_number",
                 "evo_flags", "ih_compat_bind", "ih_compat_save_hook",
                 "ih_compat_status"]:
        assert hasattr(H, name), name


def test_c_host_compiles_against_the_headers(tmp_path, ih):
    """A C99 host that replaces the reference's run_dopri853 call with the
    library (the compat layer's names) compiles and links."""
    import subprocess
    src = tmp_path / "host.c"
    src.write_text(r"""
#include <stdio.h>
#include "instaheat_compat.h"
int main(void) {
    ih_config cfg; ih_config_defaults(&cfg);
    cfg.grid.nx = cfg.grid.ny = cfg.grid.nz = 16;
    ih_sim *sim = NULL;
    if (ih_create(&cfg, &sim)) { fputs(ihg_last_error(), stderr); return 1; }
    ih_compat_bind(sim);
    evo_flags.filter = 1;
    run_dopri853();
    if (cfg.integration_method == IH_RKCK) {  /* never taken: links the name */
        ih_compat_gsl_output_number(1000);
        run_gsl_stepper();
    }
    ih_destroy(sim);
    return ih_compat_status() ? 1 : 0;
}
""")
    libdir = os.path.dirname(ih.HOST_LIB_PATH)
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I",
                        os.path.join(ROOT, "include"), str(src), "-L", libdir,
                        "-linstaheat", "-lihg", "-Wl,-rpath," + libdir,
                        "-o", str(tmp_path / "host")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
