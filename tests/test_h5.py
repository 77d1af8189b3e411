"""CPU checks of the HDF5 output file (include/instaheat_h5.h), the
reference's src/io.c format: dataset names, shapes, chunking, layouts and
string types, the buffered writes across buffer boundaries, the strided
full-field copy (against numpy slicing), the parameter datasets, and the
error paths.  No GPU: the file is written from host arrays with
ih_h5_append.  h5dump (HDF5's own tool) reads the file independently when
it is installed."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from tests.test_abi import _declared

pytestmark = pytest.mark.skipif(
    not os.path.exists(os.path.join(os.path.dirname(os.path.dirname(
        os.path.abspath(__file__))), "instaheat_amd", "libinstaheat_h5.so")),
    reason="libinstaheat_h5.so not built (no HDF5 at build time)")

H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump"
                                    if os.path.exists("/opt/conda/bin/h5dump") else None)


@pytest.fixture(scope="module")
def h5():
    from instaheat_amd import h5 as mod
    mod.h5_lib()
    return mod


def _cfg(shape=(8, 6, 16), bins=6, gw=True, filt=True):
    import instaheat_amd as ih
    cfg = ih.default_config()
    cfg.grid = ih.make_params(shape, ((0.0, 1.5), (-1.0, 2.0), (0.5, 4.5)), 2.5, gw,
                              filt, bins)
    cfg.rtol, cfg.atol = 3e-9, 4e-13
    cfg.skip = 7
    cfg.max_dt_hubble_fraction = 0.02
    return cfg


def test_header_declares_what_the_wrapper_binds(h5):
    import ctypes
    assert _declared("instaheat_h5.h") == sorted(h5.H5_SYMBOLS)
    L = ctypes.CDLL(h5.H5_LIB_PATH)
    for name in h5.H5_SYMBOLS:
        assert hasattr(L, name), name


def _slices(rng, n, shape, bins):
    N = int(np.prod(shape))
    out = []
    for i in range(n):
        out.append((0.1 * i + 0.05, 1.0 + 0.3 * i, rng.standard_normal(28 + 4 * bins),
                    rng.standard_normal(4 * N), rng.standard_normal(N)))
    return out


def test_series_values_shapes_and_strided_fields(h5, tmp_path):
    shape, bins = (8, 6, 16), 6
    stride = (2, 4, 3)
    cfg = _cfg(shape, bins)
    p = tmp_path / "run.h5"
    out = h5.OUT_DEFAULT | h5.OUT_FIELDS
    sl = _slices(np.random.default_rng(5), 7, shape, bins)
    with h5.H5Output(p, cfg, outputs=out, buf_size=3, stride=stride, seed=42,
                     commit_hash="abc1234") as w:
        for i, (t, a, obs, f, rho) in enumerate(sl):
            w.append(t, a, obs, f, rho)
            assert w.slices == i + 1
        # 6 slices written (two full buffers), 1 still buffered
        assert h5.dims(p, "time") == (6, 1)
    assert h5.dims(p, "time") == (7, 1)
    T = np.array([s[0] for s in sl])
    assert np.array_equal(h5.read(p, "time"), T)
    assert np.array_equal(h5.read(p, "a"), np.array([s[1] for s in sl]))
    obs = np.array([s[2] for s in sl])
    names = ["phi_summary", "dphi_summary", "psi_summary", "dpsi_summary",
             "rho_summary", "pressure_summary"]
    for j, n in enumerate(names):
        assert np.array_equal(h5.read(p, n), obs[:, 4 * j:4 * j + 4]), n
    assert np.array_equal(h5.read(p, "constraints"), obs[:, 24:28])
    for j, n in enumerate(["phi_power_spectrum", "psi_power_spectrum",
                           "rho_power_spectrum", "gravitational_wave_spectrum"]):
        assert np.array_equal(h5.read(p, n), obs[:, 28 + j * bins:28 + (j + 1) * bins]), n
    N = int(np.prod(shape))
    sx, sy, sz = stride
    for j, n in enumerate(["phi", "dphi", "psi", "dpsi"]):
        got = h5.read(p, n)
        want = np.array([s[3][j * N:(j + 1) * N].reshape(shape)[::sx, ::sy, ::sz].ravel()
                         for s in sl])
        assert got.shape == (7, 4 * 2 * 6)
        assert np.array_equal(got, want), n
    want = np.array([s[4].reshape(shape)[::sx, ::sy, ::sz].ravel() for s in sl])
    assert np.array_equal(h5.read(p, "rho"), want)
    # one slice by row
    assert np.array_equal(h5.read(p, "phi_summary", row=4), obs[4, :4])
    # parameters (src/io.c:122-201)
    assert h5.read(p, "mass")[0] == 2.5
    assert h5.read(p, "inflaton_mass")[0] == -1.0
    assert h5.read(p, "dimension")[0] == 3.0
    assert h5.read(p, "seed")[0] == 42.0
    assert h5.read(p, "strides_time")[0] == 7.0
    assert list(h5.read(p, "tolerances")) == [3e-9, 4e-13]
    assert list(h5.read(p, "gridpoints_internal")) == [8, 6, 16]
    assert list(h5.read(p, "gridpoints_output")) == [4, 2, 6]
    assert list(h5.read(p, "strides_space")) == [2, 4, 3]
    assert list(h5.read(p, "spatial_bounds_x")) == [0.0, 1.5]
    assert list(h5.read(p, "spatial_bounds_y")) == [-1.0, 2.0]
    assert list(h5.read(p, "spatial_bounds_z")) == [0.5, 4.5]
    assert h5.read(p, "max_dt_hubble_fraction")[0] == 0.02
    assert h5.read(p, "filter")[0] == 1.0
    assert h5.read(p, "gravitational_waves")[0] == 1.0
    assert h5.read_string(p, "method") == "hyperbolic"
    assert h5.read_string(p, "commit_hash") == "abc1234"


def test_defaults_and_gw_off(h5, tmp_path):
    cfg = _cfg(gw=False, filt=False)
    cfg.max_dt_hubble_fraction = 0.0
    p = tmp_path / "plain.h5"
    with h5.H5Output(p, cfg) as w:
        w.append(0.0, 1.0, np.arange(28 + 4 * 6, dtype=float))
    assert h5.read(p, "runtime_total")[0] >= 0.0
    for n in ("gravitational_wave_spectrum", "phi", "rho", "commit_hash", "calls_rhs"):
        with pytest.raises(h5.IhgError, match="no dataset"):
            h5.dims(p, n)
    assert h5.read(p, "gravitational_waves")[0] == 0.0
    assert h5.read(p, "filter")[0] == 0.0
    assert h5.read(p, "max_dt_hubble_fraction")[0] == -1.0
    assert h5.read(p, "phi_power_spectrum").shape == (1, 6)


@pytest.mark.skipif(H5DUMP is None, reason="h5dump not installed")
def test_layout_seen_by_h5dump(h5, tmp_path):
    """chunked, time-extensible series; compact parameters; Fortran
    (space padded) strings of 10 and 15 characters -- the creation calls of
    src/io.c:203-323"""
    shape, bins = (8, 6, 16), 6
    cfg = _cfg(shape, bins)
    p = tmp_path / "layout.h5"
    sl = _slices(np.random.default_rng(1), 2, shape, bins)
    with h5.H5Output(p, cfg, outputs=h5.OUT_DEFAULT | h5.OUT_PHI, buf_size=5,
                     commit_hash="deadbeef") as w:
        for t, a, obs, f, rho in sl:
            w.append(t, a, obs, f, rho)
        w.write_simple("steps_total", 12)
    txt = subprocess.run([H5DUMP, "-p", "-H", str(p)], capture_output=True, text=True,
                         check=True).stdout

    def block(name):
        m = re.search(r'DATASET "%s" \{(.*?)\n   \}' % re.escape(name), txt, re.S)
        assert m, name
        return m.group(1)
    b = block("phi_summary")
    assert "H5T_IEEE_F64LE" in b
    assert re.search(r"SIMPLE \{ \( 2, 4 \) / \( H5S_UNLIMITED, 4 \) \}", b)
    assert re.search(r"CHUNKED \( 5, 4 \)", b)
    b = block("phi")
    assert re.search(r"SIMPLE \{ \( 2, %d \) / \( H5S_UNLIMITED, %d \) \}" % (8 * 6 * 16, 8 * 6 * 16), b)
    b = block("time")
    assert re.search(r"SIMPLE \{ \( 2 \) / \( H5S_UNLIMITED \) \}", b)
    assert re.search(r"CHUNKED \( 5 \)", b)
    for n in ("mass", "tolerances", "steps_total"):
        assert "COMPACT" in block(n), n
    b = block("method")
    assert "STRSIZE 10" in b and "H5T_STR_SPACEPAD" in b
    b = block("commit_hash")
    assert "STRSIZE 15" in b and "H5T_STR_SPACEPAD" in b
    # values through h5dump's own reader
    d = subprocess.run([H5DUMP, "-d", "gridpoints_internal", str(p)], capture_output=True,
                       text=True, check=True).stdout
    assert re.search(r"\(0\): 8, 6, 16", d)


def test_errors_leave_the_file_usable(h5, tmp_path):
    shape, bins = (8, 6, 16), 6
    cfg = _cfg(shape, bins)
    p = tmp_path / "err.h5"
    with pytest.raises(h5.IhgError, match="buf_size"):
        h5.H5Output(p, cfg, buf_size=0)
    with pytest.raises(h5.IhgError, match="stride"):
        h5.H5Output(p, cfg, stride=(1, 0, 1))
    with pytest.raises(h5.IhgError, match="cannot create"):
        h5.H5Output(tmp_path / "no" / "such" / "dir.h5", cfg)
    w = h5.H5Output(p, cfg, outputs=h5.OUT_DEFAULT | h5.OUT_PSI)
    with pytest.raises(h5.IhgError, match="observables"):
        w.append(0.0, 1.0, np.zeros(5))
    with pytest.raises(h5.IhgError, match="without fields"):
        w.append(0.0, 1.0, np.zeros(28 + 4 * bins))
    with pytest.raises(h5.IhgError, match="followup needs"):
        w.write_followup()
    w.write_simple("steps_ok", [3.0])
    with pytest.raises(h5.IhgError, match="exist"):
        w.write_simple("steps_ok", [4.0])
    with pytest.raises(h5.IhgError, match="compact"):
        w.write_simple("too_big", np.zeros(10000))
    # a rejected append stores nothing; the next one lands as slice 0
    N = int(np.prod(shape))
    w.append(0.5, 1.1, np.ones(28 + 4 * bins), np.full(4 * N, 2.0))
    w.close()
    assert np.array_equal(h5.read(p, "time"), [0.5])
    assert np.array_equal(h5.read(p, "psi"), np.full((1, N), 2.0))
    assert h5.read(p, "steps_ok")[0] == 3.0
    with pytest.raises(h5.IhgError, match="row"):
        h5.read(p, "time", row=1)
    with pytest.raises(h5.IhgError, match="no dataset"):
        h5.read(p, "nothing")
    with pytest.raises(h5.IhgError, match="cannot open"):
        h5.read(tmp_path / "missing.h5", "time")
    with pytest.raises(h5.IhgError, match="not a string"):
        h5.read_string(p, "time")
    with pytest.raises(h5.IhgError, match="doubles"):
        h5.read(p, "method")


def test_dp_stats_datasets(h5, tmp_path):
    cfg = _cfg()
    p = tmp_path / "stats.h5"
    st = {"n_stp": 41, "n_ok": 38, "n_bad": 3, "attempts": 44, "t": 1.0,
          "dt_next": 0.01, "err_old": 1e-4, "integration_seconds": 2.5,
          "calls_rhs": 530}
    with h5.H5Output(p, cfg) as w:
        w.write_dp_stats(st)
    # main's last writes (src/main.c:99-106): runtime_total always,
    # calls_rhs from the stats here (no simulation bound)
    assert h5.read(p, "runtime_total")[0] > 0.0
    assert h5.read(p, "steps_total")[0] == 41
    assert h5.read(p, "steps_ok")[0] == 38
    assert h5.read(p, "steps_bad")[0] == 3
    assert h5.read(p, "runtime_stepper")[0] == 2.5
    assert h5.read(p, "calls_rhs")[0] == 530


C_HOST = r"""
#include <stdio.h>
#include "instaheat_h5.h"
int main(int argc, char **argv) {
    ih_config cfg; ih_config_defaults(&cfg);
    cfg.grid.nx = cfg.grid.ny = cfg.grid.nz = 16;
    cfg.grid.xa = cfg.grid.ya = cfg.grid.za = 0.0;
    cfg.grid.xb = cfg.grid.yb = cfg.grid.zb = 62.83185307179586;
    cfg.grid.mass = 1.0;
    cfg.tf = 0.5;
    cfg.skip = 5;
    ih_sim *sim = NULL;
    if (ih_create(&cfg, &sim)) { fputs(ihg_last_error(), stderr); return 1; }
    if (ih_init_gaussian(sim, 1, 1e-4, 1.0093, -0.7137, 0.4) || ih_mk_initial_psi(sim))
        return 1;
    ih_h5_opts o;
    ih_h5_opts_defaults(&o);
    o.buf_size = 5;
    o.seed = 1;
    ih_h5 *h5 = NULL;
    if (ih_h5_create(argc > 1 ? argv[1] : "run.h5", &cfg, &o, &h5)) {
        fputs(ih_h5_last_error(), stderr);
        return 1;
    }
    ih_h5_bind(h5, sim);
    ih_dp_stats st;
    long n = ih_run_dopri853(sim, NULL, ih_h5_save, h5, &st);
    if (n >= 0) ih_h5_write_dp_stats(h5, &st);
    ih_h5_write_followup(h5);
    int rc = ih_h5_close(h5);
    ih_destroy(sim);
    return (n < 0 || rc) ? 1 : 0;
}
"""


def build_c_host(tmp_path):
    """compile the C host of the integration guide: DOPRI853 with the HDF5
    writer as the save() hook; returns (returncode, stderr, binary)"""
    import instaheat_amd as ih
    src = tmp_path / "host.c"
    src.write_text(C_HOST)
    libdir = os.path.dirname(ih.HOST_LIB_PATH)
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = tmp_path / "host"
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I",
                        os.path.join(root, "include"), str(src), "-L", libdir,
                        "-linstaheat_h5", "-linstaheat", "-lihg",
                        "-Wl,-rpath," + libdir, "-o", str(exe)],
                       capture_output=True, text=True)
    return r.returncode, r.stderr, exe


def test_c_host_with_hdf5_output_compiles(h5, tmp_path):
    """The C host with the HDF5 writer as the save() hook compiles and
    links (instaheat_h5.h needs no HDF5 headers)."""
    rc, err, _ = build_c_host(tmp_path)
    assert rc == 0, err
