"""GPU parity of one mk_rhs (src/toolbox.c:75-88) with every flag
combination, of the observables of output_all (src/toolbox.c:251-270) and of
mk_initial_psi (src/setup.c:825-903), against the CPU oracle on identical
seeded states."""
import zlib

import numpy as np
import pytest

from tests.states import BOX_CONFIG, blocks, make_state, rel_l2

pytestmark = pytest.mark.gpu

# Single-RHS gate: relative L2 over the whole result vector
TOL_RHS = 1e-12
# Per block the tolerance absorbs the cancellations in dpsi' (0.5 p and
# (psi - 1/2) <p> nearly cancel) and in the TT projection; an indexing or
# sign error shows up as O(1).
TOL_BLOCK = 1e-9
TOL_OBS = 1e-9

ANISO = ((0.0, 10.0), (-3.0, 5.0), (1.0, 4.0))

CASES = [
    # shape, bounds, mass, gw, flags, a
    ((16, 16, 16), BOX_CONFIG, 1.0, False, 0, 1.0),
    ((32, 32, 32), BOX_CONFIG, 1.0, False, 0, 1.0),
    ((16, 16, 16), BOX_CONFIG, 1.0, False, 1, 1.0),
    ((16, 16, 16), BOX_CONFIG, 1.0, False, 2, 1.0),
    ((16, 16, 16), BOX_CONFIG, 1.0, False, 3, 1.0),
    ((16, 16, 16), BOX_CONFIG, 1.0, True, 0, 1.0),
    ((16, 16, 16), BOX_CONFIG, 1.0, True, 3, 1.0),
    ((32, 16, 8), ANISO, 4.0, True, 3, 1.7),
    ((8, 16, 32), ANISO, 4.0, True, 0, 1.3),
    ((64, 32, 16), BOX_CONFIG, 1.0, True, 1, 1.0),
    ((32, 32, 32), BOX_CONFIG, 1.0, True, 2, 2.0),
    # the smallest and the longest lines of every axis, and a large grid
    ((4, 4, 4), BOX_CONFIG, 1.0, True, 3, 1.0),
    ((512, 16, 8), ANISO, 4.0, True, 1, 1.1),
    ((8, 512, 16), ANISO, 4.0, True, 3, 1.1),
    ((16, 8, 512), ANISO, 4.0, True, 0, 1.1),
    ((256, 64, 32), BOX_CONFIG, 1.0, True, 1, 1.4),
    ((128, 128, 128), BOX_CONFIG, 1.0, True, 3, 1.2),
    ((256, 256, 256), BOX_CONFIG, 1.0, True, 3, 1.0),
]


def _ids(c):
    shape, bounds, mass, gw, flags, a = c
    return "{}_{}_m{}_gw{}_f{}_a{}".format("x".join(map(str, shape)),
                                          "box" if bounds is BOX_CONFIG else "aniso",
                                          mass, int(gw), flags, a)


@pytest.fixture(scope="module")
def ih():
    import instaheat_amd
    instaheat_amd.lib()
    return instaheat_amd


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_single_rhs(ih, oracle_mod, case):
    shape, bounds, mass, gw, flags, a = case
    bins = 20
    # a seed fixed by the case (str hash() changes from process to process)
    y = make_state(shape, bounds, seed=zlib.crc32(_ids(case).encode()) % 1000,
                   gw=gw, a=a)
    o = oracle_mod.Oracle(*shape, bounds=bounds, mass=mass, gw=gw, filt=True,
                          bins=bins)
    fo = y.copy()
    ro = o.mk_rhs(fo, filter=flags & 1, output=(flags >> 1) & 1, field=fo)
    g = ih.Gpu(shape, bounds=bounds, mass=mass, gw=gw, filt=True, bins=bins)
    df = g.to_device(y)
    dr = g.alloc(g.Ntot)
    g.mk_rhs(df, dr, flags=flags)
    rg = dr.download()
    fg = df.download()
    B = blocks(shape)
    live = ["phi", "dphi", "psi", "dpsi", "a"] + (
        ["h1", "h2", "dh1", "dh2"] if gw else [])
    mask = np.zeros(o.Ntot, bool)
    for k in live:
        mask[B[k]] = True
    errs = {k: rel_l2(rg[B[k]], ro[B[k]]) for k in live}
    print("block errors", errs)
    assert rel_l2(rg[mask], ro[mask]) < TOL_RHS
    for k in live:
        assert errs[k] < TOL_BLOCK, k
    # the filter mutates f in place, as in the reference
    for k in ["phi", "dphi", "psi", "dpsi"]:
        assert rel_l2(fg[B[k]], fo[B[k]]) < TOL_RHS, k
    mg, mo = g.means(), o.means()
    assert abs(mg[0] - mo[0]) <= 1e-13 * abs(mo[0])
    assert abs(mg[1] - mo[1]) <= 1e-13 * abs(mo[1]) + 1e-16
    if flags & 2:
        og = ih.unpack_obs(g.observables(), bins)
        oo = oracle_mod.unpack_obs(o.obs(), bins)
        for k in oo:
            if k == "gw" and not gw:
                continue
            e = rel_l2(og[k], oo[k])
            print("obs", k, e)
            assert e < TOL_OBS, k
    assert g.calls_rhs == 1


def test_summary_reads_field_not_f(ih, oracle_mod):
    """mk_summary reads the global `field` (src/toolbox.c:638-647), so an
    output call on another array reports the summary of `field`."""
    shape = (16, 16, 16)
    bins = 10
    y0 = make_state(shape, seed=3)
    y1 = make_state(shape, seed=4)
    o = oracle_mod.Oracle(*shape, bounds=BOX_CONFIG, bins=bins)
    f1 = y1.copy()
    field_o = y0.copy()
    o.mk_rhs(f1, output=1, field=field_o)
    g = ih.Gpu(shape, bounds=BOX_CONFIG, bins=bins)
    dfield = g.to_device(y0)
    d1 = g.to_device(y1)
    dr = g.alloc(g.Ntot)
    g.set_field(dfield)
    g.mk_rhs(d1, dr, flags=ih.IHG_OUTPUT)
    og = ih.unpack_obs(g.observables(), bins)
    oo = oracle_mod.unpack_obs(o.obs(), bins)
    assert rel_l2(og["phi_smry"], oo["phi_smry"]) < TOL_OBS
    # and it is the summary of y0, not of y1
    assert abs(og["phi_smry"][0] - y0[:o.N].mean()) < 1e-12


@pytest.mark.parametrize("gw", [False, True])
def test_initial_psi(ih, oracle_mod, gw):
    shape = (32, 32, 32)
    y = make_state(shape, seed=11, gw=gw, psi_amp=0.0)
    o = oracle_mod.Oracle(*shape, bounds=BOX_CONFIG, gw=gw)
    fo = y.copy()
    dfo = np.zeros_like(y)
    o.mk_initial_psi(fo, dfo)
    g = ih.Gpu(shape, bounds=BOX_CONFIG, gw=gw)
    df = g.to_device(y)
    ddf = g.alloc(g.Ntot)
    g.mk_initial_psi(df, ddf)
    fg = df.download()
    B = blocks(shape)
    for k in ["phi", "dphi", "psi", "dpsi"]:
        e = rel_l2(fg[B[k]], fo[B[k]])
        print(k, e)
        assert e < 1e-10, k
    assert np.abs(fg[B["psi"]]).max() > 0
    dg = ddf.download()
    assert rel_l2(dg[:4 * o.N], dfo[:4 * o.N]) < TOL_RHS
    assert g.calls_rhs == 1


def test_rhs_repeatable(ih):
    """The RHS and the reductions are two-level in fixed order: bitwise
    repeatable.  The spectra are histograms accumulated with LDS atomics, so
    only their order of summation may vary."""
    shape = (32, 32, 32)
    y = make_state(shape, seed=5, gw=True)
    g = ih.Gpu(shape, bounds=BOX_CONFIG, gw=True, bins=16)
    df = g.to_device(y)
    r1, r2 = g.alloc(g.Ntot), g.alloc(g.Ntot)
    g.mk_rhs(df, r1, flags=2)
    o1 = g.observables()
    g.mk_rhs(df, r2, flags=2)
    o2 = g.observables()
    assert np.array_equal(r1.download(), r2.download())
    assert np.array_equal(o1[:28], o2[:28])
    assert np.allclose(o1[28:], o2[28:], rtol=1e-13, atol=0)


@pytest.mark.parametrize("shape,gw", [((32, 16, 64), True), ((64, 64, 64), False)])
def test_log_binned_spectra(ih, oracle_mod, shape, gw):
    """Observables with log-spaced spectrum bins (log_bins = 1) against the
    oracle's."""
    bins = 32
    y = make_state(shape, ANISO, seed=8, gw=gw, a=1.2)
    o = oracle_mod.Oracle(*shape, bounds=ANISO, mass=4.0, gw=gw, filt=True,
                          bins=bins, log_bins=True)
    fo = y.copy()
    o.mk_rhs(fo, filter=1, output=1, field=fo)
    g = ih.Gpu(shape, bounds=ANISO, mass=4.0, gw=gw, filt=True, bins=bins,
               log_bins=True)
    df = g.to_device(y)
    dr = g.alloc(g.Ntot)
    g.mk_rhs(df, dr, flags=3)
    og = ih.unpack_obs(g.observables(), bins)
    oo = oracle_mod.unpack_obs(o.obs(), bins)
    for k in ("phi_ps", "psi_ps", "rho_ps") + (("gw",) if gw else ()):
        e = rel_l2(og[k], oo[k])
        print("log-binned", k, e)
        assert e < TOL_OBS, k


def test_api_rejects_bad_arguments(ih):
    """Misuse returns an error code with a message, and leaves the context
    usable."""
    import ctypes
    L = ih.lib()
    shape = (16, 16, 16)
    g = ih.Gpu(shape, bounds=BOX_CONFIG, gw=True, bins=8)
    f = g.to_device(make_state(shape, seed=1, gw=True))
    r = g.alloc(g.Ntot)
    assert L.ihg_mk_rhs(g.ctx, 0.0, f.ptr, f.ptr, 0) != 0           # f == result
    assert L.ihg_mk_rhs(g.ctx, 0.0, f.ptr + 8, r.ptr, 0) != 0       # misaligned
    assert "aligned" in L.ihg_last_error().decode()
    assert L.ihg_init_gaussian(g.ctx, r.ptr, 1, 1e-5, 1.0, 0.0, -1.0, 1.0) != 0
    ks = (ctypes.c_void_p * 16)(*([r.ptr] * 16))
    assert L.ihg_dp853_dense(g.ctx, 1.5, 0.1, f.ptr, r.ptr, ks, r.ptr) != 0
    assert L.ihg_dp853_dense_stage(g.ctx, 13, 0.1, f.ptr, ks, r.ptr) != 0
    assert L.ihg_dp853_stage(g.ctx, 13, 0.1, f.ptr, ks, r.ptr) != 0
    other = ih.Gpu(shape, bounds=BOX_CONFIG, gw=True, bins=8)
    assert L.ihg_free(other.ctx, r.ptr) != 0                         # not its array
    # still usable
    g.mk_rhs(f, r, flags=0)
    assert np.all(np.isfinite(r.download()))
    other.close()
