/*
 * ih_h5.c -- the reference's HDF5 output file and IC_FROM_H5_FILE readers
 * (src/io.c) over the HDF5 1.10 C API, fed by the save() hook of the
 * MI355X host (include/instaheat_h5.h).
 *
 * The file keeps the reference's format: one dataset per output, no
 * groups; the time series are (n_slices, dim) doubles, extensible in time
 * and chunked by buf_size slices, filled from host buffers of buf_size
 * slices (src/io.c:259-304, 336-432, 494-613); parameters are compact 1-D
 * datasets (src/io.c:306-323); strings are fixed-length Fortran strings
 * (src/io.c:203-256).  HDF5's own error printing is switched off inside
 * these calls; failures come back as codes with ih_h5_last_error().
 */
#define _POSIX_C_SOURCE 199309L
#include <float.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <hdf5.h>

#include "instaheat_h5.h"

#define N_SUMMARY 4      /* SUMMARY_VALUES, src/main_template.h:324 */
#define N_CONSTRAINTS 4  /* NUMBER_CONSTRAINTS, src/main_template.h:347 */
#define MAX_SERIES 24
/* HDF5 chunks must stay below 4 GiB */
#define CHUNK_BYTES_MAX 4294967295.0
/* compact datasets live in the object header (64 KiB with its fields) */
#define COMPACT_MAX_DOUBLES 8000

static char g_err[512];

static int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

const char *ih_h5_last_error(void) { return g_err; }

/* HDF5 prints its error stack by default; these calls report codes */
typedef struct quiet {
    H5E_auto2_t fn;
    void *data;
} quiet;

static void quiet_on(quiet *q)
{
    H5Eget_auto2(H5E_DEFAULT, &q->fn, &q->data);
    H5Eset_auto2(H5E_DEFAULT, NULL, NULL);
}

static void quiet_off(const quiet *q) { H5Eset_auto2(H5E_DEFAULT, q->fn, q->data); }

/* where the values of one time-series dataset come from */
enum src_kind { SRC_TIME, SRC_A, SRC_OBS, SRC_FIELD, SRC_RHO };

typedef struct series {
    const char *name;
    int rank;
    enum src_kind kind;
    size_t off; /* SRC_OBS: offset in the packed observables;
                   SRC_FIELD: 0..3 = phi, dphi, psi, dpsi */
    size_t dim;
    hid_t id;
    double *buf; /* buf_size * dim */
} series;

struct ih_h5 {
    hid_t file;
    ih_sim *sim;
    size_t N, Ntot, obs_len, out_n;
    long n[3], stride[3], outn[3];
    long buf_size, index, slices;
    unsigned outputs;
    int err;
    int nser;
    series ser[MAX_SERIES];
    double *host_fields; /* 4 out_n, a bound simulation's strided fields */
    double *host_rho;    /* out_n */
    double t_create;     /* wall clock at ih_h5_create */
    double calls_rhs;    /* the bound simulation's mon.calls_rhs at the
                            latest save, < 0 before one */
};

static double wall_time(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

void ih_h5_opts_defaults(ih_h5_opts *o)
{
    memset(o, 0, sizeof *o);
    o->outputs = IH_H5_OUT_DEFAULT;
    o->buf_size = 5;
    o->stride[0] = o->stride[1] = o->stride[2] = 1;
    o->seed = 0.0;
    o->inflaton_mass = -1.0;
    o->bunch_davies_cutoff = 0.0;
    o->threads = 1.0;
    o->fftw_threads = 1.0;
    o->commit_hash = NULL;
}

/* h5_create_dset, src/io.c:259-304 */
static int add_series(ih_h5 *h, const char *name, int rank, enum src_kind kind,
                      size_t off, size_t dim)
{
    if (h->nser >= MAX_SERIES) return fail(IHG_EINVAL, "too many datasets");
    series *s = &h->ser[h->nser];
    memset(s, 0, sizeof *s);
    s->name = name;
    s->rank = rank;
    s->kind = kind;
    s->off = off;
    s->dim = dim;
    s->id = -1;
    hsize_t rows = (hsize_t)h->buf_size;
    const double row_bytes = 8.0 * (double)dim;
    if ((double)rows * row_bytes > CHUNK_BYTES_MAX)
        rows = (hsize_t)(CHUNK_BYTES_MAX / row_bytes);
    if (rows < 1) return fail(IHG_EINVAL, "%s: one time slice exceeds an HDF5 chunk", name);
    s->buf = malloc((size_t)h->buf_size * dim * sizeof(double));
    if (!s->buf) return fail(IHG_ENOMEM, "%s: out of host memory", name);
    ++h->nser;
    hsize_t cur[2] = {0, dim};
    hsize_t max[2] = {H5S_UNLIMITED, dim};
    hsize_t chunk[2] = {rows, dim};
    hid_t space = H5Screate_simple(rank, cur, max);
    hid_t plist = H5Pcreate(H5P_DATASET_CREATE);
    int rc = 0;
    if (space < 0 || plist < 0 || H5Pset_layout(plist, H5D_CHUNKED) < 0 ||
        H5Pset_chunk(plist, rank, chunk) < 0)
        rc = fail(IHG_EINVAL, "%s: dataspace / property list failed", name);
    if (!rc) {
        s->id = H5Dcreate2(h->file, name, H5T_NATIVE_DOUBLE, space, H5P_DEFAULT,
                           plist, H5P_DEFAULT);
        if (s->id < 0) rc = fail(IHG_EINVAL, "%s: dataset creation failed", name);
    }
    if (plist >= 0) H5Pclose(plist);
    if (space >= 0) H5Sclose(space);
    return rc;
}

/* h5_write_simple, src/io.c:306-323, with the layout as an argument */
static int write_doubles(hid_t file, const char *name, const double *val,
                         size_t n, H5D_layout_t layout)
{
    hsize_t dim[1] = {n};
    hid_t space = H5Screate_simple(1, dim, dim);
    hid_t plist = H5Pcreate(H5P_DATASET_CREATE);
    hid_t dset = -1;
    int rc = 0;
    if (space < 0 || plist < 0 || H5Pset_layout(plist, layout) < 0)
        rc = fail(IHG_EINVAL, "%s: dataspace / property list failed", name);
    if (!rc) {
        dset = H5Dcreate2(file, name, H5T_NATIVE_DOUBLE, space, H5P_DEFAULT,
                          plist, H5P_DEFAULT);
        if (dset < 0)
            rc = fail(IHG_EINVAL, "%s: dataset creation failed (does it exist?)", name);
    }
    if (!rc && H5Dwrite(dset, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT,
                        val) < 0)
        rc = fail(IHG_EINVAL, "%s: write failed", name);
    if (dset >= 0) H5Dclose(dset);
    if (plist >= 0) H5Pclose(plist);
    if (space >= 0) H5Sclose(space);
    return rc;
}

static int write_compact(hid_t file, const char *name, const double *val, size_t n)
{
    if (n < 1 || n > COMPACT_MAX_DOUBLES)
        return fail(IHG_EINVAL, "%s: %zu values do not fit a compact dataset", name, n);
    return write_doubles(file, name, val, n, H5D_COMPACT);
}

/* the fixed-length string datasets of src/io.c:203-256: a Fortran string
 * of len - 1 characters in the file, written from a C string of len */
static int write_string(hid_t file, const char *name, const char *str, size_t len)
{
    char buf[64];
    if (len > sizeof buf) return fail(IHG_EINVAL, "%s: string too long", name);
    memset(buf, 0, sizeof buf);
    strncpy(buf, str, len - 1);
    hsize_t dim[1] = {1};
    hid_t ftype = H5Tcopy(H5T_FORTRAN_S1);
    hid_t mtype = H5Tcopy(H5T_C_S1);
    hid_t space = H5Screate_simple(1, dim, NULL);
    hid_t dset = -1;
    int rc = 0;
    if (ftype < 0 || mtype < 0 || space < 0 || H5Tset_size(ftype, len - 1) < 0 ||
        H5Tset_size(mtype, len) < 0)
        rc = fail(IHG_EINVAL, "%s: string type failed", name);
    if (!rc) {
        dset = H5Dcreate2(file, name, ftype, space, H5P_DEFAULT, H5P_DEFAULT,
                          H5P_DEFAULT);
        if (dset < 0) rc = fail(IHG_EINVAL, "%s: dataset creation failed", name);
    }
    if (!rc && H5Dwrite(dset, mtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf) < 0)
        rc = fail(IHG_EINVAL, "%s: write failed", name);
    if (dset >= 0) H5Dclose(dset);
    if (space >= 0) H5Sclose(space);
    if (mtype >= 0) H5Tclose(mtype);
    if (ftype >= 0) H5Tclose(ftype);
    return rc;
}

/* the parameter datasets of h5_create_empty_by_path, src/io.c:122-256 */
static int write_parameters(ih_h5 *h, const ih_config *c, const ih_h5_opts *o)
{
    const ihg_params *g = &c->grid;
    const hid_t f = h->file;
    int dim = 3;
    if (g->nz == 1) dim = g->ny == 1 ? 1 : 2;
    double v[3];
    int rc = 0;
#define PUT1(NAME, X) \
    do { v[0] = (X); if (!rc) rc = write_compact(f, NAME, v, 1); } while (0)
#define PUT2(NAME, X, Y) \
    do { v[0] = (X); v[1] = (Y); if (!rc) rc = write_compact(f, NAME, v, 2); } while (0)
#define PUT3(NAME, X, Y, Z) \
    do { v[0] = (X); v[1] = (Y); v[2] = (Z); \
         if (!rc) rc = write_compact(f, NAME, v, 3); } while (0)
    PUT1(IH_H5_MASS_NAME, g->mass);
    PUT1(IH_H5_INFLATON_MASS_NAME, o->inflaton_mass);
    PUT1(IH_H5_DIMENSION_NAME, (double)dim);
    PUT1(IH_H5_SEED_NAME, o->seed);
    PUT1(IH_H5_THREAD_NUMBER_NAME, o->threads);
    PUT1(IH_H5_FFTW_THREAD_NUMBER_NAME, o->fftw_threads);
    PUT1(IH_H5_STRIDES_TIME_NAME, (double)c->skip);
    PUT2(IH_H5_TOLERANCES_NAME, c->rtol, c->atol);
    PUT3(IH_H5_GRIDPOINTS_INTERNAL_NAME, (double)g->nx, (double)g->ny, (double)g->nz);
    PUT2(IH_H5_SPATIAL_BOUNDS_X_NAME, g->xa, g->xb);
    PUT2(IH_H5_SPATIAL_BOUNDS_Y_NAME, g->ya, g->yb);
    PUT2(IH_H5_SPATIAL_BOUNDS_Z_NAME, g->za, g->zb);
    PUT3(IH_H5_GRIDPOINTS_OUTPUT_NAME, (double)h->outn[0], (double)h->outn[1],
         (double)h->outn[2]);
    PUT3(IH_H5_STRIDES_SPACE_NAME, (double)h->stride[0], (double)h->stride[1],
         (double)h->stride[2]);
    PUT1(IH_H5_BUNCH_DAVIES_CUTOFF_NAME, o->bunch_davies_cutoff);
    PUT1(IH_H5_MAX_DT_HUBBLE_FRACTION_NAME,
         c->max_dt_hubble_fraction > 0.0 ? c->max_dt_hubble_fraction : -1.0);
    PUT1(IH_H5_ENABLE_FILTER_NAME, g->enable_filter ? 1.0 : 0.0);
    PUT1(IH_H5_ENABLE_GW_NAME, g->enable_gw ? 1.0 : 0.0);
#undef PUT1
#undef PUT2
#undef PUT3
    if (!rc && o->commit_hash) rc = write_string(f, IH_H5_COMMIT_HASH_NAME, o->commit_hash, 16);
    if (!rc) rc = write_string(f, IH_H5_METHOD_NAME, "hyperbolic", 11);
    return rc;
}

static void free_h5(ih_h5 *h)
{
    for (int i = 0; i < h->nser; ++i) {
        if (h->ser[i].id >= 0) H5Dclose(h->ser[i].id);
        free(h->ser[i].buf);
    }
    if (h->file >= 0) H5Fclose(h->file);
    free(h->host_fields);
    free(h->host_rho);
    free(h);
}

/* h5_create_empty_by_path, src/io.c:47-257 */
int ih_h5_create(const char *path, const ih_config *cfg, const ih_h5_opts *opts,
                 ih_h5 **out)
{
    if (!path || !cfg || !out) return fail(IHG_EINVAL, "NULL argument");
    *out = NULL;
    ih_h5_opts dflt;
    if (!opts) {
        ih_h5_opts_defaults(&dflt);
        opts = &dflt;
    }
    const ihg_params *g = &cfg->grid;
    if (g->nx < 1 || g->ny < 1 || g->nz < 1 || g->bins < 1)
        return fail(IHG_EINVAL, "bad grid %d x %d x %d, %d bins", g->nx, g->ny, g->nz,
                    g->bins);
    if (opts->buf_size < 1) return fail(IHG_EINVAL, "buf_size must be >= 1");
    if (opts->outputs & ~(IH_H5_OUT_FIELDS | IH_H5_OUT_DEFAULT))
        return fail(IHG_EINVAL, "unknown output bits 0x%x", opts->outputs);
    ih_h5 *h = calloc(1, sizeof *h);
    if (!h) return fail(IHG_ENOMEM, "out of host memory");
    h->file = -1;
    h->n[0] = g->nx;
    h->n[1] = g->ny;
    h->n[2] = g->nz;
    h->N = (size_t)g->nx * g->ny * g->nz;
    h->out_n = 1;
    for (int d = 0; d < 3; ++d) {
        if (opts->stride[d] < 1) {
            free(h);
            return fail(IHG_EINVAL, "stride[%d] must be >= 1", d);
        }
        h->stride[d] = opts->stride[d];
        /* src/setup.c:230-233 */
        h->outn[d] = (h->n[d] + h->stride[d] - 1) / h->stride[d];
        h->out_n *= (size_t)h->outn[d];
    }
    h->obs_len = 28 + 4 * (size_t)g->bins;
    h->t_create = wall_time();
    h->calls_rhs = -1.0;
    h->buf_size = opts->buf_size;
    h->outputs = opts->outputs;
    const size_t bins = (size_t)g->bins;
    const unsigned u = opts->outputs;
    quiet q;
    quiet_on(&q);
    int rc = 0;
    h->file = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (h->file < 0) rc = fail(IHG_EINVAL, "cannot create %s", path);
    /* creation order of src/io.c:57-118 */
    static const char *const field_names[4] = {IH_H5_PHI_NAME, IH_H5_DPHI_NAME,
                                               IH_H5_PSI_NAME, IH_H5_DPSI_NAME};
    static const unsigned field_bits[4] = {IH_H5_OUT_PHI, IH_H5_OUT_DPHI,
                                           IH_H5_OUT_PSI, IH_H5_OUT_DPSI};
    for (int j = 0; j < 4 && !rc; ++j)
        if (u & field_bits[j]) rc = add_series(h, field_names[j], 2, SRC_FIELD, j, h->out_n);
    if (!rc && (u & IH_H5_OUT_RHO)) rc = add_series(h, IH_H5_RHO_NAME, 2, SRC_RHO, 0, h->out_n);
    static const char *const smry_names[6] = {
        IH_H5_PHI_SMRY_NAME, IH_H5_DPHI_SMRY_NAME, IH_H5_PSI_SMRY_NAME,
        IH_H5_DPSI_SMRY_NAME, IH_H5_RHO_SMRY_NAME, IH_H5_PRESSURE_SMRY_NAME};
    static const unsigned smry_bits[6] = {
        IH_H5_OUT_PHI_SMRY, IH_H5_OUT_DPHI_SMRY, IH_H5_OUT_PSI_SMRY,
        IH_H5_OUT_DPSI_SMRY, IH_H5_OUT_RHO_SMRY, IH_H5_OUT_PRESSURE_SMRY};
    for (int j = 0; j < 6 && !rc; ++j)
        if (u & smry_bits[j])
            rc = add_series(h, smry_names[j], 2, SRC_OBS, N_SUMMARY * (size_t)j, N_SUMMARY);
    const size_t ps0 = 6 * N_SUMMARY + N_CONSTRAINTS;
    if (!rc && (u & IH_H5_OUT_PHI_PS))
        rc = add_series(h, IH_H5_PHI_PS_NAME, 2, SRC_OBS, ps0, bins);
    if (!rc && (u & IH_H5_OUT_PSI_PS))
        rc = add_series(h, IH_H5_PSI_PS_NAME, 2, SRC_OBS, ps0 + bins, bins);
    if (!rc && (u & IH_H5_OUT_RHO_PS))
        rc = add_series(h, IH_H5_RHO_PS_NAME, 2, SRC_OBS, ps0 + 2 * bins, bins);
    if (!rc && g->enable_gw)
        rc = add_series(h, IH_H5_GW_PS_NAME, 2, SRC_OBS, ps0 + 3 * bins, bins);
    if (!rc && (u & IH_H5_OUT_CONSTRAINTS))
        rc = add_series(h, IH_H5_CONSTRAINTS_NAME, 2, SRC_OBS, 6 * N_SUMMARY,
                        N_CONSTRAINTS);
    /* time first among the rank-1 series: its extent is the write offset */
    if (!rc) rc = add_series(h, IH_H5_TIME_NAME, 1, SRC_TIME, 0, 1);
    if (!rc) rc = add_series(h, IH_H5_A_NAME, 1, SRC_A, 0, 1);
    if (!rc) rc = write_parameters(h, cfg, opts);
    quiet_off(&q);
    if (rc) {
        free_h5(h);
        return rc;
    }
    *out = h;
    return 0;
}

static const series *time_series(const ih_h5 *h)
{
    for (int i = 0; i < h->nser; ++i)
        if (h->ser[i].kind == SRC_TIME) return &h->ser[i];
    return NULL;
}

/* h5_write_buffer, src/io.c:434-470 */
static int write_rows(const series *s, hsize_t os, hsize_t nt)
{
    hsize_t add[2] = {nt, s->dim};
    hsize_t ext[2] = {os + nt, s->dim};
    hsize_t start[2] = {os, 0};
    int rc = 0;
    if (H5Dset_extent(s->id, ext) < 0) return fail(IHG_EINVAL, "%s: extend failed", s->name);
    hid_t mem = H5Screate_simple(s->rank, add, NULL);
    hid_t space = H5Dget_space(s->id);
    if (mem < 0 || space < 0 ||
        H5Sselect_hyperslab(space, H5S_SELECT_SET, start, NULL, add, NULL) < 0 ||
        H5Dwrite(s->id, H5T_NATIVE_DOUBLE, mem, space, H5P_DEFAULT, s->buf) < 0)
        rc = fail(IHG_EINVAL, "%s: write failed", s->name);
    if (space >= 0) H5Sclose(space);
    if (mem >= 0) H5Sclose(mem);
    return rc;
}

/* h5_write_all_buffers, src/io.c:336-432: the offset is the current
 * extent of `time` */
static int flush_buffers(ih_h5 *h, long nt)
{
    const series *ts = time_series(h);
    hsize_t cur[1] = {0};
    hid_t space = H5Dget_space(ts->id);
    if (space < 0 || H5Sget_simple_extent_dims(space, cur, NULL) < 0) {
        if (space >= 0) H5Sclose(space);
        return fail(IHG_EINVAL, "time: extent unreadable");
    }
    H5Sclose(space);
    int rc = 0;
    for (int i = 0; i < h->nser && !rc; ++i) rc = write_rows(&h->ser[i], cur[0], (hsize_t)nt);
    return rc;
}

/* the LARGE_OUTPUT copy of save(), src/io.c:560-598: every stride-th
 * point of each axis, row-major with z fastest */
static void gather(const ih_h5 *h, const double *src, double *dst)
{
    const long ny = h->n[1], nz = h->n[2];
    const long oy = h->outn[1], oz = h->outn[2];
    for (long i = 0, bi = 0; i < h->n[0]; i += h->stride[0], ++bi)
        for (long j = 0, bj = 0; j < ny; j += h->stride[1], ++bj) {
            const double *row = src + ((size_t)i * ny + j) * nz;
            double *out = dst + ((size_t)bi * oy + bj) * oz;
            for (long k = 0, bk = 0; k < nz; k += h->stride[2], ++bk) out[bk] = row[k];
        }
}

/* save(), src/io.c:494-613; `packed`: fields and rho hold only the
 * output points already (out_n per field) */
static int append(ih_h5 *h, double t, double a, const double *obs, size_t nobs,
                  const double *fields, const double *rho, int packed)
{
    if (!h || !obs) return fail(IHG_EINVAL, "NULL argument");
    if (h->err) return h->err;
    if (nobs != h->obs_len)
        return fail(IHG_EINVAL, "%zu observables, expected %zu", nobs, h->obs_len);
    if ((h->outputs & (IH_H5_OUT_FIELDS & ~IH_H5_OUT_RHO)) && !fields)
        return fail(IHG_EINVAL, "full-field output without fields");
    if ((h->outputs & IH_H5_OUT_RHO) && !rho)
        return fail(IHG_EINVAL, "rho output without rho");
    for (int i = 0; i < h->nser; ++i) {
        series *s = &h->ser[i];
        double *dst = s->buf + (size_t)h->index * s->dim;
        switch (s->kind) {
        case SRC_TIME: dst[0] = t; break;
        case SRC_A: dst[0] = a; break;
        case SRC_OBS: memcpy(dst, obs + s->off, s->dim * sizeof(double)); break;
        case SRC_FIELD:
            if (packed)
                memcpy(dst, fields + s->off * h->out_n, h->out_n * sizeof(double));
            else
                gather(h, fields + s->off * h->N, dst);
            break;
        case SRC_RHO:
            if (packed)
                memcpy(dst, rho, h->out_n * sizeof(double));
            else
                gather(h, rho, dst);
            break;
        }
    }
    ++h->slices;
    if (++h->index < h->buf_size) return 0;
    quiet q;
    quiet_on(&q);
    int rc = flush_buffers(h, h->buf_size);
    quiet_off(&q);
    h->index = 0;
    if (rc) h->err = rc;
    return rc;
}

int ih_h5_append(ih_h5 *h, double t, double a, const double *obs, size_t nobs,
                 const double *fields, const double *rho)
{
    return append(h, t, a, obs, nobs, fields, rho, 0);
}

int ih_h5_bind(ih_h5 *h, ih_sim *s)
{
    if (!h || !s) return fail(IHG_EINVAL, "NULL argument");
    size_t sz[4];
    int rc = ihg_sizes(ih_gpu(s), sz);
    if (rc) return fail(rc, "ihg_sizes: %s", ihg_last_error());
    if (sz[0] != h->N) return fail(IHG_EINVAL, "simulation has %zu points, file %zu", sz[0], h->N);
    if ((h->outputs & (IH_H5_OUT_FIELDS & ~IH_H5_OUT_RHO)) && !h->host_fields) {
        h->host_fields = malloc(4 * h->out_n * sizeof(double));
        if (!h->host_fields) return fail(IHG_ENOMEM, "out of host memory");
    }
    if ((h->outputs & IH_H5_OUT_RHO) && !h->host_rho) {
        h->host_rho = malloc(h->out_n * sizeof(double));
        if (!h->host_rho) return fail(IHG_ENOMEM, "out of host memory");
    }
    h->sim = s;
    h->Ntot = sz[3];
    return 0;
}

void ih_h5_save(void *user, double t, double a, const double *obs, size_t nobs)
{
    ih_h5 *h = user;
    if (!h || h->err) return;
    const double *fields = NULL, *rho = NULL;
    int rc = 0;
    if (h->outputs & IH_H5_OUT_FIELDS) {
        /* only the output points cross PCIe: gathered on the device */
        const int stride[3] = {(int)h->stride[0], (int)h->stride[1],
                               (int)h->stride[2]};
        if (!h->sim) rc = fail(IHG_EINVAL, "full-field output needs ih_h5_bind");
        const double *y = rc ? NULL : ih_save_state(h->sim);
        if (!rc && !y) rc = fail(IHG_EINVAL, "ih_h5_save called outside save()");
        static const unsigned bits[4] = {IH_H5_OUT_PHI, IH_H5_OUT_DPHI,
                                         IH_H5_OUT_PSI, IH_H5_OUT_DPSI};
        for (int j = 0; j < 4 && !rc; ++j)
            if (h->outputs & bits[j]) {
                rc = ihg_download_strided(ih_gpu(h->sim), y + j * h->N, stride,
                                          h->host_fields + j * h->out_n);
                if (rc) rc = fail(rc, "download: %s", ihg_last_error());
            }
        fields = h->host_fields;
        if (!rc && (h->outputs & IH_H5_OUT_RHO)) {
            /* rho of the output call that produced obs */
            rc = ihg_download_strided(ih_gpu(h->sim), NULL, stride, h->host_rho);
            if (rc) rc = fail(rc, "rho: %s", ihg_last_error());
            rho = h->host_rho;
        }
    }
    if (!rc) rc = append(h, t, a, obs, nobs, fields, rho, 1);
    /* the count main writes at the end: the last save follows the run's
     * last RHS call (prepare_and_save_timeslice) */
    if (!rc && h->sim) h->calls_rhs = (double)ihg_calls_rhs(ih_gpu(h->sim));
    if (rc) h->err = rc;
}

int ih_h5_status(const ih_h5 *h) { return h ? h->err : IHG_EINVAL; }
long ih_h5_slices(const ih_h5 *h) { return h ? h->slices : -1; }

int ih_h5_write_simple(ih_h5 *h, const char *name, const double *val, size_t n)
{
    if (!h || !name || !val) return fail(IHG_EINVAL, "NULL argument");
    quiet q;
    quiet_on(&q);
    int rc = write_compact(h->file, name, val, n);
    quiet_off(&q);
    return rc;
}

/* src/dopri853.c:234-246, src/main.c:104-106 */
int ih_h5_write_dp_stats(ih_h5 *h, const ih_dp_stats *st)
{
    if (!h || !st) return fail(IHG_EINVAL, "NULL argument");
    double v[5] = {(double)st->n_stp, (double)st->n_ok, (double)st->n_bad,
                   st->integration_seconds, (double)st->calls_rhs};
    const char *names[5] = {IH_H5_STEPS_TOTAL_NAME, IH_H5_STEPS_OK_NAME,
                            IH_H5_STEPS_BAD_NAME, IH_H5_RUNTIME_STEPPER_NAME,
                            IH_H5_COUNTER_RHS};
    int rc = 0;
    for (int i = 0; i < 5 && !rc; ++i) rc = ih_h5_write_simple(h, names[i], &v[i], 1);
    return rc;
}

/* h5_write_followup, src/io.c:822-827 */
int ih_h5_write_followup(ih_h5 *h)
{
    if (!h) return fail(IHG_EINVAL, "NULL argument");
    if (!h->sim) return fail(IHG_EINVAL, "followup needs ih_h5_bind");
    double *y = malloc(h->Ntot * sizeof(double));
    if (!y) return fail(IHG_ENOMEM, "out of host memory");
    int rc = ihg_download(ih_gpu(h->sim), y, ih_field(h->sim), h->Ntot);
    if (rc) rc = fail(rc, "download: %s", ihg_last_error());
    if (!rc) {
        quiet q;
        quiet_on(&q);
        rc = write_doubles(h->file, IH_H5_FOLLOWUP_NAME, y, h->Ntot, H5D_CONTIGUOUS);
        quiet_off(&q);
    }
    free(y);
    return rc;
}

long ih_h5_buffered(void *h5)
{
    const ih_h5 *h = h5;
    return h ? h->index : 0;
}

int ih_h5_followup(void *h5) { return ih_h5_write_followup(h5); }

/* h5_close, src/io.c:472-488, after main's last writes (src/main.c:99-106):
 * runtime_total and calls_rhs, unless written already */
int ih_h5_close(ih_h5 *h)
{
    if (!h) return fail(IHG_EINVAL, "NULL argument");
    quiet q;
    quiet_on(&q);
    int rc = h->err;
    if (!rc && h->index > 0) rc = flush_buffers(h, h->index);
    if (!rc && H5Lexists(h->file, IH_H5_RUNTIME_TOTAL_NAME, H5P_DEFAULT) == 0) {
        const double v = wall_time() - h->t_create;
        rc = write_compact(h->file, IH_H5_RUNTIME_TOTAL_NAME, &v, 1);
    }
    /* from the latest save: the simulation may be gone by now */
    if (!rc && h->calls_rhs >= 0.0 &&
        H5Lexists(h->file, IH_H5_COUNTER_RHS, H5P_DEFAULT) == 0)
        rc = write_compact(h->file, IH_H5_COUNTER_RHS, &h->calls_rhs, 1);
    if (H5Fflush(h->file, H5F_SCOPE_GLOBAL) < 0 && !rc)
        rc = fail(IHG_EINVAL, "flush failed");
    for (int i = 0; i < h->nser; ++i) {
        if (h->ser[i].id >= 0 && H5Dclose(h->ser[i].id) < 0 && !rc)
            rc = fail(IHG_EINVAL, "%s: close failed", h->ser[i].name);
        h->ser[i].id = -1;
    }
    if (H5Fclose(h->file) < 0 && !rc) rc = fail(IHG_EINVAL, "file close failed");
    h->file = -1;
    free_h5(h);
    quiet_off(&q);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* readers                                                                  */
/* ------------------------------------------------------------------------ */

/* extent of an open dataset: rank (1 or 2) and dims (dims[1] = 1 for
 * rank 1) */
static int extent(hid_t dset, const char *name, int *rank, hsize_t dims[2])
{
    hid_t space = H5Dget_space(dset);
    if (space < 0) return fail(IHG_EINVAL, "%s: no dataspace", name);
    int r = H5Sget_simple_extent_ndims(space);
    dims[0] = dims[1] = 1;
    if (r < 1 || r > 2 || H5Sget_simple_extent_dims(space, dims, NULL) < 0) {
        H5Sclose(space);
        return fail(IHG_EINVAL, "%s: rank %d is not 1 or 2", name, r);
    }
    H5Sclose(space);
    if (r == 1) dims[1] = 1;
    *rank = r;
    return 0;
}

/* one row (rank 2) or element (rank 1) of an open dataset into out */
static int read_row(hid_t dset, const char *name, int rank, hsize_t row,
                    hsize_t width, double *out)
{
    hsize_t start[2] = {row, 0};
    hsize_t count[2] = {1, width};
    hsize_t mdim[1] = {width};
    hid_t space = H5Dget_space(dset);
    hid_t mem = H5Screate_simple(1, mdim, NULL);
    int rc = 0;
    if (space < 0 || mem < 0 ||
        H5Sselect_hyperslab(space, H5S_SELECT_SET, start, NULL, count, NULL) < 0 ||
        H5Dread(dset, H5T_NATIVE_DOUBLE, mem, space, H5P_DEFAULT, out) < 0)
        rc = fail(IHG_EINVAL, "%s: reading row %llu failed", name,
                  (unsigned long long)row);
    (void)rank;
    if (mem >= 0) H5Sclose(mem);
    if (space >= 0) H5Sclose(space);
    return rc;
}

static int read_in(hid_t file, const char *name, long row, double *out, size_t cap,
                   size_t dims_out[2])
{
    hid_t dset = H5Dopen2(file, name, H5P_DEFAULT);
    if (dset < 0) return fail(IHG_EINVAL, "no dataset %s", name);
    int rank = 0;
    hsize_t dims[2];
    int rc = extent(dset, name, &rank, dims);
    if (!rc && dims_out) {
        dims_out[0] = (size_t)dims[0];
        dims_out[1] = (size_t)dims[1];
    }
    if (!rc && out) {
        hid_t type = H5Dget_type(dset);
        if (type < 0 || H5Tget_class(type) != H5T_FLOAT)
            rc = fail(IHG_EINVAL, "%s does not hold doubles", name);
        if (type >= 0) H5Tclose(type);
        if (!rc && row < 0) {
            if (cap < dims[0] * dims[1])
                rc = fail(IHG_EINVAL, "%s: %llu values, room for %zu", name,
                          (unsigned long long)(dims[0] * dims[1]), cap);
            else if (dims[0] * dims[1] > 0 &&
                     H5Dread(dset, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT,
                             out) < 0)
                rc = fail(IHG_EINVAL, "%s: read failed", name);
        } else if (!rc) {
            if ((hsize_t)row >= dims[0])
                rc = fail(IHG_EINVAL, "%s: row %ld of %llu", name, row,
                          (unsigned long long)dims[0]);
            else if (cap < dims[1])
                rc = fail(IHG_EINVAL, "%s: %llu values per row, room for %zu", name,
                          (unsigned long long)dims[1], cap);
            else
                rc = read_row(dset, name, rank, (hsize_t)row, dims[1], out);
        }
    }
    H5Dclose(dset);
    return rc;
}

int ih_h5_read(const char *path, const char *name, long row, double *out, size_t cap,
               size_t dims[2])
{
    if (!path || !name) return fail(IHG_EINVAL, "NULL argument");
    quiet q;
    quiet_on(&q);
    int rc = 0;
    hid_t file = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (file < 0)
        rc = fail(IHG_EINVAL, "cannot open %s", path);
    else {
        rc = read_in(file, name, row, out, cap, dims);
        H5Fclose(file);
    }
    quiet_off(&q);
    return rc;
}

int ih_h5_read_string(const char *path, const char *name, char *buf, size_t cap)
{
    if (!path || !name || !buf || cap < 1) return fail(IHG_EINVAL, "bad argument");
    quiet q;
    quiet_on(&q);
    int rc = 0;
    hid_t file = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    hid_t dset = -1, ftype = -1, mtype = -1;
    if (file < 0) rc = fail(IHG_EINVAL, "cannot open %s", path);
    if (!rc && (dset = H5Dopen2(file, name, H5P_DEFAULT)) < 0)
        rc = fail(IHG_EINVAL, "no dataset %s", name);
    if (!rc && ((ftype = H5Dget_type(dset)) < 0 || H5Tget_class(ftype) != H5T_STRING))
        rc = fail(IHG_EINVAL, "%s is not a string", name);
    size_t len = rc ? 0 : H5Tget_size(ftype);
    if (!rc && len + 1 > cap) rc = fail(IHG_EINVAL, "%s: %zu characters, room for %zu", name, len, cap - 1);
    if (!rc) {
        mtype = H5Tcopy(H5T_C_S1);
        if (mtype < 0 || H5Tset_size(mtype, len + 1) < 0 ||
            H5Dread(dset, mtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf) < 0)
            rc = fail(IHG_EINVAL, "%s: read failed", name);
    }
    if (!rc) {
        buf[len] = '\0';
        size_t n = strlen(buf);
        while (n > 0 && (buf[n - 1] == ' ' || buf[n - 1] == '\n')) buf[--n] = '\0';
    }
    if (mtype >= 0) H5Tclose(mtype);
    if (ftype >= 0) H5Tclose(ftype);
    if (dset >= 0) H5Dclose(dset);
    if (file >= 0) H5Fclose(file);
    quiet_off(&q);
    return rc;
}

/* the `time` dataset in full: *nt values in *times (malloc'd) */
static int read_times(hid_t file, double **times, hsize_t *nt)
{
    size_t dims[2];
    int rc = read_in(file, IH_H5_TIME_NAME, -1, NULL, 0, dims);
    if (!rc && dims[1] != 1) rc = fail(IHG_EINVAL, "Could not read time properly.");
    if (!rc && dims[0] == 0) rc = fail(IHG_EINVAL, "time is empty");
    if (rc) return rc;
    *times = malloc(dims[0] * sizeof(double));
    if (!*times) return fail(IHG_ENOMEM, "out of host memory");
    rc = read_in(file, IH_H5_TIME_NAME, -1, *times, dims[0], NULL);
    if (rc) {
        free(*times);
        *times = NULL;
    }
    *nt = dims[0];
    return rc;
}

/* h5_read_followup, src/io.c:761-801 */
int ih_h5_read_followup(ih_sim *s, const char *path)
{
    if (!s || !path) return fail(IHG_EINVAL, "NULL argument");
    size_t sz[4];
    int rc = ihg_sizes(ih_gpu(s), sz);
    if (rc) return fail(rc, "ihg_sizes: %s", ihg_last_error());
    const size_t Ntot = sz[3];
    double *y = malloc(Ntot * sizeof(double));
    double *times = NULL;
    hsize_t nt = 0;
    if (!y) return fail(IHG_ENOMEM, "out of host memory");
    quiet q;
    quiet_on(&q);
    hid_t file = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (file < 0) rc = fail(IHG_EINVAL, "cannot open %s", path);
    size_t dims[2] = {0, 0};
    if (!rc) rc = read_in(file, IH_H5_FOLLOWUP_NAME, -1, NULL, 0, dims);
    if (!rc && (dims[1] != 1 || dims[0] != Ntot))
        rc = fail(IHG_EINVAL, "Could not read followup data properly (%zu values, "
                  "expected %zu). Do all specifications match?", dims[0], Ntot);
    if (!rc) rc = read_in(file, IH_H5_FOLLOWUP_NAME, -1, y, Ntot, NULL);
    if (!rc) rc = read_times(file, &times, &nt);
    if (file >= 0) H5Fclose(file);
    quiet_off(&q);
    if (!rc) {
        rc = ihg_upload(ih_gpu(s), ih_field(s), y, Ntot);
        if (!rc) rc = ihg_sync(ih_gpu(s));
        if (rc) rc = fail(rc, "upload: %s", ihg_last_error());
    }
    if (!rc) {
        /* pars.t.ti and pars.t.t (src/io.c:783-784) */
        ih_set_time(s, times[nt - 1]);
        ih_set_ti(s, times[nt - 1]);
    }
    free(times);
    free(y);
    return rc;
}

/* h5_read_timeslice and h5_read_and_fill, src/io.c:640-759 */
int ih_h5_read_timeslice(ih_sim *s, const char *path, double t, double *t_found)
{
    if (!s || !path) return fail(IHG_EINVAL, "NULL argument");
    size_t sz[4];
    int rc = ihg_sizes(ih_gpu(s), sz);
    if (rc) return fail(rc, "ihg_sizes: %s", ihg_last_error());
    const size_t N = sz[0], Ntot = sz[3];
    double *y = calloc(Ntot, sizeof(double));
    double *times = NULL;
    hsize_t nt = 0, index = 0;
    if (!y) return fail(IHG_ENOMEM, "out of host memory");
    quiet q;
    quiet_on(&q);
    hid_t file = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (file < 0) rc = fail(IHG_EINVAL, "cannot open %s", path);
    if (!rc) rc = read_times(file, &times, &nt);
    if (!rc) {
        index = nt;
        for (hsize_t i = 0; i < nt; ++i)
            if (times[i] + DBL_EPSILON >= t) {
                index = i;
                break;
            }
        if (index == nt) index = nt - 1; /* past the last slice: start there */
    }
    static const char *const names[4] = {IH_H5_PHI_NAME, IH_H5_DPHI_NAME,
                                         IH_H5_PSI_NAME, IH_H5_DPSI_NAME};
    for (int j = 0; j < 4 && !rc; ++j) {
        size_t dims[2];
        rc = read_in(file, names[j], -1, NULL, 0, dims);
        if (!rc && dims[1] != N)
            rc = fail(IHG_EINVAL, "%s: %zu values per slice, the grid has %zu "
                      "(full fields at stride 1 are needed)", names[j], dims[1], N);
        if (!rc) rc = read_in(file, names[j], (long)index, y + j * N, N, NULL);
    }
    if (!rc) {
        size_t dims[2];
        rc = read_in(file, IH_H5_A_NAME, -1, NULL, 0, dims);
        if (!rc && dims[0] != nt)
            rc = fail(IHG_EINVAL, "a has %zu slices, time %llu", dims[0],
                      (unsigned long long)nt);
        if (!rc) rc = read_in(file, IH_H5_A_NAME, (long)index, y + Ntot - 1, 1, NULL);
    }
    if (file >= 0) H5Fclose(file);
    quiet_off(&q);
    if (!rc) {
        rc = ihg_upload(ih_gpu(s), ih_field(s), y, Ntot);
        if (!rc) rc = ihg_sync(ih_gpu(s));
        if (rc) rc = fail(rc, "upload: %s", ihg_last_error());
    }
    if (!rc) {
        /* pars.t.ti and pars.t.t (src/io.c:674-675) */
        ih_set_time(s, times[index]);
        ih_set_ti(s, times[index]);
        if (t_found) *t_found = times[index];
    }
    free(times);
    free(y);
    return rc;
}
