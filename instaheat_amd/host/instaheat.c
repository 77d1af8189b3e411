/*
 * instaheat.c -- C99 host of the InstaHeat hot path (include/instaheat.h).
 *
 * The control flow is the reference's; every array operation is a libihg
 * call on device memory.  Per accepted DOPRI853 step the host reads back
 * only the two error sums of each attempt and rho_mean (for the Hubble cap).
 */
#define _POSIX_C_SOURCE 199309L
#include "instaheat.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define IH_MAX(x, y) (((x) > (y)) ? (x) : (y))
#define IH_MIN(x, y) (((x) > (y)) ? (y) : (x))

struct ih_sim {
    ih_config cfg;
    ihg_ctx *g;
    size_t N, M, Next, Ntot;
    double *field, *dfield, *field_new, *dfield_new; /* device */
    double *stage[12];   /* stage RHS arrays: [1..11] used by DOPRI853,
                            [0..3] are k1..k4 of RK4 */
    double *k_tmp;       /* stage input */
    struct ih_evolution_flags evo_flags;
    double t;
    double *obs;         /* host copy of the observables */
    size_t obs_len;
    double dp_c[13];     /* DOPRI853 nodes c_s */
    double dp_c_extra[3]; /* nodes of the dense output's extra stages */
    /* the last IH_HISTORY DOPRI853 attempts, oldest first (the step
       proposer's inputs); kept across runs, cleared by ih_set_time */
    ih_dp_record hist[IH_HISTORY];
    long n_hist;
    const double *save_y; /* the state of the save() call in progress */
    double t_created;     /* wall time of ih_create (MAX_RUNTIME's origin) */
};

static double wall_time(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

void ih_config_defaults(ih_config *cfg)
{
    memset(cfg, 0, sizeof *cfg);
    cfg->grid.nx = cfg->grid.ny = cfg->grid.nz = 32;
    cfg->grid.xa = cfg->grid.ya = cfg->grid.za = -5.0;
    cfg->grid.xb = cfg->grid.yb = cfg->grid.zb = 5.0;
    cfg->grid.mass = 4.0;
    cfg->grid.enable_gw = 1;
    cfg->grid.enable_filter = 1;
    cfg->grid.bins = 50;
    cfg->integration_method = IH_DOPRI853;
    cfg->ti = 0.0;
    cfg->tf = 1.0;
    cfg->dt = 1.0e-7;
    cfg->dt_min = 1.0e-9;
    cfg->max_steps = (long)1e15;
    cfg->beta = 0.0;
    cfg->safe = 0.9;
    cfg->minscale = 0.333;
    cfg->maxscale = 3.0;
    cfg->rtol = 1.0e-10;
    cfg->atol = 1.0e-14;
    cfg->max_dt_hubble_fraction = 1.0e-2;
    cfg->skip = 1;
    cfg->a_initial = 1.0;
    cfg->max_runtime = 172500.0;
    cfg->enable_followup = 0;
    cfg->enable_stiffness_check = 0;
}

void ih_destroy(ih_sim *s)
{
    if (!s) return;
    if (s->g) {
        double *arrs[] = {s->field, s->dfield, s->field_new, s->dfield_new,
                          s->k_tmp};
        for (size_t i = 0; i < sizeof arrs / sizeof arrs[0]; ++i)
            ihg_free(s->g, arrs[i]);
        for (int i = 0; i < 12; ++i) ihg_free(s->g, s->stage[i]);
        ihg_destroy(s->g);
    }
    free(s->obs);
    free(s);
}

int ih_create(const ih_config *cfg, ih_sim **out)
{
    if (!cfg || !out) return IHG_EINVAL;
    *out = NULL;
    if (cfg->skip < 1 || cfg->dt <= 0.0 || cfg->ti > cfg->tf) return IHG_EINVAL;
    if (cfg->integration_method < IH_RK4 ||
        cfg->integration_method > IH_DOPRI853)
        return IHG_EINVAL;
    ih_sim *s = calloc(1, sizeof *s);
    if (!s) return IHG_ENOMEM;
    s->cfg = *cfg;
    int rc = ihg_create(&cfg->grid, &s->g);
    if (rc) {
        free(s);
        return rc;
    }
    size_t sz[4];
    ihg_sizes(s->g, sz);
    s->N = sz[0];
    s->M = sz[1];
    s->Next = sz[2];
    s->Ntot = sz[3];
    double **arrs[] = {&s->field, &s->dfield, &s->field_new, &s->dfield_new,
                       &s->k_tmp};
    for (size_t i = 0; i < sizeof arrs / sizeof arrs[0] && !rc; ++i)
        rc = ihg_malloc(s->g, arrs[i], s->Ntot);
    const int nstage = cfg->integration_method == IH_RK4 ? 4 : 12;
    for (int i = 0; i < nstage && !rc; ++i)
        rc = ihg_malloc(s->g, &s->stage[i], s->Ntot);
    s->obs_len = 28 + 4 * (size_t)cfg->grid.bins;
    s->obs = calloc(s->obs_len, sizeof(double));
    if (!rc && !s->obs) rc = IHG_ENOMEM;
    if (!rc) {
        double a[13 * 12], b[12], bhh[3], er[12];
        rc = ihg_dp853_tableau(a, s->dp_c, b, bhh, er);
    }
    if (!rc) {
        double ax[3 * 16], d[4 * 16];
        rc = ihg_dp853_dense_tableau(s->dp_c_extra, ax, d);
    }
    if (rc) {
        ih_destroy(s);
        return rc;
    }
    s->t = cfg->ti;
    s->t_created = wall_time();
    ihg_set_field(s->g, s->field);
    *out = s;
    return 0;
}

ihg_ctx *ih_gpu(ih_sim *s) { return s->g; }
const ih_config *ih_cfg(const ih_sim *s) { return &s->cfg; }
double *ih_field(ih_sim *s) { return s->field; }
double *ih_dfield(ih_sim *s) { return s->dfield; }
struct ih_evolution_flags *ih_flags(ih_sim *s) { return &s->evo_flags; }
double ih_time(const ih_sim *s) { return s->t; }
void ih_set_time(ih_sim *s, double t)
{
    s->t = t;
    s->n_hist = 0;
}
void ih_set_ti(ih_sim *s, double ti) { s->cfg.ti = ti; }
void ih_set_dt(ih_sim *s, double dt) { s->cfg.dt = dt; }
void ih_set_max_steps(ih_sim *s, long max_steps) { s->cfg.max_steps = max_steps; }
void ih_set_tf(ih_sim *s, double tf) { s->cfg.tf = tf; }

int ih_mk_rhs(ih_sim *s, double t, double *f, double *result)
{
    unsigned fl = (s->evo_flags.filter ? IHG_FILTER : 0u) |
                  (s->evo_flags.output ? IHG_OUTPUT : 0u);
    return ihg_mk_rhs(s->g, t, f, result, fl);
}

/* save(), src/io.c:494-613: a is read from `field` */
static int do_save(ih_sim *s, ih_save_fn save, void *user, double t)
{
    if (!save) return 0;
    double a;
    int rc = ihg_observables(s->g, s->obs, s->obs_len);
    if (!rc) rc = ihg_download(s->g, &a, s->field + s->Ntot - 1, 1);
    if (rc) return rc;
    s->save_y = s->field;
    save(user, t, a, s->obs, s->obs_len);
    s->save_y = NULL;
    return 0;
}

/* whether a time-based output falls in (t0, t0 + h] */
static int dense_due(const ih_dp_opts *o, long next, double t0, double h)
{
    while (next < o->n_out && o->t_out[next] <= t0) ++next;
    return next < o->n_out && o->t_out[next] <= t0 + h;
}

/* DOPRI853 dense output of one accepted step from t0 to t0 + h, called
 * after the swap: field = y_new, dfield = f(t0 + h, y_new), field_new = y,
 * dfield_new = the step's FSAL derivative, stage[1..11] = stages 2..12.
 * For every t_out in (t0, t0 + h]: the three extra stages once (into the
 * arrays of stages 2..4, which the interpolant does not read), the
 * interpolated state into stage[0], an output mk_rhs on it with the
 * summaries reading it, and save(t_out).  Stage 5's array (stage[4]) is
 * neither read nor written.  With step_save (a step-based save at t0 + h
 * follows) a t_out equal to t0 + h is left to that save. */
static int dense_saves(ih_sim *s, const ih_dp_opts *o, long *next, double t0,
                       double h, int step_save, ih_save_fn save, void *user)
{
    long i = *next;
    while (i < o->n_out && o->t_out[i] <= t0) ++i;
    *next = i;
    if (i >= o->n_out || o->t_out[i] > t0 + h) return 0;
    const double *k[16];
    k[0] = s->dfield_new;
    for (int j = 1; j < 12; ++j) k[j] = s->stage[j];
    k[12] = s->dfield;
    k[13] = s->stage[1];
    k[14] = s->stage[2];
    k[15] = s->stage[3];
    int rc = 0;
    for (int st = 14; st <= 16 && !rc; ++st) {
        rc = ihg_dp853_dense_stage(s->g, st, h, s->field_new, k, s->k_tmp);
        if (!rc)
            rc = ih_mk_rhs(s, t0 + s->dp_c_extra[st - 14] * h, s->k_tmp,
                           s->stage[st - 13]);
    }
    double *yd = s->stage[0];
    for (; !rc && i < o->n_out && o->t_out[i] <= t0 + h; ++i) {
        if (step_save && o->t_out[i] == t0 + h) continue;
        double theta = (o->t_out[i] - t0) / h;
        if (theta > 1.0) theta = 1.0;
        rc = ihg_dp853_dense(s->g, theta, h, s->field_new, s->field, k, yd);
        if (rc) break;
        ihg_set_field(s->g, yd);
        s->evo_flags.output = 1;
        rc = ih_mk_rhs(s, o->t_out[i], yd, s->k_tmp);
        s->evo_flags.output = 0;
        ihg_set_field(s->g, s->field);
        if (rc || !save) continue;
        double a;
        rc = ihg_observables(s->g, s->obs, s->obs_len);
        if (!rc) rc = ihg_download(s->g, &a, yd + s->Ntot - 1, 1);
        if (rc) break;
        s->save_y = yd;
        save(user, o->t_out[i], a, s->obs, s->obs_len);
        s->save_y = NULL;
    }
    *next = i;
    return rc;
}

double *ih_dense_state(ih_sim *s) { return s->stage[0]; }
const double *ih_save_state(const ih_sim *s) { return s->save_y; }

/* prepare_and_save_timeslice, src/toolbox.c:96-104 */
static int prepare_and_save(ih_sim *s, ih_save_fn save, void *user)
{
    s->evo_flags.output = 1;
    s->evo_flags.filter = 1;
    ihg_set_field(s->g, s->field);
    int rc = ih_mk_rhs(s, s->t, s->field, s->dfield);
    s->evo_flags.filter = 0;
    s->evo_flags.output = 0;
    if (rc) return rc;
    return do_save(s, save, user, s->t);
}

int ih_prepare_and_save_timeslice(ih_sim *s, ih_save_fn save, void *user)
{
    if (!s) return IHG_EINVAL;
    return prepare_and_save(s, save, user);
}

/* ------------------------------------------------------------------------ */
/* initial data                                                             */
/* ------------------------------------------------------------------------ */

int ih_init_gaussian(ih_sim *s, uint64_t seed, double amplitude, double phi0,
                     double dphi0, double k_cut)
{
    int rc = ihg_zero(s->g, s->dfield, s->Ntot);
    if (!rc)
        rc = ihg_init_gaussian(s->g, s->field, seed, amplitude, phi0, dphi0,
                               s->cfg.a_initial, k_cut);
    s->t = s->cfg.ti;
    return rc;
}

/* read_initial_data + init_from_dat (src/io.c:801-820, src/setup.c:802-813):
 * N lines "i j k phi dphi psi dpsi" in the order of the flat index, a =
 * A_INITIAL, and with !with_psi (IC_FROM_DAT_FILE_WITHOUT_PSI) psi and dpsi
 * from mk_initial_psi.  Every other element of the state is zero.  A line
 * with fewer than seven numbers, or a file with fewer than N lines, is an
 * error here (the reference only catches a line with no number at all). */
int ih_read_dat(ih_sim *s, const char *path, int with_psi)
{
    FILE *fp = fopen(path, "r");
    if (!fp) return IHG_EINVAL;
    const size_t N = s->N;
    double *h = calloc(s->Ntot, sizeof(double));
    int rc = h ? 0 : IHG_ENOMEM;
    for (size_t i = 0; !rc && i < N; ++i) {
        int ii, jj, kk;
        if (fscanf(fp, " %d %d %d %lf %lf %lf %lf", &ii, &jj, &kk, &h[i],
                   &h[i + N], &h[i + 2 * N], &h[i + 3 * N]) != 7)
            rc = IHG_EINVAL;
    }
    fclose(fp);
    if (!rc) {
        h[s->Ntot - 1] = s->cfg.a_initial;
        rc = ihg_upload(s->g, s->field, h, s->Ntot);
    }
    if (!rc) rc = ihg_zero(s->g, s->dfield, s->Ntot);
    free(h);
    s->t = s->cfg.ti;
    if (!rc && !with_psi) rc = ih_mk_initial_psi(s);
    return rc;
}

/* the current phi, dphi, psi, dpsi in the format ih_read_dat reads, with
 * (i, j, k) the grid indices of each line and 17 significant digits (the
 * values read back are the same doubles) */
int ih_write_dat(ih_sim *s, const char *path)
{
    const size_t N = s->N;
    double *h = malloc(4 * N * sizeof(double));
    if (!h) return IHG_ENOMEM;
    int rc = ihg_download(s->g, h, s->field, 4 * N);
    FILE *fp = rc ? NULL : fopen(path, "w");
    if (!rc && !fp) rc = IHG_EINVAL;
    const int Ny = s->cfg.grid.ny, Nz = s->cfg.grid.nz;
    for (size_t i = 0; !rc && i < N; ++i) {
        const int x = (int)(i / ((size_t)Ny * Nz)), y = (int)(i / Nz % Ny),
                  z = (int)(i % Nz);
        if (fprintf(fp, "%d %d %d %.17g %.17g %.17g %.17g\n", x, y, z, h[i],
                    h[i + N], h[i + 2 * N], h[i + 3 * N]) < 0)
            rc = IHG_EINVAL;
    }
    if (fp && fclose(fp) != 0 && !rc) rc = IHG_EINVAL;
    free(h);
    return rc;
}

int ih_mk_initial_psi(ih_sim *s)
{
    s->evo_flags.filter = 0;
    s->evo_flags.output = 0;
    ihg_set_field(s->g, s->field);
    return ihg_mk_initial_psi(s->g, s->cfg.ti, s->field, s->dfield);
}

/* ------------------------------------------------------------------------ */
/* RK4, src/rk4.c:27-132                                                    */
/* ------------------------------------------------------------------------ */

long ih_run_rk4(ih_sim *s, ih_save_fn save, void *user)
{
    const ih_config *c = &s->cfg;
    double dt = c->dt, t = s->t;
    double *k1 = s->stage[0], *k2 = s->stage[1], *k3 = s->stage[2],
           *k4 = s->stage[3], *tmp = s->k_tmp;
    int rc = 0;
    long nt;
    if (!k4) return IHG_EINVAL; /* created for DOPRI853 only */
    ihg_set_field(s->g, s->field);
    for (nt = 0; t < c->tf && nt < c->max_steps; ++nt) {
        if (t + dt * 1.0001 > c->tf) dt = c->tf - t;
        s->evo_flags.filter = c->grid.enable_filter ? 1 : 0;
        if (nt % c->skip == 0) {
            s->evo_flags.output = 1;
            rc = ih_mk_rhs(s, t, s->field, k1);
            s->evo_flags.output = 0;
            s->t = t;
            if (!rc) rc = do_save(s, save, user, t);
        } else {
            rc = ih_mk_rhs(s, t, s->field, k1);
        }
        s->evo_flags.filter = 0;
        if (!rc) rc = ihg_rk4_stage(s->g, tmp, s->field, k1, dt, 1);
        if (!rc) rc = ih_mk_rhs(s, t + dt / 2.0, tmp, k2);
        if (!rc) rc = ihg_rk4_stage(s->g, tmp, s->field, k2, dt, 1);
        if (!rc) rc = ih_mk_rhs(s, t + dt / 2.0, tmp, k3);
        if (!rc) rc = ihg_rk4_stage(s->g, tmp, s->field, k3, dt, 0);
        if (!rc) rc = ih_mk_rhs(s, t + dt, tmp, k4);
        if (!rc) rc = ihg_rk4_final(s->g, s->field, k1, k2, k3, k4, dt);
        if (rc) return rc;
        t += dt;
    }
    s->t = t;
    rc = prepare_and_save(s, save, user);
    if (!rc) rc = ihg_sync(s->g);
    return rc ? rc : nt;
}

/* ------------------------------------------------------------------------ */
/* DOPRI853, src/dopri853.c                                                 */
/* ------------------------------------------------------------------------ */

int ih_success(ih_ctrl *ctl, const ih_config *cfg, double err, double *dt,
               double rho_mean)
{
    const double alpha = 1.0 / 8.0 - cfg->beta * 0.2;
    double scale;
    if (err <= 1.0) {
        if (err == 0.0) {
            scale = cfg->maxscale;
        } else {
            scale = cfg->safe * pow(err, -alpha) * pow(ctl->err_old, cfg->beta);
            if (scale < cfg->minscale) scale = cfg->minscale;
            if (scale > cfg->maxscale) scale = cfg->maxscale;
        }
        if (ctl->reject)
            ctl->dt_next = (*dt) * IH_MIN(scale, 1.0);
        else
            ctl->dt_next = (*dt) * scale;
        if (cfg->max_dt_hubble_fraction > 0.0) {
            const double minstep =
                cfg->max_dt_hubble_fraction * sqrt(3.0 / rho_mean);
            if (ctl->dt_next > minstep) ctl->dt_next = minstep;
        }
        ctl->err_old = IH_MAX(err, 1.0e-4);
        ctl->reject = 0;
        return 1;
    }
    scale = IH_MAX(cfg->safe * pow(err, -alpha), cfg->minscale);
    (*dt) *= scale;
    ctl->reject = 1;
    return 0;
}

/* check_for_stiffness, src/dopri853.c:550-580 (dtlamb keeps its last value
 * when den is 0) */
int ih_stiffness_update(ih_stiffness *st, double dt, double num, double den)
{
    if (den > 0.0) st->dtlamb = dt * sqrt(num / den);
    if (st->dtlamb > 6.1) {
        st->nonstiff = 0;
        st->stiff += 1;
        if (st->stiff == 15) return 1;
    } else {
        st->nonstiff += 1;
        if (st->nonstiff == 6) st->stiff = 0;
    }
    return 0;
}

double ih_error_norm(double e5, double e3, double dt, size_t ntot)
{
    double deno = e5 + 0.01 * e3;
    if (deno <= 0.0) deno = 1.0;
    return dt * e5 * sqrt(1.0 / (ntot * deno));
}

/* try_step (src/dopri853.c:331-458) + error (:469-486) */
static int try_step(ih_sim *s, double t, double dt, double *err)
{
    const double *k[12];
    k[0] = s->dfield;
    for (int j = 1; j < 12; ++j) k[j] = s->stage[j];
    for (int st = 2; st <= 12; ++st) {
        int rc = ihg_dp853_stage(s->g, st, dt, s->field, k, s->k_tmp);
        const double ts = (st == 12) ? t + dt : t + s->dp_c[st - 1] * dt;
        if (!rc) rc = ih_mk_rhs(s, ts, s->k_tmp, s->stage[st - 1]);
        if (rc) return rc;
    }
    double sums[2];
    int rc = ihg_dp853_finish(s->g, dt, s->field, k, s->field_new,
                              s->cfg.rtol, s->cfg.atol, sums);
    if (rc) return rc;
    *err = ih_error_norm(sums[0], sums[1], dt, s->Ntot);
    return 0;
}

long ih_run_dopri853(ih_sim *s, const ih_dp_opts *opts, ih_save_fn save,
                     void *user, ih_dp_stats *stats)
{
    const ih_config *c = &s->cfg;
    ih_dp_opts none;
    memset(&none, 0, sizeof none);
    if (!opts) opts = &none;
    if (!s->stage[11]) return IHG_EINVAL; /* created for RK4 only */
    ih_ctrl ctl = {1.0e-4, c->dt, 0};
    double dt = c->dt, last_save_t = s->t;
    long n_stp, n_ok = 0, n_bad = 0, attempts = 0, next_out = 0;
    int rc, stop = 0;
    const int filter = c->grid.enable_filter ? 1 : 0;
    ih_stiffness stiff = {0.0, 0, 0};

    rc = prepare_and_save(s, save, user);
    if (rc) return rc;
    /* features for the log and the proposer, starting with the initial state */
    const int want_feat = opts->log != NULL || opts->proposer != NULL;
    double feat[3] = {0.0, 0.0, 0.0};
    ih_dp_record *hist = s->hist;
    if (want_feat) {
        rc = ihg_features(s->g, s->field, feat);
        if (rc) return rc;
    }
    rc = ihg_sync(s->g);
    if (rc) return rc;
    const double t_start = wall_time();
    for (n_stp = 0;; ++n_stp) {
        if (opts->forced_dt) {
            if (n_stp >= opts->n_forced) {
                stop = IH_STOP_FORCED;
                break;
            }
            dt = opts->forced_dt[n_stp];
        } else if (s->t + dt * 1.0001 > c->tf) {
            dt = c->tf - s->t;
        }
        /* perform_step, src/dopri853.c:271-311 */
        double h = dt, err = 0.0, rho_mean = 0.0;
        int underflow = 0;
        for (;;) {
            rc = try_step(s, s->t, h, &err);
            if (rc) return rc;
            double means[2];
            rc = ihg_means(s->g, means); /* stage 12's rho_mean */
            if (rc) return rc;
            rho_mean = means[0];
            const double tried = h;
            const int after_reject = ctl.reject;
            int ok = 1;
            if (!opts->forced_dt) ok = ih_success(&ctl, c, err, &h, rho_mean);
            if (ok && want_feat) {
                rc = ihg_features(s->g, s->field_new, feat);
                if (rc) return rc;
            }
            ih_dp_record rec;
            rec.t = s->t;
            rec.dt = tried;
            rec.err = err;
            rec.rho_mean = rho_mean;
            rec.a = feat[0];
            rec.var_phi = feat[1];
            rec.mean_psi2 = feat[2];
            rec.accepted = ok;
            if (opts->log && attempts < opts->log_cap) opts->log[attempts] = rec;
            /* the last IH_HISTORY attempts, oldest first */
            if (s->n_hist == IH_HISTORY) {
                memmove(hist, hist + 1, (IH_HISTORY - 1) * sizeof hist[0]);
                s->n_hist--;
            }
            hist[s->n_hist++] = rec;
            attempts++;
            if (ok && opts->proposer && !opts->forced_dt && !after_reject) {
                double p = opts->proposer(opts->proposer_user, hist,
                                          s->n_hist, ctl.dt_next);
                if (p > 0.0 && isfinite(p)) {
                    if (c->max_dt_hubble_fraction > 0.0) {
                        const double cap =
                            c->max_dt_hubble_fraction * sqrt(3.0 / rho_mean);
                        if (p > cap) p = cap;
                    }
                    ctl.dt_next = p;
                }
            }
            if (ok) break;
            if (fabs(h) <= fabs(s->t) * DBL_EPSILON) {
                underflow = 1;
                break;
            }
        }
        if (underflow) {
            stop = IH_STOP_UNDERFLOW;
            break;
        }
        s->evo_flags.filter = filter;
        if ((n_stp + 1) % c->skip == 0) s->evo_flags.output = 1;
        const int output = s->evo_flags.output;
        /* A step with a step-based save and dense saves: the dense saves'
         * output calls replace the observables and rho of this output call,
         * so the call is made once more after them, on a copy of y_new as it
         * is now (before the filter changes it), in stage 5's array, which
         * the dense output leaves alone. */
        const int redo = output && opts->t_out &&
                         dense_due(opts, next_out, s->t, h);
        if (redo) {
            rc = ihg_stream_copy(s->g, s->stage[4], s->field_new, s->Ntot);
            if (rc) return rc;
        }
        ihg_set_field(s->g, s->field); /* summaries lag: they read `field` */
        rc = ih_mk_rhs(s, s->t + h, s->field_new, s->dfield_new);
        s->evo_flags.output = 0;
        s->evo_flags.filter = 0;
        if (rc) return rc;
        /* check_for_stiffness (src/dopri853.c:295-299): the FSAL derivative
         * against stage 12's, the filtered y_new against stage 12's input,
         * with the step's first trial size; n_ok before this step's count */
        if (c->enable_stiffness_check && ((n_ok % 10) == 0 || stiff.stiff > 0)) {
            double nd[2];
            rc = ihg_diff_norms(s->g, s->dfield_new, s->stage[11],
                                s->field_new, s->k_tmp, s->Ntot, nd);
            if (rc) return rc;
            if (ih_stiffness_update(&stiff, dt, nd[0], nd[1])) {
                if (stats) {
                    memset(stats, 0, sizeof *stats);
                    stats->n_stp = n_stp;
                    stats->n_ok = n_ok;
                    stats->n_bad = n_bad;
                    stats->attempts = attempts;
                    stats->t = s->t;
                    stats->stop = IH_STOP_STIFF;
                    stats->dtlamb = stiff.dtlamb;
                    stats->calls_rhs = ihg_calls_rhs(s->g);
                }
                return IH_ESTIFF;
            }
        }
        /* field = field_new, dfield = dfield_new (as a swap) */
        double *tmp = s->field;
        s->field = s->field_new;
        s->field_new = tmp;
        tmp = s->dfield;
        s->dfield = s->dfield_new;
        s->dfield_new = tmp;
        ihg_set_field(s->g, s->field);
        const double t_old = s->t;
        s->t += h;
        if (opts->t_out) {
            rc = dense_saves(s, opts, &next_out, t_old, h, output, save, user);
            if (rc) return rc;
        }
        if (redo) {
            /* the same call as above on the same input: its observables,
             * rho and filtered state are bit for bit the first call's */
            s->evo_flags.filter = filter;
            s->evo_flags.output = 1;
            ihg_set_field(s->g, s->field_new); /* the state before the step */
            rc = ih_mk_rhs(s, s->t, s->stage[4], s->k_tmp);
            s->evo_flags.output = 0;
            s->evo_flags.filter = 0;
            ihg_set_field(s->g, s->field);
            if (rc) return rc;
        }
        if (h == dt)
            ++n_ok;
        else
            ++n_bad;
        if (output) {
            rc = do_save(s, save, user, s->t);
            if (rc) return rc;
            last_save_t = s->t;
        }
        if (s->t >= c->tf) {
            stop = IH_STOP_FINAL_TIME;
            break;
        }
        if (n_stp >= c->max_steps) {
            stop = IH_STOP_MAX_STEPS;
            break;
        }
        if (c->max_runtime > 0.0 &&
            wall_time() - s->t_created > c->max_runtime) {
            stop = IH_STOP_MAX_RUNTIME;
            break;
        }
        if (!opts->forced_dt && fabs(ctl.dt_next) <= c->dt_min) {
            stop = IH_STOP_DT_MIN;
            break;
        }
        if (!opts->forced_dt) dt = ctl.dt_next;
    }
    rc = ihg_sync(s->g);
    if (rc) return rc;
    const double t_end = wall_time();
    /* wrap_up_dopri853, src/dopri853.c:223-231: the final slice when the
     * file buffer holds slices and the last one is older than t (without a
     * file: whenever the last save is older than t), then the followup */
    const int pending = opts->buffered ? opts->buffered(opts->file) != 0 : 1;
    if (pending && last_save_t < s->t) {
        rc = prepare_and_save(s, save, user);
        if (rc) return rc;
    }
    if (c->enable_followup && opts->followup) {
        rc = opts->followup(opts->file);
        if (rc) return rc;
    }
    if (stats) {
        stats->n_stp = n_stp;
        stats->n_ok = n_ok;
        stats->n_bad = n_bad;
        stats->attempts = attempts;
        stats->t = s->t;
        stats->dt_next = ctl.dt_next;
        stats->err_old = ctl.err_old;
        stats->integration_seconds = t_end - t_start;
        stats->calls_rhs = ihg_calls_rhs(s->g);
        stats->stop = stop;
        stats->dtlamb = stiff.dtlamb;
    }
    return attempts;
}

/* ------------------------------------------------------------------------ */
/* GSL steppers, src/gsl_stepper.c:41-76 over GSL's gsl_odeiv2 driver       */
/* ------------------------------------------------------------------------ */

/* the standard control's hadjust (gsl_odeiv2_control_y_new): decrease by
 * S r^(-1/q), at most 5 times; increase by S r^(-1/(q+1)), at most 5 times
 * and never below 1; S = 0.9, q the stepper's order */
int ih_gsl_hadjust(double rmax, double *h, int order)
{
    const double S = 0.9;
    const double h_old = *h;
    if (rmax > 1.1) {
        double r = S / pow(rmax, 1.0 / order);
        if (r < 0.2) r = 0.2;
        *h = r * h_old;
        return -1;
    }
    if (rmax < 0.5) {
        double r = S / pow(rmax, 1.0 / (order + 1.0));
        if (r > 5.0) r = 5.0;
        if (r < 1.0) r = 1.0;
        *h = r * h_old;
        return 1;
    }
    return 0;
}

/* gsl_odeiv2_evolve_apply: one accepted step from s->t towards t1 with the
 * suggested step *h (updated).  k[0] = f(t0, y) once; a step past t1 is cut
 * to land on t1; the control's DEC verdict undoes the attempt and retries
 * with the smaller step unless it would not move t by one ulp. */
static int gsl_evolve(ih_sim *s, int method, int stages, int order,
                      const double *cc, double t1, double *h,
                      ih_gsl_stats *st)
{
    const ih_config *c = &s->cfg;
    const double t0 = s->t, dt = t1 - t0;
    const double *k[13];
    k[0] = s->dfield;
    for (int j = 1; j < stages; ++j) k[j] = s->stage[j - 1];
    int rc = ih_mk_rhs(s, t0, s->field, s->dfield);
    double h0 = *h;
    for (;;) {
        int final_step = h0 > dt;
        if (final_step) h0 = dt;
        for (int j = 2; j <= stages && !rc; ++j) {
            rc = ihg_gsl_stage(s->g, method, j, h0, s->field, k, s->k_tmp);
            if (!rc) rc = ih_mk_rhs(s, t0 + cc[j - 1] * h0, s->k_tmp, s->stage[j - 2]);
        }
        double r = 0.0;
        if (!rc) rc = ihg_gsl_finish(s->g, method, h0, s->field, k, s->field_new,
                                     c->atol, c->rtol, &r);
        if (rc) return rc;
        ++st->attempts;
        const double t_new = final_step ? t1 : t0 + h0;
        const double h_old = h0;
        if (ih_gsl_hadjust(r, &h0, order) < 0) {
            volatile double t_curr = t_new, t_next = t_new + h0;
            if (fabs(h0) < fabs(h_old) && t_next != t_curr) {
                ++st->failed_steps;
                continue;
            }
            h0 = h_old; /* the tolerance is out of reach: keep the step */
        }
        double *tmp = s->field;
        s->field = s->field_new;
        s->field_new = tmp;
        ihg_set_field(s->g, s->field);
        s->t = t_new;
        *h = h0;
        ++st->steps;
        return 0;
    }
}

long ih_run_gsl(ih_sim *s, long n_output, ih_save_fn save, void *user,
                ih_gsl_stats *stats)
{
    const ih_config *c = &s->cfg;
    const int method = c->integration_method;
    if (method < IH_RKF45 || method > IH_DOPRI89 || n_output < 1)
        return IHG_EINVAL;
    if (!s->stage[11]) return IHG_EINVAL;
    int info[3];
    double cc[13], a[13 * 12], b[13], w[13];
    int rc = ihg_gsl_tableau(method, info, cc, a, b, w);
    if (rc) return rc;
    const int stages = info[0], order = info[1];
    ih_gsl_stats st;
    memset(&st, 0, sizeof st);
    double h = c->dt;
    rc = prepare_and_save(s, save, user);
    if (!rc) rc = ihg_sync(s->g);
    if (rc) return rc;
    const double t_start = wall_time();
    const double dt_out = (c->tf - c->ti) / (double)n_output;
    for (long i = 1; i <= n_output && st.status == IH_GSL_SUCCESS; ++i) {
        /* gsl_odeiv2_driver_apply from s->t to t1 */
        const double t1 = (double)i * dt_out;
        if (t1 - s->t < 0.0) {
            st.status = IH_GSL_EINVAL;
            break;
        }
        while (t1 - s->t > 0.0) {
            rc = gsl_evolve(s, method, stages, order, cc, t1, &h, &st);
            if (rc) return rc;
            if (fabs(h) < c->dt_min) {
                st.status = IH_GSL_ENOPROG;
                break;
            }
        }
        if (st.status != IH_GSL_SUCCESS) break;
        ++st.intervals;
        rc = prepare_and_save(s, save, user);
        if (rc) return rc;
    }
    rc = ihg_sync(s->g);
    if (rc) return rc;
    st.integration_seconds = wall_time() - t_start;
    st.t = s->t;
    st.h = h;
    st.calls_rhs = ihg_calls_rhs(s->g);
    if (stats) *stats = st;
    return st.steps;
}

/* ------------------------------------------------------------------------ */
/* learned step proposer                                                    */
/* ------------------------------------------------------------------------ */

void ih_proposer_features(const ih_dp_record *history, long n,
                          double x[IHG_MLP_IN])
{
    for (int j = 0; j < IH_HISTORY; ++j) {
        long i = n - IH_HISTORY + j;
        if (i < 0) i = 0;
        const ih_dp_record *r = &history[i];
        x[j] = log10(IH_MAX(r->err, 1.0e-12));
        x[IH_HISTORY + j] = log10(r->dt);
    }
    const ih_dp_record *r = &history[n - 1];
    x[16] = log10(r->a);
    x[17] = log10(sqrt(r->rho_mean / 3.0));
    x[18] = log10(r->var_phi + 1.0e-300);
    x[19] = log10(r->mean_psi2 + 1.0e-300);
}

double ih_mlp_propose(void *user, const ih_dp_record *history, long n,
                      double dt_classic)
{
    (void)dt_classic;
    ih_mlp_proposer *p = (ih_mlp_proposer *)user;
    if (!p || !p->mlp || n < 1) return 0.0;
    double raw[IHG_MLP_IN];
    float x[IHG_MLP_IN], y;
    ih_proposer_features(history, n, raw);
    for (int i = 0; i < IHG_MLP_IN; ++i) {
        if (!isfinite(raw[i])) return 0.0;
        x[i] = (float)((raw[i] - p->in_mean[i]) / p->in_std[i]);
    }
    if (ihg_mlp_forward(p->mlp, x, &y, 1)) return 0.0;
    double lr = (double)y * p->out_std + p->out_mean;
    const double lo = log(p->min_ratio), hi = log(p->max_ratio);
    if (!isfinite(lr)) return 0.0;
    if (lr < lo) lr = lo;
    if (lr > hi) lr = hi;
    p->calls++;
    return history[n - 1].dt * exp(lr);
}

/* ------------------------------------------------------------------------ */
/* ensembles: fixed-step RK4 of many realizations of one grid               */
/* ------------------------------------------------------------------------ */

struct ih_ens {
    ih_config cfg;
    ihg_ctx *g;
    int nreal;
    long stride;            /* doubles between realizations' states */
    size_t N, M, Next, Ntot;
    double *y, *k[4], *tmp, *tmp2; /* nreal * stride doubles each, device */
    double t;
    unsigned long calls;    /* RHS evaluations of all realizations */
};

int ih_ens_create(const ih_config *cfg, int nreal, ih_ens **out)
{
    if (!cfg || !out || nreal < 1) return IHG_EINVAL;
    *out = NULL;
    ih_ens *e = calloc(1, sizeof *e);
    if (!e) return IHG_ENOMEM;
    e->cfg = *cfg;
    e->nreal = nreal;
    int rc = ihg_create_ensemble(&cfg->grid, nreal, &e->g);
    if (rc) {
        free(e);
        return rc;
    }
    size_t sz[4];
    long ens[2];
    ihg_sizes(e->g, sz);
    ihg_ensemble(e->g, ens);
    e->N = sz[0];
    e->M = sz[1];
    e->Next = sz[2];
    e->Ntot = sz[3];
    e->stride = ens[1];
    const size_t n = (size_t)nreal * e->stride;
    rc = ihg_malloc(e->g, &e->y, n);
    for (int i = 0; !rc && i < 4; ++i) rc = ihg_malloc(e->g, &e->k[i], n);
    if (!rc) rc = ihg_malloc(e->g, &e->tmp, n);
    if (!rc) rc = ihg_malloc(e->g, &e->tmp2, n);
    if (rc) {
        ih_ens_destroy(e);
        return rc;
    }
    e->t = cfg->ti;
    *out = e;
    return 0;
}

void ih_ens_destroy(ih_ens *e)
{
    if (!e) return;
    if (e->g) {
        ihg_free(e->g, e->y);
        for (int i = 0; i < 4; ++i) ihg_free(e->g, e->k[i]);
        ihg_free(e->g, e->tmp);
        ihg_free(e->g, e->tmp2);
        ihg_destroy(e->g);
    }
    free(e);
}

ihg_ctx *ih_ens_gpu(ih_ens *e) { return e->g; }
double *ih_ens_state(ih_ens *e) { return e->y; }
long ih_ens_stride(const ih_ens *e) { return e->stride; }
int ih_ens_nreal(const ih_ens *e) { return e->nreal; }
double ih_ens_time(const ih_ens *e) { return e->t; }
unsigned long ih_ens_calls_rhs(const ih_ens *e) { return e->calls; }

int ih_ens_init(ih_ens *e, const uint64_t *seeds, const double *amplitudes,
                double phi0, double dphi0, double k_cut)
{
    int rc = 0;
    for (int r = 0; !rc && r < e->nreal; ++r) {
        double *f = e->y + (size_t)r * e->stride;
        rc = ihg_init_gaussian(e->g, f, seeds[r], amplitudes[r], phi0, dphi0,
                               e->cfg.a_initial, k_cut);
        /* mk_initial_psi (src/setup.c:825-903) of this realization */
        if (!rc) rc = ihg_set_field(e->g, f);
        if (!rc) rc = ihg_mk_initial_psi(e->g, e->cfg.ti, f, e->k[0]);
        e->calls += 1;
    }
    if (!rc) rc = ihg_set_field(e->g, NULL);
    e->t = e->cfg.ti;
    return rc;
}

/* one step of run_rk4 (src/rk4.c:78-102) for every realization; the filter
 * acts at the first stage, as there */
/* One RK4 step of run_rk4 (src/rk4.c:78-102) on every realization: the
   stage sums are folded into the RHS kernels that write the slopes
   (ihg_mk_rhs_rk4), the stage inputs alternating between tmp and tmp2; the
   values are those of mk_rhs + the separate stage sweeps, bit for bit */
static int ens_step(ih_ens *e, double t, double dt)
{
    ihg_ctx *g = e->g;
    const unsigned f1 = e->cfg.grid.enable_filter ? IHG_FILTER : 0u;
    ihg_rk4_ep st = {IHG_RK4_STAGE, 1, dt, e->y, e->tmp, NULL, NULL, NULL};
    int rc = ihg_mk_rhs_rk4(g, t, e->y, e->k[0], f1, &st);
    st.out = e->tmp2;
    if (!rc) rc = ihg_mk_rhs_rk4(g, t + dt / 2.0, e->tmp, e->k[1], 0, &st);
    st.half = 0;
    st.out = e->tmp;
    if (!rc) rc = ihg_mk_rhs_rk4(g, t + dt / 2.0, e->tmp2, e->k[2], 0, &st);
    const ihg_rk4_ep fin = {IHG_RK4_FINAL, 0, dt, e->y, NULL,
                            e->k[0], e->k[1], e->k[2]};
    if (!rc) rc = ihg_mk_rhs_rk4(g, t + dt, e->tmp, e->k[3], 0, &fin);
    return rc;
}

int ih_ens_capture_step(ih_ens *e, double dt, ihg_graph **out)
{
    /* the captured step may be replayed on another context's stream (a
       joined graph): the ensemble's earlier work must be finished first */
    int rc = ihg_sync(e->g);
    if (rc) return rc;
    rc = ihg_graph_begin(e->g);
    if (rc) return rc;
    rc = ens_step(e, e->t, dt);
    ihg_graph *g = NULL;
    const int rc2 = ihg_graph_end(e->g, &g);
    if (rc || rc2) {
        ihg_graph_destroy(g);
        return rc ? rc : rc2;
    }
    *out = g;
    return 0;
}

int ih_ens_run_rk4(ih_ens *const *es, int n, long nsteps, double dt,
                   int use_graph)
{
    if (!es || n < 1 || nsteps < 0) return IHG_EINVAL;
    int rc = 0;
    /* every ensemble's earlier work (initial data, observables) finished:
       the joined graph runs on the first ensemble's stream */
    for (int i = 0; !rc && i < n; ++i) rc = ihg_sync(es[i]->g);
    if (!use_graph) {
        for (long s = 0; !rc && s < nsteps; ++s)
            for (int i = 0; !rc && i < n; ++i) {
                rc = ens_step(es[i], es[i]->t, dt);
                es[i]->t += dt;
                es[i]->calls += 4UL * es[i]->nreal;
            }
    } else {
        /* one graph per ensemble, each replayed on its own ensemble's
           stream, so the sizes run side by side as in the eager form (one
           joined graph on a single stream ran its branches one after the
           other: 0.58x the eager rate for 48 x 16^3 + 8 x 32^3) */
        ihg_graph **parts = calloc(n, sizeof *parts);
        if (!parts) return IHG_ENOMEM;
        for (int i = 0; !rc && i < n; ++i)
            rc = ih_ens_capture_step(es[i], dt, &parts[i]);
        for (long s = 0; !rc && s < nsteps; ++s) {
            for (int i = 0; !rc && i < n; ++i) {
                rc = ihg_graph_launch(parts[i], es[i]->g);
                es[i]->t += dt;
                es[i]->calls += 4UL * es[i]->nreal;
            }
        }
        for (int i = 0; i < n; ++i) ihg_graph_destroy(parts[i]);
        free(parts);
    }
    for (int i = 0; !rc && i < n; ++i) rc = ihg_sync(es[i]->g);
    return rc;
}

/* one RK4 step of every ensemble of a group (ens_step's calls, each pass of
   all ensembles one launch); the arrays hold one pointer per ensemble */
typedef struct {
    double **y, **k[4], **tmp, **tmp2;
    ihg_rk4_ep *ep; /* one per ensemble */
} ens_arrays;

static int group_step(ihg_group *grp, const ens_arrays *a, int n, unsigned f1,
                      double t, double dt)
{
    for (int i = 0; i < n; ++i)
        a->ep[i] = (ihg_rk4_ep){IHG_RK4_STAGE, 1, dt, a->y[i], a->tmp[i],
                                NULL, NULL, NULL};
    int rc = ihg_group_mk_rhs_rk4(grp, t, a->y, a->k[0], f1, a->ep);
    for (int i = 0; i < n; ++i) a->ep[i].out = a->tmp2[i];
    if (!rc) rc = ihg_group_mk_rhs_rk4(grp, t + dt / 2.0, a->tmp, a->k[1], 0, a->ep);
    for (int i = 0; i < n; ++i) {
        a->ep[i].half = 0;
        a->ep[i].out = a->tmp[i];
    }
    if (!rc) rc = ihg_group_mk_rhs_rk4(grp, t + dt / 2.0, a->tmp2, a->k[2], 0, a->ep);
    for (int i = 0; i < n; ++i)
        a->ep[i] = (ihg_rk4_ep){IHG_RK4_FINAL, 0, dt, a->y[i], NULL,
                                a->k[0][i], a->k[1][i], a->k[2][i]};
    if (!rc) rc = ihg_group_mk_rhs_rk4(grp, t + dt, a->tmp, a->k[3], 0, a->ep);
    return rc;
}

/* the group over the ensembles' contexts and the pointer tables of their
   arrays (one allocation, freed by group_arrays_free) */
static int group_arrays(ih_ens *const *es, int n, ihg_group **grp,
                        ens_arrays *a, unsigned *f1)
{
    ihg_ctx **ctxs = calloc((size_t)n, sizeof *ctxs);
    double **ptrs = calloc((size_t)n * 7, sizeof *ptrs);
    ihg_rk4_ep *eps = calloc((size_t)n, sizeof *eps);
    if (!ctxs || !ptrs || !eps) {
        free(ctxs);
        free(ptrs);
        free(eps);
        return IHG_ENOMEM;
    }
    a->y = ptrs;
    for (int j = 0; j < 4; ++j) a->k[j] = ptrs + (size_t)(j + 1) * n;
    a->tmp = ptrs + (size_t)5 * n;
    a->tmp2 = ptrs + (size_t)6 * n;
    a->ep = eps;
    /* the filter flag of ens_step: the group applies it to the ensembles
       whose grids have the filter on */
    *f1 = 0u;
    for (int i = 0; i < n; ++i) {
        ctxs[i] = es[i]->g;
        a->y[i] = es[i]->y;
        for (int j = 0; j < 4; ++j) a->k[j][i] = es[i]->k[j];
        a->tmp[i] = es[i]->tmp;
        a->tmp2[i] = es[i]->tmp2;
        if (es[i]->cfg.grid.enable_filter) *f1 = IHG_FILTER;
    }
    *grp = NULL;
    int rc = ihg_group_create(ctxs, n, grp);
    free(ctxs);
    if (rc) {
        free(ptrs);
        free(eps);
    }
    return rc;
}

static void group_arrays_free(ihg_group *grp, ens_arrays *a)
{
    ihg_group_destroy(grp);
    free(a->y);
    free(a->ep);
}

int ih_ens_capture_step_grouped(ih_ens *const *es, int n, double dt,
                                ihg_graph **out)
{
    if (!es || n < 1 || !out) return IHG_EINVAL;
    *out = NULL;
    ihg_group *grp;
    ens_arrays a;
    unsigned f1;
    int rc = group_arrays(es, n, &grp, &a, &f1);
    if (rc) return rc;
    /* every ensemble's earlier work finished: the group runs on the first
       ensemble's stream */
    for (int i = 0; !rc && i < n; ++i) rc = ihg_sync(es[i]->g);
    if (!rc) rc = ihg_graph_begin(es[0]->g);
    if (!rc) {
        const int rc1 = group_step(grp, &a, n, f1, es[0]->t, dt);
        const int rc2 = ihg_graph_end(es[0]->g, out);
        rc = rc1 ? rc1 : rc2;
        if (rc) {
            ihg_graph_destroy(*out);
            *out = NULL;
        }
    }
    group_arrays_free(grp, &a);
    return rc;
}

int ih_ens_run_rk4_grouped(ih_ens *const *es, int n, long nsteps, double dt,
                           int use_graph)
{
    if (!es || n < 1 || nsteps < 0) return IHG_EINVAL;
    ihg_graph *graph = NULL;
    ihg_group *grp = NULL;
    ens_arrays a;
    unsigned f1;
    int rc = 0;
    if (use_graph) {
        rc = ih_ens_capture_step_grouped(es, n, dt, &graph);
    } else {
        rc = group_arrays(es, n, &grp, &a, &f1);
        for (int i = 0; !rc && i < n; ++i) rc = ihg_sync(es[i]->g);
    }
    for (long s = 0; !rc && s < nsteps; ++s) {
        rc = graph ? ihg_graph_launch(graph, es[0]->g)
                   : group_step(grp, &a, n, f1, es[0]->t, dt);
        for (int i = 0; i < n; ++i) {
            es[i]->t += dt;
            es[i]->calls += 4UL * es[i]->nreal;
        }
    }
    if (!rc) rc = ihg_sync(es[0]->g);
    ihg_graph_destroy(graph);
    if (grp) group_arrays_free(grp, &a);
    return rc;
}

int ih_ens_observables(ih_ens *e, int r, double *obs, size_t len)
{
    if (r < 0 || r >= e->nreal) return IHG_EINVAL;
    double *f = e->y + (size_t)r * e->stride;
    /* prepare_and_save_timeslice (src/toolbox.c:96-104) on a copy, so the
       filter of the output call leaves the ensemble state alone */
    int rc = ihg_copy(e->g, e->tmp, f, e->Ntot);
    if (!rc) rc = ihg_set_field(e->g, e->tmp);
    if (!rc) rc = ihg_mk_rhs(e->g, e->t, e->tmp, e->k[0], IHG_FILTER | IHG_OUTPUT);
    if (!rc) rc = ihg_observables(e->g, obs, len);
    if (!rc) rc = ihg_set_field(e->g, NULL);
    e->calls += 1;
    return rc;
}
