/*
 * instaheat_params.h -- an ih_config from the reference's compile-time
 * parameters.
 *
 * The reference configures a run by macros: `configure` copies
 * src/main_template.h to main.h and replaces its placeholders
 * (src/main_template.h:116-162) with the values of src/parameters.sh; a
 * switch ENABLE_X = "1" becomes `#define ENABLE_X`, "0" becomes
 * `#define NOT_ENABLE_X`.  A host that includes such a main.h (or defines
 * the same macros itself) and then this header gets
 *
 *     ih_config cfg;
 *     ih_config_from_macros(&cfg);
 *
 * with every parameter the hot path reads taken from the macros under the
 * reference's names.  A macro that is not defined keeps the value of
 * ih_config_defaults (the defaults of src/parameters.sh); the switches are
 * read only when a parameter set is present (GRIDPOINTS_X defined).
 * INTEGRATION_METHOD takes the reference's method names (RK4 ... DOPRI853,
 * src/main_template.h:51-56), defined here when main.h is not included.
 */
#ifndef INSTAHEAT_PARAMS_H
#define INSTAHEAT_PARAMS_H

#include "instaheat.h"

#ifndef RK4
#define RK4 (0)
#endif
#ifndef RKF45
#define RKF45 (1)
#endif
#ifndef RKCK
#define RKCK (2)
#endif
#ifndef DOPRI89
#define DOPRI89 (3)
#endif
#ifndef DOPRI853
#define DOPRI853 (4)
#endif

static inline void ih_config_from_macros(ih_config *cfg)
{
    ih_config_defaults(cfg);
#ifdef GRIDPOINTS_X
    cfg->grid.nx = (int)(GRIDPOINTS_X);
#endif
#ifdef GRIDPOINTS_Y
    cfg->grid.ny = (int)(GRIDPOINTS_Y);
#endif
#ifdef GRIDPOINTS_Z
    cfg->grid.nz = (int)(GRIDPOINTS_Z);
#endif
#ifdef SPATIAL_LOWER_BOUND_X
    cfg->grid.xa = (double)(SPATIAL_LOWER_BOUND_X);
#endif
#ifdef SPATIAL_UPPER_BOUND_X
    cfg->grid.xb = (double)(SPATIAL_UPPER_BOUND_X);
#endif
#ifdef SPATIAL_LOWER_BOUND_Y
    cfg->grid.ya = (double)(SPATIAL_LOWER_BOUND_Y);
#endif
#ifdef SPATIAL_UPPER_BOUND_Y
    cfg->grid.yb = (double)(SPATIAL_UPPER_BOUND_Y);
#endif
#ifdef SPATIAL_LOWER_BOUND_Z
    cfg->grid.za = (double)(SPATIAL_LOWER_BOUND_Z);
#endif
#ifdef SPATIAL_UPPER_BOUND_Z
    cfg->grid.zb = (double)(SPATIAL_UPPER_BOUND_Z);
#endif
#ifdef MASS
    cfg->grid.mass = (double)(MASS);
#endif
#ifdef POWER_SPECTRUM_BINS
    cfg->grid.bins = (int)(POWER_SPECTRUM_BINS);
#endif
#ifdef INTEGRATION_METHOD
    cfg->integration_method = (int)(INTEGRATION_METHOD);
#endif
#ifdef INITIAL_TIME
    cfg->ti = (double)(INITIAL_TIME);
#endif
#ifdef FINAL_TIME
    cfg->tf = (double)(FINAL_TIME);
#endif
#ifdef DELTA_T
    cfg->dt = (double)(DELTA_T);
#endif
#ifdef MINIMAL_DELTA_T
    cfg->dt_min = (double)(MINIMAL_DELTA_T);
#endif
#ifdef MAX_STEPS
    cfg->max_steps = (long)(MAX_STEPS);
#endif
#ifdef BETA
    cfg->beta = (double)(BETA);
#endif
#ifdef SAFE
    cfg->safe = (double)(SAFE);
#endif
#ifdef SMALLEST_SCALING
    cfg->minscale = (double)(SMALLEST_SCALING);
#endif
#ifdef LARGEST_SCALING
    cfg->maxscale = (double)(LARGEST_SCALING);
#endif
#ifdef RELATIVE_TOLERANCE
    cfg->rtol = (double)(RELATIVE_TOLERANCE);
#endif
#ifdef ABSOLUTE_TOLERANCE
    cfg->atol = (double)(ABSOLUTE_TOLERANCE);
#endif
#ifdef MAX_DT_HUBBLE_FRACTION
    cfg->max_dt_hubble_fraction = (double)(MAX_DT_HUBBLE_FRACTION);
#endif
#ifdef TIME_STEP_SKIPS
    cfg->skip = (long)(TIME_STEP_SKIPS);
#endif
#ifdef A_INITIAL
    cfg->a_initial = (double)(A_INITIAL);
#endif
#ifdef MAX_RUNTIME
    cfg->max_runtime = (double)(MAX_RUNTIME);
#endif
#ifdef GRIDPOINTS_X
    /* the switches of a parameter set: defined or NOT_ */
#ifdef ENABLE_GW
    cfg->grid.enable_gw = 1;
#else
    cfg->grid.enable_gw = 0;
#endif
#ifdef ENABLE_FFT_FILTER
    cfg->grid.enable_filter = 1;
#else
    cfg->grid.enable_filter = 0;
#endif
#ifdef ENABLE_FOLLOWUP
    cfg->enable_followup = 1;
#else
    cfg->enable_followup = 0;
#endif
#ifdef ENABLE_STIFFNESSCHECK
    cfg->enable_stiffness_check = 1;
#else
    cfg->enable_stiffness_check = 0;
#endif
#endif
}

#endif
