/*
 * orc_timestep.c -- LEM timestep control restated: the median interior
 * gradient, the pre-run diagnosis, the adaptive controller and the
 * convergence monitor.  TEST INFRASTRUCTURE ONLY.
 *
 * Reference: src/cpu/lsdtt_lem_timestep_cpu.cpp
 *   median_interior_gradient   :28-50 (cardinal central differences of the
 *       valid interior, src/cpu/lsdtt_raster_central_difference_cpu.hpp:43-62;
 *       median as in src/utils/lsdtt_statistics.cpp:365-376)
 *   diagnose_lem_timestep      :54-98
 *   lsdtt_lem_timestep_controller :101-213
 *   lsdtt_lem_convergence_monitor :216-302
 * Options and defaults: src/cpu/lsdtt_lem_timestep_cpu.hpp:62-160, 236-292.
 *
 * The surface is FP64 here (the LEM state of this repo); the reference
 * reads float rasters widened to double, so the arithmetic is the same.
 *
 * Pins (tests/test_oracle_timestep.py): the reference's D1-D6
 * (tests/cpp/test_lem_timestep.cpp) on the same ramp and flat surfaces; the
 * median against numpy's on random surfaces with holes; the ladder and
 * hysteresis against hand-computed sequences; the monitor's converged,
 * dt-guard, stalled and oscillating outcomes.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static int cmp_double(const void* a, const void* b)
{
    const double x = *(const double*)a, y = *(const double*)b;
    return (x > y) - (x < y);
}

double orc_median_interior_gradient(int nr, int nc, double dx, const double* z, const uint8_t* mask)
{
    if (nr < 3 || nc < 3 || !(dx > 0.0)) return 0.0;
    double* g = (double*)malloc(sizeof(double) * (size_t)nr * (size_t)nc);
    int64_t m = 0;
    for (int r = 1; r < nr - 1; ++r)
        for (int c = 1; c < nc - 1; ++c) {
            const int64_t f = (int64_t)r * nc + c;
            if (!mask[f]) continue;
            const int64_t fn = f - nc, fs = f + nc, fw = f - 1, fe = f + 1;
            if (!mask[fn] || !mask[fs] || !mask[fw] || !mask[fe]) continue;
            const double gx = (z[fe] - z[fw]) / (2.0 * dx);
            const double gy = (z[fs] - z[fn]) / (2.0 * dx);
            g[m++] = sqrt(gx * gx + gy * gy);
        }
    double med = 0.0;
    if (m > 0) {
        qsort(g, (size_t)m, sizeof(double), cmp_double);
        const int64_t mid = m / 2;
        med = (m % 2) ? g[mid] : 0.5 * (g[mid - 1] + g[mid]);
    }
    free(g);
    return med;
}

void orc_diagnose_timestep(int nr, int nc, double dx, const double* z, const uint8_t* mask, double U, double dt,
                           double target, orc_ts_diagnosis* out)
{
    memset(out, 0, sizeof(*out));
    out->reference_slope = orc_median_interior_gradient(nr, nc, dx, z, mask);
    if (!(out->reference_slope > 1e-6) || !(dx > 0.0) || !(U > 0.0) || !(dt > 0.0)) return;
    out->cell_relief_m = dx * out->reference_slope;
    out->uplift_per_step_m = U * dt;
    out->ratio = out->uplift_per_step_m / out->cell_relief_m;
    out->suggested_dt_yr = target * out->cell_relief_m / U;
    out->too_large = out->ratio > target;
}

void orc_ts_controller_init(orc_ts_controller* c, double dt_initial, const orc_ts_options* o)
{
    memset(c, 0, sizeof(*c));
    c->o = *o;
    c->dt = c->dt_raw = c->dt_initial = dt_initial;
    c->dt_min = o->dt_min_yr > 0.0 ? o->dt_min_yr : dt_initial / 100.0;
    c->dt_max = o->dt_max_yr > 0.0 ? o->dt_max_yr : dt_initial * 100.0;
}

double orc_ts_controller_update(orc_ts_controller* c, double max_dz, double slope_now, double dx)
{
    const int every = c->o.slope_refresh_interval > 1 ? c->o.slope_refresh_interval : 1;
    if (c->n_steps % every == 0) c->slope = slope_now;
    ++c->n_steps;
    const double relief = dx * c->slope;
    if (!(relief > 1e-9) || !(max_dz > 0.0)) return c->dt;
    c->last_ratio = max_dz / relief;
    c->ratio_ema = c->ratio_ema > 0.0
                       ? (1.0 - c->o.ratio_smoothing) * c->ratio_ema + c->o.ratio_smoothing * c->last_ratio
                       : c->last_ratio;
    const double rd = c->ratio_ema;
    const double before = c->dt;
    if (rd > c->o.target_ratio) {
        const double sc = c->o.shrink_factor > c->o.target_ratio / rd ? c->o.shrink_factor : c->o.target_ratio / rd;
        const double v = c->dt_raw * sc;
        c->dt_raw = v > c->dt_min ? v : c->dt_min;
        ++c->n_shrinks;
    } else if (rd < 0.5 * c->o.target_ratio) {
        const double v = c->dt_raw * c->o.grow_factor;
        c->dt_raw = v < c->dt_max ? v : c->dt_max;
        ++c->n_grows;
    }
    c->dt = c->dt_raw;
    if (c->o.ladder_ratio > 1.0 && c->dt_initial > 0.0) {
        const double q = c->o.ladder_ratio;
        if (rd > c->o.target_ratio && c->rung > c->lowest_rung - 1) --c->rung;
        while (c->dt_raw >= c->dt_initial * pow(q, c->rung + 1)) ++c->rung;
        while (c->dt_raw <= c->dt_initial * pow(q, c->rung - 1)) --c->rung;
        if (c->rung < c->lowest_rung) c->lowest_rung = c->rung;
        double d = c->dt_initial * pow(q, c->rung);
        if (d < c->dt_min) d = c->dt_min;
        if (d > c->dt_max) d = c->dt_max;
        c->dt = d;
    }
    if (c->dt != before) ++c->n_changes;
    return c->dt;
}

void orc_ts_monitor_init(orc_ts_monitor* m, double dt_initial, const orc_ts_monitor_options* o)
{
    memset(m, 0, sizeof(*m));
    m->o = *o;
    m->dt_initial = dt_initial;
    m->decay = 1.0;
    const int w = o->stall_window > 8 ? o->stall_window : 8;
    m->window = w;
    m->hist = (double*)calloc((size_t)w, sizeof(double));
}

void orc_ts_monitor_free(orc_ts_monitor* m)
{
    free(m->hist);
    m->hist = NULL;
}

/* returns 0 running, 1 converged, 2 stalled */
int orc_ts_monitor_update(orc_ts_monitor* m, double max_dz, double dt_used)
{
    ++m->n_steps;
    m->last_dz = max_dz;
    if (m->n_hist < m->window) {
        m->hist[m->n_hist++] = max_dz;
    } else {
        memmove(m->hist, m->hist + 1, sizeof(double) * (size_t)(m->window - 1));
        m->hist[m->window - 1] = max_dz;
    }
    const int dt_ok = m->dt_initial <= 0.0 || dt_used >= m->o.min_dt_fraction * m->dt_initial;
    const int below_abs = m->o.dz_tolerance_m <= 0.0 || max_dz < m->o.dz_tolerance_m;
    const int below_rate = m->o.dz_rate_tolerance_m_per_yr <= 0.0 || dt_used <= 0.0 ||
                           max_dz / dt_used < m->o.dz_rate_tolerance_m_per_yr;
    if (below_abs && below_rate && dt_ok) ++m->n_below;
    else m->n_below = 0;
    if (m->n_below >= m->o.patience_steps) {
        m->state = 1;
        return m->state;
    }
    if (m->n_hist < m->window) return m->state;
    const int q = m->window / 5 > 1 ? m->window / 5 : 1;
    double early = m->hist[0], late = m->hist[m->window - q];
    for (int i = 1; i < q; ++i) if (m->hist[i] < early) early = m->hist[i];
    for (int i = m->window - q + 1; i < m->window; ++i) if (m->hist[i] < late) late = m->hist[i];
    m->decay = early > 0.0 ? late / early : 1.0;
    if (m->decay < m->o.stall_decay_threshold) return m->state;
    long alt = 0;
    for (int i = 2; i < m->window; ++i) {
        const double d1 = m->hist[i - 1] - m->hist[i - 2], d2 = m->hist[i] - m->hist[i - 1];
        if (d1 * d2 < 0.0) ++alt;
    }
    m->oscillating = (double)alt / (double)(m->window - 2) > m->o.oscillation_alternation_fraction;
    m->state = 2;
    return m->state;
}
