/*
 * orc_steady.c -- analytic fluvial steady state on the D8 tree and on the
 * D-infinity graph restated.  TEST INFRASTRUCTURE ONLY.
 *
 * Reference: src/cpu/lsdtt_steady_state_cpu.cpp
 *   make_chi_weights                       :37-72
 *   compute_fluvial_steady_state_on_tree   :76-170
 *   compute_d8_fluvial_steady_state        :173-225
 *   compute_dinf_fluvial_steady_state      :227-436
 * Options: src/cpu/lsdtt_steady_state_cpu.hpp:62-89 (float fields).
 *
 * Float arithmetic as the reference's.  powf: with libm_pow = 1 the C
 * library's powf (the reference's call); with 0 the correctly rounded value
 * (float)pow((double)x, (double)y), which is what the device computes.
 *
 * Pins (tests/test_oracle_steady.py): the steady state satisfies
 * U = K A^m S^n on every non-outlet pixel along its receiver (the defining
 * property), a 1-D channel against the closed form of chi, relief mode
 * hits target_relief exactly, the erodibility field's weighted chi keeps
 * U = K_i A^m S^n, and the oracle's own D8 stream power leaves the surface
 * still under uplift (one implicit step of U dt + incision returns it).
 * D-infinity (tests/test_oracle_steady_dinf.py): the reference's S1, S4,
 * S5/S5b (tests/cpp/test_lem_erodibility.cpp:774-987) and C7
 * (tests/cpp/test_chi_channel_head.cpp:250-272) restated, a north-draining
 * ramp against the closed form of chi and bit-equal to the D8 result, and a
 * diagonal plane against the dx sqrt2 sum.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static float powf_of(float x, float y, int libm)
{
    return libm ? powf(x, y) : (float)pow((double)x, (double)y);
}

int orc_steady_state_d8(int nr, int nc, double cell_size, const uint8_t* valid, const int64_t* receiver,
                        const uint8_t* dir, const int64_t* topo, int64_t n_topo, const double* area,
                        const float* K_field, const orc_ss_options* o, int libm_pow, float nodata, float* z_out,
                        orc_ss_result* res)
{
    const int64_t n = (int64_t)nr * nc;
    const float dx = (float)cell_size;
    const float dxd = dx * (float)1.4142135623730950488;
    float* w = NULL;
    memset(res, 0, sizeof(*res));
    res->erodibility_field_scale = 1.0f;
    if (K_field) {
        if (!(o->K > 0.0f) || !isfinite(o->K)) return ORC_ERR_ARG;
        w = (float*)malloc(sizeof(float) * (size_t)n);
        for (int64_t s = 0; s < n; ++s) {
            w[s] = 1.0f;
            if (!valid[s]) continue;
            const float k = K_field[s];
            if (!(k > 0.0f) || !isfinite(k)) {
                free(w);
                return ORC_ERR_ARG;
            }
            w[s] = powf_of(o->K / k, 1.0f / o->n, libm_pow);
        }
    }
    float* chi = (float*)calloc((size_t)(n ? n : 1), sizeof(float));
    /* step 1: reverse sweep (outlets first) */
    for (int64_t i = n_topo - 1; i >= 0; --i) {
        const int64_t s = topo[i];
        if (!valid[s]) continue;
        const int64_t r = receiver[s];
        if (r < 0 || r == s) continue;
        double A = area[s];
        if (o->divide_area_by_dx) A /= (double)dx;
        if (A <= 0.0) A = (double)dx * dx;
        const float ws = w ? w[s] : 1.0f;
        const float amn = powf_of((float)A, -o->m / o->n, libm_pow) * ws;
        const float dxe = (dir[s] & 1u) ? dxd : dx;
        chi[s] = chi[r] + dxe * amn;
    }
    /* step 2 */
    float mx = 0.0f;
    for (int64_t s = 0; s < n; ++s)
        if (valid[s] && chi[s] > mx) mx = chi[s];
    if (mx <= 0.0f) mx = 1.0f;
    /* step 3 */
    float f;
    if (o->relief_mode) {
        res->K_used = o->U * powf(mx / o->target_relief, o->n);
        f = o->target_relief / mx;
    } else {
        res->K_used = o->K;
        f = powf(o->U / o->K, 1.0f / o->n);
    }
    /* step 4 */
    for (int64_t s = 0; s < n; ++s) z_out[s] = valid[s] ? o->baselevel_elevation + f * chi[s] : nodata;
    res->max_unit_height = mx;
    res->computed_relief = f * mx;
    if (w) res->erodibility_field_scale = res->K_used / o->K;
    if (o->channel_threshold_area > 0.0f) {
        float cc = 0.0f;
        int64_t nch = 0;
        for (int64_t s = 0; s < n; ++s) {
            if (!valid[s]) continue;
            const double q = orc_channel_threshold_quantity(area[s], (double)dx, o->divide_area_by_dx);
            if (q < (double)o->channel_threshold_area) continue;
            ++nch;
            if (chi[s] > cc) cc = chi[s];
        }
        res->n_channel_pixels = nch;
        if (nch > 0) {
            res->max_unit_height_at_channel_head = cc;
            res->chi_c_reported = 1;
        }
    }
    free(chi);
    free(w);
    return ORC_OK;
}

/* compute_dinf_fluvial_steady_state (lsdtt_steady_state_cpu.cpp:227-436) on
 * the oracle's D-inf arrays: receivers first (reverse Kahn order, which
 * orders every receiver before its donors as the reference's reverse
 * levels do); per cell A = area (/ dx), amn = powf(A, -m/n) w; two
 * receivers (neither weight >= 1 - 1e-6): dchi = (chi(e2) - chi(e1)) / dx,
 * disc = amn^2 - dchi^2, chi = chi(e1) + dx sqrt(disc) when disc >= 0, else
 * chi(e1) + dx amn; one receiver (the diagonal slot when w1 >= 1 - 1e-6):
 * chi = chi(r) + dx_eff amn, dx_eff = dx sqrt2 when the receiver is one row
 * and one column away (a periodic wrap does not count as diagonal). */
int orc_steady_state_dinf(int nr, int nc, double cell_size, const uint8_t* valid, const int64_t* rec, const float* w,
                          const uint8_t* nrec, const int64_t* topo, int64_t n_topo, const double* area,
                          const float* K_field, const orc_ss_options* o, int libm_pow, float nodata, float* z_out,
                          orc_ss_result* res)
{
    const int64_t n = (int64_t)nr * nc;
    const float dx = (float)cell_size;
    const float dxd = dx * (float)1.4142135623730950488;
    float* wk = NULL;
    memset(res, 0, sizeof(*res));
    res->erodibility_field_scale = 1.0f;
    if (K_field) {
        if (!(o->K > 0.0f) || !isfinite(o->K)) return ORC_ERR_ARG;
        wk = (float*)malloc(sizeof(float) * (size_t)n);
        for (int64_t s = 0; s < n; ++s) {
            wk[s] = 1.0f;
            if (!valid[s]) continue;
            const float k = K_field[s];
            if (!(k > 0.0f) || !isfinite(k)) {
                free(wk);
                return ORC_ERR_ARG;
            }
            wk[s] = powf_of(o->K / k, 1.0f / o->n, libm_pow);
        }
    }
    float* chi = (float*)calloc((size_t)(n ? n : 1), sizeof(float));
    for (int64_t i = n_topo - 1; i >= 0; --i) {
        const int64_t s = topo[i];
        const int k = nrec[s];
        if (!valid[s] || k == 0) continue;
        double A = area[s];
        if (o->divide_area_by_dx) A /= (double)dx;
        if (A <= 0.0) A = (double)dx * dx;
        const float ws = wk ? wk[s] : 1.0f;
        const float amn = powf_of((float)A, -o->m / o->n, libm_pow) * ws;
        const float w0 = w[2 * s], w1 = k > 1 ? w[2 * s + 1] : 0.0f;
        const int two = k > 1 && !(w0 >= 1.0f - 1e-6f) && !(w1 >= 1.0f - 1e-6f);
        const float c0 = chi[rec[2 * s]];
        const float c1 = k > 1 ? chi[rec[2 * s + 1]] : 0.0f;
        float cn;
        if (two) {
            const float dchi = (c1 - c0) / dx;
            const float disc = amn * amn - dchi * dchi;
            cn = disc >= 0.0f ? c0 + dx * sqrtf(disc) : c0 + dx * amn;
        } else {
            const int k0 = (k > 1 && w1 >= 1.0f - 1e-6f) ? 1 : 0;
            const int64_t r = rec[2 * s + k0];
            const int64_t dr = llabs(r / nc - s / nc), dc = llabs(r % nc - s % nc);
            const float dxe = (dr == 1 && dc == 1) ? dxd : dx;
            cn = (k0 ? c1 : c0) + dxe * amn;
        }
        chi[s] = cn;
    }
    float mx = 0.0f;
    for (int64_t s = 0; s < n; ++s)
        if (valid[s] && chi[s] > mx) mx = chi[s];
    if (mx <= 0.0f) mx = 1.0f;
    float f;
    if (o->relief_mode) {
        res->K_used = o->U * powf(mx / o->target_relief, o->n);
        f = o->target_relief / mx;
    } else {
        res->K_used = o->K;
        f = powf(o->U / o->K, 1.0f / o->n);
    }
    for (int64_t s = 0; s < n; ++s) z_out[s] = valid[s] ? o->baselevel_elevation + f * chi[s] : nodata;
    res->max_unit_height = mx;
    res->computed_relief = f * mx;
    if (wk) res->erodibility_field_scale = res->K_used / o->K;
    if (o->channel_threshold_area > 0.0f) {
        float cc = 0.0f;
        int64_t nch = 0;
        for (int64_t s = 0; s < n; ++s) {
            if (!valid[s]) continue;
            const double q = orc_channel_threshold_quantity(area[s], (double)dx, o->divide_area_by_dx);
            if (q < (double)o->channel_threshold_area) continue;
            ++nch;
            if (chi[s] > cc) cc = chi[s];
        }
        res->n_channel_pixels = nch;
        if (nch > 0) {
            res->max_unit_height_at_channel_head = cc;
            res->chi_c_reported = 1;
        }
    }
    free(chi);
    free(wk);
    return ORC_OK;
}
