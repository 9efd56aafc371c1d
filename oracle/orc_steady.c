/*
 * orc_steady.c -- analytic fluvial steady state on the D8 tree restated.
 * TEST INFRASTRUCTURE ONLY.
 *
 * Reference: src/cpu/lsdtt_steady_state_cpu.cpp
 *   make_chi_weights                       :37-72
 *   compute_fluvial_steady_state_on_tree   :76-170
 *   compute_d8_fluvial_steady_state        :173-225
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
