/*
 * orc_chi.c -- chi (integral of (A0/A)^theta along the flow path) on a D8
 * flow object.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates src/cpu/lsdtt_chi_extraction_cpu.cpp:
 *   chi_increment                 :154-163  step * 0.5 * (pow(A0/Au, t) + pow(A0/Ar, t))
 *   step_length_from_receiver_dir :166-179  dx (even dir) or dx * sqrt2 (odd)
 *   collect_flow_outlets          :182-211  valid cells with receiver -2, row-major
 *   compute_full_basin_chi        :283-365  s-vector slice of each outlet, chi[0] =
 *                                           seed, chi[i] = float(double(chi[rec]) + inc)
 *   compute_threshold_basin_chi   :368-497  same recursion on the pixels of the slice
 *                                           whose contributing count >= threshold
 *   compute_raster_from_seed_outlets :635-692 (first basin to reach a pixel wins)
 *   get_chi_raster                :739-765  A = s_vector_size * dx^2
 *   get_all_basin_outlet_flats    (src/cpu/lsdtt_flow_base.cpp:82-110) outlets with
 *                                           size >= threshold, row-major
 *
 * Parity: pinned by closed forms in tests/test_oracle_chi.py (theta = 0 gives
 * the float-accumulated flow length; a 1-D channel gives the trapezoid sum of
 * (A0/A)^theta with known counts) and by the invariant that thresholded chi
 * equals the full raster on the selected pixels.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static const double k_sqrt2 = 1.4142135623730950488;

static double chi_inc(double step, double au, double ar, double A0, double t)
{
    return step * 0.5 * (pow(A0 / au, t) + pow(A0 / ar, t));
}

int orc_chi_raster(int nr, int nc, double dx, const uint8_t* mask,
                   const int64_t* receiver, const uint8_t* dir,
                   const int64_t* s_vector, const int64_t* s_start,
                   const uint64_t* s_size, double m_over_n, double A0,
                   float nodata, float* chi, uint8_t* chi_mask)
{
    const int64_t n = (int64_t)nr * nc;
    const double pa = dx * dx;
    for (int64_t i = 0; i < n; ++i) {
        chi[i] = nodata;
        chi_mask[i] = 0;
    }
    for (int64_t o = 0; o < n; ++o) {
        if (!mask[o] || receiver[o] != -2) continue;
        const int64_t b = s_start[o];
        const int64_t sz = (int64_t)s_size[o];
        if (b < 0 || sz <= 0) return ORC_ERR_ARG;
        /* local chi array of the slice, indexed like the s-vector */
        float* loc = (float*)malloc(sizeof(float) * (size_t)sz);
        if (!loc) return ORC_ERR_ARG;
        loc[0] = 0.0f;
        for (int64_t i = 1; i < sz; ++i) {
            const int64_t f = s_vector[b + i];
            const int64_t r = receiver[f];
            if (r < 0) {
                free(loc);
                return ORC_ERR_ARG;
            }
            const int64_t rl = s_start[r] - b;
            if (rl < 0 || rl >= i || dir[f] == 255) {
                free(loc);
                return ORC_ERR_ARG;
            }
            const double step = (dir[f] % 2u == 0u) ? dx : dx * k_sqrt2;
            const double au = (double)s_size[f] * pa;
            const double ar = (double)s_size[r] * pa;
            loc[i] = (float)((double)loc[rl] + chi_inc(step, au, ar, A0, m_over_n));
        }
        for (int64_t i = 0; i < sz; ++i) {
            const int64_t f = s_vector[b + i];
            if (!chi_mask[f]) {
                chi[f] = loc[i];
                chi_mask[f] = 1;
            }
        }
        free(loc);
    }
    return ORC_OK;
}

int64_t orc_chi_threshold(int nr, int nc, double dx, const uint8_t* mask,
                          const int64_t* receiver, const uint8_t* dir,
                          const int64_t* s_vector, const int64_t* s_start,
                          const uint64_t* s_size, double m_over_n, double A0,
                          int64_t min_pixels, int64_t* flats, float* chi,
                          int64_t* receivers, int64_t* basin_start,
                          int64_t* n_basins)
{
    const int64_t n = (int64_t)nr * nc;
    const double pa = dx * dx;
    /* full-grid position of each selected pixel in the current basin */
    int64_t* pos = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    if (!pos) return -ORC_ERR_ARG;
    for (int64_t i = 0; i < n; ++i) pos[i] = -1;
    int64_t m = 0, nb = 0;
    for (int64_t o = 0; o < n; ++o) {
        if (!mask[o] || receiver[o] != -2) continue;
        if ((int64_t)s_size[o] < min_pixels) continue;
        const int64_t b = s_start[o];
        const int64_t sz = (int64_t)s_size[o];
        const int64_t first = m;
        basin_start[nb++] = first;
        for (int64_t i = 0; i < sz; ++i) {
            const int64_t f = s_vector[b + i];
            if ((int64_t)s_size[f] >= min_pixels) {
                pos[f] = m;
                flats[m++] = f;
            }
        }
        chi[first] = 0.0f;
        receivers[first] = -2;
        for (int64_t i = first + 1; i < m; ++i) {
            const int64_t f = flats[i];
            const int64_t r = receiver[f];
            const int64_t rl = r >= 0 ? pos[r] : -1;
            if (rl < first || rl >= i) {
                free(pos);
                return -ORC_ERR_ARG;
            }
            const double step = (dir[f] % 2u == 0u) ? dx : dx * k_sqrt2;
            const double au = (double)s_size[f] * pa;
            const double ar = (double)s_size[r] * pa;
            receivers[i] = r;
            chi[i] = (float)((double)chi[rl] + chi_inc(step, au, ar, A0, m_over_n));
        }
    }
    free(pos);
    *n_basins = nb;
    return m;
}
