/*
 * orc_dinf.c -- D-infinity flow graph restated (receivers and weights,
 * Kahn topological order, accumulation) and the drainage-area channel mask
 * the LEM hillslope step derives from it.  TEST INFRASTRUCTURE ONLY.
 *
 * Reference: src/cpu/lsdtt_flowgraph_cpu.cpp
 *   facet tables (Tarboton 1997, Table 1)      :62-67
 *   constructor: validity = mask && z != ndv && finite    :87-177
 *   parse_boundary_conditions ('b' forced, 'p' periodic)  :252-300
 *   is_forced_outlet / touches_interior_nodata             :310-352
 *   resolve_neighbor (periodic wrap, else no neighbour)    :364-396
 *   build_dinf                                             :423-663
 *   build_topological_order                                :874-985
 *   accumulate (dx^2 per valid cell, push in topo order)   :988-1033
 * src/cpu/lsdtt_lem_cpu.cpp
 *   resolve_hillslope_coupling_regime :64-77, build_channel_mask :86-134,
 *   build_hillslope_flowgraph (D-inf, interior nodata = outlet) :137-154
 * src/cpu/lsdtt_stream_power_cpu.cpp:613-619 channel_threshold_quantity
 *
 * Arithmetic is the reference's float arithmetic (no contraction).  The
 * one library call whose result is not fixed by IEEE rounding is
 * std::atan2(float, float): the C library decides its last bit (glibc
 * 2.35's atan2f differs from the correctly rounded value in about one
 * argument pair in eight; glibc 2.41 and later round correctly).  With
 * libm_atan2 = 0 the oracle uses the correctly rounded value,
 * (float)atan2((double)y, (double)x), which is what the device computes;
 * with libm_atan2 = 1 it calls atan2f as the reference does.
 *
 * Pins (tests/test_oracle_dinf.py): planes of every facet against the
 * gradient direction (receivers, weights, slope, bearing); weights sum to 1
 * and <= 2 receivers (tests/cpp/test_flow_properties.cpp P7); accumulation
 * against a dense linear solve (I - W^T) a = dx^2 on small grids; mass
 * conservation (sinks collect n_valid dx^2); the threshold quantity's
 * closed forms (tests/cpp/test_channel_threshold_units.cpp U1).
 */
#include "oracle.h"

#include <math.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif
#include <stdlib.h>
#include <string.h>

/* facet a: e1 cardinal at (r + E1R[a], c + E1C[a]), e2 diagonal at
 * (r + E2R[a], c + E2C[a]); direction = AF[a] * r + AC[a] * pi/2 */
static const int E1R[8] = {0, -1, -1, 0, 0, 1, 1, 0};
static const int E1C[8] = {1, 0, 0, -1, -1, 0, 0, 1};
static const int E2R[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
static const int E2C[8] = {1, 1, -1, -1, -1, -1, 1, 1};
static const int AC[8] = {0, 1, 1, 2, 2, 3, 3, 4};
static const int AF[8] = {1, -1, 1, -1, 1, -1, 1, -1};
/* raw 8-neighbourhood of touches_interior_nodata (order does not matter) */
static const int NBR[8] = {-1, 0, 0, 1, -1, -1, 1, 1};
static const int NBC[8] = {0, -1, 1, 0, -1, 1, -1, 1};

static int bc_ok(const char* bc)
{
    for (int i = 0; i < 4; ++i)
        if (bc[i] != 'o' && bc[i] != 'n' && bc[i] != 'p' && bc[i] != 'b') return 0;
    return (bc[0] == 'p') == (bc[2] == 'p') && (bc[1] == 'p') == (bc[3] == 'p');
}

/* resolve_neighbor: periodic edges wrap, any other edge has no neighbour */
static int64_t wrap_nb(int r, int c, int dr, int dc, int nr, int nc, const char* bc)
{
    int tr = r + dr, tc = c + dc;
    if (tr < 0) { if (bc[0] == 'p') tr = nr - 1; else return -1; }
    if (tr >= nr) { if (bc[2] == 'p') tr = 0; else return -1; }
    if (tc < 0) { if (bc[3] == 'p') tc = nc - 1; else return -1; }
    if (tc >= nc) { if (bc[1] == 'p') tc = 0; else return -1; }
    return (int64_t)tr * nc + tc;
}

int orc_dinf_graph(int nr, int nc, double dx, const float* z, const uint8_t* mask, float nodata,
                   const char* bc, int interior_nodata_is_outlet, int libm_atan2, int64_t* rec, float* w,
                   uint8_t* nrec, float* slope, float* bearing)
{
    if (!bc_ok(bc)) return ORC_ERR_BC;
    const int64_t n = (int64_t)nr * nc;
    const float cs = (float)dx;
    const float pi4 = (float)(M_PI / 4.0);
    const float sqrt2f = (float)1.4142135623730950488;
    const float rad2deg = (float)(180.0 / M_PI);
    uint8_t* valid = (uint8_t*)malloc((size_t)(n ? n : 1));
    for (int64_t f = 0; f < n; ++f) valid[f] = mask[f] && z[f] != nodata && isfinite(z[f]);
    for (int64_t f = 0; f < n; ++f) {
        rec[2 * f] = rec[2 * f + 1] = -1;
        w[2 * f] = w[2 * f + 1] = 0.0f;
        nrec[f] = 0;
        if (slope) slope[f] = nodata;
        if (bearing) bearing[f] = nodata;
    }
    for (int r = 0; r < nr; ++r)
        for (int c = 0; c < nc; ++c) {
            const int64_t f = (int64_t)r * nc + c;
            if (!valid[f]) continue;
            if ((bc[0] == 'b' && r == 0) || (bc[2] == 'b' && r == nr - 1) || (bc[1] == 'b' && c == nc - 1) ||
                (bc[3] == 'b' && c == 0))
                continue;
            if (interior_nodata_is_outlet) {
                int touch = 0;
                for (int k = 0; k < 8 && !touch; ++k) {
                    const int rr = r + NBR[k], cc = c + NBC[k];
                    if (rr < 0 || rr >= nr || cc < 0 || cc >= nc) continue;
                    if (!valid[(int64_t)rr * nc + cc]) touch = 1;
                }
                if (touch) continue;
            }
            const float e0 = z[f];
            float best_s = 0.0f, best_r = 0.0f;
            int best = -1;
            int64_t b1 = -1, b2 = -1;
            for (int a = 0; a < 8; ++a) {
                const int64_t g1 = wrap_nb(r, c, E1R[a], E1C[a], nr, nc, bc);
                const int64_t g2 = wrap_nb(r, c, E2R[a], E2C[a], nr, nc, bc);
                if (g1 < 0 || g2 < 0 || g1 == f || g2 == f) continue;
                if (!valid[g1] || !valid[g2]) continue;
                const float e1 = z[g1], e2 = z[g2];
                if (e0 <= e1 && e0 <= e2) continue;
                float rr, s;
                const float s1 = (e0 - e1) / cs;
                const float s2 = (e1 - e2) / cs;
                if (e0 > e1 && e0 > e2) {
                    rr = libm_atan2 ? atan2f(s2, s1) : (float)atan2((double)s2, (double)s1);
                    if (rr < 0.0f) {
                        rr = 0.0f;
                        s = s1;
                    } else if (rr > pi4) {
                        rr = pi4;
                        s = (e0 - e2) / (cs * sqrt2f);
                    } else {
                        s = sqrtf(s1 * s1 + s2 * s2);
                    }
                } else if (e0 > e1) {
                    rr = 0.0f;
                    s = s1;
                } else {
                    rr = pi4;
                    s = (e0 - e2) / (cs * sqrt2f);
                }
                if (s > best_s) {
                    best_s = s;
                    best_r = rr;
                    best = a;
                    b1 = g1;
                    b2 = g2;
                }
            }
            if (best < 0) continue;
            const float ang = (float)AF[best] * best_r + (float)AC[best] * ((float)M_PI * 0.5f);
            float br = 360.0f - (ang * rad2deg) + 90.0f;
            while (br >= 360.0f) br -= 360.0f;
            while (br < 0.0f) br += 360.0f;
            if (slope) slope[f] = best_s;
            if (bearing) bearing[f] = br;
            if (best_r == 0.0f) {
                nrec[f] = 1; rec[2 * f] = b1; w[2 * f] = 1.0f;
            } else if (best_r >= pi4) {
                nrec[f] = 1; rec[2 * f] = b2; w[2 * f] = 1.0f;
            } else {
                const float p2 = best_r / pi4;
                const float p1 = 1.0f - p2;
                if (b1 == b2) {
                    nrec[f] = 1; rec[2 * f] = b1; w[2 * f] = 1.0f;
                } else {
                    nrec[f] = 2;
                    rec[2 * f] = b1; w[2 * f] = p1;
                    rec[2 * f + 1] = b2; w[2 * f + 1] = p2;
                }
            }
        }
    free(valid);
    return ORC_OK;
}

int64_t orc_dinf_topo(int64_t n, const uint8_t* valid, const int64_t* rec, const float* w,
                      const uint8_t* nrec, int64_t* topo)
{
    uint8_t* rem = (uint8_t*)calloc((size_t)(n ? n : 1), 1);
    for (int64_t f = 0; f < n; ++f) {
        double ws = 0.0;
        for (int k = 0; k < nrec[f]; ++k) {
            const int64_t g = rec[2 * f + k];
            if (g < 0 || g >= n || g == f || !isfinite(w[2 * f + k]) || w[2 * f + k] < 0.0f) {
                free(rem);
                return -1;
            }
            ws += (double)w[2 * f + k];
            const int dup = k == 1 && rec[2 * f] == g;
            if (!dup && rem[g] < 255) ++rem[g];
        }
        if (nrec[f] && fabs(ws - 1.0) > 1e-5) {
            free(rem);
            return -1;
        }
    }
    int64_t head = 0, tail = 0;
    for (int64_t f = 0; f < n; ++f)
        if (valid[f] && rem[f] == 0) topo[tail++] = f;
    while (head < tail) {
        const int64_t f = topo[head++];
        for (int k = 0; k < nrec[f]; ++k) {
            const int64_t g = rec[2 * f + k];
            if (--rem[g] == 0) topo[tail++] = g;
        }
    }
    free(rem);
    return tail;
}

void orc_dinf_accumulate(int64_t n, const int64_t* rec, const float* w, const uint8_t* nrec, const int64_t* topo,
                         int64_t n_topo, double dx, double* accum)
{
    for (int64_t f = 0; f < n; ++f) accum[f] = 0.0;
    for (int64_t i = 0; i < n_topo; ++i) accum[topo[i]] = dx * dx;
    for (int64_t i = 0; i < n_topo; ++i) {
        const int64_t f = topo[i];
        const double a = accum[f];
        for (int k = 0; k < nrec[f]; ++k) accum[rec[2 * f + k]] += a * (double)w[2 * f + k];
    }
}

double orc_channel_threshold_quantity(double area, double dx, int divide_area_by_dx)
{
    if (!divide_area_by_dx) return area;
    return area / (dx > 1e-12 ? dx : 1e-12);
}

int orc_dinf_channel_mask(int nr, int nc, double dx, const float* z, const uint8_t* mask, const char* bc,
                          double threshold, int divide_area_by_dx, int libm_atan2, uint8_t* is_channel,
                          double* area_out)
{
    const int64_t n = (int64_t)nr * nc;
    int64_t* rec = (int64_t*)malloc(sizeof(int64_t) * (size_t)(2 * n + 1));
    float* w = (float*)malloc(sizeof(float) * (size_t)(2 * n + 1));
    uint8_t* nrec = (uint8_t*)malloc((size_t)(n + 1));
    uint8_t* valid = (uint8_t*)malloc((size_t)(n + 1));
    int64_t* topo = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n + 1));
    double* area = (double*)malloc(sizeof(double) * (size_t)(n + 1));
    /* the LEM's rasters carry no nodata value inside the mask: NaN never
     * compares equal */
    int st = orc_dinf_graph(nr, nc, dx, z, mask, NAN, bc, 1, libm_atan2, rec, w, nrec, NULL, NULL);
    int64_t nt = -1;
    if (st == ORC_OK) {
        for (int64_t f = 0; f < n; ++f) valid[f] = mask[f] && isfinite(z[f]);
        nt = orc_dinf_topo(n, valid, rec, w, nrec, topo);
        int64_t nv = 0;
        for (int64_t f = 0; f < n; ++f) nv += valid[f];
        if (nt != nv) st = ORC_ERR_CYCLIC;
    }
    if (st == ORC_OK) {
        orc_dinf_accumulate(n, rec, w, nrec, topo, nt, dx, area);
        for (int64_t f = 0; f < n; ++f)
            is_channel[f] = valid[f] && threshold > 0.0 &&
                            orc_channel_threshold_quantity(area[f], dx, divide_area_by_dx) >= threshold;
        if (area_out) memcpy(area_out, area, sizeof(double) * (size_t)n);
    }
    free(rec); free(w); free(nrec); free(valid); free(topo); free(area);
    return st;
}

/* compute_dinf_stream_power_incision (lsdtt_stream_power_cpu.cpp:329-602)
 * in FP64 state: receivers first (reverse Kahn order: every receiver
 * precedes its donors, as in the reference's reverse levels
 * (lsdtt_flowgraph_cpu.cpp:1176-1223), and each cell reads only its
 * receivers, so the result is the same).  Skips outlets (no receivers),
 * invalid cells and cells below the area threshold (on area/dx when
 * divide_area_by_dx).  Receiver classes (:435-441): one receiver, or w0 >=
 * 1 - 1e-6 (float): "cardinal" (rec[0], dx); w1 >= 1 - 1e-6: "diagonal"
 * (rec[1], dx sqrt2); else two receivers: Newton on
 * sbt = ((z_e1 - z)/dx)^2 + ((z_e2 - z_e1)/dx)^2 (:459-493), then the rise
 * guard and the clamp above the lower receiver.  pow(x, 2) is written x*x
 * (the same correctly rounded value).  Returns the cells whose solve ran
 * (two receivers, or one lower receiver). */
int64_t orc_stream_power_dinf(int64_t n, const uint8_t* valid, const int64_t* rec, const float* w,
                              const uint8_t* nrec, const int64_t* topo, int64_t n_topo, const double* area,
                              const double* K_field, const orc_sp_params* p, const double* z_old, double* z)
{
    const int n_is_one = fabs(p->n - 1.0) < 1e-6;
    const double dxc = p->dx, dxd = p->dx * 1.4142135623730950488;
    const double tol = p->slope_clamp * dxc;
    int64_t eroded = 0;
    for (int64_t ti = n_topo - 1; ti >= 0; --ti) {
        const int64_t f = topo[ti];
        const int k = nrec[f];
        if (!valid[f] || k == 0) continue;
        double a = area[f];
        if (p->divide_area_by_dx) a /= dxc;
        if (p->threshold > 0.0 && a < p->threshold) continue;
        const double zo = z_old[f];
        double zn = zo;
        const float w0 = w[2 * f], w1 = k > 1 ? w[2 * f + 1] : 0.0f;
        const int only_card = k == 1 || w0 >= 1.0f - 1e-6f;
        const int only_diag = k > 1 && w1 >= 1.0f - 1e-6f;
        const int two = k > 1 && !only_card && !only_diag;
        /* recv_elev (:443-448): a missing receiver reads the cell's own z_old */
        const int64_t g0 = rec[2 * f], g1 = k > 1 ? rec[2 * f + 1] : -1;
        const double ze1 = (g0 >= 0 && g0 < n) ? z[g0] : zo;
        const double ze2 = (g1 >= 0 && g1 < n) ? z[g1] : zo;
        const double Ki = K_field ? K_field[f] : p->K;
        const double kam = Ki * p->width_factor * pow(a, p->m);
        if (two) {
            double eps = tol + 1.0;
            int it = 0;
            do {
                const double t1 = (ze1 - zn) / dxc, t2 = (ze2 - ze1) / dxc;
                const double sbt = t1 * t1 + t2 * t2;
                if (sbt <= 0.0) {
                    eps = 0.0;
                } else {
                    const double fz = zn - zo + kam * pow(sbt, 0.5 * p->n) * p->dt;
                    const double fd = 1.0 - kam * p->n * (ze1 - zn) / (dxc * dxc) * pow(sbt, 0.5 * p->n - 1.0) * p->dt;
                    eps = fz / fd;
                }
                zn -= eps;
                ++it;
                if (it > 100) eps = tol / 2.0;
            } while (fabs(eps) > tol);
            if ((zn - zo) > tol || isnan(zn)) zn = zo;
            const double zmin = ze1 < ze2 ? ze1 : ze2;
            if (zn < zmin) zn = zmin + p->slope_clamp * dxc;
            ++eroded;
        } else {
            const double zr = only_diag ? ze2 : ze1;
            const double dxe = only_diag ? dxd : dxc;
            if (zr >= zo) {
                zn = zo;
            } else if (n_is_one) {
                const double spf = kam * (p->dt / dxe);
                zn = (zo + zr * spf) / (1.0 + spf);
                if ((zn - zo) > tol || isnan(zn)) zn = zo;
                ++eroded;
            } else {
                double eps = tol + 1.0;
                int it = 0;
                do {
                    const double slope = (zn - zr) / dxe;
                    if (slope <= 0.0) {
                        eps = 0.0;
                    } else {
                        const double fz = zn - zo + kam * p->dt * pow(slope, p->n);
                        const double fd = 1.0 + kam * p->dt * (p->n / dxe) * pow(slope, p->n - 1.0);
                        eps = fz / fd;
                    }
                    zn -= eps;
                    ++it;
                    if (it > 100) eps = tol / 2.0;
                } while (fabs(eps) > tol);
                if ((zn - zo) > tol || isnan(zn)) zn = zo;
                if (zn < zr) zn = zr + p->slope_clamp * dxe;
                ++eroded;
            }
        }
        z[f] = zn;
    }
    return eroded;
}
