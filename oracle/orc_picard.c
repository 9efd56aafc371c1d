/*
 * orc_picard.c -- backward-Euler Roering hillslope transport by damped
 * Picard iteration, the reference's fallback when Newton fails.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates src/cpu/lsdtt_hillslope_implicit_cpu.cpp:108-1010 (backward-Euler
 * path, time_integrator 0):
 *   classification (Dirichlet = channel or 'o' edge)       :186-198
 *   nbr_z reflecting neighbour values, get_face no-flux faces :214-244
 *   cell gradients, face D_eff = D / (1 - min(r^2, r_max^2)) :384-435
 *   max face ratio^2 (uncapped)                            :442-468
 *   system [I - dt L(D_eff)] y = z_old + dt U (+ Dirichlet) :475-500, 586-616
 *   damped Picard z += omega (y - z), stop on max|dz| < tol :512-583
 *   linear: fixed sub-steps, one solve each                :857-879
 *   nonlinear controller (semi-implicit fast path, accept on
 *   max ratio < r_max, grow / shrink / damping, floor projection) :880-993
 *   projection (Gauss-Seidel lowering to z_nb + r_max Sc h) :733-758
 * The reference solves each system with Eigen's SimplicialLDLT (exact up to
 * rounding); this restatement uses Jacobi-preconditioned conjugate
 * gradients to a relative residual of 1e-14.  Eigen is absent here, so the
 * reference's iterates are parity unpinned; the checks are closed forms
 * (tests/test_oracle_picard.py).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

typedef struct {
    int nr, nc;
    char bn, be, bs, bw;
    const uint8_t* mask;
    int64_t* dof_of;       /* cell -> dof or -1 */
    int64_t* dof_cell;     /* dof -> cell */
    int64_t ndof;
    double h;
} pg_t;

static double nbr_z(const pg_t* g, const double* z, int r, int c, int dr, int dc)
{
    int rr = r + dr, cc = c + dc;
    if (rr < 0) rr = (g->bn == 'p') ? g->nr - 1 : r;
    if (rr >= g->nr) rr = (g->bs == 'p') ? 0 : r;
    if (cc < 0) cc = (g->bw == 'p') ? g->nc - 1 : c;
    if (cc >= g->nc) cc = (g->be == 'p') ? 0 : c;
    if (!g->mask[(int64_t)rr * g->nc + cc]) return z[(int64_t)r * g->nc + c];
    return z[(int64_t)rr * g->nc + cc];
}

/* neighbour across face dir (0 N, 1 E, 2 S, 3 W) or -1 for a no-flux face */
static int64_t get_face(const pg_t* g, int r, int c, int dir)
{
    int dr = 0, dc = 0;
    char code;
    if (dir == 0) { dr = -1; code = g->bn; }
    else if (dir == 1) { dc = 1; code = g->be; }
    else if (dir == 2) { dr = 1; code = g->bs; }
    else { dc = -1; code = g->bw; }
    int rr = r + dr, cc = c + dc;
    if (rr < 0 || rr >= g->nr || cc < 0 || cc >= g->nc) {
        if (code != 'p') return -1;
        rr = (rr + g->nr) % g->nr;
        cc = (cc + g->nc) % g->nc;
    }
    const int64_t f = (int64_t)rr * g->nc + cc;
    return g->mask[f] ? f : -1;
}

/* gx, gy of every valid cell (lsdtt_hillslope_implicit_cpu.cpp:384-398) */
static void gradients(const pg_t* g, const double* z, double* gx, double* gy)
{
    for (int r = 0; r < g->nr; ++r)
        for (int c = 0; c < g->nc; ++c) {
            const int64_t f = (int64_t)r * g->nc + c;
            if (!g->mask[f]) continue;
            gx[f] = (nbr_z(g, z, r, c, 0, 1) - nbr_z(g, z, r, c, 0, -1)) / (2.0 * g->h);
            gy[f] = (nbr_z(g, z, r, c, -1, 0) - nbr_z(g, z, r, c, 1, 0)) / (2.0 * g->h);
        }
}

/* D_eff of the face between s and its neighbour sn across dir */
static double face_deff(const pg_t* g, const orc_picard_params* p, const double* z, const double* gx,
                        const double* gy, int64_t s, int64_t sn, int dir)
{
    if (!p->nonlinear) return p->D;
    const double Sc = p->Sc > 0.0 ? p->Sc : 1.0;
    const double Sn = (z[sn] - z[s]) / g->h;
    const double Sx = (dir == 1 || dir == 3) ? 0.5 * (gy[s] + gy[sn]) : 0.5 * (gx[s] + gx[sn]);
    double r2 = (Sn * Sn + Sx * Sx) / (Sc * Sc);
    const double cap = p->slope_ratio_cap * p->slope_ratio_cap;
    if (r2 > cap) r2 = cap;
    return p->D / (1.0 - r2);
}

static double max_face_ratio2(const pg_t* g, const orc_picard_params* p, const double* z, double* gx, double* gy)
{
    gradients(g, z, gx, gy);
    const double Sc = p->Sc > 0.0 ? p->Sc : 1.0;
    double mx = 0.0;
    for (int r = 0; r < g->nr; ++r)
        for (int c = 0; c < g->nc; ++c) {
            const int64_t s = (int64_t)r * g->nc + c;
            if (!g->mask[s]) continue;
            for (int dir = 1; dir <= 2; ++dir) {
                const int64_t sn = get_face(g, r, c, dir);
                if (sn < 0) continue;
                const double Sn = (z[sn] - z[s]) / g->h;
                const double Sx = dir == 1 ? 0.5 * (gy[s] + gy[sn]) : 0.5 * (gx[s] + gx[sn]);
                const double r2 = (Sn * Sn + Sx * Sx) / (Sc * Sc);
                if (r2 > mx) mx = r2;
            }
        }
    return mx;
}

/* w[4 d + dir] = beta D_eff / h^2 of DOF d's faces (0 = no face) */
static void face_weights(const pg_t* g, const orc_picard_params* p, const double* z, double* gx, double* gy,
                         double beta, double* w)
{
    if (p->nonlinear) gradients(g, z, gx, gy);
    for (int64_t d = 0; d < g->ndof; ++d) {
        const int64_t s = g->dof_cell[d];
        const int r = (int)(s / g->nc), c = (int)(s % g->nc);
        for (int dir = 0; dir < 4; ++dir) {
            const int64_t sn = get_face(g, r, c, dir);
            w[4 * d + dir] = sn < 0 ? 0.0 : beta * face_deff(g, p, z, gx, gy, s, sn, dir) / (g->h * g->h);
        }
    }
}

static void apply_A(const pg_t* g, const double* w, const int64_t* fn, const double* x, double* y)
{
    for (int64_t d = 0; d < g->ndof; ++d) {
        double acc = x[d];
        for (int k = 0; k < 4; ++k) {
            const double wk = w[4 * d + k];
            if (wk == 0.0) continue;
            acc += wk * x[d];
            const int64_t nd = fn[4 * d + k];
            if (nd >= 0) acc -= wk * x[nd];
        }
        y[d] = acc;
    }
}

/* Jacobi-preconditioned CG on the SPD system, warm-started from x */
static void solve_spd(const pg_t* g, const double* w, const int64_t* fn, const double* b, double* x)
{
    const int64_t nd = g->ndof;
    double* r = (double*)malloc(sizeof(double) * (size_t)nd);
    double* q = (double*)malloc(sizeof(double) * (size_t)nd);
    double* pp = (double*)malloc(sizeof(double) * (size_t)nd);
    double* Ap = (double*)malloc(sizeof(double) * (size_t)nd);
    double* mi = (double*)malloc(sizeof(double) * (size_t)nd);
    for (int64_t d = 0; d < nd; ++d) {
        double diag = 1.0;
        for (int k = 0; k < 4; ++k) diag += w[4 * d + k];
        mi[d] = 1.0 / diag;
    }
    apply_A(g, w, fn, x, Ap);
    double bn = 0.0, rz = 0.0;
    for (int64_t d = 0; d < nd; ++d) {
        r[d] = b[d] - Ap[d];
        q[d] = mi[d] * r[d];
        pp[d] = q[d];
        rz += r[d] * q[d];
        bn += b[d] * b[d];
    }
    const double tol2 = 1e-28 * (bn > 0.0 ? bn : 1.0);
    for (int it = 0; it < 10000; ++it) {
        double rr = 0.0;
        for (int64_t d = 0; d < nd; ++d) rr += r[d] * r[d];
        if (rr <= tol2) break;
        apply_A(g, w, fn, pp, Ap);
        double pAp = 0.0;
        for (int64_t d = 0; d < nd; ++d) pAp += pp[d] * Ap[d];
        if (pAp <= 0.0) break;
        const double alpha = rz / pAp;
        double rz_new = 0.0;
        for (int64_t d = 0; d < nd; ++d) {
            x[d] += alpha * pp[d];
            r[d] -= alpha * Ap[d];
            q[d] = mi[d] * r[d];
            rz_new += r[d] * q[d];
        }
        const double beta = rz_new / rz;
        rz = rz_new;
        for (int64_t d = 0; d < nd; ++d) pp[d] = q[d] + beta * pp[d];
    }
    free(r); free(q); free(pp); free(Ap); free(mi);
}

/* one backward-Euler stage: damped Picard on [I - beta L] y = g */
static int solve_stage(const pg_t* g, const orc_picard_params* p, double beta, const double* gfield, double omega,
                       int iter_cap, double* z, double* gx, double* gy, double* w, const int64_t* fn, double* b,
                       double* y, int* iters_out)
{
    *iters_out = 0;
    int converged = 0;
    for (int k = 0; k < iter_cap; ++k) {
        face_weights(g, p, z, gx, gy, beta, w);
        for (int64_t d = 0; d < g->ndof; ++d) {
            const int64_t s = g->dof_cell[d];
            const int r = (int)(s / g->nc), c = (int)(s % g->nc);
            double bd = 0.0;
            for (int dir = 0; dir < 4; ++dir) {
                const int64_t sn = get_face(g, r, c, dir);
                if (sn < 0 || fn[4 * d + dir] >= 0) continue;
                bd += w[4 * d + dir] * z[sn];  /* Dirichlet neighbour */
            }
            b[d] = gfield[s] + bd;
            y[d] = z[s];
        }
        solve_spd(g, w, fn, b, y);
        ++*iters_out;
        double max_dz = 0.0;
        for (int64_t d = 0; d < g->ndof; ++d) {
            const int64_t s = g->dof_cell[d];
            const double zn = z[s] + omega * (y[d] - z[s]);
            const double dz = fabs(zn - z[s]);
            if (dz > max_dz) max_dz = dz;
            z[s] = zn;
        }
        if (!p->nonlinear) { converged = 1; break; }
        if (max_dz < p->picard_tol) { converged = 1; break; }
    }
    return converged;
}

static void projection(const pg_t* g, const orc_picard_params* p, double* z)
{
    const double cap_slope = p->slope_ratio_cap * (p->Sc > 0.0 ? p->Sc : 1.0);
    for (int sweep = 0; sweep < 25; ++sweep) {
        double max_change = 0.0;
        for (int r = 0; r < g->nr; ++r)
            for (int c = 0; c < g->nc; ++c) {
                const int64_t f = (int64_t)r * g->nc + c;
                if (g->dof_of[f] < 0) continue;
                double za = INFINITY;
                for (int dir = 0; dir < 4; ++dir) {
                    const int64_t sn = get_face(g, r, c, dir);
                    if (sn < 0) continue;
                    const double v = z[sn] + cap_slope * g->h;
                    if (v < za) za = v;
                }
                if (z[f] > za) {
                    if (z[f] - za > max_change) max_change = z[f] - za;
                    z[f] = za;
                }
            }
        if (max_change < 1e-12) break;
    }
}

int orc_picard_transient(int nr, int nc, double h, const uint8_t* mask, const char* bc,
                         const uint8_t* is_channel, const orc_picard_params* p, double dt_total, double* z,
                         orc_picard_result* res)
{
    memset(res, 0, sizeof(*res));
    const int64_t n = (int64_t)nr * nc;
    pg_t g;
    g.nr = nr; g.nc = nc; g.h = h; g.mask = mask;
    g.bn = bc[0]; g.be = bc[1]; g.bs = bc[2]; g.bw = bc[3];
    g.dof_of = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
    g.dof_cell = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
    g.ndof = 0;
    for (int r = 0; r < nr; ++r)
        for (int c = 0; c < nc; ++c) {
            const int64_t f = (int64_t)r * nc + c;
            g.dof_of[f] = -1;
            if (!mask[f]) continue;
            const int edge = (r == 0 && g.bn == 'o') || (r == nr - 1 && g.bs == 'o') ||
                             (c == 0 && g.bw == 'o') || (c == nc - 1 && g.be == 'o');
            if ((is_channel && is_channel[f]) || edge) continue;
            g.dof_of[f] = g.ndof;
            g.dof_cell[g.ndof++] = f;
        }
    double* gx = (double*)calloc((size_t)n, sizeof(double));
    double* gy = (double*)calloc((size_t)n, sizeof(double));
    if (g.ndof == 0 || dt_total <= 0.0) {
        res->max_slope_ratio = sqrt(max_face_ratio2(&g, p, z, gx, gy));
        res->final_dt = dt_total;
        free(g.dof_of); free(g.dof_cell); free(gx); free(gy);
        return ORC_OK;
    }
    const int64_t nd = g.ndof;
    int64_t* fn = (int64_t*)malloc(sizeof(int64_t) * (size_t)(4 * nd));
    for (int64_t d = 0; d < nd; ++d) {
        const int64_t s = g.dof_cell[d];
        const int r = (int)(s / nc), c = (int)(s % nc);
        for (int dir = 0; dir < 4; ++dir) {
            const int64_t sn = get_face(&g, r, c, dir);
            fn[4 * d + dir] = sn < 0 ? -1 : g.dof_of[sn];
        }
    }
    double* w = (double*)malloc(sizeof(double) * (size_t)(4 * nd));
    double* b = (double*)malloc(sizeof(double) * (size_t)nd);
    double* y = (double*)malloc(sizeof(double) * (size_t)nd);
    double* gf = (double*)calloc((size_t)n, sizeof(double));
    double* z_old = (double*)malloc(sizeof(double) * (size_t)n);
    const double omega = p->relaxation_omega > 0.0 ? p->relaxation_omega : 1.0;
    const int pmax = p->picard_max_iter > 0 ? p->picard_max_iter : 50;
    double final_dt = 0.0;

    if (!p->nonlinear) {
        double dt_lin = p->dt_init > 0.0 ? fmin(p->dt_init, dt_total) : dt_total;
        int n_lin = (int)ceil(dt_total / dt_lin);
        if (n_lin < 1) n_lin = 1;
        dt_lin = dt_total / n_lin;
        for (int s = 0; s < n_lin; ++s) {
            memcpy(z_old, z, sizeof(double) * (size_t)n);
            for (int64_t d = 0; d < nd; ++d) gf[g.dof_cell[d]] = z_old[g.dof_cell[d]] + dt_lin * p->U;
            int it = 0;
            solve_stage(&g, p, dt_lin, gf, 1.0, pmax, z, gx, gy, w, fn, b, y, &it);
            res->picard_iters += it;
            ++res->substeps;
        }
        final_dt = dt_lin;
    } else {
        const double dt_max = p->dt_max > 0.0 ? p->dt_max : dt_total;
        double dt_min = p->dt_min > 0.0 ? p->dt_min : dt_total * 1e-4;
        if (dt_min > dt_max) dt_min = dt_max;
        const double grow = p->grow_factor > 1.0 ? p->grow_factor : 1.5;
        const double shrink = (p->shrink_factor > 0.0 && p->shrink_factor < 1.0) ? p->shrink_factor : 0.5;
        const double rmax = p->slope_ratio_cap;
        double cur_omega = omega;
        double dt = p->dt_init > 0.0 ? p->dt_init : dt_total;
        dt = fmin(dt, fmin(dt_max, dt_total));
        double t_rem = dt_total;
        const double t_eps = dt_total * 1e-12;
        double mr2 = max_face_ratio2(&g, p, z, gx, gy);
        int attempts = 0;
        while (t_rem > t_eps && attempts < 2000000) {
            ++attempts;
            dt = fmin(dt, t_rem);
            dt = fmin(dt, dt_max);
            if (dt < dt_min) dt = dt_min;
            memcpy(z_old, z, sizeof(double) * (size_t)n);
            const int semi = (mr2 < 0.25) && (cur_omega >= omega - 1e-12);
            const int cap = semi ? 1 : pmax;
            const double step_omega = semi ? 1.0 : cur_omega;
            for (int64_t d = 0; d < nd; ++d) gf[g.dof_cell[d]] = z_old[g.dof_cell[d]] + dt * p->U;
            int it = 0;
            int conv = solve_stage(&g, p, dt, gf, step_omega, cap, z, gx, gy, w, fn, b, y, &it);
            if (semi) conv = 1;
            double step_dz = 0.0;
            for (int64_t d = 0; d < nd; ++d) {
                const double v = fabs(z[g.dof_cell[d]] - z_old[g.dof_cell[d]]);
                if (v > step_dz) step_dz = v;
            }
            const double m2 = max_face_ratio2(&g, p, z, gx, gy);
            const int slope_ok = m2 < rmax * rmax;
            const int dz_ok = p->dz_target <= 0.0 || step_dz <= p->dz_target;
            if (conv && slope_ok && dz_ok) {
                t_rem -= dt; ++res->substeps; res->picard_iters += it; final_dt = dt;
                mr2 = m2;
                const int easy = it <= 5 && (p->dz_target <= 0.0 || step_dz <= 0.5 * p->dz_target);
                if (easy) dt = fmin(dt * grow, dt_max);
                cur_omega = fmin(omega, cur_omega * 1.1);
            } else if (dt <= dt_min * (1.0 + 1e-9)) {
                projection(&g, p, z);
                res->limiter_fired = 1;
                mr2 = max_face_ratio2(&g, p, z, gx, gy);
                t_rem -= dt; ++res->substeps; res->picard_iters += it; final_dt = dt;
            } else {
                memcpy(z, z_old, sizeof(double) * (size_t)n);
                ++res->rejected;
                dt = fmax(dt * shrink, dt_min);
                if (!conv) cur_omega = fmax(0.1, cur_omega * 0.7);
            }
        }
        if (final_dt <= 0.0) final_dt = dt;
    }
    res->final_dt = final_dt;
    res->max_slope_ratio = sqrt(max_face_ratio2(&g, p, z, gx, gy));
    free(fn); free(w); free(b); free(y); free(gf); free(z_old);
    free(g.dof_of); free(g.dof_cell); free(gx); free(gy);
    return ORC_OK;
}
