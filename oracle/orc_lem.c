/*
 * orc_lem.c -- the LEM timestep restated in FP64 state.  TEST
 * INFRASTRUCTURE ONLY.
 *
 * Reference (all paths under src/cpu/):
 *   lsdtt_stream_power_cpu.cpp:159-326   D8 implicit stream power
 *   lsdtt_hillslope_operator_cpu.hpp:83-321  domain, BC and face alphas
 *   lsdtt_hillslope_newton_cpu.cpp:137-186   dalpha_dz
 *   lsdtt_hillslope_newton_cpu.cpp:196-675   Newton solve + line search
 *   lsdtt_hillslope_newton_cpu.cpp:690-788   adaptive sub-steps
 *   lsdtt_lem_cpu.cpp:285-732                run_lem_step
 *   lsdtt_lem_cpu.cpp:735-808                build_boundary_flats
 *   src/utils/lsdtt_statistics.cpp:420-429   quantile (nth_element, floor)
 *
 * Departures, each deliberate:
 *   - state is double throughout (the reference stores float and widens);
 *   - drainage area is n_contributing * cell_size^2 (equal to the
 *     reference's double accumulation whenever cell_size^2 is an integer);
 *   - the Newton linear solve is right-Jacobi-preconditioned restarted
 *     GMRES in double to a 1e-13 relative residual (the reference uses
 *     Eigen SparseLU / BiCGSTAB+ILUT, absent here); the Newton iteration,
 *     line search, restoration mode and sub-stepping are the reference's;
 *   - convergence: ||F||inf <= max(abs_tol, rel_tol * max(||F(z^n)||inf, U))
 *     (abs_tol 1e-12 m/yr is the reference's default criterion);
 *   - the hillslope solve's per-cell loops run under OpenMP (the reference's
 *     Newton is OpenMP-parallel as well, lsdtt_hillslope_newton_cpu.cpp:216);
 *     every sum is taken over fixed 4096-element blocks added in block
 *     order, so results do not depend on the thread count.
 *
 * Parity status: stream power pinned by the n = 1 closed form and the n != 1
 * root residual (tests/test_oracle_lem.py); the hillslope solver by the
 * discrete backward-Euler decay of a Fourier mode (linear), the Roering 1-D
 * steady profile (tests/cpp/test_hillslope_implicit.cpp:38-39), divide
 * curvature -U/D (tests/cpp/test_hillslope_newton.cpp:278-311), mass
 * conservation on a closed domain, and a dense direct solve of the same
 * Newton system on small grids.  The reference Newton itself needs Eigen and
 * cannot run here, so its iterates are not pinned -- only the converged
 * solution of the same discrete equations.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int orc_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* sum of a[k]*b[k] over fixed blocks added in block order (thread-count
 * independent); part holds (n + ORC_BLK - 1) / ORC_BLK doubles */
#define ORC_BLK 4096
static double dot_blocks(const double* a, const double* b, int64_t n, double* part)
{
    const int64_t nb = (n + ORC_BLK - 1) / ORC_BLK;
    #pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < nb; ++q) {
        const int64_t k1 = (q + 1) * ORC_BLK < n ? (q + 1) * ORC_BLK : n;
        double acc = 0.0;
        for (int64_t k = q * ORC_BLK; k < k1; ++k) acc += a[k] * b[k];
        part[q] = acc;
    }
    double acc = 0.0;
    for (int64_t q = 0; q < nb; ++q) acc += part[q];
    return acc;
}

/* ----------------------------------------------------------------------- */
/* stream power                                                            */
/* ----------------------------------------------------------------------- */

static const double SQRT2 = 1.4142135623730950488;

int64_t orc_stream_power_d8(int64_t n, const uint8_t* mask,
                            const int64_t* receiver, const uint8_t* dir,
                            const int64_t* topo, int64_t n_topo,
                            const double* area, const double* K_field,
                            const orc_sp_params* p, const double* z_old,
                            double* z)
{
    const int n_is_one = fabs(p->n - 1.0) < 1e-6;
    int64_t eroded = 0;
    for (int64_t ti = n_topo - 1; ti >= 0; --ti) {
        const int64_t f = topo[ti];
        const int64_t rc = receiver[f];
        if (rc < 0 || rc >= n || !mask[f]) continue;
        const double dx_eff = (dir[f] & 1u) ? p->dx * SQRT2 : p->dx;
        double a = area[f];
        if (p->divide_area_by_dx) a /= p->dx;
        if (p->threshold > 0.0 && a < p->threshold) continue;
        const double zo = z_old[f];
        const double zr = z[rc];
        if (zr >= zo) continue;
        const double tol = p->slope_clamp * dx_eff;
        const double Ki = K_field ? K_field[f] : p->K;
        const double kam = Ki * p->width_factor * pow(a, p->m);
        double zn = zo;
        if (n_is_one) {
            const double spf = kam * (p->dt / dx_eff);
            zn = (zo + zr * spf) / (1.0 + spf);
            if ((zn - zo) > tol || isnan(zn)) zn = zo;
        } else {
            double eps = tol + 1.0;
            int it = 0;
            do {
                const double slope = (zn - zr) / dx_eff;
                if (slope <= 0.0) {
                    eps = 0.0;
                } else {
                    const double fz = zn - zo + kam * p->dt * pow(slope, p->n);
                    const double fd = 1.0 + kam * p->dt * (p->n / dx_eff) * pow(slope, p->n - 1.0);
                    eps = fz / fd;
                }
                zn -= eps;
                ++it;
                if (it > 100) eps = tol / 2.0;
            } while (fabs(eps) > tol);
            if ((zn - zo) > tol || isnan(zn)) zn = zo;
            if (zn < zr) zn = zr + p->slope_clamp * dx_eff;
        }
        z[f] = zn;
        ++eroded;
    }
    return eroded;
}

/* ----------------------------------------------------------------------- */
/* hillslope operator                                                      */
/* ----------------------------------------------------------------------- */

typedef struct {
    int nr, nc;
    char bn, be, bs, bw;
    const uint8_t* mask;
    double h;
} dom_t;

static void nbr(const dom_t* d, int r, int c, int dr, int dc, int* orr, int* occ)
{
    int rr = r + dr, cc = c + dc;
    if (rr < 0) rr = (d->bn == 'p') ? d->nr - 1 : r;
    else if (rr >= d->nr) rr = (d->bs == 'p') ? 0 : r;
    if (cc < 0) cc = (d->bw == 'p') ? d->nc - 1 : c;
    else if (cc >= d->nc) cc = (d->be == 'p') ? 0 : c;
    if (!d->mask[(int64_t)rr * d->nc + cc]) { rr = r; cc = c; }
    *orr = rr;
    *occ = cc;
}

static double at(const dom_t* d, const double* z, int r, int c, int dr, int dc)
{
    int rr, cc;
    nbr(d, r, c, dr, dc, &rr, &cc);
    return z[(int64_t)rr * d->nc + cc];
}

static double slope_x(const dom_t* d, const double* z, int r, int c)
{
    return (at(d, z, r, c, 0, 1) - at(d, z, r, c, 0, -1)) / (2.0 * d->h);
}
static double slope_y(const dom_t* d, const double* z, int r, int c)
{
    return (at(d, z, r, c, 1, 0) - at(d, z, r, c, -1, 0)) / (2.0 * d->h);
}

static const int NESW_R[4] = {-1, 0, 1, 0};
static const int NESW_C[4] = {0, 1, 0, -1};

/* face_alphas, lsdtt_hillslope_operator_cpu.hpp:290-321 */
static void face_alphas(const dom_t* d, const double* z, int r, int c, double Sc,
                        int nonlinear, double cap, double a[4], double sn[4])
{
    const double zc = z[(int64_t)r * d->nc + c];
    const double sxc = nonlinear ? slope_x(d, z, r, c) : 0.0;
    const double syc = nonlinear ? slope_y(d, z, r, c) : 0.0;
    for (int k = 0; k < 4; ++k) {
        int rr, cc;
        nbr(d, r, c, NESW_R[k], NESW_C[k], &rr, &cc);
        sn[k] = (z[(int64_t)rr * d->nc + cc] - zc) / d->h;
        if (!nonlinear) { a[k] = 1.0; continue; }
        const double st = (k == 0 || k == 2) ? 0.5 * (sxc + slope_x(d, z, rr, cc))
                                             : 0.5 * (syc + slope_y(d, z, rr, cc));
        double s2 = (sn[k] * sn[k] + st * st) / (Sc * Sc);
        if (s2 > cap) s2 = cap;
        a[k] = 1.0 / (1.0 - s2);
    }
}

typedef struct {
    dom_t d;
    int64_t ndof;
    int64_t* dof_of;   /* cell -> dof or -1 */
    int* rc;           /* dof -> (r, c) */
    double D, U, Sc, inv_dt, cap;
    int nonlinear;
    const double* zn;
} hs_t;

static double residual(const hs_t* s, const double* z, double* F)
{
    const double DA = s->D / (s->d.h * s->d.h);
    double norm = 0.0;
    #pragma omp parallel for schedule(static) reduction(max:norm)
    for (int64_t k = 0; k < s->ndof; ++k) {
        const int r = s->rc[2 * k], c = s->rc[2 * k + 1];
        double a[4], sn[4];
        face_alphas(&s->d, z, r, c, s->Sc, s->nonlinear, s->cap, a, sn);
        double flux = 0.0;
        for (int q = 0; q < 4; ++q) flux += a[q] * sn[q] * s->d.h;
        double f = -s->U - DA * flux;
        if (s->inv_dt > 0.0) {
            const int64_t i = (int64_t)r * s->d.nc + c;
            f += s->inv_dt * (z[i] - s->zn[i]);
        }
        F[k] = f;
        if (fabs(f) > norm) norm = fabs(f);
    }
    return norm;
}

static double max_ratio(const hs_t* s, const double* z)
{
    double m = 0.0;
    #pragma omp parallel for schedule(static) reduction(max:m)
    for (int64_t k = 0; k < s->ndof; ++k) {
        const int r = s->rc[2 * k], c = s->rc[2 * k + 1];
        const double sx = slope_x(&s->d, z, r, c), sy = slope_y(&s->d, z, r, c);
        const double v = sqrt(sx * sx + sy * sy) / s->Sc;
        if (v > m) m = v;
    }
    return m;
}

/* sparse rows: up to MAXE merged (col, val) per row */
#define MAXE 24
typedef struct { int64_t col[MAXE]; double val[MAXE]; int n; } row_t;

static void row_add(const hs_t* s, row_t* row, int r, int c, double w)
{
    if (w == 0.0) return;
    const int64_t dd = s->dof_of[(int64_t)r * s->d.nc + c];
    if (dd < 0) return;
    for (int i = 0; i < row->n; ++i)
        if (row->col[i] == dd) { row->val[i] += w; return; }
    row->col[row->n] = dd;
    row->val[row->n] = w;
    ++row->n;
}

/* Jacobian assembly, lsdtt_hillslope_newton_cpu.cpp:415-471 with dalpha_dz
 * :137-186 */
static void assemble(const hs_t* s, const double* z, row_t* J)
{
    const double DA = s->D / (s->d.h * s->d.h);
    const double h = s->d.h, Sc2 = s->Sc * s->Sc;
    #pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < s->ndof; ++k) {
        const int r = s->rc[2 * k], c = s->rc[2 * k + 1];
        row_t* row = &J[k];
        row->n = 0;
        double a[4], sn[4];
        face_alphas(&s->d, z, r, c, s->Sc, s->nonlinear, s->cap, a, sn);
        if (s->inv_dt > 0.0) row_add(s, row, r, c, s->inv_dt);
        for (int q = 0; q < 4; ++q) {
            int rr, cc;
            nbr(&s->d, r, c, NESW_R[q], NESW_C[q], &rr, &cc);
            row_add(s, row, rr, cc, -DA * a[q]);
            row_add(s, row, r, c, +DA * a[q]);
            if (!s->nonlinear) continue;
            const int tx = (q == 0 || q == 2);
            const double st = tx ? 0.5 * (slope_x(&s->d, z, r, c) + slope_x(&s->d, z, rr, cc))
                                 : 0.5 * (slope_y(&s->d, z, r, c) + slope_y(&s->d, z, rr, cc));
            const double rraw = (sn[q] * sn[q] + st * st) / Sc2;
            if (rraw >= s->cap) continue;
            const double pref = a[q] * a[q] * 2.0 / Sc2;
            const double dzf = sn[q] * h;
            int er[6], ec[6];
            double ew[6];
            er[0] = r; ec[0] = c; ew[0] = pref * sn[q] * (-1.0 / h);
            er[1] = rr; ec[1] = cc; ew[1] = pref * sn[q] * (+1.0 / h);
            const double wt = pref * st * (1.0 / (4.0 * h));
            if (tx) {
                nbr(&s->d, r, c, 0, 1, &er[2], &ec[2]);
                nbr(&s->d, r, c, 0, -1, &er[3], &ec[3]);
                nbr(&s->d, rr, cc, 0, 1, &er[4], &ec[4]);
                nbr(&s->d, rr, cc, 0, -1, &er[5], &ec[5]);
            } else {
                nbr(&s->d, r, c, 1, 0, &er[2], &ec[2]);
                nbr(&s->d, r, c, -1, 0, &er[3], &ec[3]);
                nbr(&s->d, rr, cc, 1, 0, &er[4], &ec[4]);
                nbr(&s->d, rr, cc, -1, 0, &er[5], &ec[5]);
            }
            ew[2] = +wt; ew[3] = -wt; ew[4] = +wt; ew[5] = -wt;
            for (int e = 0; e < 6; ++e) row_add(s, row, er[e], ec[e], -DA * dzf * ew[e]);
        }
    }
}

static void matvec(const row_t* J, int64_t n, const double* x, double* y)
{
    #pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < n; ++k) {
        double a = 0.0;
        for (int i = 0; i < J[k].n; ++i) a += J[k].val[i] * x[J[k].col[i]];
        y[k] = a;
    }
}

/* right-Jacobi-preconditioned restarted GMRES(m); returns 0 on success */
static int gmres(const row_t* J, int64_t n, const double* b, double* x, double rtol, int maxit)
{
    const int m = 60;
    double* dinv = (double*)malloc(sizeof(double) * (size_t)n);
    double* part = (double*)malloc(sizeof(double) * (size_t)((n + ORC_BLK - 1) / ORC_BLK));
    #pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < n; ++k) {
        double dg = 0.0;
        for (int i = 0; i < J[k].n; ++i) if (J[k].col[i] == k) dg = J[k].val[i];
        dinv[k] = dg != 0.0 ? 1.0 / dg : 1.0;
    }
    double* V = (double*)malloc(sizeof(double) * (size_t)n * (m + 1));
    double* Hm = (double*)calloc((size_t)(m + 1) * m, sizeof(double));
    double* cs = (double*)malloc(sizeof(double) * m);
    double* sn = (double*)malloc(sizeof(double) * m);
    double* g = (double*)malloc(sizeof(double) * (m + 1));
    double* w = (double*)malloc(sizeof(double) * (size_t)n);
    double* t = (double*)malloc(sizeof(double) * (size_t)n);
    double bn = sqrt(dot_blocks(b, b, n, part));
    memset(x, 0, sizeof(double) * (size_t)n);
    int ok = (bn == 0.0);
    int it = 0;
    while (!ok && it < maxit) {
        matvec(J, n, x, w);
        #pragma omp parallel for schedule(static)
        for (int64_t k = 0; k < n; ++k) V[k] = b[k] - w[k];
        const double beta = sqrt(dot_blocks(V, V, n, part));
        if (beta <= rtol * bn) { ok = 1; break; }
        #pragma omp parallel for schedule(static)
        for (int64_t k = 0; k < n; ++k) V[k] /= beta;
        memset(g, 0, sizeof(double) * (m + 1));
        g[0] = beta;
        int j;
        for (j = 0; j < m && it < maxit; ++j, ++it) {
            double* vj = V + (size_t)j * n;
            double* vn = V + (size_t)(j + 1) * n;
            #pragma omp parallel for schedule(static)
            for (int64_t k = 0; k < n; ++k) t[k] = vj[k] * dinv[k];
            matvec(J, n, t, vn);
            for (int i = 0; i <= j; ++i) {
                const double* vi = V + (size_t)i * n;
                const double hij = dot_blocks(vn, vi, n, part);
                Hm[i * m + j] = hij;
                #pragma omp parallel for schedule(static)
                for (int64_t k = 0; k < n; ++k) vn[k] -= hij * vi[k];
            }
            const double hn = sqrt(dot_blocks(vn, vn, n, part));
            Hm[(j + 1) * m + j] = hn;
            if (hn > 0.0) {
                #pragma omp parallel for schedule(static)
                for (int64_t k = 0; k < n; ++k) vn[k] /= hn;
            }
            for (int i = 0; i < j; ++i) {
                const double a0 = Hm[i * m + j], a1 = Hm[(i + 1) * m + j];
                Hm[i * m + j] = cs[i] * a0 + sn[i] * a1;
                Hm[(i + 1) * m + j] = -sn[i] * a0 + cs[i] * a1;
            }
            const double a0 = Hm[j * m + j], a1 = Hm[(j + 1) * m + j];
            const double rr = sqrt(a0 * a0 + a1 * a1);
            cs[j] = rr > 0 ? a0 / rr : 1.0;
            sn[j] = rr > 0 ? a1 / rr : 0.0;
            Hm[j * m + j] = rr;
            Hm[(j + 1) * m + j] = 0.0;
            g[j + 1] = -sn[j] * g[j];
            g[j] = cs[j] * g[j];
            if (fabs(g[j + 1]) <= rtol * bn) { ++j; ++it; break; }
        }
        /* back substitution, x += M^-1 V y */
        double* y = (double*)malloc(sizeof(double) * (size_t)(j > 0 ? j : 1));
        for (int i = j - 1; i >= 0; --i) {
            double s = g[i];
            for (int q = i + 1; q < j; ++q) s -= Hm[i * m + q] * y[q];
            y[i] = Hm[i * m + i] != 0.0 ? s / Hm[i * m + i] : 0.0;
        }
        #pragma omp parallel for schedule(static)
        for (int64_t k = 0; k < n; ++k) {
            double s = 0.0;
            for (int i = 0; i < j; ++i) s += V[(size_t)i * n + k] * y[i];
            x[k] += s * dinv[k];
        }
        free(y);
    }
    if (!ok) {   /* final check */
        matvec(J, n, x, w);
        #pragma omp parallel for schedule(static)
        for (int64_t k = 0; k < n; ++k) t[k] = b[k] - w[k];
        ok = sqrt(dot_blocks(t, t, n, part)) <= rtol * bn * 10.0;
    }
    free(part);
    free(dinv); free(V); free(Hm); free(cs); free(sn); free(g); free(w); free(t);
    return ok ? 0 : 1;
}

/* compute_hillslope_newton (one backward-Euler solve); z in/out.
 * Returns iterations (>= 0) or -1 when not converged. */
int orc_hillslope_newton(int nr, int nc, double h, const uint8_t* mask,
                         const char* bc, const uint8_t* is_channel,
                         const orc_hs_params* p, double inv_dt,
                         const double* zn, double* z, orc_hs_result* res)
{
    hs_t s;
    s.d.nr = nr; s.d.nc = nc; s.d.bn = bc[0]; s.d.be = bc[1]; s.d.bs = bc[2]; s.d.bw = bc[3];
    s.d.mask = mask; s.d.h = h;
    s.D = p->D; s.U = p->U; s.Sc = p->Sc; s.inv_dt = inv_dt; s.cap = p->slope_ratio_cap;
    s.nonlinear = p->nonlinear; s.zn = zn;
    const int64_t n = (int64_t)nr * nc;
    s.dof_of = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
    s.rc = (int*)malloc(sizeof(int) * 2 * (size_t)n);
    s.ndof = 0;
    for (int r = 0; r < nr; ++r)
        for (int c = 0; c < nc; ++c) {
            const int64_t i = (int64_t)r * nc + c;
            s.dof_of[i] = -1;
            if (!mask[i]) continue;
            if (is_channel && is_channel[i]) continue;
            if ((r == 0 && bc[0] == 'o') || (r == nr - 1 && bc[2] == 'o') ||
                (c == 0 && bc[3] == 'o') || (c == nc - 1 && bc[1] == 'o')) continue;
            s.dof_of[i] = s.ndof;
            s.rc[2 * s.ndof] = r;
            s.rc[2 * s.ndof + 1] = c;
            ++s.ndof;
        }
    memset(res, 0, sizeof(*res));
    if (s.ndof == 0) {
        free(s.dof_of); free(s.rc);
        res->converged = 1;
        return 0;
    }
    const int64_t nd = s.ndof;
    double* F = (double*)malloc(sizeof(double) * (size_t)nd);
    double* Ft = (double*)malloc(sizeof(double) * (size_t)nd);
    double* step = (double*)malloc(sizeof(double) * (size_t)nd);
    double* rhs = (double*)malloc(sizeof(double) * (size_t)nd);
    double* zt = (double*)malloc(sizeof(double) * (size_t)n);
    row_t* J = (row_t*)malloc(sizeof(row_t) * (size_t)nd);
    double fnorm = residual(&s, z, F);
    res->initial_residual = fnorm;
    const double ref = fnorm > fabs(p->U) ? fnorm : fabs(p->U);
    double tol = p->rel_tol * ref;
    if (p->abs_tol > tol) tol = p->abs_tol;
    res->tolerance = tol;
    const double feasible = sqrt(s.cap);
    int it = 0, converged = 0;
    for (; it < p->max_iterations; ++it) {
        if (fnorm <= tol) { converged = 1; break; }
        assemble(&s, z, J);
        #pragma omp parallel for schedule(static)
        for (int64_t k = 0; k < nd; ++k) rhs[k] = -F[k];
        if (gmres(J, nd, rhs, step, 1e-13, 20000) != 0) break;
        double lambda = 1.0, fnorm_try = fnorm, max_dz = 0.0;
        int accepted = 0;
        const double ratio_now = s.nonlinear ? max_ratio(&s, z) : 0.0;
        const int restoring = s.nonlinear && ratio_now > feasible;
        for (int ls = 0; ls < p->max_line_search; ++ls) {
            memcpy(zt, z, sizeof(double) * (size_t)n);
            max_dz = 0.0;
            #pragma omp parallel for schedule(static) reduction(max:max_dz)
            for (int64_t k = 0; k < nd; ++k) {
                const double dz = lambda * step[k];
                zt[(int64_t)s.rc[2 * k] * nc + s.rc[2 * k + 1]] += dz;
                if (fabs(dz) > max_dz) max_dz = fabs(dz);
            }
            if (s.nonlinear) {
                const double rt = max_ratio(&s, zt);
                if (restoring) {
                    if (rt >= ratio_now) { lambda *= 0.5; ++res->line_search_steps; continue; }
                    fnorm_try = residual(&s, zt, Ft);
                    accepted = 1;
                    break;
                }
                if (rt > feasible) { lambda *= 0.5; ++res->line_search_steps; continue; }
            }
            fnorm_try = residual(&s, zt, Ft);
            if (fnorm_try < fnorm) { accepted = 1; break; }
            lambda *= 0.5;
            ++res->line_search_steps;
        }
        if (!accepted) break;
        memcpy(z, zt, sizeof(double) * (size_t)n);
        memcpy(F, Ft, sizeof(double) * (size_t)nd);
        fnorm = fnorm_try;
        res->final_max_dz = max_dz;
        if (fnorm <= tol) { converged = 1; ++it; break; }
    }
    res->iterations = it;
    res->final_residual = fnorm;
    res->converged = converged;
    res->max_slope_ratio = max_ratio(&s, z);
    free(F); free(Ft); free(step); free(rhs); free(zt); free(J);
    free(s.dof_of); free(s.rc);
    return converged ? it : -1;
}

/* compute_hillslope_newton_transient, lsdtt_hillslope_newton_cpu.cpp:690-788 */
int orc_hillslope_transient(int nr, int nc, double h, const uint8_t* mask,
                            const char* bc, const uint8_t* is_channel,
                            const orc_hs_params* p, double dt_total,
                            double dt_init, double* z, orc_hs_result* total)
{
    const int64_t n = (int64_t)nr * nc;
    memset(total, 0, sizeof(*total));
    if (!(dt_total > 0.0)) { total->converged = 1; return 0; }
    double dt = dt_init > 0.0 ? dt_init : dt_total;
    const double dt_min = dt_total * 1e-9, dt_max = dt_total;
    if (dt > dt_max) dt = dt_max;
    double* zn = (double*)malloc(sizeof(double) * (size_t)n);
    double* cur = (double*)malloc(sizeof(double) * (size_t)n);
    memcpy(zn, z, sizeof(double) * (size_t)n);
    memcpy(cur, z, sizeof(double) * (size_t)n);
    double remaining = dt_total;
    int rejects = 0, ok = 1;
    while (remaining > 0.0) {
        const double step = dt < remaining ? dt : remaining;
        double* trial = (double*)malloc(sizeof(double) * (size_t)n);
        memcpy(trial, cur, sizeof(double) * (size_t)n);
        orc_hs_result r;
        const int itn = orc_hillslope_newton(nr, nc, h, mask, bc, is_channel, p, 1.0 / step, zn, trial, &r);
        total->iterations += r.iterations;
        total->line_search_steps += r.line_search_steps;
        if (total->substeps == 0 && rejects == 0) total->initial_residual = r.initial_residual;
        if (itn < 0) {
            free(trial);
            ++rejects;
            if (rejects > 40 || step <= dt_min) { ok = 0; break; }
            dt = step * 0.5 > dt_min ? step * 0.5 : dt_min;
            continue;
        }
        memcpy(cur, trial, sizeof(double) * (size_t)n);
        memcpy(zn, trial, sizeof(double) * (size_t)n);
        free(trial);
        remaining -= step;
        ++total->substeps;
        total->final_residual = r.final_residual;
        total->tolerance = r.tolerance;
        total->max_slope_ratio = r.max_slope_ratio;
        if (r.iterations <= 4) dt = step * 1.7 < dt_max ? step * 1.7 : dt_max;
        else if (r.iterations >= 15) dt = step * 0.5 > dt_min ? step * 0.5 : dt_min;
    }
    total->converged = ok;
    total->rejected = rejects;
    if (ok) memcpy(z, cur, sizeof(double) * (size_t)n);
    free(zn);
    free(cur);
    return ok ? 0 : -1;
}

/* ----------------------------------------------------------------------- */
/* LEM step                                                                */
/* ----------------------------------------------------------------------- */

int64_t orc_boundary_flats(int nr, int nc, const uint8_t* mask, const char* bc,
                           int64_t* out)
{
    const int fn = bc[0] == 'o' || bc[0] == 'b', fe = bc[1] == 'o' || bc[1] == 'b';
    const int fs = bc[2] == 'o' || bc[2] == 'b', fw = bc[3] == 'o' || bc[3] == 'b';
    int64_t k = 0;
    for (int r = 0; r < nr; ++r)
        for (int c = 0; c < nc; ++c) {
            const int64_t f = (int64_t)r * nc + c;
            if (!mask[f]) continue;
            int edge = 0;
            for (int dr = -1; dr <= 1 && !edge; ++dr)
                for (int dc = -1; dc <= 1; ++dc) {
                    if (dr == 0 && dc == 0) continue;
                    const int r2 = r + dr, c2 = c + dc;
                    if (r2 < 0) { if (fn) edge = 1; }
                    else if (r2 >= nr) { if (fs) edge = 1; }
                    else if (c2 < 0) { if (fw) edge = 1; }
                    else if (c2 >= nc) { if (fe) edge = 1; }
                    else if (!mask[(int64_t)r2 * nc + c2]) {
                        if ((dr == -1 && fn) || (dr == 1 && fs) || (dc == -1 && fw) || (dc == 1 && fe))
                            edge = 1;
                    }
                }
            if (edge) out[k++] = f;
        }
    return k;
}

static int cmp_float(const void* a, const void* b)
{
    const float x = *(const float*)a, y = *(const float*)b;
    return (x > y) - (x < y);
}

/* smooth_drainage_area, lsdtt_stream_power_cpu.cpp:34-139 */
static int wrap_or_none(int i, int size, int periodic_neg, int periodic_pos)
{
    if (i < 0 && !periodic_neg) return -1;
    if (i >= size && !periodic_pos) return -1;
    return (i % size + size) % size;
}

int orc_smooth_area(int nr, int nc, const double* area, const uint8_t* mask,
                    int kernel_size, const char* bc, double* out)
{
    if (kernel_size < 1 || kernel_size % 2 == 0) return ORC_ERR_ARG;
    const int64_t n = (int64_t)nr * nc;
    if (kernel_size == 1) {
        memcpy(out, area, sizeof(double) * (size_t)n);
        return ORC_OK;
    }
    const int rad = kernel_size / 2;
    const int pn = bc[0] == 'p' || bc[0] == 'P', pe = bc[1] == 'p' || bc[1] == 'P';
    const int ps = bc[2] == 'p' || bc[2] == 'P', pw = bc[3] == 'p' || bc[3] == 'P';
    double* hs = (double*)calloc((size_t)n, sizeof(double));
    int64_t* hc = (int64_t*)calloc((size_t)n, sizeof(int64_t));
    for (int r = 0; r < nr; ++r) {
        double sum = 0.0;
        int64_t cnt = 0;
        for (int o = -rad; o <= rad; ++o) {
            const int c = wrap_or_none(o, nc, pw, pe);
            if (c >= 0 && mask[(int64_t)r * nc + c]) { sum += area[(int64_t)r * nc + c]; ++cnt; }
        }
        for (int c = 0; c < nc; ++c) {
            const int64_t f = (int64_t)r * nc + c;
            hs[f] = sum;
            hc[f] = cnt;
            if (c + 1 < nc) {
                const int cm = wrap_or_none(c - rad, nc, pw, pe);
                if (cm >= 0 && mask[(int64_t)r * nc + cm]) { sum += -1.0 * area[(int64_t)r * nc + cm]; --cnt; }
                const int cp = wrap_or_none(c + rad + 1, nc, pw, pe);
                if (cp >= 0 && mask[(int64_t)r * nc + cp]) { sum += area[(int64_t)r * nc + cp]; ++cnt; }
            }
        }
    }
    for (int64_t i = 0; i < n; ++i) out[i] = 0.0;
    for (int c = 0; c < nc; ++c) {
        double sum = 0.0;
        int64_t cnt = 0;
        for (int o = -rad; o <= rad; ++o) {
            const int r = wrap_or_none(o, nr, pn, ps);
            if (r >= 0) { sum += hs[(int64_t)r * nc + c]; cnt += hc[(int64_t)r * nc + c]; }
        }
        for (int r = 0; r < nr; ++r) {
            const int64_t f = (int64_t)r * nc + c;
            if (mask[f] && cnt != 0) out[f] = sum / (double)cnt;
            if (r + 1 < nr) {
                const int rm = wrap_or_none(r - rad, nr, pn, ps);
                if (rm >= 0) { sum += -1.0 * hs[(int64_t)rm * nc + c]; cnt -= hc[(int64_t)rm * nc + c]; }
                const int rp = wrap_or_none(r + rad + 1, nr, pn, ps);
                if (rp >= 0) { sum += hs[(int64_t)rp * nc + c]; cnt += hc[(int64_t)rp * nc + c]; }
            }
        }
    }
    free(hs);
    free(hc);
    return ORC_OK;
}

int orc_lem_step(int nr, int nc, double dx, const uint8_t* mask,
                 const orc_lem_params* p, const int64_t* boundary_flats,
                 int64_t n_bf, const double* K_field, double* z,
                 orc_lem_stats* st)
{
    return orc_lem_step_ex(nr, nc, dx, mask, p, boundary_flats, n_bf, K_field, z, st, NULL);
}

/* run_lem_step; erosion_rate (optional, float[n]): the per-pixel field of
 * compute_erosion_rate_field (lsdtt_lem_cpu.cpp:666-696) */
int orc_lem_step_ex(int nr, int nc, double dx, const uint8_t* mask,
                    const orc_lem_params* p, const int64_t* boundary_flats,
                    int64_t n_bf, const double* K_field, double* z,
                    orc_lem_stats* st, float* erosion_rate)
{
    const int64_t n = (int64_t)nr * nc;
    memset(st, 0, sizeof(*st));
    double* z_before = (double*)malloc(sizeof(double) * (size_t)n);
    memcpy(z_before, z, sizeof(double) * (size_t)n);
    double* bel = (double*)malloc(sizeof(double) * (size_t)(n_bf ? n_bf : 1));
    for (int64_t i = 0; i < n_bf; ++i) if (boundary_flats[i] < n) bel[i] = z[boundary_flats[i]];
    int status = ORC_OK;

    if (p->enable_fluvial && p->fluvial_dinf) {
        /* run_lem_step's D-inf branch (lsdtt_lem_cpu.cpp:357-391): the
         * D-inf graph of the input surface (a float raster in the reference),
         * its accumulation, the smoothing, then the incision */
        float* zf = (float*)malloc(sizeof(float) * (size_t)n);
        int64_t* rec = (int64_t*)malloc(sizeof(int64_t) * (size_t)(2 * n));
        float* w = (float*)malloc(sizeof(float) * (size_t)(2 * n));
        uint8_t* nrec = (uint8_t*)malloc((size_t)n);
        int64_t* topo = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
        double* area = (double*)malloc(sizeof(double) * (size_t)n);
        for (int64_t i = 0; i < n; ++i) zf[i] = (float)z_before[i];
        if (orc_dinf_graph(nr, nc, dx, zf, mask, NAN, p->bc, 1, 0, rec, w, nrec, NULL, NULL) != ORC_OK) {
            status = ORC_ERR_BC;
        } else {
            const int64_t nt = orc_dinf_topo(n, mask, rec, w, nrec, topo);
            orc_dinf_accumulate(n, rec, w, nrec, topo, nt, dx, area);
            if (p->area_kernel > 1) {
                double* sm = (double*)malloc(sizeof(double) * (size_t)n);
                if (orc_smooth_area(nr, nc, area, mask, p->area_kernel, p->bc, sm) != ORC_OK) status = ORC_ERR_ARG;
                memcpy(area, sm, sizeof(double) * (size_t)n);
                free(sm);
            }
            orc_sp_params sp = {p->K, p->m, p->n, p->dt, dx, p->threshold,
                                p->divide_area_by_dx, p->slope_clamp, p->width_factor};
            st->cells_eroded = orc_stream_power_dinf(n, mask, rec, w, nrec, topo, nt, area, K_field, &sp, z_before, z);
        }
        free(zf); free(rec); free(w); free(nrec); free(topo); free(area);
    } else if (p->enable_fluvial) {
        int64_t* rcv = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
        uint8_t* dir = (uint8_t*)malloc((size_t)n);
        uint8_t* don = (uint8_t*)malloc((size_t)n);
        int64_t* outl = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
        int64_t* topo = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
        int64_t* cnt = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
        double* area = (double*)malloc(sizeof(double) * (size_t)n);
        orc_d8_receivers(nr, nc, dx, z_before, mask, p->bc, 1, 1, NULL, rcv, dir, NULL);
        orc_donors_outlets(nr, nc, mask, rcv, don, outl);
        const int64_t nt = orc_topo_kahn(nr, nc, mask, rcv, don, topo);
        orc_upstream_counts(n, mask, rcv, topo, nt, cnt);
        for (int64_t i = 0; i < n; ++i) area[i] = (double)cnt[i] * (dx * dx);
        if (p->area_kernel > 1) {
            double* sm = (double*)malloc(sizeof(double) * (size_t)n);
            if (orc_smooth_area(nr, nc, area, mask, p->area_kernel, p->bc, sm) != ORC_OK) status = ORC_ERR_ARG;
            memcpy(area, sm, sizeof(double) * (size_t)n);
            free(sm);
        }
        orc_sp_params sp = {p->K, p->m, p->n, p->dt, dx, p->threshold,
                            p->divide_area_by_dx, p->slope_clamp, p->width_factor};
        st->cells_eroded = orc_stream_power_d8(n, mask, rcv, dir, topo, nt, area, K_field, &sp, z_before, z);
        free(rcv); free(dir); free(don); free(outl); free(topo); free(cnt); free(area);
    }

    /* the hillslope's Dirichlet channel set (lsdtt_lem_cpu.cpp:86-134,
     * 397-432): none in the coupled regime, else the explicit mask, else the
     * drainage-area threshold on the D-inf graph of the post-incision
     * surface (a float raster in the reference) */
    const uint8_t* channel = NULL;
    uint8_t* derived = NULL;
    if (p->enable_hillslope) {
        const int regime = p->regime != 0 ? p->regime : ((p->threshold > 0.0 || p->channel) ? 2 : 1);
        if (regime == 2 && p->channel) {
            channel = p->channel;
        } else if (regime == 2 && p->threshold > 0.0) {
            float* zf = (float*)malloc(sizeof(float) * (size_t)n);
            for (int64_t i = 0; i < n; ++i) zf[i] = (float)z[i];
            derived = (uint8_t*)malloc((size_t)n);
            if (orc_dinf_channel_mask(nr, nc, dx, zf, mask, p->bc, p->threshold, p->divide_area_by_dx, 0, derived,
                                      NULL) != ORC_OK)
                status = ORC_ERR_ARG;
            channel = derived;
            free(zf);
        }
    }

    if (p->enable_hillslope) {
        /* Newton first, Picard when it fails or is forced (lsdtt_lem_cpu.cpp:531-582) */
        int done = 0;
        if (!p->force_picard) {
            double* keep = (double*)malloc(sizeof(double) * (size_t)n);
            memcpy(keep, z, sizeof(double) * (size_t)n);
            orc_hs_result hr;
            const int rc = orc_hillslope_transient(nr, nc, dx, mask, p->bc, channel, &p->hs, p->dt, p->dt, z, &hr);
            st->newton_iterations = hr.iterations;
            st->newton_substeps = hr.substeps;
            st->hillslope_residual = hr.final_residual;
            st->hillslope_initial_residual = hr.initial_residual;
            if (rc == 0) done = 1;
            else memcpy(z, keep, sizeof(double) * (size_t)n);   /* the Newton result is discarded */
            free(keep);
        }
        if (!done) {
            orc_picard_params pp = p->picard;
            pp.D = p->hs.D;
            pp.U = 0.0;
            pp.Sc = p->hs.Sc;
            pp.nonlinear = p->hs.nonlinear;
            orc_picard_result pr;
            orc_picard_transient(nr, nc, dx, mask, p->bc, channel, &pp, p->dt, z, &pr);
            st->hillslope_method = 1;
            st->picard_iterations = pr.picard_iters;
            st->picard_substeps = pr.substeps;
        }
    }

    if (p->enable_uplift) {
        const double udz = p->U * p->dt;
        for (int64_t i = 0; i < n; ++i) if (mask[i]) z[i] += udz;
    }
    const double bdz = p->baselevel_fall * p->dt;
    for (int64_t i = 0; i < n_bf; ++i) if (boundary_flats[i] < n) z[boundary_flats[i]] = bel[i] - bdz;

    double sb = 0, sa = 0, mb = -1e30, ma = -1e30, ml = 0, se = 0, me = 0;
    int64_t nv = 0;
    float* lc = (float*)malloc(sizeof(float) * (size_t)(n ? n : 1));
    const double udz = p->enable_uplift ? p->U * p->dt : 0.0;
    const double inv_dt_rate = 1.0 / p->dt;
    if (erosion_rate)
        for (int64_t i = 0; i < n; ++i) erosion_rate[i] = 0.0f;
    for (int64_t i = 0; i < n; ++i) {
        if (!mask[i]) continue;
        const double zb = z_before[i], za = z[i];
        sb += zb; sa += za;
        if (zb > mb) mb = zb;
        if (za > ma) ma = za;
        const double l = fabs(za - zb);
        if (l > ml) ml = l;
        lc[nv++] = (float)l;
        const double ero = zb - (za - udz);
        /* the field keeps the sign: deposition is a negative rate */
        if (erosion_rate) erosion_rate[i] = (float)(ero * inv_dt_rate);
        if (ero > 0.0) { se += ero; if (ero > me) me = ero; }
    }
    const double inv_dt = 1.0 / p->dt, inv_n = nv > 0 ? 1.0 / (double)nv : 0.0;
    st->delta_max_elevation = fabs(ma - mb);
    st->delta_mean_elevation = fabs((sa - sb) * inv_n);
    st->max_local_elevation_change = ml;
    if (nv > 0) {
        qsort(lc, (size_t)nv, sizeof(float), cmp_float);
        st->p999_local_elevation_change = lc[(int64_t)floor(0.999 * (double)(nv - 1))];
    }
    st->mean_erosion_rate = se * inv_n * inv_dt;
    st->max_erosion_rate = me * inv_dt;
    st->n_valid = nv;
    free(lc);
    free(z_before);
    free(bel);
    free(derived);
    return status;
}
