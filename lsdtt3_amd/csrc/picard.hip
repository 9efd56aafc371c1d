// picard.hip -- damped Picard backward-Euler hillslope solver (see picard.hpp).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "picard.hpp"
#include "reduce.hpp"

namespace lsdtt_gpu {

namespace {

constexpr int kB = 256;

struct PgArgs {
    int nr, nc;
    BC bc;
    const uint8_t* mask;
    const uint8_t* dof;
    double h, D, Sc, cap;  // cap = r_max^2
    int nonlinear;
};

// reflecting / periodic neighbour value (lsdtt_hillslope_implicit_cpu.cpp:
// 214-223): off-grid non-periodic and masked neighbours give the cell's own
__device__ __forceinline__ double nbr_z(const PgArgs& a, const double* z, int r, int c, int dr, int dc)
{
    int rr = r + dr, cc = c + dc;
    if (rr < 0) rr = (a.bc.n == 'p') ? a.nr - 1 : r;
    if (rr >= a.nr) rr = (a.bc.s == 'p') ? 0 : r;
    if (cc < 0) cc = (a.bc.w == 'p') ? a.nc - 1 : c;
    if (cc >= a.nc) cc = (a.bc.e == 'p') ? 0 : c;
    const long long f = (long long)rr * a.nc + cc;
    return a.mask[f] ? z[f] : z[(long long)r * a.nc + c];
}

// neighbour across face dir (0 N, 1 E, 2 S, 3 W), -1 for a no-flux face
// (:228-244)
__device__ __forceinline__ long long get_face(const PgArgs& a, int r, int c, int dir)
{
    const int dr = dir == 0 ? -1 : (dir == 2 ? 1 : 0);
    const int dc = dir == 1 ? 1 : (dir == 3 ? -1 : 0);
    const char code = dir == 0 ? a.bc.n : (dir == 1 ? a.bc.e : (dir == 2 ? a.bc.s : a.bc.w));
    int rr = r + dr, cc = c + dc;
    if (rr < 0 || rr >= a.nr || cc < 0 || cc >= a.nc) {
        if (code != 'p') return -1;
        rr = (rr + a.nr) % a.nr;
        cc = (cc + a.nc) % a.nc;
    }
    const long long f = (long long)rr * a.nc + cc;
    return a.mask[f] ? f : -1;
}

__device__ __forceinline__ double grad_x(const PgArgs& a, const double* z, int r, int c)
{
    return (nbr_z(a, z, r, c, 0, 1) - nbr_z(a, z, r, c, 0, -1)) / (2.0 * a.h);
}
__device__ __forceinline__ double grad_y(const PgArgs& a, const double* z, int r, int c)
{
    return (nbr_z(a, z, r, c, -1, 0) - nbr_z(a, z, r, c, 1, 0)) / (2.0 * a.h);
}

// squared face slope ratio (uncapped) of the face s -> sn across dir
__device__ __forceinline__ double face_r2(const PgArgs& a, const double* z, long long s, long long sn, int dir)
{
    const int r = static_cast<int>(s / a.nc), c = static_cast<int>(s - (long long)r * a.nc);
    const int rn = static_cast<int>(sn / a.nc), cn = static_cast<int>(sn - (long long)rn * a.nc);
    const double Sn = (z[sn] - z[s]) / a.h;
    const double Sx = (dir == 1 || dir == 3) ? 0.5 * (grad_y(a, z, r, c) + grad_y(a, z, rn, cn))
                                             : 0.5 * (grad_x(a, z, r, c) + grad_x(a, z, rn, cn));
    return (Sn * Sn + Sx * Sx) / (a.Sc * a.Sc);
}

// beta D_eff / h^2 of the E and S faces of every valid cell (:400-435)
__global__ void __launch_bounds__(kB) k_faces(PgArgs a, const double* __restrict__ z, double beta,
                                              double* __restrict__ wE, double* __restrict__ wS)
{
    const long long n = (long long)a.nr * a.nc;
    for (long long s = blockIdx.x * (long long)blockDim.x + threadIdx.x; s < n;
         s += (long long)gridDim.x * blockDim.x) {
        double we = 0.0, ws = 0.0;
        if (a.mask[s]) {
            const int r = static_cast<int>(s / a.nc), c = static_cast<int>(s - (long long)r * a.nc);
            const long long se = get_face(a, r, c, 1), ss = get_face(a, r, c, 2);
            if (se >= 0) {
                double deff = a.D;
                if (a.nonlinear) {
                    double r2 = face_r2(a, z, s, se, 1);
                    if (r2 > a.cap) r2 = a.cap;
                    deff = a.D / (1.0 - r2);
                }
                we = beta * deff / (a.h * a.h);
            }
            if (ss >= 0) {
                double deff = a.D;
                if (a.nonlinear) {
                    double r2 = face_r2(a, z, s, ss, 2);
                    if (r2 > a.cap) r2 = a.cap;
                    deff = a.D / (1.0 - r2);
                }
                ws = beta * deff / (a.h * a.h);
            }
        }
        wE[s] = we;
        wS[s] = ws;
    }
}

// the four face weights of cell s (N, E, S, W) and their neighbours
__device__ __forceinline__ void faces4(const PgArgs& a, const double* wE, const double* wS, long long s, double w[4],
                                       long long nb[4])
{
    const int r = static_cast<int>(s / a.nc), c = static_cast<int>(s - (long long)r * a.nc);
#pragma unroll
    for (int d = 0; d < 4; ++d) nb[d] = get_face(a, r, c, d);
    w[0] = nb[0] >= 0 ? wS[nb[0]] : 0.0;  // N face = the north cell's S face
    w[1] = nb[1] >= 0 ? wE[s] : 0.0;
    w[2] = nb[2] >= 0 ? wS[s] : 0.0;
    w[3] = nb[3] >= 0 ? wE[nb[3]] : 0.0;  // W face = the west cell's E face
}

// b = z_old + beta U + fixed-neighbour terms, y = z (warm start), 1/diag
// (:475-500, 586-616)
__global__ void __launch_bounds__(kB) k_rhs(PgArgs a, const double* __restrict__ z,
                                            const double* __restrict__ z_old, double betaU,
                                            const double* __restrict__ wE, const double* __restrict__ wS,
                                            double* __restrict__ b, double* __restrict__ y, double* __restrict__ dinv)
{
    const long long n = (long long)a.nr * a.nc;
    for (long long s = blockIdx.x * (long long)blockDim.x + threadIdx.x; s < n;
         s += (long long)gridDim.x * blockDim.x) {
        if (!a.dof[s]) {
            b[s] = 0.0;
            y[s] = z[s];
            dinv[s] = 0.0;
            continue;
        }
        double w[4];
        long long nb[4];
        faces4(a, wE, wS, s, w, nb);
        double diag = 1.0, bd = 0.0;
#pragma unroll
        for (int d = 0; d < 4; ++d) {
            if (nb[d] < 0) continue;
            diag += w[d];
            if (!a.dof[nb[d]]) bd += w[d] * z[nb[d]];
        }
        b[s] = (z_old[s] + betaU) + bd;
        y[s] = z[s];
        dinv[s] = 1.0 / diag;
    }
}

// (A x)_s over the unknowns: x_s (1 + sum w) - sum over unknown neighbours
__device__ __forceinline__ double apply_A(const PgArgs& a, const double* wE, const double* wS, const double* x,
                                          long long s)
{
    double w[4];
    long long nb[4];
    faces4(a, wE, wS, s, w, nb);
    double acc = x[s];
#pragma unroll
    for (int d = 0; d < 4; ++d) {
        if (w[d] == 0.0) continue;
        acc += w[d] * x[s];
        if (a.dof[nb[d]]) acc -= w[d] * x[nb[d]];
    }
    return acc;
}

enum { S_RZ = 0, S_BB, S_RR, S_ALPHA, S_BETA, S_MAXDZ, S_RATIO2, S_STEPDZ, S_CHANGE, S_COUNT };

struct CgArgs {
    PgArgs a;
    const double *wE, *wS, *b, *dinv;
    double *y, *r, *q, *p, *Ap;
    double* scal;
    double* partials;
    unsigned* ticket;
    int* flags;  // [0] done, [1] iterations
    int max_it;
};

__global__ void __launch_bounds__(kB) k_cg_init(CgArgs g)
{
    const long long n = (long long)g.a.nr * g.a.nc;
    double acc[3] = {0.0, 0.0, 0.0};
    for (long long s = blockIdx.x * (long long)blockDim.x + threadIdx.x; s < n;
         s += (long long)gridDim.x * blockDim.x) {
        double rv = 0.0, qv = 0.0;
        if (g.a.dof[s]) {
            rv = g.b[s] - apply_A(g.a, g.wE, g.wS, g.y, s);
            qv = g.dinv[s] * rv;
            acc[0] += rv * qv;
            acc[1] += g.b[s] * g.b[s];
            acc[2] += rv * rv;
        }
        g.r[s] = rv;
        g.q[s] = qv;
        g.p[s] = qv;
    }
    const bool mx[3] = {false, false, false};
    double out[3];
    if (grid_reduce<3>(acc, mx, g.partials, g.ticket, out)) {
        g.scal[S_RZ] = out[0];
        g.scal[S_BB] = out[1];
        g.scal[S_RR] = out[2];
        g.flags[0] = (out[2] <= 1e-28 * (out[1] > 0.0 ? out[1] : 1.0)) ? 1 : 0;
        g.flags[1] = 0;
    }
}

__global__ void __launch_bounds__(kB) k_cg_Ap(CgArgs g)
{
    if (g.flags[0]) return;
    const long long n = (long long)g.a.nr * g.a.nc;
    double acc[1] = {0.0};
    for (long long s = blockIdx.x * (long long)blockDim.x + threadIdx.x; s < n;
         s += (long long)gridDim.x * blockDim.x) {
        double v = 0.0;
        if (g.a.dof[s]) {
            v = apply_A(g.a, g.wE, g.wS, g.p, s);
            acc[0] += g.p[s] * v;
        }
        g.Ap[s] = v;
    }
    const bool mx[1] = {false};
    double out[1];
    if (grid_reduce<1>(acc, mx, g.partials, g.ticket, out)) {
        if (out[0] <= 0.0) g.flags[0] = 1;
        else g.scal[S_ALPHA] = g.scal[S_RZ] / out[0];
    }
}

__global__ void __launch_bounds__(kB) k_cg_update(CgArgs g)
{
    if (g.flags[0]) return;
    const long long n = (long long)g.a.nr * g.a.nc;
    const double al = g.scal[S_ALPHA];
    double acc[2] = {0.0, 0.0};
    for (long long s = blockIdx.x * (long long)blockDim.x + threadIdx.x; s < n;
         s += (long long)gridDim.x * blockDim.x) {
        if (!g.a.dof[s]) continue;
        g.y[s] += al * g.p[s];
        const double rv = g.r[s] - al * g.Ap[s];
        const double qv = g.dinv[s] * rv;
        g.r[s] = rv;
        g.q[s] = qv;
        acc[0] += rv * qv;
        acc[1] += rv * rv;
    }
    const bool mx[2] = {false, false};
    double out[2];
    if (grid_reduce<2>(acc, mx, g.partials, g.ticket, out)) {
        g.scal[S_BETA] = out[0] / g.scal[S_RZ];
        g.scal[S_RZ] = out[0];
        g.scal[S_RR] = out[1];
        const int it = g.flags[1] + 1;
        g.flags[1] = it;
        if (out[1] <= 1e-28 * (g.scal[S_BB] > 0.0 ? g.scal[S_BB] : 1.0) || it >= g.max_it) g.flags[0] = 1;
    }
}

__global__ void __launch_bounds__(kB) k_cg_p(CgArgs g)
{
    if (g.flags[0]) return;
    const long long n = (long long)g.a.nr * g.a.nc;
    const double be = g.scal[S_BETA];
    for (long long s = blockIdx.x * (long long)blockDim.x + threadIdx.x; s < n;
         s += (long long)gridDim.x * blockDim.x)
        if (g.a.dof[s]) g.p[s] = g.q[s] + be * g.p[s];
}

// z += omega (y - z) on the unknowns; max |dz| (:568-578)
__global__ void __launch_bounds__(kB) k_update(PgArgs a, double* __restrict__ z, const double* __restrict__ y,
                                               double omega, double* partials, unsigned* ticket, double* out_slot)
{
    const long long n = (long long)a.nr * a.nc;
    double acc[1] = {0.0};
    for (long long s = blockIdx.x * (long long)blockDim.x + threadIdx.x; s < n;
         s += (long long)gridDim.x * blockDim.x) {
        if (!a.dof[s]) continue;
        const double zn = z[s] + omega * (y[s] - z[s]);
        acc[0] = fmax(acc[0], fabs(zn - z[s]));
        z[s] = zn;
    }
    const bool mx[1] = {true};
    double out[1];
    if (grid_reduce<1>(acc, mx, partials, ticket, out)) *out_slot = out[0];
}

// max uncapped face ratio^2 over the E and S faces of valid cells (:442-468)
__global__ void __launch_bounds__(kB) k_ratio2(PgArgs a, const double* __restrict__ z, double* partials,
                                               unsigned* ticket, double* out_slot)
{
    const long long n = (long long)a.nr * a.nc;
    double acc[1] = {0.0};
    for (long long s = blockIdx.x * (long long)blockDim.x + threadIdx.x; s < n;
         s += (long long)gridDim.x * blockDim.x) {
        if (!a.mask[s]) continue;
        const int r = static_cast<int>(s / a.nc), c = static_cast<int>(s - (long long)r * a.nc);
        const long long se = get_face(a, r, c, 1), ss = get_face(a, r, c, 2);
        if (se >= 0) acc[0] = fmax(acc[0], face_r2(a, z, s, se, 1));
        if (ss >= 0) acc[0] = fmax(acc[0], face_r2(a, z, s, ss, 2));
    }
    const bool mx[1] = {true};
    double out[1];
    if (grid_reduce<1>(acc, mx, partials, ticket, out)) *out_slot = out[0];
}

__global__ void __launch_bounds__(kB) k_step_dz(const uint8_t* __restrict__ dof, const double* __restrict__ z,
                                                const double* __restrict__ z_old, long long n, double* partials,
                                                unsigned* ticket, double* out_slot)
{
    double acc[1] = {0.0};
    for (long long s = blockIdx.x * (long long)blockDim.x + threadIdx.x; s < n;
         s += (long long)gridDim.x * blockDim.x)
        if (dof[s]) acc[0] = fmax(acc[0], fabs(z[s] - z_old[s]));
    const bool mx[1] = {true};
    double out[1];
    if (grid_reduce<1>(acc, mx, partials, ticket, out)) *out_slot = out[0];
}

// one Jacobi sweep of the slope projection (:733-758): an unknown above
// min over its faces of (z_nb + cap_slope h) is lowered to it
__global__ void __launch_bounds__(kB) k_project(PgArgs a, const double* __restrict__ zin, double* __restrict__ zout,
                                                double lift, double* partials, unsigned* ticket, double* out_slot)
{
    const long long n = (long long)a.nr * a.nc;
    double acc[1] = {0.0};
    for (long long s = blockIdx.x * (long long)blockDim.x + threadIdx.x; s < n;
         s += (long long)gridDim.x * blockDim.x) {
        double v = zin[s];
        if (a.dof[s]) {
            const int r = static_cast<int>(s / a.nc), c = static_cast<int>(s - (long long)r * a.nc);
            double za = INFINITY;
            for (int d = 0; d < 4; ++d) {
                const long long sn = get_face(a, r, c, d);
                if (sn >= 0) za = fmin(za, zin[sn] + lift);
            }
            if (v > za) {
                acc[0] = fmax(acc[0], v - za);
                v = za;
            }
        }
        zout[s] = v;
    }
    const bool mx[1] = {true};
    double out[1];
    if (grid_reduce<1>(acc, mx, partials, ticket, out)) *out_slot = out[0];
}

__global__ void k_copy_d(const double* __restrict__ a, double* __restrict__ b, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        b[i] = a[i];
}

}  // namespace

double PicardSolver::read_scalar(int slot, hipStream_t s)
{
    LSDTT_HIP(hipMemcpyAsync(hpin.p, scal.p + slot, sizeof(double), hipMemcpyDeviceToHost, s));
    LSDTT_HIP(hipStreamSynchronize(s));
    return hpin.p[0];
}

static PgArgs pg_args(int nr, int nc, BC bc, const uint8_t* mask, const uint8_t* dof, double h,
                      const PicardOptions& o)
{
    return PgArgs{nr, nc, bc, mask, dof, h, o.D, o.Sc > 0.0 ? o.Sc : 1.0, o.r_max * o.r_max, o.nonlinear};
}

double PicardSolver::max_ratio2(const double* z, hipStream_t s)
{
    k_ratio2<<<grid_, kB, 0, s>>>(pg_args(nr_, nc_, bc_, mask_, dof_, h_, o_), z, partials.p, ticket.p,
                                  scal.p + S_RATIO2);
    LSDTT_LAUNCH_CHECK();
    return read_scalar(S_RATIO2, s);
}

double PicardSolver::step_dz(const double* z, hipStream_t s)
{
    k_step_dz<<<grid_, kB, 0, s>>>(dof_, z, z_old.p, n_, partials.p, ticket.p, scal.p + S_STEPDZ);
    LSDTT_LAUNCH_CHECK();
    return read_scalar(S_STEPDZ, s);
}

int PicardSolver::cg(hipStream_t s)
{
    CgArgs g{pg_args(nr_, nc_, bc_, mask_, dof_, h_, o_), wE.p, wS.p, b.p, dinv.p, y.p, r.p, q.p, p.p, Ap.p,
             scal.p, partials.p, ticket.p, flags.p, 10000};
    k_cg_init<<<grid_, kB, 0, s>>>(g);
    LSDTT_LAUNCH_CHECK();
    int* hf = reinterpret_cast<int*>(hpin.p + 2);
    hf[0] = hf[1] = 0;
    for (int chunk = 4;; chunk = std::min(2 * chunk, 64)) {
        for (int k = 0; k < chunk; ++k) {
            k_cg_Ap<<<grid_, kB, 0, s>>>(g);
            k_cg_update<<<grid_, kB, 0, s>>>(g);
            k_cg_p<<<grid_, kB, 0, s>>>(g);
        }
        LSDTT_LAUNCH_CHECK();
        LSDTT_HIP(hipMemcpyAsync(hf, flags.p, sizeof(int) * 2, hipMemcpyDeviceToHost, s));
        LSDTT_HIP(hipStreamSynchronize(s));
        if (hf[0]) break;
    }
    return hf[1];
}

bool PicardSolver::solve_stage(double beta, double omega, int iter_cap, double* z, int* iters, PicardStats* st,
                               hipStream_t s)
{
    const PgArgs a = pg_args(nr_, nc_, bc_, mask_, dof_, h_, o_);
    *iters = 0;
    for (int k = 0; k < iter_cap; ++k) {
        k_faces<<<grid_, kB, 0, s>>>(a, z, beta, wE.p, wS.p);
        k_rhs<<<grid_, kB, 0, s>>>(a, z, z_old.p, beta * o_.U, wE.p, wS.p, b.p, y.p, dinv.p);
        LSDTT_LAUNCH_CHECK();
        st->cg_iters += cg(s);
        ++*iters;
        k_update<<<grid_, kB, 0, s>>>(a, z, y.p, omega, partials.p, ticket.p, scal.p + S_MAXDZ);
        LSDTT_LAUNCH_CHECK();
        if (!o_.nonlinear) return true;  // linear: one solve is exact
        if (read_scalar(S_MAXDZ, s) < o_.picard_tol) return true;
    }
    return false;
}

void PicardSolver::projection(double* z, hipStream_t s)
{
    // Jacobi sweeps to the fixed point the reference's Gauss-Seidel sweeps
    // approach (the largest field below z with every face slope <= r_max Sc)
    const PgArgs a = pg_args(nr_, nc_, bc_, mask_, dof_, h_, o_);
    const double lift = o_.r_max * (o_.Sc > 0.0 ? o_.Sc : 1.0) * h_;
    zp.reserve(n_);
    for (int sweep = 0; sweep < 4096; ++sweep) {
        k_project<<<grid_, kB, 0, s>>>(a, z, zp.p, lift, partials.p, ticket.p, scal.p + S_CHANGE);
        k_copy_d<<<grid_, kB, 0, s>>>(zp.p, z, n_);
        LSDTT_LAUNCH_CHECK();
        if (read_scalar(S_CHANGE, s) < 1e-12) break;
    }
}

void PicardSolver::run(int nr, int nc, BC bc, double h, const uint8_t* mask, const uint8_t* dof,
                       const PicardOptions& o, double dt_total, double* z, PicardStats* st, hipStream_t s)
{
    nr_ = nr;
    nc_ = nc;
    n_ = (long long)nr * nc;
    bc_ = bc;
    h_ = h;
    mask_ = mask;
    dof_ = dof;
    o_ = o;
    grid_ = std::max(1, std::min(2048, div_up(n_, kB)));
    for (auto* v : {&z_old, &wE, &wS, &b, &y, &r, &q, &p, &Ap, &dinv}) v->reserve(n_);
    partials.reserve((size_t)grid_ * 4 + 8);
    scal.reserve(S_COUNT + 4);
    ticket.reserve(2);
    flags.reserve(4);
    hpin.reserve(8);
    LSDTT_HIP(hipMemsetAsync(ticket.p, 0, sizeof(unsigned) * 2, s));
    *st = PicardStats{};
    const double omega = o.omega > 0.0 ? o.omega : 1.0;
    const int pmax = o.picard_max_iter > 0 ? o.picard_max_iter : 50;
    auto save = [&] {
        k_copy_d<<<grid_, kB, 0, s>>>(z, z_old.p, n_);
        LSDTT_LAUNCH_CHECK();
    };
    double final_dt = 0.0;
    if (!o.nonlinear) {
        // linear: fixed sub-steps, one exact solve each (:857-879)
        double dt_lin = o.dt_init > 0.0 ? std::min(o.dt_init, dt_total) : dt_total;
        int n_lin = static_cast<int>(std::ceil(dt_total / dt_lin));
        if (n_lin < 1) n_lin = 1;
        dt_lin = dt_total / n_lin;
        for (int k = 0; k < n_lin; ++k) {
            save();
            int it = 0;
            solve_stage(dt_lin, 1.0, pmax, z, &it, st, s);
            st->picard_iters += it;
            ++st->substeps;
        }
        final_dt = dt_lin;
    } else {
        // adaptive controller (:880-993)
        const double dt_max = o.dt_max > 0.0 ? o.dt_max : dt_total;
        double dt_min = o.dt_min > 0.0 ? o.dt_min : dt_total * 1e-4;
        dt_min = std::min(dt_min, dt_max);
        const double grow = o.grow > 1.0 ? o.grow : 1.5;
        const double shrink = (o.shrink > 0.0 && o.shrink < 1.0) ? o.shrink : 0.5;
        double cur_omega = omega;
        double dt = o.dt_init > 0.0 ? o.dt_init : dt_total;
        dt = std::min(dt, std::min(dt_max, dt_total));
        double t_rem = dt_total;
        const double t_eps = dt_total * 1e-12;
        double mr2 = max_ratio2(z, s);
        for (int attempts = 0; t_rem > t_eps && attempts < 2000000; ++attempts) {
            dt = std::min(dt, t_rem);
            dt = std::min(dt, dt_max);
            if (dt < dt_min) dt = dt_min;
            save();
            const bool semi = (mr2 < 0.25) && (cur_omega >= omega - 1e-12);
            int it = 0;
            bool conv = solve_stage(dt, semi ? 1.0 : cur_omega, semi ? 1 : pmax, z, &it, st, s);
            if (semi) conv = true;
            const double sdz = step_dz(z, s);
            const double m2 = max_ratio2(z, s);
            const bool slope_ok = m2 < o.r_max * o.r_max;
            const bool dz_ok = o.dz_target <= 0.0 || sdz <= o.dz_target;
            if (conv && slope_ok && dz_ok) {
                t_rem -= dt;
                ++st->substeps;
                st->picard_iters += it;
                final_dt = dt;
                mr2 = m2;
                const bool easy = it <= 5 && (o.dz_target <= 0.0 || sdz <= 0.5 * o.dz_target);
                if (easy) dt = std::min(dt * grow, dt_max);
                cur_omega = std::min(omega, cur_omega * 1.1);
            } else if (dt <= dt_min * (1.0 + 1e-9)) {
                projection(z, s);
                st->limiter_fired = 1;
                mr2 = max_ratio2(z, s);
                t_rem -= dt;
                ++st->substeps;
                st->picard_iters += it;
                final_dt = dt;
            } else {
                k_copy_d<<<grid_, kB, 0, s>>>(z_old.p, z, n_);
                LSDTT_LAUNCH_CHECK();
                ++st->rejected;
                dt = std::max(dt * shrink, dt_min);
                if (!conv) cur_omega = std::max(0.1, cur_omega * 0.7);
            }
        }
        if (final_dt <= 0.0) final_dt = dt;
    }
    st->final_dt = final_dt;
    st->max_slope_ratio = std::sqrt(max_ratio2(z, s));
}

}  // namespace lsdtt_gpu
