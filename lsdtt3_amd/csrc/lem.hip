// lem.hip -- the LEM timestep kernels and host orchestration (see lem.hpp).
//
// Reference semantics restated (src/cpu/):
//   lsdtt_stream_power_cpu.cpp:159-326  D8 implicit stream power: skip
//       outlets, cells below the area threshold and cells whose (already
//       updated) receiver is not lower; n == 1 closed form with the rise
//       guard; n != 1 per-node Newton (<= 100 iterations, stop at
//       slope_clamp * dx_eff) with the rise guard and over-excavation clamp
//   lsdtt_hillslope_operator_cpu.hpp:83-321  'o' edges Dirichlet, 'n'/'b'
//       and masked neighbours resolve to the cell itself, 'p' wraps; face
//       alpha = 1 / (1 - min((Sn^2 + St^2) / Sc^2, cap)), St = mean of the
//       two cells' centred tangential slopes
//   lsdtt_hillslope_newton_cpu.cpp:137-186  dalpha/dz (9-point Jacobian)
//   lsdtt_hillslope_newton_cpu.cpp:283-308  residual
//   lsdtt_hillslope_newton_cpu.cpp:553-633  line search (halving; the
//       centred-gradient feasibility test; restoration mode when the start
//       is super-critical)
//   lsdtt_hillslope_newton_cpu.cpp:690-788  sub-steps (grow x1.7 at <= 4
//       iterations, shrink x0.5 at >= 15, reject and halve on failure)
//   lsdtt_lem_cpu.cpp:614-722  uplift, boundary restore, statistics
//   src/utils/lsdtt_statistics.cpp:420-429  p99.9 = element floor(0.999 (n-1))
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "lem.hpp"
#include "reduce.hpp"

namespace lsdtt_gpu {

namespace {

constexpr int kB = 256;

// scalar slots (double) of the hillslope solver
enum {
    S_RHO = 0, S_RHO_OLD, S_ALPHA, S_OMEGA, S_BNORM2, S_RNORM2, S_FNORM, S_RATIO, S_TMP0, S_TMP1,
    S_SUM_B, S_SUM_A, S_MAX_B, S_MAX_A, S_MAX_L, S_SUM_E, S_MAX_E, S_NV, S_COUNT
};

inline int grid_for(long long n, int cap) { return std::max(1, std::min(cap, div_up(n, kB))); }

__global__ void k_copy(const double* __restrict__ a, double* __restrict__ b, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        b[i] = a[i];
}

__global__ void k_gather(const double* __restrict__ z, const int* __restrict__ idx, long long m,
                         double* __restrict__ out)
{
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i < m) out[i] = z[idx[i]];
}

// ---------------------------------------------------------------------------
// stream power
// ---------------------------------------------------------------------------
struct SpArgs {
    const int* rcv;
    const uint8_t* dir;
    const int* cnt;
    const double* z_old;
    double* z;
    const double* K;  // per-cell field or null
    double K0, m, nexp, dt, dx, pixel_area, threshold, slope_clamp;
    int divide, n_is_one;
    unsigned long long* eroded;
};

__device__ __forceinline__ bool sp_node(const SpArgs& a, int v)
{
    const int rc = a.rcv[v];
    const double dx_eff = (a.dir[v] & 1u) ? a.dx * kSqrt2 : a.dx;
    double area = static_cast<double>(a.cnt[v]) * a.pixel_area;
    if (a.divide) area /= a.dx;
    if (a.threshold > 0.0 && area < a.threshold) return false;
    const double zo = a.z_old[v];
    const double zr = a.z[rc];
    if (zr >= zo) return false;
    const double tol = a.slope_clamp * dx_eff;
    const double Ki = a.K ? a.K[v] : a.K0;
    const double kam = Ki * pow(area, a.m);
    double zn = zo;
    if (a.n_is_one) {
        const double spf = kam * (a.dt / dx_eff);
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
                const double fz = zn - zo + kam * a.dt * pow(slope, a.nexp);
                const double fd = 1.0 + kam * a.dt * (a.nexp / dx_eff) * pow(slope, a.nexp - 1.0);
                eps = fz / fd;
            }
            zn -= eps;
            ++it;
            if (it > 100) eps = tol / 2.0;
        } while (fabs(eps) > tol);
        if ((zn - zo) > tol || isnan(zn)) zn = zo;
        if (zn < zr) zn = zr + a.slope_clamp * dx_eff;
    }
    a.z[v] = zn;
    return true;
}

__device__ __forceinline__ void count_eroded(unsigned long long* ctr, bool e)
{
    const unsigned long long m = __ballot(e);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(ctr, static_cast<unsigned long long>(__popcll(m)));
}

__global__ void k_sp_level(const int* __restrict__ order, int begin, int end, SpArgs a)
{
    const int i = begin + blockIdx.x * blockDim.x + threadIdx.x;
    bool e = false;
    if (i < end) e = sp_node(a, order[i]);
    count_eroded(a.eroded, e);
}

// one workgroup walks levels lo..hi (outlets up): a receiver is always one
// level shallower, so it is final when its donors read it
__global__ void __launch_bounds__(1024) k_sp_tail(const int* __restrict__ order, const int* __restrict__ lstart,
                                                  int lo, int hi, SpArgs a)
{
    for (int L = lo; L <= hi; ++L) {
        const int b = lstart[L], e = lstart[L + 1];
        for (int i0 = b; i0 < e; i0 += blockDim.x) {
            const int i = i0 + threadIdx.x;
            bool er = false;
            if (i < e) er = sp_node(a, order[i]);
            count_eroded(a.eroded, er);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// hillslope operator
// ---------------------------------------------------------------------------
struct HsDom {
    int nr, nc;
    BC bc;
    const uint8_t* mask;
    const uint8_t* dof;
    double h, D, U, Sc, cap, inv_dt;
    int nonlinear;
    const double* zn;
};

// lsdtt_hillslope_domain::neighbour (operator .hpp:146-159)
__device__ __forceinline__ int hs_nbr(const HsDom& d, int r, int c, int dr, int dc)
{
    int rr = r + dr, cc = c + dc;
    if (rr < 0) rr = (d.bc.n == 'p') ? d.nr - 1 : r;
    else if (rr >= d.nr) rr = (d.bc.s == 'p') ? 0 : r;
    if (cc < 0) cc = (d.bc.w == 'p') ? d.nc - 1 : c;
    else if (cc >= d.nc) cc = (d.bc.e == 'p') ? 0 : c;
    int f = rr * d.nc + cc;
    if (!d.mask[f]) f = r * d.nc + c;
    return f;
}

// elevation view: z, or z + lambda * dz for a line-search trial
struct ZView {
    const double* z;
    const float* dz;
    double lam;
    __device__ __forceinline__ double operator()(int i) const
    {
        return dz ? z[i] + lam * static_cast<double>(dz[i]) : z[i];
    }
};

__device__ __forceinline__ double hs_sx(const HsDom& d, const ZView& Z, int r, int c)
{
    return (Z(hs_nbr(d, r, c, 0, 1)) - Z(hs_nbr(d, r, c, 0, -1))) / (2.0 * d.h);
}
__device__ __forceinline__ double hs_sy(const HsDom& d, const ZView& Z, int r, int c)
{
    return (Z(hs_nbr(d, r, c, 1, 0)) - Z(hs_nbr(d, r, c, -1, 0))) / (2.0 * d.h);
}

__constant__ static const int nesw_r[4] = {-1, 0, 1, 0};
__constant__ static const int nesw_c[4] = {0, 1, 0, -1};

// face alphas and normal slopes (operator .hpp:290-321)
__device__ __forceinline__ void hs_faces(const HsDom& d, const ZView& Z, int r, int c, double a[4], double sn[4],
                                         int nb[4])
{
    const double zc = Z(r * d.nc + c);
    const double sxc = d.nonlinear ? hs_sx(d, Z, r, c) : 0.0;
    const double syc = d.nonlinear ? hs_sy(d, Z, r, c) : 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        nb[k] = hs_nbr(d, r, c, nesw_r[k], nesw_c[k]);
        sn[k] = (Z(nb[k]) - zc) / d.h;
        if (!d.nonlinear) {
            a[k] = 1.0;
            continue;
        }
        const int rr = nb[k] / d.nc, cc = nb[k] - rr * d.nc;
        const double st = (k == 0 || k == 2) ? 0.5 * (sxc + hs_sx(d, Z, rr, cc)) : 0.5 * (syc + hs_sy(d, Z, rr, cc));
        double s2 = (sn[k] * sn[k] + st * st) / (d.Sc * d.Sc);
        if (s2 > d.cap) s2 = d.cap;
        a[k] = 1.0 / (1.0 - s2);
    }
}

// residual (newton .cpp:283-308)
__device__ __forceinline__ double hs_residual(const HsDom& d, const ZView& Z, int r, int c)
{
    const double DA = d.D / (d.h * d.h);
    double a[4], sn[4];
    int nb[4];
    hs_faces(d, Z, r, c, a, sn, nb);
    double flux = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) flux += a[k] * sn[k] * d.h;
    double f = -d.U - DA * flux;
    if (d.inv_dt > 0.0) {
        const int i = r * d.nc + c;
        f += d.inv_dt * (Z(i) - d.zn[i]);
    }
    return f;
}

__device__ __forceinline__ double hs_ratio(const HsDom& d, const ZView& Z, int r, int c)
{
    const double sx = hs_sx(d, Z, r, c), sy = hs_sy(d, Z, r, c);
    return sqrt(sx * sx + sy * sy) / d.Sc;
}

// map a resolved target cell to its slot in the 3x3 stencil of (r, c)
__device__ __forceinline__ int hs_slot(const HsDom& d, int r, int c, int t)
{
    const int tr = t / d.nc, tc = t - tr * d.nc;
    int dr = tr - r, dc = tc - c;
    if (dr > 1) dr -= d.nr;
    else if (dr < -1) dr += d.nr;
    if (dc > 1) dc -= d.nc;
    else if (dc < -1) dc += d.nc;
    return (dr + 1) * 3 + (dc + 1);
}

// stencil target of slot k for the matrix-vector product: periodic axes
// wrap, others clamp (a clamped slot always carries a zero coefficient)
__device__ __forceinline__ int hs_target(const HsDom& d, int r, int c, int k)
{
    int rr = r + k / 3 - 1, cc = c + k % 3 - 1;
    if (rr < 0) rr = (d.bc.n == 'p') ? d.nr - 1 : 0;
    else if (rr >= d.nr) rr = (d.bc.s == 'p') ? 0 : d.nr - 1;
    if (cc < 0) cc = (d.bc.w == 'p') ? d.nc - 1 : 0;
    else if (cc >= d.nc) cc = (d.bc.e == 'p') ? 0 : d.nc - 1;
    return rr * d.nc + cc;
}

// Jacobian row and right-hand side (newton .cpp:415-471): 9 FP32
// coefficients per DOF (accumulated in FP64), b = -F, Jacobi inverse
__global__ void k_hs_assemble(HsDom d, const double* __restrict__ z, long long n, float* __restrict__ coef,
                              float* __restrict__ dinv, float* __restrict__ b)
{
    for (long long f = blockIdx.x * (long long)blockDim.x + threadIdx.x; f < n;
         f += (long long)gridDim.x * blockDim.x) {
        const int r = static_cast<int>(f / d.nc), c = static_cast<int>(f - (long long)r * d.nc);
        double c9[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) c9[k] = 0.0;
        float rhs = 0.0f;
        if (d.dof[f]) {
            const ZView Z{z, nullptr, 0.0};
            const double DA = d.D / (d.h * d.h), h = d.h, Sc2 = d.Sc * d.Sc;
            double a[4], sn[4];
            int nb[4];
            hs_faces(d, Z, r, c, a, sn, nb);
            double flux = 0.0;
            for (int k = 0; k < 4; ++k) flux += a[k] * sn[k] * h;
            double F = -d.U - DA * flux;
            if (d.inv_dt > 0.0) F += d.inv_dt * (z[f] - d.zn[f]);
            rhs = static_cast<float>(-F);
            auto add = [&](int t, double w) {
                if (w == 0.0 || !d.dof[t]) return;
                c9[hs_slot(d, r, c, t)] += w;
            };
            const int self = static_cast<int>(f);
            if (d.inv_dt > 0.0) add(self, d.inv_dt);
            for (int q = 0; q < 4; ++q) {
                add(nb[q], -DA * a[q]);
                add(self, +DA * a[q]);
                if (!d.nonlinear) continue;
                const int rr = nb[q] / d.nc, cc = nb[q] - rr * d.nc;
                const bool tx = (q == 0 || q == 2);
                const double st = tx ? 0.5 * (hs_sx(d, Z, r, c) + hs_sx(d, Z, rr, cc))
                                     : 0.5 * (hs_sy(d, Z, r, c) + hs_sy(d, Z, rr, cc));
                const double rraw = (sn[q] * sn[q] + st * st) / Sc2;
                if (rraw >= d.cap) continue;
                const double pref = a[q] * a[q] * 2.0 / Sc2;
                const double dzf = sn[q] * h;
                const double wt = pref * st * (1.0 / (4.0 * h));
                add(self, -DA * dzf * (pref * sn[q] * (-1.0 / h)));
                add(nb[q], -DA * dzf * (pref * sn[q] * (+1.0 / h)));
                if (tx) {
                    add(hs_nbr(d, r, c, 0, 1), -DA * dzf * (+wt));
                    add(hs_nbr(d, r, c, 0, -1), -DA * dzf * (-wt));
                    add(hs_nbr(d, rr, cc, 0, 1), -DA * dzf * (+wt));
                    add(hs_nbr(d, rr, cc, 0, -1), -DA * dzf * (-wt));
                } else {
                    add(hs_nbr(d, r, c, 1, 0), -DA * dzf * (+wt));
                    add(hs_nbr(d, r, c, -1, 0), -DA * dzf * (-wt));
                    add(hs_nbr(d, rr, cc, 1, 0), -DA * dzf * (+wt));
                    add(hs_nbr(d, rr, cc, -1, 0), -DA * dzf * (-wt));
                }
            }
        }
#pragma unroll
        for (int k = 0; k < 9; ++k) coef[k * n + f] = static_cast<float>(c9[k]);
        dinv[f] = (d.dof[f] && c9[4] != 0.0) ? static_cast<float>(1.0 / c9[4]) : 0.0f;
        b[f] = rhs;
    }
}

__device__ __forceinline__ float hs_matvec_row(const HsDom& d, const float* __restrict__ coef, long long n,
                                               const float* __restrict__ x, int f, int r, int c)
{
    float acc = 0.0f;
#pragma unroll
    for (int k = 0; k < 9; ++k) acc += coef[k * n + f] * x[hs_target(d, r, c, k)];
    return acc;
}

// ---- BiCGSTAB (right Jacobi preconditioning), FP32 vectors, FP64 sums ----
struct BiArgs {
    HsDom d;
    long long n;
    const float* coef;
    const float* dinv;
    const float* b;
    float *x, *r, *rhat, *p, *phat, *v, *s, *shat, *t;
    double* scal;
    double* partials;
    unsigned* ticket;
    int* flags;  // [0] done, [1] iterations, [2] breakdown
    double eta2; // (relative tolerance)^2
    int max_it;
};

__global__ void k_bi_init(BiArgs a)
{
    double acc[2] = {0.0, 0.0};
    for (long long f = blockIdx.x * (long long)blockDim.x + threadIdx.x; f < a.n;
         f += (long long)gridDim.x * blockDim.x) {
        const float bv = a.b[f];
        a.r[f] = bv;
        a.rhat[f] = bv;
        a.p[f] = 0.0f;
        a.v[f] = 0.0f;
        a.x[f] = 0.0f;
        acc[0] += static_cast<double>(bv) * bv;
    }
    acc[1] = acc[0];
    const bool mx[2] = {false, false};
    double out[2];
    if (grid_reduce<2>(acc, mx, a.partials, a.ticket, out)) {
        a.scal[S_RHO] = out[0];
        a.scal[S_BNORM2] = out[1];
        a.scal[S_RHO_OLD] = 1.0;
        a.scal[S_ALPHA] = 1.0;
        a.scal[S_OMEGA] = 1.0;
        a.flags[0] = (out[1] == 0.0) ? 1 : 0;
        a.flags[1] = 0;
        a.flags[2] = 0;
    }
}

__global__ void k_bi_p(BiArgs a)
{
    if (a.flags[0]) return;
    const double beta = (a.scal[S_RHO] / a.scal[S_RHO_OLD]) * (a.scal[S_ALPHA] / a.scal[S_OMEGA]);
    const float bt = static_cast<float>(beta), om = static_cast<float>(a.scal[S_OMEGA]);
    for (long long f = blockIdx.x * (long long)blockDim.x + threadIdx.x; f < a.n;
         f += (long long)gridDim.x * blockDim.x) {
        const float pv = a.r[f] + bt * (a.p[f] - om * a.v[f]);
        a.p[f] = pv;
        a.phat[f] = a.dinv[f] * pv;
    }
}

__global__ void k_bi_v(BiArgs a)
{
    if (a.flags[0]) return;
    double acc[1] = {0.0};
    for (long long f = blockIdx.x * (long long)blockDim.x + threadIdx.x; f < a.n;
         f += (long long)gridDim.x * blockDim.x) {
        const int r = static_cast<int>(f / a.d.nc), c = static_cast<int>(f - (long long)r * a.d.nc);
        const float vv = hs_matvec_row(a.d, a.coef, a.n, a.phat, static_cast<int>(f), r, c);
        a.v[f] = vv;
        acc[0] += static_cast<double>(a.rhat[f]) * vv;
    }
    const bool mx[1] = {false};
    double out[1];
    if (grid_reduce<1>(acc, mx, a.partials, a.ticket, out)) {
        if (out[0] == 0.0) {
            a.flags[0] = 1;
            a.flags[2] = 1;
        } else {
            a.scal[S_ALPHA] = a.scal[S_RHO] / out[0];
        }
    }
}

__global__ void k_bi_s(BiArgs a)
{
    if (a.flags[0]) return;
    const float al = static_cast<float>(a.scal[S_ALPHA]);
    for (long long f = blockIdx.x * (long long)blockDim.x + threadIdx.x; f < a.n;
         f += (long long)gridDim.x * blockDim.x) {
        const float sv = a.r[f] - al * a.v[f];
        a.s[f] = sv;
        a.shat[f] = a.dinv[f] * sv;
    }
}

__global__ void k_bi_t(BiArgs a)
{
    if (a.flags[0]) return;
    double acc[2] = {0.0, 0.0};
    for (long long f = blockIdx.x * (long long)blockDim.x + threadIdx.x; f < a.n;
         f += (long long)gridDim.x * blockDim.x) {
        const int r = static_cast<int>(f / a.d.nc), c = static_cast<int>(f - (long long)r * a.d.nc);
        const float tv = hs_matvec_row(a.d, a.coef, a.n, a.shat, static_cast<int>(f), r, c);
        a.t[f] = tv;
        acc[0] += static_cast<double>(tv) * a.s[f];
        acc[1] += static_cast<double>(tv) * tv;
    }
    const bool mx[2] = {false, false};
    double out[2];
    if (grid_reduce<2>(acc, mx, a.partials, a.ticket, out)) a.scal[S_OMEGA] = out[1] > 0.0 ? out[0] / out[1] : 0.0;
}

__global__ void k_bi_x(BiArgs a)
{
    if (a.flags[0]) return;
    const float al = static_cast<float>(a.scal[S_ALPHA]), om = static_cast<float>(a.scal[S_OMEGA]);
    double acc[2] = {0.0, 0.0};
    for (long long f = blockIdx.x * (long long)blockDim.x + threadIdx.x; f < a.n;
         f += (long long)gridDim.x * blockDim.x) {
        a.x[f] += al * a.phat[f] + om * a.shat[f];
        const float rv = a.s[f] - om * a.t[f];
        a.r[f] = rv;
        acc[0] += static_cast<double>(a.rhat[f]) * rv;
        acc[1] += static_cast<double>(rv) * rv;
    }
    const bool mx[2] = {false, false};
    double out[2];
    if (grid_reduce<2>(acc, mx, a.partials, a.ticket, out)) {
        a.scal[S_RHO_OLD] = a.scal[S_RHO];
        a.scal[S_RHO] = out[0];
        a.scal[S_RNORM2] = out[1];
        const int it = a.flags[1] + 1;
        a.flags[1] = it;
        const bool conv = out[1] <= a.eta2 * a.scal[S_BNORM2];
        const bool stall = (a.scal[S_OMEGA] == 0.0) || (out[0] == 0.0);
        if (conv || it >= a.max_it || stall) a.flags[0] = 1;
        if (stall && !conv) a.flags[2] = 1;
    }
}

// line-search trial: max |F| and max centred-gradient ratio of z + lam*dz
// over the DOFs (lam = 0: the current iterate)
__global__ void k_hs_trial(HsDom d, const double* __restrict__ z, const float* __restrict__ dz, double lam,
                           long long n, double* partials, unsigned* ticket, double* out_scal)
{
    const ZView Z{z, lam != 0.0 ? dz : nullptr, lam};
    double acc[2] = {0.0, 0.0};
    for (long long f = blockIdx.x * (long long)blockDim.x + threadIdx.x; f < n;
         f += (long long)gridDim.x * blockDim.x) {
        if (!d.dof[f]) continue;
        const int r = static_cast<int>(f / d.nc), c = static_cast<int>(f - (long long)r * d.nc);
        acc[0] = fmax(acc[0], fabs(hs_residual(d, Z, r, c)));
        if (d.nonlinear) acc[1] = fmax(acc[1], hs_ratio(d, Z, r, c));
    }
    const bool mx[2] = {true, true};
    double out[2];
    if (grid_reduce<2>(acc, mx, partials, ticket, out)) {
        out_scal[0] = out[0];
        out_scal[1] = out[1];
    }
}

__global__ void k_hs_apply(const uint8_t* __restrict__ dof, double* __restrict__ z, const float* __restrict__ dz,
                           double lam, long long n)
{
    for (long long f = blockIdx.x * (long long)blockDim.x + threadIdx.x; f < n;
         f += (long long)gridDim.x * blockDim.x)
        if (dof[f]) z[f] = z[f] + lam * static_cast<double>(dz[f]);
}

// ---------------------------------------------------------------------------
// uplift, restore, statistics
// ---------------------------------------------------------------------------
__global__ void k_uplift(const uint8_t* __restrict__ mask, double* __restrict__ z, double udz, long long n)
{
    for (long long f = blockIdx.x * (long long)blockDim.x + threadIdx.x; f < n;
         f += (long long)gridDim.x * blockDim.x)
        if (mask[f]) z[f] += udz;
}

__global__ void k_restore(double* __restrict__ z, const int* __restrict__ idx, const double* __restrict__ bel,
                          long long m, double bdz)
{
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i < m) z[idx[i]] = bel[i] - bdz;
}

// lsdtt_lem_cpu.cpp:655-722 sums and maxima; |dz| as float for p99.9
// (invalid cells get +inf, which sorts after every valid value)
__global__ void k_stats(const uint8_t* __restrict__ mask, const double* __restrict__ zb, const double* __restrict__ za,
                        long long n, double udz, float* __restrict__ absdz, double* partials, unsigned* ticket,
                        double* out)
{
    double acc[8] = {0.0, 0.0, -1e30, -1e30, 0.0, 0.0, 0.0, 0.0};
    for (long long f = blockIdx.x * (long long)blockDim.x + threadIdx.x; f < n;
         f += (long long)gridDim.x * blockDim.x) {
        if (!mask[f]) {
            absdz[f] = __int_as_float(0x7f800000);
            continue;
        }
        const double b = zb[f], a = za[f];
        acc[0] += b;
        acc[1] += a;
        acc[2] = fmax(acc[2], b);
        acc[3] = fmax(acc[3], a);
        const double l = fabs(a - b);
        acc[4] = fmax(acc[4], l);
        absdz[f] = static_cast<float>(l);
        const double ero = b - (a - udz);
        if (ero > 0.0) {
            acc[5] += ero;
            acc[6] = fmax(acc[6], ero);
        }
        acc[7] += 1.0;
    }
    const bool mx[8] = {false, false, true, true, true, false, true, false};
    double res[8];
    if (grid_reduce<8>(acc, mx, partials, ticket, res))
        for (int i = 0; i < 8; ++i) out[i] = res[i];
}

// exact k-th smallest of the float bit patterns (all >= 0): three radix
// passes of 11, 11 and 10 bits from the top
constexpr int kRadixBins = 2048;

__global__ void k_radix_hist(const float* __restrict__ v, long long n, const unsigned* __restrict__ sel, int shift,
                             unsigned* __restrict__ hist)
{
    __shared__ unsigned h[kRadixBins];
    for (int i = threadIdx.x; i < kRadixBins; i += blockDim.x) h[i] = 0;
    __syncthreads();
    const unsigned prefix = sel[0], pmask = sel[1];
    for (long long f = blockIdx.x * (long long)blockDim.x + threadIdx.x; f < n;
         f += (long long)gridDim.x * blockDim.x) {
        const unsigned u = __float_as_uint(v[f]);
        if ((u & pmask) == prefix) atomicAdd(&h[(u >> shift) & (kRadixBins - 1)], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kRadixBins; i += blockDim.x)
        if (h[i]) atomicAdd(&hist[i], h[i]);
}

// sel = {prefix, prefix mask, k}; picks the bin holding rank k
__global__ void k_radix_pick(unsigned* __restrict__ hist, unsigned* __restrict__ sel, int shift, int bits)
{
    if (threadIdx.x != 0) return;
    unsigned k = sel[2];
    unsigned cum = 0;
    int bin = 0;
    for (; bin < (1 << bits); ++bin) {
        if (cum + hist[bin] > k) break;
        cum += hist[bin];
    }
    sel[0] |= static_cast<unsigned>(bin) << shift;
    sel[1] |= ((1u << bits) - 1u) << shift;
    sel[2] = k - cum;
    for (int i = 0; i < kRadixBins; ++i) hist[i] = 0;
}

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static FlowOptions lem_flow_options(int nr, int nc, double dx, BC bc)
{
    FlowOptions o;
    o.nr = nr;
    o.nc = nc;
    o.dx = dx;
    o.bc = bc;
    o.interior_nodata_is_outlet = true;   // lsdtt_lem_cpu.cpp:331-332
    o.allow_pits_as_outlets = true;
    return o;
}

LemEngine::LemEngine(int nr_, int nc_, double dx_, const uint8_t* mask_host, const LemParams& p_)
    : nr(nr_), nc(nc_), n((long long)nr_ * nc_), dx(dx_), p(p_), flow(lem_flow_options(nr_, nc_, dx_, p_.bc))
{
    flow.timer = &timer;
    mask.reserve(n);
    dof.reserve(n);
    z.reserve(n);
    z_before.reserve(n);
    LSDTT_HIP(hipMemcpy(mask.p, mask_host, n, hipMemcpyHostToDevice));
    // hillslope DOFs: valid, not on an 'o' (Dirichlet) edge
    // (lsdtt_hillslope_operator_cpu.hpp:136-142; no channel set in the
    // coupled regime, lsdtt_lem_cpu.cpp:64-77)
    std::vector<uint8_t> d(static_cast<size_t>(n));
    for (int r = 0; r < nr; ++r)
        for (int c = 0; c < nc; ++c) {
            const long long f = (long long)r * nc + c;
            const bool edge = (r == 0 && p.bc.n == 'o') || (r == nr - 1 && p.bc.s == 'o') ||
                              (c == 0 && p.bc.w == 'o') || (c == nc - 1 && p.bc.e == 'o');
            d[static_cast<size_t>(f)] = (mask_host[f] && !edge) ? 1 : 0;
        }
    LSDTT_HIP(hipMemcpy(dof.p, d.data(), n, hipMemcpyHostToDevice));
    hs_grid = grid_for(n, 2048);
    partials.reserve(static_cast<size_t>(hs_grid) * 8 + 8);
    scal.reserve(S_COUNT + 8);
    tickets.reserve(4);
    flags.reserve(8);
    LSDTT_HIP(hipMemset(tickets.p, 0, sizeof(unsigned) * 4));
    LSDTT_HIP(hipMemset(scal.p, 0, sizeof(double) * (S_COUNT + 8)));
}

void LemEngine::set_elevation(const double* zh)
{
    LSDTT_HIP(hipMemcpy(z.p, zh, sizeof(double) * n, hipMemcpyHostToDevice));
}

void LemEngine::get_elevation(double* zh)
{
    LSDTT_HIP(hipDeviceSynchronize());
    LSDTT_HIP(hipMemcpy(zh, z.p, sizeof(double) * n, hipMemcpyDeviceToHost));
}

void LemEngine::set_erodibility(const double* Kh)
{
    if (!Kh) {
        has_K = false;
        return;
    }
    K_field.reserve(n);
    LSDTT_HIP(hipMemcpy(K_field.p, Kh, sizeof(double) * n, hipMemcpyHostToDevice));
    has_K = true;
}

void LemEngine::set_boundary_flats(const long long* flats, long long m)
{
    n_bflats = m;
    bflats.reserve(m ? m : 1);
    bel.reserve(m ? m : 1);
    std::vector<int> v(static_cast<size_t>(m));
    for (long long i = 0; i < m; ++i) {
        if (flats[i] < 0 || flats[i] >= n) throw std::invalid_argument("boundary flat index out of range");
        v[static_cast<size_t>(i)] = static_cast<int>(flats[i]);
    }
    if (m) LSDTT_HIP(hipMemcpy(bflats.p, v.data(), sizeof(int) * m, hipMemcpyHostToDevice));
}

void LemEngine::stream_power(hipStream_t s)
{
    flow.route<double>(z_before.p, mask.p, nullptr, s);
    flow.upstream_counts_fast(s);
    SpArgs a;
    a.rcv = flow.rcv.p;
    a.dir = flow.dir.p;
    a.cnt = flow.cnt.p;
    a.z_old = z_before.p;
    a.z = z.p;
    a.K = has_K ? K_field.p : nullptr;
    a.K0 = p.K;
    a.m = p.m;
    a.nexp = p.n;
    a.dt = p.dt;
    a.dx = dx;
    a.pixel_area = dx * dx;
    a.threshold = p.threshold;
    a.slope_clamp = p.slope_clamp;
    a.divide = p.divide_area_by_dx;
    a.n_is_one = std::fabs(p.n - 1.0) < 1e-6;
    a.eroded = reinterpret_cast<unsigned long long*>(flags.p + 4);
    if (flow.n_valid == 0) return;
    const int nl = flow.max_depth + 1;
    const int split = std::max(flow.tail_from, 1);
    for (int L = 1; L < split && L < nl; ++L) {
        const int b = flow.level_start_h[L], e = flow.level_start_h[L + 1];
        if (e > b) {
            timer.begin("stream_power", s);
            k_sp_level<<<div_up(e - b, kB), kB, 0, s>>>(flow.order.p, b, e, a);
            timer.end(s, 37.0 * (e - b));
            LSDTT_LAUNCH_CHECK();
        }
    }
    if (split < nl) {
        timer.begin("stream_power_tail", s);
        k_sp_tail<<<1, 1024, 0, s>>>(flow.order.p, flow.level_start.p, split, nl - 1, a);
        timer.end(s, 37.0 * (flow.level_start_h[nl] - flow.level_start_h[split]));
        LSDTT_LAUNCH_CHECK();
    }
}

int LemEngine::bicgstab(hipStream_t s)
{
    BiArgs a;
    a.d = HsDom{nr, nc, p.bc, mask.p, dof.p, dx, p.D, 0.0, p.Sc, p.slope_ratio_cap, 0.0, p.nonlinear, zn_hs.p};
    a.n = n;
    a.coef = coef.p;
    a.dinv = dinv.p;
    a.b = b.p;
    a.x = x.p;
    a.r = r.p;
    a.rhat = rhat.p;
    a.p = pv.p;
    a.phat = phat.p;
    a.v = v.p;
    a.s = sv.p;
    a.shat = shat.p;
    a.t = t.p;
    a.scal = scal.p;
    a.partials = partials.p;
    a.ticket = tickets.p;
    a.flags = flags.p;
    a.eta2 = p.krylov_rel_tol * p.krylov_rel_tol;
    a.max_it = p.krylov_max_iterations;
    const int g = hs_grid;
    k_bi_init<<<g, kB, 0, s>>>(a);
    LSDTT_LAUNCH_CHECK();
    constexpr int kChunk = 4;
    int hf[3] = {0, 0, 0};
    for (int done = 0; !done;) {
        for (int k = 0; k < kChunk; ++k) {
            timer.begin("bicgstab_p", s);
            k_bi_p<<<g, kB, 0, s>>>(a);
            timer.end(s, 24.0 * n);
            timer.begin("bicgstab_matvec", s);
            k_bi_v<<<g, kB, 0, s>>>(a);
            timer.end(s, 48.0 * n);
            timer.begin("bicgstab_s", s);
            k_bi_s<<<g, kB, 0, s>>>(a);
            timer.end(s, 20.0 * n);
            timer.begin("bicgstab_matvec", s);
            k_bi_t<<<g, kB, 0, s>>>(a);
            timer.end(s, 48.0 * n);
            timer.begin("bicgstab_x", s);
            k_bi_x<<<g, kB, 0, s>>>(a);
            timer.end(s, 32.0 * n);
            LSDTT_LAUNCH_CHECK();
        }
        LSDTT_HIP(hipMemcpyAsync(hf, flags.p, sizeof(int) * 3, hipMemcpyDeviceToHost, s));
        LSDTT_HIP(hipStreamSynchronize(s));
        done = hf[0];
    }
    return hf[1];
}

void LemEngine::trial(double lambda, double* fnorm, double* ratio, hipStream_t s)
{
    HsDom d{nr, nc, p.bc, mask.p, dof.p, dx, p.D, 0.0, p.Sc, p.slope_ratio_cap, cur_inv_dt, p.nonlinear, zn_hs.p};
    timer.begin("hillslope_residual", s);
    k_hs_trial<<<hs_grid, kB, 0, s>>>(d, z.p, x.p, lambda, n, partials.p, tickets.p, scal.p + S_TMP0);
    timer.end(s, 22.0 * n);
    LSDTT_LAUNCH_CHECK();
    double h[2];
    LSDTT_HIP(hipMemcpyAsync(h, scal.p + S_TMP0, sizeof(double) * 2, hipMemcpyDeviceToHost, s));
    LSDTT_HIP(hipStreamSynchronize(s));
    *fnorm = h[0];
    *ratio = h[1];
}

// compute_hillslope_newton on the resident z (z^n in zn_hs); returns the
// iteration count, or -1 when it did not converge
int LemEngine::newton_solve(double inv_dt, LemStats* st, hipStream_t s)
{
    cur_inv_dt = inv_dt;
    double fnorm, ratio_now;
    trial(0.0, &fnorm, &ratio_now, s);
    // relative to ||F(z^n)||inf; the hillslope's own uplift term is zero
    // (uplift is applied after the hillslope, lsdtt_lem_cpu.cpp:506)
    const double tol = std::max(p.newton_abs_tol, p.newton_rel_tol * fnorm);
    if (st->substeps == 0 && st->rejected == 0) st->hillslope_initial_residual = fnorm;
    st->hillslope_tolerance = tol;
    const double feasible = std::sqrt(p.slope_ratio_cap);
    HsDom d{nr, nc, p.bc, mask.p, dof.p, dx, p.D, 0.0, p.Sc, p.slope_ratio_cap, inv_dt, p.nonlinear, zn_hs.p};
    int it = 0;
    bool converged = false;
    for (; it < p.newton_max_iterations; ++it) {
        if (fnorm <= tol) {
            converged = true;
            break;
        }
        timer.begin("hillslope_assemble", s);
        k_hs_assemble<<<hs_grid, kB, 0, s>>>(d, z.p, n, coef.p, dinv.p, b.p);
        timer.end(s, 62.0 * n);
        LSDTT_LAUNCH_CHECK();
        st->krylov_iterations += bicgstab(s);
        // line search (newton .cpp:553-633)
        const bool restoring = p.nonlinear && ratio_now > feasible;
        double lambda = 1.0, fnorm_try = fnorm, ratio_try = 0.0;
        bool accepted = false;
        for (int ls = 0; ls < p.newton_max_line_search; ++ls) {
            trial(lambda, &fnorm_try, &ratio_try, s);
            if (p.nonlinear) {
                if (restoring) {
                    if (ratio_try >= ratio_now) {
                        lambda *= 0.5;
                        continue;
                    }
                    accepted = true;
                    break;
                }
                if (ratio_try > feasible) {
                    lambda *= 0.5;
                    continue;
                }
            }
            if (fnorm_try < fnorm) {
                accepted = true;
                break;
            }
            lambda *= 0.5;
        }
        if (!accepted) break;
        timer.begin("hillslope_update", s);
        k_hs_apply<<<hs_grid, kB, 0, s>>>(dof.p, z.p, x.p, lambda, n);
        timer.end(s, 21.0 * n);
        LSDTT_LAUNCH_CHECK();
        fnorm = fnorm_try;
        ratio_now = ratio_try;
        if (fnorm <= tol) {
            converged = true;
            ++it;
            break;
        }
    }
    st->newton_iterations += it;
    st->hillslope_residual = fnorm;
    st->max_slope_ratio = ratio_now;
    return converged ? it : -1;
}

// compute_hillslope_newton_transient (newton .cpp:690-788)
int LemEngine::hillslope(LemStats* st, hipStream_t s)
{
    const size_t nn = static_cast<size_t>(n);
    coef.reserve(9 * nn);
    for (auto* bf : {&dinv, &b, &x, &r, &rhat, &pv, &phat, &v, &sv, &shat, &t}) bf->reserve(nn);
    zn_hs.reserve(nn);
    z_save.reserve(nn);
    const double total = p.dt;
    double dt = total;
    const double dt_min = total * 1e-9, dt_max = total;
    double remaining = total;
    const int g = grid_for(n, 4096);
    k_copy<<<g, kB, 0, s>>>(z.p, zn_hs.p, n);
    LSDTT_LAUNCH_CHECK();
    while (remaining > 0.0) {
        const double step = std::min(dt, remaining);
        k_copy<<<g, kB, 0, s>>>(z.p, z_save.p, n);
        LSDTT_LAUNCH_CHECK();
        const int its = newton_solve(1.0 / step, st, s);
        if (its < 0) {
            k_copy<<<g, kB, 0, s>>>(z_save.p, z.p, n);
            LSDTT_LAUNCH_CHECK();
            ++st->rejected;
            if (st->rejected > 40 || step <= dt_min) {
                st->hillslope_converged = 0;
                return 1;
            }
            dt = std::max(dt_min, step * 0.5);
            continue;
        }
        k_copy<<<g, kB, 0, s>>>(z.p, zn_hs.p, n);
        LSDTT_LAUNCH_CHECK();
        remaining -= step;
        ++st->substeps;
        if (its <= 4) dt = std::min(dt_max, step * 1.7);
        else if (its >= 15) dt = std::max(dt_min, step * 0.5);
    }
    return 0;
}

void LemEngine::stats(LemStats* st, hipStream_t s)
{
    absdz.reserve(n);
    hist.reserve(kRadixBins + 8);
    const double udz = p.enable_uplift ? p.U * p.dt : 0.0;
    timer.begin("step_stats", s);
    k_stats<<<hs_grid, kB, 0, s>>>(mask.p, z_before.p, z.p, n, udz, absdz.p, partials.p, tickets.p, scal.p + S_SUM_B);
    timer.end(s, 21.0 * n);
    LSDTT_LAUNCH_CHECK();
    double h[8];
    LSDTT_HIP(hipMemcpyAsync(h, scal.p + S_SUM_B, sizeof(h), hipMemcpyDeviceToHost, s));
    LSDTT_HIP(hipStreamSynchronize(s));
    const long long nv = static_cast<long long>(h[7]);
    const double inv_dt = 1.0 / p.dt, inv_n = nv > 0 ? 1.0 / static_cast<double>(nv) : 0.0;
    st->n_valid = nv;
    st->delta_max_elevation = std::fabs(h[3] - h[2]);
    st->delta_mean_elevation = std::fabs((h[1] - h[0]) * inv_n);
    st->max_local_elevation_change = h[4];
    st->mean_erosion_rate = h[5] * inv_n * inv_dt;
    st->max_erosion_rate = h[6] * inv_dt;
    st->p999_local_elevation_change = 0.0;
    if (nv > 0) {
        const unsigned k = static_cast<unsigned>(std::floor(0.999 * static_cast<double>(nv - 1)));
        unsigned sel[3] = {0u, 0u, k};
        unsigned* dsel = hist.p + kRadixBins;
        LSDTT_HIP(hipMemsetAsync(hist.p, 0, sizeof(unsigned) * kRadixBins, s));
        LSDTT_HIP(hipMemcpyAsync(dsel, sel, sizeof(sel), hipMemcpyHostToDevice, s));
        const int shifts[3] = {21, 10, 0}, bits[3] = {11, 11, 10};
        for (int pass = 0; pass < 3; ++pass) {
            k_radix_hist<<<grid_for(n, 2048), kB, 0, s>>>(absdz.p, n, dsel, shifts[pass], hist.p);
            k_radix_pick<<<1, 64, 0, s>>>(hist.p, dsel, shifts[pass], bits[pass]);
            LSDTT_LAUNCH_CHECK();
        }
        LSDTT_HIP(hipMemcpyAsync(sel, dsel, sizeof(sel), hipMemcpyDeviceToHost, s));
        LSDTT_HIP(hipStreamSynchronize(s));
        float pv;
        std::memcpy(&pv, &sel[0], sizeof(float));
        st->p999_local_elevation_change = pv;
    }
}

int LemEngine::step(LemStats* st, hipStream_t s)
{
    *st = LemStats{};
    const int g = grid_for(n, 4096);
    k_copy<<<g, kB, 0, s>>>(z.p, z_before.p, n);
    LSDTT_LAUNCH_CHECK();
    if (n_bflats) {
        k_gather<<<div_up(n_bflats, kB), kB, 0, s>>>(z.p, bflats.p, n_bflats, bel.p);
        LSDTT_LAUNCH_CHECK();
    }
    LSDTT_HIP(hipMemsetAsync(flags.p + 4, 0, sizeof(unsigned long long), s));
    if (p.enable_fluvial) stream_power(s);
    int rc = 0;
    if (p.enable_hillslope) rc = hillslope(st, s);
    if (p.enable_uplift) {
        k_uplift<<<g, kB, 0, s>>>(mask.p, z.p, p.U * p.dt, n);
        LSDTT_LAUNCH_CHECK();
    }
    if (n_bflats) {
        k_restore<<<div_up(n_bflats, kB), kB, 0, s>>>(z.p, bflats.p, bel.p, n_bflats, p.baselevel_fall * p.dt);
        LSDTT_LAUNCH_CHECK();
    }
    stats(st, s);
    unsigned long long er = 0;
    LSDTT_HIP(hipMemcpyAsync(&er, flags.p + 4, sizeof(er), hipMemcpyDeviceToHost, s));
    LSDTT_HIP(hipStreamSynchronize(s));
    st->cells_eroded = static_cast<long long>(er);
    return rc;
}

}  // namespace lsdtt_gpu
