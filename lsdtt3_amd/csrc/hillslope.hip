// hillslope.hip -- hillslope operator kernels (see hillslope.hpp).
//
// Reference (src/cpu/): lsdtt_hillslope_operator_cpu.hpp:146-159 (neighbour
// resolution), :290-321 (face alphas); lsdtt_hillslope_newton_cpu.cpp:
// 137-186 (dalpha_dz), :283-308 (residual), :415-471 (Jacobian entries),
// :313-328 (max centred-gradient ratio).
//
// Every kernel has two paths with identical arithmetic: a fast path for a
// DOF whose 3x3 block is interior and fully valid (all neighbour
// resolutions are plain offsets, values come from nine registers), and the
// general path that resolves every neighbour by the boundary rules.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "hillslope.hpp"
#include "ktimer.hpp"
#include "reduce.hpp"

namespace lsdtt_gpu {

namespace {

constexpr int kB = 256;

// ---------------------------------------------------------------------------
// general path (any cell): the reference's neighbour resolution
// ---------------------------------------------------------------------------
__device__ __forceinline__ int nbr(const HsDom& d, int r, int c, int dr, int dc)
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

// elevation z, or z + lam * dz for a line-search trial
struct ZView {
    const double* z;
    const float* dz;
    double lam;
    __device__ __forceinline__ double operator()(int i) const
    {
        return dz ? z[i] + lam * static_cast<double>(dz[i]) : z[i];
    }
};

__device__ __forceinline__ double sx_g(const HsDom& d, const ZView& Z, int r, int c)
{
    return (Z(nbr(d, r, c, 0, 1)) - Z(nbr(d, r, c, 0, -1))) / (2.0 * d.h);
}
__device__ __forceinline__ double sy_g(const HsDom& d, const ZView& Z, int r, int c)
{
    return (Z(nbr(d, r, c, 1, 0)) - Z(nbr(d, r, c, -1, 0))) / (2.0 * d.h);
}

constexpr int nesw_r[4] = {-1, 0, 1, 0};
constexpr int nesw_c[4] = {0, 1, 0, -1};

// c9[slot] += w with a runtime slot, as predicated register updates (a
// runtime index into a register array would move it to scratch memory)
__device__ __forceinline__ void add_slot(double c9[9], int slot, double w)
{
#pragma unroll
    for (int k = 0; k < 9; ++k) c9[k] += (k == slot) ? w : 0.0;
}

// slot of a resolved target in the 3x3 block of (r, c)
__device__ __forceinline__ int slot_of(const HsDom& d, int r, int c, int t)
{
    const int tr = t / d.nc, tc = t - tr * d.nc;
    int dr = tr - r, dc = tc - c;
    if (dr > 1) dr -= d.nr;
    else if (dr < -1) dr += d.nr;
    if (dc > 1) dc -= d.nc;
    else if (dc < -1) dc += d.nc;
    return (dr + 1) * 3 + (dc + 1);
}

// cell of slot k of (r, c) for a product: periodic axes wrap, others clamp
// (a clamped slot always carries a zero coefficient)
__device__ __forceinline__ int slot_cell(const HsDom& d, int r, int c, int k)
{
    int rr = r + k / 3 - 1, cc = c + k % 3 - 1;
    if (rr < 0) rr = (d.bc.n == 'p') ? d.nr - 1 : 0;
    else if (rr >= d.nr) rr = (d.bc.s == 'p') ? 0 : d.nr - 1;
    if (cc < 0) cc = (d.bc.w == 'p') ? d.nc - 1 : 0;
    else if (cc >= d.nc) cc = (d.bc.e == 'p') ? 0 : d.nc - 1;
    return rr * d.nc + cc;
}

// F (FP64) and the 9 Jacobian coefficients (FP64, into c9 when non-null) of
// DOF (r, c) by the general rules.  Only cells on the grid edge or next to
// a masked cell take it, in companion kernels of their own, so the hot
// kernels do not carry its registers.
__device__ double cell_general(const HsDom& d, const ZView& Z, int r, int c, double* c9)
{
    const double h = d.h, DA = d.D / (h * h), Sc2 = d.Sc * d.Sc;
    const int self = r * d.nc + c;
    const double zc = Z(self);
    const double sxc = d.nonlinear ? sx_g(d, Z, r, c) : 0.0;
    const double syc = d.nonlinear ? sy_g(d, Z, r, c) : 0.0;
    double a[4], sn[4], st[4];
    int nb[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        nb[k] = nbr(d, r, c, nesw_r[k], nesw_c[k]);
        sn[k] = (Z(nb[k]) - zc) / h;
        st[k] = 0.0;
        if (!d.nonlinear) {
            a[k] = 1.0;
            continue;
        }
        const int rr = nb[k] / d.nc, cc = nb[k] - rr * d.nc;
        st[k] = (k == 0 || k == 2) ? 0.5 * (sxc + sx_g(d, Z, rr, cc)) : 0.5 * (syc + sy_g(d, Z, rr, cc));
        double s2 = (sn[k] * sn[k] + st[k] * st[k]) / Sc2;
        if (s2 > d.cap) s2 = d.cap;
        a[k] = 1.0 / (1.0 - s2);
    }
    double flux = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) flux += a[k] * sn[k] * h;
    double F = -d.U - DA * flux;
    if (d.inv_dt > 0.0) F += d.inv_dt * (zc - d.zn[self]);
    if (!c9) return F;
#pragma unroll
    for (int k = 0; k < 9; ++k) c9[k] = 0.0;
    auto add = [&](int t, double w) {
        if (w == 0.0 || !d.dof[t]) return;
        add_slot(c9, slot_of(d, r, c, t), w);
    };
    if (d.inv_dt > 0.0) add(self, d.inv_dt);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        add(nb[q], -DA * a[q]);
        add(self, +DA * a[q]);
        if (!d.nonlinear) continue;
        const double rraw = (sn[q] * sn[q] + st[q] * st[q]) / Sc2;
        if (rraw >= d.cap) continue;
        const int rr = nb[q] / d.nc, cc = nb[q] - rr * d.nc;
        const double pref = a[q] * a[q] * 2.0 / Sc2;
        const double dzf = sn[q] * h;
        const double wt = pref * st[q] * (1.0 / (4.0 * h));
        add(self, -DA * dzf * (pref * sn[q] * (-1.0 / h)));
        add(nb[q], -DA * dzf * (pref * sn[q] * (+1.0 / h)));
        if (q == 0 || q == 2) {
            add(nbr(d, r, c, 0, 1), -DA * dzf * (+wt));
            add(nbr(d, r, c, 0, -1), -DA * dzf * (-wt));
            add(nbr(d, rr, cc, 0, 1), -DA * dzf * (+wt));
            add(nbr(d, rr, cc, 0, -1), -DA * dzf * (-wt));
        } else {
            add(nbr(d, r, c, 1, 0), -DA * dzf * (+wt));
            add(nbr(d, r, c, -1, 0), -DA * dzf * (-wt));
            add(nbr(d, rr, cc, 1, 0), -DA * dzf * (+wt));
            add(nbr(d, rr, cc, -1, 0), -DA * dzf * (-wt));
        }
    }
    return F;
}

// squared centred slope; the trial's largest slope ratio is
// sqrt(max) / Sc, formed once in its finisher
__device__ __forceinline__ double slope2_general(const HsDom& d, const ZView& Z, int r, int c)
{
    const double sx = sx_g(d, Z, r, c), sy = sy_g(d, Z, r, c);
    return sx * sx + sy * sy;
}

// ---------------------------------------------------------------------------
// fast path: the 3x3 block z[0..8] (row-major, centre 4) is interior and
// valid, so every resolution is a plain offset
// ---------------------------------------------------------------------------
// NESW neighbour slots, and per face the tangential slots (own +/-, then
// the neighbour's +/-)
// (constexpr, so that unrolled loops index registers, never scratch)
constexpr int f_nb[4] = {1, 5, 7, 3};
constexpr int f_tp[4] = {5, 7, 5, 7};
constexpr int f_tm[4] = {3, 1, 3, 1};
constexpr int f_np[4] = {2, 8, 8, 6};
constexpr int f_nm[4] = {0, 2, 6, 0};

template <class T>
__device__ __forceinline__ void faces_fast(const HsDom& d, const double z[9], T sn[4], T st[4], T a[4], bool capped[4])
{
    const T inv_h = static_cast<T>(1.0 / d.h), inv_2h = static_cast<T>(0.5 / d.h);
    const T inv_Sc2 = static_cast<T>(1.0 / (d.Sc * d.Sc)), cap = static_cast<T>(d.cap);
    // differences in FP64 (elevations are ~1e3 m, slopes need ~1e-6), then T
    const T sxc = static_cast<T>(z[5] - z[3]) * inv_2h;
    const T syc = static_cast<T>(z[7] - z[1]) * inv_2h;
    const T sxn[4] = {static_cast<T>(z[2] - z[0]) * inv_2h, static_cast<T>(z[8] - z[2]) * inv_2h,
                      static_cast<T>(z[8] - z[6]) * inv_2h, static_cast<T>(z[6] - z[0]) * inv_2h};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        sn[k] = static_cast<T>(z[f_nb[k]] - z[4]) * inv_h;
        if (!d.nonlinear) {
            st[k] = T(0);
            a[k] = T(1);
            capped[k] = true;
            continue;
        }
        st[k] = T(0.5) * (((k & 1) ? syc : sxc) + sxn[k]);
        const T raw = (sn[k] * sn[k] + st[k] * st[k]) * inv_Sc2;
        capped[k] = raw >= cap;
        if constexpr (sizeof(T) == 4)
            a[k] = __builtin_amdgcn_rcpf(T(1) - (raw > cap ? cap : raw));  // ~1 ulp: Krylov operator only
        else
            a[k] = T(1) / (T(1) - (raw > cap ? cap : raw));
    }
}

// FP64 residual on the fast path (cell_general's expressions, with the
// constant divisions by h, 2h and Sc^2 as products by their reciprocals),
// and the Jacobian diagonal J_PP in FP64 (the FP32 row's centre entry)
__device__ __forceinline__ double eval_fast(const HsDom& d, const double z[9], double zn_c, double* diag)
{
    const double DA = d.D * d.inv_h * d.inv_h;
    const double sxc = (z[5] - z[3]) * d.inv_2h;
    const double syc = (z[7] - z[1]) * d.inv_2h;
    const double sxn[4] = {(z[2] - z[0]) * d.inv_2h, (z[8] - z[2]) * d.inv_2h, (z[8] - z[6]) * d.inv_2h,
                           (z[6] - z[0]) * d.inv_2h};
    double flux = 0.0, jd = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const double dz = z[f_nb[k]] - z[4];
        if (!d.nonlinear) {
            flux += dz;
            jd += 1.0;
            continue;
        }
        const double sn = dz * d.inv_h;
        const double st = 0.5 * (((k & 1) ? syc : sxc) + sxn[k]);
        const double raw = (sn * sn + st * st) * d.inv_Sc2;
        const double a = 1.0 / (1.0 - (raw > d.cap ? d.cap : raw));
        flux += a * dz;
        jd += a;
        if (raw < d.cap) jd += 2.0 * a * a * sn * sn * d.inv_Sc2;
    }
    *diag = d.inv_dt + DA * jd;
    double F = -d.U - DA * flux;
    if (d.inv_dt > 0.0) F += d.inv_dt * (z[4] - zn_c);
    return F;
}

// Jacobian row on the fast path in precision T
template <class T>
__device__ __forceinline__ void J_fast(const HsDom& d, const double z[9], T c9[9])
{
    T sn[4], st[4], a[4];
    bool capped[4];
    faces_fast<T>(d, z, sn, st, a, capped);
    const T h = static_cast<T>(d.h), DA = static_cast<T>(d.D / (d.h * d.h));
    const T two_inv_Sc2 = static_cast<T>(2.0 / (d.Sc * d.Sc)), inv_h = static_cast<T>(1.0 / d.h);
    const T inv_4h = static_cast<T>(0.25 / d.h);
#pragma unroll
    for (int k = 0; k < 9; ++k) c9[k] = T(0);
    c9[4] = static_cast<T>(d.inv_dt);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        c9[f_nb[q]] -= DA * a[q];
        c9[4] += DA * a[q];
        if (capped[q]) continue;
        const T pref = a[q] * a[q] * two_inv_Sc2;
        const T g = -DA * (sn[q] * h);
        const T wt = pref * st[q] * inv_4h;
        c9[4] -= g * (pref * sn[q] * inv_h);
        c9[f_nb[q]] += g * (pref * sn[q] * inv_h);
        c9[f_tp[q]] += g * wt;
        c9[f_tm[q]] -= g * wt;
        c9[f_np[q]] += g * wt;
        c9[f_nm[q]] -= g * wt;
    }
}

// ---------------------------------------------------------------------------
// LDS tiles for the stencil kernels: a workgroup walks 64 x 16 cell tiles
// (persistent grid); per tile it stages z (or z + lam dz) and, for the
// Jacobian products, the input vector over the tile plus a one-cell halo
// with coalesced row loads, then each wavefront evaluates one tile row per
// step from LDS.  Halo cells outside the grid are never read: a fast cell's
// 3 x 3 block is interior.
// ---------------------------------------------------------------------------
constexpr int kTW = 64, kTH = 16, kSW = kTW + 2, kSH = kTH + 2;

// Branch-free: every thread issues all of its loads (indices clamped into
// the grid; a clamped value lands in a halo slot no fast cell reads) before
// the first LDS store, so kStage loads are in flight at once (a load under
// a branch would be followed by a full wait for it).
template <bool HAS_DZ, bool HAS_X>
__device__ __forceinline__ void stage_tile(const double* __restrict__ z, const float* __restrict__ dz, double lam,
                                           const float* __restrict__ x, int nr, int nc, int r0, int c0,
                                           double* __restrict__ zs, float* __restrict__ xs)
{
    constexpr int kStage = (kSH * kSW + kB - 1) / kB;
    double zv[kStage];
    float dv[kStage], xv[kStage];
#pragma unroll
    for (int j = 0; j < kStage; ++j) {
        const int i = min(static_cast<int>(threadIdx.x) + j * kB, kSH * kSW - 1);
        const int rr = min(max(r0 - 1 + i / kSW, 0), nr - 1), cc = min(max(c0 - 1 + i % kSW, 0), nc - 1);
        const long long g = (long long)rr * nc + cc;
        zv[j] = z[g];
        if (HAS_DZ) dv[j] = dz[g];
        if (HAS_X) xv[j] = x[g];
    }
#pragma unroll
    for (int j = 0; j < kStage; ++j) {
        const int i = threadIdx.x + j * kB;
        if (i >= kSH * kSW) continue;
        zs[i] = HAS_DZ ? zv[j] + lam * static_cast<double>(dv[j]) : zv[j];  // ZView's expression
        if (HAS_X) xs[i] = xv[j];
    }
}

__device__ __forceinline__ void tile_z9(const double* __restrict__ zs, int rl, int cl, double zz[9])
{
#pragma unroll
    for (int q = 0; q < 9; ++q) zz[q] = zs[(rl + q / 3) * kSW + cl + q % 3];
}

// y = J x (FP32) on a fast row, from the staged tile (jx_fast's arithmetic)
__device__ __forceinline__ float jx_tile(const HsDom& d, const double* __restrict__ zs, const float* __restrict__ xs,
                                         int rl, int cl)
{
    double zz[9];
    tile_z9(zs, rl, cl, zz);
    float c9[9];
    J_fast<float>(d, zz, c9);
    float acc = 0.0f;
#pragma unroll
    for (int k = 0; k < 9; ++k) acc += c9[k] * xs[(rl + k / 3) * kSW + cl + k % 3];
    return acc;
}

// ---------------------------------------------------------------------------
// kernels.  Each hot kernel covers the fast cells (and writes zeros on
// non-DOF cells); a companion kernel over the list of general DOF cells
// runs first and leaves its block partials after the hot kernel's, which
// the hot kernel's one-block finisher folds in after the hot kernel's own.
// ---------------------------------------------------------------------------
constexpr int kGenBlocks = 64;
// partial-sum regions of one BiCGSTAB iteration (in doubles, times the grid
// for the 2-value regions): v dots, t dots, x dots; a region is folded by
// the next kernel while that kernel writes its own
constexpr int kPartV = 0, kPartT = 2, kPartX = 4;

__host__ __device__ inline int gen_grid(int n_gen)
{
    return n_gen > 0 ? ((n_gen + kB - 1) / kB < kGenBlocks ? (n_gen + kB - 1) / kB : kGenBlocks) : 0;
}

__global__ void __launch_bounds__(256) k_fast_flags(HsDom d, uint8_t* __restrict__ fast)
{
    const long long n = (long long)d.nr * d.nc;
    for (long long f = blockIdx.x * (long long)blockDim.x + threadIdx.x; f < n;
         f += (long long)gridDim.x * blockDim.x) {
        const int r = static_cast<int>(f / d.nc), c = static_cast<int>(f - (long long)r * d.nc);
        uint8_t ok = d.dof[f] && r > 0 && r < d.nr - 1 && c > 0 && c < d.nc - 1;
        if (ok)
            for (int k = 0; k < 9; ++k) ok &= d.mask[f + (k / 3 - 1) * d.nc + (k % 3 - 1)] ? 1 : 0;
        fast[f] = ok;
    }
}

// y = J x on general row i (x is zero on non-DOF cells), from the stored row
__device__ __forceinline__ float jx_gen(const HsDom& d, const float* __restrict__ x, int i, int f)
{
    const int r = f / d.nc, c = f - r * d.nc;
    float acc = 0.0f;
#pragma unroll
    for (int k = 0; k < 9; ++k) acc += d.c9g[(size_t)k * d.n_gen + i] * x[slot_cell(d, r, c, k)];
    return acc;
}

// The vector updates stream four cells per thread per load (float4: the
// solver vectors are separate allocations, 16-byte aligned), with a scalar
// tail when n is not a multiple of 4.
__device__ __forceinline__ float4 ld4(const float* __restrict__ p, long long i)
{
    return reinterpret_cast<const float4*>(p)[i];
}
__device__ __forceinline__ void st4(float* __restrict__ p, long long i, float4 v)
{
    reinterpret_cast<float4*>(p)[i] = v;
}

// One BiCGSTAB iteration (right Jacobi preconditioning M^-1 = dinv) in three
// kernels.  The vectors phat = M^-1 p, s = r - alpha v and shat = M^-1 s are
// never stored: each kernel forms them where it needs them, around its
// staged tiles too, with the same float expressions as stored vectors would
// hold, so the iterates are the same as with separate update kernels.
//   k_bi_pv: p = r + beta (p_old - omega v_old) (first: p = b); v = J phat;
//            partial sums of rhat . v
//   k_bi_st: alpha = rho / (rhat . v); t = J shat; partial sums of t . s, t . t
//   k_bi_x:  omega = (t . s) / (t . t); x += alpha phat + omega shat;
//            r = s - omega t; partial sums of rhat . r, r . r

// the staged input of a Jacobian product: one value per staged cell, from
// the loads V::load issued before any is used
struct PhatView {
    const float *b, *r, *p, *v, *dinv;
    float bt, om;
    bool first;
    struct L {
        float di, a0, a1, a2;
    };
    __device__ __forceinline__ L load(long long g) const
    {
        L l;
        l.di = dinv[g];
        if (first) {
            l.a0 = b[g];
            l.a1 = l.a2 = 0.0f;
        } else {
            l.a0 = r[g];
            l.a1 = p[g];
            l.a2 = v[g];
        }
        return l;
    }
    __device__ __forceinline__ float pval(const L& l) const { return first ? l.a0 : l.a0 + bt * (l.a1 - om * l.a2); }
    __device__ __forceinline__ float value(const L& l) const { return l.di * pval(l); }
    __device__ __forceinline__ float raw(const L& l) const { return pval(l); }
};

struct ShatView {
    const float *r, *v, *dinv;
    float al;
    struct L {
        float di, a0, a1;
    };
    __device__ __forceinline__ L load(long long g) const
    {
        L l;
        l.di = dinv[g];
        l.a0 = r[g];
        l.a1 = v[g];
        return l;
    }
    __device__ __forceinline__ float sval(const L& l) const { return l.a0 - al * l.a1; }
    __device__ __forceinline__ float value(const L& l) const { return l.di * sval(l); }
    __device__ __forceinline__ float raw(const L& l) const { return sval(l); }
};

// stage z and the view's values over the tile plus halo (stage_tile's
// clamped, branch-free loads); rs receives the view's vector before the
// preconditioner (p or s), which the tile's own cells read back from LDS
// rather than loading their inputs a second time
template <class V>
__device__ __forceinline__ void stage_tile_view(const double* __restrict__ z, const V& view, int nr, int nc, int r0,
                                                int c0, double* __restrict__ zs, float* __restrict__ xs,
                                                float* __restrict__ rs)
{
    constexpr int kStage = (kSH * kSW + kB - 1) / kB;
    double zv[kStage];
    typename V::L lv[kStage];
#pragma unroll
    for (int j = 0; j < kStage; ++j) {
        const int i = min(static_cast<int>(threadIdx.x) + j * kB, kSH * kSW - 1);
        const int rr = min(max(r0 - 1 + i / kSW, 0), nr - 1), cc = min(max(c0 - 1 + i % kSW, 0), nc - 1);
        const long long g = (long long)rr * nc + cc;
        zv[j] = z[g];
        lv[j] = view.load(g);
    }
#pragma unroll
    for (int j = 0; j < kStage; ++j) {
        const int i = threadIdx.x + j * kB;
        if (i >= kSH * kSW) continue;
        zs[i] = zv[j];
        xs[i] = view.value(lv[j]);
        rs[i] = view.raw(lv[j]);
    }
}

// y = J x on general row i, x given by a view (zero on non-DOF cells)
template <class V>
__device__ __forceinline__ float jx_gen_view(const HsDom& d, const V& view, int i, int f)
{
    const int r = f / d.nc, c = f - r * d.nc;
    float acc = 0.0f;
#pragma unroll
    for (int k = 0; k < 9; ++k) acc += d.c9g[(size_t)k * d.n_gen + i] * view.value(view.load(slot_cell(d, r, c, k)));
    return acc;
}

__device__ __forceinline__ PhatView phat_view(const BiArgs& a)
{
    PhatView pv;
    pv.b = a.b;
    pv.r = a.r;
    pv.p = a.p_old;
    pv.v = a.v_old;
    pv.dinv = a.dinv;
    pv.first = a.first != 0;
    if (pv.first) {
        pv.bt = pv.om = 0.0f;
    } else {
        const double beta = (a.scal[HS_RHO] / a.scal[HS_RHO_OLD]) * (a.scal[HS_ALPHA] / a.scal[HS_OMEGA]);
        pv.bt = static_cast<float>(beta);
        pv.om = static_cast<float>(a.scal[HS_OMEGA]);
    }
    return pv;
}

// p = r + beta (p_old - omega v_old); v = J phat; alpha's denominator
__global__ void __launch_bounds__(256) k_bi_pv(BiArgs a)
{
    if (a.flags[0]) return;
    __shared__ double zs[kSH * kSW];
    __shared__ float xs[kSH * kSW], ps[kSH * kSW];
    const int nr = a.d.nr, nc = a.d.nc, tx_n = (nc + kTW - 1) / kTW, nt = tx_n * ((nr + kTH - 1) / kTH);
    const int cl = threadIdx.x & 63;
    const uint8_t* __restrict__ fast = a.d.fast;
    const uint8_t* __restrict__ dof = a.d.dof;
    const float* __restrict__ rhat = a.rhat;
    float* __restrict__ vout = a.v;
    float* __restrict__ pout = a.p;
    const PhatView view = phat_view(a);
    double acc[1] = {0.0};
    // tiles in rounds of gridDim.x, each round's tiles dealt XCD by XCD in
    // contiguous bands (shared halos stay in one L2)
    for (int base = 0; base < nt; base += gridDim.x) {
        const int tile = base + xcd_tile(blockIdx.x, gridDim.x);
        if (tile >= nt) continue;
        const int r0 = (tile / tx_n) * kTH, c0 = (tile % tx_n) * kTW;
        __syncthreads();
        stage_tile_view(a.z, view, nr, nc, r0, c0, zs, xs, ps);
        // per-cell inputs of this thread's four rows, loaded before the
        // barrier (they do not depend on the tile)
        long long f[kTH / 4];
        bool in[kTH / 4];
        uint8_t fl[kTH / 4], dl[kTH / 4];
        float rh[kTH / 4];
#pragma unroll
        for (int k = 0; k < kTH / 4; ++k) {
            const int rr = r0 + (threadIdx.x >> 6) + 4 * k, cc = c0 + cl;
            in[k] = rr < nr && cc < nc;
            f[k] = (long long)min(rr, nr - 1) * nc + min(cc, nc - 1);
            fl[k] = fast[f[k]];
            dl[k] = dof[f[k]];
            rh[k] = rhat[f[k]];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kTH / 4; ++k) {
            if (!in[k]) continue;
            pout[f[k]] = ps[((threadIdx.x >> 6) + 4 * k + 1) * kSW + cl + 1];
            if (fl[k]) {
                const float vv = jx_tile(a.d, zs, xs, (threadIdx.x >> 6) + 4 * k, cl);
                vout[f[k]] = vv;
                acc[0] += static_cast<double>(rh[k]) * vv;
            } else if (!dl[k]) {
                vout[f[k]] = 0.0f;
            }
        }
    }
    // the general rows (grid edge, next to a masked cell): stored Jacobian rows
    for (int gi = blockIdx.x * blockDim.x + threadIdx.x; gi < a.d.n_gen; gi += gridDim.x * blockDim.x) {
        const int f = a.d.gen[gi];
        const float y = jx_gen_view(a.d, view, gi, f);
        vout[f] = y;
        acc[0] += static_cast<double>(rhat[f]) * y;
    }
    const bool mx[1] = {false};
    block_partial<1>(acc, mx, a.partials + kPartV, blockIdx.x);
}

// alpha = rho / (rhat . v); t = J shat (s = r - alpha v); omega's sums
__global__ void __launch_bounds__(256) k_bi_st(BiArgs a)
{
    if (a.flags[0]) return;
    // alpha from k_bi_pv's partials: every block folds them in the same
    // order (so all agree), block 0 keeps it for k_bi_x
    double rv1[1];
    const bool mx1[1] = {false};
    fold_partials<1>(a.partials + kPartV, gridDim.x, mx1, rv1);
    __shared__ double alpha_s;
    __shared__ int brk;
    if (threadIdx.x == 0) {
        brk = rv1[0] == 0.0;
        alpha_s = brk ? 0.0 : a.scal[HS_RHO] / rv1[0];
        if (blockIdx.x == 0) {
            if (brk) {
                a.flags[0] = 1;
                a.flags[2] = 1;
            } else {
                a.scal[HS_ALPHA] = alpha_s;
            }
        }
    }
    __syncthreads();
    if (brk) return;
    __shared__ double zs[kSH * kSW];
    __shared__ float xs[kSH * kSW], ss[kSH * kSW];
    const int nr = a.d.nr, nc = a.d.nc, tx_n = (nc + kTW - 1) / kTW, nt = tx_n * ((nr + kTH - 1) / kTH);
    const int cl = threadIdx.x & 63;
    const uint8_t* __restrict__ fast = a.d.fast;
    const uint8_t* __restrict__ dof = a.d.dof;
    float* __restrict__ tout = a.t;
    ShatView view;
    view.r = a.first ? a.b : a.r;
    view.v = a.v;
    view.dinv = a.dinv;
    view.al = static_cast<float>(alpha_s);
    double acc[2] = {0.0, 0.0};
    // tiles in rounds of gridDim.x, each round's tiles dealt XCD by XCD in
    // contiguous bands (shared halos stay in one L2)
    for (int base = 0; base < nt; base += gridDim.x) {
        const int tile = base + xcd_tile(blockIdx.x, gridDim.x);
        if (tile >= nt) continue;
        const int r0 = (tile / tx_n) * kTH, c0 = (tile % tx_n) * kTW;
        __syncthreads();
        stage_tile_view(a.z, view, nr, nc, r0, c0, zs, xs, ss);
        long long f[kTH / 4];
        bool in[kTH / 4];
        uint8_t fl[kTH / 4], dl[kTH / 4];
#pragma unroll
        for (int k = 0; k < kTH / 4; ++k) {
            const int rr = r0 + (threadIdx.x >> 6) + 4 * k, cc = c0 + cl;
            in[k] = rr < nr && cc < nc;
            f[k] = (long long)min(rr, nr - 1) * nc + min(cc, nc - 1);
            fl[k] = fast[f[k]];
            dl[k] = dof[f[k]];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kTH / 4; ++k) {
            if (!in[k]) continue;
            if (fl[k]) {
                const float tv = jx_tile(a.d, zs, xs, (threadIdx.x >> 6) + 4 * k, cl);
                tout[f[k]] = tv;
                acc[0] += static_cast<double>(tv) * ss[((threadIdx.x >> 6) + 4 * k + 1) * kSW + cl + 1];
                acc[1] += static_cast<double>(tv) * tv;
            } else if (!dl[k]) {
                tout[f[k]] = 0.0f;
            }
        }
    }
    for (int gi = blockIdx.x * blockDim.x + threadIdx.x; gi < a.d.n_gen; gi += gridDim.x * blockDim.x) {
        const int f = a.d.gen[gi];
        const float tv = jx_gen_view(a.d, view, gi, f);
        tout[f] = tv;
        acc[0] += static_cast<double>(tv) * view.sval(view.load(f));
        acc[1] += static_cast<double>(tv) * tv;
    }
    const bool mx[2] = {false, false};
    block_partial<2>(acc, mx, a.partials + kPartT * gridDim.x, blockIdx.x);
}

// x += alpha phat + omega shat; r = s - omega t; convergence on ||r||
__global__ void __launch_bounds__(256) k_bi_x(BiArgs a)
{
    if (a.flags[0]) return;
    // omega = (t . s) / (t . t) from k_bi_st's partials, folded by every block
    double tt[2];
    const bool mx2[2] = {false, false};
    fold_partials<2>(a.partials + kPartT * gridDim.x, gridDim.x, mx2, tt);
    __shared__ double omega_s;
    if (threadIdx.x == 0) {
        omega_s = tt[1] > 0.0 ? tt[0] / tt[1] : 0.0;
        if (blockIdx.x == 0) a.scal[HS_OMEGA] = omega_s;
    }
    __syncthreads();
    const float al = static_cast<float>(a.scal[HS_ALPHA]), om = static_cast<float>(omega_s);
    const float* __restrict__ pv = a.p;
    const float* __restrict__ dinv = a.dinv;
    const float* __restrict__ rin = a.first ? a.b : a.r;
    const float* __restrict__ vv = a.v;
    const float* __restrict__ tv = a.t;
    const float* __restrict__ rhat = a.rhat;
    float* __restrict__ x = a.x;
    float* __restrict__ r = a.r;
    double acc[2] = {0.0, 0.0};
    const long long n4 = a.n >> 2, stride = (long long)gridDim.x * blockDim.x;
    const long long t0 = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    for (long long i = t0; i < n4; i += stride) {
        const float4 p4 = ld4(pv, i), di = ld4(dinv, i), r4 = ld4(rin, i), v4 = ld4(vv, i), t4 = ld4(tv, i),
                     rh = ld4(rhat, i);
        const float4 xo = a.first ? make_float4(0.f, 0.f, 0.f, 0.f) : ld4(x, i);
        const float4 s4 = make_float4(r4.x - al * v4.x, r4.y - al * v4.y, r4.z - al * v4.z, r4.w - al * v4.w);
        const float4 ph = make_float4(di.x * p4.x, di.y * p4.y, di.z * p4.z, di.w * p4.w);
        const float4 sh = make_float4(di.x * s4.x, di.y * s4.y, di.z * s4.z, di.w * s4.w);
        const float4 dx = make_float4(al * ph.x + om * sh.x, al * ph.y + om * sh.y, al * ph.z + om * sh.z,
                                      al * ph.w + om * sh.w);
        st4(x, i, a.first ? dx : make_float4(xo.x + dx.x, xo.y + dx.y, xo.z + dx.z, xo.w + dx.w));
        const float4 rv = make_float4(s4.x - om * t4.x, s4.y - om * t4.y, s4.z - om * t4.z, s4.w - om * t4.w);
        st4(r, i, rv);
        acc[0] += static_cast<double>(rh.x) * rv.x;
        acc[0] += static_cast<double>(rh.y) * rv.y;
        acc[0] += static_cast<double>(rh.z) * rv.z;
        acc[0] += static_cast<double>(rh.w) * rv.w;
        acc[1] += static_cast<double>(rv.x) * rv.x;
        acc[1] += static_cast<double>(rv.y) * rv.y;
        acc[1] += static_cast<double>(rv.z) * rv.z;
        acc[1] += static_cast<double>(rv.w) * rv.w;
    }
    for (long long f = 4 * n4 + t0; f < a.n; f += stride) {
        const float s1 = rin[f] - al * vv[f];
        const float dx = al * (dinv[f] * pv[f]) + om * (dinv[f] * s1);
        x[f] = a.first ? dx : x[f] + dx;
        const float rv = s1 - om * tv[f];
        r[f] = rv;
        acc[0] += static_cast<double>(rhat[f]) * rv;
        acc[1] += static_cast<double>(rv) * rv;
    }
    const bool mx[2] = {false, false};
    block_partial<2>(acc, mx, a.partials + kPartX * gridDim.x, blockIdx.x);
}

// rho, ||r||^2 and the convergence / breakdown flags
__global__ void __launch_bounds__(256) k_bi_x_fin(BiArgs a, unsigned count)
{
    if (a.flags[0]) return;
    const bool mx[2] = {false, false};
    double out[2];
    fold_partials<2>(a.partials + kPartX * count, count, mx, out);
    if (threadIdx.x == 0) {
        a.scal[HS_RHO_OLD] = a.scal[HS_RHO];
        a.scal[HS_RHO] = out[0];
        a.scal[HS_RNORM2] = out[1];
        const int it = a.flags[1] + 1;
        a.flags[1] = it;
        const bool conv = out[1] <= a.eta2 * a.scal[HS_BNORM2];
        const bool stall = (a.scal[HS_OMEGA] == 0.0) || (out[0] == 0.0);
        if (conv || it >= a.max_it || stall) a.flags[0] = 1;
        if (stall && !conv) a.flags[2] = 1;
    }
}

// general DOFs of a trial: F, the ratio, b, dinv and the trial values; block
// partials (max |F|, max ratio, sum b^2) into slots base + block
__global__ void __launch_bounds__(256) k_trial_gen(HsDom d, const double* __restrict__ z, const float* __restrict__ dz,
                                                   double lam, double* __restrict__ z_out, float* __restrict__ b,
                                                   float* __restrict__ dinv, double* partials, int base)
{
    const ZView Z{z, lam != 0.0 ? dz : nullptr, lam};
    double acc[3] = {0.0, 0.0, 0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < d.n_gen; i += gridDim.x * blockDim.x) {
        const int f = d.gen[i];
        const int r = f / d.nc, c = f - r * d.nc;
        double c9[9];
        const double F = cell_general(d, Z, r, c, c9);
        acc[0] = fmax(acc[0], fabs(F));
        if (d.nonlinear) acc[1] = fmax(acc[1], slope2_general(d, Z, r, c));
        const float bf = static_cast<float>(-F);
        b[f] = bf;
        dinv[f] = c9[4] != 0.0 ? static_cast<float>(1.0 / c9[4]) : 0.0f;
        acc[2] += static_cast<double>(bf) * bf;
        if (lam != 0.0) z_out[f] = Z(f);
#pragma unroll
        for (int k = 0; k < 9; ++k) d.c9g[(size_t)k * d.n_gen + i] = static_cast<float>(c9[k]);
    }
    const bool mx[3] = {true, true, false};
    block_partial<3>(acc, mx, partials, base + blockIdx.x);
}

__global__ void __launch_bounds__(256) k_trial(HsDom d, const double* __restrict__ z, const float* __restrict__ dz,
                                               double lam, long long n, double* __restrict__ z_out,
                                               float* __restrict__ b, float* __restrict__ dinv, double* partials,
                                               unsigned* ticket, double* scal, int* flags)
{
    (void)n;
    __shared__ double zs[kSH * kSW];
    const int nr = d.nr, nc = d.nc, tx_n = (nc + kTW - 1) / kTW, nt = tx_n * ((nr + kTH - 1) / kTH);
    const int cl = threadIdx.x & 63;
    double acc[3] = {0.0, 0.0, 0.0};
    // tiles in rounds of gridDim.x, each round's tiles dealt XCD by XCD in
    // contiguous bands (shared halos stay in one L2)
    for (int base = 0; base < nt; base += gridDim.x) {
        const int tile = base + xcd_tile(blockIdx.x, gridDim.x);
        if (tile >= nt) continue;
        const int r0 = (tile / tx_n) * kTH, c0 = (tile % tx_n) * kTW;
        __syncthreads();
        if (lam != 0.0) stage_tile<true, false>(z, dz, lam, nullptr, nr, nc, r0, c0, zs, nullptr);
        else stage_tile<false, false>(z, nullptr, 0.0, nullptr, nr, nc, r0, c0, zs, nullptr);
        long long f[kTH / 4];
        bool in[kTH / 4];
        uint8_t fl[kTH / 4], dl[kTH / 4];
        double znl[kTH / 4];
#pragma unroll
        for (int k = 0; k < kTH / 4; ++k) {
            const int rr = r0 + (threadIdx.x >> 6) + 4 * k, cc = c0 + cl;
            in[k] = rr < nr && cc < nc;
            f[k] = (long long)min(rr, nr - 1) * nc + min(cc, nc - 1);
            fl[k] = d.fast[f[k]];
            dl[k] = d.dof[f[k]];
            znl[k] = d.zn[f[k]];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kTH / 4; ++k) {
            if (!in[k]) continue;
            if (!fl[k]) {
                if (!dl[k]) {
                    b[f[k]] = 0.0f;
                    dinv[f[k]] = 0.0f;
                    // a trial point is a whole surface: cells off the
                    // DOFs keep the iterate's value
                    if (lam != 0.0) z_out[f[k]] = z[f[k]];
                }
                continue;
            }
            double zz[9];
            tile_z9(zs, (threadIdx.x >> 6) + 4 * k, cl, zz);
            double diag;
            const double F = eval_fast(d, zz, znl[k], &diag);
            acc[0] = fmax(acc[0], fabs(F));
            if (d.nonlinear) {
                const double sx = (zz[5] - zz[3]) * d.inv_2h, sy = (zz[7] - zz[1]) * d.inv_2h;
                acc[1] = fmax(acc[1], sx * sx + sy * sy);  // squared slope; ratio in the finisher
            }
            const float bf = static_cast<float>(-F);
            b[f[k]] = bf;
            // Jacobi preconditioner only: the hardware reciprocal suffices
            dinv[f[k]] = diag != 0.0 ? __builtin_amdgcn_rcpf(static_cast<float>(diag)) : 0.0f;
            acc[2] += static_cast<double>(bf) * bf;
            if (lam != 0.0) z_out[f[k]] = zz[4];
        }
    }
    const bool mx[3] = {true, true, false};
    block_partial<3>(acc, mx, partials, blockIdx.x);
}

// max |F|, max slope ratio and the Krylov start of the trial's system
__global__ void __launch_bounds__(256) k_trial_fin(double* partials, unsigned count, double inv_Sc, double* scal,
                                                   int* flags)
{
    const bool mx[3] = {true, true, false};
    double out[3];
    fold_partials<3>(partials, count, mx, out);
    if (threadIdx.x == 0) {
        scal[HS_TRIAL_F] = out[0];
        scal[HS_TRIAL_RATIO] = sqrt(out[1]) * inv_Sc;
        // Krylov start for the system at this point: r = rhat = b, x = 0
        scal[HS_RHO] = out[2];
        scal[HS_BNORM2] = out[2];
        scal[HS_RHO_OLD] = 1.0;
        scal[HS_ALPHA] = 1.0;
        scal[HS_OMEGA] = 1.0;
        flags[0] = (out[2] == 0.0) ? 1 : 0;
        flags[1] = 0;
        flags[2] = 0;
    }
}

// ---------------------------------------------------------------------------
// batched solves (BatchHs): one block per 64 x 16 stencil tile of any grid,
// or per chunk of 256 general DOFs; partial slot = block (tiles first, then
// the chunks); per-grid finishers fold a grid's slots in that order
// ---------------------------------------------------------------------------
__device__ __forceinline__ HsDom batch_dom(const BatchHs& B, const BatchHsGrid& g, double inv_dt, int zn)
{
    HsDom d;
    d.nr = g.nr;
    d.nc = g.nc;
    d.bc = g.bc;
    d.mask = B.mask + g.off;
    d.dof = B.dof + g.off;
    d.fast = B.fast + g.off;
    d.gen = B.gen + g.gen_off;
    d.n_gen = g.n_gen;
    d.h = g.h;
    d.D = g.D;
    d.U = 0.0;
    d.Sc = g.Sc;
    d.cap = g.cap;
    d.inv_dt = inv_dt;
    d.nonlinear = g.nonlinear;
    d.zn = B.zbuf[zn] + g.off;
    d.inv_h = 1.0 / d.h;
    d.inv_2h = 0.5 / d.h;
    d.inv_Sc = 1.0 / d.Sc;
    d.inv_Sc2 = 1.0 / (d.Sc * d.Sc);
    d.c9g = B.c9g + 9ll * g.gen_off;
    return d;
}

__device__ __forceinline__ bool trial_phase(int ph) { return ph == HS_START || ph == HS_SEARCH || ph == HS_WARM; }

// this block's stencil tile: grid and corner
// a batched work item is a 64 x 64 tile, swept as kBands bands of the
// stencil kernels' 64 x 16 LDS tiles: the item's lookups, its reduction
// and its partial are shared by 4096 cells
constexpr int kBTH = kBatchTileRows, kBands = kBTH / kTH;
static_assert(kBatchTileCols == kTW && kBTH % kTH == 0, "batched items are whole columns of LDS tiles");

struct BTileAt {
    int gi, r0, c0;
};
__device__ __forceinline__ BTileAt btile(const BatchHs& B, int slot)
{
    const int gi = B.st_grid[slot];
    const BatchHsGrid& g = B.grids[gi];
    const int t = slot - g.stbase;
    return BTileAt{gi, (t / g.tx_n) * kBTH, (t % g.tx_n) * kTW};
}

// Work lists: the instances a kernel covers this round (those in its
// phase, as the control kernel found them), each with the running count of
// its items (stencil tiles or general-DOF chunks).  Kernels run a fixed
// grid whose blocks walk the items, so a round in which few instances are
// still solving costs launches over their tiles only.  Item w of list L
// is the tile (or chunk) slot of its instance; partials go to that slot,
// so the per-instance reductions are the same whatever the list.
__device__ __forceinline__ int wl_total(const BatchHs& B, int list) { return B.wl_hdr[2 * list + 1]; }

__device__ __forceinline__ int wl_slot(const BatchHs& B, int list, int w, bool chunks)
{
    const int* pre = B.wl_pre + (size_t)list * (B.G + 1);
    const int* inst = B.wl_inst + (size_t)list * B.G;
    int lo = 0, hi = B.wl_hdr[2 * list] - 1;
    while (lo < hi) {  // the last instance whose first item is <= w
        const int mid = (lo + hi + 1) >> 1;
        if (pre[mid] <= w) lo = mid;
        else hi = mid - 1;
    }
    const BatchHsGrid& g = B.grids[inst[lo]];
    return (chunks ? g.gcbase : g.stbase) + (w - pre[lo]);
}

// trial of the START / SEARCH grids on their fast cells (k_trial's tile body)
__device__ __forceinline__ void trial_tile(const BatchHs& B, int slot)
{
    const BTileAt T = btile(B, slot);
    const HsCtl& c = B.ctl[T.gi];
    const int ph = c.phase;
    if (!trial_phase(ph)) return;
    const BatchHsGrid& g = B.grids[T.gi];
    const double lam = ph == HS_START ? 0.0 : c.lambda;
    const bool has_dz = lam != 0.0 && !c.zero_step;
    const HsDom d = batch_dom(B, g, c.inv_dt, c.zn);
    const double* z = B.zbuf[c.zc] + g.off;
    double* z_out = B.zbuf[c.zt] + g.off;
    const float* dz = (ph == HS_WARM ? B.ws_dz : B.x) + g.off;
    float* b = B.b + g.off;
    float* dinv = B.dinv + g.off;
    __shared__ double zs[kSH * kSW];
    const int nr = d.nr, nc = d.nc, c0 = T.c0;
    const int cl = threadIdx.x & 63;
    double acc[3] = {0.0, 0.0, 0.0};
    for (int band = 0; band < kBands; ++band) {
        const int r0 = T.r0 + band * kTH;
        if (r0 >= nr) break;
        __syncthreads();
        if (has_dz) stage_tile<true, false>(z, dz, lam, nullptr, nr, nc, r0, c0, zs, nullptr);
        else stage_tile<false, false>(z, nullptr, 0.0, nullptr, nr, nc, r0, c0, zs, nullptr);
        long long f[kTH / 4];
        bool in[kTH / 4];
        uint8_t fl[kTH / 4], dl[kTH / 4];
        double znl[kTH / 4];
#pragma unroll
        for (int k = 0; k < kTH / 4; ++k) {
            const int rr = r0 + (threadIdx.x >> 6) + 4 * k, cc = c0 + cl;
            in[k] = rr < nr && cc < nc;
            f[k] = (long long)min(rr, nr - 1) * nc + min(cc, nc - 1);
            fl[k] = d.fast[f[k]];
            dl[k] = d.dof[f[k]];
            znl[k] = d.zn[f[k]];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kTH / 4; ++k) {
            if (!in[k]) continue;
            if (!fl[k]) {
                if (!dl[k]) {
                    b[f[k]] = 0.0f;
                    dinv[f[k]] = 0.0f;
                    if (lam != 0.0) z_out[f[k]] = z[f[k]];
                }
                continue;
            }
            double zz[9];
            tile_z9(zs, (threadIdx.x >> 6) + 4 * k, cl, zz);
            double diag;
            const double F = eval_fast(d, zz, znl[k], &diag);
            acc[0] = fmax(acc[0], fabs(F));
            if (d.nonlinear) {
                const double sx = (zz[5] - zz[3]) * d.inv_2h, sy = (zz[7] - zz[1]) * d.inv_2h;
                acc[1] = fmax(acc[1], sx * sx + sy * sy);
            }
            const float bf = static_cast<float>(-F);
            b[f[k]] = bf;
            dinv[f[k]] = diag != 0.0 ? __builtin_amdgcn_rcpf(static_cast<float>(diag)) : 0.0f;
            acc[2] += static_cast<double>(bf) * bf;
            if (lam != 0.0) z_out[f[k]] = zz[4];
        }
    }
    const bool mx[3] = {true, true, false};
    block_partial<3>(acc, mx, B.partials, slot);
}

__global__ void __launch_bounds__(256) kb_trial(BatchHs B)
{
    const int total = wl_total(B, WL_TTILE);
    for (int w = blockIdx.x; w < total; w += gridDim.x) {
        __syncthreads();
        trial_tile(B, wl_slot(B, WL_TTILE, w, false));
    }
}

// trial of the general DOFs (k_trial_gen's body), one chunk of 256 per item
__device__ __forceinline__ void trial_chunk(const BatchHs& B, int slot)
{
    const int gi = B.gc_grid[slot];
    const HsCtl& c = B.ctl[gi];
    const int ph = c.phase;
    if (!trial_phase(ph)) return;
    const BatchHsGrid& g = B.grids[gi];
    const double lam = ph == HS_START ? 0.0 : c.lambda;
    const bool has_dz = lam != 0.0 && !c.zero_step;
    const HsDom d = batch_dom(B, g, c.inv_dt, c.zn);
    const double* z = B.zbuf[c.zc] + g.off;
    double* z_out = B.zbuf[c.zt] + g.off;
    const ZView Z{z, has_dz ? (ph == HS_WARM ? B.ws_dz : B.x) + g.off : nullptr, lam};
    float* b = B.b + g.off;
    float* dinv = B.dinv + g.off;
    double acc[3] = {0.0, 0.0, 0.0};
    const int i = (slot - g.gcbase) * 256 + threadIdx.x;
    if (i < d.n_gen) {
        const int f = d.gen[i];
        const int r = f / d.nc, cc = f - r * d.nc;
        double c9[9];
        const double F = cell_general(d, Z, r, cc, c9);
        acc[0] = fabs(F);
        if (d.nonlinear) acc[1] = slope2_general(d, Z, r, cc);
        const float bf = static_cast<float>(-F);
        b[f] = bf;
        dinv[f] = c9[4] != 0.0 ? static_cast<float>(1.0 / c9[4]) : 0.0f;
        acc[2] = static_cast<double>(bf) * bf;
        if (lam != 0.0) z_out[f] = Z(f);
#pragma unroll
        for (int k = 0; k < 9; ++k) d.c9g[(size_t)k * d.n_gen + i] = static_cast<float>(c9[k]);
    }
    const bool mx[3] = {true, true, false};
    block_partial<3>(acc, mx, B.partials, B.NT + slot);
}

__global__ void __launch_bounds__(256) kb_trial_gen(BatchHs B)
{
    const int total = wl_total(B, WL_TCHUNK);
    for (int w = blockIdx.x; w < total; w += gridDim.x) {
        __syncthreads();
        trial_chunk(B, wl_slot(B, WL_TCHUNK, w, true));
    }
}

// folds grid g's partial slots (its tiles, then -- for kernels with chunk
// blocks -- its general chunks) in slot order; result in thread 0
template <int NV>
__device__ __forceinline__ void fold_grid(const BatchHs& B, const BatchHsGrid& g, bool with_chunks,
                                          const bool (&is_max)[NV], double (&w)[NV])
{
#pragma unroll
    for (int i = 0; i < NV; ++i) w[i] = is_max[i] ? -1.7976931348623157e308 : 0.0;
    const int n = g.nst + (with_chunks ? g.ngc : 0);
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        const int slot = k < g.nst ? g.stbase + k : B.NT + g.gcbase + (k - g.nst);
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            const double x = B.partials[(size_t)slot * NV + i];
            w[i] = is_max[i] ? (w[i] > x ? w[i] : x) : w[i] + x;
        }
    }
    block_reduce<NV>(w, is_max);
}

// max |F|, max slope ratio and the Krylov start of each trial's system
__global__ void __launch_bounds__(256) kb_trial_fin(BatchHs B)
{
    HsCtl& c = B.ctl[blockIdx.x];
    if (!trial_phase(c.phase)) return;
    const BatchHsGrid& g = B.grids[blockIdx.x];
    const bool mx[3] = {true, true, false};
    double out[3];
    fold_grid<3>(B, g, true, mx, out);
    if (threadIdx.x == 0) {
        c.trial_f = out[0];
        c.trial_ratio = sqrt(out[1]) * (1.0 / g.Sc);
        c.rho = out[2];
        c.bnorm2 = out[2];
        c.rho_old = 1.0;
        c.alpha = 1.0;
        c.omega = 1.0;
        c.kdone = out[2] == 0.0 ? 1 : 0;
        c.kits = 0;
        c.kbreak = 0;
    }
}

__device__ __forceinline__ bool krylov_active(const HsCtl& c) { return c.phase == HS_KRYLOV && !c.kdone; }

// one row-major pass over this block's tile: fn(f) for each cell
template <class F>
__device__ __forceinline__ void tile_cells(const BatchHsGrid& g, int r0, int c0, F&& fn)
{
    const int cc = c0 + (threadIdx.x & 63);
    if (cc >= g.nc) return;
#pragma unroll 4
    for (int k = 0; k < kBTH / 4; ++k) {
        const int rr = r0 + (threadIdx.x >> 6) + 4 * k;
        if (rr < g.nr) fn((long long)rr * g.nc + cc);
    }
}

// a grid's Krylov vectors for its current iteration: p and v are read from
// the pair of the previous iteration and written to the other
struct BatchKrylov {
    const float *p_old, *v_old;
    float *p, *v;
};
__device__ __forceinline__ BatchKrylov batch_krylov(const BatchHs& B, const HsCtl& c, long long off)
{
    const bool odd = c.kits & 1;
    return BatchKrylov{(odd ? B.p2 : B.p) + off, (odd ? B.v2 : B.v) + off, (odd ? B.p : B.p2) + off,
                       (odd ? B.v : B.v2) + off};
}

// the views of the single-grid kernels on one grid of the batch
__device__ __forceinline__ PhatView batch_phat_view(const BatchHs& B, const HsCtl& c, const BatchHsGrid& g)
{
    const BatchKrylov k = batch_krylov(B, c, g.off);
    PhatView pv;
    pv.b = B.b + g.off;
    pv.r = B.r + g.off;
    pv.p = k.p_old;
    pv.v = k.v_old;
    pv.dinv = B.dinv + g.off;
    pv.first = c.kits == 0;
    if (pv.first) {
        pv.bt = pv.om = 0.0f;
    } else {
        pv.bt = static_cast<float>((c.rho / c.rho_old) * (c.alpha / c.omega));
        pv.om = static_cast<float>(c.omega);
    }
    return pv;
}

__device__ __forceinline__ ShatView batch_shat_view(const BatchHs& B, const HsCtl& c, const BatchHsGrid& g)
{
    ShatView sv;
    sv.r = (c.kits == 0 ? B.b : B.r) + g.off;
    sv.v = batch_krylov(B, c, g.off).v;
    sv.dinv = B.dinv + g.off;
    sv.al = static_cast<float>(c.alpha);
    return sv;
}

// y = J x on the tile's fast rows (0 off the DOFs) and dot(w, y); the
// general rows come from the chunk blocks
// y = J x on the item (its fast rows from the tile, or its chunk of general
// rows); partials w . y (NV = 1) or y . w and y . y (NV = 2).  w is loaded
// with the other per-cell inputs, before the barrier.
template <int NV>
__device__ __forceinline__ void dot_acc(double (&acc)[NV], float y, float w)
{
    if constexpr (NV == 1) {
        acc[0] += static_cast<double>(w) * y;
    } else {
        acc[0] += static_cast<double>(y) * w;
        acc[1] += static_cast<double>(y) * y;
    }
}

// y = J x with x given by the grid's view (PhatView: also writes p over
// the tile; ShatView: the dot products take s from the view); the general
// rows come from the chunk items
template <int NV, bool PV>
__device__ __forceinline__ void batch_matvec(const BatchHs& B, bool chunk, int slot, double (&acc)[NV])
{
    __shared__ double zs[kSH * kSW];
    __shared__ float xs[kSH * kSW], rs[kSH * kSW];
    using View = typename std::conditional<PV, PhatView, ShatView>::type;
    auto make_view = [&](const HsCtl& c, const BatchHsGrid& g) {
        if constexpr (PV) return batch_phat_view(B, c, g);
        else return batch_shat_view(B, c, g);
    };
    if (!chunk) {
        const BTileAt T = btile(B, slot);
        const HsCtl& c = B.ctl[T.gi];
        if (!krylov_active(c)) return;
        const BatchHsGrid& g = B.grids[T.gi];
        const HsDom d = batch_dom(B, g, c.inv_dt, c.zn);
        const double* z = B.zbuf[c.zc] + g.off;
        const View view = make_view(c, g);
        const BatchKrylov kv = batch_krylov(B, c, g.off);
        const float* rhat = B.b + g.off;
        float* y = PV ? kv.v : B.t + g.off;
        const int nr = g.nr, nc = g.nc, c0 = T.c0, cl = threadIdx.x & 63;
        for (int band = 0; band < kBands; ++band) {
            const int r0 = T.r0 + band * kTH;
            if (r0 >= nr) break;
            __syncthreads();
            stage_tile_view(z, view, nr, nc, r0, c0, zs, xs, rs);
            long long f[kTH / 4];
            bool in[kTH / 4];
            uint8_t fl[kTH / 4], dl[kTH / 4];
            float wl[kTH / 4];
#pragma unroll
            for (int k = 0; k < kTH / 4; ++k) {
                const int rr = r0 + (threadIdx.x >> 6) + 4 * k, cc = c0 + cl;
                in[k] = rr < nr && cc < nc;
                f[k] = (long long)min(rr, nr - 1) * nc + min(cc, nc - 1);
                fl[k] = d.fast[f[k]];
                dl[k] = d.dof[f[k]];
                if constexpr (PV) wl[k] = rhat[f[k]];
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < kTH / 4; ++k) {
                if (!in[k]) continue;
                const float own = rs[((threadIdx.x >> 6) + 4 * k + 1) * kSW + cl + 1];  // p or s of the cell
                float w;
                if constexpr (PV) {
                    kv.p[f[k]] = own;
                    w = wl[k];
                } else {
                    w = own;
                }
                if (fl[k]) {
                    const float yv = jx_tile(d, zs, xs, (threadIdx.x >> 6) + 4 * k, cl);
                    y[f[k]] = yv;
                    dot_acc<NV>(acc, yv, w);
                } else if (!dl[k]) {
                    y[f[k]] = 0.0f;
                }
            }
        }
    } else {
        const int gi = B.gc_grid[slot];
        const HsCtl& c = B.ctl[gi];
        if (!krylov_active(c)) return;
        const BatchHsGrid& g = B.grids[gi];
        const HsDom d = batch_dom(B, g, c.inv_dt, c.zn);
        const int i = (slot - g.gcbase) * 256 + threadIdx.x;
        if (i < d.n_gen) {
            const View view = make_view(c, g);
            const int f = d.gen[i];
            float w;
            if constexpr (PV) w = B.b[g.off + f];
            else w = view.sval(view.load(f));
            const float yv = jx_gen_view(d, view, i, f);
            if constexpr (PV) batch_krylov(B, c, g.off).v[f] = yv;
            else B.t[g.off + f] = yv;
            dot_acc<NV>(acc, yv, w);
        }
    }
    const bool mx[NV] = {};
    block_partial<NV>(acc, mx, B.partials, chunk ? B.NT + slot : slot);
}

// the items of a product: the Krylov tiles, then the Krylov chunks
template <class F>
__device__ __forceinline__ void matvec_items(const BatchHs& B, F&& fn)
{
    const int nt = wl_total(B, WL_KTILE), total = nt + wl_total(B, WL_KCHUNK);
    for (int w = blockIdx.x; w < total; w += gridDim.x) {
        __syncthreads();
        if (w < nt) fn(false, wl_slot(B, WL_KTILE, w, false));
        else fn(true, wl_slot(B, WL_KCHUNK, w - nt, true));
    }
}

// p = r + beta (p_old - omega v_old); v = J phat, partial rhat . v (rhat = b)
__global__ void __launch_bounds__(256) kb_bi_v(BatchHs B)
{
    matvec_items(B, [&](bool chunk, int slot) {
        double acc[1] = {0.0};
        batch_matvec<1, true>(B, chunk, slot, acc);
    });
}

// t = J shat (s = r - alpha v), partials t . s and t . t
__global__ void __launch_bounds__(256) kb_bi_t(BatchHs B)
{
    matvec_items(B, [&](bool chunk, int slot) {
        double acc[2] = {0.0, 0.0};
        batch_matvec<2, false>(B, chunk, slot, acc);
    });
}

// alpha = rho / (rhat . v) per grid (breakdown stops that grid's solve)
__global__ void __launch_bounds__(256) kb_bi_vfin(BatchHs B)
{
    HsCtl& c = B.ctl[blockIdx.x];
    if (!krylov_active(c)) return;
    const bool mx[1] = {false};
    double out[1];
    fold_grid<1>(B, B.grids[blockIdx.x], true, mx, out);
    if (threadIdx.x == 0) {
        if (out[0] == 0.0) {
            c.kdone = 1;
            c.kbreak = 1;
        } else {
            c.alpha = c.rho / out[0];
        }
    }
}

// omega = (t . s) / (t . t) per grid
__global__ void __launch_bounds__(256) kb_bi_tfin(BatchHs B)
{
    HsCtl& c = B.ctl[blockIdx.x];
    if (!krylov_active(c)) return;
    const bool mx[2] = {false, false};
    double out[2];
    fold_grid<2>(B, B.grids[blockIdx.x], true, mx, out);
    if (threadIdx.x == 0) c.omega = out[1] > 0.0 ? out[0] / out[1] : 0.0;
}

// x += alpha phat + omega shat; r = s - omega t; partials rhat . r, r . r
__device__ __forceinline__ void bi_x_tile(const BatchHs& B, int slot)
{
    const BTileAt T = btile(B, slot);
    const HsCtl& c = B.ctl[T.gi];
    if (!krylov_active(c)) return;
    const BatchHsGrid& g = B.grids[T.gi];
    const float al = static_cast<float>(c.alpha), om = static_cast<float>(c.omega);
    const bool first = c.kits == 0;
    const BatchKrylov kv = batch_krylov(B, c, g.off);
    const float* pv = kv.p;
    const float* vv = kv.v;
    const float* dinv = B.dinv + g.off;
    const float* rin = (first ? B.b : B.r) + g.off;
    const float* tv = B.t + g.off;
    const float* rhat = B.b + g.off;
    float* x = B.x + g.off;
    float* r = B.r + g.off;
    double acc[2] = {0.0, 0.0};
    tile_cells(g, T.r0, T.c0, [&](long long f) {
        // phat = M^-1 p, s = r - alpha v, shat = M^-1 s: the expressions of
        // the products' views
        const float s1 = rin[f] - al * vv[f];
        const float dx = al * (dinv[f] * pv[f]) + om * (dinv[f] * s1);
        x[f] = first ? dx : x[f] + dx;
        const float rv = s1 - om * tv[f];
        r[f] = rv;
        acc[0] += static_cast<double>(rhat[f]) * rv;
        acc[1] += static_cast<double>(rv) * rv;
    });
    const bool mx[2] = {false, false};
    block_partial<2>(acc, mx, B.partials, slot);
}

__global__ void __launch_bounds__(256) kb_bi_x(BatchHs B)
{
    const int total = wl_total(B, WL_KTILE);
    for (int w = blockIdx.x; w < total; w += gridDim.x) {
        __syncthreads();
        bi_x_tile(B, wl_slot(B, WL_KTILE, w, false));
    }
}

// rho, ||r||^2 and the convergence / breakdown flags per grid
__global__ void __launch_bounds__(256) kb_bi_xfin(BatchHs B)
{
    HsCtl& c = B.ctl[blockIdx.x];
    if (!krylov_active(c)) return;
    const BatchHsGrid& g = B.grids[blockIdx.x];
    const bool mx[2] = {false, false};
    double out[2];
    fold_grid<2>(B, g, false, mx, out);  // kb_bi_x has tile blocks only
    if (threadIdx.x == 0) {
        c.rho_old = c.rho;
        c.rho = out[0];
        c.rnorm2 = out[1];
        const int it = c.kits + 1;
        c.kits = it;
        const bool conv = out[1] <= c.eta * c.eta * c.bnorm2;
        const bool stall = (c.omega == 0.0) || (out[0] == 0.0);
        if (conv || it >= g.krylov_max || stall) c.kdone = 1;
        if (stall && !conv) c.kbreak = 1;
    }
}

// the trial buffer after zc and zn in the order 1, 2, 3
__device__ __forceinline__ int free_buffer(int zc, int zn)
{
    for (int q = 1; q <= 3; ++q)
        if (q != zc && q != zn) return q;
    return 1;
}

// a sub-step's Newton solve ended (its accepted iterations; converged or
// not): the sub-step controller of compute_hillslope_newton_transient
// (lsdtt_hillslope_newton_cpu.cpp:690-788)
__device__ __forceinline__ void end_solve(HsCtl& c, int its, bool converged)
{
    c.newton_total += its;
    if (!converged) {
        c.zc = c.zn;
        c.zt = free_buffer(c.zc, c.zn);
        ++c.rejected;
        if (c.rejected > 40 || c.step <= c.dt_min) {
            c.phase = HS_FAILED;
            return;
        }
        c.dt = fmax(c.dt_min, c.step * 0.5);
    } else {
        c.remaining -= c.step;
        if (c.remaining > 0.0) {
            c.zn = c.zc;
            c.zt = free_buffer(c.zc, c.zn);
        }
        ++c.substeps;
        if (its <= 4) c.dt = fmin(c.dt_max, c.step * 1.7);
        else if (its >= 15) c.dt = fmax(c.dt_min, c.step * 0.5);
        if (!(c.remaining > 0.0)) {
            c.phase = HS_DONE;
            return;
        }
    }
    c.step = fmin(c.dt, c.remaining);
    c.inv_dt = 1.0 / c.step;
    c.phase = HS_START;
}

// the host loop's decisions, per grid (one thread each).  stage 0 follows
// the Krylov iteration: a finished solve starts its line search.  stage 1
// follows the trials: a START trial sets the tolerance, a SEARCH trial is
// accepted or halves lambda (compute_hillslope_newton's line search,
// lsdtt_hillslope_newton_cpu.cpp:553-633, with restoration mode while the
// iterate's slope ratio is above sqrt(cap)).
__device__ __forceinline__ void control_one(const BatchHs& B, int gi, int stage)
{
    HsCtl& c = B.ctl[gi];
    const BatchHsGrid& g = B.grids[gi];
    const double feasible = sqrt(g.cap);
    if (stage == 0) {
        if (c.phase == HS_START || c.phase == HS_KRYLOV || c.phase == HS_SEARCH || c.phase == HS_WARM) ++c.rounds;
        if (c.phase == HS_KRYLOV && c.kdone) {
            c.krylov_total += c.kits;
            // no iteration ran (b = 0): the correction is zero.  (Each trial
            // resets kits for the system at its own point, so the line
            // search keeps this instead.)
            c.zero_step = c.kits == 0 ? 1 : 0;
            c.restoring = (g.nonlinear && c.ratio > feasible) ? 1 : 0;
            if (c.restoring) ++c.restoration_its;
            c.lambda = 1.0;
            c.ls = 0;
            c.phase = HS_SEARCH;
        }
        return;
    }
    if (c.phase == HS_START) {
        c.fnorm = c.trial_f;
        c.ratio = c.trial_ratio;
        c.tol = fmax(g.abs_tol, g.rel_tol * c.fnorm);
        if (c.first_solve) {
            c.f_init = c.fnorm;
            c.first_solve = 0;
        }
        c.it = 0;
        c.eta = sqrt(g.eta2);
        if (c.fnorm <= c.tol) {
            end_solve(c, 0, true);
        } else if (g.newton_max <= 0) {
            end_solve(c, 0, false);
        } else if (g.ws_mode == 3 && !c.ws_tried && c.substeps == 0 && c.rejected == 0 && B.ws_ok[gi]) {
            // first try z^n + the last step's hillslope change (a trial at
            // lambda 1 of that change); plain Newton from z^n is the fallback
            c.ws_tried = 1;
            c.lambda = 1.0;
            c.zero_step = 0;
            c.phase = HS_WARM;
        } else {
            c.phase = HS_KRYLOV;
        }
        return;
    }
    if (c.phase == HS_WARM) {
        const double ft = c.trial_f, rt = c.trial_ratio;
        c.ws_f = ft;
        if (ft < c.fnorm && (!g.nonlinear || rt <= feasible)) {
            // the predicted start is the first iterate; its system came with the trial
            c.zc = c.zt;
            c.zt = free_buffer(c.zc, c.zn);
            c.fnorm = ft;
            c.ratio = rt;
            c.ws_accepted = 1;
            if (c.fnorm <= c.tol) end_solve(c, 0, true);
            else c.phase = HS_KRYLOV;
        } else {
            c.phase = HS_START;  // the system at z^n again
        }
        return;
    }
    if (c.phase != HS_SEARCH) return;
    const double ft = c.trial_f, rt = c.trial_ratio;
    bool accept;
    if (g.nonlinear && c.restoring) accept = rt < c.ratio;
    else if (g.nonlinear && rt > feasible) accept = false;
    else accept = ft < c.fnorm;
    if (!accept) {
        c.lambda *= 0.5;
        if (++c.ls >= g.ls_max) end_solve(c, c.it, false);
        return;
    }
    // the accepted trial point is the new iterate; its system (b, dinv,
    // Krylov start) came with the trial
    c.zc = c.zt;
    c.zt = free_buffer(c.zc, c.zn);
    const double fprev = c.fnorm;
    c.fnorm = ft;
    c.ratio = rt;
    const int its = ++c.it;
    if (c.fnorm <= c.tol) {
        end_solve(c, its, true);
    } else if (its >= g.newton_max) {
        end_solve(c, its, false);
    } else {
        if (g.forcing == 1) c.eta = forcing_eta(c.eta, c.fnorm, fprev, c.tol);
        c.phase = HS_KRYLOV;
    }
}


// the instances of one phase as work list `list` (their tiles, or their
// general-DOF chunks): one block, block-wide scans over the instances
template <class Pred>
__device__ void build_list(const BatchHs& B, int list, bool chunks, Pred&& on_list)
{
    __shared__ int s_on[256], s_it[256];
    int* inst = B.wl_inst + (size_t)list * B.G;
    int* pre = B.wl_pre + (size_t)list * (B.G + 1);
    int run_n = 0, run_items = 0;
    for (int base = 0; base < B.G; base += blockDim.x) {
        const int gi = base + threadIdx.x;
        int items = 0;
        if (gi < B.G && on_list(B.ctl[gi])) items = chunks ? B.grids[gi].ngc : B.grids[gi].nst;
        const int on = items > 0 ? 1 : 0;
        s_on[threadIdx.x] = on;
        s_it[threadIdx.x] = items;
        __syncthreads();
        for (int o = 1; o < (int)blockDim.x; o <<= 1) {  // inclusive scans
            const int a = threadIdx.x >= (unsigned)o ? s_on[threadIdx.x - o] : 0;
            const int b = threadIdx.x >= (unsigned)o ? s_it[threadIdx.x - o] : 0;
            __syncthreads();
            s_on[threadIdx.x] += a;
            s_it[threadIdx.x] += b;
            __syncthreads();
        }
        if (on) {
            const int k = run_n + s_on[threadIdx.x] - 1;
            inst[k] = gi;
            pre[k] = run_items + s_it[threadIdx.x] - items;
        }
        run_n += s_on[blockDim.x - 1];
        run_items += s_it[blockDim.x - 1];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        pre[run_n] = run_items;
        B.wl_hdr[2 * list] = run_n;
        B.wl_hdr[2 * list + 1] = run_items;
    }
}

__device__ __forceinline__ bool trial_listed(const HsCtl& c) { return trial_phase(c.phase); }
__device__ __forceinline__ bool krylov_listed(const HsCtl& c) { return krylov_active(c); }

// One block: every instance's decisions, then the work lists of what
// follows (stage 0: the trials of this round; stage 1: the Krylov
// iteration of the next round).
__global__ void __launch_bounds__(256) kb_control(BatchHs B, int stage)
{
    for (int gi = threadIdx.x; gi < B.G; gi += blockDim.x) control_one(B, gi, stage);
    __syncthreads();
    if (stage == 0) {
        build_list(B, WL_TTILE, false, trial_listed);
        build_list(B, WL_TCHUNK, true, trial_listed);
    } else {
        build_list(B, WL_KTILE, false, krylov_listed);
        build_list(B, WL_KCHUNK, true, krylov_listed);
    }
}

// a new step: every grid with a hillslope solve starts its first sub-step;
// no Krylov work is listed yet
__global__ void __launch_bounds__(256) kb_hs_start(BatchHs B)
{
    for (int gi = threadIdx.x; gi < B.G; gi += blockDim.x) {
        HsCtl& c = B.ctl[gi];
        const BatchHsGrid& g = B.grids[gi];
        c = HsCtl{};
        c.zc = c.zn = 0;
        c.zt = 1;
        c.dt = c.remaining = c.dt_max = g.dt_total;
        c.dt_min = g.dt_total * 1e-9;
        c.step = fmin(c.dt, c.remaining);
        c.inv_dt = 1.0 / c.step;
        c.first_solve = 1;
        c.phase = g.enabled ? HS_START : HS_IDLE;
    }
    if (threadIdx.x < 2 * WL_N) B.wl_hdr[threadIdx.x] = 0;
}

}  // namespace

int hs_general_blocks(int n_gen) { return gen_grid(n_gen); }

namespace {
// blocks of 256 threads of `k` that one CU holds at once, times the CUs
template <class K>
int resident_blocks(K k)
{
    int dev = 0, cus = 0, per = 0;
    LSDTT_HIP(hipGetDevice(&dev));
    LSDTT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    LSDTT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k, kB, 0));
    return std::max(1, per) * std::max(1, cus);
}
}  // namespace

int hs_bicgstab_grid(long long n, int cap)
{
    // one grid for the three kernels (each folds the previous one's block
    // partials): the blocks that are all resident at once, so no block
    // waits for a second round behind the others
    int g = resident_blocks(k_bi_pv);
    for (int v : {resident_blocks(k_bi_st), resident_blocks(k_bi_x)}) g = std::min(g, v);
    return std::max(1, std::min({g, cap, (int)std::min<long long>(div_up(n, kB), 1ll << 30)}));
}

void hs_fast_flags(const HsDom& d, uint8_t* fast, hipStream_t s)
{
    const long long n = (long long)d.nr * d.nc;
    k_fast_flags<<<std::max(1, std::min(4096, div_up(n, kB))), kB, 0, s>>>(d, fast);
    LSDTT_LAUNCH_CHECK();
}

void hs_bicgstab_iteration(const BiArgs& a, int grid, hipStream_t s, KernelTimer* t)
{
    const double n = static_cast<double>(a.n);
    if (t) t->begin("bicgstab_pv", s);
    k_bi_pv<<<grid, kB, 0, s>>>(a);
    // z 8, r 4, p_old 4, v_old 4, dinv 4, fast 1, dof 1, rhat 4 in; p 4, v 4 out
    if (t) t->end(s, 38.0 * n);
    if (t) t->begin("bicgstab_st", s);
    k_bi_st<<<grid, kB, 0, s>>>(a);
    // z 8, r 4, v 4, dinv 4, fast 1, dof 1 in; t 4 out
    if (t) t->end(s, 26.0 * n);
    if (t) t->begin("bicgstab_x", s);
    k_bi_x<<<grid, kB, 0, s>>>(a);
    k_bi_x_fin<<<1, kB, 0, s>>>(a, (unsigned)grid);
    // p 4, dinv 4, r 4, v 4, t 4, rhat 4, x 4 in; x 4, r 4 out
    if (t) t->end(s, 36.0 * n);
    LSDTT_LAUNCH_CHECK();
}

void hs_trial(const HsDom& d, const double* z, const float* dz, double lam, long long n, double* z_out, float* b,
              float* dinv, double* partials, unsigned* ticket, double* scal, int* flags, int grid, hipStream_t s,
              hipStream_t aux, hipEvent_t fork, hipEvent_t join)
{
    const bool side = d.n_gen > 0 && aux && fork && join;
    if (side) {
        LSDTT_HIP(hipEventRecord(fork, s));
        LSDTT_HIP(hipStreamWaitEvent(aux, fork, 0));
        k_trial_gen<<<gen_grid(d.n_gen), kB, 0, aux>>>(d, z, dz, lam, z_out, b, dinv, partials, grid);
        LSDTT_HIP(hipEventRecord(join, aux));
    } else if (d.n_gen > 0) {
        k_trial_gen<<<gen_grid(d.n_gen), kB, 0, s>>>(d, z, dz, lam, z_out, b, dinv, partials, grid);
    }
    k_trial<<<grid, kB, 0, s>>>(d, z, dz, lam, n, z_out, b, dinv, partials, ticket, scal, flags);
    if (side) LSDTT_HIP(hipStreamWaitEvent(s, join, 0));
    k_trial_fin<<<1, kB, 0, s>>>(partials, (unsigned)(grid + gen_grid(d.n_gen)), d.inv_Sc, scal, flags);
    LSDTT_LAUNCH_CHECK();
}

void hsb_start(const BatchHs& B, int n_grids, hipStream_t s)
{
    (void)n_grids;
    kb_hs_start<<<1, kB, 0, s>>>(B);
    LSDTT_LAUNCH_CHECK();
}

void hsb_round(const BatchHs& B, int n_grids, hipStream_t s, int tiles, int chunks)
{
    // one block per tile (chunk) of the batch, enough for every list; the
    // blocks beyond a list's length exit at once, and the item lookups of
    // the others overlap across the grid
    const int nt = tiles >= 0 ? std::min(tiles, B.NT) : B.NT;
    const int ng = chunks >= 0 ? std::min(chunks, B.NGC) : B.NGC;
    const int gt = std::max(1, nt), gc = std::max(1, ng), gm = std::max(1, nt + ng);
    kb_bi_v<<<gm, kB, 0, s>>>(B);
    kb_bi_vfin<<<n_grids, kB, 0, s>>>(B);
    kb_bi_t<<<gm, kB, 0, s>>>(B);
    kb_bi_tfin<<<n_grids, kB, 0, s>>>(B);
    kb_bi_x<<<gt, kB, 0, s>>>(B);
    kb_bi_xfin<<<n_grids, kB, 0, s>>>(B);
    kb_control<<<1, kB, 0, s>>>(B, 0);
    kb_trial<<<gt, kB, 0, s>>>(B);
    if (B.NGC > 0) kb_trial_gen<<<gc, kB, 0, s>>>(B);
    kb_trial_fin<<<n_grids, kB, 0, s>>>(B);
    kb_control<<<1, kB, 0, s>>>(B, 1);
    LSDTT_LAUNCH_CHECK();
}

}  // namespace lsdtt_gpu
