// hillslope.hpp -- backward-Euler hillslope diffusion operator (linear or
// Roering) and its Newton-Krylov pieces on the device.
//
// The discrete operator is the reference's single operator
// (src/cpu/lsdtt_hillslope_operator_cpu.hpp:83-321):
//   F_P = inv_dt (z_P - zn_P) - U - (D / h^2) sum_f alpha_f (z_nb - z_P)
// over the degrees of freedom (valid cells off the 'o' edges), with
// neighbours resolved by the boundary rules ('p' wraps, other edges and
// masked cells resolve to the cell itself).  Its Jacobian is 9-point and
// non-symmetric (src/cpu/lsdtt_hillslope_newton_cpu.cpp:137-186, 415-471).
//
// Precision: F, the line-search residuals and the state are FP64; Jacobian
// products are formed matrix-free in FP32 from FP64 elevation differences,
// so each product streams the elevation (8 B), the vector (4 B) and the result
// (4 B) per cell instead of an assembled matrix.
#pragma once

#include "common.hpp"

namespace lsdtt_gpu {

struct HsDom {
    int nr, nc;
    BC bc;
    const uint8_t* mask;
    const uint8_t* dof;
    const uint8_t* fast;  // DOF whose 3x3 block is interior and valid
    const int* gen;       // the other DOFs (grid edge or next to a masked cell)
    int n_gen;
    double h, D, U, Sc, cap, inv_dt;
    int nonlinear;
    const double* zn;
    double inv_h, inv_2h, inv_Sc, inv_Sc2;  // 1/h, 1/(2h), 1/Sc, 1/Sc^2
    // Jacobian rows of the general DOFs (FP32, 9 x n_gen, slot-major),
    // written by the trial that produced the current iterate
    float* c9g;
};

// fills the reciprocal fields of a domain
inline HsDom with_reciprocals(HsDom d)
{
    d.inv_h = 1.0 / d.h;
    d.inv_2h = 0.5 / d.h;
    d.inv_Sc = 1.0 / d.Sc;
    d.inv_Sc2 = 1.0 / (d.Sc * d.Sc);
    return d;
}

struct BiArgs {
    HsDom d;
    long long n;
    const double* z;     // Newton iterate (the Jacobian is taken here)
    const float* dinv;
    const float* b;
    float *x, *r, *rhat, *p, *phat, *v, *s, *shat, *t;
    double* scal;
    double* partials;
    unsigned* ticket;
    int* flags;  // [0] done, [1] iterations, [2] breakdown
    double eta2;
    int max_it;
    int first;  // first iteration: r = rhat = b, p = v = x = 0 (not read)
};

// scalar slots used by the solver (double array)
enum {
    HS_RHO = 0, HS_RHO_OLD, HS_ALPHA, HS_OMEGA, HS_BNORM2, HS_RNORM2, HS_TRIAL_F, HS_TRIAL_RATIO, HS_NSCAL
};

void hs_fast_flags(const HsDom& d, uint8_t* fast, hipStream_t s);
// companion-kernel blocks for n_gen general cells (their partials follow the
// hot kernel's `grid` partials, so `partials` holds grid + this many slots)
int hs_general_blocks(int n_gen);
// grid (at most cap blocks) of the BiCGSTAB kernels for n cells: as many
// blocks as the device holds at once
int hs_bicgstab_grid(long long n, int cap);
void hs_bicgstab_iteration(const BiArgs& a, int grid, hipStream_t s, class KernelTimer* timer);
// One evaluation at the trial point zt = z + lam dz (dz unused when lam is
// 0): scal[HS_TRIAL_F] = max |F(zt)|, scal[HS_TRIAL_RATIO] = max centred-
// gradient ratio over the DOFs; and the next linear system at zt, should
// the point be accepted: b = -F(zt), dinv = 1 / J_PP(zt) (FP32), all of
// zt into z_out when lam != 0 (the DOFs moved, every other cell z's value)
// and the Krylov start (rho =
// ||b||^2, flags cleared).  With `aux` (a second stream of the same device)
// and two events, the general cells' companion kernel runs on aux beside
// the main kernel instead of before it.
void hs_trial(const HsDom& d, const double* z, const float* dz, double lam, long long n, double* z_out, float* b,
              float* dinv, double* partials, unsigned* ticket, double* scal, int* flags, int grid, hipStream_t s,
              hipStream_t aux = nullptr, hipEvent_t fork = nullptr, hipEvent_t join = nullptr);

}  // namespace lsdtt_gpu
