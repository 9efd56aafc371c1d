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
};

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
};

// scalar slots used by the solver (double array)
enum {
    HS_RHO = 0, HS_RHO_OLD, HS_ALPHA, HS_OMEGA, HS_BNORM2, HS_RNORM2, HS_TRIAL_F, HS_TRIAL_RATIO, HS_NSCAL
};

void hs_fast_flags(const HsDom& d, uint8_t* fast, hipStream_t s);
// companion-kernel blocks for n_gen general cells (their partials follow the
// hot kernel's `grid` partials, so `partials` holds grid + this many slots)
int hs_general_blocks(int n_gen);
// b = -F(z) (FP32), dinv = 1 / J_PP (FP32)
void hs_prepare(const HsDom& d, const double* z, long long n, float* dinv, float* b, int grid, hipStream_t s);
void hs_bicgstab_init(const BiArgs& a, int grid, hipStream_t s);
void hs_bicgstab_iteration(const BiArgs& a, int grid, hipStream_t s, class KernelTimer* timer);
// max |F(z + lam dz)| and max centred-gradient ratio over the DOFs
void hs_trial(const HsDom& d, const double* z, const float* dz, double lam, long long n, double* partials,
              unsigned* ticket, double* out2, int grid, hipStream_t s);
void hs_apply(const uint8_t* dof, double* z, const float* dz, double lam, long long n, int grid, hipStream_t s);

}  // namespace lsdtt_gpu
