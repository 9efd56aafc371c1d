// picard.hpp -- backward-Euler Roering hillslope transport by damped Picard
// iteration on the device: the reference's fallback when the Newton solve
// fails, or when it is forced (lsdtt_lem_options::hillslope_force_picard).
//
// Reference: src/cpu/lsdtt_hillslope_implicit_cpu.cpp:108-1010 (backward-
// Euler path).  Each Picard iteration lags the face diffusivities
// D_eff = D / (1 - min(r^2, r_max^2)) at the current iterate and solves the
// symmetric positive definite 5-point system [I - dt L(D_eff)] y = z_old +
// dt U (+ fixed neighbours), then z += omega (y - z).  The reference factors
// the system with a sparse Cholesky (Eigen SimplicialLDLT); here it is
// solved matrix-free by Jacobi-preconditioned conjugate gradients in FP64 to
// a relative residual of 1e-14 (the systems are diagonally dominant, so a
// handful of iterations).  The sub-step controller (semi-implicit fast
// path, accept on max face ratio < r_max, grow / shrink, damping, floor
// projection) runs on the host from device reductions.
#pragma once

#include "common.hpp"
#include "dbuf.hpp"

namespace lsdtt_gpu {

struct PicardOptions {
    double D = 1e-3, U = 0.0, Sc = 0.8;
    int nonlinear = 1;
    double picard_tol = 1e-4;   // max |dz| between iterates (m)
    int picard_max_iter = 50;
    double omega = 1.0;         // relaxation (<= 1)
    double r_max = 0.99;        // slope-ratio cap
    double dt_init = 0.0, dt_min = 0.0, dt_max = 0.0;  // 0 = auto
    double grow = 1.5, shrink = 0.5, dz_target = 0.0;
};

struct PicardStats {
    int substeps = 0, rejected = 0, picard_iters = 0, cg_iters = 0, limiter_fired = 0;
    double final_dt = 0.0, max_slope_ratio = 0.0;
};

class PicardSolver {
public:
    // z (device, n cells) advanced in place over dt_total; dof marks the
    // unknowns (valid, not an 'o' edge, not a channel pixel)
    void run(int nr, int nc, BC bc, double h, const uint8_t* mask, const uint8_t* dof, const PicardOptions& o,
             double dt_total, double* z, PicardStats* st, hipStream_t s);

private:
    int nr_ = 0, nc_ = 0;
    long long n_ = 0;
    BC bc_{'o', 'o', 'o', 'o'};
    double h_ = 1.0;
    const uint8_t* mask_ = nullptr;
    const uint8_t* dof_ = nullptr;
    PicardOptions o_;
    int grid_ = 1;
    dbuf<double> z_old, wE, wS, b, y, r, q, p, Ap, dinv, partials, scal, zp;
    dbuf<unsigned> ticket;
    dbuf<int> flags;
    hbuf<double> hpin;  // pinned read-back slots

    double max_ratio2(const double* z, hipStream_t s);
    // one backward-Euler stage: damped Picard on [I - beta L] y = z_old + beta U
    bool solve_stage(double beta, double omega, int iter_cap, double* z, int* iters, PicardStats* st,
                     hipStream_t s);
    int cg(hipStream_t s);
    double step_dz(const double* z, hipStream_t s);
    void projection(double* z, hipStream_t s);
    double read_scalar(int slot, hipStream_t s);
};

}  // namespace lsdtt_gpu
