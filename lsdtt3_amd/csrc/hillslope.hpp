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
    // p and v alternate between two buffers from one iteration to the next:
    // an iteration reads the previous p and v (p_old, v_old) around every
    // tile it stages while other tiles already write the new ones
    float *x, *r, *rhat, *p_old, *p, *v_old, *v, *t;
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

// ---------------------------------------------------------------------------
// Batched hillslope solves: the Newton-Krylov solves of many grids advance
// together, one launch per kernel for all of them.  Each grid's state
// machine lives on the device (HsCtl): a sub-step starts with a trial at
// lambda = 0 (F, the tolerance, the first linear system), BiCGSTAB
// iterations run until that grid's solve has converged, then line-search
// trials until one is accepted; a control kernel takes each grid's
// decisions (acceptance, restoration mode, convergence, sub-step growth,
// rejection) exactly as the single-grid host loop does.  A kernel skips the
// grids whose state does not call for it, so launches can be enqueued in a
// fixed pattern (and captured in a graph) while each grid follows its own
// path.  Reductions are per grid: per-tile partials folded by a per-grid
// finisher in tile order, so a grid's arithmetic does not depend on the
// other grids of the batch.
// ---------------------------------------------------------------------------
enum HsPhase : int {
    HS_IDLE = 0,    // no hillslope this step (or not started)
    HS_START = 1,   // needs the sub-step's first trial (lambda = 0)
    HS_KRYLOV = 2,  // BiCGSTAB iterations of the current Newton iteration
    HS_SEARCH = 3,  // line-search trial at ctl.lambda
    HS_DONE = 4,    // converged over the whole dt; result in buffer zc
    HS_FAILED = 5,  // Newton failed at the minimum sub-step (Picard fallback)
    HS_WARM = 6,    // trial of the predicted start z^n + (last step's hillslope change)
};

struct HsCtl {
    int phase;
    int zc, zn, zt;  // buffer ids: 0 = the post-incision surface, 1..3 trial buffers
    int it, ls, restoring;
    int substeps, rejected, newton_total, krylov_total, restoration_its;
    int kdone, kits, kbreak;  // BiCGSTAB: converged or stopped, iterations, breakdown
    int first_solve;          // the next START is the step's first solve
    int rounds;               // rounds in which the grid was still solving
    int stats_done;           // the step's statistics have been taken
    int method;               // 0 Newton, 1 damped Picard (after a failure)
    int zero_step;            // the line search's Newton step is zero (no Krylov iteration ran)
    double inv_dt, step, dt, remaining, dt_min, dt_max;
    double fnorm, ratio, tol, lambda, f_init;
    double trial_f, trial_ratio;
    double rho, rho_old, alpha, omega, bnorm2, rnorm2;
    double eta;  // relative tolerance of the current linear solve
    int ws_tried, ws_accepted;  // this step's warm-start trial ran / was kept
    double ws_f;                // ||F||inf at the predicted start
};

// Relative tolerance of the next Newton iteration's linear solve after one
// that took the residual from fprev to fnorm: Eisenstat-Walker choice 2
// (SIAM J. Sci. Comput. 17, 1996), eta = 0.9 (fnorm / fprev)^2, kept at
// least 0.9 eta_prev^2 when that exceeds 0.1, no tighter than half of what
// reaching tol from fnorm needs, within [1e-6, 0.1].
__host__ __device__ inline double forcing_eta(double eta_prev, double fnorm, double fprev, double tol)
{
    const double ratio = fnorm / fprev;
    double e = 0.9 * ratio * ratio;
    const double guard = 0.9 * eta_prev * eta_prev;
    if (guard > 0.1) e = e > guard ? e : guard;
    const double need = 0.5 * tol / fnorm;
    e = e > need ? e : need;
    e = e > 1e-6 ? e : 1e-6;
    return e < 0.1 ? e : 0.1;
}

// work item of the batched kernels: a tile of kBatchTileRows x
// kBatchTileCols cells
constexpr int kBatchTileRows = 64, kBatchTileCols = 64;

// one grid of a batched hillslope solve (cells of the concatenated arrays
// from `off`; its work-item tiles in slots stbase .. stbase + nst - 1;
// its general DOFs (grid edge, next to a masked cell) at gen_off in the
// concatenated lists, handled in chunks of 256 by slots gcbase ..)
struct BatchHsGrid {
    int enabled;  // the grid's step has a hillslope solve
    int nr, nc;
    BC bc;
    long long off;
    int stbase, nst, tx_n;
    int gen_off, n_gen, gcbase, ngc;
    double h, D, Sc, cap;
    int nonlinear;
    int ws_mode;       // 3: start each step from the last step's hillslope change
    double eta2;       // Krylov relative tolerance squared (the first Newton iteration's)
    int forcing;       // 1: later iterations by forcing_eta, 0: eta2 throughout
    int krylov_max;
    int newton_max, ls_max;
    double abs_tol, rel_tol, dt_total;
};

// work lists of the batched kernels (hillslope.hip): Krylov tiles and
// chunks, trial tiles and chunks
enum { WL_KTILE = 0, WL_KCHUNK = 1, WL_TTILE = 2, WL_TCHUNK = 3, WL_N = 4 };

struct BatchHs {
    int G;  // grids
    int* wl_inst;  // WL_N x G: the listed grids
    int* wl_pre;   // WL_N x (G + 1): their first item
    int* wl_hdr;   // WL_N x 2: grids listed, items
    const BatchHsGrid* grids;
    const int* st_grid;  // stencil-tile slot -> grid
    const int* gc_grid;  // general-DOF chunk slot -> grid
    int NT, NGC;         // stencil tiles, general-DOF chunks
    HsCtl* ctl;
    const uint8_t* mask;
    const uint8_t* dof;
    const uint8_t* fast;
    const int* gen;  // per grid: its general DOFs (cell indices within the grid), row-major
    float* c9g;      // per grid: 9 x n_gen Jacobian rows at 9 * gen_off, slot-major
    double* zbuf[4];
    // p and v each in two buffers: a grid's iteration k reads the pair
    // written at k - 1 (p2, v2 when k is odd) and writes the other
    float *b, *dinv, *x, *r, *p, *p2, *v, *v2, *t;
    double* partials;  // (NT + NGC) slots of up to 3 values
    const float* ws_dz;  // per cell: the last step's hillslope change (warm start)
    const int* ws_ok;    // per grid: ws_dz holds a converged step's change
};

// Batched launches over every grid of B (n_grids of them).
// hsb_start: each enabled grid enters START with z (buffer 0) as z^n and the
// whole dt as its first sub-step; the others IDLE.
void hsb_start(const BatchHs& B, int n_grids, hipStream_t s);
// One round: a BiCGSTAB iteration of the KRYLOV grids, the control step
// that sends finished solves to their line search, the trials of the START
// and SEARCH grids and the control step that judges them.
// one solver round; `tiles` / `chunks` (when >= 0) bound the work items
// listed in it (the kernels stride over their lists, so any launch size is
// correct; a bound only trims empty blocks)
void hsb_round(const BatchHs& B, int n_grids, hipStream_t s, int tiles = -1, int chunks = -1);

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
