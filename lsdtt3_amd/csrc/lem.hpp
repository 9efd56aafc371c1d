// lem.hpp -- the LEM timestep on the device, FP64 state.
//
// One step (src/cpu/lsdtt_lem_cpu.cpp:285-732, order is the contract):
//   1. D8 routing of the *input* surface (pits and interior nodata are
//      outlets), upstream cell counts -> area = count * dx^2
//   2. implicit stream power, level by level from the outlets up
//      (src/cpu/lsdtt_stream_power_cpu.cpp:159-326)
//   3. implicit hillslope diffusion (linear or Roering) by Newton with the
//      reference's line search and sub-stepping
//      (src/cpu/lsdtt_hillslope_newton_cpu.cpp:196-788); the Newton residual
//      and state are FP64, the linear solves run BiCGSTAB on an assembled
//      9-point Jacobian in FP32 (iterative refinement by the FP64 Newton loop)
//   4. uplift U*dt on valid cells, then boundary cells restored to their
//      pre-step value minus base-level fall
//   5. step statistics, p99.9 by exact radix selection
#pragma once

#include <memory>
#include <vector>

#include "common.hpp"
#include "dbuf.hpp"
#include "dinf.hpp"
#include "flow.hpp"
#include "fluvial.hpp"
#include "hillslope.hpp"
#include "ktimer.hpp"
#include "picard.hpp"
#include "warmstart.hpp"

namespace lsdtt_gpu {

struct LemParams {
    double dt = 100.0;
    double K = 1e-3, m = 0.45, n = 1.0, threshold = 0.0, slope_clamp = 1e-5;
    int divide_area_by_dx = 1;
    double D = 1e-3, Sc = 0.6, slope_ratio_cap = 0.9999;
    int nonlinear = 0;
    double U = 1e-4, baselevel_fall = 0.0;
    int enable_fluvial = 1, enable_hillslope = 0, enable_uplift = 1;
    int fluvial_dinf = 0;  // with enable_fluvial: D-infinity incision instead of D8
    BC bc{'o', 'o', 'o', 'o'};
    double newton_abs_tol = 1e-12, newton_rel_tol = 1e-8;
    int newton_max_iterations = 50, newton_max_line_search = 30;
    double krylov_rel_tol = 1e-4;
    int krylov_max_iterations = 500;
    int area_kernel = 1;        // drainage-area moving average (odd; 1 = none)
    double width_factor = 1.0;  // w / dx or 1
    int regime = 0;             // 0 automatic, 1 coupled, 2 channels_as_boundaries
    int force_picard = 0;       // hillslope_force_picard
    int erosion_rate_field = 0; // compute_erosion_rate_field
    int krylov_forcing = 1;     // 0 fixed krylov_rel_tol, 1 Eisenstat-Walker
    PicardOptions picard;       // fallback settings (D, Sc, nonlinear from above; U = 0)
};

struct LemStats {
    double delta_max_elevation = 0, delta_mean_elevation = 0, max_local_elevation_change = 0;
    double p999_local_elevation_change = 0, mean_erosion_rate = 0, max_erosion_rate = 0;
    long long n_valid = 0, cells_eroded = 0;
    int newton_iterations = 0, krylov_iterations = 0, substeps = 0, rejected = 0;
    int hillslope_converged = 1;
    double hillslope_initial_residual = 0, hillslope_residual = 0, hillslope_tolerance = 0;
    double max_slope_ratio = 0;
    int count_passes = 0, stream_power_passes = 0;
    long long count_tile_runs = 0, stream_power_tile_runs = 0;
    int hillslope_method = 0;  // 0 Newton, 1 Picard
    PicardStats picard;
    // warm start of the first Newton solve (0 none, 1 explicit Euler
    // predictor, 2 learned); accepted when it lowers ||F||inf below F(z^n)
    int warm_start_mode = 0, warm_start_accepted = 0;
    double warm_start_residual = 0;  // ||F||inf at the predicted start
    // Newton iterations whose line search ran in restoration mode (the
    // iterate's slope ratio above sqrt(cap), newton .cpp:553-633)
    int restoration_iterations = 0;
};

class LemEngine {
public:
    LemEngine(int nr, int nc, double dx, const uint8_t* mask_host, const LemParams& p);
    ~LemEngine();
    LemEngine(const LemEngine&) = delete;
    LemEngine& operator=(const LemEngine&) = delete;
    // explicit Dirichlet channel mask for the hillslope (null clears it)
    void set_channel_mask(const uint8_t* channel_host);

    void set_elevation(const double* z_host);
    void get_elevation(double* z_host);
    void set_erodibility(const double* K_host);  // null clears the field
    void set_boundary_flats(const long long* flats, long long n);
    // the state array: its address stays the same for the life of the
    // engine (every step writes its result into it)
    double* elevation_device() { return z.p; }

    // one timestep on stream `s`; returns 0, or LSDTT_GPU_ERR_NOT_CONVERGED
    // when the hillslope solve failed at the minimum sub-step (the state is
    // then the one after stream power + uplift, as the reference would have
    // to fall back to Picard)
    int step(LemStats* st, hipStream_t s);

    // Times `steps` steps (after one untimed step) with each tile height of
    // FluvialTiles::kTileHeights from the current surface, which is
    // restored after each trial, and keeps the fastest (the current height
    // unless another is more than 3% faster); returns it.
    // ms_per_step (optional, 3 entries) receives each height's time.  The
    // results of a step do not depend on the tile height.
    int autotune_tiles(int steps, double* ms_per_step, hipStream_t s);

    int nr, nc;
    long long n;
    double dx;
    LemParams p;
    FlowEngine flow;
    FluvialTiles tiles;
    int last_count_passes = 0, last_sp_passes = 0;
    bool counts_warm = false;  // flow.cnt holds the previous step's counts
    long long last_count_tiles = 0, last_sp_tiles = 0;
    long long last_eroded = 0;

    dbuf<uint8_t> mask, dof, fast;
    dbuf<uint8_t> cflag;  // per cell: bit 0 valid, bit 1 uplifted (valid and not a boundary flat)
    // channel set derived each step from the D-infinity drainage area of the
    // post-incision surface (lsdtt_lem_cpu.cpp:86-154): the DOF arrays
    // without channels, the receivers, the area and the set itself
    dbuf<uint8_t> dof_base, fast_base, chan;
    dbuf<int> gen_base, drec;
    dbuf<float> dw;
    dbuf<double> darea;
    int n_gen_base = 0;
    dbuf<unsigned long long> grad_keys;  // median interior gradient (timestep control)
    bool darea_warm = false;
    int last_dinf_passes = 0;
    dbuf<double> z, z_before, K_field, bel, area_sm;
    std::vector<uint8_t> mask_h, channel_h;  // host copies (DOF set rebuilds)
    long long n_mask = 0;                     // valid cells
    dbuf<int> bflats, gen;
    int n_gen = 0;
    long long n_bflats = 0;
    bool has_K = false;

    // hillslope work: three trial buffers; after a step z_hs0 holds the
    // post-incision surface (what the Picard fallback and the captured
    // training pairs start from)
    dbuf<double> zn_hs, z_trial, z_hs0;
    PicardSolver picard;
    // during the hillslope: z keeps the post-incision surface; zn is the
    // sub-step's start (backward Euler's z^n), zc the iterate, zt the next
    // trial point, each z or one of the three trial buffers (zt never z,
    // zn or zc), so no surface is copied while the solve runs
    double* zn = nullptr;
    double* zc = nullptr;
    double* zt = nullptr;
    double* trial_target() const;
    dbuf<float> dinv, b, x, r, pv, pv2, v, v2, t;  // BiCGSTAB (p and v in two buffers each)
    dbuf<float> c9g;  // general DOFs' Jacobian rows (9 x n_gen)
    dbuf<double> partials, scal;  // reduction partials; scalars
    hbuf<double> hpin;            // pinned host slots for read-backs
    dbuf<unsigned> tickets;
    dbuf<int> flags;
    dbuf<float> absdz;
    dbuf<unsigned> hist;
    // per-pixel erosion rate of the last step (p.erosion_rate_field)
    dbuf<float> ero_rate;
    bool ero_valid = false;
    // where the last hillslope solve left its result (z or a trial buffer)
    const double* hs_result = nullptr;

    int hs_grid = 1024;   // BiCGSTAB kernels and the step statistics
    int hs_tgrid = 1024;  // hillslope trial kernel
    // side stream for the trial's general-cell kernel (default priority)
    hipStream_t aux_stream = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // warm start of the first Newton solve of a step: 0 off, 1 explicit
    // Euler (dz = dt (-F(z^n))), 2 the learned predictor `ws`, 3 the
    // previous step's hillslope change (ws_dz, kept when that step's solve
    // converged with the same dt and the surface was not set since)
    int ws_mode = 0;
    WarmStartPredictor* ws = nullptr;
    dbuf<float> ws_dz;
    bool ws_dz_valid = false;
    double ws_dz_dt = 0.0;
    // training capture: after each step, cap_dz = hillslope change of z
    // (post-hillslope minus post-incision, FP32) and cap_kappa = dt D / h^2;
    // z_hs0 holds the post-incision surface
    bool capture = false, cap_valid = false;
    dbuf<float> cap_dz;
    double cap_kappa = 0.0;
    KernelTimer timer;
    double cur_inv_dt = 0.0;
    // BiCGSTAB iterations used by each Newton iteration of the last solve
    static constexpr int kKrylovHist = 8;
    int krylov_hist[kKrylovHist] = {2, 2, 2, 2, 2, 2, 2, 2};

private:
    void stream_power(hipStream_t s);
    void dinf_stream_power(const SpParams& a, hipStream_t s);
    int hillslope(LemStats* st, hipStream_t s);
    int newton_solve(double inv_dt, LemStats* st, hipStream_t s);
    int bicgstab(hipStream_t s, int newton_it, double eta);
    void trial(double lambda, double* fnorm, double* ratio, hipStream_t s);
    HsDom hs_domain(double inv_dt) const;
    // hillslope DOFs (valid, not on an 'o' edge, not a channel pixel of an
    // explicit mask outside the coupled regime), fast flags, general list
    void build_dofs();
    void derive_channels(hipStream_t s);

public:
    // the hillslope has a Dirichlet channel set (channels_as_boundaries)
    bool channels_active() const;
    // channels from the drainage-area threshold (no explicit mask, not coupled)
    bool threshold_channels() const;

private:
    void stats(LemStats* st, const double* zres, hipStream_t s);
    // first_pass_done: the first radix histogram is already in hist (the
    // statistics pass builds it)
    void enqueue_p999(long long nv, unsigned* sel, hipStream_t s, bool first_pass_done);
};

}  // namespace lsdtt_gpu
