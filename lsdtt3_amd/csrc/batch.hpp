// batch.hpp -- many independent LEM timesteps advanced together on one GPU.
//
// An ensemble of grids of any sizes and parameters (a parameter sweep) is
// laid out as one set of concatenated device arrays: grid i's cells are the
// slice [off_i, off_i + n_i) of every per-cell array, its receivers hold
// cell indices within the grid.  Every kernel of the step is launched once
// for the whole batch over a list of (grid, tile) work items, so a batch of
// small grids fills the GPU like one large grid:
//   receivers (D8 of the pre-step surface, which is copied as it is read),
//   per-tile level order, tile passes of the upstream counts and of the
//   stream power (lsdtt_stream_power_cpu.cpp:159-326), the hillslope
//   Newton-Krylov solve (lsdtt_hillslope_newton_cpu.cpp:196-788) with every
//   decision taken on the device by each grid's state machine (hillslope.hpp
//   HsCtl), uplift, boundary restore and the step statistics with the exact
//   p99.9 per grid (lsdtt_lem_cpu.cpp:614-722).
// A step is a fixed pattern of launches (pass counts and solver rounds from
// the previous steps plus slack); grids that finish early skip the rest, so
// the pattern is captured once into a hipGraph and replayed each step.  One
// read-back per step checks that every sweep settled and every grid
// finished; a grid that needs more rounds (or passes) continues with
// launches outside the graph, and the caps grow for the next capture.
// A grid's arithmetic does not depend on the other grids of the batch
// (per-grid reductions fold per-tile partials in tile order), so a grid run
// alone in a batch of one gives bit-identical results.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "common.hpp"
#include "dbuf.hpp"
#include "fluvial.hpp"
#include "hillslope.hpp"
#include "lem.hpp"
#include "picard.hpp"

namespace lsdtt_gpu {

// per grid: what the step's statistics need
struct BatchStepGrid {
    double udz, bdz, inv_dt;  // uplift and base-level fall over the step, 1 / dt
    unsigned p999_k;          // rank of the p99.9 among the grid's valid cells
    int rtbase, nrt;          // routing tiles
    int erosion_field;        // keep the per-pixel erosion rate
};

struct BatchSpec {
    int nr = 0, nc = 0;
    double dx = 1.0;
    const uint8_t* mask = nullptr;  // host, nr * nc
    LemParams p;
};

class BatchEngine {
public:
    explicit BatchEngine(const std::vector<BatchSpec>& specs);
    ~BatchEngine();
    BatchEngine(const BatchEngine&) = delete;
    BatchEngine& operator=(const BatchEngine&) = delete;

    int size() const { return static_cast<int>(grids_h.size()); }
    long long cells(int i) const { return (long long)grids_h[i].nr * grids_h[i].nc; }
    long long valid_cells(int i) const { return n_valid_h[i]; }
    void set_elevation(int i, const double* z_host);
    void get_elevation(int i, double* z_host);
    void set_boundary_flats(int i, const long long* flats, long long n);
    // the last step's per-pixel erosion rate of grid i (needs
    // compute_erosion_rate_field in its options)
    void get_erosion_rate(int i, float* out);

    // one step of every grid; st receives one entry per grid.  Returns the
    // number of grids whose hillslope Newton failed at its minimum sub-step
    // (they took the damped Picard fallback, as run_lem_step does).
    int step(std::vector<LemStats>& st);

    // warm start of each grid's first Newton solve: 0 off, 3 the grid's
    // previous step's hillslope change (kept when that solve converged by
    // Newton; set_elevation, a restarted step and autotune_tiles forget it)
    void set_warm_start(int mode);
    int warm_start() const { return ws_mode; }

    // graph replay of the step pattern (default on); off: the same launches
    // enqueued directly
    void set_graph(bool on);
    // routing tile rows (8, 16 or 32; results do not depend on it)
    void set_tile_height(int th);
    int tile_height() const { return th; }
    // times `steps` steps per tile height from the current surfaces (restored
    // after each trial) and keeps the fastest; ms_per_step[3] optional
    int autotune_tiles(int steps, double* ms_per_step);
    // the engine's stream (non-blocking, default priority): every launch of
    // the step goes there, so it can be captured
    hipStream_t stream() const { return s_; }

    struct Counters {
        long long graph_launches = 0, graph_captures = 0, eager_rounds = 0, eager_passes = 0, refills = 0;
        int count_passes = 0, sp_passes = 0, rounds = 0;  // caps of the current pattern
    } counters;

private:
    struct Plan {
        int count_passes = 4, sp_passes = 4, rounds = 8;
        bool operator==(const Plan& o) const
        {
            return count_passes == o.count_passes && sp_passes == o.sp_passes && rounds == o.rounds;
        }
    };

    // tables (host copies and device)
    std::vector<BatchGrid> grids_h;        // routing view of each grid
    std::vector<BatchHsGrid> hsgrids_h;    // hillslope view
    std::vector<LemParams> params;
    std::vector<long long> n_valid_h;
    std::vector<uint8_t> mask_all;         // host copy of the concatenated masks
    dbuf<BatchGrid> grids_d;
    dbuf<BatchTile> rtiles_d;
    dbuf<BatchHsGrid> hsgrids_d;
    dbuf<int> st_grid_d, gc_grid_d;
    dbuf<HsCtl> ctl;
    dbuf<int> wl_inst, wl_pre, wl_hdr;  // work lists of the batched hillslope kernels
    int G = 0, RT = 0, NT = 0, NGC = 0, th = 16;
    long long N = 0;  // cells of the concatenated arrays
    unsigned sp_modes = 0;
    bool any_hillslope = false;
    int ws_mode = 0;

    // per-cell arrays (concatenated)
    dbuf<double> z, z_before, zA, zB, zC;
    dbuf<uint8_t> mask, dof, fast, cflag, dir, xdon;
    dbuf<int> rcv, cnt, gen;
    dbuf<float> c9g, b, dinv, x, r, pv, pv2, v, v2, t, absdz, ero_rate;
    dbuf<float> ws_dz;  // warm start: the last step's hillslope change (allocated with the mode)
    dbuf<int> ws_ok;    // per grid: ws_dz is valid
    dbuf<BatchStepGrid> step_grids;
    // per routing tile
    dbuf<short> ord, lstart;
    dbuf<int> maxdep, act[3], runs, eroded;
    dbuf<int> pflags;  // pass flags: [0, kMaxPass) counts, [kMaxPass, 2 kMaxPass) stream power
    // reductions and statistics
    dbuf<double> partials, spartials, stats_out;  // stats_out: G x 8
    dbuf<long long> eroded_out;                   // G
    dbuf<unsigned> hist, sel;                     // G x 2048, G x 4
    dbuf<long long> bflat;                        // concatenated boundary flats (cells of the arrays)
    dbuf<int> bflat_grid;
    dbuf<double> bel;
    std::vector<std::vector<long long>> bflats_h;
    long long n_bflats = 0;

    // host read-back
    hbuf<HsCtl> ctl_h;
    hbuf<double> stats_h;
    hbuf<long long> eroded_h;
    hbuf<unsigned> sel_h;
    hbuf<int> pflags_h;

    // graph
    hipStream_t s_ = nullptr;
    bool use_graph = true;
    hipGraphExec_t exec = nullptr;
    Plan captured, plan;
    PicardSolver picard;

    static constexpr int kMaxPass = 4096;
    // caps of the captured launch pattern, so that its graph stays a few
    // thousand nodes (a solver round is 14 launches): sweeps that need more
    // passes restart eagerly, solver rounds beyond the cap run eagerly after
    // the replay
    static constexpr int kMaxPatternPasses = 64, kMaxPatternRounds = 160;

    void build_layout(const std::vector<BatchSpec>& specs);
    void rebuild_routing_tables();
    BatchTiles btiles() const { return BatchTiles{grids_d.p, rtiles_d.p}; }
    BatchHs bhs();
    // the step's launches with the given caps
    void enqueue_step(const Plan& pl, hipStream_t s);
    void enqueue_counts(int first_pass, int passes, hipStream_t s);
    void enqueue_sp(int first_pass, int passes, hipStream_t s);
    void enqueue_stats(hipStream_t s);
    void read_status(hipStream_t s);
    void finish_step(std::vector<LemStats>& st, hipStream_t s, int* picard_count);
    void release_graph();
};

}  // namespace lsdtt_gpu
