// fluvial.hpp -- upstream counts and implicit stream power by tiles.
//
// Both sweeps follow the D8 receiver graph: counts from the headwaters down
// (count(v) = 1 + sum of its donors' counts), stream power from the outlets
// up (z(v) is a function of its receiver's new elevation,
// src/cpu/lsdtt_stream_power_cpu.cpp:218-319).  Instead of global depth
// levels (random gathers), the grid is cut into tiles of 32 columns by 8,
// 16 (default) or 32 rows held in LDS:
// inside a tile the cells are ordered by their hop count to the tile's exit
// (pointer jumping in LDS, once per step) and processed level by level by
// the tile's one wavefront; values that cross a tile edge are read from HBM.
// Passes repeat over the tiles that read a value another tile changed since
// (the tile of a changed cell's receiver for counts, of its donors for
// stream power) until a pass marks no tile.  Every cell's final value is the same function of
// its receiver's (or donors') final values as in the serial reference order,
// so the result does not depend on the pass schedule.
#pragma once

#include <vector>

#include "common.hpp"
#include "dbuf.hpp"
#include "ktimer.hpp"

namespace lsdtt_gpu {

struct SpParams {
    double K0, m, nexp, dt, dx, pixel_area, threshold, slope_clamp;
    int divide, n_is_one;
    int m_half, n1_half;  // m == 0.5, n - 1 == 0.5: sqrt in place of pow
    double width;         // w / dx or 1 (multiplies K)
};

// Pass bookkeeping of the tile sweeps.  Pass p reads act[p % 3], marks
// act[(p + 1) % 3] and clears act[(p + 2) % 3] for pass p + 1; it does
// nothing when pass p - 1 marked no tile (the sweep has settled).
struct PassArgs {
    const int* act;
    int* act_next;
    int* act_clear;
    const int* prev_marked;  // null for the first pass
    int* marked;
    int* runs;  // per tile: times run in this settle (one writer per pass)
    int first_filtered = 0;  // the first pass too runs only the tiles in act (else every tile)
};

// inputs and outputs of a stream-power sweep
struct SpTileArgs {
    const int* rcv;
    const uint8_t* dir;
    const uint8_t* mask;
    const int* cnt;
    const double* area;  // null: count * pixel_area
    const short* ord;
    const short* lstart;
    const int* tile_maxdep;
    const uint8_t* xdon;
    const double* z_old;
    const double* K;
    double* z;
    int* tile_eroded;
    SpParams p;
};

// exponent mode of a stream-power kernel: m == 0.5 (sqrt), and n == 1
// (closed form), n - 1 == 0.5 (sqrt) or general (pow)
__host__ __device__ constexpr int sp_mode_of(bool m_half, int n_mode) { return (m_half ? 3 : 0) + n_mode; }

// A batch of grids swept together: each grid's per-cell arrays are one
// slice of concatenated arrays (receivers hold indices within the grid),
// each of its 32-column tiles one slot of the per-tile arrays.
struct BatchGrid {
    int nr, nc;
    BC bc;
    int ntx;        // tile columns
    int tbase;      // slot of the grid's first routing tile
    long long off;  // first cell in the concatenated arrays
    SpParams sp;    // stream power (K as a scalar, no area field)
    int sp_mode;    // sp_mode_of for sp
};
struct BatchTile {
    int inst, tx, ty;
};
struct BatchTiles {
    const BatchGrid* grids;
    const BatchTile* tiles;  // slot -> grid and position
};

void batch_tile_order(int th, const BatchTiles& B, int ntiles, const int* rcv, const uint8_t* mask, short* ord,
                      short* lstart, int* maxdep, uint8_t* xdon, hipStream_t s);
void batch_tile_counts(int th, const BatchTiles& B, int ntiles, const int* rcv, const uint8_t* mask,
                       const short* ord, const short* lstart, const int* maxdep, const uint8_t* xdon, int* cnt,
                       const PassArgs& P, hipStream_t s);
// one pass of every exponent mode in `modes` (bit sp_mode_of)
void batch_tile_sp(int th, unsigned modes, const BatchTiles& B, int ntiles, const SpTileArgs& A, const PassArgs& P,
                   hipStream_t s);

class FluvialTiles {
public:
    static constexpr int TW = 32;
    static constexpr int kTileHeights[3] = {8, 16, 32};
    static bool valid_tile_height(int th);

    void resize(int nr, int nc, BC bc);
    // tile rows (one of kTileHeights); re-sizes the tile buffers.  Results
    // do not depend on it, only the speed does.  The next sweep must be
    // preceded by depths().
    void set_tile_height(int th);
    int tile_height() const { return th; }
    // each tile's cells in level order, once per set of receivers; with
    // only_dirty, only the tiles flagged in `dirty` (the receivers of the
    // others are those of the last depths(), which needs d8_order_valid())
    void depths(const int* rcv, const uint8_t* mask, hipStream_t s, bool only_dirty = false);
    // counts (int32) of valid cells from receivers; warm: cnt holds counts
    // of earlier receivers on the same mask (the starting values); with
    // only_dirty the first pass runs only the flagged tiles (cnt must hold
    // the counts of the last receivers, which changed only there); returns
    // passes used
    int counts(const int* rcv, const uint8_t* mask, int* cnt, bool warm, hipStream_t s, bool only_dirty = false);
    // the D8 level order of every tile is that of the receivers last given
    // to depths() (the D-infinity orders share its buffers)
    bool d8_order_valid() const { return d8_valid; }
    // per tile: a receiver in it, or one leading into it, changed since the
    // last depths() (set by the receiver kernels, cleared by clear_dirty)
    dbuf<int> dirty;
    void clear_dirty(hipStream_t s);
    // stream power in place on z (z == z_old on entry); returns passes used
    // (area: per-cell drainage area in m^2, or null for count * dx^2)
    int stream_power(const int* rcv, const uint8_t* dir, const uint8_t* mask, const int* cnt, const double* area,
                     const double* z_old, const double* K, const SpParams& sp, double* z, long long* eroded,
                     hipStream_t s);
    // D-infinity accumulation (rec / w: two receivers and weights per cell,
    // -1 = none): dinf_order once per graph, then dinf_accumulate (acc =
    // cell_area per valid cell plus its donors' weighted values; warm: acc
    // holds any earlier values of the same mask, else it is cleared);
    // returns passes used.  The order shares the buffers of depths(), which
    // the D8 sweeps need again afterwards.
    void dinf_order(const int* rec, const uint8_t* mask, hipStream_t s);
    int dinf_accumulate(const int* rec, const float* w, const uint8_t* mask, double cell_area, double* acc,
                        bool warm, hipStream_t s);
    // D-infinity stream power in place on z (z == z_old on entry; area in
    // m^2 per cell): dinf_up_order once per graph (shares the buffers of
    // depths()), then dinf_stream_power; returns passes used
    void dinf_up_order(const int* rec, const uint8_t* mask, hipStream_t s);
    int dinf_stream_power(const int* rec, const float* w, const uint8_t* mask, const double* area,
                          const double* z_old, const double* K, const SpParams& sp, double* z, long long* eroded,
                          hipStream_t s);
    // steady-state chi of the D-infinity graph after dinf_up_order (amn:
    // powf(A, -m/n) times the chi weight per valid cell; chi float[n],
    // 0 where no receiver leads out); returns passes used
    int dinf_chi(const int* rec, const float* w, const uint8_t* mask, const float* amn, float dx, float* chi,
                 hipStream_t s);
    // smooth_drainage_area of count * pixel_area: rolling row sums, then
    // rolling column sums, in the reference's order (one thread per row,
    // then per column), so the averages are bit-identical
    void smooth_area(const int* cnt, const uint8_t* mask, double pixel_area, int kernel_size, double* out,
                     hipStream_t s);
    // the same of areas given in m^2
    void smooth_area(const double* area, const uint8_t* mask, int kernel_size, double* out, hipStream_t s);

    int nr = 0, nc = 0, ntx = 0, nty = 0, ntiles = 0, max_passes = 0;
    long long tile_runs = 0;  // tiles processed by the last counts / stream_power
    BC bc_{'o', 'o', 'o', 'o'};
    dbuf<int> act[3], flags, tile_eroded, tile_maxdep, tile_run;
    hbuf<int> flags_h;
    KernelTimer* timer = nullptr;  // optional: each pass timed as one launch
    dbuf<short> ord, lstart;  // tile-major level order
    dbuf<uint8_t> xdon;       // per cell: directions of donors in other tiles
    dbuf<double> hsum;        // smoothing: row sums
    dbuf<int> hcnt;           // smoothing: row counts

private:
    int th = 16;
    bool d8_valid = false;
    // passes of the last call of each sweep: the next call's first batch
    int last_passes = 0, last_count_passes = 0, last_sp_passes = 0, last_dinf_passes = 0, last_chi_passes = 0;
    // runs `pass(PassArgs)` until a pass marks no tile
    // start: the tiles active in the first pass (null: all)
    template <class Pass, class After>
    int settle(const char* name, double first_pass_bytes, Pass&& pass, After&& after_batch, hipStream_t s,
               const int* start = nullptr);
    template <class T>
    void smooth(const T* v, const uint8_t* mask, double pa, int kernel_size, double* out, hipStream_t s);
};

}  // namespace lsdtt_gpu
