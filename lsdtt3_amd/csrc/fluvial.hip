// fluvial.hip -- tile-local upstream counts and implicit stream power (see
// fluvial.hpp).  Reference: src/cpu/lsdtt_flowinfo_cpu.cpp:1979-2046 (counts
// = accumulate with unit weights), src/cpu/lsdtt_stream_power_cpu.cpp:
// 159-326 (stream power), applied in an order that respects every
// receiver-before-donor (stream power) or donor-before-receiver (counts)
// dependency, which is all the serial reference order guarantees.
//
// One wavefront owns one 32 x 32 tile.  Loads and stores touch cells
// i = lane + 64 k (k = 0..15): lane l covers column l % 32 of tile rows
// 2k + l / 32, so every HBM access is a coalesced 32-cell row segment, and
// each lane issues all of its loads before it uses any of them.  Inside the
// tile the inputs are staged in LDS and each level's cells are spread over
// the 64 lanes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <type_traits>
#include <utility>

#include "fluvial.hpp"

namespace lsdtt_gpu {

namespace {

// Tiles are TW = 32 columns by TH rows, TH one of FluvialTiles::kTileHeights
// (a template parameter of every tile kernel); a tile's TN = TW * TH cells
// are spread over one wavefront, PER = TN / WL per lane.
constexpr int TW = FluvialTiles::TW;
constexpr int WL = 64;  // lanes: one wavefront per tile

template <int TH>
struct TileShape {
    static constexpr int TN = TW * TH;
    static constexpr int PER = TN / WL;
    // pointer-jumping rounds: 2^rounds >= TN bounds any in-tile path
    static constexpr int kJumpRounds = TN <= 256 ? 8 : (TN <= 512 ? 9 : 10);
    static_assert(TW == 32 && TN % WL == 0 && TH >= 2, "lane layout assumes 32-wide tiles of even height");
    static_assert(TH <= 32, "ring slots cover at most 32 rows");
    static_assert((1 << kJumpRounds) >= TN, "pointer jumping rounds too few");
};

// D8 offsets of common.hpp's direction order (N, NE, E, SE, S, SW, W, NW),
// two bits per direction, biased by one
constexpr unsigned kRowBits = 0x1A90u, kColBits = 0x01A9u;

template <int TH>
__device__ __forceinline__ int local_index(int g, int nc, int r0, int c0)
{
    const int gr = g / nc, gc = g - gr * nc;
    return (gr >= r0 && gr < r0 + TH && gc >= c0 && gc < c0 + TW) ? (gr - r0) * TW + (gc - c0) : -1;
}

// local index of the receiver g of this lane's cell k (-1 outside the
// tile), from its D8 code d: the geometric neighbour when g is that cell
// (no division), else (a step across a periodic edge) from g itself
template <int TH>
__device__ __forceinline__ int local_receiver(int g, int d, int lane, int k, int nc, int r0, int c0)
{
    const int lr = 2 * k + (lane >> 5), lc = lane & 31;
    const int rr = lr + static_cast<int>((kRowBits >> (2 * (d & 7))) & 3u) - 1;
    const int cc = lc + static_cast<int>((kColBits >> (2 * (d & 7))) & 3u) - 1;
    if ((r0 + rr) * nc + (c0 + cc) == g)
        return (rr >= 0 && rr < TH && cc >= 0 && cc < TW) ? rr * TW + cc : -1;
    return local_index<TH>(g, nc, r0, c0);
}

template <int TH>
__device__ __forceinline__ int tile_of(int g, int nc, int ntx)
{
    const int gr = g / nc;
    return (gr / TH) * ntx + (g - gr * nc) / TW;
}

// neighbour of cell (r, c) in D8 direction d under the boundary
// conditions, -1 if there is none
__device__ __forceinline__ int neighbour(int r, int c, int d, int nr, int nc, BC bc)
{
    const int rr = r + static_cast<int>((kRowBits >> (2 * d)) & 3u) - 1;
    const int cc = c + static_cast<int>((kColBits >> (2 * d)) & 3u) - 1;
    if (rr >= 0 && rr < nr && cc >= 0 && cc < nc) return rr * nc + cc;
    const Step st = resolve_step(r, c, d, nr, nc, bc);
    return st.kind == 1 ? st.r * nc + st.c : -1;
}

__device__ __forceinline__ int wave_max(int v)
{
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
    return v;
}

__device__ __forceinline__ int wave_exclusive_scan(int v, int* total)
{
    const int lane = threadIdx.x & 63;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    *total = __shfl(x, 63);
    return x - v;
}

// this lane's cell k of the tile at (r0, c0): flat index (0 outside the
// grid, so that loads stay in bounds) and whether it is in the grid
struct LaneCell {
    int f, r, c;
    bool in;
};

__device__ __forceinline__ LaneCell lane_cell(int k, int r0, int c0, int nr, int nc)
{
    const int lane = threadIdx.x;
    LaneCell q;
    q.r = r0 + 2 * k + (lane >> 5);
    q.c = c0 + (lane & 31);
    q.in = q.r < nr && q.c < nc;
    q.f = q.in ? q.r * nc + q.c : 0;
    return q;
}

// Where a tile sits: its slot in the per-tile arrays, the slot of its
// grid's first tile, and its column and row among that grid's tiles.  A
// single-grid launch takes them from the block index; a batched launch
// (many grids at once, their per-cell arrays concatenated) from its tile
// list.
struct TileAt {
    int tile, tbase, tx, ty, ntx;
};

__device__ __forceinline__ TileAt grid_tile()
{
    // XCD by XCD in contiguous bands of tiles (the hardware deals
    // consecutive workgroups to the XCDs in turn): the halos and cross-tile
    // neighbours a band reads stay in its XCD's L2
    const int ntx = static_cast<int>(gridDim.x);
    const int t = xcd_tile(static_cast<int>(blockIdx.y * gridDim.x + blockIdx.x), ntx * static_cast<int>(gridDim.y));
    return TileAt{t, 0, t % ntx, t / ntx, ntx};
}

__device__ __forceinline__ TileAt batch_tile(const BatchTiles& B, BatchGrid& g)
{
    const BatchTile t = B.tiles[blockIdx.x];
    g = B.grids[t.inst];
    return TileAt{static_cast<int>(blockIdx.x), g.tbase, t.tx, t.ty, g.ntx};
}

// ---------------------------------------------------------------------------
// once per set of receivers: each tile's cells sorted by in-tile hop count
// to the tile's exit (pointer jumping in LDS, then a counting sort), stored
// tile-major: ord[t*TN + p] = local cell, lstart[t*(TN+1) + L] = first p of
// level L.  Also, per cell, the directions of its valid donors that lie in
// another tile (only cells on the ring of the tile's part of the grid can
// have them; a periodic wrap on a narrow grid can reach one cell by two
// directions, which then counts once).
// ---------------------------------------------------------------------------
template <int TH>
__device__ __forceinline__ void tile_order_body(const TileAt& T, const int* __restrict__ rcv,
                                                const uint8_t* __restrict__ mask, int nr, int nc, BC bc,
                                                short* __restrict__ ordg, short* __restrict__ lstartg,
                                                int* __restrict__ tile_maxdep, uint8_t* __restrict__ xdon)
{
    constexpr int TN = TileShape<TH>::TN, PER = TileShape<TH>::PER;
    __shared__ int jmp[TN];
    __shared__ int dep[TN];
    __shared__ int hist[TN + 1];
    __shared__ uint8_t xd[TN];
    const int tx = T.tx, ty = T.ty, lane = threadIdx.x;
    const int tile = T.tile;
    const int r0 = ty * TH, c0 = tx * TW;
    const int h = min(TH, nr - r0), w = min(TW, nc - c0);
    bool valid[PER];
    {
        uint8_t m[PER];
        int g[PER];
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const LaneCell q = lane_cell(k, r0, c0, nr, nc);
            m[k] = mask[q.f];
            g[k] = rcv[q.f];
            valid[k] = q.in && m[k];
        }
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int i = lane + WL * k;
            const int l = (valid[k] && g[k] >= 0) ? local_index<TH>(g[k], nc, r0, c0) : -1;
            jmp[i] = l;
            dep[i] = l >= 0 ? 1 : 0;
            hist[i] = 0;
            xd[i] = 0;
        }
    }
    if (lane == 0) hist[TN] = 0;
    __syncthreads();
    // ring slots: 0..31 top row, 32..63 bottom row, 64..95 left column,
    // 96..127 right column (corners belong to the rows)
#pragma unroll
    for (int half = 0; half < 2; ++half) {
        const int s = lane + WL * half;
        int lr, lc;
        bool ok;
        if (s < 32) {
            lr = 0, lc = s, ok = lc < w;
        } else if (s < 64) {
            lr = h - 1, lc = s - 32, ok = lc < w && h > 1;
        } else if (s < 96) {
            lr = s - 64, lc = 0, ok = lr >= 1 && lr <= h - 2;
        } else {
            lr = s - 96, lc = w - 1, ok = lr >= 1 && lr <= h - 2 && w > 1;
        }
        const int r = r0 + lr, c = c0 + lc;
        const int f = ok ? r * nc + c : 0;
        int gs[8], rv[8];
#pragma unroll
        for (int d = 0; d < 8; ++d) {
            int g = ok ? neighbour(r, c, d, nr, nc, bc) : -1;
            if (g == f || (g >= 0 && local_index<TH>(g, nc, r0, c0) >= 0)) g = -1;
            gs[d] = g;
        }
#pragma unroll
        for (int d = 0; d < 8; ++d) rv[d] = rcv[gs[d] >= 0 ? gs[d] : 0];
        unsigned bits = 0;
#pragma unroll
        for (int d = 0; d < 8; ++d) {
            bool dup = false;
#pragma unroll
            for (int e = 0; e < d; ++e) dup |= gs[e] == gs[d];
            // masked cells hold a negative receiver, never == f
            if (gs[d] >= 0 && !dup && rv[d] == f) bits |= 1u << d;
        }
        if (ok && mask[f]) xd[lr * TW + lc] = static_cast<uint8_t>(bits);
    }
    // pointer jumping; a round after which no pointer is left stops it
    // early (the tile is one wavefront, so the ballot is the tile's)
    for (int round = 0; round < TileShape<TH>::kJumpRounds; ++round) {
        int nj[PER], nd[PER];
        bool live = false;
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int i = lane + WL * k;
            const int j = jmp[i];
            nd[k] = dep[i];
            nj[k] = j;
            if (j >= 0) {
                nd[k] += dep[j];
                nj[k] = jmp[j];
            }
            live |= nj[k] >= 0;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int i = lane + WL * k;
            jmp[i] = nj[k];
            dep[i] = nd[k];
        }
        __syncthreads();
        if (__ballot(live) == 0) break;
    }
    int mx = 0;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int i = lane + WL * k;
        const LaneCell q = lane_cell(k, r0, c0, nr, nc);
        if (q.in) xdon[q.f] = xd[i];
        if (valid[k]) {
            const int d = dep[i];
            atomicAdd(&hist[d], 1);
            mx = max(mx, d);
        }
    }
    mx = wave_max(mx);
    __syncthreads();
    // exclusive scan of hist[0..TN), PER consecutive bins per lane
    int loc[PER], sum = 0;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        loc[k] = hist[lane * PER + k];
        sum += loc[k];
    }
    int total;
    int base = wave_exclusive_scan(sum, &total);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        hist[lane * PER + k] = base;
        lstartg[(size_t)tile * (TN + 1) + lane * PER + k] = static_cast<short>(base);
        base += loc[k];
        jmp[lane * PER + k] = 0;  // per-level cursors
    }
    if (lane == 0) {
        hist[TN] = total;
        lstartg[(size_t)tile * (TN + 1) + TN] = static_cast<short>(total);
        tile_maxdep[tile] = mx;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < PER; ++k)
        if (valid[k]) {
            const int i = lane + WL * k;
            const int d = dep[i];
            ordg[(size_t)tile * TN + hist[d] + atomicAdd(&jmp[d], 1)] = static_cast<short>(i);
        }
}

// dirty: only the flagged tiles (null: all)
template <int TH>
__global__ void __launch_bounds__(WL) k_tile_order(const int* __restrict__ rcv, const uint8_t* __restrict__ mask,
                                                   int nr, int nc, BC bc, short* __restrict__ ordg,
                                                   short* __restrict__ lstartg, int* __restrict__ tile_maxdep,
                                                   uint8_t* __restrict__ xdon, const int* __restrict__ dirty)
{
    const TileAt T = grid_tile();
    if (dirty && !dirty[T.tile]) return;
    tile_order_body<TH>(T, rcv, mask, nr, nc, bc, ordg, lstartg, tile_maxdep, xdon);
}

template <int TH>
__global__ void __launch_bounds__(WL) kb_tile_order(BatchTiles B, const int* __restrict__ rcv,
                                                    const uint8_t* __restrict__ mask, short* __restrict__ ordg,
                                                    short* __restrict__ lstartg, int* __restrict__ tile_maxdep,
                                                    uint8_t* __restrict__ xdon)
{
    BatchGrid g;
    const TileAt T = batch_tile(B, g);
    tile_order_body<TH>(T, rcv + g.off, mask + g.off, g.nr, g.nc, g.bc, ordg, lstartg, tile_maxdep, xdon + g.off);
}

// Pass bookkeeping shared by both sweeps.  Pass p reads act[p % 3], marks
// act[(p + 1) % 3] and clears act[(p + 2) % 3] for pass p + 1; it does
// nothing when pass p - 1 marked no tile (the sweep has settled).

// level starts of levels 0..maxdep + 1 into LDS
template <int TH>
__device__ __forceinline__ int load_levels(const short* __restrict__ lstartg, const int* __restrict__ tile_maxdep,
                                           int tile, short* ls)
{
    constexpr int TN = TileShape<TH>::TN;
    const int maxdep = tile_maxdep[tile];
    for (int L = threadIdx.x; L <= maxdep + 1; L += WL) ls[L] = lstartg[(size_t)tile * (TN + 1) + L];
    return maxdep;
}

// a tile's level starts held in registers: lane l keeps entries l, l + 64,
// ... of the tile's lstart row, and a level's start is read back with a
// readlane (the level index is the same in every lane).  Walking the levels
// then waits on no memory access: reading them from global memory one level
// ahead left a cache round trip per level exposed whenever a level had
// little work.
template <int TH>
struct LevelStarts {
    static constexpr int TN = TileShape<TH>::TN;
    static constexpr int NJ = (TN + 1 + WL - 1) / WL;
    int v[NJ];
    __device__ __forceinline__ void load(const short* __restrict__ row)
    {
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
            const int idx = static_cast<int>(threadIdx.x) + WL * j;
            v[j] = idx <= TN ? row[idx] : 0;
        }
    }
    // L: 0..TN, uniform across the wavefront
    __device__ __forceinline__ int at(int L) const
    {
        const int j = L >> 6, l = L & 63;
        int r = 0;
#pragma unroll
        for (int jj = 0; jj < NJ; ++jj)
            if (jj == j) r = __builtin_amdgcn_readlane(v[jj], l);
        return r;
    }
};

// ---------------------------------------------------------------------------
// counts: deep levels first; count = 1 + in-tile donors (LDS) + donors in
// other tiles (HBM, their current values)
// ---------------------------------------------------------------------------
template <int TH>
__device__ __forceinline__ void tile_counts_body(const TileAt& T, const int* __restrict__ rcv,
                                                 const uint8_t* __restrict__ mask, const short* __restrict__ ordg,
                                                 const short* __restrict__ lstartg,
                                                 const int* __restrict__ tile_maxdep,
                                                 const uint8_t* __restrict__ xdon, int nr, int nc, BC bc,
                                                 int* __restrict__ cnt, const PassArgs& P)
{
    constexpr int TN = TileShape<TH>::TN, PER = TileShape<TH>::PER;
    const int tx = T.tx, ty = T.ty, ntx = T.ntx;
    const int tile = T.tile;
    const int lane = threadIdx.x;
    if (lane == 0) P.act_clear[tile] = 0;
    if ((P.prev_marked || P.first_filtered) && !P.act[tile]) return;  // the first pass may run every tile
    __shared__ int acc[TN];
    __shared__ short lrcv[TN];
    __shared__ short ord[TN];
    __shared__ short ls[TN + 1];
    const int r0 = ty * TH, c0 = tx * TW;
    int old[PER], g[PER];
    bool valid[PER];
    short ordv[PER];
    {
        // every load is issued before any is used (a use of a later load
        // waits for all earlier ones)
        uint8_t m[PER];
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const LaneCell q = lane_cell(k, r0, c0, nr, nc);
            ordv[k] = ordg[(size_t)tile * TN + lane + WL * k];
            m[k] = mask[q.f];
            g[k] = rcv[q.f];
            old[k] = cnt[q.f];  // only this tile writes its cells
            valid[k] = q.in;
        }
#pragma unroll
        for (int k = 0; k < PER; ++k) valid[k] = valid[k] && m[k];
    }
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int i = lane + WL * k;
        acc[i] = valid[k] ? 1 : 0;
        lrcv[i] = static_cast<short>((valid[k] && g[k] >= 0) ? local_index<TH>(g[k], nc, r0, c0) : -1);
        ord[i] = ordv[k];
    }
    const int maxdep = load_levels<TH>(lstartg, tile_maxdep, tile, ls);
    __syncthreads();
    // inflow from other tiles into the ring cells: ring slots as in
    // k_tile_order (two per lane); every donor's count is loaded at once
    // (addresses of all 8 neighbours, clamped to the cell itself when off
    // the grid) and selected by the cross-tile donor bits afterwards
    {
        const int h = min(TH, nr - r0), w = min(TW, nc - c0);
#pragma unroll
        for (int half = 0; half < 2; ++half) {
            const int sl = lane + WL * half;
            int lr, lc;
            bool ok;
            if (sl < 32) {
                lr = 0, lc = sl, ok = lc < w;
            } else if (sl < 64) {
                lr = h - 1, lc = sl - 32, ok = lc < w && h > 1;
            } else if (sl < 96) {
                lr = sl - 64, lc = 0, ok = lr >= 1 && lr <= h - 2;
            } else {
                lr = sl - 96, lc = w - 1, ok = lr >= 1 && lr <= h - 2 && w > 1;
            }
            const int r = r0 + lr, c = c0 + lc;
            const int f = ok ? r * nc + c : 0;
            const uint8_t bits = xdon[f];
            int v[8];
#pragma unroll
            for (int d = 0; d < 8; ++d) {
                const int g2 = ok ? neighbour(r, c, d, nr, nc, bc) : -1;
                v[d] = __hip_atomic_load(&cnt[g2 >= 0 ? g2 : f], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            int a = 0;
#pragma unroll
            for (int d = 0; d < 8; ++d) a += ((bits >> d) & 1u) ? v[d] : 0;
            if (ok && a) acc[lr * TW + lc] += a;
        }
    }
    __syncthreads();
    for (int L = maxdep; L >= 0; --L) {
        const int b = ls[L], e = ls[L + 1];
        for (int p = b + lane; p < e; p += WL) {
            const int i = ord[p];
            const int l = lrcv[i];
            if (l >= 0) atomicAdd(&acc[l], acc[i]);
        }
        __syncthreads();
    }
    // a changed count whose receiver lies in another tile re-runs that tile
    bool marked = false;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        if (!valid[k]) continue;
        const int i = lane + WL * k;
        const int v = acc[i];
        if (v != old[k]) {
            const LaneCell q = lane_cell(k, r0, c0, nr, nc);
            __hip_atomic_store(&cnt[q.f], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (lrcv[i] < 0 && g[k] >= 0) {
                P.act_next[T.tbase + tile_of<TH>(g[k], nc, ntx)] = 1;
                marked = true;
            }
        }
    }
    const bool wmarked = __ballot(marked) != 0;
    if (lane == 0) {
        if (wmarked) *P.marked = 1;
        // per-tile counter (a single-address atomic from every tile would
        // serialise); an atomic whose result is unused is not waited on
        __hip_atomic_fetch_add(&P.runs[tile], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

template <int TH>
__global__ void __launch_bounds__(WL) k_tile_counts(const int* __restrict__ rcv, const uint8_t* __restrict__ mask,
                                                    const short* __restrict__ ordg, const short* __restrict__ lstartg,
                                                    const int* __restrict__ tile_maxdep,
                                                    const uint8_t* __restrict__ xdon, int nr, int nc, BC bc,
                                                    int* __restrict__ cnt, PassArgs P)
{
    if (P.prev_marked && *P.prev_marked == 0) return;
    tile_counts_body<TH>(grid_tile(), rcv, mask, ordg, lstartg, tile_maxdep, xdon, nr, nc, bc, cnt, P);
}

template <int TH>
__global__ void __launch_bounds__(WL) kb_tile_counts(BatchTiles B, const int* __restrict__ rcv,
                                                     const uint8_t* __restrict__ mask,
                                                     const short* __restrict__ ordg,
                                                     const short* __restrict__ lstartg,
                                                     const int* __restrict__ tile_maxdep,
                                                     const uint8_t* __restrict__ xdon, int* __restrict__ cnt,
                                                     PassArgs P)
{
    if (P.prev_marked && *P.prev_marked == 0) return;
    BatchGrid g;
    const TileAt T = batch_tile(B, g);
    tile_counts_body<TH>(T, rcv + g.off, mask + g.off, ordg, lstartg, tile_maxdep, xdon + g.off, g.nr, g.nc, g.bc,
                         cnt + g.off, P);
}

// ---------------------------------------------------------------------------
// stream power: outlet side first; a cell's new elevation needs its
// receiver's new elevation (LDS in the tile, HBM across the edge)
// ---------------------------------------------------------------------------

// Newton / closed form of lsdtt_stream_power_cpu.cpp:257-316 for a cell
// that is not skipped
// NM: 0 n == 1, 1 n - 1 == 0.5 (sqrt), 2 general (pow): a kernel per mode so
// that the common ones carry no pow registers
// Newton (n != 1): the per-node constants are formed once, the slope is a
// product by 1/dx_eff and the correction a product by the reciprocal of
// f'(z) (one refinement of the hardware estimate, ~1e-16 relative): the
// iterates stop at the same |eps| <= tol as the reference's divisions
// would, within rounding of the root.
template <int NM>
__device__ __forceinline__ double sp_solve(const SpParams& a, double zo, double zr, double kam, double dx_eff,
                                           double inv_dx_eff)
{
    const double tol = a.slope_clamp * dx_eff;
    double zn = zo;
    if (NM == 0) {
        const double spf = kam * (a.dt / dx_eff);
        zn = (zo + zr * spf) / (1.0 + spf);
        if ((zn - zo) > tol || isnan(zn)) zn = zo;
    } else {
        const double kdt = kam * a.dt;
        const double kdtn = kdt * (a.nexp * inv_dx_eff);
        double eps = tol + 1.0;
        int it = 0;
        do {
            const double slope = (zn - zr) * inv_dx_eff;
            if (slope <= 0.0) {
                eps = 0.0;
            } else {
                // S^n = S^(n-1) * S: one pow per iteration; sqrt is the
                // correctly rounded pow(S, 0.5)
                const double pn1 = NM == 1 ? sqrt(slope) : pow(slope, a.nexp - 1.0);
                const double fz = zn - zo + kdt * (pn1 * slope);
                const double fd = 1.0 + kdtn * pn1;
                double rf = __builtin_amdgcn_rcp(fd);
                rf = fma(fma(-fd, rf, 1.0), rf, rf);
                eps = fz * rf;
            }
            zn -= eps;
            ++it;
            if (it > 100) eps = tol / 2.0;
        } while (fabs(eps) > tol);
        if ((zn - zo) > tol || isnan(zn)) zn = zo;
        if (zn < zr) zn = zr + a.slope_clamp * dx_eff;
    }
    return zn;
}

constexpr short kNoRcv = -1, kOutRcv = -2;  // lrcv codes besides a local cell
// packed local receiver + D8 code of the stream power tiles
constexpr int kNoRcvP = 0xFFF, kOutRcvP = 0xFFE;
static_assert(TW * 32 <= kOutRcvP, "local indices must fit below the packed codes");
__device__ __forceinline__ short pack_lrd(short l, uint8_t d)
{
    const int lr = l == kNoRcv ? kNoRcvP : (l == kOutRcv ? kOutRcvP : static_cast<int>(l));
    return static_cast<short>(lr | ((d & 7) << 12));
}


// RR: a re-run (any pass after the first).  The tile's cells then hold the
// results of its previous run, computed from the receivers' values of that
// run, so a cell is solved again only when its receiver changed since: its
// in-tile receiver changed in this run, or its receiver lies in another tile
// (those cells are always solved again, and count as changed only if the
// result differs).  Every other cell keeps its value without a Newton solve,
// which is what makes the passes after the first cheap: a change at a tile's
// mouth decays within a few cells upstream.
template <int TH, bool MH, int NM, bool RR>
__device__ __forceinline__ void tile_sp_body(const TileAt& T, const SpTileArgs& A, int nr, int nc, BC bc,
                                             const PassArgs& P)
{
    constexpr int TN = TileShape<TH>::TN, PER = TileShape<TH>::PER;
    const int tx = T.tx, ty = T.ty, ntx = T.ntx;
    const int tile = T.tile;
    const int lane = threadIdx.x;
    if (lane == 0) P.act_clear[tile] = 0;
    // the first pass runs every tile
    if (RR && !P.act[tile]) return;
    // 20 bytes of LDS per cell (10 KB for a 32 x 16 tile), 29 in a re-run:
    // the D8 code shares a word with the local receiver
    // first run: z_old until the cell's level, then its new z; re-run: the
    // previous run's z until the cell's level, then its new z
    __shared__ double val[TN];
    __shared__ double kam[TN];   // K A^m; < 0: the reference skips the cell
    __shared__ short lrd[TN];    // local receiver (low 12 bits: index, kNoRcvP, kOutRcvP) | D8 code << 12
    __shared__ short ord[TN];
    // re-run only: z_old, and whether the cell's value changed in this run
    __shared__ double zos[RR ? TN : 1];
    __shared__ uint8_t chg[RR ? TN : 1];
    const int r0 = ty * TH, c0 = tx * TW;
    const SpParams& p = A.p;
    // every load of the tile is issued before any is used: a use of a later
    // load waits for all earlier ones, so loads interleaved with their uses
    // cost one memory round trip each
    LevelStarts<TH> lvl;
    lvl.load(A.lstart + (size_t)tile * (TN + 1));
    bool valid[PER];
    uint8_t x[PER], m[PER], dr[PER];
    short ordv[PER];
    int g[PER], ct[PER];
    double zo[PER], Kf[PER], ar[PER];
    double zcur[PER];  // the cells' current z (only this tile writes them)
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const LaneCell q = lane_cell(k, r0, c0, nr, nc);
        ordv[k] = A.ord[(size_t)tile * TN + lane + WL * k];
        m[k] = A.mask[q.f];
        g[k] = A.rcv[q.f];
        dr[k] = A.dir[q.f];
        ct[k] = A.cnt[q.f];
        zo[k] = A.z_old[q.f];
        Kf[k] = A.K ? A.K[q.f] : p.K0;
        ar[k] = A.area ? A.area[q.f] : 0.0;
        x[k] = A.xdon[q.f];
        zcur[k] = A.z[q.f];
        valid[k] = q.in;
    }
    const int maxdep = A.tile_maxdep[tile];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int i = lane + WL * k;
        valid[k] = valid[k] && m[k];
        double km = -1.0;
        short l = kNoRcv;
        if (valid[k] && g[k] >= 0) {
            double area = A.area ? ar[k] : static_cast<double>(ct[k]) * p.pixel_area;
            if (p.divide) area /= p.dx;
            if (!(p.threshold > 0.0 && area < p.threshold)) {
                // K_i * width_factor * pow(area, m) (lsdtt_stream_power_cpu.cpp:258-260)
                km = Kf[k] * p.width * (MH ? sqrt(area) : pow(area, p.m));
                const int li = local_receiver<TH>(g[k], dr[k], lane, k, nc, r0, c0);
                l = li >= 0 ? static_cast<short>(li) : kOutRcv;
            }
        }
        if (RR) {
            val[i] = zcur[k];
            zos[i] = zo[k];
            chg[i] = 0;
        } else {
            val[i] = zo[k];
        }
        kam[i] = km;
        lrd[i] = pack_lrd(l, dr[k]);
        ord[i] = ordv[k];
    }
    int lb = lvl.at(0);
    const double dxd = p.dx * kSqrt2, inv_dx = 1.0 / p.dx, inv_dxd = 1.0 / dxd;
    __syncthreads();
    int eroded = 0;
    for (int L = 0; L <= maxdep; ++L) {
        const int le = lvl.at(L + 1);
        for (int q = lb + lane; q < le; q += WL) {
            const int i = ord[q];
            const int w = lrd[i];
            const int lr = w & 0xFFF;
            if (lr == kNoRcvP) continue;
            const int dcode = (w >> 12) & 7;
            const double z0 = RR ? zos[i] : val[i];
            double zr;
            if (lr != kOutRcvP) {
                zr = val[lr];
                if (RR && !chg[lr]) {
                    // receiver unchanged: so is the cell (and its erosion)
                    eroded += zr < z0;
                    continue;
                }
            } else {
                const int g = neighbour(r0 + i / TW, c0 + i % TW, dcode, nr, nc, bc);
                zr = __hip_atomic_load(&A.z[g], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            if (RR) {
                const double v =
                    zr < z0 ? sp_solve<NM>(p, z0, zr, kam[i], (dcode & 1) ? dxd : p.dx, (dcode & 1) ? inv_dxd : inv_dx)
                            : z0;
                eroded += zr < z0;
                if (__double_as_longlong(v) != __double_as_longlong(val[i])) {
                    val[i] = v;
                    chg[i] = 1;
                }
            } else if (zr < z0) {
                val[i] = sp_solve<NM>(p, z0, zr, kam[i], (dcode & 1) ? dxd : p.dx, (dcode & 1) ? inv_dxd : inv_dx);
                ++eroded;
            }
        }
        lb = le;
        __syncthreads();
    }
    // a changed elevation with donors in other tiles re-runs those tiles
    bool marked = false;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        if (!valid[k]) continue;
        const int i = lane + WL * k;
        const LaneCell q = lane_cell(k, r0, c0, nr, nc);
        const double v = val[i];
        const double o = zcur[k];
        if (__double_as_longlong(o) != __double_as_longlong(v)) {
            __hip_atomic_store(&A.z[q.f], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            for (unsigned b = x[k]; b; b &= b - 1) {
                P.act_next[T.tbase + tile_of<TH>(neighbour(q.r, q.c, __builtin_ctz(b), nr, nc, bc), nc, ntx)] = 1;
                marked = true;
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) eroded += __shfl_xor(eroded, o);
    const bool wmarked = __ballot(marked) != 0;
    if (lane == 0) {
        A.tile_eroded[tile] = eroded;
        if (wmarked) *P.marked = 1;
        // per-tile counter (a single-address atomic from every tile would
        // serialise); an atomic whose result is unused is not waited on
        __hip_atomic_fetch_add(&P.runs[tile], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

template <int TH, bool MH, int NM, bool RR>
__global__ void __launch_bounds__(WL) k_tile_sp(SpTileArgs A, int nr, int nc, BC bc, PassArgs P)
{
    if (P.prev_marked && *P.prev_marked == 0) return;
    tile_sp_body<TH, MH, NM, RR>(grid_tile(), A, nr, nc, bc, P);
}

// batched: A holds the concatenated arrays; each tile's grid supplies its
// offset and stream-power parameters; tiles of grids whose exponent mode
// is not this kernel's are left to that mode's launch
template <int TH, bool MH, int NM, bool RR>
__global__ void __launch_bounds__(WL) kb_tile_sp(BatchTiles B, SpTileArgs A, PassArgs P)
{
    if (P.prev_marked && *P.prev_marked == 0) return;
    BatchGrid g;
    const TileAt T = batch_tile(B, g);
    if (g.sp_mode != sp_mode_of(MH, NM)) return;
    SpTileArgs a = A;
    a.rcv += g.off;
    a.dir += g.off;
    a.mask += g.off;
    a.cnt += g.off;
    a.xdon += g.off;
    a.z_old += g.off;
    a.z += g.off;
    a.area = nullptr;
    a.K = nullptr;
    a.p = g.sp;
    tile_sp_body<TH, MH, NM, RR>(T, a, g.nr, g.nc, g.bc, P);
}

// ---------------------------------------------------------------------------
// D-infinity accumulation (lsdtt_flowgraph_cpu.cpp:988-1033) by the same
// tile passes.  A cell's value is its own area plus, in D8 direction order
// (N, NE, ..., NW; a neighbour reached twice by a periodic wrap counts
// once), each donor's value times the donor's weight toward it; a donor is
// a neighbour one of whose (at most two) receivers is the cell.  The graph
// is acyclic, so the values are the unique solution whatever the pass
// schedule or tiling.  (The reference adds the donors in its Kahn order
// instead; the two sums differ only by rounding.)
// ---------------------------------------------------------------------------

// once per graph: each tile's cells sorted by in-tile depth, depth(v) =
// 1 + the largest depth of v's in-tile donors (0 without), found by
// relaxation to the least fixed point (LDS atomic max, until a sweep changes
// nothing); stored like k_tile_order's order
template <int TH>
__global__ void __launch_bounds__(WL) k_dinf_tile_order(const int* __restrict__ rec, const uint8_t* __restrict__ mask,
                                                        int nr, int nc, short* __restrict__ ordg,
                                                        short* __restrict__ lstartg, int* __restrict__ tile_maxdep)
{
    constexpr int TN = TileShape<TH>::TN, PER = TileShape<TH>::PER;
    __shared__ int dep[TN];
    __shared__ int hist[TN + 1];
    __shared__ int cur[TN];
    __shared__ short lrec[2][TN];
    const TileAt T = grid_tile();
    const int tx = T.tx, ty = T.ty, lane = threadIdx.x;
    const int tile = T.tile;
    const int r0 = ty * TH, c0 = tx * TW;
    bool valid[PER];
    {
        uint8_t m[PER];
        int g0[PER], g1[PER];
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const LaneCell q = lane_cell(k, r0, c0, nr, nc);
            m[k] = mask[q.f];
            g0[k] = rec[2 * (size_t)q.f];
            g1[k] = rec[2 * (size_t)q.f + 1];
            valid[k] = q.in && m[k];
        }
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int i = lane + WL * k;
            lrec[0][i] = static_cast<short>((valid[k] && g0[k] >= 0) ? local_index<TH>(g0[k], nc, r0, c0) : -1);
            lrec[1][i] = static_cast<short>((valid[k] && g1[k] >= 0) ? local_index<TH>(g1[k], nc, r0, c0) : -1);
            dep[i] = 0;
            hist[i] = 0;
        }
    }
    if (lane == 0) hist[TN] = 0;
    __syncthreads();
    // one wavefront per tile: the ballot is the whole tile's.  An acyclic
    // graph settles within TN sweeps (depths < TN); the cap and the clamp
    // below keep a malformed graph in bounds.
    for (int sweep = 0; sweep < TN; ++sweep) {
        bool ch = false;
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            if (!valid[k]) continue;
            const int i = lane + WL * k;
            const int d1 = dep[i] + 1;
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const int l = lrec[j][i];
                if (l >= 0 && atomicMax(&dep[l], d1) < d1) ch = true;
            }
        }
        const bool any = __ballot(ch) != 0;
        __syncthreads();
        if (!any) break;
    }
    int mx = 0;
#pragma unroll
    for (int k = 0; k < PER; ++k)
        if (valid[k]) {
            const int i = lane + WL * k;
            const int d = min(dep[i], TN - 1);
            dep[i] = d;
            atomicAdd(&hist[d], 1);
            mx = max(mx, d);
        }
    mx = wave_max(mx);
    __syncthreads();
    int loc[PER], sum = 0;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        loc[k] = hist[lane * PER + k];
        sum += loc[k];
    }
    int total;
    int base = wave_exclusive_scan(sum, &total);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        hist[lane * PER + k] = base;
        lstartg[(size_t)tile * (TN + 1) + lane * PER + k] = static_cast<short>(base);
        base += loc[k];
        cur[lane * PER + k] = 0;
    }
    if (lane == 0) {
        hist[TN] = total;
        lstartg[(size_t)tile * (TN + 1) + TN] = static_cast<short>(total);
        tile_maxdep[tile] = mx;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < PER; ++k)
        if (valid[k]) {
            const int i = lane + WL * k;
            const int d = dep[i];
            ordg[(size_t)tile * TN + hist[d] + atomicAdd(&cur[d], 1)] = static_cast<short>(i);
        }
}

struct DinfTileArgs {
    const int* rec;      // 2 per cell, -1 none
    const float* w;      // 2 per cell
    const uint8_t* mask;
    const short* ord;
    const short* lstart;
    const int* tile_maxdep;
    double cell_area;    // each valid cell's own contribution (dx^2)
    double* acc;
};

// ring slot of a tile cell (as in k_tile_order: 0..31 top row, 32..63
// bottom row, 64..95 left column, 96..127 right column), -1 inside
__device__ __forceinline__ int ring_slot(int lrow, int lcol, int h, int w)
{
    if (lrow == 0) return lcol;
    if (lrow == h - 1) return 32 + lcol;
    if (lcol == 0) return 64 + lrow;
    if (lcol == w - 1) return 96 + lrow;
    return -1;
}

// (lr, lc) of ring slot sl and whether the slot is a cell of the tile
__device__ __forceinline__ bool slot_cell(int sl, int h, int w, int* lr, int* lc)
{
    if (sl < 32) {
        *lr = 0, *lc = sl;
        return sl < w;
    }
    if (sl < 64) {
        *lr = h - 1, *lc = sl - 32;
        return *lc < w && h > 1;
    }
    if (sl < 96) {
        *lr = sl - 64, *lc = 0;
        return *lr >= 1 && *lr <= h - 2;
    }
    *lr = sl - 96, *lc = w - 1;
    return *lr >= 1 && *lr <= h - 2 && w > 1;
}

template <int TH>
__global__ void __launch_bounds__(WL) k_dinf_tile_accum(DinfTileArgs A, int nr, int nc, BC bc, PassArgs P)
{
    constexpr int TN = TileShape<TH>::TN, PER = TileShape<TH>::PER;
    if (P.prev_marked && *P.prev_marked == 0) return;
    const TileAt T = grid_tile();
    const int tx = T.tx, ty = T.ty, ntx = T.ntx;
    const int tile = T.tile;
    const int lane = threadIdx.x;
    if (lane == 0) P.act_clear[tile] = 0;
    if (P.prev_marked && !P.act[tile]) return;  // the first pass runs every tile
    __shared__ double val[TN];
    __shared__ float wl[2][TN];
    __shared__ short lr[2][TN];  // local receivers (-1: none or in another tile)
    __shared__ short ord[TN];
    __shared__ short ls[TN + 1];
    // ring cells: per direction the neighbour's local index (-1: none, the
    // cell itself, a repeat of an earlier direction, or another tile) and
    // the contributions of donors in other tiles, loaded before the sweep
    // into a pool: slot sl's donors, in direction order, from xoff[sl].
    // Each cell of the one-cell frame around the tile (2 TW + 2 TH + 4 at
    // most) is a donor of at most two ring cells, which bounds the pool.
    constexpr int kPool = 2 * (2 * TW + 2 * TH + 4);
    __shared__ double xv[kPool];
    __shared__ short xoff[128];
    __shared__ short nbl[128][8];
    __shared__ uint8_t xbits[128];
    const int r0 = ty * TH, c0 = tx * TW;
    const int h = min(TH, nr - r0), wd = min(TW, nc - c0);
    int g[PER][2];
    double old[PER];
    bool valid[PER];
    {
        // every load is issued before any is used (a use of a later load
        // waits for all earlier ones)
        float wv[PER][2];
        short ordv[PER];
        uint8_t mk[PER];
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const LaneCell q = lane_cell(k, r0, c0, nr, nc);
            ordv[k] = A.ord[(size_t)tile * TN + lane + WL * k];
            mk[k] = A.mask[q.f];
            g[k][0] = A.rec[2 * (size_t)q.f];
            g[k][1] = A.rec[2 * (size_t)q.f + 1];
            wv[k][0] = A.w[2 * (size_t)q.f];
            wv[k][1] = A.w[2 * (size_t)q.f + 1];
            old[k] = A.acc[q.f];  // only this tile writes its cells
            valid[k] = q.in;
        }
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            valid[k] = valid[k] && mk[k];
            const int i = lane + WL * k;
            lr[0][i] = static_cast<short>((valid[k] && g[k][0] >= 0) ? local_index<TH>(g[k][0], nc, r0, c0) : -1);
            lr[1][i] = static_cast<short>((valid[k] && g[k][1] >= 0) ? local_index<TH>(g[k][1], nc, r0, c0) : -1);
            wl[0][i] = wv[k][0];
            wl[1][i] = wv[k][1];
            val[i] = 0.0;
            ord[i] = ordv[k];
        }
    }
    const int maxdep = load_levels<TH>(A.lstart, A.tile_maxdep, tile, ls);
    int gsx[2][8];
    unsigned bitsx[2], slot1[2];  // donor directions; which receiver slot of the donor
#pragma unroll
    for (int half = 0; half < 2; ++half) {
        const int sl = lane + WL * half;
        int lrw, lcl;
        const bool ok = slot_cell(sl, h, wd, &lrw, &lcl);
        const int r = r0 + lrw, c = c0 + lcl, v = ok ? r * nc + c : 0;
        int* gs = gsx[half];
        bool ext[8];
#pragma unroll
        for (int d = 0; d < 8; ++d) {
            int u = -1;
            if (ok) {
                const Step st = resolve_step(r, c, d, nr, nc, bc);
                if (st.kind == 1) u = st.r * nc + st.c;
            }
            bool dup = u == v;
#pragma unroll
            for (int e = 0; e < d; ++e) dup |= gs[e] == u;
            if (dup) u = -1;
            gs[d] = u;
            const int lu = u >= 0 ? local_index<TH>(u, nc, r0, c0) : -1;
            ext[d] = u >= 0 && lu < 0;
            nbl[sl][d] = static_cast<short>(lu);
        }
        unsigned bits = 0, s1 = 0;
#pragma unroll
        for (int d = 0; d < 8; ++d) {
            const int u = ext[d] ? gs[d] : 0;
            const int h0 = ext[d] ? A.rec[2 * (size_t)u] : -1;
            const int h1 = ext[d] ? A.rec[2 * (size_t)u + 1] : -1;
            if (ext[d] && (h0 == v || h1 == v)) {
                bits |= 1u << d;
                if (h0 != v) s1 |= 1u << d;
            }
        }
        bitsx[half] = bits;
        slot1[half] = s1;
        xbits[sl] = static_cast<uint8_t>(bits);
    }
    {
        const int c0n = __popc(bitsx[0]), c1n = __popc(bitsx[1]);
        int total;
        const int base = wave_exclusive_scan(c0n + c1n, &total);
        const int off[2] = {base, base + c0n};
#pragma unroll
        for (int half = 0; half < 2; ++half) {
            const int sl = lane + WL * half;
            xoff[sl] = static_cast<short>(off[half]);
            int e = off[half];
#pragma unroll
            for (int d = 0; d < 8; ++d) {
                if (!((bitsx[half] >> d) & 1u)) continue;
                const int u = gsx[half][d];
                const double au = __hip_atomic_load(&A.acc[u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                xv[e++] = au * static_cast<double>(A.w[2 * (size_t)u + ((slot1[half] >> d) & 1u)]);
            }
        }
    }
    __syncthreads();
    for (int L = 0; L <= maxdep; ++L) {
        const int b = ls[L], e = ls[L + 1];
        for (int p = b + lane; p < e; p += WL) {
            const int i = ord[p];
            const int lrow = i / TW, lcol = i - lrow * TW;
            double a = A.cell_area;
            const int sl = ring_slot(lrow, lcol, h, wd);
            if (sl < 0) {
                // all eight neighbours are distinct cells of this tile
#pragma unroll
                for (int d = 0; d < 8; ++d) {
                    const int u = i + d8_row[d] * TW + d8_col[d];
                    if (lr[0][u] == i) a += val[u] * static_cast<double>(wl[0][u]);
                    else if (lr[1][u] == i) a += val[u] * static_cast<double>(wl[1][u]);
                }
            } else {
                const unsigned xb = xbits[sl];
                const int xo = xoff[sl];
#pragma unroll
                for (int d = 0; d < 8; ++d) {
                    if ((xb >> d) & 1u) {
                        a += xv[xo + __popc(xb & ((1u << d) - 1u))];
                    } else {
                        const int lu = nbl[sl][d];
                        if (lu >= 0) {
                            if (lr[0][lu] == i) a += val[lu] * static_cast<double>(wl[0][lu]);
                            else if (lr[1][lu] == i) a += val[lu] * static_cast<double>(wl[1][lu]);
                        }
                    }
                }
            }
            val[i] = a;
        }
        __syncthreads();
    }
    // a changed value whose receiver lies in another tile re-runs that tile
    bool marked = false;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        if (!valid[k]) continue;
        const int i = lane + WL * k;
        const double v = val[i];
        if (__double_as_longlong(v) != __double_as_longlong(old[k])) {
            const LaneCell q = lane_cell(k, r0, c0, nr, nc);
            __hip_atomic_store(&A.acc[q.f], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
            for (int j = 0; j < 2; ++j)
                if (g[k][j] >= 0 && lr[j][i] < 0) {
                    P.act_next[tile_of<TH>(g[k][j], nc, ntx)] = 1;
                    marked = true;
                }
        }
    }
    const bool wmarked = __ballot(marked) != 0;
    if (lane == 0) {
        if (wmarked) *P.marked = 1;
        __hip_atomic_fetch_add(&P.runs[tile], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// ---------------------------------------------------------------------------
// D-infinity stream power (lsdtt_stream_power_cpu.cpp:329-602) by the tile
// passes: a cell's new elevation is a function of its (at most two)
// receivers' new elevations, so a tile evaluates its cells in order of
// in-tile height above the tile's exit, height(v) = 1 + the largest height
// of v's in-tile receivers (0 without); receivers in other tiles are read
// from HBM, and a changed cell re-runs the tiles of its donors in other
// tiles (directions in xdon).
// ---------------------------------------------------------------------------
template <int TH>
__global__ void __launch_bounds__(WL) k_dinf_tile_up_order(const int* __restrict__ rec,
                                                           const uint8_t* __restrict__ mask, int nr, int nc,
                                                           BC bc, short* __restrict__ ordg,
                                                           short* __restrict__ lstartg, int* __restrict__ tile_maxdep,
                                                           uint8_t* __restrict__ xdon)
{
    constexpr int TN = TileShape<TH>::TN, PER = TileShape<TH>::PER;
    __shared__ int hgt[TN];
    __shared__ int hist[TN + 1];
    __shared__ int cur[TN];
    __shared__ short lrec[2][TN];
    __shared__ uint8_t xd[TN];
    const TileAt T = grid_tile();
    const int tx = T.tx, ty = T.ty, lane = threadIdx.x;
    const int tile = T.tile;
    const int r0 = ty * TH, c0 = tx * TW;
    bool valid[PER];
    {
        uint8_t m[PER];
        int g0[PER], g1[PER];
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const LaneCell q = lane_cell(k, r0, c0, nr, nc);
            m[k] = mask[q.f];
            g0[k] = rec[2 * (size_t)q.f];
            g1[k] = rec[2 * (size_t)q.f + 1];
            valid[k] = q.in && m[k];
        }
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int i = lane + WL * k;
            lrec[0][i] = static_cast<short>((valid[k] && g0[k] >= 0) ? local_index<TH>(g0[k], nc, r0, c0) : -1);
            lrec[1][i] = static_cast<short>((valid[k] && g1[k] >= 0) ? local_index<TH>(g1[k], nc, r0, c0) : -1);
            hgt[i] = 0;
            hist[i] = 0;
            xd[i] = 0;
        }
    }
    if (lane == 0) hist[TN] = 0;
    __syncthreads();
    // donors in other tiles: only ring cells can have them (ring slots as in
    // k_tile_order); a neighbour reached by two directions counts once
    {
        const int h = min(TH, nr - r0), w = min(TW, nc - c0);
#pragma unroll
        for (int half = 0; half < 2; ++half) {
            const int sl = lane + WL * half;
            int lr, lc;
            bool ok;
            if (sl < 32) {
                lr = 0, lc = sl, ok = lc < w;
            } else if (sl < 64) {
                lr = h - 1, lc = sl - 32, ok = lc < w && h > 1;
            } else if (sl < 96) {
                lr = sl - 64, lc = 0, ok = lr >= 1 && lr <= h - 2;
            } else {
                lr = sl - 96, lc = w - 1, ok = lr >= 1 && lr <= h - 2 && w > 1;
            }
            const int r = r0 + lr, c = c0 + lc;
            const int f = ok ? r * nc + c : 0;
            int gs[8];
#pragma unroll
            for (int d = 0; d < 8; ++d) {
                int g = ok ? neighbour(r, c, d, nr, nc, bc) : -1;
                if (g == f || (g >= 0 && local_index<TH>(g, nc, r0, c0) >= 0)) g = -1;
                gs[d] = g;
            }
            unsigned bits = 0;
#pragma unroll
            for (int d = 0; d < 8; ++d) {
                bool dup = false;
#pragma unroll
                for (int e = 0; e < d; ++e) dup |= gs[e] == gs[d];
                if (gs[d] >= 0 && !dup) {
                    const int a0 = rec[2 * (size_t)gs[d]], a1 = rec[2 * (size_t)gs[d] + 1];
                    if (a0 == f || a1 == f) bits |= 1u << d;
                }
            }
            if (ok && mask[f]) xd[lr * TW + lc] = static_cast<uint8_t>(bits);
        }
    }
    // heights by relaxation (each lane raises only its own cells)
    for (int sweep = 0; sweep < TN; ++sweep) {
        bool ch = false;
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            if (!valid[k]) continue;
            const int i = lane + WL * k;
            int hv = hgt[i];
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const int l = lrec[j][i];
                if (l >= 0) hv = max(hv, hgt[l] + 1);
            }
            if (hv > hgt[i]) {
                hgt[i] = hv;
                ch = true;
            }
        }
        const bool any = __ballot(ch) != 0;
        __syncthreads();
        if (!any) break;
    }
    int mx = 0;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int i = lane + WL * k;
        const LaneCell q = lane_cell(k, r0, c0, nr, nc);
        if (q.in) xdon[q.f] = xd[i];
        if (valid[k]) {
            const int d = min(hgt[i], TN - 1);
            hgt[i] = d;
            atomicAdd(&hist[d], 1);
            mx = max(mx, d);
        }
    }
    mx = wave_max(mx);
    __syncthreads();
    int loc[PER], sum = 0;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        loc[k] = hist[lane * PER + k];
        sum += loc[k];
    }
    int total;
    int base = wave_exclusive_scan(sum, &total);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        hist[lane * PER + k] = base;
        lstartg[(size_t)tile * (TN + 1) + lane * PER + k] = static_cast<short>(base);
        base += loc[k];
        cur[lane * PER + k] = 0;
    }
    if (lane == 0) {
        hist[TN] = total;
        lstartg[(size_t)tile * (TN + 1) + TN] = static_cast<short>(total);
        tile_maxdep[tile] = mx;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < PER; ++k)
        if (valid[k]) {
            const int i = lane + WL * k;
            const int d = hgt[i];
            ordg[(size_t)tile * TN + hist[d] + atomicAdd(&cur[d], 1)] = static_cast<short>(i);
        }
}

struct DinfSpArgs {
    const int* rec;
    const float* w;
    const uint8_t* mask;
    const double* area;  // m^2 per cell (after any smoothing)
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

// the reference's receiver classes (lsdtt_stream_power_cpu.cpp:435-441)
constexpr uint8_t kDinfSkip = 0, kDinfCard = 1, kDinfDiag = 2, kDinfTwo = 3;

template <int TH>
__global__ void __launch_bounds__(WL) k_dinf_tile_sp(DinfSpArgs A, int nr, int nc, BC bc, PassArgs P)
{
    constexpr int TN = TileShape<TH>::TN, PER = TileShape<TH>::PER;
    if (P.prev_marked && *P.prev_marked == 0) return;
    const TileAt T = grid_tile();
    const int tx = T.tx, ty = T.ty, ntx = T.ntx;
    const int tile = T.tile;
    const int lane = threadIdx.x;
    if (lane == 0) P.act_clear[tile] = 0;
    if (P.prev_marked && !P.act[tile]) return;  // the first pass runs every tile
    __shared__ double val[TN];  // z_old until the cell's level, then its new z
    __shared__ double kam[TN];
    __shared__ short lr[2][TN]; // local receivers, -1 outside the tile or none
    __shared__ uint8_t kind[TN];
    __shared__ short ord[TN];
    __shared__ double zx[128][2];  // ring cells: receivers in other tiles, loaded before the sweep
    const int r0 = ty * TH, c0 = tx * TW;
    const int h = min(TH, nr - r0), wd = min(TW, nc - c0);
    const SpParams& p = A.p;
    bool valid[PER];
    uint8_t x[PER];
    double zcur[PER];
    // every load is issued before any is used (a use of a later load waits
    // for all earlier ones)
    uint8_t mk[PER];
    int gv[PER][2];
    float wv[PER][2];
    double zov[PER], Kv[PER], arv[PER];
    short ordv[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const LaneCell q = lane_cell(k, r0, c0, nr, nc);
        ordv[k] = A.ord[(size_t)tile * TN + lane + WL * k];
        mk[k] = A.mask[q.f];
        gv[k][0] = A.rec[2 * (size_t)q.f];
        gv[k][1] = A.rec[2 * (size_t)q.f + 1];
        wv[k][0] = A.w[2 * (size_t)q.f];
        wv[k][1] = A.w[2 * (size_t)q.f + 1];
        zov[k] = A.z_old[q.f];
        Kv[k] = A.K ? A.K[q.f] : p.K0;
        arv[k] = A.area[q.f];
        x[k] = A.xdon[q.f];
        zcur[k] = A.z[q.f];
        valid[k] = q.in;
    }
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int g0 = gv[k][0], g1 = gv[k][1];
        const float w0 = wv[k][0], w1 = wv[k][1];
        const double zo = zov[k], Kf = Kv[k], ar = arv[k];
        valid[k] = valid[k] && mk[k];
        const int i = lane + WL * k;
        const int nrec = (g0 >= 0) + (g1 >= 0);
        uint8_t kd = kDinfSkip;
        double km = 0.0;
        if (valid[k] && nrec > 0) {
            double area = ar;
            if (p.divide) area /= p.dx;
            if (!(p.threshold > 0.0 && area < p.threshold)) {
                const bool only_card = nrec == 1 || w0 >= 1.0f - 1e-6f;
                const bool only_diag = nrec > 1 && w1 >= 1.0f - 1e-6f;
                kd = only_card ? kDinfCard : (only_diag ? kDinfDiag : kDinfTwo);
                // pow(area, 0.5) is sqrt(area): both correctly rounded
                km = Kf * p.width * (p.m_half ? sqrt(area) : pow(area, p.m));
            }
        }
        val[i] = zo;
        kam[i] = km;
        kind[i] = kd;
        lr[0][i] = static_cast<short>((valid[k] && g0 >= 0) ? local_index<TH>(g0, nc, r0, c0) : -1);
        lr[1][i] = static_cast<short>((valid[k] && g1 >= 0) ? local_index<TH>(g1, nc, r0, c0) : -1);
        ord[i] = ordv[k];
    }
#pragma unroll
    for (int half = 0; half < 2; ++half) {
        const int sl = lane + WL * half;
        int lrw, lcl;
        const bool ok = slot_cell(sl, h, wd, &lrw, &lcl);
        const int f = ok ? (r0 + lrw) * nc + c0 + lcl : 0;
        const int g0 = A.rec[2 * (size_t)f], g1 = A.rec[2 * (size_t)f + 1];
        const bool x0 = ok && g0 >= 0 && local_index<TH>(g0, nc, r0, c0) < 0;
        const bool x1 = ok && g1 >= 0 && local_index<TH>(g1, nc, r0, c0) < 0;
        zx[sl][0] = x0 ? __hip_atomic_load(&A.z[g0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0.0;
        zx[sl][1] = x1 ? __hip_atomic_load(&A.z[g1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0.0;
    }
    LevelStarts<TH> lvl;
    lvl.load(A.lstart + (size_t)tile * (TN + 1));
    const int maxdep = A.tile_maxdep[tile];
    int lb = lvl.at(0);
    __syncthreads();
    const double dxc = p.dx, dxd = p.dx * kSqrt2, tol = p.slope_clamp * p.dx;
    int eroded = 0;
    for (int L = 0; L <= maxdep; ++L) {
        const int le = lvl.at(L + 1);
        for (int q = lb + lane; q < le; q += WL) {
            const int i = ord[q];
            const uint8_t kd = kind[i];
            if (kd == kDinfSkip) continue;
            const double zo = val[i];
            // recv_elev (:443-448): the receivers' current elevations
            double ze[2];
            const int sl = ring_slot(i / TW, i % TW, h, wd);
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                // a receiver outside the tile is a ring cell's, preloaded;
                // a missing receiver's value is never used by its class
                const int l = lr[j][i];
                ze[j] = l >= 0 ? val[l] : (sl >= 0 ? zx[sl][j] : zo);
            }
            const double km = kam[i];
            double zn = zo;
            if (kd == kDinfTwo) {
                // :459-493
                double eps = tol + 1.0;
                int it = 0;
                do {
                    const double t1 = (ze[0] - zn) / dxc, t2 = (ze[1] - ze[0]) / dxc;
                    const double sbt = t1 * t1 + t2 * t2;
                    if (sbt <= 0.0) {
                        eps = 0.0;
                    } else {
                        // one pow per iteration: sbt^(n/2 - 1) = sbt^(n/2) / sbt
                        // (the reference evaluates both powers; they agree to
                        // an ulp or two)
                        const double pw = pow(sbt, 0.5 * p.nexp);
                        const double fz = zn - zo + km * pw * p.dt;
                        const double fd = 1.0 - km * p.nexp * (ze[0] - zn) / (dxc * dxc) * (pw / sbt) * p.dt;
                        eps = fz / fd;
                    }
                    zn -= eps;
                    ++it;
                    if (it > 100) eps = tol / 2.0;
                } while (fabs(eps) > tol);
                if ((zn - zo) > tol || isnan(zn)) zn = zo;
                const double zmin = ze[0] < ze[1] ? ze[0] : ze[1];
                if (zn < zmin) zn = zmin + p.slope_clamp * dxc;
                ++eroded;
            } else {
                // :497-566
                const double zr = kd == kDinfDiag ? ze[1] : ze[0];
                const double dxe = kd == kDinfDiag ? dxd : dxc;
                if (zr < zo) {
                    if (p.n_is_one) {
                        const double spf = km * (p.dt / dxe);
                        zn = (zo + zr * spf) / (1.0 + spf);
                        if ((zn - zo) > tol || isnan(zn)) zn = zo;
                    } else {
                        double eps = tol + 1.0;
                        int it = 0;
                        do {
                            const double slope = (zn - zr) / dxe;
                            if (slope <= 0.0) {
                                eps = 0.0;
                            } else {
                                // S^n = S^(n-1) * S: one pow per iteration
                                const double pn1 = pow(slope, p.nexp - 1.0);
                                const double fz = zn - zo + km * p.dt * (pn1 * slope);
                                const double fd = 1.0 + km * p.dt * (p.nexp / dxe) * pn1;
                                eps = fz / fd;
                            }
                            zn -= eps;
                            ++it;
                            if (it > 100) eps = tol / 2.0;
                        } while (fabs(eps) > tol);
                        if ((zn - zo) > tol || isnan(zn)) zn = zo;
                        if (zn < zr) zn = zr + p.slope_clamp * dxe;
                    }
                    ++eroded;
                }
            }
            val[i] = zn;
        }
        lb = le;
        __syncthreads();
    }
    bool marked = false;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        if (!valid[k]) continue;
        const int i = lane + WL * k;
        const LaneCell q = lane_cell(k, r0, c0, nr, nc);
        const double v = val[i];
        if (__double_as_longlong(zcur[k]) != __double_as_longlong(v)) {
            __hip_atomic_store(&A.z[q.f], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            for (unsigned b = x[k]; b; b &= b - 1) {
                P.act_next[tile_of<TH>(neighbour(q.r, q.c, __builtin_ctz(b), nr, nc, bc), nc, ntx)] = 1;
                marked = true;
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) eroded += __shfl_xor(eroded, o);
    const bool wmarked = __ballot(marked) != 0;
    if (lane == 0) {
        A.tile_eroded[tile] = eroded;
        if (wmarked) *P.marked = 1;
        __hip_atomic_fetch_add(&P.runs[tile], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// ---------------------------------------------------------------------------
// D-infinity steady-state chi (compute_dinf_fluvial_steady_state,
// lsdtt_steady_state_cpu.cpp:282-355) by the same tile passes and up order
// as the D-infinity stream power: a cell's chi is a function of its
// receivers' chi, in float.  Two receivers (neither weight >= 1 - 1e-6):
// dchi = (chi(e2) - chi(e1)) / dx, disc = amn^2 - dchi^2, chi = chi(e1) +
// dx sqrt(disc), or chi(e1) + dx amn when disc < 0; one receiver (slot 1
// when its weight is >= 1 - 1e-6, else slot 0): chi = chi(r) + dx_eff amn,
// dx_eff = dx sqrt2 when r is one row and one column away (by index
// difference, as the reference: a periodic wrap is not diagonal).
// ---------------------------------------------------------------------------
struct DinfChiArgs {
    const int* rec;
    const float* w;
    const uint8_t* mask;
    const float* amn;  // powf(A, -m/n) times the chi weight per valid cell
    const short* ord;
    const short* lstart;
    const int* tile_maxdep;
    const uint8_t* xdon;
    float dx, dxd;
    float* chi;
};

// 0: no receiver (chi stays 0); 1: two receivers; 2 / 3: one receiver in
// slot 0 / 1, plus 4 when it is a diagonal step
constexpr uint8_t kChiSkip = 0, kChiTwo = 1, kChiOne0 = 2, kChiOne1 = 3, kChiDiag = 4;

template <int TH>
__global__ void __launch_bounds__(WL) k_dinf_tile_chi(DinfChiArgs A, int nr, int nc, BC bc, PassArgs P)
{
    constexpr int TN = TileShape<TH>::TN, PER = TileShape<TH>::PER;
    if (P.prev_marked && *P.prev_marked == 0) return;
    const TileAt T = grid_tile();
    const int tx = T.tx, ty = T.ty, ntx = T.ntx;
    const int tile = T.tile;
    const int lane = threadIdx.x;
    if (lane == 0) P.act_clear[tile] = 0;
    if (P.prev_marked && !P.act[tile]) return;  // the first pass runs every tile
    __shared__ float val[TN];
    __shared__ float amn[TN];
    __shared__ short lr[2][TN];  // local receivers, -1 outside the tile or none
    __shared__ uint8_t kind[TN];
    __shared__ short ord[TN];
    __shared__ float cx[128][2];  // ring cells: chi of receivers in other tiles
    const int r0 = ty * TH, c0 = tx * TW;
    const int h = min(TH, nr - r0), wd = min(TW, nc - c0);
    bool valid[PER];
    uint8_t x[PER];
    float ccur[PER];
    // every load is issued before any is used (a use of a later load waits
    // for all earlier ones)
    uint8_t mk[PER];
    int gv[PER][2];
    float wv[PER][2], amv[PER];
    short ordv[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const LaneCell q = lane_cell(k, r0, c0, nr, nc);
        ordv[k] = A.ord[(size_t)tile * TN + lane + WL * k];
        mk[k] = A.mask[q.f];
        gv[k][0] = A.rec[2 * (size_t)q.f];
        gv[k][1] = A.rec[2 * (size_t)q.f + 1];
        wv[k][0] = A.w[2 * (size_t)q.f];
        wv[k][1] = A.w[2 * (size_t)q.f + 1];
        amv[k] = A.amn[q.f];
        x[k] = A.xdon[q.f];
        ccur[k] = A.chi[q.f];
        valid[k] = q.in;
    }
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const LaneCell q = lane_cell(k, r0, c0, nr, nc);
        const int g0 = gv[k][0], g1 = gv[k][1];
        const float w0 = wv[k][0], w1 = wv[k][1];
        valid[k] = valid[k] && mk[k];
        const int nrec = (g0 >= 0) + (g1 >= 0);
        uint8_t kd = kChiSkip;
        if (valid[k] && nrec > 0) {
            const bool one1 = nrec > 1 && w1 >= 1.0f - 1e-6f;
            if (nrec > 1 && !(w0 >= 1.0f - 1e-6f) && !one1) {
                kd = kChiTwo;
            } else {
                const int g = one1 ? g1 : g0;
                const int dr = abs(g / nc - q.r), dc = abs(g % nc - q.c);
                kd = static_cast<uint8_t>((one1 ? kChiOne1 : kChiOne0) | ((dr == 1 && dc == 1) ? kChiDiag : 0));
            }
        }
        const int i = lane + WL * k;
        val[i] = ccur[k];
        amn[i] = amv[k];
        kind[i] = kd;
        lr[0][i] = static_cast<short>((valid[k] && g0 >= 0) ? local_index<TH>(g0, nc, r0, c0) : -1);
        lr[1][i] = static_cast<short>((valid[k] && g1 >= 0) ? local_index<TH>(g1, nc, r0, c0) : -1);
        ord[i] = ordv[k];
    }
#pragma unroll
    for (int half = 0; half < 2; ++half) {
        const int sl = lane + WL * half;
        int lrw, lcl;
        const bool ok = slot_cell(sl, h, wd, &lrw, &lcl);
        const int f = ok ? (r0 + lrw) * nc + c0 + lcl : 0;
        const int g0 = A.rec[2 * (size_t)f], g1 = A.rec[2 * (size_t)f + 1];
        const bool x0 = ok && g0 >= 0 && local_index<TH>(g0, nc, r0, c0) < 0;
        const bool x1 = ok && g1 >= 0 && local_index<TH>(g1, nc, r0, c0) < 0;
        cx[sl][0] = x0 ? __hip_atomic_load(&A.chi[g0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0.0f;
        cx[sl][1] = x1 ? __hip_atomic_load(&A.chi[g1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0.0f;
    }
    LevelStarts<TH> lvl;
    lvl.load(A.lstart + (size_t)tile * (TN + 1));
    const int maxdep = A.tile_maxdep[tile];
    int lb = lvl.at(0);
    __syncthreads();
    const float dx = A.dx, dxd = A.dxd;
    for (int L = 0; L <= maxdep; ++L) {
        const int le = lvl.at(L + 1);
        for (int q = lb + lane; q < le; q += WL) {
            const int i = ord[q];
            const uint8_t kd = kind[i];
            if (kd == kChiSkip) continue;
            const int sl = ring_slot(i / TW, i % TW, h, wd);
            float c[2];
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const int l = lr[j][i];
                c[j] = l >= 0 ? val[l] : (sl >= 0 ? cx[sl][j] : 0.0f);
            }
            const float a = amn[i];
            float cn;
            if (kd == kChiTwo) {
                const float dchi = (c[1] - c[0]) / dx;
                const float disc = a * a - dchi * dchi;
                cn = disc >= 0.0f ? c[0] + dx * sqrtf(disc) : c[0] + dx * a;
            } else {
                const float dxe = (kd & kChiDiag) ? dxd : dx;
                cn = ((kd & 3) == kChiOne1 ? c[1] : c[0]) + dxe * a;
            }
            val[i] = cn;
        }
        lb = le;
        __syncthreads();
    }
    bool marked = false;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        if (!valid[k]) continue;
        const int i = lane + WL * k;
        const LaneCell q = lane_cell(k, r0, c0, nr, nc);
        const float v = val[i];
        if (__float_as_uint(ccur[k]) != __float_as_uint(v)) {
            __hip_atomic_store(&A.chi[q.f], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            for (unsigned b = x[k]; b; b &= b - 1) {
                P.act_next[tile_of<TH>(neighbour(q.r, q.c, __builtin_ctz(b), nr, nc, bc), nc, ntx)] = 1;
                marked = true;
            }
        }
    }
    const bool wmarked = __ballot(marked) != 0;
    if (lane == 0) {
        if (wmarked) *P.marked = 1;
        __hip_atomic_fetch_add(&P.runs[tile], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// smooth_drainage_area (lsdtt_stream_power_cpu.cpp:34-139), row pass: one
// thread per row keeps the reference's rolling sum (window -r..r, then per
// column: record, subtract column c - r, add column c + r + 1)
__device__ __forceinline__ int wrap_or_none(int i, int size, bool neg, bool pos)
{
    if (i < 0 && !neg) return -1;
    if (i >= size && !pos) return -1;
    return (i % size + size) % size;
}

// (cnt: upstream counts times pa = pixel area, or areas in m^2 with pa = 1)
template <class T>
__global__ void k_smooth_rows(const T* __restrict__ cnt, const uint8_t* __restrict__ mask, int nr, int nc,
                              double pa, int rad, bool pw, bool pe, double* __restrict__ hsum, int* __restrict__ hcnt)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nr) return;
    const long long row = (long long)r * nc;
    double sum = 0.0;
    int count = 0;
    for (int o = -rad; o <= rad; ++o) {
        const int c = wrap_or_none(o, nc, pw, pe);
        if (c >= 0 && mask[row + c]) {
            sum += static_cast<double>(cnt[row + c]) * pa;
            ++count;
        }
    }
    for (int c = 0; c < nc; ++c) {
        hsum[row + c] = sum;
        hcnt[row + c] = count;
        if (c + 1 < nc) {
            const int cm = wrap_or_none(c - rad, nc, pw, pe);
            if (cm >= 0 && mask[row + cm]) {
                sum += -1.0 * (static_cast<double>(cnt[row + cm]) * pa);
                --count;
            }
            const int cp = wrap_or_none(c + rad + 1, nc, pw, pe);
            if (cp >= 0 && mask[row + cp]) {
                sum += static_cast<double>(cnt[row + cp]) * pa;
                ++count;
            }
        }
    }
}

// column pass over the row sums: one thread per column (a wavefront reads
// 64 consecutive columns of a row: coalesced)
__global__ void k_smooth_cols(const uint8_t* __restrict__ mask, int nr, int nc, int rad, bool pn, bool ps,
                              const double* __restrict__ hsum, const int* __restrict__ hcnt, double* __restrict__ out)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    double sum = 0.0;
    long long count = 0;
    for (int o = -rad; o <= rad; ++o) {
        const int r = wrap_or_none(o, nr, pn, ps);
        if (r >= 0) {
            sum += hsum[(long long)r * nc + c];
            count += hcnt[(long long)r * nc + c];
        }
    }
    for (int r = 0; r < nr; ++r) {
        const long long f = (long long)r * nc + c;
        out[f] = (mask[f] && count != 0) ? sum / static_cast<double>(count) : 0.0;
        if (r + 1 < nr) {
            const int rm = wrap_or_none(r - rad, nr, pn, ps);
            if (rm >= 0) {
                sum += -1.0 * hsum[(long long)rm * nc + c];
                count -= hcnt[(long long)rm * nc + c];
            }
            const int rp = wrap_or_none(r + rad + 1, nr, pn, ps);
            if (rp >= 0) {
                sum += hsum[(long long)rp * nc + c];
                count += hcnt[(long long)rp * nc + c];
            }
        }
    }
}

__global__ void k_fill_i32(int* __restrict__ a, long long n, int v)
{
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ void k_copy_i32(const int* __restrict__ a, long long n, int* __restrict__ b)
{
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

// one-block sum (deterministic): kSumThreads threads, loads unrolled so that
// each thread keeps several in flight
constexpr int kSumThreads = 1024;
__global__ void __launch_bounds__(kSumThreads) k_sum_i32(const int* __restrict__ a, int n, long long* __restrict__ out)
{
    __shared__ long long s[kSumThreads];
    long long acc = 0;
    int i = threadIdx.x;
    for (; i + 3 * kSumThreads < n; i += 4 * kSumThreads) {
        const int v0 = a[i], v1 = a[i + kSumThreads], v2 = a[i + 2 * kSumThreads], v3 = a[i + 3 * kSumThreads];
        acc += (long long)v0 + v1 + v2 + v3;
    }
    for (; i < n; i += kSumThreads) acc += a[i];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int o = kSumThreads / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = s[0];
}

// calls f(std::integral_constant<int, TH>) for the runtime tile height
template <class F>
void with_tile_height(int th, F&& f)
{
    switch (th) {
    case 8: f(std::integral_constant<int, 8>{}); break;
    case 16: f(std::integral_constant<int, 16>{}); break;
    case 32: f(std::integral_constant<int, 32>{}); break;
    default: throw std::invalid_argument("tile height must be 8, 16 or 32");
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// batched sweeps: one launch covers the tiles of every grid of a batch
// ---------------------------------------------------------------------------
void batch_tile_order(int th, const BatchTiles& B, int ntiles, const int* rcv, const uint8_t* mask, short* ord,
                      short* lstart, int* maxdep, uint8_t* xdon, hipStream_t s)
{
    with_tile_height(th, [&](auto H) {
        kb_tile_order<decltype(H)::value><<<ntiles, WL, 0, s>>>(B, rcv, mask, ord, lstart, maxdep, xdon);
    });
    LSDTT_LAUNCH_CHECK();
}

void batch_tile_counts(int th, const BatchTiles& B, int ntiles, const int* rcv, const uint8_t* mask,
                       const short* ord, const short* lstart, const int* maxdep, const uint8_t* xdon, int* cnt,
                       const PassArgs& P, hipStream_t s)
{
    with_tile_height(th, [&](auto H) {
        kb_tile_counts<decltype(H)::value><<<ntiles, WL, 0, s>>>(B, rcv, mask, ord, lstart, maxdep, xdon, cnt, P);
    });
    LSDTT_LAUNCH_CHECK();
}

void batch_tile_sp(int th, unsigned modes, const BatchTiles& B, int ntiles, const SpTileArgs& A, const PassArgs& P,
                   hipStream_t s)
{
    with_tile_height(th, [&](auto H) {
        constexpr int h = decltype(H)::value;
        // the first pass of a settle runs every tile from z_old; later passes
        // are re-runs
        auto launch = [&](auto RR) {
            constexpr bool rr = decltype(RR)::value;
            if (modes & (1u << sp_mode_of(false, 0))) kb_tile_sp<h, false, 0, rr><<<ntiles, WL, 0, s>>>(B, A, P);
            if (modes & (1u << sp_mode_of(false, 1))) kb_tile_sp<h, false, 1, rr><<<ntiles, WL, 0, s>>>(B, A, P);
            if (modes & (1u << sp_mode_of(false, 2))) kb_tile_sp<h, false, 2, rr><<<ntiles, WL, 0, s>>>(B, A, P);
            if (modes & (1u << sp_mode_of(true, 0))) kb_tile_sp<h, true, 0, rr><<<ntiles, WL, 0, s>>>(B, A, P);
            if (modes & (1u << sp_mode_of(true, 1))) kb_tile_sp<h, true, 1, rr><<<ntiles, WL, 0, s>>>(B, A, P);
            if (modes & (1u << sp_mode_of(true, 2))) kb_tile_sp<h, true, 2, rr><<<ntiles, WL, 0, s>>>(B, A, P);
        };
        if (P.prev_marked) launch(std::true_type{});
        else launch(std::false_type{});
    });
    LSDTT_LAUNCH_CHECK();
}

bool FluvialTiles::valid_tile_height(int th)
{
    for (int h : kTileHeights)
        if (h == th) return true;
    return false;
}

void FluvialTiles::set_tile_height(int th_)
{
    if (!valid_tile_height(th_)) throw std::invalid_argument("tile height must be 8, 16 or 32");
    if (th_ == th) return;
    th = th_;
    // the pass heuristics were learnt on the other tiling
    last_passes = last_count_passes = last_sp_passes = last_dinf_passes = 0;
    if (nr > 0) resize(nr, nc, bc_);
}

void FluvialTiles::resize(int nr_, int nc_, BC bc)
{
    bc_ = bc;
    nr = nr_;
    nc = nc_;
    const int TN = TW * th;
    ntx = div_up(nc, TW);
    nty = div_up(nr, th);
    ntiles = ntx * nty;
    max_passes = ntiles + 2;
    d8_valid = false;
    dirty.reserve(ntiles);
    for (auto& a : act) a.reserve(ntiles);
    tile_eroded.reserve(ntiles);
    tile_maxdep.reserve(ntiles);
    ord.reserve((size_t)ntiles * TN);
    xdon.reserve((size_t)nr * nc);
    lstart.reserve((size_t)ntiles * (TN + 1));
    // [0, 2) eroded-cell sum and [2, 4) tile-run sum (long long each), then
    // per pass its marked flag
    flags.reserve(4 + (size_t)max_passes);
    flags_h.reserve(4 + (size_t)max_passes);
    tile_run.reserve(ntiles);
}

template <class T>
void FluvialTiles::smooth(const T* cnt, const uint8_t* mask, double pixel_area, int kernel_size, double* out,
                          hipStream_t s)
{
    const long long n = (long long)nr * nc;
    hsum.reserve(n);
    hcnt.reserve(n);
    const int rad = kernel_size / 2;
    k_smooth_rows<T><<<div_up(nr, 64), 64, 0, s>>>(cnt, mask, nr, nc, pixel_area, rad, bc_.w == 'p', bc_.e == 'p',
                                                   hsum.p, hcnt.p);
    LSDTT_LAUNCH_CHECK();
    k_smooth_cols<<<div_up(nc, 64), 64, 0, s>>>(mask, nr, nc, rad, bc_.n == 'p', bc_.s == 'p', hsum.p, hcnt.p, out);
    LSDTT_LAUNCH_CHECK();
}

void FluvialTiles::smooth_area(const int* cnt, const uint8_t* mask, double pixel_area, int kernel_size,
                               double* out, hipStream_t s)
{
    smooth<int>(cnt, mask, pixel_area, kernel_size, out, s);
}

void FluvialTiles::smooth_area(const double* area, const uint8_t* mask, int kernel_size, double* out, hipStream_t s)
{
    smooth<double>(area, mask, 1.0, kernel_size, out, s);
}

void FluvialTiles::clear_dirty(hipStream_t s)
{
    k_fill_i32<<<div_up(ntiles, 256), 256, 0, s>>>(dirty.p, ntiles, 0);
    LSDTT_LAUNCH_CHECK();
}

void FluvialTiles::depths(const int* rcv, const uint8_t* mask, hipStream_t s, bool only_dirty)
{
    const int* d = only_dirty && d8_valid ? dirty.p : nullptr;
    with_tile_height(th, [&](auto H) {
        k_tile_order<decltype(H)::value><<<dim3(ntx, nty), WL, 0, s>>>(rcv, mask, nr, nc, bc_, ord.p, lstart.p,
                                                                       tile_maxdep.p, xdon.p, d);
    });
    LSDTT_LAUNCH_CHECK();
    d8_valid = true;
}

// Runs passes in batches with one host read-back per batch: a pass after
// the settling one exits at once, so a batch that overshoots costs only
// empty launches.  `after_batch` may enqueue work whose result is read back
// with the flags (the eroded-cell sum).
template <class Pass, class After>
int FluvialTiles::settle(const char* name, double first_pass_bytes, Pass&& pass, After&& after_batch, hipStream_t s,
                         const int* start)
{
    const size_t nflags = 4 + (size_t)max_passes;
    k_fill_i32<<<div_up(nflags, 256), 256, 0, s>>>(flags.p, nflags, 0);
    LSDTT_LAUNCH_CHECK();
    k_fill_i32<<<div_up(ntiles, 256), 256, 0, s>>>(tile_run.p, ntiles, 0);
    LSDTT_LAUNCH_CHECK();
    if (start) k_copy_i32<<<div_up(ntiles, 256), 256, 0, s>>>(start, ntiles, act[0].p);
    else k_fill_i32<<<div_up(ntiles, 256), 256, 0, s>>>(act[0].p, ntiles, 1);
    LSDTT_LAUNCH_CHECK();
    k_fill_i32<<<div_up(ntiles, 256), 256, 0, s>>>(act[1].p, ntiles, 0);
    LSDTT_LAUNCH_CHECK();
    int launched = 0;
    int batch = std::max(2, last_passes + 1);
    for (;;) {
        const int end = std::min(max_passes, launched + batch);
        for (int q = launched; q < end; ++q) {
            PassArgs P{act[q % 3].p, act[(q + 1) % 3].p, act[(q + 2) % 3].p, q > 0 ? flags.p + 4 + q - 1 : nullptr,
                       flags.p + 4 + q, tile_run.p, start != nullptr};
            if (timer) timer->begin(name, s);
            pass(P);
            if (timer) timer->end(s, q == 0 ? first_pass_bytes : 0.0);
        }
        launched = end;
        after_batch();
        k_sum_i32<<<1, kSumThreads, 0, s>>>(tile_run.p, ntiles, reinterpret_cast<long long*>(flags.p + 2));
        LSDTT_LAUNCH_CHECK();
        LSDTT_HIP(hipMemcpyAsync(flags_h.data(), flags.p, (4 + (size_t)launched) * sizeof(int),
                                 hipMemcpyDeviceToHost, s));
        LSDTT_HIP(hipStreamSynchronize(s));
        int settled = -1;
        for (int q = 0; q < launched; ++q)
            if (!flags_h[4 + q]) {
                settled = q;
                break;
            }
        if (settled >= 0) {
            long long runs = 0;
            std::memcpy(&runs, flags_h.data() + 2, sizeof(long long));
            tile_runs = runs;
            last_passes = settled + 1;
            return last_passes;
        }
        if (launched >= max_passes) throw std::runtime_error("tile passes did not settle");
        batch *= 2;
    }
}

int FluvialTiles::counts(const int* rcv, const uint8_t* mask, int* cnt, bool warm, hipStream_t s, bool only_dirty)
{
    // The counts are the unique solution of cnt(v) = 1 + sum of the donors'
    // counts, and a pass re-runs every tile whose inputs differ from what it
    // read, so the passes reach that solution from any starting values.
    // Starting from the previous receivers' counts, only the tiles on routes
    // that changed need a second pass.
    const long long n = (long long)nr * nc;
    if (!warm) {
        k_fill_i32<<<div_up(n, 256), 256, 0, s>>>(cnt, n, 0);
        LSDTT_LAUNCH_CHECK();
    }
    last_passes = last_count_passes;
    // algorithmic bytes of the sweep, counted once: rcv 4, mask 1 in; cnt 4 out
    const int passes = settle(
        "upstream_counts", 9.0 * n,
        [&](const PassArgs& P) {
            with_tile_height(th, [&](auto H) {
                k_tile_counts<decltype(H)::value><<<dim3(ntx, nty), WL, 0, s>>>(rcv, mask, ord.p, lstart.p, tile_maxdep.p,
                                                                               xdon.p, nr, nc, bc_, cnt, P);
            });
            LSDTT_LAUNCH_CHECK();
        },
        [] {}, s, warm && only_dirty ? dirty.p : nullptr);
    last_count_passes = passes;
    return passes;
}

int FluvialTiles::stream_power(const int* rcv, const uint8_t* dir, const uint8_t* mask, const int* cnt,
                               const double* area, const double* z_old, const double* K, const SpParams& sp,
                               double* z, long long* eroded, hipStream_t s)
{
    SpTileArgs A{rcv, dir, mask, cnt, area, ord.p, lstart.p, tile_maxdep.p, xdon.p, z_old, K, z, tile_eroded.p, sp};
    long long* d_sum = reinterpret_cast<long long*>(flags.p);
    last_passes = last_sp_passes;
    // algorithmic bytes of the sweep, counted once: rcv 4, dir 1, mask 1,
    // cnt 4, z_old 8 (, K 8) in; z 8 read and 8 written
    const double bytes = (K ? 42.0 : 34.0) * (double)nr * nc;
    const int passes = settle(
        "stream_power", bytes,
        [&](const PassArgs& P) {
            const dim3 g(ntx, nty);
            const int nm = sp.n_is_one ? 0 : (sp.n1_half ? 1 : 2);
            // the first pass runs every tile from z_old; later passes are
            // re-runs that solve only the cells whose receiver changed
            auto launch_sp = [&](auto k) { k<<<g, WL, 0, s>>>(A, nr, nc, bc_, P); };
            auto launch = [&](auto H, auto RR) {
                constexpr int h = decltype(H)::value;
                constexpr bool rr = decltype(RR)::value;
                if (sp.m_half) {
                    if (nm == 0) launch_sp(k_tile_sp<h, true, 0, rr>);
                    else if (nm == 1) launch_sp(k_tile_sp<h, true, 1, rr>);
                    else launch_sp(k_tile_sp<h, true, 2, rr>);
                } else {
                    if (nm == 0) launch_sp(k_tile_sp<h, false, 0, rr>);
                    else if (nm == 1) launch_sp(k_tile_sp<h, false, 1, rr>);
                    else launch_sp(k_tile_sp<h, false, 2, rr>);
                }
            };
            with_tile_height(th, [&](auto H) {
                if (P.prev_marked) launch(H, std::true_type{});
                else launch(H, std::false_type{});
            });
            LSDTT_LAUNCH_CHECK();
        },
        [&] {
            k_sum_i32<<<1, kSumThreads, 0, s>>>(tile_eroded.p, ntiles, d_sum);
            LSDTT_LAUNCH_CHECK();
        },
        s);
    last_sp_passes = passes;
    std::memcpy(eroded, flags_h.data(), sizeof(long long));
    return passes;
}

void FluvialTiles::dinf_up_order(const int* rec, const uint8_t* mask, hipStream_t s)
{
    d8_valid = false;  // the D8 order shares these buffers
    with_tile_height(th, [&](auto H) {
        k_dinf_tile_up_order<decltype(H)::value><<<dim3(ntx, nty), WL, 0, s>>>(rec, mask, nr, nc, bc_, ord.p,
                                                                               lstart.p, tile_maxdep.p, xdon.p);
    });
    LSDTT_LAUNCH_CHECK();
}

int FluvialTiles::dinf_stream_power(const int* rec, const float* w, const uint8_t* mask, const double* area,
                                    const double* z_old, const double* K, const SpParams& sp, double* z,
                                    long long* eroded, hipStream_t s)
{
    DinfSpArgs A{rec, w, mask, area, ord.p, lstart.p, tile_maxdep.p, xdon.p, z_old, K, z, tile_eroded.p, sp};
    long long* d_sum = reinterpret_cast<long long*>(flags.p);
    last_passes = last_sp_passes;
    // algorithmic bytes of the sweep, counted once: rec 8, w 8, mask 1,
    // area 8, z_old 8 (, K 8) in; z 8 read and 8 written
    const double bytes = (K ? 57.0 : 49.0) * (double)nr * nc;
    const int passes = settle(
        "stream_power", bytes,
        [&](const PassArgs& P) {
            with_tile_height(th, [&](auto H) {
                k_dinf_tile_sp<decltype(H)::value><<<dim3(ntx, nty), WL, 0, s>>>(A, nr, nc, bc_, P);
            });
            LSDTT_LAUNCH_CHECK();
        },
        [&] {
            k_sum_i32<<<1, kSumThreads, 0, s>>>(tile_eroded.p, ntiles, d_sum);
            LSDTT_LAUNCH_CHECK();
        },
        s);
    last_sp_passes = passes;
    std::memcpy(eroded, flags_h.data(), sizeof(long long));
    return passes;
}

int FluvialTiles::dinf_chi(const int* rec, const float* w, const uint8_t* mask, const float* amn, float dx,
                           float* chi, hipStream_t s)
{
    // chi is the unique solution of the receiver recursion; starting from
    // zeros, the passes reach it whatever their schedule
    const long long n = (long long)nr * nc;
    k_fill_i32<<<div_up(n, 256), 256, 0, s>>>(reinterpret_cast<int*>(chi), n, 0);
    LSDTT_LAUNCH_CHECK();
    DinfChiArgs A{rec, w, mask, amn, ord.p, lstart.p, tile_maxdep.p, xdon.p, dx, dx * static_cast<float>(kSqrt2),
                  chi};
    last_passes = last_chi_passes;
    // algorithmic bytes of the sweep, counted once: rec 8, w 8, mask 1,
    // amn 4 in; chi 4 out
    const int passes = settle(
        "dinf_chi", 25.0 * n,
        [&](const PassArgs& P) {
            with_tile_height(th, [&](auto H) {
                k_dinf_tile_chi<decltype(H)::value><<<dim3(ntx, nty), WL, 0, s>>>(A, nr, nc, bc_, P);
            });
            LSDTT_LAUNCH_CHECK();
        },
        [] {}, s);
    last_chi_passes = passes;
    return passes;
}

void FluvialTiles::dinf_order(const int* rec, const uint8_t* mask, hipStream_t s)
{
    d8_valid = false;  // the D8 order shares these buffers
    with_tile_height(th, [&](auto H) {
        k_dinf_tile_order<decltype(H)::value><<<dim3(ntx, nty), WL, 0, s>>>(rec, mask, nr, nc, ord.p, lstart.p,
                                                                            tile_maxdep.p);
    });
    LSDTT_LAUNCH_CHECK();
}

int FluvialTiles::dinf_accumulate(const int* rec, const float* w, const uint8_t* mask, double cell_area,
                                  double* acc, bool warm, hipStream_t s)
{
    const long long n = (long long)nr * nc;
    if (!warm) {
        LSDTT_HIP(hipMemsetAsync(acc, 0, sizeof(double) * n, s));
    }
    DinfTileArgs A{rec, w, mask, ord.p, lstart.p, tile_maxdep.p, cell_area, acc};
    last_passes = last_dinf_passes;
    // algorithmic bytes of the sweep, counted once: rec 8, w 8, mask 1 in;
    // acc 8 out
    const int passes = settle(
        "dinf_accumulate", 25.0 * n,
        [&](const PassArgs& P) {
            with_tile_height(th, [&](auto H) {
                k_dinf_tile_accum<decltype(H)::value><<<dim3(ntx, nty), WL, 0, s>>>(A, nr, nc, bc_, P);
            });
            LSDTT_LAUNCH_CHECK();
        },
        [] {}, s);
    last_dinf_passes = passes;
    return passes;
}

}  // namespace lsdtt_gpu
