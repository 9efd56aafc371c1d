// flow.hip -- D8 flow engine kernels (see flow.hpp for the pipeline).
//
// Reference semantics restated here (src/cpu/lsdtt_flowinfo_cpu.cpp):
//   build_receiver_graph_d8 :1077-1329   classification order: masked ->
//       'b'-edge outlet -> forced outlet -> steepest descent (strict `>`,
//       dz in double from the stored z, slope = dz / (dx or dx*sqrt2)) ->
//       outlet if an off-grid 'o'/'b' step or (interior_nodata_is_outlet) a
//       masked neighbour exists -> outlet if allow_pits_as_outlets -> error
//   build_donor_counts_and_outlets_d8 :1340-1408
//   build_topological_order :1428-1522 (Kahn FIFO, sources in row-major order)
//   build_s_vector :1544-1731 (preorder DFS, outlets row-major, donors by
//       ascending flat index)
//   accumulate :1979-2046 (accum[rec] += accum[flat] in topological order)
#include <hip/hip_runtime.h>

#include <rocprim/block/block_reduce.hpp>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/iterator/transform_iterator.hpp>

#include <algorithm>

#include "flow.hpp"

namespace lsdtt_gpu {

namespace {

constexpr int kBlock = 256;

struct TScope {
    KernelTimer* t;
    hipStream_t s;
    double bytes;
    TScope(KernelTimer* t_, const char* name, hipStream_t s_, double b) : t(t_), s(s_), bytes(b)
    {
        if (t) t->begin(name, s);
    }
    ~TScope()
    {
        if (t) t->end(s, bytes);
    }
};
constexpr int kLdsLevels = 2048;

__device__ __forceinline__ unsigned long long ld_relaxed(const unsigned long long* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_relaxed(unsigned long long* p, unsigned long long v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------------
// receivers.  Interior cells (not on the grid's outer rows or columns):
// a thread walks kRecvRows consecutive rows of one column, rows along x so
// a wavefront reads 64 consecutive cells of a row; the rows r0 - 1 ..
// r0 + kRecvRows of the columns c - 1, c, c + 1 are all loaded before the
// first cell is decided (3 z and 3 mask loads per cell instead of 9 + 9,
// all in flight at once).  The outer ring, where steps wrap or leave the
// grid, is a second small launch; keeping its general path out of the
// interior kernel keeps that kernel's code small.
// ---------------------------------------------------------------------------
constexpr int kRecvRows = 4;

__device__ __forceinline__ void mark_tile(const TileMark& m, int g)
{
    const int r = g / m.nc, c = g - r * m.nc;
    m.dirty[(r / m.th) * m.ntx + c / 32] = 1;
}

// old: the receiver stored before (loaded with the kernel's other inputs
// when mark.dirty is set)
__device__ __forceinline__ void write_receiver(int f, int r, int d, int* __restrict__ rcv, uint8_t* __restrict__ dir,
                                               const TileMark& mark, int old)
{
    if (mark.dirty) {
        if (old != r) {
            mark_tile(mark, f);
            if (old >= 0) mark_tile(mark, old);
            if (r >= 0) mark_tile(mark, r);
        }
    }
    rcv[f] = r;
    dir[f] = static_cast<uint8_t>(d);
}

// the steepest descent of an interior cell from its 8 neighbours (D8 order
// N, NE, E, SE, S, SW, W, NW), the first of equal slopes winning
template <class T>
__device__ __forceinline__ void decide_interior(const uint8_t* __restrict__ forced, int nc, int row, int col,
                                                double card, double diag, double rcard, double rdiag,
                                                int nodata_outlet, int pits_outlet,
                                                uint8_t mc, T zcell, const uint8_t (&mk)[8], const T (&zn)[8],
                                                int* __restrict__ rcv, uint8_t* __restrict__ dir,
                                                int* __restrict__ fail, const TileMark& mark, int old)
{
    const int f = row * nc + col;
    if (!mc) {
        write_receiver(f, kNodataReceiver, kNoDirection, rcv, dir, mark, old);
        return;
    }
    if (forced && forced[f]) {
        write_receiver(f, kOutletReceiver, kNoDirection, rcv, dir, mark, old);
        return;
    }
    constexpr int kdr[8] = {-1, -1, 0, 1, 1, 1, 0, -1};
    constexpr int kdc[8] = {0, 1, 1, 1, 0, -1, -1, -1};
    const double zc = static_cast<double>(zcell);
    double best = 0.0;
    int bdir = kNoDirection;
    bool allowed = false;
#pragma unroll
    for (int d = 0; d < 8; ++d) {
        if (!mk[d]) {
            allowed |= (nodata_outlet != 0);
            continue;
        }
        const double dz = zc - static_cast<double>(zn[d]);
        if (dz <= 0.0) continue;
        const double sl = (d & 1) ? div_by_rcp(dz, diag, rdiag) : div_by_rcp(dz, card, rcard);
        if (sl > best) {
            best = sl;
            bdir = d;
        }
    }
    if (bdir != kNoDirection) {
        write_receiver(f, f + kdr[bdir] * nc + kdc[bdir], bdir, rcv, dir, mark, old);
    } else {
        write_receiver(f, kOutletReceiver, kNoDirection, rcv, dir, mark, old);
        if (!allowed && !pits_outlet) atomicMin(fail, f);
    }
}

template <class T>
__global__ void __launch_bounds__(256) k_receivers(const T* __restrict__ z, const uint8_t* __restrict__ mask,
                                                   const uint8_t* __restrict__ forced, int nr, int nc,
                                                   double card, double diag, double rcard, double rdiag,
                                                   int nodata_outlet, int pits_outlet,
                                                   int* __restrict__ rcv, uint8_t* __restrict__ dir,
                                                   int* __restrict__ fail, T* __restrict__ zcopy, TileMark mark,
                                                   int ntx)
{
    // row-major tile order (an XCD-banded order measured slower here)
    const int tile = blockIdx.x;
    const int col = (tile % ntx) * blockDim.x + threadIdx.x;
    const int r0 = ((tile / ntx) * blockDim.y + threadIdx.y) * kRecvRows;
    if (col >= nc || r0 >= nr) return;
    // the window: rows r0 - 1 .. r0 + kRecvRows, clamped into the grid (the
    // clamped values belong to ring cells, which this kernel skips)
    const int cm = col > 0 ? col - 1 : col, cp = col < nc - 1 ? col + 1 : col;
    T zw[kRecvRows + 2][3];
    uint8_t mw[kRecvRows + 2][3];
    int oldr[kRecvRows + 2];
#pragma unroll
    for (int k = 0; k < kRecvRows + 2; ++k) {
        const int rr = min(max(r0 - 1 + k, 0), nr - 1);
        const size_t b = (size_t)rr * nc;
        zw[k][0] = z[b + cm];
        zw[k][1] = z[b + col];
        zw[k][2] = z[b + cp];
        mw[k][0] = mask[b + cm];
        mw[k][1] = mask[b + col];
        mw[k][2] = mask[b + cp];
        oldr[k] = (mark.dirty && k >= 1 && k <= kRecvRows) ? rcv[b + col] : 0;
    }
    if (col == 0 || col == nc - 1) return;
#pragma unroll
    for (int k = 1; k <= kRecvRows; ++k) {
        const int row = r0 - 1 + k;
        if (row >= nr - 1) break;
        if (row == 0) continue;
        if (zcopy) zcopy[(size_t)row * nc + col] = zw[k][1];
        const uint8_t mk[8] = {mw[k - 1][1], mw[k - 1][2], mw[k][2], mw[k + 1][2],
                               mw[k + 1][1], mw[k + 1][0], mw[k][0], mw[k - 1][0]};
        const T zn[8] = {zw[k - 1][1], zw[k - 1][2], zw[k][2], zw[k + 1][2],
                         zw[k + 1][1], zw[k + 1][0], zw[k][0], zw[k - 1][0]};
        decide_interior<T>(forced, nc, row, col, card, diag, rcard, rdiag, nodata_outlet, pits_outlet, mw[k][1],
                           zw[k][1], mk, zn, rcv, dir, fail, mark, oldr[k]);
    }
}

// the outer ring (2 nc + 2 (nr - 2) cells, fewer on 1- and 2-wide grids):
// every step is resolved against the boundary conditions
template <class T>
__global__ void __launch_bounds__(256) k_receivers_ring(const T* __restrict__ z, const uint8_t* __restrict__ mask,
                                                        const uint8_t* __restrict__ forced, int nr, int nc,
                                                        double card, double diag, double rcard, double rdiag,
                                                        uint32_t bcw, int nodata_outlet,
                                                        int pits_outlet, int* __restrict__ rcv,
                                                        uint8_t* __restrict__ dir, int* __restrict__ fail,
                                                        T* __restrict__ zcopy, TileMark mark)
{
    const BC bc = bc_unpack(bcw);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    // top row, bottom row (when nr > 1), then the side columns of rows 1 .. nr - 2
    const int nbottom = nr > 1 ? nc : 0;
    const int nside = nr > 2 ? (nc > 1 ? 2 : 1) * (nr - 2) : 0;
    int row, col;
    if (i < nc) {
        row = 0, col = i;
    } else if (i < nc + nbottom) {
        row = nr - 1, col = i - nc;
    } else if (i < nc + nbottom + nside) {
        const int k = i - nc - nbottom;
        const int per = nc > 1 ? 2 : 1;
        row = 1 + k / per;
        col = (k % per) ? nc - 1 : 0;
    } else {
        return;
    }
    const int f = row * nc + col;
    const int old = mark.dirty ? rcv[f] : 0;
    if (zcopy) zcopy[f] = z[f];
    if (!mask[f]) {
        write_receiver(f, kNodataReceiver, kNoDirection, rcv, dir, mark, old);
        return;
    }
    const bool bedge = (row == 0 && bc.n == 'b') || (col == nc - 1 && bc.e == 'b') ||
                       (row == nr - 1 && bc.s == 'b') || (col == 0 && bc.w == 'b');
    if (bedge || (forced && forced[f])) {
        write_receiver(f, kOutletReceiver, kNoDirection, rcv, dir, mark, old);
        return;
    }
    const double zc = static_cast<double>(z[f]);
    double best = 0.0;
    int brcv = kOutletReceiver, bdir = kNoDirection;
    bool allowed = false;
    for (int d = 0; d < 8; ++d) {
        const Step st = resolve_step(row, col, d, nr, nc, bc);
        if (st.kind == 2) {
            allowed = true;
            continue;
        }
        if (st.kind == 0) continue;
        const int g = st.r * nc + st.c;
        if (!mask[g]) {
            allowed |= (nodata_outlet != 0);
            continue;
        }
        if (g == f) continue;
        const double dz = zc - static_cast<double>(z[g]);
        if (dz <= 0.0) continue;
        const double sl = (d & 1) ? div_by_rcp(dz, diag, rdiag) : div_by_rcp(dz, card, rcard);
        if (sl > best) {
            best = sl;
            brcv = g;
            bdir = d;
        }
    }
    if (bdir != kNoDirection) {
        write_receiver(f, brcv, bdir, rcv, dir, mark, old);
    } else {
        write_receiver(f, kOutletReceiver, kNoDirection, rcv, dir, mark, old);
        if (!allowed && !pits_outlet) atomicMin(fail, f);
    }
}

// both receiver launches; zcopy (or null) receives a copy of z
template <class T>
void launch_receivers(const T* z, const uint8_t* msk, const uint8_t* forced, int nr, int nc, double dx, BC bc,
                      bool nodata_outlet, bool pits_outlet, int* rcv, uint8_t* dir, int* fail, hipStream_t s,
                      T* zcopy = nullptr, const TileMark& mark = TileMark{})
{
    const dim3 b2(64, 4);
    const int ntx = div_up(nc, 64), tiles = ntx * div_up(nr, 4 * kRecvRows);
    // slopes divide by the two step lengths through their reciprocals
    const double card = dx, diag = dx * kSqrt2;
    k_receivers<T><<<tiles, b2, 0, s>>>(z, msk, forced, nr, nc, card, diag, 1.0 / card, 1.0 / diag, nodata_outlet,
                                        pits_outlet, rcv, dir, fail, zcopy, mark, ntx);
    LSDTT_LAUNCH_CHECK();
    const int nring = nc + (nr > 1 ? nc : 0) + (nr > 2 ? (nc > 1 ? 2 : 1) * (nr - 2) : 0);
    k_receivers_ring<T><<<div_up(nring, 256), 256, 0, s>>>(z, msk, forced, nr, nc, card, diag, 1.0 / card,
                                                          1.0 / diag, bc_pack(bc), nodata_outlet, pits_outlet, rcv,
                                                          dir, fail, zcopy, mark);
    LSDTT_LAUNCH_CHECK();
}

// donor count = number of distinct valid neighbours whose receiver is f;
// dmask bit i set when the i-th of distinct_neighbours(f) is one.  Blocks of
// 32 x 8 cells stage their receivers with a one-cell halo in LDS; interior
// cells (neighbours in direction order, none wrapped) read them there,
// raster-border cells take the general path.  Masked cells hold the no-data
// code, never a cell index.
constexpr int kDonW = 32, kDonH = 8;
__global__ void __launch_bounds__(kDonW * kDonH) k_donors(const int* __restrict__ rcv,
                                                          const uint8_t* __restrict__ mask, int nr, int nc, BC bc,
                                                          int ntx, int tiles, uint8_t* __restrict__ donors,
                                                          uint8_t* __restrict__ dmask)
{
    constexpr int TW = kDonW + 2, TH = kDonH + 2;
    __shared__ int sr[TH * TW];
    const int t = xcd_tile(blockIdx.x, tiles);
    const int c0 = (t % ntx) * kDonW, r0 = (t / ntx) * kDonH;
    const int tid = threadIdx.y * kDonW + threadIdx.x;
    for (int i = tid; i < TH * TW; i += kDonW * kDonH) {
        const int rr = r0 - 1 + i / TW, cc = c0 - 1 + i % TW;
        sr[i] = (rr >= 0 && rr < nr && cc >= 0 && cc < nc) ? rcv[rr * nc + cc] : kNodataReceiver;
    }
    __syncthreads();
    const int r = r0 + threadIdx.y, c = c0 + threadIdx.x;
    if (r >= nr || c >= nc) return;
    const int f = r * nc + c;
    unsigned bits = 0;
    if (mask[f]) {
        if (r > 0 && r < nr - 1 && c > 0 && c < nc - 1) {
            const int li = (threadIdx.y + 1) * TW + threadIdx.x + 1;
#pragma unroll
            for (int d = 0; d < 8; ++d) bits |= (sr[li + d8_row[d] * TW + d8_col[d]] == f) ? (1u << d) : 0u;
        } else {
            int nb[8];
            const int k = distinct_neighbours(f, nr, nc, bc, nb);
            for (int i = 0; i < k; ++i) bits |= (mask[nb[i]] && rcv[nb[i]] == f) ? (1u << i) : 0u;
        }
    }
    donors[f] = static_cast<uint8_t>(__popc(bits));
    dmask[f] = static_cast<uint8_t>(bits);
}

// ---------------------------------------------------------------------------
// pointer jumping on packed words (acc << 32 | next).  Invariant kept by
// every update: acc(v) = number of hops on the path from v up to, but not
// including, next(v).  Roots (outlets, masked cells) point at themselves.
// 64-bit loads/stores are single accesses, so updates in place are safe.
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned long long pack(unsigned acc, unsigned nxt)
{
    return (static_cast<unsigned long long>(acc) << 32) | nxt;
}

// Depth (unit weights): the packed words start already followed inside
// 32 x 32 tiles.  In LDS every cell holds (hops << 16 | local cell) and
// points at its receiver while that lies in the tile; pointer doubling
// (10 rounds cover any in-tile chain of the 1024 cells)
// takes it to the last cell L of its path in the tile, and the word becomes
// (hops to L, L) when L is a root, else (hops to L + 1, L's receiver).  A
// word read while its owner updates it is still a valid (hops, cell) pair
// of the same path.
constexpr int kPjT = 32;
__global__ void __launch_bounds__(kPjT * 8) k_pj_init_tiles(const int* __restrict__ rcv,
                                                            const uint8_t* __restrict__ mask, int nr, int nc,
                                                            int ntx, int tiles, unsigned long long* __restrict__ P)
{
    __shared__ unsigned st[kPjT * kPjT];
    __shared__ int rg[kPjT * kPjT];  // receiver (global; negative for roots and masked cells)
    const int tile = xcd_tile(blockIdx.x, tiles);
    const int c0 = (tile % ntx) * kPjT, r0 = (tile / ntx) * kPjT;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int ty = threadIdx.y + 8 * j, tx = threadIdx.x, t = ty * kPjT + tx;
        const int r = r0 + ty, c = c0 + tx;
        int g = -1;
        unsigned w = static_cast<unsigned>(t);
        if (r < nr && c < nc) {
            const int f = r * nc + c;
            g = mask[f] ? rcv[f] : -1;
            if (g >= 0) {
                const int gr = g / nc, gc = g - gr * nc;
                if (gr >= r0 && gr < r0 + kPjT && gc >= c0 && gc < c0 + kPjT)
                    w = (1u << 16) | static_cast<unsigned>((gr - r0) * kPjT + (gc - c0));
            }
        }
        rg[t] = g;
        st[t] = w;
    }
    __syncthreads();
    for (int it = 0; it < 10; ++it) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int t = (threadIdx.y + 8 * j) * kPjT + threadIdx.x;
            const unsigned w = st[t];
            const int p = static_cast<int>(w & 0xffffu);
            if (p != t) {
                const unsigned wp = st[p];
                st[t] = (((w >> 16) + (wp >> 16)) << 16) | (wp & 0xffffu);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int ty = threadIdx.y + 8 * j, tx = threadIdx.x, t = ty * kPjT + tx;
        const int r = r0 + ty, c = c0 + tx;
        if (r >= nr || c >= nc) continue;
        const unsigned w = st[t];
        const int L = static_cast<int>(w & 0xffffu);
        const unsigned hops = w >> 16;
        const int gL = rg[L];
        const int fL = (r0 + L / kPjT) * nc + (c0 + L % kPjT);
        P[r * nc + c] = gL >= 0 ? pack(hops + 1u, static_cast<unsigned>(gL)) : pack(hops, static_cast<unsigned>(fL));
    }
}

__global__ void k_pj_round(unsigned long long* __restrict__ P, long long n, const int* __restrict__ prev,
                           int* __restrict__ cur)
{
    // grid-stride over a few thousand blocks: a round after convergence
    // returns at once, and costs little as a launch
    if (*prev == 0) return;
    bool ch = false;
    for (long long f = blockIdx.x * (long long)blockDim.x + threadIdx.x; f < n;
         f += (long long)gridDim.x * blockDim.x) {
        const unsigned long long w = ld_relaxed(&P[f]);
        const unsigned nx = static_cast<unsigned>(w);
        if (nx == f) continue;
        const unsigned long long wn = ld_relaxed(&P[nx]);
        const unsigned nn = static_cast<unsigned>(wn);
        if (nn != nx) {
            st_relaxed(&P[f], pack(static_cast<unsigned>(w >> 32) + static_cast<unsigned>(wn >> 32), nn));
            ch = true;
        }
    }
    // one flag store per wavefront
    const unsigned long long m = __ballot(ch);
    if (ch && (threadIdx.x & 63) == __ffsll(m) - 1) *cur = 1;
}

// value(v) = acc(v) + acc(root) for non-roots; acc(v) for roots.  Also block
// max of values (depth) and valid-cell count when `stats` is set.
// grid-stride (a few thousand blocks) so the two global atomics per block
// stay uncontended
// gate: when not null, the extraction runs only if *gate == 0 (the last
// launched round changed nothing)
// stats (when not null): [0] max value, [1] valid cells, [2] valid cells
// without donors (sources)
__global__ void k_pj_extract(const unsigned long long* __restrict__ P, const uint8_t* __restrict__ mask,
                             const uint8_t* __restrict__ donors, long long n, int* __restrict__ value,
                             int* __restrict__ stats, const int* __restrict__ gate)
{
    if (gate && *gate != 0) return;
    int v = -1, valid = 0, sources = 0;
    for (long long f = blockIdx.x * (long long)blockDim.x + threadIdx.x; f < n;
         f += (long long)gridDim.x * blockDim.x) {
        int vf = -1;
        if (mask[f]) {
            const unsigned long long w = P[f];
            const unsigned nx = static_cast<unsigned>(w);
            unsigned acc = static_cast<unsigned>(w >> 32);
            if (nx != f) acc += static_cast<unsigned>(P[nx] >> 32);
            vf = static_cast<int>(acc);
            ++valid;
            if (stats) sources += donors[f] == 0;
        }
        value[f] = vf;
        v = max(v, vf);
    }
    if (stats) {
        using BR = rocprim::block_reduce<int, kBlock>;
        __shared__ typename BR::storage_type s1;
        int m = v, c = valid;
        BR().reduce(m, m, s1, rocprim::maximum<int>());
        __syncthreads();
        BR().reduce(c, c, s1);
        __syncthreads();
        int so = sources;
        BR().reduce(so, so, s1);
        if (threadIdx.x == 0) {
            atomicMax(&stats[0], m);
            atomicAdd(&stats[1], c);
            atomicAdd(&stats[2], so);
        }
    }
}

// small device-side initialisations (no copies from pageable host memory):
// p[0..2] = a, b, c and p[3..count) = rest
__global__ void k_init_ints(int* __restrict__ p, int count, int a, int b, int c, int rest)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = i == 0 ? a : i == 1 ? b : i == 2 ? c : rest;
}

// a few single ints gathered into one block (after the round flags), for
// one read-back
__global__ void k_collect_ints(FlowEngine::ReadInts extra, int* __restrict__ out)
{
    const int i = threadIdx.x;
    if (i < extra.count) out[i] = *extra.src[i];
}

// ---------------------------------------------------------------------------
// level buckets: counting sort of valid cells by depth
// ---------------------------------------------------------------------------
// Wave-aggregated add of 1 to h[d] for every active lane: lanes holding the
// same d share one atomic (most cells sit at a handful of shallow levels, so
// per-lane atomics would serialise on one bin).  Returns each lane's rank
// among the lanes of the wave that added to its bin, plus that bin's base.
__device__ __forceinline__ int wave_agg_add(int* h, int d, bool active)
{
    const int lane = threadIdx.x & 63;
    unsigned long long act = __ballot(active);
    int rank = 0;
    while (act) {
        const int leader = __ffsll(static_cast<long long>(act)) - 1;
        const int dl = __shfl(d, leader);
        const bool mine = active && d == dl;
        const unsigned long long same = __ballot(mine);
        int base = 0;
        if (lane == leader) base = atomicAdd(&h[dl], __popcll(same));
        base = __shfl(base, leader);
        if (mine) rank = base + __popcll(same & ((1ull << lane) - 1ull));
        act &= ~same;
    }
    return rank;
}

constexpr int kCellsPerThread = 4;

// Level counters are split by block group for the shallow levels (the
// first `lg`): block b counts into slot d * kLevelGroups + b % kLevelGroups,
// so the few crowded levels do not serialise every block's atomics on one
// address; deeper levels keep one slot each after them.  In slot order the
// counters are level-major, so their exclusive scan gives every group its
// range inside its level's.
constexpr int kLevelGroups = 64;
__device__ __forceinline__ int level_slot(int d, int g, int lg)
{
    return d < lg ? d * kLevelGroups + g : lg * kLevelGroups + (d - lg);
}

__global__ void __launch_bounds__(1024) k_level_hist(const int* __restrict__ depth, long long n, int lg,
                                                     int* __restrict__ count)
{
    __shared__ int h[kLdsLevels];
    for (int i = threadIdx.x; i < kLdsLevels; i += blockDim.x) h[i] = 0;
    __syncthreads();
    const int g = blockIdx.x % kLevelGroups;
    const long long base = blockIdx.x * (long long)blockDim.x * kCellsPerThread;
    for (int j = 0; j < kCellsPerThread; ++j) {
        const long long f = base + j * (long long)blockDim.x + threadIdx.x;
        const int d = f < n ? depth[f] : -1;
        wave_agg_add(h, d < kLdsLevels ? d : 0, d >= 0 && d < kLdsLevels);
        if (d >= kLdsLevels) atomicAdd(&count[level_slot(d, g, lg)], 1);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kLdsLevels; i += blockDim.x)
        if (h[i]) atomicAdd(&count[level_slot(i, g, lg)], h[i]);
}

// the start of every level (and the total after the last)
__global__ void k_level_starts(const int* __restrict__ sub, int levels, int lg, int* __restrict__ start)
{
    const int L = blockIdx.x * blockDim.x + threadIdx.x;
    if (L < levels) start[L] = sub[level_slot(L, 0, lg)];
}

__global__ void __launch_bounds__(1024) k_level_scatter(const int* __restrict__ depth, long long n, int lg,
                                                        const int* __restrict__ sub, int* __restrict__ cursor,
                                                        int* __restrict__ order)
{
    __shared__ int h[kLdsLevels];
    __shared__ int base_s[kLdsLevels];
    for (int i = threadIdx.x; i < kLdsLevels; i += blockDim.x) h[i] = 0;
    __syncthreads();
    const int g = blockIdx.x % kLevelGroups;
    const long long base = blockIdx.x * (long long)blockDim.x * kCellsPerThread;
    int d[kCellsPerThread], rank[kCellsPerThread];
    for (int j = 0; j < kCellsPerThread; ++j) {
        const long long f = base + j * (long long)blockDim.x + threadIdx.x;
        d[j] = f < n ? depth[f] : -1;
        rank[j] = wave_agg_add(h, d[j] < kLdsLevels ? d[j] : 0, d[j] >= 0 && d[j] < kLdsLevels);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kLdsLevels; i += blockDim.x)
        if (h[i]) base_s[i] = atomicAdd(&cursor[level_slot(i, g, lg)], h[i]);
    __syncthreads();
    for (int j = 0; j < kCellsPerThread; ++j) {
        const int dj = d[j];
        if (dj < 0) continue;
        const long long f = base + j * (long long)blockDim.x + threadIdx.x;
        const int slot = level_slot(dj, g, lg);
        int pos;
        if (dj < kLdsLevels) pos = sub[slot] + base_s[dj] + rank[j];
        else pos = sub[slot] + atomicAdd(&cursor[slot], 1);
        order[pos] = static_cast<int>(f);
    }
}

// ---------------------------------------------------------------------------
// upstream sweep (levels deep -> shallow).  A node's donors all sit one level
// deeper, so they are final when the node is processed.
// ---------------------------------------------------------------------------
struct UpArgs {
    const int* rcv;
    const uint8_t* mask;
    const uint8_t* dmask;  // donor bits per cell, in distinct_neighbours order
    int nr, nc;
    BC bc;
    int* cnt;
    unsigned long long* key;  // null: counts only
    const double* w;          // null: no weighted accumulation
    double* acc;
    int* off;                 // null: no s-vector sibling offsets
};

__device__ __forceinline__ void up_node(const UpArgs& a, int v)
{
    // the donors (from the routing's donor bits), then every donor's count
    // and key, issued before any is used
    const unsigned dm = a.dmask[v];
    if (dm == 0) {  // a source
        a.cnt[v] = 1;
        if (a.key) a.key[v] = static_cast<unsigned long long>(static_cast<unsigned>(v));
        if (a.acc) a.acc[v] = a.w[v];
        return;
    }
    int nb[8];
    distinct_neighbours(v, a.nr, a.nc, a.bc, nb);
    int cv[8];
    unsigned long long kv[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        cv[i] = 0;
        kv[i] = 0;
        if ((dm >> i) & 1u) {
            cv[i] = a.cnt[nb[i]];
            if (a.key) kv[i] = a.key[nb[i]];
        }
    }
    int c = 1;
    unsigned long long mk = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        c += cv[i];
        mk = max(mk, kv[i]);
    }
    a.cnt[v] = c;
    if (a.key)
        a.key[v] = dm ? ((((mk >> 32) + 1ull) << 32) | (mk & 0xffffffffull))
                      : static_cast<unsigned long long>(static_cast<unsigned>(v));
    if (a.off) {
        // s-vector sibling offsets: donors are visited in ascending flat
        // order, each after the subtrees of the donors before it
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            if (!((dm >> i) & 1u)) continue;
            int o = 1;
#pragma unroll
            for (int j = 0; j < 8; ++j)
                if (((dm >> j) & 1u) && nb[j] < nb[i]) o += cv[j];
            a.off[nb[i]] = o;
        }
    }
    if (a.acc) {
        // donors in topological order = ascending Kahn key
        int dn[8];
        unsigned long long dk[8];
        int nd = 0;
        for (int i = 0; i < 8; ++i)
            if ((dm >> i) & 1u) {
                dn[nd] = nb[i];
                dk[nd] = kv[i];
                ++nd;
            }
        for (int i = 1; i < nd; ++i) {
            const unsigned long long kk = dk[i];
            const int vv = dn[i];
            int j = i;
            while (j > 0 && dk[j - 1] > kk) {
                dk[j] = dk[j - 1];
                dn[j] = dn[j - 1];
                --j;
            }
            dk[j] = kk;
            dn[j] = vv;
        }
        double s = a.w[v];
        for (int i = 0; i < nd; ++i) s += a.acc[dn[i]];
        a.acc[v] = s;
    }
}

__global__ void k_up_level(const int* __restrict__ order, int begin, int end, UpArgs a)
{
    const int i = begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (i < end) up_node(a, order[i]);
}

// one workgroup walks the levels [lo, hi] from hi down to lo
__global__ void __launch_bounds__(1024) k_up_tail(const int* __restrict__ order,
                                                  const int* __restrict__ lstart, int lo, int hi, UpArgs a)
{
    for (int L = hi; L >= lo; --L) {
        const int b = lstart[L], e = lstart[L + 1];
        for (int i = b + threadIdx.x; i < e; i += blockDim.x) up_node(a, order[i]);
        __syncthreads();
    }
}

// order-free counts by atomics: cnt starts at 1; each node adds its final
// count to its receiver
__global__ void k_cnt_init(const uint8_t* __restrict__ mask, long long n, int* __restrict__ cnt)
{
    const long long f = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (f < n) cnt[f] = mask[f] ? 1 : 0;
}
__global__ void k_cnt_level(const int* __restrict__ order, int begin, int end, const int* __restrict__ rcv,
                            int* __restrict__ cnt)
{
    const int i = begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (i < end) {
        const int v = order[i];
        atomicAdd(&cnt[rcv[v]], cnt[v]);
    }
}
__global__ void __launch_bounds__(1024) k_cnt_tail(const int* __restrict__ order, const int* __restrict__ lstart,
                                                   int lo, int hi, const int* __restrict__ rcv,
                                                   int* __restrict__ cnt)
{
    for (int L = hi; L >= lo; --L) {
        const int b = lstart[L], e = lstart[L + 1];
        for (int i = b + threadIdx.x; i < e; i += blockDim.x) {
            const int v = order[i];
            atomicAdd(&cnt[rcv[v]], __hip_atomic_load(&cnt[v], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        }
        __syncthreads();
    }
}

// compact Kahn keys: (height << nbits) | last-donor ancestor, in 32 bits
// when they fit
template <class K>
__global__ void k_topo_keys(const int* __restrict__ cells, int m, const unsigned long long* __restrict__ key,
                            int nbits, K* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) {
        const unsigned long long k = key[cells[i]];
        out[i] = static_cast<K>(((k >> 32) << nbits) | (k & 0xffffffffull));
    }
}

// Stable three-way split of the cells by a class (0, 1, or 2 = dropped),
// classes 0 and 1 each written in flat order, in three launches: per-chunk
// counts, one scan of the chunk counts (which also writes both totals), the
// writes.  Each thread takes four consecutive cells per step, their loads
// issued together.  Used for the topological order (0: sources -- no
// donors, Kahn key (0, own index) -- 1: the other valid cells) and for the
// outlets (0: cells whose receiver is the outlet code).
constexpr int kPartChunks = 1024;
__device__ __forceinline__ long long part_chunk_len(long long n)
{
    const long long step = 4ll * kBlock;
    return ((n + kPartChunks - 1) / kPartChunks + step - 1) / step * step;
}
struct SourceClass {
    const uint8_t* mask;
    const uint8_t* donors;
    __device__ int operator()(long long f) const
    {
        const uint8_t mv = mask[f], dv = donors[f];
        return !mv ? 2 : (dv == 0 ? 0 : 1);
    }
};
struct OutletClass {
    const int* rcv;  // masked cells hold the no-data code, never the outlet code
    __device__ int operator()(long long f) const { return rcv[f] == kOutletReceiver ? 0 : 2; }
};
template <class C>
__global__ void __launch_bounds__(kBlock) k_part_count(C cls, long long n, int* __restrict__ counts)
{
    const long long len = part_chunk_len(n);
    const long long b0 = blockIdx.x * len, b1 = std::min(n, b0 + len);
    int c0 = 0, c1 = 0;
    for (long long f = b0 + 4ll * threadIdx.x; f < b1; f += 4ll * blockDim.x) {
        int k[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) k[j] = f + j < b1 ? cls(f + j) : 2;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            c0 += k[j] == 0;
            c1 += k[j] == 1;
        }
    }
    using BR = rocprim::block_reduce<int, kBlock>;
    __shared__ typename BR::storage_type sr;
    BR().reduce(c0, c0, sr);
    __syncthreads();
    BR().reduce(c1, c1, sr);
    if (threadIdx.x == 0) {
        counts[blockIdx.x] = c0;
        counts[kPartChunks + blockIdx.x] = c1;
    }
}
// exclusive scans of both count rows (one block of kPartChunks threads);
// totals[0..1] = the two class sizes
__global__ void __launch_bounds__(kPartChunks) k_part_scan(int* __restrict__ counts, int* __restrict__ totals)
{
    __shared__ int sc[2][kPartChunks];
    const int t = threadIdx.x;
    const int a0 = counts[t], a1 = counts[kPartChunks + t];
    sc[0][t] = a0;
    sc[1][t] = a1;
    __syncthreads();
    for (int o = 1; o < kPartChunks; o <<= 1) {
        const int u0 = t >= o ? sc[0][t - o] : 0, u1 = t >= o ? sc[1][t - o] : 0;
        __syncthreads();
        sc[0][t] += u0;
        sc[1][t] += u1;
        __syncthreads();
    }
    counts[t] = sc[0][t] - a0;
    counts[kPartChunks + t] = sc[1][t] - a1;
    if (t == kPartChunks - 1) {
        totals[0] = sc[0][t];
        totals[1] = sc[1][t];
    }
}
// exclusive prefix over the block of per-thread values (in thread order);
// returns the prefix and sets `total`
__device__ __forceinline__ int block_prefix(int v, int& total)
{
    __shared__ int ws[kBlock / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int x = v;
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) ws[wave] = x;
    __syncthreads();
    int before = 0;
    total = 0;
    for (int w = 0; w < kBlock / 64; ++w) {
        before += w < wave ? ws[w] : 0;
        total += ws[w];
    }
    __syncthreads();
    return before + x - v;
}
template <class C>
__global__ void __launch_bounds__(kBlock) k_part_write(C cls, long long n, const int* __restrict__ offs,
                                                       int* __restrict__ out0, int* __restrict__ out1)
{
    const long long len = part_chunk_len(n);
    const long long b0 = blockIdx.x * len, b1 = std::min(n, b0 + len);
    int base0 = offs[blockIdx.x], base1 = offs[kPartChunks + blockIdx.x];
    for (long long f0 = b0; f0 < b1; f0 += 4ll * blockDim.x) {
        const long long f = f0 + 4ll * threadIdx.x;
        int k[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) k[j] = f + j < b1 ? cls(f + j) : 2;
        int n0 = 0, n1 = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            n0 += k[j] == 0;
            n1 += k[j] == 1;
        }
        int t0, t1;
        int p0 = block_prefix(n0, t0);
        int p1 = block_prefix(n1, t1);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (k[j] == 0) out0[base0 + p0++] = static_cast<int>(f + j);
            if (k[j] == 1) out1[base1 + p1++] = static_cast<int>(f + j);
        }
        base0 += t0;
        base1 += t1;
    }
}


// s-vector sibling offset: 1 + sizes of the receiver's donors with a smaller
// flat index (donors are visited in ascending flat order); outlets take the
// exclusive prefix of outlet sizes in row-major order.
__global__ void k_sv_offsets(const int* __restrict__ rcv, const uint8_t* __restrict__ mask,
                             const int* __restrict__ cnt, int nr, int nc, BC bc, int* __restrict__ off)
{
    const long long f = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (f >= (long long)nr * nc) return;
    const int r = rcv[f];
    if (!mask[f] || r < 0) return;  // roots start at their outlet's prefix
    int nb[8];
    const int k = distinct_neighbours(r, nr, nc, bc, nb);
    int s = 1;
    for (int i = 0; i < k; ++i) {
        const int u = nb[i];
        if (u < f && mask[u] && rcv[u] == r) s += cnt[u];
    }
    off[f] = s;
}

// an outlet's subtree size
struct CountOf {
    const int* cnt;
    __device__ int operator()(int v) const { return cnt[v]; }
};
// s-vector positions, outlets first (their subtrees in row-major order),
// then level by level away from them: a cell's subtree starts at its
// receiver's position plus its sibling offset
__global__ void k_sv_outlet_start(const int* __restrict__ outlets, int no, const int* __restrict__ pref,
                                  int* __restrict__ start, int* __restrict__ svec)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < no) {
        const int v = outlets[i], st = pref[i];
        start[v] = st;
        svec[st] = v;
    }
}
struct SvArgs {
    const int* rcv;
    const int* off;
    int* start;
    int* svec;
};
__device__ __forceinline__ void sv_node(const SvArgs& a, int v)
{
    const int st = a.start[a.rcv[v]] + a.off[v];
    a.start[v] = st;
    a.svec[st] = v;
}
__global__ void k_sv_level(const int* __restrict__ order, int begin, int end, SvArgs a)
{
    const int i = begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (i < end) sv_node(a, order[i]);
}
// one workgroup walks levels lo..hi (each reads the level before it)
__global__ void __launch_bounds__(1024) k_sv_tail(const int* __restrict__ order, const int* __restrict__ lstart,
                                                  int lo, int hi, SvArgs a)
{
    for (int L = lo; L <= hi; ++L) {
        const int b = lstart[L], e = lstart[L + 1];
        for (int i = b + threadIdx.x; i < e; i += blockDim.x) sv_node(a, order[i]);
        __syncthreads();
    }
}

inline int bit_length(unsigned long long x)
{
    int b = 0;
    while (x) {
        ++b;
        x >>= 1;
    }
    return b;
}

}  // namespace

FlowEngine::FlowEngine(const FlowOptions& o) : opt(o)
{
    if (o.nr <= 0 || o.nc <= 0) throw std::invalid_argument("grid must have at least one row and column");
    n = static_cast<long long>(o.nr) * o.nc;
    if (n >= (1ll << 31) - 1) throw std::invalid_argument("grids of 2^31 cells or more are not supported");
    rcv.reserve(n);
    dir.reserve(n);
    donors.reserve(n);
    dmask.reserve(n);
    depth.reserve(n);
    pj.reserve(n);
    flags.reserve(n);
    scan_out.reserve(n);
    outlets.reserve(n);
    misc.reserve(8);
    pjflags.reserve(kReadSlot + 8);
    hread.reserve(kReadSlot + 8);
}

template <class T>
long long FlowEngine::route(const T* z, const uint8_t* msk, const uint8_t* forced, hipStream_t s)
{
    mask = msk;
    off_ready = false;
    const int nr = opt.nr, nc = opt.nc;
    // misc[0] = first unresolved flat, misc[1] = max depth, misc[2] = n_valid,
    // misc[3] = valid cells without donors
    k_init_ints<<<1, 64, 0, s>>>(misc.p, 4, 0x7fffffff, -1, 0, 0);
    LSDTT_LAUNCH_CHECK();
    {
        TScope ts_(timer, "receivers", s, (sizeof(T) + 6.0) * n);
        launch_receivers<T>(z, msk, forced, nr, nc, opt.dx, opt.bc, opt.interior_nodata_is_outlet,
                            opt.allow_pits_as_outlets, rcv, dir, misc.p, s);
    }
    {
        TScope ts_(timer, "donors", s, 6.0 * n);
        const int ntx = div_up(nc, kDonW), tiles = ntx * div_up(nr, kDonH);
        k_donors<<<tiles, dim3(kDonW, kDonH), 0, s>>>(rcv, msk, nr, nc, opt.bc, ntx, tiles, donors, dmask);
    }
    LSDTT_LAUNCH_CHECK();
    // the outlets in row-major order
    part_counts.reserve(2 * kPartChunks + 2);
    const OutletClass ocls{rcv};
    k_part_count<<<kPartChunks, kBlock, 0, s>>>(ocls, n, part_counts.p);
    k_part_scan<<<1, kPartChunks, 0, s>>>(part_counts.p, part_counts.p + 2 * kPartChunks);
    k_part_write<<<kPartChunks, kBlock, 0, s>>>(ocls, n, part_counts.p, outlets.p, nullptr);
    LSDTT_LAUNCH_CHECK();

    // one read-back with the pointer jumping's: misc and the outlet total
    ReadInts extra;
    extra.src[0] = misc.p;
    extra.src[1] = misc.p + 1;
    extra.src[2] = misc.p + 2;
    extra.src[3] = part_counts.p + 2 * kPartChunks;
    extra.src[4] = misc.p + 3;
    extra.count = 5;
    pointer_jump(depth, s, extra);
    const int* h = hread.p + kReadSlot;
    n_outlets = h[3];
    if (h[0] != 0x7fffffff) return -(static_cast<long long>(h[0]) + 1);
    max_depth = h[1];
    n_valid = h[2];
    n_sources = h[4];

    level_start_h.assign(static_cast<size_t>(std::max(max_depth, 0)) + 2, 0);
    if (n_valid == 0) {
        tail_from = 0;
        return 0;
    }
    const int nl = max_depth + 1;
    // level counters: shallow levels split by block group (see
    // level_slot), then one slot per deeper level
    const int lg = std::min(nl + 1, kLdsLevels);
    const long long slots = (long long)lg * kLevelGroups + (nl + 1 - lg);
    level_count.reserve(slots);
    level_sub.reserve(slots);
    level_start.reserve(nl + 1);
    cursor.reserve(slots);
    order.reserve(n_valid);
    LSDTT_HIP(hipMemsetAsync(level_count.p, 0, sizeof(int) * slots, s));
    LSDTT_HIP(hipMemsetAsync(cursor.p, 0, sizeof(int) * slots, s));
    const int g1k = div_up(n, 1024 * kCellsPerThread);
    {
        TScope ts_(timer, "level_hist", s, 4.0 * n);
        k_level_hist<<<g1k, 1024, 0, s>>>(depth, n, lg, level_count);
    }
    LSDTT_LAUNCH_CHECK();
    size_t tb = 0;
    LSDTT_HIP(rocprim::exclusive_scan(nullptr, tb, level_count.p, level_sub.p, 0, (size_t)slots,
                                      rocprim::plus<int>(), s));
    LSDTT_HIP(rocprim::exclusive_scan(tmp.need(tb), tb, level_count.p, level_sub.p, 0, (size_t)slots,
                                      rocprim::plus<int>(), s));
    k_level_starts<<<div_up(nl + 1, kBlock), kBlock, 0, s>>>(level_sub, nl + 1, lg, level_start);
    {
        TScope ts_(timer, "level_scatter", s, 8.0 * n);
        k_level_scatter<<<g1k, 1024, 0, s>>>(depth, n, lg, level_sub, cursor, order);
    }
    LSDTT_LAUNCH_CHECK();
    LSDTT_HIP(hipMemcpyAsync(level_start_h.data(), level_start.p, sizeof(int) * (nl + 1), hipMemcpyDeviceToHost, s));
    LSDTT_HIP(hipStreamSynchronize(s));
    // levels from tail_from on are all small enough for one workgroup
    tail_from = nl;
    while (tail_from > 0 && level_start_h[tail_from] - level_start_h[tail_from - 1] < kTailThreshold) --tail_from;
    return 0;
}

template <class T>
void FlowEngine::receivers_only(const T* z, const uint8_t* msk, hipStream_t s, T* zcopy, const TileMark& mark)
{
    mask = msk;
    off_ready = false;
    TScope ts_(timer, "receivers", s, (sizeof(T) + 6.0) * n);
    launch_receivers<T>(z, msk, nullptr, opt.nr, opt.nc, opt.dx, opt.bc, opt.interior_nodata_is_outlet,
                        opt.allow_pits_as_outlets, rcv, dir, misc.p, s, zcopy, mark);
}

template void FlowEngine::receivers_only<double>(const double*, const uint8_t*, hipStream_t, double*,
                                                 const TileMark&);
template void FlowEngine::receivers_only<float>(const float*, const uint8_t*, hipStream_t, float*, const TileMark&);

template long long FlowEngine::route<float>(const float*, const uint8_t*, const uint8_t*, hipStream_t);
template long long FlowEngine::route<double>(const double*, const uint8_t*, const uint8_t*, hipStream_t);

void FlowEngine::pointer_jump(int* value_out, hipStream_t s, const ReadInts& extra)
{
    const int g1 = div_up(n, kBlock), gs = std::min(g1, 2048);
    {
        const int ntx = div_up(opt.nc, kPjT), tiles = ntx * div_up(opt.nr, kPjT);
        k_pj_init_tiles<<<tiles, dim3(kPjT, 8), 0, s>>>(rcv, mask, opt.nr, opt.nc, ntx, tiles, pj);
    }
    LSDTT_LAUNCH_CHECK();
    // chg[r] == 1 while round r-1 changed something; a round after
    // convergence returns at once
    int* chg = pjflags.p;
    k_init_ints<<<1, 64, 0, s>>>(chg, kPjRounds + 1, 1, 0, 0, 0);
    LSDTT_LAUNCH_CHECK();
    int* stats = misc.p + 1;
    auto rounds = [&](int r0, int r1) {
        for (int r = r0; r < r1; ++r) {
            {
                TScope ts_(timer, "pointer_jump", s, 24.0 * n);
                k_pj_round<<<gs, kBlock, 0, s>>>(pj, n, chg + r, chg + r + 1);
            }
            LSDTT_LAUNCH_CHECK();
        }
    };
    auto read_back = [&]() {
        static_assert(kPjRounds + 1 <= kReadSlot && kReadSlot + 8 <= 64, "read-back slots");
        if (extra.count > 0) {
            k_collect_ints<<<1, 64, 0, s>>>(extra, pjflags.p + kReadSlot);
            LSDTT_LAUNCH_CHECK();
        }
        LSDTT_HIP(hipMemcpyAsync(hread.p, pjflags.p, sizeof(int) * (kReadSlot + 8), hipMemcpyDeviceToHost, s));
        LSDTT_HIP(hipStreamSynchronize(s));
    };
    // a first batch as long as the last call needed (plus a spare): the
    // extraction runs only if the batch's last round changed nothing, else
    // the remaining rounds and the extraction follow the read-back
    const int first = std::min(kPjRounds, pj_hint);
    rounds(0, first);
    k_pj_extract<<<gs, kBlock, 0, s>>>(pj, mask, donors, n, value_out, stats, chg + first);
    LSDTT_LAUNCH_CHECK();
    read_back();
    if (hread.p[first] != 0) {
        if (first == kPjRounds) throw std::runtime_error("pointer jumping did not converge");
        rounds(first, kPjRounds);
        k_pj_extract<<<gs, kBlock, 0, s>>>(pj, mask, donors, n, value_out, stats, nullptr);
        LSDTT_LAUNCH_CHECK();
        read_back();
        if (hread.p[kPjRounds] != 0) throw std::runtime_error("pointer jumping did not converge");
    }
    int changed = 0;
    while (changed < kPjRounds && hread.p[changed + 1] != 0) ++changed;
    pj_hint = changed + 2;
}

void FlowEngine::upstream(bool keys, const double* w, double* acc, hipStream_t s, bool sv_offsets)
{
    cnt.reserve(n);
    if (keys || w) key.reserve(n);
    if (sv_offsets) off.reserve(n);
    off_ready = sv_offsets;
    if (n_valid == 0) return;
    UpArgs a{rcv, mask, dmask, opt.nr, opt.nc, opt.bc, cnt, (keys || w) ? key.p : nullptr, w, w ? acc : nullptr,
             sv_offsets ? off.p : nullptr};
    const int nl = max_depth + 1;
    if (tail_from < nl) {
        k_up_tail<<<1, 1024, 0, s>>>(order, level_start, tail_from, nl - 1, a);
        LSDTT_LAUNCH_CHECK();
    }
    for (int L = tail_from - 1; L >= 0; --L) {
        const int b = level_start_h[L], e = level_start_h[L + 1];
        if (e > b) {
            k_up_level<<<div_up(e - b, kBlock), kBlock, 0, s>>>(order, b, e, a);
            LSDTT_LAUNCH_CHECK();
        }
    }
}

void FlowEngine::upstream_counts_fast(hipStream_t s)
{
    cnt.reserve(n);
    k_cnt_init<<<div_up(n, kBlock), kBlock, 0, s>>>(mask, n, cnt);
    LSDTT_LAUNCH_CHECK();
    if (n_valid == 0) return;
    const int nl = max_depth + 1;
    // level 0 holds the roots; they have no receiver to add to
    const int lo = std::max(tail_from, 1);
    if (lo < nl) {
        {
            TScope ts_(timer, "upstream_counts_tail", s, 16.0 * (level_start_h[nl] - level_start_h[lo]));
            k_cnt_tail<<<1, 1024, 0, s>>>(order, level_start, lo, nl - 1, rcv, cnt);
        }
        LSDTT_LAUNCH_CHECK();
    }
    for (int L = lo - 1; L >= 1; --L) {
        const int b = level_start_h[L], e = level_start_h[L + 1];
        if (e > b) {
            {
                TScope ts_(timer, "upstream_counts", s, 16.0 * (e - b));
                k_cnt_level<<<div_up(e - b, kBlock), kBlock, 0, s>>>(order, b, e, rcv, cnt);
            }
            LSDTT_LAUNCH_CHECK();
        }
    }
}

void FlowEngine::topological_order(hipStream_t s)
{
    topo.reserve(n_valid ? n_valid : 1);
    if (n_valid == 0) return;
    sort_keys_a.reserve(n_valid);
    sort_keys_b.reserve(n_valid);
    // Sources sort first (height 0), by their own index: a stable partition
    // puts them in place; only the other cells go through the radix sort.
    const int ns = n_sources, m = n_valid - n_sources;
    topo_inner.reserve(n_valid);
    int* inner = topo_inner.p;
    part_counts.reserve(2 * kPartChunks + 2);
    const SourceClass cls{mask, donors.p};
    k_part_count<<<kPartChunks, kBlock, 0, s>>>(cls, n, part_counts.p);
    k_part_scan<<<1, kPartChunks, 0, s>>>(part_counts.p, part_counts.p + 2 * kPartChunks);
    k_part_write<<<kPartChunks, kBlock, 0, s>>>(cls, n, part_counts.p, topo.p, inner);
    LSDTT_LAUNCH_CHECK();
    if (m == 0) return;
    size_t tb = 0;
    const int nbits = std::max(1, bit_length(static_cast<unsigned long long>(n - 1)));
    const int end_bit = std::min(64, nbits + bit_length(static_cast<unsigned long long>(max_depth) + 1));
    int* out = topo.p + ns;
    if (end_bit <= 32) {
        // half the key bytes on every radix pass
        unsigned* ka = reinterpret_cast<unsigned*>(sort_keys_a.p);
        unsigned* kb = reinterpret_cast<unsigned*>(sort_keys_b.p);
        k_topo_keys<unsigned><<<div_up(m, kBlock), kBlock, 0, s>>>(inner, m, key, nbits, ka);
        LSDTT_LAUNCH_CHECK();
        LSDTT_HIP(rocprim::radix_sort_pairs(nullptr, tb, ka, kb, inner, out, (size_t)m, 0, end_bit, s));
        LSDTT_HIP(rocprim::radix_sort_pairs(tmp.need(tb), tb, ka, kb, inner, out, (size_t)m, 0, end_bit, s));
        return;
    }
    k_topo_keys<unsigned long long><<<div_up(m, kBlock), kBlock, 0, s>>>(inner, m, key, nbits, sort_keys_a.p);
    LSDTT_LAUNCH_CHECK();
    LSDTT_HIP(rocprim::radix_sort_pairs(nullptr, tb, sort_keys_a.p, sort_keys_b.p, inner, out, (size_t)m, 0, end_bit,
                                        s));
    LSDTT_HIP(rocprim::radix_sort_pairs(tmp.need(tb), tb, sort_keys_a.p, sort_keys_b.p, inner, out, (size_t)m, 0,
                                        end_bit, s));
}

void FlowEngine::s_vector(hipStream_t s)
{
    start.reserve(n);
    svec.reserve(n_valid ? n_valid : 1);
    off.reserve(n);
    if (n_valid == 0) return;
    const int g1 = div_up(n, kBlock);
    if (!off_ready) {
        k_sv_offsets<<<g1, kBlock, 0, s>>>(rcv, mask, cnt, opt.nr, opt.nc, opt.bc, off);
        LSDTT_LAUNCH_CHECK();
    }
    // outlet starts: exclusive scan of outlet sizes in row-major order
    const auto sizes = rocprim::make_transform_iterator(outlets.p, CountOf{cnt.p});
    size_t tb = 0;
    LSDTT_HIP(rocprim::exclusive_scan(nullptr, tb, sizes, scan_out.p, 0, (size_t)n_outlets, rocprim::plus<int>(),
                                      s));
    LSDTT_HIP(rocprim::exclusive_scan(tmp.need(tb), tb, sizes, scan_out.p, 0, (size_t)n_outlets,
                                      rocprim::plus<int>(), s));
    k_sv_outlet_start<<<div_up(n_outlets, kBlock), kBlock, 0, s>>>(outlets, n_outlets, scan_out, start, svec);
    LSDTT_LAUNCH_CHECK();
    // the other cells, level by level (levels from tail_from on in one
    // workgroup)
    const SvArgs a{rcv, off, start, svec};
    const int nl = max_depth + 1;
    const int split = std::max(tail_from, 1);
    for (int L = 1; L < split && L < nl; ++L) {
        const int b = level_start_h[L], e = level_start_h[L + 1];
        if (e > b) {
            k_sv_level<<<div_up(e - b, kBlock), kBlock, 0, s>>>(order, b, e, a);
            LSDTT_LAUNCH_CHECK();
        }
    }
    if (split < nl) {
        k_sv_tail<<<1, 1024, 0, s>>>(order, level_start, split, nl - 1, a);
        LSDTT_LAUNCH_CHECK();
    }
}

}  // namespace lsdtt_gpu
