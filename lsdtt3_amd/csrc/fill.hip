// fill.hip -- epsilon-free depression filling as a minimax path problem on
// the basin graph, solved by Boruvka merging.
//
// The reference (src/cpu/lsdtt_raster_flow_preprocessing_cpu.cpp) fills at
// min_slope <= 0 either by Vincent 1993 reconstruction (:1203-1394) or by the
// priority flood (:893-920, core :416-511); its property P9 pins the two
// equal.  Both compute the spill level
//     filled(p) = z(p)                                  p an exit
//     filled(p) = max(z(p), min_{q in N8(p), valid} filled(q))   otherwise
// whose greatest solution is the minimax path height to an exit; cells no
// exit reaches keep z (the flood never visits them).  Exits (:518-589,
// :1208-1245): every valid cell on a non-periodic raster border, plus -- only
// when some interior cell is invalid -- every valid cell with an in-bounds
// invalid 8-neighbour.  The n/o/b codes are ignored; a periodic axis wraps
// neighbours (:727-782).  Only min and max of the elevations are taken, so
// the result is bit-identical whatever the evaluation order.
//
// Device scheme.  Elevations are order-preserving int keys.
//  1. Basins.  Every cell points at its valid 8-neighbour lowest in
//     (z, flat index) order when that is below its own (z, index), else at
//     itself (exits always at themselves); pointer jumping takes each cell to
//     the root of its non-ascending path.  Such a path never rises, so
//     filled(p) = max(z(p), level(root)): cells draining to an exit keep z,
//     and the cells of a pit basin share its root's level L.
//  2. Basin levels.  L is the minimax exit path over the basin graph, where
//     stepping from cell p of one basin to its neighbour q in another costs
//     w = max(z(p), z(q)) (inside a basin the non-ascending paths cost
//     nothing more).  Boruvka rounds: the exit basins form one marker
//     component that never moves; every other component takes its cheapest
//     outgoing edge (ties broken by edge id, so a round's choices form trees
//     plus mutual pairs, the lower id of a pair staying put) and hooks onto
//     the component across it.  A component with cheapest outgoing edge m
//     holds no exit, so every exit path from its cells pays at least m; and
//     inside a component the hook edges met on the way to its own cheapest
//     edge are no dearer than the largest cheapest-edge weight of the
//     components holding the cell so far.  Hence
//         L(basin b) = the largest hook weight on b's chain of hooks up to
//                      the marker,
//     computed once the rounds end by walking every basin's chain in place
//     (each round at least halves the components that have not reached the
//     marker).
//     Work per round follows the component boundaries, not the grid: the
//     first round scans every cell and lists those with a neighbour in
//     another component; later rounds scan only that list and keep the
//     cells still on a boundary.  Cells keep their basin label; the basin's
//     component is found through the hook parents (flattened every round,
//     compressed when read), so no pass relabels the grid.
//  3. Output: max(z, L(root)) for pit-basin cells whose chain reaches the
//     marker, z otherwise.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <climits>
#include <cmath>
#include <stdexcept>
#include <utility>

#include "common.hpp"

namespace lsdtt_gpu {

namespace {

constexpr int kFixed = INT_MAX;  // z key of exits and invalid cells
constexpr unsigned long long kNoEdge = ~0ull;
constexpr int kJumpRounds = 30;      // descent pointer jumping (2^30 > any grid the fill takes)
constexpr int kMaxRounds = 40;       // Boruvka rounds (each at least halves the components)
constexpr int kWalkSteps = 1 << 22;  // bound on one pointer walk (hook trees, chains): never reached
constexpr int kScanGrid = 2048;      // most blocks of a kernel over a list
constexpr int kChunkBlocks = 1024;   // blocks of a list compaction (one contiguous chunk each)
constexpr int kB = 256;
constexpr int kStrideGrid = 2048;    // blocks of the grid-stride full-grid kernels

// flag slots (ints): [0] interior invalid, [1] a pointer walk hit its bound,
// [2, 2+kJumpRounds] descent jumps, [64, 64+kMaxRounds] round go flags,
// [128, 128+kMaxRounds) live components after each round, [600] number of
// basins, [603..604] lengths of the two boundary-cell lists, [640, 640+2
// kChunkBlocks) per-chunk counts and offsets of a compaction
constexpr int kFlagStuck = 1, kFlagJump = 2, kFlagRound = 64, kFlagLive = 128, kFlagBasins = 600;
constexpr int kFlagBoundary = 603, kFlagChunks = 640;
constexpr int kFlagInts = kFlagChunks + 2 * kChunkBlocks;
static_assert(kFlagJump + kJumpRounds + 1 <= kFlagRound, "flag slots");
static_assert(kFlagRound + kMaxRounds + 1 <= kFlagLive, "flag slots");
static_assert(kFlagLive + kMaxRounds <= kFlagBasins, "flag slots");
static_assert(kFlagBoundary + 2 <= kFlagChunks, "flag slots");

__device__ __forceinline__ int fkey(float f)
{
    const int i = __float_as_int(f);
    return i >= 0 ? i : (i ^ 0x7fffffff);
}
__device__ __forceinline__ float kfloat(int k)
{
    return __int_as_float(k >= 0 ? k : (k ^ 0x7fffffff));
}
__device__ __forceinline__ bool is_nodata(float v, double nodata, bool nan_nodata)
{
    return nan_nodata ? isnan(v) : (static_cast<double>(v) == nodata);
}

// neighbour of (r, c) in direction d with periodic wrap, or -1 outside
__device__ __forceinline__ int nbr_flat(int r, int c, int d, int nr, int nc, int per_ns, int per_ew)
{
    int rr = r + d8_row[d], cc = c + d8_col[d];
    if (per_ns) rr = (rr + nr) % nr;
    if (per_ew) cc = (cc + nc) % nc;
    if (rr < 0 || rr >= nr || cc < 0 || cc >= nc) return -1;
    return rr * nc + cc;
}
// one flag store per wavefront when any lane of it sets the flag
__device__ __forceinline__ void wave_flag(bool on, int* flag)
{
    const unsigned long long m = __ballot(on);
    if (on && (int)(threadIdx.x & 63) == __ffsll(m) - 1) *flag = 1;
}
// wk: key of every valid cell (kFixed when invalid); flags an invalid cell
// off the raster border (then cells beside invalid cells are exits too)
__global__ void k_fill_init(const float* __restrict__ z, const uint8_t* __restrict__ mask, int nr, int nc,
                            double nodata, int nan_nodata, int* __restrict__ interior_invalid, int* __restrict__ wk)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y * blockDim.y + threadIdx.y;
    const bool in = r < nr && c < nc;
    bool inner_invalid = false;
    if (in) {
        const long long k = (long long)r * nc + c;
        const float zv = z[k];
        const bool valid = mask[k] && !is_nodata(zv, nodata, nan_nodata);
        wk[k] = valid ? fkey(zv) : kFixed;
        inner_invalid = !valid && r >= 1 && r < nr - 1 && c >= 1 && c < nc - 1;
    }
    wave_flag(inner_invalid, interior_invalid);
}

// Exits and descent pointers over 32 x 32 tiles staged in LDS with a
// one-cell halo.  zk: the key of non-exits, kFixed for exits and invalid
// cells.  Descent pointer: the valid 8-neighbour lowest in (key, flat
// index) order when below the cell's own (key, index), else the cell
// itself -- so equal-height plateaus chain into one basin, and (key, index)
// strictly falls along every pointer (no cycles; paths never rise in the
// key order that every min and max below uses).  Exits and invalid cells
// point at themselves.  Before leaving the tile, every pointer is followed
// inside it (pointer doubling in LDS) to the first cell outside the tile or
// to its root, so the global jumps only cross tiles.
constexpr int kDT = 32;
__global__ void __launch_bounds__(kDT * 8) k_fill_descent(const int* __restrict__ wk, int nr, int nc, int per_ns,
                                                          int per_ew, const int* __restrict__ interior_invalid,
                                                          int ntx, int tiles, int* __restrict__ zk,
                                                          int* __restrict__ ptr)
{
    constexpr int TW = kDT + 2;
    __shared__ int sk[TW * TW];
    __shared__ int sq[TW * TW];  // flat index of the staged cell, -1 outside the grid
    __shared__ int lp[kDT * kDT];
    const int tile = xcd_tile(blockIdx.x, tiles);
    const int c0 = (tile % ntx) * kDT, r0 = (tile / ntx) * kDT;
    const int tid = threadIdx.y * kDT + threadIdx.x;
    for (int i = tid; i < TW * TW; i += kDT * 8) {
        int rr = r0 - 1 + i / TW, cc = c0 - 1 + i % TW;
        if (per_ns) rr = (rr + nr) % nr;
        if (per_ew) cc = (cc + nc) % nc;
        const bool ok = rr >= 0 && rr < nr && cc >= 0 && cc < nc;
        const int q = ok ? rr * nc + cc : -1;
        sq[i] = q;
        sk[i] = ok ? wk[q] : kFixed;
    }
    __syncthreads();
    const bool inv_inside = *interior_invalid != 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int ty = threadIdx.y + 8 * j, tx = threadIdx.x;
        const int r = r0 + ty, c = c0 + tx;
        const int li = (ty + 1) * TW + tx + 1;
        const int t = ty * kDT + tx;
        if (r >= nr || c >= nc) {
            lp[t] = t;
            continue;
        }
        const int k = r * nc + c;
        const int key = sk[li];
        int best = k;
        if (key != kFixed) {
            bool exit_cell = (!per_ns && (r == 0 || r == nr - 1)) || (!per_ew && (c == 0 || c == nc - 1));
            if (!exit_cell && inv_inside) {
                // an in-bounds invalid neighbour (no wrap for this rule)
                for (int d = 0; d < 8 && !exit_cell; ++d) {
                    const int rr = r + d8_row[d], cc = c + d8_col[d];
                    if (rr < 0 || rr >= nr || cc < 0 || cc >= nc) continue;
                    exit_cell = sk[li + d8_row[d] * TW + d8_col[d]] == kFixed;
                }
            }
            zk[k] = exit_cell ? kFixed : key;
            if (!exit_cell) {
                int bk = key;
                for (int d = 0; d < 8; ++d) {
                    const int lq = li + d8_row[d] * TW + d8_col[d];
                    const int q = sq[lq];
                    if (q < 0) continue;
                    const int kq = sk[lq];
                    if (kq == kFixed) continue;  // invalid
                    if (kq < bk || (kq == bk && q < best)) {
                        bk = kq;
                        best = q;
                    }
                }
            }
        } else {
            zk[k] = kFixed;
        }
        // local index of the target when it lies in the tile, else -(flat + 1)
        const int br = best / nc, bc = best - br * nc;
        const bool inside = br >= r0 && br < r0 + kDT && bc >= c0 && bc < c0 + kDT;
        lp[t] = inside ? (br - r0) * kDT + (bc - c0) : -(best + 1);
    }
    __syncthreads();
    // pointer doubling inside the tile: 10 rounds cover any chain of its
    // 1024 cells; a pointer read while its owner updates it is still a cell
    // of the same path
    for (int it = 0; it < 10; ++it) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int t = (threadIdx.y + 8 * j) * kDT + threadIdx.x;
            const int p = lp[t];
            if (p >= 0 && p != t) lp[t] = lp[p];
        }
        __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int ty = threadIdx.y + 8 * j, tx = threadIdx.x;
        const int r = r0 + ty, c = c0 + tx;
        if (r >= nr || c >= nc) continue;
        const int p = lp[ty * kDT + tx];
        ptr[r * nc + c] = p >= 0 ? (r0 + p / kDT) * nc + (c0 + p % kDT) : -(p + 1);
    }
}

// ptr[k] <- ptr[ptr[k]] until every pointer reaches a root (in place: a
// pointer only ever moves further down its own descent path); grid-stride,
// so that a launch after the last change costs little
__global__ void k_fill_jump(int* __restrict__ ptr, long long n, const int* __restrict__ prev, int* __restrict__ cur)
{
    if (*prev == 0) return;
    bool ch = false;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < n; k += stride) {
        const int p = __hip_atomic_load(&ptr[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const int q = __hip_atomic_load(&ptr[p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (q != p) {
            __hip_atomic_store(&ptr[k], q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            ch = true;
        }
    }
    wave_flag(ch, cur);
}

__device__ __forceinline__ int ld_rel(const int* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_rel(int* p, int v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void k_zero_ints(int* __restrict__ p, int count)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) p[i] = 0;
}

// Labels, in place over the descent roots: the root (a pit basin's cell),
// X (= n, the marker) when the root is an exit, -1 when invalid.  Also
// counts the pit basins (lab[k] == k) of each contiguous chunk, the first
// launch of their numbering.
__global__ void __launch_bounds__(kB) k_bv_label(const int* __restrict__ zk, const int* __restrict__ wk,
                                                 int* __restrict__ lab, long long n, int* __restrict__ chunk_count)
{
    const long long len = (n + gridDim.x - 1) / gridDim.x;
    const long long b0 = blockIdx.x * len, b1 = std::min(n, b0 + len);
    int c = 0;
    for (long long k0 = b0 + threadIdx.x; k0 < b1; k0 += 4ll * blockDim.x) {
        int w[4], r[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const long long k = k0 + j * (long long)blockDim.x;
            w[j] = k < b1 ? wk[k] : kFixed;
            r[j] = k < b1 ? lab[k] : 0;
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const long long k = k0 + j * (long long)blockDim.x;
            if (k >= b1) continue;
            int v = -1;
            if (w[j] != kFixed) v = zk[r[j]] == kFixed ? static_cast<int>(n) : r[j];
            lab[k] = v;
            c += v == k;
        }
    }
    __shared__ int red[kB / 64];
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int wv = 0; wv < kB / 64; ++wv) t += red[wv];
        chunk_count[blockIdx.x] = t;
    }
}

// exclusive scan of the chunk counts (one block of kChunkBlocks threads);
// the total is the new list's length
__global__ void __launch_bounds__(kChunkBlocks) k_list_scan(const int* __restrict__ chunk_count,
                                                            int* __restrict__ chunk_off, int* __restrict__ total,
                                                            const int* __restrict__ go)
{
    if (go && *go == 0) return;
    __shared__ int sc[kChunkBlocks];
    const int t = threadIdx.x;
    sc[t] = chunk_count[t];
    __syncthreads();
    for (int o = 1; o < kChunkBlocks; o <<= 1) {
        const int a = t >= o ? sc[t - o] : 0;
        __syncthreads();
        sc[t] += a;
        __syncthreads();
    }
    chunk_off[t] = sc[t] - chunk_count[t];
    if (t == kChunkBlocks - 1) *total = sc[t];
}

// Compactions in three launches (a count per contiguous chunk, one scan,
// the writes in order), so no counter is shared by the whole grid.
// kRanks: over the cells, the pit basins' roots (lab[c] == c) numbered in
// flat order (rank_of[root] = its number); kKeepCells / kKeepList: the
// cells / list entries whose keep byte is set, into a list.
enum { kRanks = 0, kKeepCells = 2, kKeepList = 3 };

template <int MODE>
__device__ __forceinline__ int list_item(int i, const int* __restrict__ lab, const int* __restrict__ list,
                                         const uint8_t* __restrict__ keep)
{
    if (MODE == kRanks) return lab[i] == i ? i : -1;
    if (MODE == kKeepCells) return keep[i] ? i : -1;
    return keep[i] ? list[i] : -1;
}

template <int MODE>
__global__ void __launch_bounds__(kB) k_list_count(const int* __restrict__ list, const uint8_t* __restrict__ keep,
                                                   long long m_fixed, const int* __restrict__ m_dev,
                                                   int* __restrict__ chunk_count, const int* __restrict__ go)
{
    if (go && *go == 0) return;
    const long long m = m_dev ? *m_dev : m_fixed;
    const long long len = (m + gridDim.x - 1) / gridDim.x;
    const long long b0 = blockIdx.x * len, b1 = std::min(m, b0 + len);
    int c = 0;
    // four items per thread per step, their loads issued together
    for (long long i = b0 + threadIdx.x; i < b1; i += 4 * blockDim.x) {
        int v[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const long long ij = i + j * (long long)blockDim.x;
            v[j] = ij < b1 ? (keep[ij] ? 1 : -1) : -1;
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) c += v[j] >= 0;
    }
    __shared__ int red[kB / 64];
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < kB / 64; ++w) t += red[w];
        chunk_count[blockIdx.x] = t;
    }
}

// kRanks also starts each basin as its own component with no cheapest edge
template <int MODE>
__global__ void __launch_bounds__(kB) k_list_write(const int* __restrict__ lab, const int* __restrict__ list,
                                                   const uint8_t* __restrict__ keep, long long m_fixed,
                                                   const int* __restrict__ m_dev, const int* __restrict__ chunk_off,
                                                   int* __restrict__ out, int* __restrict__ par,
                                                   unsigned long long* __restrict__ best, const int* __restrict__ go)
{
    if (go && *go == 0) return;
    const long long m = m_dev ? *m_dev : m_fixed;
    const long long len = (m + gridDim.x - 1) / gridDim.x;
    const long long b0 = blockIdx.x * len, b1 = std::min(m, b0 + len);
    __shared__ int wsum[kB / 64];
    int base = chunk_off[blockIdx.x];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // four rows of blockDim.x items per step, their loads issued together,
    // then written row by row in order
    for (long long i0 = b0; i0 < b1; i0 += 4 * blockDim.x) {
        int v[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const long long i = i0 + j * (long long)blockDim.x + threadIdx.x;
            v[j] = i < b1 ? list_item<MODE>((int)i, lab, list, keep) : -1;
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const unsigned long long bal = __ballot(v[j] >= 0);
            if (lane == 0) wsum[wave] = __popcll(bal);
            __syncthreads();
            int before = 0, all = 0;
            for (int w = 0; w < kB / 64; ++w) {
                before += w < wave ? wsum[w] : 0;
                all += wsum[w];
            }
            if (v[j] >= 0) {
                const int pos = base + before + __popcll(bal & ((1ull << lane) - 1ull));
                if (MODE == kRanks) {
                    out[v[j]] = pos;
                    par[pos] = pos;
                    best[pos] = kNoEdge;
                } else {
                    out[pos] = v[j];
                }
            }
            base += all;
            __syncthreads();
        }
    }
}

// labels become basin numbers: a component is named by the number of the
// basin it started from, so every per-component array is dense
__global__ void k_relabel(int* __restrict__ lab, long long n, int X, const int* __restrict__ rank_of)
{
    const long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int l = lab[k];
    if (l >= 0 && l != X) lab[k] = rank_of[l];
}

// the key of the edge from cell k (key kp) to its neighbour q in direction d
// (key kq): (w key with the sign bit flipped) << 32 | edge id, the edge id
// min(k, q) * 8 + direction from the lower to the higher index
__device__ __forceinline__ unsigned long long edge_key(int k, int q, int d, int kp, int kq)
{
    const unsigned uw = static_cast<unsigned>(max(kp, kq)) ^ 0x80000000u;
    const unsigned eid = k < q ? static_cast<unsigned>(k) * 8u + d : static_cast<unsigned>(q) * 8u + ((d + 4) & 7);
    return (static_cast<unsigned long long>(uw) << 32) | eid;
}

// Offers of a block's cells (every thread calls it with its index in the
// block; wavefronts are runs of 64 consecutive indices).  Within a
// wavefront, runs of one component fold their minima towards the run's
// first lane; the run heads then meet in an LDS table keyed by component,
// so the block issues one global atomic per component it saw (large
// components late in the rounds would otherwise take one per run from
// every wavefront)
__device__ __forceinline__ void offer_block(int A, unsigned long long m, unsigned long long* __restrict__ best,
                                            int tid)
{
    __shared__ int tkey[kB];
    __shared__ unsigned long long tval[kB];
    const int lane = tid & 63;
    tkey[tid] = -1;
    tval[tid] = kNoEdge;
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned long long om = __shfl_down(m, o);
        const int oa = __shfl_down(A, o);
        if (lane + o < 64 && oa == A && om < m) m = om;
    }
    const int left = __shfl_up(A, 1);
    const bool head = (lane == 0 || left != A) && A >= 0 && m != kNoEdge;
    __syncthreads();
    if (head) {
        // at most one head per thread, so the table never fills
        int slot = static_cast<int>((static_cast<unsigned>(A) * 2654435761u) >> 24) & (kB - 1);
        for (int probe = 0; probe < kB; ++probe) {
            const int prev = atomicCAS(&tkey[slot], -1, A);
            if (prev == -1 || prev == A) {
                atomicMin(&tval[slot], m);
                break;
            }
            slot = (slot + 1) & (kB - 1);
        }
    }
    __syncthreads();
    const int key = tkey[tid];
    const unsigned long long v = tval[tid];
    if (key >= 0 && v < best[key]) atomicMin(&best[key], v);
    __syncthreads();
}

// Early rounds, over the whole grid: each cell of a non-marker component
// offers its cheapest edge to another component, and (keep_out set) the
// cells with such an edge are flagged for the boundary list.  A block covers
// 64 x 4 cells; their components and keys, with a one-cell halo (periodic
// axes wrap; -1 outside the grid), are staged in LDS first.  In the first
// round the components are the basins themselves (no hook yet), so the
// labels are read directly; later a cell's component is its basin's parent.
constexpr int kScanW = 64, kScanH = 4;
template <bool VIA_PARENT>
__global__ void __launch_bounds__(kScanW * kScanH) k_bv_scan(const int* __restrict__ wk, const int* __restrict__ lab,
                                                             const int* __restrict__ par, int nr, int nc, int per_ns,
                                                             int per_ew, int X, unsigned long long* __restrict__ best,
                                                             uint8_t* __restrict__ keep_out, int ntx, int tiles,
                                                             const int* __restrict__ go)
{
    if (go && *go == 0) return;
    constexpr int TW = kScanW + 2, TH = kScanH + 2;
    __shared__ int sl[TH * TW], sk[TH * TW];
    const int tile = xcd_tile(blockIdx.x, tiles);
    const int c0 = (tile % ntx) * kScanW, r0 = (tile / ntx) * kScanH;
    const int tid = threadIdx.y * kScanW + threadIdx.x;
    for (int i = tid; i < TH * TW; i += kScanW * kScanH) {
        int rr = r0 - 1 + i / TW, cc = c0 - 1 + i % TW;
        if (per_ns) rr = (rr + nr) % nr;
        if (per_ew) cc = (cc + nc) % nc;
        const bool ok = rr >= 0 && rr < nr && cc >= 0 && cc < nc;
        const int q = ok ? rr * nc + cc : 0;
        int l = ok ? lab[q] : -1;
        if (VIA_PARENT && l >= 0) l = par[l];
        sl[i] = l;
        sk[i] = ok ? wk[q] : 0;
    }
    __syncthreads();
    const int c = c0 + threadIdx.x, r = r0 + threadIdx.y;
    const bool in = r < nr && c < nc;
    const int k = in ? r * nc + c : 0;
    const int li = (threadIdx.y + 1) * TW + threadIdx.x + 1;
    int A = in ? sl[li] : -1;
    if (A == X) A = -1;
    unsigned long long m = kNoEdge;
    bool keep = false;
    if (A >= 0) {
        const int kp = sk[li];
        for (int d = 0; d < 8; ++d) {
            const int lq = li + d8_row[d] * TW + d8_col[d];
            const int B = sl[lq];
            if (B < 0 || B == A) continue;
            keep = true;
            const int q = nbr_flat(r, c, d, nr, nc, per_ns, per_ew);
            const unsigned long long key = edge_key(k, q, d, kp, sk[lq]);
            m = key < m ? key : m;
        }
    }
    if (keep_out && in) keep_out[k] = keep;
    offer_block(A, m, best, tid);
}

// Later rounds: the cells of the boundary list offer their cheapest edges
// to other components; those that still have one are flagged to stay on
// the list.  A cell's component is its basin's parent; the loads of all
// nine cells are issued together, one step at a time.
__global__ void __launch_bounds__(kB) k_bv_scan_list(const int* __restrict__ wk, const int* __restrict__ lab,
                                                     const int* __restrict__ par, int nr, int nc, int per_ns,
                                                     int per_ew, int X, const int* __restrict__ list,
                                                     const int* __restrict__ count, uint8_t* __restrict__ keep_out,
                                                     unsigned long long* __restrict__ best,
                                                     const int* __restrict__ go)
{
    if (*go == 0) return;
    const int cnt = *count;
    const int stride = gridDim.x * blockDim.x;
    for (int i0 = blockIdx.x * blockDim.x; i0 < cnt; i0 += stride) {
        const int i = i0 + threadIdx.x;
        const bool in = i < cnt;
        const int k = in ? list[i] : 0;
        const int r = k / nc, c = k - r * nc;
        int q[8], lq[8], kq[8];
#pragma unroll
        for (int d = 0; d < 8; ++d) q[d] = in ? nbr_flat(r, c, d, nr, nc, per_ns, per_ew) : -1;
#pragma unroll
        for (int d = 0; d < 8; ++d) {
            lq[d] = q[d] >= 0 ? lab[q[d]] : -1;
            kq[d] = q[d] >= 0 ? wk[q[d]] : 0;
        }
        const int lk = in ? lab[k] : -1;
        const int kp = in ? wk[k] : 0;
        int A = -1;
        unsigned long long m = kNoEdge;
        bool keep = false;
        if (lk >= 0) {
            int cq[8];
            A = par[lk];
#pragma unroll
            for (int d = 0; d < 8; ++d) cq[d] = par[lq[d] >= 0 ? lq[d] : lk];
            if (A == X) {
                A = -1;
            } else {
#pragma unroll
                for (int d = 0; d < 8; ++d) {
                    if (lq[d] < 0 || cq[d] == A) continue;
                    keep = true;
                    const unsigned long long key = edge_key(k, q[d], d, kp, kq[d]);
                    m = key < m ? key : m;
                }
            }
        }
        if (in) keep_out[i] = keep;
        offer_block(A, m, best, threadIdx.x);
    }
}

// Each component still in the rounds (its own parent) with an outgoing edge
// records the component across it; targets are read before any component of
// the round hooks
__global__ void k_bv_target(const int* __restrict__ lab, const int* __restrict__ par,
                            const unsigned long long* __restrict__ best, int nr, int nc, int per_ns, int per_ew,
                            const int* __restrict__ nb_dev, int* __restrict__ hp, const int* __restrict__ go)
{
    if (*go == 0) return;
    const int nb = *nb_dev;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nb; c += gridDim.x * blockDim.x) {
        if (par[c] != c) continue;
        const unsigned long long b = best[c];
        if (b == kNoEdge) continue;
        const unsigned eid = static_cast<unsigned>(b);
        const int p = static_cast<int>(eid >> 3), d = static_cast<int>(eid & 7);
        const int q = nbr_flat(p / nc, p % nc, d, nr, nc, per_ns, per_ew);
        const int A = par[lab[p]], B = par[lab[q]];
        hp[c] = A == c ? B : A;
    }
}

// Each such component hooks onto its target, except the lower number of a
// pair that chose the same edge (only a component's own thread reads or
// writes its parent here)
__global__ void k_bv_hook(const unsigned long long* __restrict__ best, int X, const int* __restrict__ nb_dev,
                          int* __restrict__ par, const int* __restrict__ hp, int* __restrict__ hw,
                          const int* __restrict__ go, int* __restrict__ hooked)
{
    if (*go == 0) return;
    const int nb = *nb_dev;
    const int stride = gridDim.x * blockDim.x;
    for (int c0 = blockIdx.x * blockDim.x; c0 < nb; c0 += stride) {
        const int c = c0 + threadIdx.x;
        bool h = false;
        if (c < nb && par[c] == c) {
            const unsigned long long b = best[c];
            if (b != kNoEdge) {
                const int t = hp[c];
                const bool stays = t != X && best[t] == b && c < t;
                if (!stays) {
                    par[c] = t;
                    hw[c] = static_cast<int>(static_cast<unsigned>(b >> 32) ^ 0x80000000u);
                    h = true;
                }
            }
        }
        wave_flag(h, hooked);
    }
}

// Every basin's parent made its component, in place: components still in
// the rounds and the marker are their own parents, every other parent lies
// on the way to its component and only ever moves towards it, so walks that
// overlap stay correct (this also flattens the round's hook trees).  The
// scans and hook targets then read a basin's component as par[basin].
// Components still in the rounds clear their cheapest edge for the next
// round and are counted.
__global__ void __launch_bounds__(kB) k_bv_compress(int* __restrict__ par, unsigned long long* __restrict__ best,
                                                    const int* __restrict__ nb_dev, const int* __restrict__ go,
                                                    int* __restrict__ live, int* __restrict__ stuck)
{
    if (*go == 0) return;
    const int nb = *nb_dev;
    int alive = 0;
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += gridDim.x * blockDim.x) {
        const int a = ld_rel(&par[b]);
        if (a == b) {
            best[b] = kNoEdge;
            ++alive;
            continue;
        }
        int p = ld_rel(&par[a]);
        if (p == a) continue;
        for (int s = 0;; ++s) {
            const int q = ld_rel(&par[p]);
            if (q == p) break;
            if (s == kWalkSteps) {
                *stuck = 1;
                break;
            }
            p = q;
        }
        st_rel(&par[b], p);
    }
    __shared__ int red[kB / 64];
    for (int o = 32; o > 0; o >>= 1) alive += __shfl_xor(alive, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = alive;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < kB / 64; ++w) t += red[w];
        if (t) atomicAdd(live, t);
    }
}

// Hook chains as packed words, one per basin: hook weight key << 32 | hook
// parent, or parent -1 when the basin's component never hooked (it is
// still its own parent)
__global__ void k_chain_pack(const int* __restrict__ nb_dev, const int* __restrict__ par,
                             const int* __restrict__ hp, const int* __restrict__ hw,
                             unsigned long long* __restrict__ P)
{
    const int nb = *nb_dev;
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += gridDim.x * blockDim.x) {
        P[b] = par[b] == b ? 0xffffffffull
                           : (static_cast<unsigned long long>(static_cast<unsigned>(hw[b])) << 32) |
                                 static_cast<unsigned>(hp[b]);
    }
}

// One step of the chain maxima below for every basin (pointer doubling when
// all basins step together); the walk then finishes the few long chains
__global__ void k_chain_step(const int* __restrict__ nb_dev, unsigned long long* __restrict__ P, int X)
{
    const int nb = *nb_dev;
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += gridDim.x * blockDim.x) {
        const unsigned long long w = __hip_atomic_load(&P[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const int u = static_cast<int>(static_cast<unsigned>(w));
        if (u < 0 || u == X) continue;
        const unsigned long long wu = __hip_atomic_load(&P[u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const int m = max(static_cast<int>(w >> 32), static_cast<int>(wu >> 32));
        __hip_atomic_store(&P[b], (static_cast<unsigned long long>(static_cast<unsigned>(m)) << 32) |
                                      (wu & 0xffffffffull),
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Chain maxima in place: a basin's word (m, u) becomes (max(m, m_u), u_u)
// from its parent u's word until u is the marker or hooks nowhere.  Every
// stored word is (the largest hook weight from the basin up to u, u), so
// words read while their owners update them still compose correctly.
__global__ void k_chain_walk(const int* __restrict__ nb_dev, unsigned long long* __restrict__ P, int X,
                             int* __restrict__ stuck)
{
    const int nb = *nb_dev;
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += gridDim.x * blockDim.x) {
        unsigned long long w = __hip_atomic_load(&P[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        for (int s = 0;; ++s) {
            const int u = static_cast<int>(static_cast<unsigned>(w));
            if (u < 0 || u == X) break;
            if (s == kWalkSteps) {
                *stuck = 1;
                break;
            }
            const unsigned long long wu = __hip_atomic_load(&P[u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const int m = max(static_cast<int>(w >> 32), static_cast<int>(wu >> 32));
            w = (static_cast<unsigned long long>(static_cast<unsigned>(m)) << 32) | (wu & 0xffffffffull);
            __hip_atomic_store(&P[b], w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// filled = max(z, L(basin)) in a pit basin whose chain reached the marker,
// z otherwise (its key read back: the key map is a bijection on float bits);
// nodata for invalid cells
__global__ void k_fill_out(const int* __restrict__ wk, const int* __restrict__ lab,
                           const unsigned long long* __restrict__ P, long long n, int X, float nodata_f,
                           float* __restrict__ out)
{
    const long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int key = wk[k];
    if (key == kFixed) {
        out[k] = nodata_f;
        return;
    }
    const int b = lab[k];
    int v = key;
    if (b != X) {
        const unsigned long long w = P[b];
        if (static_cast<int>(static_cast<unsigned>(w)) == X) v = max(key, static_cast<int>(w >> 32));
    }
    out[k] = kfloat(v);
}

// the flags whose launches run unconditionally (the first descent jump and
// the first round), and the marker's slot: its own component, no edge
__global__ void k_fill_start_flags(int* __restrict__ flags, int X, int* __restrict__ par,
                                   unsigned long long* __restrict__ best)
{
    const int i = threadIdx.x;
    if (i == 0) flags[kFlagJump] = 1;
    if (i == 1) flags[kFlagRound] = 1;
    if (i == 2) par[X] = X;
    if (i == 3) best[X] = kNoEdge;
}

// descent jumps the last fill of this process needed, plus a spare: the
// length of the next fill's first batch (a hint only; more follow when it
// falls short)
std::atomic<int> g_jump_hint{12};

}  // namespace

// ints of scratch: the cheapest edges (2 per slot, first, 8-byte aligned),
// zk (then each root's basin number), wk, the labels, par, hp, hw, two
// boundary-cell lists, a keep byte per cell and the flags
size_t fill_scratch_ints(int nr, int nc)
{
    const long long n = (long long)nr * nc;
    return static_cast<size_t>(2 * (n + 1) + 3 * n + 3 * (n + 1) + 2 * n + (n + 3) / 4 + kFlagInts + 8);
}

// Enqueues the fill on `s` and returns the number of Boruvka rounds run.
// `scratch` holds fill_scratch_ints(nr, nc) ints, 8-byte aligned.  Blocks on
// one read-back after the descent jumps, one per batch of rounds and one at
// the end.
int fill_device(const float* z, const uint8_t* mask, int nr, int nc, double nodata, BC bc, float* out,
                int* scratch, hipStream_t s)
{
    const int per_ns = bc.n == 'p', per_ew = bc.e == 'p';
    const bool nan_nodata = std::isnan(nodata);
    const long long n = (long long)nr * nc;
    if (n >= (1ll << 29)) throw std::invalid_argument("the fill supports grids of fewer than 2^29 cells");
    const int X = static_cast<int>(n);
    // per-component arrays are indexed by basin number (dense, < number of
    // basins) and by X for the marker
    unsigned long long* best = reinterpret_cast<unsigned long long*>(scratch);
    int* zk = scratch + 2 * (n + 1);
    int* rank_of = zk;  // once the labels are made
    int* wk = zk + n;
    int* lab = wk + n;  // descent pointers, then labels, then basin numbers
    int* par = lab + n;
    int* hp = par + (n + 1);
    int* hw = hp + (n + 1);
    int* blist[2] = {hw + (n + 1), hw + (n + 1) + n};
    uint8_t* keep = reinterpret_cast<uint8_t*>(blist[1] + n);
    int* flags = blist[1] + n + (n + 3) / 4;
    int* nb_dev = flags + kFlagBasins;
    int* bcount = flags + kFlagBoundary;  // boundary lists
    dim3 b2(64, 4), g2(div_up(nc, 64), div_up(nr, 4));
    const int g1 = div_up(n, kB);
    const int gs = std::min(g1, kStrideGrid);  // grid-stride launches
    const dim3 bsc(kScanW, kScanH);
    const int scan_ntx = div_up(nc, kScanW), scan_tiles = scan_ntx * div_up(nr, kScanH);
    k_zero_ints<<<div_up(kFlagInts, kB), kB, 0, s>>>(flags, kFlagInts);
    k_fill_init<<<g2, b2, 0, s>>>(z, mask, nr, nc, nodata, nan_nodata, flags, wk);
    LSDTT_LAUNCH_CHECK();
    {
        const int ntx = div_up(nc, kDT), tiles = ntx * div_up(nr, kDT);
        k_fill_descent<<<tiles, dim3(kDT, 8), 0, s>>>(wk, nr, nc, per_ns, per_ew, flags, ntx, tiles, zk, lab);
    }
    LSDTT_LAUNCH_CHECK();
    k_fill_start_flags<<<1, 64, 0, s>>>(flags, X, par, best);
    LSDTT_LAUNCH_CHECK();
    // descent jumps: a first batch as long as the last fill needed (plus a
    // spare), the rest only when its last jump still moved a pointer
    int fh[kFlagChunks];
    auto read_flags = [&]() {
        LSDTT_HIP(hipMemcpyAsync(fh, flags, sizeof(fh), hipMemcpyDeviceToHost, s));
        LSDTT_HIP(hipStreamSynchronize(s));
    };
    auto jumps = [&](int r0, int r1) {
        for (int r = r0; r < r1; ++r) {
            k_fill_jump<<<gs, kB, 0, s>>>(lab, n, flags + kFlagJump + r, flags + kFlagJump + r + 1);
            LSDTT_LAUNCH_CHECK();
        }
    };
    const int first_jumps = std::min(kJumpRounds, g_jump_hint.load(std::memory_order_relaxed));
    jumps(0, first_jumps);
    read_flags();
    if (fh[kFlagJump + first_jumps]) {
        jumps(first_jumps, kJumpRounds);
        read_flags();
        if (fh[kFlagJump + kJumpRounds]) throw std::runtime_error("fill: descent pointers did not converge");
    }
    {
        int moved = 0;
        while (moved < kJumpRounds && fh[kFlagJump + moved + 1]) ++moved;
        g_jump_hint.store(moved + 2, std::memory_order_relaxed);
    }
    // labels, then the pit basins numbered in flat order (each starting as
    // its own component), then every label as its basin's number
    int* chunk_count = flags + kFlagChunks;
    int* chunk_off = chunk_count + kChunkBlocks;
    k_bv_label<<<kChunkBlocks, kB, 0, s>>>(zk, wk, lab, n, chunk_count);
    k_list_scan<<<1, kChunkBlocks, 0, s>>>(chunk_count, chunk_off, nb_dev, nullptr);
    k_list_write<kRanks><<<kChunkBlocks, kB, 0, s>>>(lab, nullptr, nullptr, n, nullptr, chunk_off, rank_of, par, best,
                                                     nullptr);
    k_relabel<<<g1, kB, 0, s>>>(lab, n, X, rank_of);
    LSDTT_LAUNCH_CHECK();
    // Boruvka rounds in batches with one read-back of their flags; a round
    // after the last hook returns at once.  Every round: the scan offers
    // cheapest edges, then targets, hooks and the compression of the
    // parents over all basins.  The scans of rounds 0 and 1 cover the grid
    // (most cells are still on a boundary); from round 1 on a scan flags the
    // cells that stay on a boundary and compacts them, in order, into
    // boundary list r & 1, which the next round scans instead of the grid.
    // After a batch the live components size the next: each round at least
    // halves the components that can still hook, so at most log2(live) + 1
    // more rounds hook, and the grids follow the lengths.
    auto grid_for = [](long long items) {
        return static_cast<int>(std::clamp<long long>(div_up(items, kB), 1, kScanGrid));
    };
    int basin_grid = kScanGrid, scan_grid = kScanGrid;
    int launched = 0, rounds = 0, batch = 4;
    for (;;) {
        const int end = std::min(kMaxRounds, launched + batch);
        for (int r = launched; r < end; ++r) {
            const int* go = flags + kFlagRound + r;
            int* hooked = flags + kFlagRound + r + 1;
            int* b_out = blist[r & 1];
            int* b_out_count = bcount + (r & 1);
            if (r == 0) {
                k_bv_scan<false><<<scan_tiles, bsc, 0, s>>>(wk, lab, par, nr, nc, per_ns, per_ew, X, best, nullptr,
                                                             scan_ntx, scan_tiles, nullptr);
            } else if (r == 1) {
                k_bv_scan<true><<<scan_tiles, bsc, 0, s>>>(wk, lab, par, nr, nc, per_ns, per_ew, X, best, keep, scan_ntx,
                                                            scan_tiles, go);
                k_list_count<kKeepCells><<<kChunkBlocks, kB, 0, s>>>(nullptr, keep, n, nullptr, chunk_count, go);
                k_list_scan<<<1, kChunkBlocks, 0, s>>>(chunk_count, chunk_off, b_out_count, go);
                k_list_write<kKeepCells><<<kChunkBlocks, kB, 0, s>>>(nullptr, nullptr, keep, n, nullptr, chunk_off,
                                                                     b_out, nullptr, nullptr, go);
            } else {
                const int* b_in = blist[(r + 1) & 1];
                const int* b_in_count = bcount + ((r + 1) & 1);
                k_bv_scan_list<<<scan_grid, kB, 0, s>>>(wk, lab, par, nr, nc, per_ns, per_ew, X, b_in, b_in_count,
                                                         keep, best, go);
                k_list_count<kKeepList><<<kChunkBlocks, kB, 0, s>>>(b_in, keep, 0, b_in_count, chunk_count, go);
                k_list_scan<<<1, kChunkBlocks, 0, s>>>(chunk_count, chunk_off, b_out_count, go);
                k_list_write<kKeepList><<<kChunkBlocks, kB, 0, s>>>(nullptr, b_in, keep, 0, b_in_count, chunk_off,
                                                                    b_out, nullptr, nullptr, go);
            }
            k_bv_target<<<basin_grid, kB, 0, s>>>(lab, par, best, nr, nc, per_ns, per_ew, nb_dev, hp, go);
            k_bv_hook<<<basin_grid, kB, 0, s>>>(best, X, nb_dev, par, hp, hw, go, hooked);
            k_bv_compress<<<basin_grid, kB, 0, s>>>(par, best, nb_dev, go, flags + kFlagLive + r, flags + kFlagStuck);
            LSDTT_LAUNCH_CHECK();
        }
        read_flags();
        if (fh[kFlagStuck]) throw std::runtime_error("fill: a hook tree deeper than the walk bound");
        const int* h = fh + kFlagRound;
        launched = end;
        if (!h[end]) {
            rounds = 0;
            while (rounds < end && h[rounds + 1]) ++rounds;
            break;
        }
        if (launched >= kMaxRounds) throw std::runtime_error("fill: Boruvka rounds did not finish");
        const int live = fh[kFlagLive + end - 1];
        int lg = 0;
        while ((1 << (lg + 1)) <= live) ++lg;
        batch = std::min(4, lg + 2);
        basin_grid = grid_for(fh[kFlagBasins]);
        scan_grid = grid_for(fh[kFlagBoundary + ((end - 1) & 1)]);
    }
    // chain maxima over the basins, in the cheapest-edge words (free now)
    unsigned long long* P = best;
    const int chain_grid = grid_for(fh[kFlagBasins]);
    k_chain_pack<<<chain_grid, kB, 0, s>>>(nb_dev, par, hp, hw, P);
    for (int j = 0; j < 5; ++j) k_chain_step<<<chain_grid, kB, 0, s>>>(nb_dev, P, X);
    k_chain_walk<<<chain_grid, kB, 0, s>>>(nb_dev, P, X, flags + kFlagStuck);
    LSDTT_LAUNCH_CHECK();
    k_fill_out<<<g1, kB, 0, s>>>(wk, lab, P, n, X, static_cast<float>(nodata), out);
    LSDTT_LAUNCH_CHECK();
    int stuck = 0;
    LSDTT_HIP(hipMemcpyAsync(&stuck, flags + kFlagStuck, sizeof(int), hipMemcpyDeviceToHost, s));
    LSDTT_HIP(hipStreamSynchronize(s));
    if (stuck) throw std::runtime_error("fill: a hook chain longer than the walk bound");
    return rounds;
}

}  // namespace lsdtt_gpu
