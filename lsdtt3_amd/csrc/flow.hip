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
                                                double card, double diag, int nodata_outlet, int pits_outlet,
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
        const double sl = dz / ((d & 1) ? diag : card);
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
                                                   double card, double diag, int nodata_outlet, int pits_outlet,
                                                   int* __restrict__ rcv, uint8_t* __restrict__ dir,
                                                   int* __restrict__ fail, T* __restrict__ zcopy, TileMark mark)
{
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    const int r0 = (blockIdx.y * blockDim.y + threadIdx.y) * kRecvRows;
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
        decide_interior<T>(forced, nc, row, col, card, diag, nodata_outlet, pits_outlet, mw[k][1], zw[k][1], mk, zn,
                           rcv, dir, fail, mark, oldr[k]);
    }
}

// the outer ring (2 nc + 2 (nr - 2) cells, fewer on 1- and 2-wide grids):
// every step is resolved against the boundary conditions
template <class T>
__global__ void __launch_bounds__(256) k_receivers_ring(const T* __restrict__ z, const uint8_t* __restrict__ mask,
                                                        const uint8_t* __restrict__ forced, int nr, int nc,
                                                        double card, double diag, uint32_t bcw, int nodata_outlet,
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
        const double sl = dz / ((d & 1) ? diag : card);
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
    const dim3 b2(64, 4), g2(div_up(nc, 64), div_up(nr, 4 * kRecvRows));
    k_receivers<T><<<g2, b2, 0, s>>>(z, msk, forced, nr, nc, dx, dx * kSqrt2, nodata_outlet, pits_outlet, rcv, dir,
                                     fail, zcopy, mark);
    LSDTT_LAUNCH_CHECK();
    const int nring = nc + (nr > 1 ? nc : 0) + (nr > 2 ? (nc > 1 ? 2 : 1) * (nr - 2) : 0);
    k_receivers_ring<T><<<div_up(nring, 256), 256, 0, s>>>(z, msk, forced, nr, nc, dx, dx * kSqrt2, bc_pack(bc),
                                                          nodata_outlet, pits_outlet, rcv, dir, fail, zcopy, mark);
    LSDTT_LAUNCH_CHECK();
}

// donor count = number of distinct valid neighbours whose receiver is f
__global__ void k_donors(const int* __restrict__ rcv, const uint8_t* __restrict__ mask, int nr, int nc,
                         BC bc, uint8_t* __restrict__ donors, int* __restrict__ outlet_flag)
{
    const long long f = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (f >= (long long)nr * nc) return;
    if (!mask[f]) {
        donors[f] = 0;
        outlet_flag[f] = 0;
        return;
    }
    int nb[8];
    const int k = distinct_neighbours(static_cast<int>(f), nr, nc, bc, nb);
    int c = 0;
    if (k == 8) {
        int rv[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) rv[i] = rcv[nb[i]];  // masked cells hold -1, never == f
#pragma unroll
        for (int i = 0; i < 8; ++i) c += (rv[i] == f) ? 1 : 0;
    } else {
        for (int i = 0; i < k; ++i) c += (mask[nb[i]] && rcv[nb[i]] == f) ? 1 : 0;
    }
    donors[f] = static_cast<uint8_t>(c);
    outlet_flag[f] = (rcv[f] == kOutletReceiver) ? 1 : 0;
}

__global__ void k_scatter_flagged(const int* __restrict__ flag, const int* __restrict__ pos, long long n,
                                  int* __restrict__ out)
{
    const long long f = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (f >= n) return;
    if (flag[f]) out[pos[f]] = static_cast<int>(f);
}

// ---------------------------------------------------------------------------
// pointer jumping on packed words (acc << 32 | next).  Invariant kept by
// every update: acc(v) = sum of weights on the path from v up to, but not
// including, next(v).  Roots (outlets, masked cells) point at themselves.
// 64-bit loads/stores are single accesses, so updates in place are safe.
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned long long pack(unsigned acc, unsigned nxt)
{
    return (static_cast<unsigned long long>(acc) << 32) | nxt;
}

__global__ void k_pj_init(const int* __restrict__ rcv, const uint8_t* __restrict__ mask,
                          const int* __restrict__ weight, long long n, unsigned long long* __restrict__ P)
{
    const long long f = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (f >= n) return;
    const int r = rcv[f];
    if (!mask[f] || r < 0) {
        // root: acc holds the root's own value (0 for depth, start for s-vector)
        P[f] = pack(weight ? static_cast<unsigned>(weight[f]) : 0u, static_cast<unsigned>(f));
    } else {
        P[f] = pack(weight ? static_cast<unsigned>(weight[f]) : 1u, static_cast<unsigned>(r));
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
__global__ void k_pj_extract(const unsigned long long* __restrict__ P, const uint8_t* __restrict__ mask,
                             long long n, int* __restrict__ value, int* __restrict__ root,
                             int* __restrict__ stats)
{
    int v = -1, valid = 0;
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
            if (root) root[f] = static_cast<int>(nx);
        } else if (root) {
            root[f] = -1;
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
        if (threadIdx.x == 0) {
            atomicMax(&stats[0], m);
            atomicAdd(&stats[1], c);
        }
    }
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

__global__ void __launch_bounds__(1024) k_level_hist(const int* __restrict__ depth, long long n,
                                                     int* __restrict__ count)
{
    __shared__ int h[kLdsLevels];
    for (int i = threadIdx.x; i < kLdsLevels; i += blockDim.x) h[i] = 0;
    __syncthreads();
    const long long base = blockIdx.x * (long long)blockDim.x * kCellsPerThread;
    for (int j = 0; j < kCellsPerThread; ++j) {
        const long long f = base + j * (long long)blockDim.x + threadIdx.x;
        const int d = f < n ? depth[f] : -1;
        wave_agg_add(h, d < kLdsLevels ? d : 0, d >= 0 && d < kLdsLevels);
        if (d >= kLdsLevels) atomicAdd(&count[d], 1);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kLdsLevels; i += blockDim.x)
        if (h[i]) atomicAdd(&count[i], h[i]);
}

__global__ void __launch_bounds__(1024) k_level_scatter(const int* __restrict__ depth, long long n,
                                                        const int* __restrict__ start, int* __restrict__ cursor,
                                                        int* __restrict__ order)
{
    __shared__ int h[kLdsLevels];
    __shared__ int base_s[kLdsLevels];
    for (int i = threadIdx.x; i < kLdsLevels; i += blockDim.x) h[i] = 0;
    __syncthreads();
    const long long base = blockIdx.x * (long long)blockDim.x * kCellsPerThread;
    int d[kCellsPerThread], rank[kCellsPerThread];
    for (int j = 0; j < kCellsPerThread; ++j) {
        const long long f = base + j * (long long)blockDim.x + threadIdx.x;
        d[j] = f < n ? depth[f] : -1;
        rank[j] = wave_agg_add(h, d[j] < kLdsLevels ? d[j] : 0, d[j] >= 0 && d[j] < kLdsLevels);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kLdsLevels; i += blockDim.x)
        if (h[i]) base_s[i] = atomicAdd(&cursor[i], h[i]);
    __syncthreads();
    for (int j = 0; j < kCellsPerThread; ++j) {
        const int dj = d[j];
        if (dj < 0) continue;
        const long long f = base + j * (long long)blockDim.x + threadIdx.x;
        int pos;
        if (dj < kLdsLevels) pos = start[dj] + base_s[dj] + rank[j];
        else pos = start[dj] + atomicAdd(&cursor[dj], 1);
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
    int nr, nc;
    BC bc;
    int* cnt;
    unsigned long long* key;  // null: counts only
    const double* w;          // null: no weighted accumulation
    double* acc;
};

__device__ __forceinline__ void up_node(const UpArgs& a, int v)
{
    int nb[8];
    const int k = distinct_neighbours(v, a.nr, a.nc, a.bc, nb);
    int dn[8];
    int nd = 0;
    for (int i = 0; i < k; ++i)
        if (a.mask[nb[i]] && a.rcv[nb[i]] == v) dn[nd++] = nb[i];
    int c = 1;
    unsigned long long mk = 0;
    for (int i = 0; i < nd; ++i) {
        c += a.cnt[dn[i]];
        if (a.key) mk = max(mk, a.key[dn[i]]);
    }
    a.cnt[v] = c;
    if (a.key)
        a.key[v] = nd ? ((((mk >> 32) + 1ull) << 32) | (mk & 0xffffffffull))
                      : static_cast<unsigned long long>(static_cast<unsigned>(v));
    if (a.acc) {
        // donors in topological order = ascending Kahn key
        unsigned long long dk[8];
        for (int i = 0; i < nd; ++i) dk[i] = a.key[dn[i]];
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

// compact Kahn keys: (height << nbits) | last-donor ancestor
__global__ void k_topo_keys(const int* __restrict__ order, int nv, const unsigned long long* __restrict__ key,
                            int nbits, unsigned long long* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nv) {
        const unsigned long long k = key[order[i]];
        out[i] = ((k >> 32) << nbits) | (k & 0xffffffffull);
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
    if (!mask[f] || r < 0) return;  // roots are written by k_sv_outlet_start
    int nb[8];
    const int k = distinct_neighbours(r, nr, nc, bc, nb);
    int s = 1;
    for (int i = 0; i < k; ++i) {
        const int u = nb[i];
        if (u < f && mask[u] && rcv[u] == r) s += cnt[u];
    }
    off[f] = s;
}

__global__ void k_gather_outlet_sizes(const int* __restrict__ outlets, int no, const int* __restrict__ cnt,
                                      int* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < no) out[i] = cnt[outlets[i]];
}
__global__ void k_sv_outlet_start(const int* __restrict__ outlets, int no, const int* __restrict__ pref,
                                  int* __restrict__ off)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < no) off[outlets[i]] = pref[i];
}
__global__ void k_sv_scatter(const int* __restrict__ start, const uint8_t* __restrict__ mask, long long n,
                             int* __restrict__ svec)
{
    const long long f = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (f < n && mask[f]) svec[start[f]] = static_cast<int>(f);
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
    depth.reserve(n);
    root.reserve(n);
    pj.reserve(n);
    flags.reserve(n);
    scan_out.reserve(n);
    outlets.reserve(n);
    misc.reserve(8);
    pjflags.reserve(40);
}

template <class T>
long long FlowEngine::route(const T* z, const uint8_t* msk, const uint8_t* forced, hipStream_t s)
{
    mask = msk;
    const int nr = opt.nr, nc = opt.nc;
    const int g1 = div_up(n, kBlock);
    // misc[0] = first unresolved flat, misc[1] = max depth, misc[2] = n_valid
    int init[3] = {0x7fffffff, -1, 0};
    LSDTT_HIP(hipMemcpyAsync(misc.p, init, sizeof(init), hipMemcpyHostToDevice, s));
    {
        TScope ts_(timer, "receivers", s, (sizeof(T) + 6.0) * n);
        launch_receivers<T>(z, msk, forced, nr, nc, opt.dx, opt.bc, opt.interior_nodata_is_outlet,
                            opt.allow_pits_as_outlets, rcv, dir, misc.p, s);
    }
    {
        TScope ts_(timer, "donors", s, 9.0 * n);
        k_donors<<<g1, kBlock, 0, s>>>(rcv, msk, nr, nc, opt.bc, donors, flags);
    }
    LSDTT_LAUNCH_CHECK();
    size_t tb = 0;
    LSDTT_HIP(rocprim::exclusive_scan(nullptr, tb, flags.p, scan_out.p, 0, (size_t)n, rocprim::plus<int>(), s));
    LSDTT_HIP(rocprim::exclusive_scan(tmp.need(tb), tb, flags.p, scan_out.p, 0, (size_t)n, rocprim::plus<int>(), s));
    k_scatter_flagged<<<g1, kBlock, 0, s>>>(flags, scan_out, n, outlets);
    LSDTT_LAUNCH_CHECK();

    pointer_jump(nullptr, depth, root, s);

    int h[3];
    int last_sum[2];
    LSDTT_HIP(hipMemcpyAsync(h, misc.p, sizeof(h), hipMemcpyDeviceToHost, s));
    LSDTT_HIP(hipMemcpyAsync(last_sum, scan_out.p + (n - 1), sizeof(int), hipMemcpyDeviceToHost, s));
    LSDTT_HIP(hipMemcpyAsync(last_sum + 1, flags.p + (n - 1), sizeof(int), hipMemcpyDeviceToHost, s));
    LSDTT_HIP(hipStreamSynchronize(s));
    n_outlets = last_sum[0] + last_sum[1];
    if (h[0] != 0x7fffffff) return -(static_cast<long long>(h[0]) + 1);
    max_depth = h[1];
    n_valid = h[2];

    level_start_h.assign(static_cast<size_t>(std::max(max_depth, 0)) + 2, 0);
    if (n_valid == 0) {
        tail_from = 0;
        return 0;
    }
    const int nl = max_depth + 1;
    level_count.reserve(nl + 1);
    level_start.reserve(nl + 1);
    cursor.reserve(nl + 1);
    order.reserve(n_valid);
    LSDTT_HIP(hipMemsetAsync(level_count.p, 0, sizeof(int) * (nl + 1), s));
    LSDTT_HIP(hipMemsetAsync(cursor.p, 0, sizeof(int) * (nl + 1), s));
    const int g1k = div_up(n, 1024 * kCellsPerThread);
    {
        TScope ts_(timer, "level_hist", s, 4.0 * n);
        k_level_hist<<<g1k, 1024, 0, s>>>(depth, n, level_count);
    }
    LSDTT_LAUNCH_CHECK();
    LSDTT_HIP(rocprim::exclusive_scan(nullptr, tb, level_count.p, level_start.p, 0, (size_t)(nl + 1),
                                      rocprim::plus<int>(), s));
    LSDTT_HIP(rocprim::exclusive_scan(tmp.need(tb), tb, level_count.p, level_start.p, 0, (size_t)(nl + 1),
                                      rocprim::plus<int>(), s));
    {
        TScope ts_(timer, "level_scatter", s, 8.0 * n);
        k_level_scatter<<<g1k, 1024, 0, s>>>(depth, n, level_start, cursor, order);
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
    TScope ts_(timer, "receivers", s, (sizeof(T) + 6.0) * n);
    launch_receivers<T>(z, msk, nullptr, opt.nr, opt.nc, opt.dx, opt.bc, opt.interior_nodata_is_outlet,
                        opt.allow_pits_as_outlets, rcv, dir, misc.p, s, zcopy, mark);
}

template void FlowEngine::receivers_only<double>(const double*, const uint8_t*, hipStream_t, double*,
                                                 const TileMark&);
template void FlowEngine::receivers_only<float>(const float*, const uint8_t*, hipStream_t, float*, const TileMark&);

template long long FlowEngine::route<float>(const float*, const uint8_t*, const uint8_t*, hipStream_t);
template long long FlowEngine::route<double>(const double*, const uint8_t*, const uint8_t*, hipStream_t);

void FlowEngine::pointer_jump(const int* weight, int* value_out, int* root_out, hipStream_t s)
{
    const int g1 = div_up(n, kBlock);
    k_pj_init<<<g1, kBlock, 0, s>>>(rcv, mask, weight, n, pj);
    LSDTT_LAUNCH_CHECK();
    // flags[r] == 1 while round r-1 changed something; enough rounds for a
    // path of 2^31 cells
    constexpr int kRounds = 34;
    int* chg = pjflags.p;
    LSDTT_HIP(hipMemsetAsync(chg, 0, sizeof(int) * (kRounds + 1), s));
    const int one = 1;
    LSDTT_HIP(hipMemcpyAsync(chg, &one, sizeof(int), hipMemcpyHostToDevice, s));
    for (int r = 0; r < kRounds; ++r) {
        {
            TScope ts_(timer, "pointer_jump", s, 24.0 * n);
            k_pj_round<<<std::min(g1, 2048), kBlock, 0, s>>>(pj, n, chg + r, chg + r + 1);
        }
        LSDTT_LAUNCH_CHECK();
    }
    k_pj_extract<<<std::min(g1, 2048), kBlock, 0, s>>>(pj, mask, n, value_out, root_out,
                                                       weight ? nullptr : misc.p + 1);
    LSDTT_LAUNCH_CHECK();
    int last = 0;
    LSDTT_HIP(hipMemcpyAsync(&last, chg + kRounds, sizeof(int), hipMemcpyDeviceToHost, s));
    LSDTT_HIP(hipStreamSynchronize(s));
    if (last != 0) throw std::runtime_error("pointer jumping did not converge");
}

void FlowEngine::upstream(bool keys, const double* w, double* acc, hipStream_t s)
{
    cnt.reserve(n);
    if (keys || w) key.reserve(n);
    if (n_valid == 0) return;
    UpArgs a{rcv, mask, opt.nr, opt.nc, opt.bc, cnt, (keys || w) ? key.p : nullptr, w, w ? acc : nullptr};
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
    const int nbits = std::max(1, bit_length(static_cast<unsigned long long>(n - 1)));
    k_topo_keys<<<div_up(n_valid, kBlock), kBlock, 0, s>>>(order, n_valid, key, nbits, sort_keys_a);
    LSDTT_LAUNCH_CHECK();
    const int end_bit = std::min(64, nbits + bit_length(static_cast<unsigned long long>(max_depth) + 1));
    size_t tb = 0;
    LSDTT_HIP(rocprim::radix_sort_pairs(nullptr, tb, sort_keys_a.p, sort_keys_b.p, order.p, topo.p,
                                        (size_t)n_valid, 0, end_bit, s));
    LSDTT_HIP(rocprim::radix_sort_pairs(tmp.need(tb), tb, sort_keys_a.p, sort_keys_b.p, order.p, topo.p,
                                        (size_t)n_valid, 0, end_bit, s));
}

void FlowEngine::s_vector(hipStream_t s)
{
    start.reserve(n);
    svec.reserve(n_valid ? n_valid : 1);
    off.reserve(n);
    if (n_valid == 0) return;
    const int g1 = div_up(n, kBlock);
    // outlet starts: exclusive scan of outlet sizes in row-major order
    k_gather_outlet_sizes<<<div_up(n_outlets, kBlock), kBlock, 0, s>>>(outlets, n_outlets, cnt, flags);
    LSDTT_LAUNCH_CHECK();
    size_t tb = 0;
    LSDTT_HIP(rocprim::exclusive_scan(nullptr, tb, flags.p, scan_out.p, 0, (size_t)n_outlets,
                                      rocprim::plus<int>(), s));
    LSDTT_HIP(rocprim::exclusive_scan(tmp.need(tb), tb, flags.p, scan_out.p, 0, (size_t)n_outlets,
                                      rocprim::plus<int>(), s));
    k_sv_outlet_start<<<div_up(n_outlets, kBlock), kBlock, 0, s>>>(outlets, n_outlets, scan_out, off);
    LSDTT_LAUNCH_CHECK();
    k_sv_offsets<<<g1, kBlock, 0, s>>>(rcv, mask, cnt, opt.nr, opt.nc, opt.bc, off);
    LSDTT_LAUNCH_CHECK();
    pointer_jump(off, start, nullptr, s);
    k_sv_scatter<<<g1, kBlock, 0, s>>>(start, mask, n, svec);
    LSDTT_LAUNCH_CHECK();
}

}  // namespace lsdtt_gpu
