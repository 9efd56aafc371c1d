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
//     computed by pointer jumping with max once the rounds end (each round at
//     least halves the components that have not reached the marker).
//  3. Output: max(z, L(root)) for pit-basin cells whose chain reaches the
//     marker, z otherwise.
#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <stdexcept>
#include <utility>

#include "common.hpp"

namespace lsdtt_gpu {

namespace {

constexpr int kFixed = INT_MAX;  // z key of exits and invalid cells
constexpr unsigned long long kNoEdge = ~0ull;
constexpr int kJumpRounds = 30;    // descent pointer jumping (2^30 > any grid the fill takes)
constexpr int kMaxRounds = 40;     // Boruvka rounds (each at least halves the components)
constexpr int kParJumps = 6;       // per round: launches of the hook-tree flattening
constexpr int kLevelJumps = 32;    // hook-chain maxima (2^32 > any chain)
constexpr int kListGrid = 512;     // blocks of the kernels over a component list
constexpr int kChunkBlocks = 1024; // blocks of a list compaction (one contiguous chunk each)
constexpr int kB = 256;
constexpr int kStrideGrid = 2048;  // blocks of the grid-stride full-grid kernels

// flag slots (ints): [0] interior invalid, [2, 2+kJumpRounds] descent jumps,
// [64, 64+kMaxRounds] round go flags, [128, 128+kMaxRounds*(kParJumps+1))
// flattening flags, [512, 512+kLevelJumps] chain flags, [600..601] list
// lengths, [602] the last chain launch that ran, [640, 640+2 kChunkBlocks)
// per-chunk counts and offsets of a compaction
constexpr int kFlagJump = 2, kFlagRound = 64, kFlagPar = 128, kFlagLevel = 512, kFlagCount = 600;
constexpr int kFlagLast = 602, kFlagChunks = 640;
constexpr int kFlagInts = kFlagChunks + 2 * kChunkBlocks;
static_assert(kFlagJump + kJumpRounds + 1 <= kFlagRound, "flag slots");
static_assert(kFlagRound + kMaxRounds + 1 <= kFlagPar, "flag slots");
static_assert(kFlagPar + kMaxRounds * (kParJumps + 1) <= kFlagLevel, "flag slots");
static_assert(kFlagLevel + kLevelJumps + 1 <= kFlagCount, "flag slots");

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
__global__ void k_fill_interior_invalid(const float* __restrict__ z, const uint8_t* __restrict__ mask, int nr,
                                        int nc, double nodata, int nan_nodata, int* __restrict__ flag)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y * blockDim.y + threadIdx.y;
    if (r < 1 || r >= nr - 1 || c < 1 || c >= nc - 1) return;
    const long long k = (long long)r * nc + c;
    if (!mask[k] || is_nodata(z[k], nodata, nan_nodata)) *flag = 1;
}

// wk: key of every valid cell (kFixed when invalid); zk: the same for
// non-exits, kFixed for exits and invalid cells.
__global__ void k_fill_init(const float* __restrict__ z, const uint8_t* __restrict__ mask, int nr, int nc,
                            double nodata, int nan_nodata, int per_ns, int per_ew,
                            const int* __restrict__ interior_invalid, int* __restrict__ zk,
                            int* __restrict__ wk)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y * blockDim.y + threadIdx.y;
    if (r >= nr || c >= nc) return;
    const long long k = (long long)r * nc + c;
    const bool valid = mask[k] && !is_nodata(z[k], nodata, nan_nodata);
    if (!valid) {
        zk[k] = kFixed;
        wk[k] = kFixed;
        return;
    }
    bool exit_cell = (!per_ns && (r == 0 || r == nr - 1)) || (!per_ew && (c == 0 || c == nc - 1));
    if (!exit_cell && *interior_invalid) {
        for (int d = 0; d < 8 && !exit_cell; ++d) {
            const int rr = r + d8_row[d], cc = c + d8_col[d];
            if (rr < 0 || rr >= nr || cc < 0 || cc >= nc) continue;
            const long long q = (long long)rr * nc + cc;
            if (!mask[q] || is_nodata(z[q], nodata, nan_nodata)) exit_cell = true;
        }
    }
    const int key = fkey(z[k]);
    zk[k] = exit_cell ? kFixed : key;
    wk[k] = key;
}

// Descent pointer: the valid 8-neighbour lowest in (key, flat index) order
// when below the cell's own (key, index), else the cell itself -- so
// equal-height plateaus chain into one basin, and (key, index) strictly
// falls along every pointer (no cycles; paths never rise in the key order
// that every min and max below uses).  Exits and invalid cells point at
// themselves.
__global__ void k_fill_descent(const int* __restrict__ zk, const int* __restrict__ wk, int nr, int nc, int per_ns,
                               int per_ew, int* __restrict__ ptr)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y * blockDim.y + threadIdx.y;
    if (r >= nr || c >= nc) return;
    const int k = r * nc + c;
    int best = k;
    if (zk[k] != kFixed) {
        int bk = zk[k];
        for (int d = 0; d < 8; ++d) {
            const int q = nbr_flat(r, c, d, nr, nc, per_ns, per_ew);
            if (q < 0) continue;
            const int kq = wk[q];
            if (kq == kFixed) continue;  // invalid
            if (kq < bk || (kq == bk && q < best)) {
                bk = kq;
                best = q;
            }
        }
    }
    ptr[k] = best;
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

// Components: lab[k] = X (= n, the marker) when the cell's root is an exit,
// its root otherwise, -1 when invalid.  Every slot c <= n starts with no hook
// (hp -1, hw INT_MIN), par[c] = c and no cheapest edge.
__global__ void k_bv_label(const int* __restrict__ zk, const int* __restrict__ wk, const int* __restrict__ ptr,
                           long long n, int* __restrict__ lab, int* __restrict__ par, int* __restrict__ hp,
                           int* __restrict__ hw, unsigned long long* __restrict__ best)
{
    const long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (k > n) return;
    par[k] = (int)k;
    hp[k] = -1;
    hw[k] = INT_MIN;
    best[k] = kNoEdge;
    if (k == n) return;  // the marker's slots
    if (wk[k] == kFixed) {
        lab[k] = -1;
        return;
    }
    const int r = ptr[k];
    lab[k] = zk[r] == kFixed ? (int)n : r;
}

// List compaction in three launches (a count per contiguous chunk, one scan,
// the writes in order), so no counter is shared by the whole grid.
// kRoots: over the cells, those that are pit roots (lab[c] == c); kSurvivors:
// over a component list, the entries that did not hook (par[c] == c), whose
// cheapest edge is cleared for the next round.
enum { kRoots = 0, kSurvivors = 1 };

template <int MODE>
__device__ __forceinline__ int list_item(int i, const int* __restrict__ lab, const int* __restrict__ par,
                                         const int* __restrict__ list)
{
    if (MODE == kRoots) return lab[i] == i ? i : -1;
    const int c = list[i];
    return par[c] == c ? c : -1;
}

template <int MODE>
__global__ void __launch_bounds__(kB) k_list_count(const int* __restrict__ lab, const int* __restrict__ par,
                                                   const int* __restrict__ list, long long m_fixed,
                                                   const int* __restrict__ m_dev, int* __restrict__ chunk_count,
                                                   const int* __restrict__ go)
{
    if (go && *go == 0) return;
    const long long m = m_dev ? *m_dev : m_fixed;
    const long long len = (m + gridDim.x - 1) / gridDim.x;
    const long long b0 = blockIdx.x * len, b1 = std::min(m, b0 + len);
    int c = 0;
    for (long long i = b0 + threadIdx.x; i < b1; i += blockDim.x) c += list_item<MODE>((int)i, lab, par, list) >= 0;
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

template <int MODE>
__global__ void __launch_bounds__(kB) k_list_write(const int* __restrict__ lab, const int* __restrict__ par,
                                                   const int* __restrict__ list, long long m_fixed,
                                                   const int* __restrict__ m_dev, const int* __restrict__ chunk_off,
                                                   int* __restrict__ out, unsigned long long* __restrict__ best,
                                                   const int* __restrict__ go)
{
    if (go && *go == 0) return;
    const long long m = m_dev ? *m_dev : m_fixed;
    const long long len = (m + gridDim.x - 1) / gridDim.x;
    const long long b0 = blockIdx.x * len, b1 = std::min(m, b0 + len);
    __shared__ int wsum[kB / 64];
    int base = chunk_off[blockIdx.x];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (long long i0 = b0; i0 < b1; i0 += blockDim.x) {
        const long long i = i0 + threadIdx.x;
        const int v = i < b1 ? list_item<MODE>((int)i, lab, par, list) : -1;
        const unsigned long long bal = __ballot(v >= 0);
        if (lane == 0) wsum[wave] = __popcll(bal);
        __syncthreads();
        int before = 0, all = 0;
        for (int w = 0; w < kB / 64; ++w) {
            before += w < wave ? wsum[w] : 0;
            all += wsum[w];
        }
        if (v >= 0) {
            out[base + before + __popcll(bal & ((1ull << lane) - 1ull))] = v;
            if (MODE == kSurvivors) best[v] = kNoEdge;
        }
        base += all;
        __syncthreads();
    }
}

// Each cell of a non-marker component offers its cheapest edge to another
// component: key = (w key with the sign bit flipped) << 32 | edge id, the
// edge id min(p, q) * 8 + direction from the lower to the higher index.  A
// block covers 64 x 4 cells; their labels and keys, with a one-cell halo
// (periodic axes wrap; -1 outside the grid), are staged in LDS first.
constexpr int kScanW = 64, kScanH = 4;
__global__ void __launch_bounds__(kScanW * kScanH) k_bv_scan(const int* __restrict__ wk, const int* __restrict__ lab,
                                                             int nr, int nc, int per_ns, int per_ew, int X,
                                                             unsigned long long* __restrict__ best,
                                                             const int* __restrict__ go)
{
    if (*go == 0) return;
    constexpr int TW = kScanW + 2, TH = kScanH + 2;
    __shared__ int sl[TH * TW], sk[TH * TW];
    const int c0 = blockIdx.x * kScanW, r0 = blockIdx.y * kScanH;
    const int tid = threadIdx.y * kScanW + threadIdx.x;
    for (int i = tid; i < TH * TW; i += kScanW * kScanH) {
        int rr = r0 - 1 + i / TW, cc = c0 - 1 + i % TW;
        if (per_ns) rr = (rr + nr) % nr;
        if (per_ew) cc = (cc + nc) % nc;
        const bool ok = rr >= 0 && rr < nr && cc >= 0 && cc < nc;
        const int q = ok ? rr * nc + cc : 0;
        sl[i] = ok ? lab[q] : -1;
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
    if (A >= 0) {
        const int kp = sk[li];
        for (int d = 0; d < 8; ++d) {
            const int lq = li + d8_row[d] * TW + d8_col[d];
            const int B = sl[lq];
            if (B < 0 || B == A) continue;
            const int q = nbr_flat(r, c, d, nr, nc, per_ns, per_ew);
            const int w = max(kp, sk[lq]);
            const unsigned uw = static_cast<unsigned>(w) ^ 0x80000000u;
            const unsigned eid =
                k < q ? static_cast<unsigned>(k) * 8u + d : static_cast<unsigned>(q) * 8u + ((d + 4) & 7);
            const unsigned long long key = (static_cast<unsigned long long>(uw) << 32) | eid;
            m = key < m ? key : m;
        }
    }
    // the wavefront is one row segment: runs of one component fold their
    // minima towards the run's first lane, which alone issues the atomic
    const int lane = threadIdx.x & 63;
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned long long om = __shfl_down(m, o);
        const int oa = __shfl_down(A, o);
        if (lane + o < 64 && oa == A && om < m) m = om;
    }
    const int left = __shfl_up(A, 1);
    const bool head = lane == 0 || left != A;
    if (head && A >= 0 && m != kNoEdge && m < best[A]) atomicMin(&best[A], m);
}

// Each listed component with an outgoing edge hooks onto the component across
// its cheapest edge, except the lower id of a pair that chose the same edge.
__global__ void k_bv_hook(const int* __restrict__ lab, const unsigned long long* __restrict__ best, int nr, int nc,
                          int per_ns, int per_ew, int X, const int* __restrict__ list,
                          const int* __restrict__ count, int* __restrict__ par, int* __restrict__ hp,
                          int* __restrict__ hw, const int* __restrict__ go, int* __restrict__ hooked)
{
    if (*go == 0) return;
    const int cnt = *count;
    const int stride = gridDim.x * blockDim.x;
    for (int i0 = blockIdx.x * blockDim.x; i0 < cnt; i0 += stride) {
        const int i = i0 + threadIdx.x;
        bool h = false;
        if (i < cnt) {
            const int c = list[i];
            const unsigned long long b = best[c];
            if (b != kNoEdge) {
                const unsigned eid = static_cast<unsigned>(b);
                const int p = static_cast<int>(eid >> 3), d = static_cast<int>(eid & 7);
                const int q = nbr_flat(p / nc, p % nc, d, nr, nc, per_ns, per_ew);
                const int A = lab[p], B = lab[q];
                const int t = A == c ? B : A;
                const bool stays = t != X && best[t] == b && c < t;
                if (!stays) {
                    par[c] = t;
                    hp[c] = t;
                    hw[c] = static_cast<int>(static_cast<unsigned>(b >> 32) ^ 0x80000000u);
                    h = true;
                }
            }
        }
        wave_flag(h, hooked);
    }
}

// flattening of this round's hook trees: each listed component walks up to 16
// parents (in place: a parent pointer only ever moves towards its tree's root)
__global__ void k_bv_flatten(int* __restrict__ par, const int* __restrict__ list, const int* __restrict__ count,
                             const int* __restrict__ go, const int* __restrict__ prev, int* __restrict__ cur)
{
    if (*go == 0 || *prev == 0) return;
    const int cnt = *count;
    const int stride = gridDim.x * blockDim.x;
    for (int i0 = blockIdx.x * blockDim.x; i0 < cnt; i0 += stride) {
        const int i = i0 + threadIdx.x;
        bool ch = false;
        if (i < cnt) {
            const int c = list[i];
            int p = __hip_atomic_load(&par[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const int p0 = p;
            for (int s = 0; s < 16; ++s) {
                const int q = __hip_atomic_load(&par[p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (q == p) break;
                p = q;
            }
            if (p != p0) __hip_atomic_store(&par[c], p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const int top = __hip_atomic_load(&par[p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            ch = top != p;  // not at a root yet
        }
        wave_flag(ch, cur);
    }
}

// every cell takes its component's new representative (grid-stride)
__global__ void k_bv_relabel(int* __restrict__ lab, long long n, const int* __restrict__ par,
                             const int* __restrict__ go)
{
    if (*go == 0) return;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < n; k += stride) {
        const int a = lab[k];
        if (a >= 0) {
            const int b = par[a];
            if (b != a) lab[k] = b;
        }
    }
}

// Chain maxima: (up, mx) -> (up2, mx2) with mx2 = max(mx, mx(up)), up2 =
// up(up), until every chain has reached the marker (or ends unhooked).  A
// launch after the last one that moved returns at once; the one that ran
// last records its index j (its output pair is j's parity).
__global__ void k_bv_level(const int* __restrict__ up, const int* __restrict__ mx, int* __restrict__ up2,
                           int* __restrict__ mx2, long long n, int X, int j, const int* __restrict__ prev,
                           int* __restrict__ cur, int* __restrict__ last)
{
    if (*prev == 0) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) *last = j;
    bool ch = false;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < n; k += stride) {
        const int u = up[k];
        int m = mx[k], nu = u;
        if (u >= 0 && u != X) {
            m = max(m, mx[u]);
            nu = up[u];
            ch |= nu >= 0 && nu != X;
        }
        up2[k] = nu;
        mx2[k] = m;
    }
    wave_flag(ch, cur);
}

// filled = max(z, L(root)) in a pit basin whose chain reached the marker,
// z otherwise; nodata for invalid cells
__global__ void k_fill_out(const float* __restrict__ z, const uint8_t* __restrict__ mask, const int* __restrict__ zk,
                           const int* __restrict__ ptr, const int* __restrict__ up_even,
                           const int* __restrict__ mx_even, const int* __restrict__ up_odd,
                           const int* __restrict__ mx_odd, const int* __restrict__ last, long long n, int X,
                           double nodata, int nan_nodata, float* __restrict__ out)
{
    const long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (k >= n) return;
    const bool odd = (*last & 1) != 0;
    const int* __restrict__ up = odd ? up_odd : up_even;
    const int* __restrict__ mx = odd ? mx_odd : mx_even;
    const float zv = z[k];
    const bool valid = mask[k] && !is_nodata(zv, nodata, nan_nodata);
    if (!valid) {
        out[k] = static_cast<float>(nodata);
        return;
    }
    const int r = ptr[k];
    float v = zv;
    if (zk[r] != kFixed && up[r] == X) v = kfloat(max(zk[k], mx[r]));
    out[k] = v;
}

// the flags whose launches run unconditionally: the first descent jump, the
// first round, the first flattening launch of every round, the first chain
// launch
__global__ void k_fill_start_flags(int* __restrict__ flags)
{
    const int i = threadIdx.x;
    if (i == 0) flags[kFlagJump] = 1;
    if (i == 1) flags[kFlagRound] = 1;
    if (i == 2) flags[kFlagLevel] = 1;
    if (i < kMaxRounds) flags[kFlagPar + i * (kParJumps + 1)] = 1;
}

}  // namespace

// ints of scratch: the cheapest edges (2 per slot, first, 8-byte aligned),
// zk, wk, ptr, lab, two component lists, par, hp, hw and the flags
size_t fill_scratch_ints(int nr, int nc)
{
    const long long n = (long long)nr * nc;
    return static_cast<size_t>(2 * (n + 1) + 6 * n + 3 * (n + 1) + kFlagInts + 8);
}

// Enqueues the fill on `s` and returns the number of Boruvka rounds run.
// `scratch` holds fill_scratch_ints(nr, nc) ints, 8-byte aligned.  Blocks on
// one read-back per batch of rounds.
int fill_device(const float* z, const uint8_t* mask, int nr, int nc, double nodata, BC bc, float* out,
                int* scratch, hipStream_t s)
{
    const int per_ns = bc.n == 'p', per_ew = bc.e == 'p';
    const bool nan_nodata = std::isnan(nodata);
    const long long n = (long long)nr * nc;
    if (n >= (1ll << 29)) throw std::invalid_argument("the fill supports grids of fewer than 2^29 cells");
    const int X = static_cast<int>(n);
    unsigned long long* best = reinterpret_cast<unsigned long long*>(scratch);
    int* zk = scratch + 2 * (n + 1);
    int* wk = zk + n;
    int* ptr = wk + n;
    int* lab = ptr + n;
    int* lists[2] = {lab + n, lab + 2 * n};
    int* par = lab + 3 * n;
    int* hp = par + (n + 1);
    int* hw = hp + (n + 1);
    int* flags = hw + (n + 1);
    int* counts = flags + kFlagCount;
    dim3 b2(64, 4), g2(div_up(nc, 64), div_up(nr, 4));
    const int g1 = div_up(n, kB), g1x = div_up(n + 1, kB);
    const int gs = std::min(g1, kStrideGrid);  // grid-stride launches
    const dim3 bsc(kScanW, kScanH), gsc(div_up(nc, kScanW), div_up(nr, kScanH));
    LSDTT_HIP(hipMemsetAsync(flags, 0, sizeof(int) * kFlagInts, s));
    k_fill_interior_invalid<<<g2, b2, 0, s>>>(z, mask, nr, nc, nodata, nan_nodata, flags);
    LSDTT_LAUNCH_CHECK();
    k_fill_init<<<g2, b2, 0, s>>>(z, mask, nr, nc, nodata, nan_nodata, per_ns, per_ew, flags, zk, wk);
    LSDTT_LAUNCH_CHECK();
    k_fill_descent<<<g2, b2, 0, s>>>(zk, wk, nr, nc, per_ns, per_ew, ptr);
    LSDTT_LAUNCH_CHECK();
    static_assert(kMaxRounds <= 64, "one thread per round flag");
    k_fill_start_flags<<<1, 64, 0, s>>>(flags);
    LSDTT_LAUNCH_CHECK();
    for (int r = 0; r < kJumpRounds; ++r) {
        k_fill_jump<<<gs, kB, 0, s>>>(ptr, n, flags + kFlagJump + r, flags + kFlagJump + r + 1);
        LSDTT_LAUNCH_CHECK();
    }
    int* chunk_count = flags + kFlagChunks;
    int* chunk_off = chunk_count + kChunkBlocks;
    k_bv_label<<<g1x, kB, 0, s>>>(zk, wk, ptr, n, lab, par, hp, hw, best);
    k_list_count<kRoots><<<kChunkBlocks, kB, 0, s>>>(lab, par, nullptr, n, nullptr, chunk_count, nullptr);
    k_list_scan<<<1, kChunkBlocks, 0, s>>>(chunk_count, chunk_off, counts, nullptr);
    k_list_write<kRoots><<<kChunkBlocks, kB, 0, s>>>(lab, par, nullptr, n, nullptr, chunk_off, lists[0], best,
                                                     nullptr);
    LSDTT_LAUNCH_CHECK();
    // Boruvka rounds in batches of 4 with one read-back of their flags; a
    // round after the last hook returns at once
    int launched = 0, rounds = 0;
    for (const int batch = 4;;) {
        const int end = std::min(kMaxRounds, launched + batch);
        for (int r = launched; r < end; ++r) {
            const int* go = flags + kFlagRound + r;
            int* hooked = flags + kFlagRound + r + 1;
            const int cur = r & 1;
            k_bv_scan<<<gsc, bsc, 0, s>>>(wk, lab, nr, nc, per_ns, per_ew, X, best, go);
            k_bv_hook<<<kListGrid, kB, 0, s>>>(lab, best, nr, nc, per_ns, per_ew, X, lists[cur], counts + cur, par,
                                               hp, hw, go, hooked);
            int* pf = flags + kFlagPar + r * (kParJumps + 1);
            for (int j = 0; j < kParJumps; ++j)
                k_bv_flatten<<<kListGrid, kB, 0, s>>>(par, lists[cur], counts + cur, go, pf + j, pf + j + 1);
            k_bv_relabel<<<gs, kB, 0, s>>>(lab, n, par, go);
            k_list_count<kSurvivors><<<kChunkBlocks, kB, 0, s>>>(lab, par, lists[cur], 0, counts + cur, chunk_count,
                                                                  go);
            k_list_scan<<<1, kChunkBlocks, 0, s>>>(chunk_count, chunk_off, counts + (cur ^ 1), go);
            k_list_write<kSurvivors><<<kChunkBlocks, kB, 0, s>>>(lab, par, lists[cur], 0, counts + cur, chunk_off,
                                                                  lists[cur ^ 1], best, go);
            LSDTT_LAUNCH_CHECK();
        }
        int fh[kFlagInts];
        LSDTT_HIP(hipMemcpyAsync(fh, flags, sizeof(int) * kFlagInts, hipMemcpyDeviceToHost, s));
        LSDTT_HIP(hipStreamSynchronize(s));
        const int* h = fh + kFlagRound;
        // a round whose last flattening launch still found a pointer short
        // of its root would have relabelled with unflattened trees
        for (int r = launched; r < end; ++r)
            if (h[r] && fh[kFlagPar + r * (kParJumps + 1) + kParJumps])
                throw std::runtime_error("fill: hook trees deeper than the flattening covers");
        launched = end;
        if (!h[end]) {
            rounds = 0;
            while (rounds < end && h[rounds + 1]) ++rounds;
            break;
        }
        if (launched >= kMaxRounds) throw std::runtime_error("fill: Boruvka rounds did not finish");
    }
    // chain maxima, ping-pong between (hp, hw) and (lab, par) (both free
    // now): launch j reads the pair j - 1 wrote and writes (lab, par) when j
    // is even, (hp, hw) when odd
    for (int j = 0; j < kLevelJumps; ++j) {
        const int* up = (j & 1) ? lab : hp;
        const int* mx = (j & 1) ? par : hw;
        int* up2 = (j & 1) ? hp : lab;
        int* mx2 = (j & 1) ? hw : par;
        k_bv_level<<<gs, kB, 0, s>>>(up, mx, up2, mx2, n, X, j, flags + kFlagLevel + j, flags + kFlagLevel + j + 1,
                                     flags + kFlagLast);
        LSDTT_LAUNCH_CHECK();
    }
    k_fill_out<<<g1, kB, 0, s>>>(z, mask, zk, ptr, lab, par, hp, hw, flags + kFlagLast, n, X, nodata, nan_nodata,
                                 out);
    LSDTT_LAUNCH_CHECK();
    return rounds;
}

}  // namespace lsdtt_gpu
