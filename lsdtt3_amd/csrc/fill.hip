// fill.hip -- epsilon-free depression filling as a parallel min/max fixed
// point.
//
// The reference (src/cpu/lsdtt_raster_flow_preprocessing_cpu.cpp) fills at
// min_slope <= 0 either by Vincent 1993 reconstruction (:1203-1394) or by the
// priority flood (:893-920, core :416-511); its property P9 pins the two
// equal.  Both compute the spill level
//     filled(p) = z(p)                                  p an exit
//     filled(p) = max(z(p), min_{q in N8(p), valid} filled(q))   otherwise
// whose greatest solution is the minimax path height to an exit.  Exits
// (:518-589, :1208-1245): every valid cell on a non-periodic raster border,
// plus -- only when some interior cell is invalid -- every valid cell with an
// in-bounds invalid 8-neighbour.  The n/o/b codes are ignored; a periodic
// axis wraps neighbours (:727-782).  Only min and max are used, so the
// result is bit-identical whatever the evaluation order.
//
// Device scheme: values are held as order-preserving int keys so LDS
// atomicMin can lower them.  Each 64x64 tile (plus a 1-cell halo) is relaxed
// to its own fixed point in LDS by four concurrent directional sweeps (one
// wavefront each: rows left->right, rows right->left, columns down, columns
// up), then written back; launches repeat until no tile changes.  Values only
// ever decrease towards the fixed point, so concurrent updates are safe.
#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <utility>

#include "common.hpp"

namespace lsdtt_gpu {

namespace {

constexpr int T = 64;       // tile edge
constexpr int H = T + 2;    // with halo
constexpr int kFixed = INT_MAX;  // z key of exits and invalid cells: never updated

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

__global__ void k_fill_interior_invalid(const float* __restrict__ z, const uint8_t* __restrict__ mask, int nr,
                                        int nc, double nodata, int nan_nodata, int* __restrict__ flag)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y * blockDim.y + threadIdx.y;
    if (r < 1 || r >= nr - 1 || c < 1 || c >= nc - 1) return;
    const long long k = (long long)r * nc + c;
    if (!mask[k] || is_nodata(z[k], nodata, nan_nodata)) *flag = 1;
}

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
    wk[k] = exit_cell ? key : kFixed;
}

// Strict-descent pointer of every relaxable cell: its lowest strictly lower
// valid 8-neighbour (periodic axes wrap), or itself.  Exits and invalid cells
// point at themselves.  `wk` distinguishes valid from invalid neighbours
// (invalid cells hold kFixed in both zk and wk).
__global__ void k_fill_descent(const float* __restrict__ z, const int* __restrict__ zk,
                               const int* __restrict__ wk, int nr, int nc, int per_ns, int per_ew,
                               int* __restrict__ ptr)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y * blockDim.y + threadIdx.y;
    if (r >= nr || c >= nc) return;
    const int k = r * nc + c;
    int best = k;
    if (zk[k] != kFixed) {
        float bz = z[k];
        for (int d = 0; d < 8; ++d) {
            int rr = r + d8_row[d], cc = c + d8_col[d];
            if (per_ns) rr = (rr + nr) % nr;
            if (per_ew) cc = (cc + nc) % nc;
            if (rr < 0 || rr >= nr || cc < 0 || cc >= nc) continue;
            const int q = rr * nc + cc;
            if (zk[q] == kFixed && wk[q] == kFixed) continue;  // invalid
            const float zq = z[q];
            if (zq < bz) {
                bz = zq;
                best = q;
            }
        }
    }
    ptr[k] = best;
}

// ptr[k] <- ptr[ptr[k]] until every pointer reaches a root (in place: a
// pointer only ever moves further down its own descent path)
__global__ void k_fill_jump(int* __restrict__ ptr, long long n, const int* __restrict__ prev, int* __restrict__ cur)
{
    if (*prev == 0) return;
    const long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int p = __hip_atomic_load(&ptr[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int q = __hip_atomic_load(&ptr[p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const bool ch = q != p;
    if (ch) __hip_atomic_store(&ptr[k], q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // one flag store per wavefront (stores of every thread to one address
    // would queue at its memory channel)
    const unsigned long long m = __ballot(ch);
    if (ch && (threadIdx.x & 63) == __ffsll(m) - 1) *cur = 1;
}

// a cell whose descent path ends at an exit already has its spill level:
// filled = z (a non-ascending path out exists, and filled >= z always)
__global__ void k_fill_seed(const int* __restrict__ ptr, const int* __restrict__ zk, long long n,
                            int* __restrict__ wk, const float* __restrict__ z)
{
    const long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (k >= n || zk[k] == kFixed) return;
    const int root = ptr[k];
    if (zk[root] == kFixed && wk[root] != kFixed) wk[k] = fkey(z[k]);
}

// one relaxation pass: pass q reads act, marks act_next and clears
// act_clear (the next pass's act_next); it does nothing when the previous
// pass of its batch (prev) changed no tile
__global__ void __launch_bounds__(256) k_fill_tile(int* __restrict__ wk, const int* __restrict__ zk, int nr,
                                                   int nc, int per_ns, int per_ew, const int* __restrict__ act,
                                                   int* __restrict__ act_next, int* __restrict__ act_clear,
                                                   const int* __restrict__ prev, int* __restrict__ changed)
{
    const int tx = blockIdx.x, ty = blockIdx.y, ntx = gridDim.x, nty = gridDim.y;
    if (prev && *prev == 0) return;
    if (threadIdx.x == 0) act_clear[ty * ntx + tx] = 0;
    if (!act[ty * ntx + tx]) return;
    __shared__ int W[H * H];
    __shared__ int Z[T * T];
    __shared__ int lchg, tchg;
    const int r0 = ty * T, c0 = tx * T;
    const int tid = threadIdx.x;
    for (int i = tid; i < H * H; i += blockDim.x) {
        int r = i / H - 1 + r0, c = i % H - 1 + c0;
        if (per_ns) r = (r + nr) % nr;
        if (per_ew) c = (c + nc) % nc;
        W[i] = (r >= 0 && r < nr && c >= 0 && c < nc) ? wk[(long long)r * nc + c] : kFixed;
    }
    for (int i = tid; i < T * T; i += blockDim.x) {
        const int r = i / T + r0, c = i % T + c0;
        Z[i] = (r < nr && c < nc) ? zk[(long long)r * nc + c] : kFixed;
    }
    if (tid == 0) tchg = 0;
    __syncthreads();

    const int dir = tid >> 6, lane = tid & 63;
    for (;;) {
        if (tid == 0) lchg = 0;
        __syncthreads();
        int mine = 0;
        for (int s = 0; s < T; ++s) {
            int r, c;
            if (dir == 0) { r = lane; c = s; }
            else if (dir == 1) { r = lane; c = T - 1 - s; }
            else if (dir == 2) { r = s; c = lane; }
            else { r = T - 1 - s; c = lane; }
            const int zv = Z[r * T + c];
            const int b = (r + 1) * H + (c + 1);
            if (zv == kFixed || W[b] == zv) continue;   // fixed, or already at its floor
            int m = W[b - H - 1];
            m = min(m, W[b - H]);
            m = min(m, W[b - H + 1]);
            m = min(m, W[b - 1]);
            m = min(m, W[b + 1]);
            m = min(m, W[b + H - 1]);
            m = min(m, W[b + H]);
            m = min(m, W[b + H + 1]);
            const int v = max(zv, m);
            if (v < W[b]) {
                atomicMin(&W[b], v);
                mine = 1;
            }
        }
        if (mine) lchg = 1;
        __syncthreads();
        if (!lchg) break;
        if (tid == 0) tchg = 1;
        __syncthreads();
    }
    if (!tchg) return;
    for (int i = tid; i < T * T; i += blockDim.x) {
        const int r = i / T + r0, c = i % T + c0;
        if (r < nr && c < nc) {
            const int lr = i / T + 1, lc = i % T + 1;
            const int v = W[lr * H + lc];
            int* g = &wk[(long long)r * nc + c];
            if (v < *g) atomicMin(g, v);
        }
    }
    if (tid < 9) {
        // this tile and its 8 neighbours may move on the next pass
        int y = ty + tid / 3 - 1, x = tx + tid % 3 - 1;
        if (per_ns) y = (y + nty) % nty;
        if (per_ew) x = (x + ntx) % ntx;
        if (y >= 0 && y < nty && x >= 0 && x < ntx) act_next[y * ntx + x] = 1;
    }
    if (tid == 0) *changed = 1;
}

__global__ void k_fill_out(const int* __restrict__ wk, const int* __restrict__ zk, const float* __restrict__ z,
                           const uint8_t* __restrict__ mask, long long n, double nodata, int nan_nodata,
                           float* __restrict__ out)
{
    const long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (k >= n) return;
    const bool valid = mask[k] && !is_nodata(z[k], nodata, nan_nodata);
    out[k] = valid ? kfloat(wk[k]) : static_cast<float>(nodata);
}

__global__ void k_fill_ones(int* __restrict__ a, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 1;
}

}  // namespace

constexpr int kFillBatchMax = 64;  // relaxation passes per host read-back, at most

size_t fill_scratch_ints(int nr, int nc)
{
    const long long n = (long long)nr * nc;
    const long long nt = (long long)div_up(nc, T) * div_up(nr, T);
    return static_cast<size_t>(3 * n + 3 * nt + 64 + kFillBatchMax);
}

// Enqueues the fill on `s` and returns the number of relaxation passes.
// `scratch` holds fill_scratch_ints(nr, nc) ints.  Blocks on one flag
// read-back per batch of passes.
int fill_device(const float* z, const uint8_t* mask, int nr, int nc, double nodata, BC bc, float* out,
                int* scratch, hipStream_t s)
{
    const int per_ns = bc.n == 'p', per_ew = bc.e == 'p';
    const bool nan_nodata = std::isnan(nodata);
    const long long n = (long long)nr * nc;
    const int ntx = div_up(nc, T), nty = div_up(nr, T), nt = ntx * nty;
    int* zk = scratch;
    int* wk = zk + n;
    int* ptr = wk + n;
    int* act[3] = {ptr + n, ptr + n + nt, ptr + n + 2 * nt};
    int* flags = ptr + n + 3 * nt;   // [0] interior invalid, [2..] jump flags
    int* pf = flags + 64;            // per-pass changed flags of one batch
    dim3 b2(64, 4), g2(div_up(nc, 64), div_up(nr, 4));
    const int g1 = div_up(n, 256);
    LSDTT_HIP(hipMemsetAsync(flags, 0, sizeof(int) * 64, s));
    k_fill_interior_invalid<<<g2, b2, 0, s>>>(z, mask, nr, nc, nodata, nan_nodata, flags);
    LSDTT_LAUNCH_CHECK();
    k_fill_init<<<g2, b2, 0, s>>>(z, mask, nr, nc, nodata, nan_nodata, per_ns, per_ew, flags, zk, wk);
    LSDTT_LAUNCH_CHECK();
    k_fill_descent<<<g2, b2, 0, s>>>(z, zk, wk, nr, nc, per_ns, per_ew, ptr);
    LSDTT_LAUNCH_CHECK();
    // the first jump round runs unconditionally (its "previous round
    // changed" flag is set on the device, no host copy)
    k_fill_ones<<<1, 64, 0, s>>>(flags + 2, 1);
    LSDTT_LAUNCH_CHECK();
    constexpr int kRounds = 34;
    for (int r = 0; r < kRounds; ++r) {
        k_fill_jump<<<g1, 256, 0, s>>>(ptr, n, flags + 2 + r, flags + 3 + r);
        LSDTT_LAUNCH_CHECK();
    }
    k_fill_seed<<<g1, 256, 0, s>>>(ptr, zk, n, wk, z);
    LSDTT_LAUNCH_CHECK();
    k_fill_ones<<<div_up(nt, 256), 256, 0, s>>>(act[0], nt);
    LSDTT_LAUNCH_CHECK();
    LSDTT_HIP(hipMemsetAsync(act[1], 0, sizeof(int) * nt, s));
    dim3 gt(ntx, nty);
    // passes are launched in batches (8, then doubling) with one read-back
    // of their flags; a pass after the settling one returns at once
    int passes = 0, launched = 0;
    int batch = 8;
    for (;;) {
        const int nb = std::min(batch, kFillBatchMax);
        LSDTT_HIP(hipMemsetAsync(pf, 0, sizeof(int) * nb, s));
        for (int q = 0; q < nb; ++q) {
            const int g = launched + q;
            k_fill_tile<<<gt, 256, 0, s>>>(wk, zk, nr, nc, per_ns, per_ew, act[g % 3], act[(g + 1) % 3],
                                           act[(g + 2) % 3], q > 0 ? pf + q - 1 : nullptr, pf + q);
            LSDTT_LAUNCH_CHECK();
        }
        int h[kFillBatchMax];
        LSDTT_HIP(hipMemcpyAsync(h, pf, sizeof(int) * nb, hipMemcpyDeviceToHost, s));
        LSDTT_HIP(hipStreamSynchronize(s));
        int settled = -1;
        for (int q = 0; q < nb; ++q)
            if (!h[q]) {
                settled = q;
                break;
            }
        if (settled >= 0) {
            passes = launched + settled + 1;
            break;
        }
        launched += nb;
        if (launched > n + 2) throw std::runtime_error("fill relaxation did not converge");
        batch *= 2;
    }
    k_fill_out<<<g1, 256, 0, s>>>(wk, zk, z, mask, n, nodata, nan_nodata, out);
    LSDTT_LAUNCH_CHECK();
    return passes;
}

}  // namespace lsdtt_gpu
