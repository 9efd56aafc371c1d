// batch.hip -- the batched LEM step (see batch.hpp).
//
// Reference semantics (src/cpu/): the D8 receivers of the pre-step surface
// (lsdtt_flowinfo_cpu.cpp:1077-1329 with interior nodata and pits as
// outlets, lsdtt_lem_cpu.cpp:329-339), D8 stream power
// (lsdtt_stream_power_cpu.cpp:159-326), the roering_implicit Newton
// hillslope (lsdtt_hillslope_newton_cpu.cpp:196-788), uplift, the boundary
// restore and the statistics (lsdtt_lem_cpu.cpp:614-722; p99.9 as
// lsdtt_statistics.cpp:420-429) -- per grid, as LemEngine runs them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>

#include "batch.hpp"
#include "reduce.hpp"

namespace lsdtt_gpu {

namespace {

constexpr int kB = 256;
// work items: the batched hillslope kernels' tiles
constexpr int kTW = kBatchTileCols, kTH = kBatchTileRows;
constexpr int kBins = 2048;

__global__ void kb_fill_i32(int* __restrict__ a, long long n, int v)
{
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ void kb_copy(const double* __restrict__ a, double* __restrict__ b, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        b[i] = a[i];
}

struct StTile {
    int gi, r0, c0;
};
__device__ __forceinline__ StTile st_tile(const BatchHsGrid* grids, const int* st_grid, int slot)
{
    const int gi = st_grid[slot];
    const BatchHsGrid& g = grids[gi];
    const int t = slot - g.stbase;
    return StTile{gi, (t / g.tx_n) * kTH, (t % g.tx_n) * kTW};
}

// D8 receiver of cell (row, col) of one grid (steepest descent over N, NE,
// ..., NW, strict >, slopes in double; no receiver -> outlet, as with
// interior_nodata_is_outlet and allow_pits_as_outlets), and the pre-step
// copy of its elevation
__device__ __forceinline__ void d8_cell(const double* __restrict__ z, const uint8_t* __restrict__ mask, int nr,
                                        int nc, BC bc, double card, double diag, int row, int col,
                                        int* __restrict__ rcv, uint8_t* __restrict__ dir,
                                        double* __restrict__ zcopy)
{
    const int f = row * nc + col;
    const double zc = z[f];
    zcopy[f] = zc;
    if (!mask[f]) {
        rcv[f] = kNodataReceiver;
        dir[f] = kNoDirection;
        return;
    }
    const bool bedge = (row == 0 && bc.n == 'b') || (col == nc - 1 && bc.e == 'b') ||
                       (row == nr - 1 && bc.s == 'b') || (col == 0 && bc.w == 'b');
    const bool interior = row > 0 && row < nr - 1 && col > 0 && col < nc - 1;
    double best = 0.0;
    int brcv = kOutletReceiver, bdir = kNoDirection;
    if (!bedge) {
        for (int d = 0; d < 8; ++d) {
            int g;
            if (interior) {
                g = f + d8_row[d] * nc + d8_col[d];
            } else {
                const Step st = resolve_step(row, col, d, nr, nc, bc);
                if (st.kind != 1) continue;  // off-grid: an outlet when nothing is lower
                g = st.r * nc + st.c;
            }
            if (!mask[g] || g == f) continue;  // masked: an outlet when nothing is lower
            const double dz = zc - z[g];
            if (dz <= 0.0) continue;
            const double sl = dz / ((d & 1) ? diag : card);
            if (sl > best) {
                best = sl;
                brcv = g;
                bdir = d;
            }
        }
    }
    rcv[f] = bdir != kNoDirection ? brcv : kOutletReceiver;
    dir[f] = static_cast<uint8_t>(bdir);
}

// one stencil tile per block; a thread walks 4 consecutive rows of one
// column with the 6 x 3 window of z and mask around them loaded first
// (4.5 loads per cell); cells on the grid's outer ring resolve their
// steps by the boundary rules in d8_cell
constexpr int kRows = 4;
__global__ void __launch_bounds__(256) kb_receivers(const BatchHsGrid* __restrict__ grids,
                                                    const int* __restrict__ st_grid, const double* __restrict__ z,
                                                    const uint8_t* __restrict__ mask, int* __restrict__ rcv,
                                                    uint8_t* __restrict__ dir, double* __restrict__ zcopy)
{
    // four blocks per item, one 16-row band each
    const StTile T = st_tile(grids, st_grid, blockIdx.x / 4);
    const BatchHsGrid& g = grids[T.gi];
    const int nr = g.nr, nc = g.nc;
    const int col = T.c0 + (threadIdx.x & 63);
    if (col >= nc) return;
    {
    const int r0 = T.r0 + 16 * (blockIdx.x % 4) + kRows * (threadIdx.x >> 6);
    if (r0 >= nr) return;
    const long long o = g.off;
    const double* zg = z + o;
    const uint8_t* mg = mask + o;
    const double card = g.h, diag = g.h * kSqrt2;
    const int cm = col > 0 ? col - 1 : col, cp = col < nc - 1 ? col + 1 : col;
    double zw[kRows + 2][3];
    uint8_t mw[kRows + 2][3];
#pragma unroll
    for (int k = 0; k < kRows + 2; ++k) {
        const int rr = min(max(r0 - 1 + k, 0), nr - 1);
        const long long b = (long long)rr * nc;
        zw[k][0] = zg[b + cm];
        zw[k][1] = zg[b + col];
        zw[k][2] = zg[b + cp];
        mw[k][0] = mg[b + cm];
        mw[k][1] = mg[b + col];
        mw[k][2] = mg[b + cp];
    }
    const bool edge_col = col == 0 || col == nc - 1;
#pragma unroll
    for (int k = 1; k <= kRows; ++k) {
        const int row = r0 - 1 + k;
        if (row >= nr) break;
        if (edge_col || row == 0 || row == nr - 1) {
            d8_cell(zg, mg, nr, nc, g.bc, card, diag, row, col, rcv + o, dir + o, zcopy + o);
            continue;
        }
        const int f = row * nc + col;
        const double zc = zw[k][1];
        zcopy[o + f] = zc;
        if (!mw[k][1]) {
            rcv[o + f] = kNodataReceiver;
            dir[o + f] = kNoDirection;
            continue;
        }
        // neighbours in D8 order N, NE, E, SE, S, SW, W, NW
        const double zn[8] = {zw[k - 1][1], zw[k - 1][2], zw[k][2], zw[k + 1][2],
                              zw[k + 1][1], zw[k + 1][0], zw[k][0], zw[k - 1][0]};
        const uint8_t mk[8] = {mw[k - 1][1], mw[k - 1][2], mw[k][2], mw[k + 1][2],
                               mw[k + 1][1], mw[k + 1][0], mw[k][0], mw[k - 1][0]};
        double best = 0.0;
        int bdir = kNoDirection;
#pragma unroll
        for (int d = 0; d < 8; ++d) {
            if (!mk[d]) continue;
            const double dz = zc - zn[d];
            if (dz <= 0.0) continue;
            const double sl = dz / ((d & 1) ? diag : card);
            if (sl > best) {
                best = sl;
                bdir = d;
            }
        }
        rcv[o + f] = bdir != kNoDirection ? f + d8_row[bdir] * nc + d8_col[bdir] : kOutletReceiver;
        dir[o + f] = static_cast<uint8_t>(bdir);
    }
    }
}

__global__ void kb_gather(const double* __restrict__ z, const long long* __restrict__ idx, long long m,
                          double* __restrict__ out)
{
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i < m) out[i] = z[idx[i]];
}

// statistics still to take this step: the grid's surface is final
__device__ __forceinline__ bool stats_pending(const HsCtl& c)
{
    return !c.stats_done && (c.phase == HS_IDLE || c.phase == HS_DONE);
}

// boundary flats take their base level in the grid's result surface
__global__ void kb_restore(const long long* __restrict__ idx, const int* __restrict__ gidx, long long m,
                           const double* __restrict__ bel, const HsCtl* __restrict__ ctl,
                           const BatchStepGrid* __restrict__ sg, double* zb0, double* zb1, double* zb2, double* zb3)
{
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int gi = gidx[i];
    const HsCtl& c = ctl[gi];
    if (!stats_pending(c)) return;
    double* zr = c.zc == 0 ? zb0 : (c.zc == 1 ? zb1 : (c.zc == 2 ? zb2 : zb3));
    zr[idx[i]] = bel[i] - sg[gi].bdz;
}

// uplift, the grid's result into z, |dz| and the sums of the statistics
// (LemEngine's k_stats per cell), one stencil tile per block
__global__ void __launch_bounds__(256) kb_stats(const BatchHsGrid* __restrict__ grids, const int* __restrict__ st_grid,
                                                const BatchStepGrid* __restrict__ sg, const HsCtl* __restrict__ ctl,
                                                const uint8_t* __restrict__ cflag, const double* __restrict__ zbefore,
                                                const double* zb0, const double* zb1, const double* zb2,
                                                const double* zb3, double* z, float* __restrict__ absdz,
                                                float* __restrict__ ero_rate, float* __restrict__ ws_dz,
                                                double* spartials)
{
    const StTile T = st_tile(grids, st_grid, blockIdx.x);
    const HsCtl& c = ctl[T.gi];
    if (!stats_pending(c)) return;
    const BatchHsGrid& g = grids[T.gi];
    const BatchStepGrid& q = sg[T.gi];
    const long long o = g.off;
    const double* zsrc = (c.zc == 0 ? zb0 : (c.zc == 1 ? zb1 : (c.zc == 2 ? zb2 : zb3))) + o;
    double* za = z + o;
    const bool moved = c.zc != 0;
    const double udz = q.udz, inv_dt = q.inv_dt;
    float* er = q.erosion_field ? ero_rate + o : nullptr;
    // warm start of the next step: the hillslope change (result - input)
    float* wd = g.ws_mode == 3 && c.method == 0 ? ws_dz + o : nullptr;
    double acc[8] = {0.0, 0.0, -1e30, -1e30, 0.0, 0.0, 0.0, 0.0};
    const int col = T.c0 + (threadIdx.x & 63);
    for (int k = 0; k < kTH / 4; ++k) {
        const int row = T.r0 + (threadIdx.x >> 6) + 4 * k;
        if (row >= g.nr || col >= g.nc) break;
        const long long f = (long long)row * g.nc + col;
        const uint8_t m = cflag[o + f];
        const double bb = zbefore[o + f];
        double a = zsrc[f];
        if (wd) wd[f] = moved ? static_cast<float>(a - za[f]) : 0.0f;
        if ((m & 2) && udz != 0.0) {
            a += udz;
            za[f] = a;
        } else if (moved) {
            za[f] = a;
        }
        if (!(m & 1)) {
            absdz[o + f] = __int_as_float(0x7f800000);
            if (er) er[f] = 0.0f;
            continue;
        }
        acc[0] += bb;
        acc[1] += a;
        acc[2] = fmax(acc[2], bb);
        acc[3] = fmax(acc[3], a);
        const double l = fabs(a - bb);
        acc[4] = fmax(acc[4], l);
        absdz[o + f] = static_cast<float>(l);
        const double ero = bb - (a - udz);
        if (er) er[f] = static_cast<float>(ero * inv_dt);
        if (ero > 0.0) {
            acc[5] += ero;
            acc[6] = fmax(acc[6], ero);
        }
        acc[7] += 1.0;
    }
    const bool mx[8] = {false, false, true, true, true, false, true, false};
    block_partial<8>(acc, mx, spartials, blockIdx.x);
}

// per grid: the statistics sums (its tiles in order), the cells the stream
// power updated, and the start of its p99.9 selection
__global__ void __launch_bounds__(256) kb_stats_fin(const BatchHsGrid* __restrict__ grids,
                                                    const BatchStepGrid* __restrict__ sg, HsCtl* ctl,
                                                    const double* spartials, const int* __restrict__ eroded,
                                                    double* out, long long* eroded_out, unsigned* sel,
                                                    int* __restrict__ ws_ok)
{
    const int gi = blockIdx.x;
    HsCtl& c = ctl[gi];
    if (!stats_pending(c)) return;
    const BatchHsGrid& g = grids[gi];
    const bool mx[8] = {false, false, true, true, true, false, true, false};
    double w[8];
    fold_partials<8>(spartials + (size_t)g.stbase * 8, g.nst, mx, w);
    __shared__ long long es[kB];
    long long e = 0;
    const int* er = eroded + sg[gi].rtbase;
    const int nrt = sg[gi].nrt;
    int t = threadIdx.x;
    for (; t + 7 * kB < nrt; t += 8 * kB) {  // eight loads in flight per thread
        int v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = er[t + j * kB];
#pragma unroll
        for (int j = 0; j < 8; ++j) e += v[j];
    }
    for (; t < nrt; t += kB) e += er[t];
    es[threadIdx.x] = e;
    __syncthreads();
    for (int s = kB / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) es[threadIdx.x] += es[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        for (int i = 0; i < 8; ++i) out[gi * 8 + i] = w[i];
        eroded_out[gi] = es[0];
        sel[gi * 4 + 0] = 0u;
        sel[gi * 4 + 1] = 0u;
        sel[gi * 4 + 2] = sg[gi].p999_k;
        sel[gi * 4 + 3] = w[7] > 0.0 ? 1u : 0u;
        ws_ok[gi] = g.ws_mode == 3 && c.phase == HS_DONE && c.method == 0 ? 1 : 0;
        c.stats_done = 1;
    }
}

// one radix pass of the p99.9 selections: a histogram per grid of this
// pass's digit of the |dz| bit patterns that match the grid's prefix
__global__ void __launch_bounds__(256) kb_radix_hist(const BatchHsGrid* __restrict__ grids,
                                                     const int* __restrict__ st_grid, const float* __restrict__ absdz,
                                                     const unsigned* __restrict__ sel, int shift, int bits,
                                                     unsigned* __restrict__ hist)
{
    const StTile T = st_tile(grids, st_grid, blockIdx.x);
    const unsigned* sl = sel + T.gi * 4;
    if (!sl[3]) return;
    __shared__ unsigned h[kBins];
    for (int i = threadIdx.x; i < kBins; i += blockDim.x) h[i] = 0;
    __syncthreads();
    const BatchHsGrid& g = grids[T.gi];
    const unsigned prefix = sl[0], pmask = sl[1];
    const int lane = threadIdx.x & 63;
    const int col = T.c0 + lane;
    for (int k = 0; k < kTH / 4; ++k) {
        const int row = T.r0 + (threadIdx.x >> 6) + 4 * k;
        bool want = false;
        unsigned bin = 0;
        if (row < g.nr && col < g.nc) {
            const unsigned u = __float_as_uint(absdz[g.off + (long long)row * g.nc + col]);
            want = (u & pmask) == prefix && u != 0x7f800000u;
            bin = (u >> shift) & ((1u << bits) - 1u);
        }
        unsigned long long pending = __ballot(want);
        while (pending) {
            const int leader = __ffsll(static_cast<long long>(pending)) - 1;
            const unsigned lb = __shfl(bin, leader);
            const unsigned long long same = __ballot(want && bin == lb);
            if (lane == leader) atomicAdd(&h[lb], static_cast<unsigned>(__popcll(same)));
            if (bin == lb) want = false;
            pending &= ~same;
        }
    }
    __syncthreads();
    unsigned* hg = hist + (size_t)T.gi * kBins;
    for (int i = threadIdx.x; i < kBins; i += blockDim.x)
        if (h[i]) atomicAdd(&hg[i], h[i]);
}

// the bin holding each grid's rank k (block per grid); clears the histogram
__global__ void __launch_bounds__(256) kb_radix_pick(unsigned* __restrict__ hist, unsigned* __restrict__ sel,
                                                     int shift, int bits)
{
    constexpr int kPer = kBins / kB;
    unsigned* sl = sel + blockIdx.x * 4;
    if (!sl[3]) return;
    unsigned* hg = hist + (size_t)blockIdx.x * kBins;
    __shared__ unsigned wtot[kB / 64];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int nb = 1 << bits;
    unsigned cv[kPer], sum = 0;
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        const int bb = t * kPer + j;
        cv[j] = bb < nb ? hg[bb] : 0u;
        sum += cv[j];
    }
    unsigned xs = sum;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned y = __shfl_up(xs, o);
        if (lane >= o) xs += y;
    }
    if (lane == 63) wtot[w] = xs;
    __syncthreads();
    unsigned before = xs - sum;
    for (int i = 0; i < w; ++i) before += wtot[i];
    const unsigned k = sl[2];
    __syncthreads();
    unsigned cum = before;
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        if (cum <= k && k < cum + cv[j]) {
            const int bb = t * kPer + j;
            sl[0] |= static_cast<unsigned>(bb) << shift;
            sl[1] |= ((1u << bits) - 1u) << shift;
            sl[2] = k - cum;
        }
        cum += cv[j];
    }
#pragma unroll
    for (int j = 0; j < kPer; ++j) hg[t * kPer + j] = 0u;
}

// after the host's Picard fallback of grid gi: its result is in z
__global__ void kb_picard_done(HsCtl* ctl, int gi)
{
    if (threadIdx.x == 0) {
        ctl[gi].zc = 0;
        ctl[gi].phase = HS_DONE;
        ctl[gi].method = 1;
    }
}

inline int blocks_for(long long n) { return std::max(1, div_up(n, kB)); }

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
BatchEngine::BatchEngine(const std::vector<BatchSpec>& specs)
{
    if (specs.empty()) throw std::invalid_argument("a batch needs at least one grid");
    build_layout(specs);
    LSDTT_HIP(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking));
}

BatchEngine::~BatchEngine()
{
    (void)hipDeviceSynchronize();
    release_graph();
    if (s_) (void)hipStreamDestroy(s_);
}

void BatchEngine::release_graph()
{
    if (exec) (void)hipGraphExecDestroy(exec);
    exec = nullptr;
}

void BatchEngine::build_layout(const std::vector<BatchSpec>& specs)
{
    G = static_cast<int>(specs.size());
    grids_h.resize(G);
    hsgrids_h.resize(G);
    params.resize(G);
    n_valid_h.assign(G, 0);
    bflats_h.assign(G, {});
    // cell offsets (each slice starts on a 64-cell boundary), stencil
    // tiles and general DOFs
    std::vector<int> gen_h, st_grid_h, gc_grid_h;
    long long off = 0;
    NT = 0;
    NGC = 0;
    int gen_off = 0;
    for (int i = 0; i < G; ++i) {
        const BatchSpec& sp = specs[static_cast<size_t>(i)];
        const LemParams& p = sp.p;
        if (sp.nr <= 0 || sp.nc <= 0 || !sp.mask) throw std::invalid_argument("grid " + std::to_string(i) + ": empty");
        if (p.fluvial_dinf) throw std::invalid_argument("batched steps support D8 incision only (grid " +
                                                        std::to_string(i) + ")");
        if (p.area_kernel > 1 || p.threshold > 0.0 || p.regime == 2 || p.force_picard)
            throw std::invalid_argument("batched steps support no area smoothing, incision threshold, channel set "
                                        "or forced Picard (grid " + std::to_string(i) + ")");
        for (char c : {p.bc.n, p.bc.e, p.bc.s, p.bc.w})
            if (c != 'o' && c != 'n' && c != 'p' && c != 'b')
                throw std::invalid_argument("boundary conditions must be four of o/n/p/b");
        const long long n = (long long)sp.nr * sp.nc;
        if (off + n >= (1ll << 31)) throw std::invalid_argument("a batch holds fewer than 2^31 cells");
        params[static_cast<size_t>(i)] = p;
        BatchHsGrid& h = hsgrids_h[static_cast<size_t>(i)];
        h = BatchHsGrid{};
        h.enabled = p.enable_hillslope;
        h.nr = sp.nr;
        h.nc = sp.nc;
        h.bc = p.bc;
        h.off = off;
        h.tx_n = div_up(sp.nc, kTW);
        h.stbase = NT;
        h.nst = h.tx_n * div_up(sp.nr, kTH);
        for (int k = 0; k < h.nst; ++k) st_grid_h.push_back(i);
        NT += h.nst;
        h.h = sp.dx;
        h.D = p.D;
        h.Sc = p.Sc;
        h.cap = p.slope_ratio_cap;
        h.nonlinear = p.nonlinear;
        h.eta2 = p.krylov_rel_tol * p.krylov_rel_tol;
        h.forcing = p.krylov_forcing;
        h.krylov_max = p.krylov_max_iterations;
        h.newton_max = p.newton_max_iterations;
        h.ls_max = p.newton_max_line_search;
        h.abs_tol = p.newton_abs_tol;
        h.rel_tol = p.newton_rel_tol;
        h.dt_total = p.dt;
        // DOFs: valid, off the 'o' edges; fast: interior with a valid 3x3
        // block; the rest in the grid's general list (LemEngine::build_dofs)
        const int nr = sp.nr, nc = sp.nc;
        h.gen_off = gen_off;
        int ng = 0;
        for (int r = 0; r < nr; ++r)
            for (int c = 0; c < nc; ++c) {
                const long long f = (long long)r * nc + c;
                const bool edge = (r == 0 && p.bc.n == 'o') || (r == nr - 1 && p.bc.s == 'o') ||
                                  (c == 0 && p.bc.w == 'o') || (c == nc - 1 && p.bc.e == 'o');
                if (!sp.mask[f] || edge) continue;
                bool ok = r > 0 && r < nr - 1 && c > 0 && c < nc - 1;
                for (int k = 0; ok && k < 9; ++k) ok = sp.mask[f + (k / 3 - 1) * nc + (k % 3 - 1)] != 0;
                if (!ok) {
                    gen_h.push_back(static_cast<int>(f));
                    ++ng;
                }
            }
        h.n_gen = ng;
        gen_off += ng;
        h.gcbase = NGC;
        h.ngc = div_up(ng, 256);
        for (int k = 0; k < h.ngc; ++k) gc_grid_h.push_back(i);
        NGC += h.ngc;
        BatchGrid& g = grids_h[static_cast<size_t>(i)];
        g = BatchGrid{};
        g.nr = nr;
        g.nc = nc;
        g.bc = p.bc;
        g.off = off;
        SpParams a{};
        a.K0 = p.K;
        a.m = p.m;
        a.nexp = p.n;
        a.dt = p.dt;
        a.dx = sp.dx;
        a.pixel_area = sp.dx * sp.dx;
        a.threshold = p.threshold;
        a.slope_clamp = p.slope_clamp;
        a.divide = p.divide_area_by_dx;
        a.n_is_one = std::fabs(p.n - 1.0) < 1e-6;
        a.m_half = p.m == 0.5;
        a.n1_half = p.n - 1.0 == 0.5;
        a.width = p.width_factor;
        g.sp = a;
        const int nm = a.n_is_one ? 0 : (a.n1_half ? 1 : 2);
        g.sp_mode = p.enable_fluvial ? sp_mode_of(a.m_half != 0, nm) : -1;
        if (g.sp_mode >= 0) sp_modes |= 1u << g.sp_mode;
        any_hillslope |= p.enable_hillslope != 0;
        long long nv = 0;
        for (long long f = 0; f < n; ++f) nv += sp.mask[f] != 0;
        n_valid_h[static_cast<size_t>(i)] = nv;
        off += (n + 63) / 64 * 64;
    }
    N = off;
    const size_t nn = static_cast<size_t>(N);
    // per-cell arrays
    for (auto* a : {&z, &z_before}) a->reserve(nn);
    if (any_hillslope)
        for (auto* a : {&zA, &zB, &zC}) a->reserve(nn);
    for (auto* a : {&mask, &dof, &fast, &cflag, &dir, &xdon}) a->reserve(nn);
    rcv.reserve(nn);
    cnt.reserve(nn);
    absdz.reserve(nn);
    if (any_hillslope)
        for (auto* a : {&b, &dinv, &x, &r, &pv, &pv2, &v, &v2, &t}) a->reserve(nn);
    gen.reserve(std::max<size_t>(1, gen_h.size()));
    c9g.reserve(std::max<size_t>(9, 9 * gen_h.size()));
    if (!gen_h.empty()) LSDTT_HIP(hipMemcpy(gen.p, gen_h.data(), sizeof(int) * gen_h.size(), hipMemcpyHostToDevice));
    // masks, DOF and fast flags, cell flags
    mask_all.assign(nn, 0);
    std::vector<uint8_t> dof_h(nn, 0), fast_h(nn, 0), cflag_h(nn, 0);
    for (int i = 0; i < G; ++i) {
        const BatchSpec& sp = specs[static_cast<size_t>(i)];
        const BatchHsGrid& h = hsgrids_h[static_cast<size_t>(i)];
        const LemParams& p = sp.p;
        const int nr = sp.nr, nc = sp.nc;
        for (int rr = 0; rr < nr; ++rr)
            for (int c = 0; c < nc; ++c) {
                const long long f = (long long)rr * nc + c, F = h.off + f;
                const uint8_t m = sp.mask[f] ? 1 : 0;
                mask_all[static_cast<size_t>(F)] = m;
                cflag_h[static_cast<size_t>(F)] = m ? 3 : 0;
                const bool edge = (rr == 0 && p.bc.n == 'o') || (rr == nr - 1 && p.bc.s == 'o') ||
                                  (c == 0 && p.bc.w == 'o') || (c == nc - 1 && p.bc.e == 'o');
                if (!m || edge) continue;
                dof_h[static_cast<size_t>(F)] = 1;
                bool ok = rr > 0 && rr < nr - 1 && c > 0 && c < nc - 1;
                for (int k = 0; ok && k < 9; ++k) ok = sp.mask[f + (k / 3 - 1) * nc + (k % 3 - 1)] != 0;
                fast_h[static_cast<size_t>(F)] = ok ? 1 : 0;
            }
    }
    LSDTT_HIP(hipMemcpy(mask.p, mask_all.data(), nn, hipMemcpyHostToDevice));
    LSDTT_HIP(hipMemcpy(dof.p, dof_h.data(), nn, hipMemcpyHostToDevice));
    LSDTT_HIP(hipMemcpy(fast.p, fast_h.data(), nn, hipMemcpyHostToDevice));
    LSDTT_HIP(hipMemcpy(cflag.p, cflag_h.data(), nn, hipMemcpyHostToDevice));
    LSDTT_HIP(hipMemset(z.p, 0, sizeof(double) * nn));
    LSDTT_HIP(hipMemset(cnt.p, 0, sizeof(int) * nn));
    // hillslope tables
    hsgrids_d.reserve(G);
    LSDTT_HIP(hipMemcpy(hsgrids_d.p, hsgrids_h.data(), sizeof(BatchHsGrid) * G, hipMemcpyHostToDevice));
    st_grid_d.reserve(NT);
    LSDTT_HIP(hipMemcpy(st_grid_d.p, st_grid_h.data(), sizeof(int) * NT, hipMemcpyHostToDevice));
    gc_grid_d.reserve(std::max(1, NGC));
    if (NGC) LSDTT_HIP(hipMemcpy(gc_grid_d.p, gc_grid_h.data(), sizeof(int) * NGC, hipMemcpyHostToDevice));
    ctl.reserve(G);
    LSDTT_HIP(hipMemset(ctl.p, 0, sizeof(HsCtl) * G));
    wl_inst.reserve(static_cast<size_t>(WL_N) * G);
    wl_pre.reserve(static_cast<size_t>(WL_N) * (G + 1));
    wl_hdr.reserve(2 * WL_N);
    LSDTT_HIP(hipMemset(wl_hdr.p, 0, sizeof(int) * 2 * WL_N));
    partials.reserve(static_cast<size_t>(NT + NGC) * 3);
    spartials.reserve(static_cast<size_t>(NT) * 8);
    stats_out.reserve(static_cast<size_t>(G) * 8);
    eroded_out.reserve(G);
    ws_ok.reserve(G);
    LSDTT_HIP(hipMemset(ws_ok.p, 0, sizeof(int) * G));
    hist.reserve(static_cast<size_t>(G) * kBins);
    LSDTT_HIP(hipMemset(hist.p, 0, sizeof(unsigned) * G * kBins));
    sel.reserve(static_cast<size_t>(G) * 4);
    LSDTT_HIP(hipMemset(sel.p, 0, sizeof(unsigned) * G * 4));
    pflags.reserve(2 * kMaxPass);
    ctl_h.reserve(G);
    stats_h.reserve(static_cast<size_t>(G) * 8);
    eroded_h.reserve(G);
    sel_h.reserve(static_cast<size_t>(G) * 4);
    pflags_h.reserve(2 * kMaxPass);
    rebuild_routing_tables();
    LSDTT_HIP(hipDeviceSynchronize());
}

// routing tiles (32 x th) of every grid and the per-step grid table
void BatchEngine::rebuild_routing_tables()
{
    std::vector<BatchTile> tiles;
    RT = 0;
    for (int i = 0; i < G; ++i) {
        BatchGrid& g = grids_h[static_cast<size_t>(i)];
        g.ntx = div_up(g.nc, FluvialTiles::TW);
        const int nty = div_up(g.nr, th);
        g.tbase = RT;
        for (int ty = 0; ty < nty; ++ty)
            for (int tx = 0; tx < g.ntx; ++tx) tiles.push_back(BatchTile{i, tx, ty});
        RT += g.ntx * nty;
    }
    grids_d.reserve(G);
    LSDTT_HIP(hipMemcpy(grids_d.p, grids_h.data(), sizeof(BatchGrid) * G, hipMemcpyHostToDevice));
    rtiles_d.reserve(RT);
    LSDTT_HIP(hipMemcpy(rtiles_d.p, tiles.data(), sizeof(BatchTile) * RT, hipMemcpyHostToDevice));
    const int TN = FluvialTiles::TW * th;
    ord.reserve(static_cast<size_t>(RT) * TN);
    lstart.reserve(static_cast<size_t>(RT) * (TN + 1));
    for (auto* a : {&maxdep, &act[0], &act[1], &act[2], &runs, &eroded}) a->reserve(RT);
    LSDTT_HIP(hipMemset(eroded.p, 0, sizeof(int) * RT));
    // per-step grid table
    std::vector<BatchStepGrid> sgh(static_cast<size_t>(G));
    for (int i = 0; i < G; ++i) {
        const LemParams& p = params[static_cast<size_t>(i)];
        BatchStepGrid& q = sgh[static_cast<size_t>(i)];
        q.udz = p.enable_uplift ? p.U * p.dt : 0.0;
        q.bdz = p.baselevel_fall * p.dt;
        q.inv_dt = 1.0 / p.dt;
        const long long nv = n_valid_h[static_cast<size_t>(i)];
        q.p999_k = nv > 0 ? static_cast<unsigned>(std::floor(0.999 * static_cast<double>(nv - 1))) : 0u;
        q.rtbase = grids_h[static_cast<size_t>(i)].tbase;
        const BatchGrid& g = grids_h[static_cast<size_t>(i)];
        q.nrt = g.ntx * div_up(g.nr, th);
        q.erosion_field = p.erosion_rate_field;
    }
    step_grids.reserve(G);
    LSDTT_HIP(hipMemcpy(step_grids.p, sgh.data(), sizeof(BatchStepGrid) * G, hipMemcpyHostToDevice));
    bool ero = false;
    for (const auto& p : params) ero |= p.erosion_rate_field != 0;
    if (ero) ero_rate.reserve(static_cast<size_t>(N));
    release_graph();
}

BatchHs BatchEngine::bhs()
{
    BatchHs B;
    B.G = G;
    B.wl_inst = wl_inst.p;
    B.wl_pre = wl_pre.p;
    B.wl_hdr = wl_hdr.p;
    B.grids = hsgrids_d.p;
    B.st_grid = st_grid_d.p;
    B.gc_grid = gc_grid_d.p;
    B.NT = NT;
    B.NGC = NGC;
    B.ctl = ctl.p;
    B.mask = mask.p;
    B.dof = dof.p;
    B.fast = fast.p;
    B.gen = gen.p;
    B.c9g = c9g.p;
    B.zbuf[0] = z.p;
    B.zbuf[1] = zA.p;
    B.zbuf[2] = zB.p;
    B.zbuf[3] = zC.p;
    B.b = b.p;
    B.dinv = dinv.p;
    B.x = x.p;
    B.r = r.p;
    B.p = pv.p;
    B.p2 = pv2.p;
    B.v = v.p;
    B.v2 = v2.p;
    B.t = t.p;
    B.partials = partials.p;
    B.ws_dz = ws_dz.p;
    B.ws_ok = ws_ok.p;
    return B;
}

void BatchEngine::set_elevation(int i, const double* zh)
{
    if (i < 0 || i >= G) throw std::invalid_argument("grid index out of range");
    const BatchHsGrid& h = hsgrids_h[static_cast<size_t>(i)];
    LSDTT_HIP(hipDeviceSynchronize());
    LSDTT_HIP(hipMemcpy(z.p + h.off, zh, sizeof(double) * cells(i), hipMemcpyHostToDevice));
    LSDTT_HIP(hipMemset(ws_ok.p + i, 0, sizeof(int)));  // no previous step for this surface
}

void BatchEngine::set_warm_start(int mode)
{
    if (mode != 0 && mode != 3) throw std::invalid_argument("ensemble warm start mode must be 0 or 3");
    LSDTT_HIP(hipDeviceSynchronize());
    if (mode == 3 && !ws_dz.p) {
        ws_dz.reserve(static_cast<size_t>(N));
        LSDTT_HIP(hipMemset(ws_dz.p, 0, sizeof(float) * N));
    }
    for (auto& h : hsgrids_h) h.ws_mode = mode;
    LSDTT_HIP(hipMemcpy(hsgrids_d.p, hsgrids_h.data(), sizeof(BatchHsGrid) * G, hipMemcpyHostToDevice));
    LSDTT_HIP(hipMemset(ws_ok.p, 0, sizeof(int) * G));
    ws_mode = mode;
    release_graph();  // the pattern holds the change buffer's address
}

void BatchEngine::get_elevation(int i, double* zh)
{
    if (i < 0 || i >= G) throw std::invalid_argument("grid index out of range");
    const BatchHsGrid& h = hsgrids_h[static_cast<size_t>(i)];
    LSDTT_HIP(hipDeviceSynchronize());
    LSDTT_HIP(hipMemcpy(zh, z.p + h.off, sizeof(double) * cells(i), hipMemcpyDeviceToHost));
}

void BatchEngine::get_erosion_rate(int i, float* out)
{
    if (i < 0 || i >= G) throw std::invalid_argument("grid index out of range");
    if (!params[static_cast<size_t>(i)].erosion_rate_field) throw std::invalid_argument("compute_erosion_rate_field is off");
    const BatchHsGrid& h = hsgrids_h[static_cast<size_t>(i)];
    LSDTT_HIP(hipDeviceSynchronize());
    LSDTT_HIP(hipMemcpy(out, ero_rate.p + h.off, sizeof(float) * cells(i), hipMemcpyDeviceToHost));
}

void BatchEngine::set_boundary_flats(int i, const long long* flats, long long m)
{
    if (i < 0 || i >= G) throw std::invalid_argument("grid index out of range");
    const long long n = cells(i);
    for (long long k = 0; k < m; ++k)
        if (flats[k] < 0 || flats[k] >= n) throw std::invalid_argument("boundary flat index out of range");
    bflats_h[static_cast<size_t>(i)].assign(flats, flats + m);
    // concatenated list (cells of the arrays) and the cell flags
    std::vector<long long> all;
    std::vector<int> gi;
    std::vector<uint8_t> cf(static_cast<size_t>(N), 0);
    for (long long F = 0; F < N; ++F) cf[static_cast<size_t>(F)] = mask_all[static_cast<size_t>(F)] ? 3 : 0;
    for (int j = 0; j < G; ++j) {
        const long long o = hsgrids_h[static_cast<size_t>(j)].off;
        for (long long f : bflats_h[static_cast<size_t>(j)]) {
            all.push_back(o + f);
            gi.push_back(j);
            cf[static_cast<size_t>(o + f)] &= 1;
        }
    }
    n_bflats = static_cast<long long>(all.size());
    bflat.reserve(std::max<size_t>(1, all.size()));
    bflat_grid.reserve(std::max<size_t>(1, all.size()));
    bel.reserve(std::max<size_t>(1, all.size()));
    LSDTT_HIP(hipDeviceSynchronize());
    if (n_bflats) {
        LSDTT_HIP(hipMemcpy(bflat.p, all.data(), sizeof(long long) * all.size(), hipMemcpyHostToDevice));
        LSDTT_HIP(hipMemcpy(bflat_grid.p, gi.data(), sizeof(int) * gi.size(), hipMemcpyHostToDevice));
    }
    LSDTT_HIP(hipMemcpy(cflag.p, cf.data(), static_cast<size_t>(N), hipMemcpyHostToDevice));
    release_graph();  // the pattern holds the list's launch size
}

void BatchEngine::set_graph(bool on)
{
    use_graph = on;
    if (!on) release_graph();
}

void BatchEngine::set_tile_height(int th_)
{
    if (!FluvialTiles::valid_tile_height(th_)) throw std::invalid_argument("tile height must be 8, 16 or 32");
    if (th_ == th) return;
    LSDTT_HIP(hipDeviceSynchronize());
    th = th_;
    rebuild_routing_tables();
}

// count passes first_pass .. first_pass + passes - 1 (pass flags in
// pflags[0, kMaxPass)); pass 0 starts with every tile active
void BatchEngine::enqueue_counts(int first_pass, int passes, hipStream_t s)
{
    if (first_pass == 0) {
        kb_fill_i32<<<blocks_for(RT), kB, 0, s>>>(act[0].p, RT, 1);
        kb_fill_i32<<<blocks_for(RT), kB, 0, s>>>(act[1].p, RT, 0);
        kb_fill_i32<<<blocks_for(RT), kB, 0, s>>>(runs.p, RT, 0);
    }
    kb_fill_i32<<<blocks_for(passes), kB, 0, s>>>(pflags.p + first_pass, passes, 0);
    LSDTT_LAUNCH_CHECK();
    for (int q = first_pass; q < first_pass + passes; ++q) {
        PassArgs P{act[q % 3].p, act[(q + 1) % 3].p, act[(q + 2) % 3].p, q > 0 ? pflags.p + q - 1 : nullptr,
                   pflags.p + q, runs.p};
        batch_tile_counts(th, btiles(), RT, rcv.p, mask.p, ord.p, lstart.p, maxdep.p, xdon.p, cnt.p, P, s);
    }
}

void BatchEngine::enqueue_sp(int first_pass, int passes, hipStream_t s)
{
    int* fl = pflags.p + kMaxPass;
    if (first_pass == 0) {
        kb_fill_i32<<<blocks_for(RT), kB, 0, s>>>(act[0].p, RT, 1);
        kb_fill_i32<<<blocks_for(RT), kB, 0, s>>>(act[1].p, RT, 0);
        kb_fill_i32<<<blocks_for(RT), kB, 0, s>>>(runs.p, RT, 0);
        kb_fill_i32<<<blocks_for(RT), kB, 0, s>>>(eroded.p, RT, 0);
    }
    kb_fill_i32<<<blocks_for(passes), kB, 0, s>>>(fl + first_pass, passes, 0);
    LSDTT_LAUNCH_CHECK();
    SpTileArgs A{rcv.p, dir.p, mask.p, cnt.p, nullptr, ord.p, lstart.p, maxdep.p, xdon.p, z_before.p, nullptr, z.p,
                 eroded.p, SpParams{}};
    for (int q = first_pass; q < first_pass + passes; ++q) {
        PassArgs P{act[q % 3].p, act[(q + 1) % 3].p, act[(q + 2) % 3].p, q > 0 ? fl + q - 1 : nullptr, fl + q,
                   runs.p};
        batch_tile_sp(th, sp_modes, btiles(), RT, A, P, s);
    }
}

void BatchEngine::enqueue_stats(hipStream_t s)
{
    if (n_bflats) {
        kb_restore<<<blocks_for(n_bflats), kB, 0, s>>>(bflat.p, bflat_grid.p, n_bflats, bel.p, ctl.p, step_grids.p,
                                                      z.p, zA.p, zB.p, zC.p);
    }
    kb_stats<<<NT, kB, 0, s>>>(hsgrids_d.p, st_grid_d.p, step_grids.p, ctl.p, cflag.p, z_before.p, z.p, zA.p, zB.p,
                               zC.p, z.p, absdz.p, ero_rate.p, ws_dz.p, spartials.p);
    kb_stats_fin<<<G, kB, 0, s>>>(hsgrids_d.p, step_grids.p, ctl.p, spartials.p, eroded.p, stats_out.p, eroded_out.p,
                                  sel.p, ws_ok.p);
    const int shifts[3] = {21, 10, 0}, bits[3] = {11, 11, 10};
    for (int pass = 0; pass < 3; ++pass) {
        kb_radix_hist<<<NT, kB, 0, s>>>(hsgrids_d.p, st_grid_d.p, absdz.p, sel.p, shifts[pass], bits[pass], hist.p);
        kb_radix_pick<<<G, kB, 0, s>>>(hist.p, sel.p, shifts[pass], bits[pass]);
    }
    LSDTT_LAUNCH_CHECK();
}

// the whole step with fixed caps: only kernel launches (capturable)
void BatchEngine::enqueue_step(const Plan& pl, hipStream_t s)
{
    kb_receivers<<<4 * NT, kB, 0, s>>>(hsgrids_d.p, st_grid_d.p, z.p, mask.p, rcv.p, dir.p, z_before.p);
    if (n_bflats) kb_gather<<<blocks_for(n_bflats), kB, 0, s>>>(z.p, bflat.p, n_bflats, bel.p);
    LSDTT_LAUNCH_CHECK();
    batch_tile_order(th, btiles(), RT, rcv.p, mask.p, ord.p, lstart.p, maxdep.p, xdon.p, s);
    enqueue_counts(0, pl.count_passes, s);
    if (sp_modes) enqueue_sp(0, pl.sp_passes, s);
    const BatchHs B = bhs();
    hsb_start(B, G, s);
    if (any_hillslope)
        for (int k = 0; k < pl.rounds; ++k) hsb_round(B, G, s);
    enqueue_stats(s);
}

void BatchEngine::read_status(hipStream_t s)
{
    LSDTT_HIP(hipMemcpyAsync(ctl_h.data(), ctl.p, sizeof(HsCtl) * G, hipMemcpyDeviceToHost, s));
    LSDTT_HIP(hipMemcpyAsync(pflags_h.data(), pflags.p, sizeof(int) * 2 * kMaxPass, hipMemcpyDeviceToHost, s));
    LSDTT_HIP(hipStreamSynchronize(s));
}

namespace {
// first pass that marked no tile among [0, launched), or -1
int settled_at(const int* flags, int launched)
{
    for (int q = 0; q < launched; ++q)
        if (!flags[q]) return q;
    return -1;
}
bool solving(int ph) { return ph == HS_START || ph == HS_KRYLOV || ph == HS_SEARCH || ph == HS_WARM; }
}  // namespace

int BatchEngine::step(std::vector<LemStats>& st)
{
    const hipStream_t s = s_;
    // launch pattern: this step's caps
    if (use_graph) {
        if (!exec || !(captured == plan)) {
            release_graph();
            hipGraph_t graph = nullptr;
            LSDTT_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            try {
                enqueue_step(plan, s);
            } catch (...) {
                (void)hipStreamEndCapture(s, &graph);
                if (graph) (void)hipGraphDestroy(graph);
                throw;
            }
            LSDTT_HIP(hipStreamEndCapture(s, &graph));
            const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            LSDTT_HIP(e);
            captured = plan;
            ++counters.graph_captures;
        }
        LSDTT_HIP(hipGraphLaunch(exec, s));
        ++counters.graph_launches;
    } else {
        enqueue_step(plan, s);
    }
    counters.count_passes = plan.count_passes;
    counters.sp_passes = plan.sp_passes;
    counters.rounds = plan.rounds;
    read_status(s);
    int cset = settled_at(pflags_h.data(), plan.count_passes);
    int sset = sp_modes ? settled_at(pflags_h.data() + kMaxPass, plan.sp_passes) : 0;
    if (cset < 0 || sset < 0) {
        // a sweep did not settle within the pattern: the later phases ran on
        // unsettled values, so every grid restarts from its pre-step surface
        // and the sweeps run until they settle
        ++counters.refills;
        kb_copy<<<std::min(blocks_for(N), 8192), kB, 0, s>>>(z_before.p, z.p, N);
        LSDTT_LAUNCH_CHECK();
        // the abandoned pass recorded its own hillslope changes: the
        // restarted step starts every solve from z^n
        LSDTT_HIP(hipMemsetAsync(ws_ok.p, 0, sizeof(int) * G, s));
        kb_receivers<<<4 * NT, kB, 0, s>>>(hsgrids_d.p, st_grid_d.p, z.p, mask.p, rcv.p, dir.p, z_before.p);
        batch_tile_order(th, btiles(), RT, rcv.p, mask.p, ord.p, lstart.p, maxdep.p, xdon.p, s);
        int launched = 0;
        for (int batch = std::max(4, plan.count_passes);; batch *= 2) {
            if (launched + batch > kMaxPass) throw std::runtime_error("tile passes did not settle");
            enqueue_counts(launched, batch, s);
            launched += batch;
            counters.eager_passes += batch;
            read_status(s);
            cset = settled_at(pflags_h.data(), launched);
            if (cset >= 0) break;
        }
        launched = 0;
        if (sp_modes) {
            for (int batch = std::max(4, plan.sp_passes);; batch *= 2) {
                if (launched + batch > kMaxPass) throw std::runtime_error("tile passes did not settle");
                enqueue_sp(launched, batch, s);
                launched += batch;
                counters.eager_passes += batch;
                read_status(s);
                sset = settled_at(pflags_h.data() + kMaxPass, launched);
                if (sset >= 0) break;
            }
        }
        hsb_start(bhs(), G, s);
        read_status(s);
    }
    plan.count_passes = std::min(kMaxPatternPasses, std::max(plan.count_passes, cset + 3));
    plan.sp_passes = std::min(kMaxPatternPasses, std::max(plan.sp_passes, sset + 3));
    int picard_count = 0;
    finish_step(st, s, &picard_count);
    return picard_count;
}

// grids still solving continue round by round; failed Newton solves take
// the damped Picard fallback; statistics of the grids not yet counted
void BatchEngine::finish_step(std::vector<LemStats>& st, hipStream_t s, int* picard_count)
{
    const BatchHs B = bhs();
    int more = 0;
    for (;;) {
        // launches sized for the grids still solving (their lists only shrink)
        int tiles = 0, chunks = 0;
        for (int i = 0; i < G; ++i)
            if (solving(ctl_h[static_cast<size_t>(i)].phase)) {
                tiles += hsgrids_h[static_cast<size_t>(i)].nst;
                chunks += hsgrids_h[static_cast<size_t>(i)].ngc;
            }
        if (tiles + chunks == 0) break;
        // 2, 4, 8, ... rounds between status reads, at most 64 (a read is
        // one synchronisation; rounds after a grid finishes are wasted)
        const int k = std::max(2, more);
        for (int j = 0; j < k; ++j) hsb_round(B, G, s, tiles, chunks);
        counters.eager_rounds += k;
        more = std::min(2 * k, 64);
        read_status(s);
    }
    int max_rounds = 0;
    bool pending = false;
    for (int i = 0; i < G; ++i) {
        const HsCtl& c = ctl_h[static_cast<size_t>(i)];
        max_rounds = std::max(max_rounds, c.rounds);
        if (c.phase == HS_FAILED) {
            // compute_roering_hillslope_transport_implicit from the
            // post-incision surface (buffer 0), as run_lem_step falls back
            const LemParams& p = params[static_cast<size_t>(i)];
            const BatchHsGrid& h = hsgrids_h[static_cast<size_t>(i)];
            PicardOptions po = p.picard;
            po.D = p.D;
            po.U = 0.0;
            po.Sc = p.Sc;
            po.nonlinear = p.nonlinear;
            PicardStats ps;
            picard.run(h.nr, h.nc, h.bc, h.h, mask.p + h.off, dof.p + h.off, po, p.dt, z.p + h.off, &ps, s);
            kb_picard_done<<<1, 64, 0, s>>>(ctl.p, i);
            LSDTT_LAUNCH_CHECK();
            ++*picard_count;
        }
        if (!c.stats_done) pending = true;
    }
    if (pending) enqueue_stats(s);
    // rounds of the next pattern: what the slowest grid needed plus a
    // quarter (a changed pattern is captured again, so it grows with a
    // margin and shrinks only when well above the need; an empty round
    // costs a few short launches)
    const int want = max_rounds + 2;
    if (want > plan.rounds) plan.rounds = want + std::max(2, want / 4);
    else if (want < plan.rounds / 2) plan.rounds = std::max(want + 2, plan.rounds * 3 / 4);
    plan.rounds = std::min(plan.rounds, kMaxPatternRounds);
    LSDTT_HIP(hipMemcpyAsync(ctl_h.data(), ctl.p, sizeof(HsCtl) * G, hipMemcpyDeviceToHost, s));
    LSDTT_HIP(hipMemcpyAsync(stats_h.data(), stats_out.p, sizeof(double) * 8 * G, hipMemcpyDeviceToHost, s));
    LSDTT_HIP(hipMemcpyAsync(eroded_h.data(), eroded_out.p, sizeof(long long) * G, hipMemcpyDeviceToHost, s));
    LSDTT_HIP(hipMemcpyAsync(sel_h.data(), sel.p, sizeof(unsigned) * 4 * G, hipMemcpyDeviceToHost, s));
    LSDTT_HIP(hipStreamSynchronize(s));
    st.assign(static_cast<size_t>(G), LemStats{});
    const int cp = settled_at(pflags_h.data(), kMaxPass) + 1;
    const int spp = sp_modes ? settled_at(pflags_h.data() + kMaxPass, kMaxPass) + 1 : 0;
    for (int i = 0; i < G; ++i) {
        LemStats& o = st[static_cast<size_t>(i)];
        const HsCtl& c = ctl_h[static_cast<size_t>(i)];
        const LemParams& p = params[static_cast<size_t>(i)];
        const double* h = stats_h.data() + 8 * i;
        const long long nv = static_cast<long long>(h[7]);
        const double inv_dt = 1.0 / p.dt, inv_n = nv > 0 ? 1.0 / static_cast<double>(nv) : 0.0;
        o.n_valid = nv;
        o.delta_max_elevation = std::fabs(h[3] - h[2]);
        o.delta_mean_elevation = std::fabs((h[1] - h[0]) * inv_n);
        o.max_local_elevation_change = h[4];
        o.mean_erosion_rate = h[5] * inv_n * inv_dt;
        o.max_erosion_rate = h[6] * inv_dt;
        o.p999_local_elevation_change = 0.0;
        if (nv > 0) {
            float pvv;
            std::memcpy(&pvv, &sel_h[static_cast<size_t>(4 * i)], sizeof(float));
            o.p999_local_elevation_change = pvv;
        }
        o.cells_eroded = p.enable_fluvial ? eroded_h[static_cast<size_t>(i)] : 0;
        o.warm_start_mode = c.ws_tried ? ws_mode : 0;
        o.warm_start_accepted = c.ws_accepted;
        o.warm_start_residual = c.ws_f;
        o.count_passes = cp;
        o.stream_power_passes = spp;
        if (p.enable_hillslope) {
            o.newton_iterations = c.newton_total;
            o.krylov_iterations = c.krylov_total;
            o.substeps = c.substeps;
            o.rejected = c.rejected;
            o.hillslope_converged = c.method == 0 ? 1 : 0;
            o.hillslope_initial_residual = c.f_init;
            o.hillslope_residual = c.fnorm;
            o.hillslope_tolerance = c.tol;
            o.max_slope_ratio = c.ratio;
            o.hillslope_method = c.method;
            o.restoration_iterations = c.restoration_its;
        }
    }
}

int BatchEngine::autotune_tiles(int steps, double* ms_per_step)
{
    const hipStream_t s = s_;
    if (steps < 1) throw std::invalid_argument("autotune needs at least one step per tile height");
    dbuf<double> saved;
    saved.reserve(static_cast<size_t>(N));
    LSDTT_HIP(hipMemcpyAsync(saved.p, z.p, sizeof(double) * N, hipMemcpyDeviceToDevice, s));
    hipEvent_t e0, e1;
    LSDTT_HIP(hipEventCreate(&e0));
    LSDTT_HIP(hipEventCreate(&e1));
    constexpr double kMargin = 0.03;
    const int cur = th;
    int order[3] = {cur, 0, 0};
    for (int c = 0, k = 1; c < 3; ++c)
        if (FluvialTiles::kTileHeights[c] != cur) order[k++] = FluvialTiles::kTileHeights[c];
    int best = cur;
    double best_ms = 0.0, cur_ms = 0.0;
    std::vector<LemStats> st;
    try {
        for (int c = 0; c < 3; ++c) {
            set_tile_height(order[c]);
            step(st);  // settles the pattern (and captures it)
            LSDTT_HIP(hipMemcpyAsync(z.p, saved.p, sizeof(double) * N, hipMemcpyDeviceToDevice, s));
            LSDTT_HIP(hipEventRecord(e0, s));
            for (int k = 0; k < steps; ++k) step(st);
            LSDTT_HIP(hipEventRecord(e1, s));
            LSDTT_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            LSDTT_HIP(hipEventElapsedTime(&ms, e0, e1));
            const double per = static_cast<double>(ms) / steps;
            if (ms_per_step) ms_per_step[order[c] == 8 ? 0 : (order[c] == 16 ? 1 : 2)] = per;
            if (c == 0) {
                cur_ms = best_ms = per;
            } else if (per < best_ms && per < (1.0 - kMargin) * cur_ms) {
                best_ms = per;
                best = order[c];
            }
            LSDTT_HIP(hipMemcpyAsync(z.p, saved.p, sizeof(double) * N, hipMemcpyDeviceToDevice, s));
        }
    } catch (...) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    set_tile_height(best);
    // the trials' hillslope changes belong to the discarded surfaces
    LSDTT_HIP(hipMemsetAsync(ws_ok.p, 0, sizeof(int) * G, s));
    LSDTT_HIP(hipStreamSynchronize(s));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return best;
}

}  // namespace lsdtt_gpu
