// lem.hip -- the LEM timestep kernels and host orchestration (see lem.hpp).
//
// Reference semantics restated (src/cpu/):
//   lsdtt_stream_power_cpu.cpp:159-326  D8 implicit stream power: skip
//       outlets, cells below the area threshold and cells whose (already
//       updated) receiver is not lower; n == 1 closed form with the rise
//       guard; n != 1 per-node Newton (<= 100 iterations, stop at
//       slope_clamp * dx_eff) with the rise guard and over-excavation clamp
//   lsdtt_hillslope_operator_cpu.hpp:83-321  'o' edges Dirichlet, 'n'/'b'
//       and masked neighbours resolve to the cell itself, 'p' wraps; face
//       alpha = 1 / (1 - min((Sn^2 + St^2) / Sc^2, cap)), St = mean of the
//       two cells' centred tangential slopes
//   lsdtt_hillslope_newton_cpu.cpp:137-186  dalpha/dz (9-point Jacobian)
//   lsdtt_hillslope_newton_cpu.cpp:283-308  residual
//   lsdtt_hillslope_newton_cpu.cpp:553-633  line search (halving; the
//       centred-gradient feasibility test; restoration mode when the start
//       is super-critical)
//   lsdtt_hillslope_newton_cpu.cpp:690-788  sub-steps (grow x1.7 at <= 4
//       iterations, shrink x0.5 at >= 15, reject and halve on failure)
//   lsdtt_lem_cpu.cpp:614-722  uplift, boundary restore, statistics
//   src/utils/lsdtt_statistics.cpp:420-429  p99.9 = element floor(0.999 (n-1))
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "hillslope.hpp"
#include "lem.hpp"
#include "reduce.hpp"

namespace lsdtt_gpu {

namespace {

constexpr int kB = 256;

// scalar slots (double) of the hillslope solver
// statistics occupy 8 slots after the hillslope solver's scalars
constexpr int kStatSlot = HS_NSCAL;
constexpr int kScalCount = HS_NSCAL + 8;

inline int grid_for(long long n, int cap) { return std::max(1, std::min(cap, div_up(n, kB))); }

__global__ void k_copy(const double* __restrict__ a, double* __restrict__ b, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        b[i] = a[i];
}

// out = float(a - b)
__global__ void k_diff_f(const double* __restrict__ a, const double* __restrict__ b, float* __restrict__ out,
                         long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        out[i] = static_cast<float>(a[i] - b[i]);
}

// out = a * c (FP32)
__global__ void k_scale_f(const float* __restrict__ a, float c, float* __restrict__ out, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        out[i] = a[i] * c;
}

// channel set from the D-inf area (build_channel_mask, lsdtt_lem_cpu.cpp:
// 117-131; channel_threshold_quantity, lsdtt_stream_power_cpu.cpp:613-619),
// and the DOF and fast sets without it
__global__ void k_channel_dofs(const uint8_t* __restrict__ mask, const double* __restrict__ area, double thr,
                               double dx, int divide, const uint8_t* __restrict__ dof0,
                               const uint8_t* __restrict__ fast0, uint8_t* __restrict__ chan,
                               uint8_t* __restrict__ dof, uint8_t* __restrict__ fast, long long n)
{
    const double den = dx > 1e-12 ? dx : 1e-12;
    for (long long f = blockIdx.x * (long long)blockDim.x + threadIdx.x; f < n;
         f += (long long)gridDim.x * blockDim.x) {
        const uint8_t m = mask[f], d = dof0[f], q0 = fast0[f];
        const double a = area[f];
        const double q = divide ? a / den : a;
        const uint8_t c = (m && q >= thr) ? 1 : 0;
        chan[f] = c;
        dof[f] = c ? 0 : d;
        fast[f] = c ? 0 : q0;
    }
}

// the general DOF list without channel cells, in its original (row-major)
// order: one block, a running offset over chunks of blockDim.x entries;
// *count receives the new length
__global__ void __launch_bounds__(1024) k_gen_filter(const int* __restrict__ gen0, int n0,
                                                     const uint8_t* __restrict__ chan, int* __restrict__ gen,
                                                     int* __restrict__ count)
{
    __shared__ int wsum[16];
    __shared__ int base;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if (threadIdx.x == 0) base = 0;
    __syncthreads();
    for (int s0 = 0; s0 < n0; s0 += blockDim.x) {
        const int i = s0 + threadIdx.x;
        const int f = i < n0 ? gen0[i] : 0;
        const bool keep = i < n0 && !chan[f];
        const unsigned long long bal = __ballot(keep);
        const int pre = __popcll(bal & ((1ull << lane) - 1ull));
        if (lane == 0) wsum[wv] = __popcll(bal);
        __syncthreads();
        int off = base;
        for (int k = 0; k < wv; ++k) off += wsum[k];
        if (keep) gen[off + pre] = f;
        __syncthreads();
        if (threadIdx.x == 0) {
            int t = 0;
            for (int k = 0; k < nw; ++k) t += wsum[k];
            base += t;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) *count = base;
}

__global__ void k_gather(const double* __restrict__ z, const int* __restrict__ idx, long long m,
                         double* __restrict__ out)
{
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i < m) out[i] = z[idx[i]];
}

// ---------------------------------------------------------------------------
// uplift, restore, statistics
// ---------------------------------------------------------------------------
// cell flags: bit 0 the mask, bit 1 uplift (valid cells that are not
// boundary flats, whose elevation the step sets instead)
__global__ void k_cflag_init(const uint8_t* __restrict__ mask, uint8_t* __restrict__ cflag, long long n)
{
    for (long long f = blockIdx.x * (long long)blockDim.x + threadIdx.x; f < n;
         f += (long long)gridDim.x * blockDim.x)
        cflag[f] = mask[f] ? 3 : 0;
}

__global__ void k_cflag_flats(const int* __restrict__ idx, long long m, uint8_t* __restrict__ cflag)
{
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i < m) cflag[idx[i]] &= 1;
}

__global__ void k_restore(double* __restrict__ z, const int* __restrict__ idx, const double* __restrict__ bel,
                          long long m, double bdz)
{
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i < m) z[idx[i]] = bel[i] - bdz;
}

// exact k-th smallest of the float bit patterns (all >= 0): three radix
// passes of 11, 11 and 10 bits from the top
constexpr int kRadixBins = 2048;
constexpr int kRadixTopShift = 21;  // the first pass's digit: bits 21..31

// adds 1 to h[bin] for every lane with `want`; lanes that share a bin add
// once through their leader (|dz| values of one magnitude crowd into few
// bins)
__device__ __forceinline__ void wave_hist_add(unsigned* h, unsigned bin, bool want)
{
    const int lane = threadIdx.x & 63;
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

__global__ void k_clear_u32(unsigned* __restrict__ a, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0u;
}

// one cell of the statistics pass below; returns the cell's |dz| as stored
// (+inf for an invalid cell)
__device__ __forceinline__ float stats_cell(const uint8_t* __restrict__ cflag, const double* __restrict__ zb,
                                            const double* zsrc, double* za, long long f, bool moved, double udz,
                                            double inv_dt, float* __restrict__ absdz, float* __restrict__ ero_rate,
                                            double (&acc)[8])
{
    // loads first, then the flag tests (no wait between them)
    const uint8_t m = cflag[f];
    const double b = zb[f];
    double a = zsrc[f];
    if ((m & 2) && udz != 0.0) {
        a += udz;
        za[f] = a;
    } else if (moved) {
        za[f] = a;
    }
    if (!(m & 1)) {
        const float inf = __int_as_float(0x7f800000);
        absdz[f] = inf;
        if (ero_rate) ero_rate[f] = 0.0f;
        return inf;
    }
    acc[0] += b;
    acc[1] += a;
    acc[2] = fmax(acc[2], b);
    acc[3] = fmax(acc[3], a);
    const double l = fabs(a - b);
    acc[4] = fmax(acc[4], l);
    const float lf = static_cast<float>(l);
    absdz[f] = lf;
    const double ero = b - (a - udz);
    if (ero_rate) ero_rate[f] = static_cast<float>(ero * inv_dt);
    if (ero > 0.0) {
        acc[5] += ero;
        acc[6] = fmax(acc[6], ero);
    }
    acc[7] += 1.0;
    return lf;
}

// uplift (lsdtt_lem_cpu.cpp:617-625: z += U dt on the cells with bit 1
// of cflag, udz = 0 without uplift) fused with the step statistics
// (:655-722 sums and maxima; |dz| as float for p99.9, invalid cells +inf,
// which sorts after every valid value).  The reference uplifts every valid
// cell and then resets the boundary flats (:645-652); here the flats are
// reset first and carry no uplift bit, which leaves the same values.
// The step's surface is read from zsrc (the hillslope result, wherever the
// solve left it) and every cell is written to za, the context's state
// array, whose address therefore never changes between steps.  ero_rate
// (optional): the per-pixel erosion rate (erosion * inv_dt as float, 0 off
// the mask, lsdtt_lem_cpu.cpp:666-696).  Also the p99.9 selection's first
// radix histogram (the top 11 bits of every |dz|) into hist0.
__global__ void __launch_bounds__(256) k_stats(const uint8_t* __restrict__ cflag, const double* __restrict__ zb,
                                               const double* zsrc, double* za, long long n, double udz,
                                               double inv_dt, float* __restrict__ absdz,
                                               float* __restrict__ ero_rate, double* partials,
                                               unsigned* __restrict__ hist0)
{
    __shared__ unsigned h[kRadixBins];
    for (int i = threadIdx.x; i < kRadixBins; i += blockDim.x) h[i] = 0;
    __syncthreads();
    const bool moved = zsrc != za;
    double acc[8] = {0.0, 0.0, -1e30, -1e30, 0.0, 0.0, 0.0, 0.0};
    // wavefront-uniform trip count (the histogram is added per wavefront)
    for (long long base = blockIdx.x * (long long)blockDim.x + (threadIdx.x & ~63); base < n;
         base += (long long)gridDim.x * blockDim.x) {
        const long long f = base + (threadIdx.x & 63);
        float adz = 0.0f;
        const bool in = f < n;
        if (in) adz = stats_cell(cflag, zb, zsrc, za, f, moved, udz, inv_dt, absdz, ero_rate, acc);
        wave_hist_add(h, __float_as_uint(adz) >> kRadixTopShift, in);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kRadixBins; i += blockDim.x)
        if (h[i]) atomicAdd(&hist0[i], h[i]);
    const bool mx[8] = {false, false, true, true, true, false, true, false};
    block_partial<8>(acc, mx, partials, blockIdx.x);
}


__global__ void __launch_bounds__(256) k_stats_fin(const double* partials, unsigned count, double* out)
{
    const bool mx[8] = {false, false, true, true, true, false, true, false};
    double res[8];
    fold_partials<8>(partials, count, mx, res);
    if (threadIdx.x == 0)
        for (int i = 0; i < 8; ++i) out[i] = res[i];
}


__global__ void k_radix_hist(const float* __restrict__ v, long long n, const unsigned* __restrict__ sel, int shift,
                             int bits, unsigned* __restrict__ hist)
{
    __shared__ unsigned h[kRadixBins];
    for (int i = threadIdx.x; i < kRadixBins; i += blockDim.x) h[i] = 0;
    __syncthreads();
    const unsigned prefix = sel[0], pmask = sel[1];
    const int lane = threadIdx.x & 63;
    // wavefront-uniform trip count; lanes that share a bin add once through
    // their leader (|dz| values of one magnitude crowd into few bins)
    for (long long base = blockIdx.x * (long long)blockDim.x + (threadIdx.x & ~63); base < n;
         base += (long long)gridDim.x * blockDim.x) {
        const long long f = base + lane;
        unsigned bin = 0;
        bool want = false;
        if (f < n) {
            const unsigned u = __float_as_uint(v[f]);
            want = (u & pmask) == prefix;
            bin = (u >> shift) & ((1u << bits) - 1u);  // this pass's digit only
        }
        wave_hist_add(h, bin, want);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kRadixBins; i += blockDim.x)
        if (h[i]) atomicAdd(&hist[i], h[i]);
}

// sel = {prefix, prefix mask, k}; picks the bin holding rank k: one block
// of 256 threads, 8 bins each, block-wide exclusive scan of the counts, the
// bin whose cumulative range contains k wins; the histogram is cleared
constexpr int kPickThreads = 256, kPickPer = kRadixBins / kPickThreads;
__global__ void __launch_bounds__(kPickThreads) k_radix_pick(unsigned* __restrict__ hist, unsigned* __restrict__ sel,
                                                             int shift, int bits)
{
    __shared__ unsigned warp_tot[kPickThreads / 64];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int nb = 1 << bits;
    unsigned c[kPickPer], sum = 0;
#pragma unroll
    for (int j = 0; j < kPickPer; ++j) {
        const int b = t * kPickPer + j;
        c[j] = b < nb ? hist[b] : 0u;
        sum += c[j];
    }
    // inclusive scan of the per-thread sums across the wavefront
    unsigned x = sum;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) warp_tot[w] = x;
    __syncthreads();
    unsigned before = x - sum;
    for (int i = 0; i < w; ++i) before += warp_tot[i];
    const unsigned k = sel[2];
    __syncthreads();
    unsigned cum = before;
#pragma unroll
    for (int j = 0; j < kPickPer; ++j) {
        if (cum <= k && k < cum + c[j]) {  // exactly one bin holds rank k
            const int b = t * kPickPer + j;
            sel[0] |= static_cast<unsigned>(b) << shift;
            sel[1] |= ((1u << bits) - 1u) << shift;
            sel[2] = k - cum;
        }
        cum += c[j];
    }
#pragma unroll
    for (int j = 0; j < kPickPer; ++j) hist[t * kPickPer + j] = 0u;
}

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static FlowOptions lem_flow_options(int nr, int nc, double dx, BC bc)
{
    FlowOptions o;
    o.nr = nr;
    o.nc = nc;
    o.dx = dx;
    o.bc = bc;
    o.interior_nodata_is_outlet = true;   // lsdtt_lem_cpu.cpp:331-332
    o.allow_pits_as_outlets = true;
    return o;
}

LemEngine::LemEngine(int nr_, int nc_, double dx_, const uint8_t* mask_host, const LemParams& p_)
    : nr(nr_), nc(nc_), n((long long)nr_ * nc_), dx(dx_), p(p_), flow(lem_flow_options(nr_, nc_, dx_, p_.bc))
{
    flow.timer = &timer;
    tiles.timer = &timer;
    tiles.resize(nr, nc, p.bc);
    mask.reserve(n);
    dof.reserve(n);
    z.reserve(n);
    z_before.reserve(n);
    LSDTT_HIP(hipMemcpy(mask.p, mask_host, n, hipMemcpyHostToDevice));
    mask_h.assign(mask_host, mask_host + n);
    n_mask = 0;
    for (long long i = 0; i < n; ++i) n_mask += mask_host[i] != 0;
    cflag.reserve(n);
    k_cflag_init<<<grid_for(n, 4096), kB>>>(mask.p, cflag.p, n);
    LSDTT_LAUNCH_CHECK();
    LSDTT_HIP(hipDeviceSynchronize());  // the step's streams may not wait on the null stream
    build_dofs();
    hs_grid = hs_bicgstab_grid(n, 2048);
    hs_tgrid = grid_for(n, 2048);  // measured faster than the resident-only grid
    scal.reserve(kScalCount + 8);
    hpin.reserve(32);
    tickets.reserve(4);
    flags.reserve(12);  // [8] the filtered general count
    LSDTT_HIP(hipMemset(tickets.p, 0, sizeof(unsigned) * 4));
    LSDTT_HIP(hipMemset(scal.p, 0, sizeof(double) * (kScalCount + 8)));
}

LemEngine::~LemEngine()
{
    if (aux_stream) {
        (void)hipStreamSynchronize(aux_stream);
        (void)hipEventDestroy(ev_fork);
        (void)hipEventDestroy(ev_join);
        (void)hipStreamDestroy(aux_stream);
    }
}

// resolve_hillslope_coupling_regime (lsdtt_lem_cpu.cpp:64-77): a channel
// set exists when the regime is channels_as_boundaries, or automatic with a
// threshold or an explicit mask
bool LemEngine::channels_active() const
{
    const bool has_mask = !channel_h.empty();
    const int regime = p.regime != 0 ? p.regime : ((p.threshold > 0.0 || has_mask) ? 2 : 1);
    return regime == 2;
}

bool LemEngine::threshold_channels() const
{
    return p.enable_hillslope && channels_active() && channel_h.empty() && p.threshold > 0.0;
}

void LemEngine::build_dofs()
{
    // hillslope DOFs: valid, not on an 'o' (Dirichlet) edge
    // (lsdtt_hillslope_operator_cpu.hpp:136-142), not a channel pixel
    // (build_channel_mask, lsdtt_lem_cpu.cpp:85-134: an explicit mask here;
    // a threshold-derived set is removed each step by derive_channels)
    const bool use_channels = channels_active() && !channel_h.empty();
    std::vector<uint8_t> d(static_cast<size_t>(n));
    for (int r = 0; r < nr; ++r)
        for (int c = 0; c < nc; ++c) {
            const long long f = (long long)r * nc + c;
            const bool edge = (r == 0 && p.bc.n == 'o') || (r == nr - 1 && p.bc.s == 'o') ||
                              (c == 0 && p.bc.w == 'o') || (c == nc - 1 && p.bc.e == 'o');
            const bool ch = use_channels && channel_h[static_cast<size_t>(f)] != 0;
            d[static_cast<size_t>(f)] = (mask_h[static_cast<size_t>(f)] && !edge && !ch) ? 1 : 0;
        }
    dof.reserve(n);
    LSDTT_HIP(hipMemcpy(dof.p, d.data(), n, hipMemcpyHostToDevice));
    // fast DOFs: interior, whole 3x3 block valid; the rest go on the
    // general list handled by the companion kernels
    std::vector<uint8_t> fh(static_cast<size_t>(n), 0);
    std::vector<int> gh;
    for (int r = 0; r < nr; ++r)
        for (int c = 0; c < nc; ++c) {
            const long long f = (long long)r * nc + c;
            if (!d[static_cast<size_t>(f)]) continue;
            bool ok = r > 0 && r < nr - 1 && c > 0 && c < nc - 1;
            for (int k = 0; ok && k < 9; ++k) ok = mask_h[static_cast<size_t>(f + (k / 3 - 1) * nc + (k % 3 - 1))] != 0;
            if (ok) fh[static_cast<size_t>(f)] = 1;
            else gh.push_back(static_cast<int>(f));
        }
    fast.reserve(n);
    LSDTT_HIP(hipMemcpy(fast.p, fh.data(), n, hipMemcpyHostToDevice));
    n_gen = static_cast<int>(gh.size());
    gen.reserve(gh.size() ? gh.size() : 1);
    if (n_gen) LSDTT_HIP(hipMemcpy(gen.p, gh.data(), sizeof(int) * gh.size(), hipMemcpyHostToDevice));
    partials.reserve(static_cast<size_t>(grid_for(n, 2048) + hs_general_blocks(n_gen) + 1) * 8);
    c9g.reserve(static_cast<size_t>(9) * (n_gen ? n_gen : 1));
    if (p.fluvial_dinf || threshold_channels()) {
        drec.reserve(2 * (size_t)n);
        dw.reserve(2 * (size_t)n);
        darea.reserve(n);
    }
    if (threshold_channels()) {
        // the channel-free sets, from which each step removes its channels
        dof_base.reserve(n);
        fast_base.reserve(n);
        LSDTT_HIP(hipMemcpy(dof_base.p, dof.p, n, hipMemcpyDeviceToDevice));
        LSDTT_HIP(hipMemcpy(fast_base.p, fast.p, n, hipMemcpyDeviceToDevice));
        n_gen_base = n_gen;
        gen_base.reserve(gh.size() ? gh.size() : 1);
        if (n_gen) LSDTT_HIP(hipMemcpy(gen_base.p, gen.p, sizeof(int) * gh.size(), hipMemcpyDeviceToDevice));
        chan.reserve(n);
    }
}

void LemEngine::derive_channels(hipStream_t s)
{
    // D-inf graph of the post-incision surface as a float raster, interior
    // nodata as outlets (build_hillslope_flowgraph, lsdtt_lem_cpu.cpp:137-154)
    DinfOptions o;
    o.dx = dx;
    o.bc = p.bc;
    timer.begin("dinf_receivers", s);
    dinf_receivers<double>(z.p, mask.p, nr, nc, o, drec.p, dw.p, nullptr, nullptr, s);
    // z 8, mask 1 in; rec 8, w 8 out
    timer.end(s, 25.0 * n);
    timer.begin("dinf_order", s);
    tiles.dinf_order(drec.p, mask.p, s);
    timer.end(s, 9.0 * n);
    last_dinf_passes = tiles.dinf_accumulate(drec.p, dw.p, mask.p, dx * dx, darea.p, darea_warm, s);
    darea_warm = true;
    timer.begin("channel_dofs", s);
    const int g = grid_for(n, 4096);
    k_channel_dofs<<<g, kB, 0, s>>>(mask.p, darea.p, p.threshold, dx, p.divide_area_by_dx, dof_base.p,
                                    fast_base.p, chan.p, dof.p, fast.p, n);
    LSDTT_LAUNCH_CHECK();
    int* hn = reinterpret_cast<int*>(hpin.p + 30);
    k_gen_filter<<<1, 1024, 0, s>>>(gen_base.p, n_gen_base, chan.p, gen.p, flags.p + 8);
    LSDTT_LAUNCH_CHECK();
    LSDTT_HIP(hipMemcpyAsync(hn, flags.p + 8, sizeof(int), hipMemcpyDeviceToHost, s));
    LSDTT_HIP(hipStreamSynchronize(s));
    timer.end(s, 13.0 * n);
    n_gen = *hn;
}

void LemEngine::set_channel_mask(const uint8_t* ch)
{
    if (ch) channel_h.assign(ch, ch + n);
    else channel_h.clear();
    build_dofs();
}

void LemEngine::set_elevation(const double* zh)
{
    LSDTT_HIP(hipMemcpy(z.p, zh, sizeof(double) * n, hipMemcpyHostToDevice));
    ws_dz_valid = false;  // the previous step's hillslope change no longer applies
}

void LemEngine::get_elevation(double* zh)
{
    LSDTT_HIP(hipDeviceSynchronize());
    LSDTT_HIP(hipMemcpy(zh, z.p, sizeof(double) * n, hipMemcpyDeviceToHost));
}

void LemEngine::set_erodibility(const double* Kh)
{
    if (!Kh) {
        has_K = false;
        return;
    }
    K_field.reserve(n);
    LSDTT_HIP(hipMemcpy(K_field.p, Kh, sizeof(double) * n, hipMemcpyHostToDevice));
    has_K = true;
}

void LemEngine::set_boundary_flats(const long long* flats, long long m)
{
    n_bflats = m;
    bflats.reserve(m ? m : 1);
    bel.reserve(m ? m : 1);
    std::vector<int> v(static_cast<size_t>(m));
    for (long long i = 0; i < m; ++i) {
        if (flats[i] < 0 || flats[i] >= n) throw std::invalid_argument("boundary flat index out of range");
        v[static_cast<size_t>(i)] = static_cast<int>(flats[i]);
    }
    if (m) LSDTT_HIP(hipMemcpy(bflats.p, v.data(), sizeof(int) * m, hipMemcpyHostToDevice));
    k_cflag_init<<<grid_for(n, 4096), kB>>>(mask.p, cflag.p, n);
    LSDTT_LAUNCH_CHECK();
    if (m) {
        k_cflag_flats<<<div_up(m, kB), kB>>>(bflats.p, m, cflag.p);
        LSDTT_LAUNCH_CHECK();
    }
    LSDTT_HIP(hipDeviceSynchronize());
}

void LemEngine::dinf_stream_power(const SpParams& a, hipStream_t s)
{
    // run_lem_step's D-inf branch (lsdtt_lem_cpu.cpp:357-391): the D-inf
    // graph of the pre-step surface as a float raster, its accumulation,
    // the smoothing, then the sweep from the outlets up
    DinfOptions o;
    o.dx = dx;
    o.bc = p.bc;
    timer.begin("dinf_receivers", s);
    dinf_receivers<double>(z_before.p, mask.p, nr, nc, o, drec.p, dw.p, nullptr, nullptr, s);
    timer.end(s, 25.0 * n);
    timer.begin("dinf_order", s);
    tiles.dinf_order(drec.p, mask.p, s);
    timer.end(s, 9.0 * n);
    last_count_passes = tiles.dinf_accumulate(drec.p, dw.p, mask.p, dx * dx, darea.p, darea_warm, s);
    last_count_tiles = tiles.tile_runs;
    darea_warm = true;
    const double* area = darea.p;
    if (p.area_kernel > 1) {
        area_sm.reserve(n);
        tiles.smooth_area(darea.p, mask.p, p.area_kernel, area_sm.p, s);
        area = area_sm.p;
    }
    timer.begin("tile_depths", s);
    tiles.dinf_up_order(drec.p, mask.p, s);
    timer.end(s, 9.0 * n);
    long long er = 0;
    last_sp_passes = tiles.dinf_stream_power(drec.p, dw.p, mask.p, area, z_before.p, has_K ? K_field.p : nullptr, a,
                                             z.p, &er, s);
    last_sp_tiles = tiles.tile_runs;
    last_eroded = er;
}

void LemEngine::stream_power(hipStream_t s)
{
    if (p.fluvial_dinf) {
        SpParams a;
        a.K0 = p.K;
        a.m = p.m;
        a.nexp = p.n;
        a.dt = p.dt;
        a.dx = dx;
        a.pixel_area = dx * dx;
        a.threshold = p.threshold;
        a.slope_clamp = p.slope_clamp;
        a.divide = p.divide_area_by_dx;
        a.n_is_one = std::fabs(p.n - 1.0) < 1e-6;
        a.m_half = p.m == 0.5;
        a.n1_half = 0;
        a.width = p.width_factor;
        dinf_stream_power(a, s);
        return;
    }
    // receivers of the pre-step surface, then counts and stream power by
    // tile passes (fluvial.hip)
    // the receivers read the pre-step surface from z and leave its copy in
    // z_before (the step skips its own copy when this path runs)
    // Once the tiles hold the level order of the last receivers and the
    // counts are those receivers' counts, only the tiles where a receiver
    // changed (or a changed receiver leads in or out) need a new order, and
    // the counts' passes start from those tiles alone: the others' counts
    // hold until a changed inflow re-runs them, which the passes do.
    const bool incremental = counts_warm && tiles.d8_order_valid();
    TileMark mark;
    if (incremental) {
        tiles.clear_dirty(s);
        mark = TileMark{tiles.dirty.p, nc, tiles.tile_height(), tiles.ntx};
    }
    flow.receivers_only<double>(z.p, mask.p, s, z_before.p, mark);
    timer.begin("tile_depths", s);
    tiles.depths(flow.rcv.p, mask.p, s, incremental);
    timer.end(s, 9.0 * n);  // rcv 4, mask 1 in; order 2, level starts 2 out
    flow.cnt.reserve(n);
    last_count_passes = tiles.counts(flow.rcv.p, mask.p, flow.cnt.p, counts_warm, s, incremental);
    counts_warm = true;
    last_count_tiles = tiles.tile_runs;
    SpParams a;
    a.K0 = p.K;
    a.m = p.m;
    a.nexp = p.n;
    a.dt = p.dt;
    a.dx = dx;
    a.pixel_area = dx * dx;
    a.threshold = p.threshold;
    a.slope_clamp = p.slope_clamp;
    a.divide = p.divide_area_by_dx;
    a.n_is_one = std::fabs(p.n - 1.0) < 1e-6;
    a.m_half = p.m == 0.5;
    a.n1_half = p.n - 1.0 == 0.5;
    a.width = p.width_factor;
    const double* area = nullptr;
    if (p.area_kernel > 1) {  // smooth_drainage_area (lsdtt_stream_power_cpu.cpp:34-139)
        area_sm.reserve(n);
        tiles.smooth_area(flow.cnt.p, mask.p, dx * dx, p.area_kernel, area_sm.p, s);
        area = area_sm.p;
    }
    long long er = 0;
    last_sp_passes = tiles.stream_power(flow.rcv.p, flow.dir.p, mask.p, flow.cnt.p, area, z_before.p,
                                        has_K ? K_field.p : nullptr, a, z.p, &er, s);
    last_sp_tiles = tiles.tile_runs;
    last_eroded = er;
}

HsDom LemEngine::hs_domain(double inv_dt) const
{
    HsDom d{nr, nc, p.bc, mask.p, dof.p, fast.p, gen.p, n_gen, dx, p.D, 0.0, p.Sc, p.slope_ratio_cap, inv_dt,
            p.nonlinear, zn};
    d.c9g = c9g.p;
    return with_reciprocals(d);
}

int LemEngine::bicgstab(hipStream_t s, int newton_it, double eta)
{
    BiArgs a;
    a.d = hs_domain(cur_inv_dt);
    a.n = n;
    a.z = zc;
    a.dinv = dinv.p;
    a.b = b.p;
    a.x = x.p;
    a.r = r.p;
    a.rhat = b.p;  // rhat = r_0 = b, unchanged during the solve
    a.t = t.p;
    a.scal = scal.p;
    a.partials = partials.p;
    a.ticket = tickets.p;
    a.flags = flags.p;
    a.eta2 = eta * eta;
    a.max_it = p.krylov_max_iterations;
    a.first = 0;
    // the start (rho = ||b||^2, flags) was set by the trial that produced b
    // iterations are enqueued in chunks between flag read-backs; an
    // iteration after convergence returns at once (but still costs its
    // launches).  The first chunk is what the same Newton iteration of the
    // previous solve needed (these well-conditioned systems repeat their
    // counts from step to step), then one more at a time, growing.
    const int slot = std::min(newton_it, kKrylovHist - 1);
    int chunk = std::max(1, krylov_hist[slot]);
    int* hf = reinterpret_cast<int*>(hpin.p + 16);
    hf[0] = hf[1] = hf[2] = 0;
    int it_launched = 0;
    for (int done = 0, round = 0; !done; ++round, chunk = round == 1 ? 1 : std::min(2 * chunk, 16)) {
        for (int k = 0; k < chunk; ++k) {
            BiArgs ai = a;
            // p and v alternate between two buffers per iteration
            const bool odd = it_launched & 1;
            ai.p_old = odd ? pv.p : pv2.p;
            ai.p = odd ? pv2.p : pv.p;
            ai.v_old = odd ? v.p : v2.p;
            ai.v = odd ? v2.p : v.p;
            ai.first = (it_launched++ == 0) ? 1 : 0;
            hs_bicgstab_iteration(ai, hs_grid, s, timer.enabled ? &timer : nullptr);
        }
        LSDTT_HIP(hipMemcpyAsync(hf, flags.p, sizeof(int) * 3, hipMemcpyDeviceToHost, s));
        LSDTT_HIP(hipStreamSynchronize(s));
        done = hf[0];
    }
    krylov_hist[slot] = hf[1];
    if (hf[1] == 0)  // no iteration ran (b = 0): the correction is zero
        LSDTT_HIP(hipMemsetAsync(x.p, 0, sizeof(float) * n, s));
    return hf[1];
}

void LemEngine::trial(double lambda, double* fnorm, double* ratio, hipStream_t s)
{
    timer.begin("hillslope_residual", s);
    if (!aux_stream) {
        LSDTT_HIP(hipStreamCreateWithFlags(&aux_stream, hipStreamNonBlocking));
        LSDTT_HIP(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
        LSDTT_HIP(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
    }
    hs_trial(hs_domain(cur_inv_dt), zc, x.p, lambda, n, zt, b.p, dinv.p, partials.p, tickets.p, scal.p, flags.p,
             hs_tgrid, s, aux_stream, ev_fork, ev_join);
    // z 8, zn 8, fast 1, dof 1 in; b 4, dinv 4 out; a trial step adds dz 4 in, z_out 8 out
    timer.end(s, (lambda != 0.0 ? 38.0 : 26.0) * n);
    double* h = hpin.p;
    LSDTT_HIP(hipMemcpyAsync(h, scal.p + HS_TRIAL_F, sizeof(double) * 2, hipMemcpyDeviceToHost, s));
    LSDTT_HIP(hipStreamSynchronize(s));
    *fnorm = h[0];
    *ratio = h[1];
}

// compute_hillslope_newton on the resident z (z^n in zn_hs); returns the
// iteration count, or -1 when it did not converge
int LemEngine::newton_solve(double inv_dt, LemStats* st, hipStream_t s)
{
    cur_inv_dt = inv_dt;
    double fnorm, ratio_now;
    trial(0.0, &fnorm, &ratio_now, s);
    // relative to ||F(z^n)||inf; the hillslope's own uplift term is zero
    // (uplift is applied after the hillslope, lsdtt_lem_cpu.cpp:506)
    const double tol = std::max(p.newton_abs_tol, p.newton_rel_tol * fnorm);
    if (st->substeps == 0 && st->rejected == 0) st->hillslope_initial_residual = fnorm;
    st->hillslope_tolerance = tol;
    const double feasible = std::sqrt(p.slope_ratio_cap);
    const bool ws_usable = ws_mode != 3 || (ws_dz_valid && ws_dz_dt == p.dt);
    if (ws_mode != 0 && ws_usable && st->substeps == 0 && st->rejected == 0 && fnorm > tol) {
        // predicted start z^n + dz; plain Newton from z^n stays the fallback
        timer.begin("warm_start", s);
        if (ws_mode == 1) {
            // explicit Euler: b = -F(z^n) from the trial above
            k_scale_f<<<grid_for(n, 4096), kB, 0, s>>>(b.p, static_cast<float>(1.0 / inv_dt), x.p, n);
            LSDTT_LAUNCH_CHECK();
        } else if (ws_mode == 3) {
            // the previous step's hillslope change over the same dt
            LSDTT_HIP(hipMemcpyAsync(x.p, ws_dz.p, sizeof(float) * n, hipMemcpyDeviceToDevice, s));
        } else {
            if (!ws) throw std::logic_error("learned warm start without a predictor");
            ws->predict(zc, mask.p, dof.p, nr, nc, dx, p.D / (inv_dt * dx * dx), p.Sc, x.p, s);
        }
        timer.end(s, 0.0);
        double fw = 0.0, rw = 0.0;
        trial(1.0, &fw, &rw, s);
        st->warm_start_mode = ws_mode;
        st->warm_start_residual = fw;
        if (fw < fnorm && (!p.nonlinear || rw <= feasible)) {
            zc = zt;
            zt = trial_target();
            fnorm = fw;
            ratio_now = rw;
            st->warm_start_accepted = 1;
        } else {
            trial(0.0, &fnorm, &ratio_now, s);  // the system at z^n again
        }
    }
    int it = 0;
    bool converged = false;
    double eta = p.krylov_rel_tol, fprev = 0.0;
    for (; it < p.newton_max_iterations; ++it) {
        if (fnorm <= tol) {
            converged = true;
            break;
        }
        if (p.krylov_forcing == 1 && it > 0) eta = forcing_eta(eta, fnorm, fprev, tol);
        fprev = fnorm;
        // b and dinv at the iterate come from the trial that accepted it
        st->krylov_iterations += bicgstab(s, it, eta);
        // line search (newton .cpp:553-633)
        const bool restoring = p.nonlinear && ratio_now > feasible;
        if (restoring) ++st->restoration_iterations;
        double lambda = 1.0, fnorm_try = fnorm, ratio_try = 0.0;
        bool accepted = false;
        for (int ls = 0; ls < p.newton_max_line_search; ++ls) {
            trial(lambda, &fnorm_try, &ratio_try, s);
            if (p.nonlinear) {
                if (restoring) {
                    if (ratio_try >= ratio_now) {
                        lambda *= 0.5;
                        continue;
                    }
                    accepted = true;
                    break;
                }
                if (ratio_try > feasible) {
                    lambda *= 0.5;
                    continue;
                }
            }
            if (fnorm_try < fnorm) {
                accepted = true;
                break;
            }
            lambda *= 0.5;
        }
        if (!accepted) break;
        zc = zt;  // the accepted trial point is the new iterate
        zt = trial_target();
        fnorm = fnorm_try;
        ratio_now = ratio_try;
        if (fnorm <= tol) {
            converged = true;
            ++it;
            break;
        }
    }
    st->newton_iterations += it;
    st->hillslope_residual = fnorm;
    st->max_slope_ratio = ratio_now;
    return converged ? it : -1;
}

double* LemEngine::trial_target() const
{
    for (double* q : {z_trial.p, zn_hs.p, z_hs0.p})
        if (q != zc && q != zn) return q;
    throw std::logic_error("no free hillslope trial buffer");
}

// compute_hillslope_newton_transient (newton .cpp:690-788)
int LemEngine::hillslope(LemStats* st, hipStream_t s)
{
    const size_t nn = static_cast<size_t>(n);
    for (auto* bf : {&dinv, &b, &x, &r, &pv, &pv2, &v, &v2, &t}) bf->reserve(nn);
    zn_hs.reserve(nn);
    z_trial.reserve(nn);
    z_hs0.reserve(nn);
    const double total = p.dt;
    double dt = total;
    const double dt_min = total * 1e-9, dt_max = total;
    double remaining = total;
    // a trial writes every cell of its target (the DOFs from the line
    // search, the others from the iterate), so the trial buffers need no
    // initial copy; a rejected sub-step goes back to its start and an
    // accepted one starts the next by moving pointers
    zn = zc = z.p;
    zt = trial_target();
    int rc = 0;
    while (remaining > 0.0) {
        const double step = std::min(dt, remaining);
        const int its = newton_solve(1.0 / step, st, s);
        if (its < 0) {
            zc = zn;
            zt = trial_target();
            ++st->rejected;
            if (st->rejected > 40 || step <= dt_min) {
                st->hillslope_converged = 0;
                rc = 1;
                break;
            }
            dt = std::max(dt_min, step * 0.5);
            continue;
        }
        remaining -= step;
        if (remaining > 0.0) {
            zn = zc;
            zt = trial_target();
        }
        ++st->substeps;
        if (its <= 4) dt = std::min(dt_max, step * 1.7);
        else if (its >= 15) dt = std::max(dt_min, step * 0.5);
    }
    // z itself is never a trial target, so it still holds the post-incision
    // surface; the result stays where the solve left it (the statistics
    // pass writes it into z)
    hs_result = rc == 0 ? zc : z.p;
    zn = zc = zt = nullptr;
    return rc;
}

// the p99.9 selection of the k-th smallest |dz| (k from the valid count),
// enqueued with its read-back into sel
void LemEngine::enqueue_p999(long long nv, unsigned* sel, hipStream_t s, bool first_pass_done)
{
    sel[0] = 0u;
    sel[1] = 0u;
    sel[2] = static_cast<unsigned>(std::floor(0.999 * static_cast<double>(nv - 1)));
    unsigned* dsel = hist.p + kRadixBins;
    if (!first_pass_done) {
        k_clear_u32<<<div_up(kRadixBins, kB), kB, 0, s>>>(hist.p, kRadixBins);
        LSDTT_LAUNCH_CHECK();
    }
    LSDTT_HIP(hipMemcpyAsync(dsel, sel, sizeof(unsigned) * 3, hipMemcpyHostToDevice, s));
    const int shifts[3] = {kRadixTopShift, 10, 0}, bits[3] = {11, 11, 10};
    for (int pass = 0; pass < 3; ++pass) {
        // the first histogram may come from the statistics pass
        if (pass > 0 || !first_pass_done)
            k_radix_hist<<<grid_for(n, 2048), kB, 0, s>>>(absdz.p, n, dsel, shifts[pass], bits[pass], hist.p);
        k_radix_pick<<<1, kPickThreads, 0, s>>>(hist.p, dsel, shifts[pass], bits[pass]);
        LSDTT_LAUNCH_CHECK();
    }
    LSDTT_HIP(hipMemcpyAsync(sel, dsel, sizeof(unsigned) * 3, hipMemcpyDeviceToHost, s));
}

void LemEngine::stats(LemStats* st, const double* zres, hipStream_t s)
{
    absdz.reserve(n);
    hist.reserve(kRadixBins + 8);
    if (p.erosion_rate_field) ero_rate.reserve(n);
    const double udz = p.enable_uplift ? p.U * p.dt : 0.0;
    timer.begin("step_stats", s);
    k_clear_u32<<<div_up(kRadixBins, kB), kB, 0, s>>>(hist.p, kRadixBins);
    k_stats<<<hs_grid, kB, 0, s>>>(cflag.p, z_before.p, zres, z.p, n, udz, 1.0 / p.dt, absdz.p,
                                   p.erosion_rate_field ? ero_rate.p : nullptr, partials.p, hist.p);
    k_stats_fin<<<1, kB, 0, s>>>(partials.p, (unsigned)hs_grid, scal.p + kStatSlot);
    timer.end(s, 21.0 * n);
    LSDTT_LAUNCH_CHECK();
    double* h = hpin.p + 4;
    LSDTT_HIP(hipMemcpyAsync(h, scal.p + kStatSlot, sizeof(double) * 8, hipMemcpyDeviceToHost, s));
    // the valid cells are the mask's, fixed for the context: the selection
    // is enqueued behind the sums and both come back with one wait
    unsigned* sel = reinterpret_cast<unsigned*>(hpin.p + 24);
    if (n_mask > 0) enqueue_p999(n_mask, sel, s, true);
    LSDTT_HIP(hipStreamSynchronize(s));
    const long long nv = static_cast<long long>(h[7]);
    if (nv != n_mask && nv > 0) {  // not expected: the statistics count the mask's cells
        enqueue_p999(nv, sel, s, false);
        LSDTT_HIP(hipStreamSynchronize(s));
    }
    const double inv_dt = 1.0 / p.dt, inv_n = nv > 0 ? 1.0 / static_cast<double>(nv) : 0.0;
    st->n_valid = nv;
    st->delta_max_elevation = std::fabs(h[3] - h[2]);
    st->delta_mean_elevation = std::fabs((h[1] - h[0]) * inv_n);
    st->max_local_elevation_change = h[4];
    st->mean_erosion_rate = h[5] * inv_n * inv_dt;
    st->max_erosion_rate = h[6] * inv_dt;
    st->p999_local_elevation_change = 0.0;
    if (nv > 0) {
        float pv;
        std::memcpy(&pv, &sel[0], sizeof(float));
        st->p999_local_elevation_change = pv;
    }
}

int LemEngine::step(LemStats* st, hipStream_t s)
{
    *st = LemStats{};
    const int g = grid_for(n, 4096);
    if (!(p.enable_fluvial && !p.fluvial_dinf)) {  // the D8 receivers copy z themselves
        k_copy<<<g, kB, 0, s>>>(z.p, z_before.p, n);
        LSDTT_LAUNCH_CHECK();
    }
    if (n_bflats) {
        k_gather<<<div_up(n_bflats, kB), kB, 0, s>>>(z.p, bflats.p, n_bflats, bel.p);
        LSDTT_LAUNCH_CHECK();
    }
    if (p.enable_fluvial) {
        stream_power(s);
        st->count_passes = last_count_passes;
        st->stream_power_passes = last_sp_passes;
        st->count_tile_runs = last_count_tiles;
        st->stream_power_tile_runs = last_sp_tiles;
    }
    int rc = 0;
    if (threshold_channels()) derive_channels(s);
    // the step's surface after the hillslope: z, or the trial buffer that
    // holds the accepted Newton iterate
    const double* zres = z.p;
    if (p.enable_hillslope) {
        // Newton first; damped Picard when it fails (from the post-incision
        // surface, the Newton result discarded) or when it is forced
        // (lsdtt_lem_cpu.cpp:531-582)
        bool done = false;
        if (!p.force_picard) {
            done = hillslope(st, s) == 0;  // z is left post-incision when it fails
            if (done) zres = hs_result;
        }
        if (ws_mode == 3) {
            // this step's hillslope change, the next step's predicted start
            // (z still holds the post-incision surface)
            ws_dz_valid = done;
            if (done) {
                ws_dz.reserve(static_cast<size_t>(n));
                k_diff_f<<<g, kB, 0, s>>>(zres, z.p, ws_dz.p, n);
                LSDTT_LAUNCH_CHECK();
                ws_dz_dt = p.dt;
            }
        }
        if (capture) {
            // the training pair keeps the post-incision surface (z) in z_hs0;
            // when the result sits in z_hs0 the two internal buffers trade
            // storage first
            z_hs0.reserve(static_cast<size_t>(n));
            if (zres == z_hs0.p) {
                z_trial.reserve(static_cast<size_t>(n));
                std::swap(z_hs0, z_trial);
            }
            k_copy<<<g, kB, 0, s>>>(z.p, z_hs0.p, n);
            LSDTT_LAUNCH_CHECK();
        }
        if (!done) {
            PicardOptions po = p.picard;
            po.D = p.D;
            po.U = 0.0;
            po.Sc = p.Sc;
            po.nonlinear = p.nonlinear;
            timer.begin("picard_hillslope", s);
            picard.run(nr, nc, p.bc, dx, mask.p, dof.p, po, p.dt, z.p, &st->picard, s);
            timer.end(s, 0.0);
            st->hillslope_method = 1;
            zres = z.p;
        }
    }
    if (p.enable_hillslope && capture) {
        cap_dz.reserve(static_cast<size_t>(n));
        k_diff_f<<<g, kB, 0, s>>>(zres, z_hs0.p, cap_dz.p, n);
        LSDTT_LAUNCH_CHECK();
        cap_kappa = p.D * p.dt / (dx * dx);
        cap_valid = true;
    }
    // boundary flats take their base level first; the uplift of every
    // other valid cell runs inside the statistics pass, which also writes
    // the result into z
    if (n_bflats) {
        k_restore<<<div_up(n_bflats, kB), kB, 0, s>>>(const_cast<double*>(zres), bflats.p, bel.p, n_bflats,
                                                     p.baselevel_fall * p.dt);
        LSDTT_LAUNCH_CHECK();
    }
    stats(st, zres, s);
    ero_valid = p.erosion_rate_field != 0;
    st->cells_eroded = p.enable_fluvial ? last_eroded : 0;
    return rc;
}

int LemEngine::autotune_tiles(int steps, double* ms_per_step, hipStream_t s)
{
    if (steps < 1) throw std::invalid_argument("autotune needs at least one step per tile height");
    dbuf<double> saved;
    saved.reserve(static_cast<size_t>(n));
    LSDTT_HIP(hipMemcpyAsync(saved.p, z.p, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    const bool cap = capture;
    capture = false;  // the trial steps are not training samples
    hipEvent_t e0, e1;
    LSDTT_HIP(hipEventCreate(&e0));
    LSDTT_HIP(hipEventCreate(&e1));
    // the current height first; another replaces it only when it is faster
    // by more than kMargin (trials of a few steps carry a few per cent of
    // noise)
    constexpr double kMargin = 0.03;
    const int cur = tiles.tile_height();
    int order[3] = {cur, 0, 0};
    for (int c = 0, k = 1; c < 3; ++c)
        if (FluvialTiles::kTileHeights[c] != cur) order[k++] = FluvialTiles::kTileHeights[c];
    int best = cur;
    double best_ms = 0.0, cur_ms = 0.0;
    LemStats st;
    try {
        for (int c = 0; c < 3; ++c) {
            const int th = order[c];
            tiles.set_tile_height(th);
            // one untimed step settles the pass and Krylov heuristics
            step(&st, s);
            LSDTT_HIP(hipMemcpyAsync(z.p, saved.p, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
            LSDTT_HIP(hipEventRecord(e0, s));
            for (int k = 0; k < steps; ++k) step(&st, s);
            LSDTT_HIP(hipEventRecord(e1, s));
            LSDTT_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            LSDTT_HIP(hipEventElapsedTime(&ms, e0, e1));
            const double per = static_cast<double>(ms) / steps;
            if (ms_per_step) ms_per_step[th == 8 ? 0 : (th == 16 ? 1 : 2)] = per;
            if (c == 0) {
                cur_ms = best_ms = per;
            } else if (per < best_ms && per < (1.0 - kMargin) * cur_ms) {
                best_ms = per;
                best = th;
            }
            LSDTT_HIP(hipMemcpyAsync(z.p, saved.p, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        }
    } catch (...) {
        capture = cap;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    capture = cap;
    tiles.set_tile_height(best);
    LSDTT_HIP(hipStreamSynchronize(s));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return best;
}

}  // namespace lsdtt_gpu
