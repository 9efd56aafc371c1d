// timestep.hip -- LEM timestep control (see timestep.hpp).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <iomanip>
#include <sstream>

#include "timestep.hpp"

namespace lsdtt_gpu {

namespace {

constexpr int kB = 256;
constexpr int kSelBins = 2048;  // 11-bit digits

// cardinal_central_gradient + gradient_magnitude
// (src/cpu/lsdtt_raster_central_difference_cpu.hpp:43-62) of the valid
// interior, as the FP64 bit patterns (order-preserving for values >= 0);
// every other cell gets the largest key.  count: the candidates (one atomic
// per block).
__global__ void __launch_bounds__(kB) k_grad_keys(const double* __restrict__ z, const uint8_t* __restrict__ mask,
                                                  int nr, int nc, double two_dx,
                                                  unsigned long long* __restrict__ keys,
                                                  unsigned long long* __restrict__ count)
{
    __shared__ unsigned wc[kB / 64];
    const long long n = (long long)nr * nc;
    unsigned c = 0;
    for (long long f = blockIdx.x * (long long)kB + threadIdx.x; f < n; f += (long long)gridDim.x * kB) {
        const int r = static_cast<int>(f / nc), col = static_cast<int>(f - (long long)r * nc);
        unsigned long long key = ~0ull;
        if (r > 0 && r < nr - 1 && col > 0 && col < nc - 1) {
            const uint8_t m = mask[f], mn = mask[f - nc], ms = mask[f + nc], mw = mask[f - 1], me = mask[f + 1];
            const double zn = z[f - nc], zs = z[f + nc], zw = z[f - 1], ze = z[f + 1];
            if (m && mn && ms && mw && me) {
                const double gx = (ze - zw) / two_dx, gy = (zs - zn) / two_dx;
                key = static_cast<unsigned long long>(__double_as_longlong(sqrt(gx * gx + gy * gy)));
                ++c;
            }
        }
        keys[f] = key;
    }
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
    if ((threadIdx.x & 63) == 0) wc[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned t = 0;
        for (int i = 0; i < kB / 64; ++i) t += wc[i];
        if (t) atomicAdd(count, static_cast<unsigned long long>(t));
    }
}

// sel: [0] prefix, [1] prefix mask, [2] rank still to find among the keys
// that match the prefix
__global__ void __launch_bounds__(kB) k_sel_hist(const unsigned long long* __restrict__ keys, long long n,
                                                 const unsigned long long* __restrict__ sel, int shift, int bits,
                                                 unsigned* __restrict__ hist)
{
    __shared__ unsigned h[kSelBins];
    for (int i = threadIdx.x; i < kSelBins; i += kB) h[i] = 0u;
    __syncthreads();
    const unsigned long long prefix = sel[0], pmask = sel[1];
    for (long long f = blockIdx.x * (long long)kB + threadIdx.x; f < n; f += (long long)gridDim.x * kB) {
        const unsigned long long k = keys[f];
        if ((k & pmask) == prefix) atomicAdd(&h[(k >> shift) & ((1u << bits) - 1u)], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kSelBins; i += kB)
        if (h[i]) atomicAdd(&hist[i], h[i]);
}

// the bin holding rank sel[2]; extends the prefix, lowers the rank and
// clears the histogram for the next pass
__global__ void __launch_bounds__(kB) k_sel_pick(unsigned* __restrict__ hist, unsigned long long* __restrict__ sel,
                                                 int shift, int bits)
{
    constexpr int kPer = kSelBins / kB;
    __shared__ unsigned long long wt[kB / 64];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int nb = 1 << bits;
    unsigned long long c[kPer], sum = 0;
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        const int b = t * kPer + j;
        c[j] = b < nb ? hist[b] : 0u;
        sum += c[j];
    }
    unsigned long long x = sum;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned long long y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) wt[w] = x;
    __syncthreads();
    unsigned long long before = x - sum;
    for (int i = 0; i < w; ++i) before += wt[i];
    const unsigned long long k = sel[2];
    __syncthreads();
    unsigned long long cum = before;
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        if (cum <= k && k < cum + c[j]) {
            const unsigned long long b = static_cast<unsigned long long>(t * kPer + j);
            sel[0] |= b << shift;
            sel[1] |= static_cast<unsigned long long>(nb - 1) << shift;
            sel[2] = k - cum;
        }
        cum += c[j];
    }
#pragma unroll
    for (int j = 0; j < kPer; ++j) hist[t * kPer + j] = 0u;
}

// key of rank k among n keys (k < number of keys): six digit passes
unsigned long long select_key(const unsigned long long* keys, long long n, unsigned long long k, unsigned* hist,
                              unsigned long long* dsel, hipStream_t s)
{
    static const int shifts[6] = {53, 42, 31, 20, 9, 0};
    static const int bits[6] = {11, 11, 11, 11, 11, 9};
    const unsigned long long init[3] = {0ull, 0ull, k};
    LSDTT_HIP(hipMemcpyAsync(dsel, init, sizeof(init), hipMemcpyHostToDevice, s));
    const int grid = std::max(1, std::min(2048, div_up(n, kB)));
    for (int p = 0; p < 6; ++p) {
        k_sel_hist<<<grid, kB, 0, s>>>(keys, n, dsel, shifts[p], bits[p], hist);
        LSDTT_LAUNCH_CHECK();
        k_sel_pick<<<1, kB, 0, s>>>(hist, dsel, shifts[p], bits[p]);
        LSDTT_LAUNCH_CHECK();
    }
    unsigned long long out[3];
    LSDTT_HIP(hipMemcpyAsync(out, dsel, sizeof(out), hipMemcpyDeviceToHost, s));
    LSDTT_HIP(hipStreamSynchronize(s));
    return out[0];
}

double key_value(unsigned long long k)
{
    double v;
    std::memcpy(&v, &k, sizeof v);
    return v;
}

}  // namespace

double median_interior_gradient(const double* z, const uint8_t* mask, int nr, int nc, double dx,
                                dbuf<unsigned long long>& keys, hipStream_t s)
{
    // median_interior_gradient (lsdtt_lem_timestep_cpu.cpp:28-50)
    if (nr < 3 || nc < 3 || !(dx > 0.0)) return 0.0;
    const long long n = (long long)nr * nc;
    keys.reserve(static_cast<size_t>(n) + 3);
    dbuf<unsigned> hist;
    hist.reserve(kSelBins);
    LSDTT_HIP(hipMemsetAsync(hist.p, 0, sizeof(unsigned) * kSelBins, s));
    unsigned long long* dcount = keys.p + n;  // 3 words after the keys: the count, then the selection state
    LSDTT_HIP(hipMemsetAsync(dcount, 0, sizeof(unsigned long long), s));
    const int grid = std::max(1, std::min(4096, div_up(n, kB)));
    k_grad_keys<<<grid, kB, 0, s>>>(z, mask, nr, nc, 2.0 * dx, keys.p, dcount);
    LSDTT_LAUNCH_CHECK();
    unsigned long long m = 0;
    LSDTT_HIP(hipMemcpyAsync(&m, dcount, sizeof m, hipMemcpyDeviceToHost, s));
    LSDTT_HIP(hipStreamSynchronize(s));
    if (m == 0) return 0.0;
    // compute_median_in_place (src/utils/lsdtt_statistics.cpp:365-376):
    // the upper middle, and for an even count the mean with the lower one
    dbuf<unsigned long long> sel;
    sel.reserve(3);
    const unsigned long long mid = m / 2;
    const double upper = key_value(select_key(keys.p, n, mid, hist.p, sel.p, s));
    if (m % 2) return upper;
    const double lower = key_value(select_key(keys.p, n, mid - 1, hist.p, sel.p, s));
    return 0.5 * (lower + upper);
}

TimestepDiagnosis diagnose_timestep(double reference_slope, double dx, double uplift_rate, double timestep_yr,
                                    double target_ratio)
{
    // diagnose_lem_timestep (lsdtt_lem_timestep_cpu.cpp:54-98)
    TimestepDiagnosis d;
    d.reference_slope = reference_slope;
    if (!(reference_slope > 1e-6) || !(dx > 0.0) || !(uplift_rate > 0.0) || !(timestep_yr > 0.0)) return d;
    d.cell_relief_m = dx * reference_slope;
    d.uplift_per_step_m = uplift_rate * timestep_yr;
    d.ratio = d.uplift_per_step_m / d.cell_relief_m;
    d.suggested_dt_yr = target_ratio * d.cell_relief_m / uplift_rate;
    d.too_large = d.ratio > target_ratio;
    if (d.too_large) {
        std::ostringstream m;
        m.setf(std::ios::fixed);
        m << std::setprecision(2) << "one timestep of uplift (" << d.uplift_per_step_m << " m) is " << d.ratio
          << "x the relief across a single cell (" << d.cell_relief_m << " m at median gradient "
          << d.reference_slope << "); above ~" << target_ratio
          << " the operator split stops resolving incision against diffusion. Suggested timestep: <= "
          << std::setprecision(0) << d.suggested_dt_yr << " yr, or an adaptive timestep.";
        d.warning = m.str();
    }
    return d;
}

TimestepController::TimestepController(double dt_initial, const TimestepOptions& o)
    : o_(o), dt_(dt_initial), dt_raw_(dt_initial), dt_initial_(dt_initial)
{
    dt_min_ = o.dt_min_yr > 0.0 ? o.dt_min_yr : dt_initial / 100.0;
    dt_max_ = o.dt_max_yr > 0.0 ? o.dt_max_yr : dt_initial * 100.0;
}

double TimestepController::decide(double max_dz, double dx)
{
    // lsdtt_lem_timestep_controller::update (lsdtt_lem_timestep_cpu.cpp:
    // 116-206): a proportional step on the smoothed ratio of the achieved
    // change to one cell's relief, shrinking faster than it grows; the
    // continuous value is snapped to dt_initial * ladder^k with one rung of
    // hysteresis upward and an immediate rung down on an overshoot
    const double relief = dx * slope_;
    if (!(relief > 1e-9) || !(max_dz > 0.0)) return dt_;
    last_ratio_ = max_dz / relief;
    ratio_ema_ = ratio_ema_ > 0.0 ? (1.0 - o_.ratio_smoothing) * ratio_ema_ + o_.ratio_smoothing * last_ratio_
                                  : last_ratio_;
    const double rd = ratio_ema_;
    const double before = dt_;
    if (rd > o_.target_ratio) {
        dt_raw_ = std::max(dt_min_, dt_raw_ * std::max(o_.shrink_factor, o_.target_ratio / rd));
        ++n_shrinks_;
    } else if (rd < 0.5 * o_.target_ratio) {
        dt_raw_ = std::min(dt_max_, dt_raw_ * o_.grow_factor);
        ++n_grows_;
    }
    dt_ = dt_raw_;
    if (o_.ladder_ratio > 1.0 && dt_initial_ > 0.0) {
        const double q = o_.ladder_ratio;
        if (rd > o_.target_ratio && rung_ > lowest_rung_ - 1) --rung_;
        while (dt_raw_ >= dt_initial_ * std::pow(q, rung_ + 1)) ++rung_;
        while (dt_raw_ <= dt_initial_ * std::pow(q, rung_ - 1)) --rung_;
        lowest_rung_ = std::min(lowest_rung_, rung_);
        dt_ = std::min(dt_max_, std::max(dt_min_, dt_initial_ * std::pow(q, rung_)));
    }
    if (dt_ != before) ++n_changes_;
    return dt_;
}

ConvergenceMonitor::ConvergenceMonitor(double dt_initial, const ConvergenceOptions& o)
    : o_(o), dt_initial_(dt_initial), window_(static_cast<size_t>(std::max(8, o.stall_window)))
{
    hist_.reserve(window_ + 1);
}

ConvergenceMonitor::State ConvergenceMonitor::update(double max_dz, double dt_used)
{
    // lsdtt_lem_convergence_monitor::update (lsdtt_lem_timestep_cpu.cpp:
    // 227-302): converged after `patience` steps below the tolerance(s)
    // while dt has not collapsed; stalled when the best of the window's last
    // fifth is not below the best of its first fifth by the decay
    // threshold, oscillating when most step-to-step changes alternate
    ++n_steps_;
    last_dz_ = max_dz;
    hist_.push_back(max_dz);
    if (hist_.size() > window_) hist_.erase(hist_.begin());
    const bool dt_ok = dt_initial_ <= 0.0 || dt_used >= o_.min_dt_fraction * dt_initial_;
    const bool below_abs = o_.dz_tolerance_m <= 0.0 || max_dz < o_.dz_tolerance_m;
    const bool below_rate = o_.dz_rate_tolerance_m_per_yr <= 0.0 || dt_used <= 0.0 ||
                            max_dz / dt_used < o_.dz_rate_tolerance_m_per_yr;
    if (below_abs && below_rate && dt_ok) ++n_below_;
    else n_below_ = 0;
    if (n_below_ >= o_.patience_steps) {
        state_ = Converged;
        std::ostringstream m;
        m << "steady state for " << o_.patience_steps << " consecutive steps (|dz|";
        if (o_.dz_tolerance_m > 0.0) m << " < " << o_.dz_tolerance_m << " m/step";
        if (o_.dz_rate_tolerance_m_per_yr > 0.0)
            m << (o_.dz_tolerance_m > 0.0 ? " and" : "") << " < " << o_.dz_rate_tolerance_m_per_yr << " m/yr";
        m << "; last " << last_dz_ << " m/step)";
        message_ = m.str();
        return state_;
    }
    if (hist_.size() < window_) return state_;
    const size_t q = std::max<size_t>(1, window_ / 5);
    const double early = *std::min_element(hist_.begin(), hist_.begin() + static_cast<long>(q));
    const double late = *std::min_element(hist_.end() - static_cast<long>(q), hist_.end());
    decay_ = early > 0.0 ? late / early : 1.0;
    if (decay_ < o_.stall_decay_threshold) return state_;
    long alt = 0;
    for (size_t i = 2; i < hist_.size(); ++i) {
        const double d1 = hist_[i - 1] - hist_[i - 2], d2 = hist_[i] - hist_[i - 1];
        if (d1 * d2 < 0.0) ++alt;
    }
    oscillating_ = static_cast<double>(alt) / static_cast<double>(hist_.size() - 2) > o_.oscillation_alternation_fraction;
    state_ = Stalled;
    message_ = oscillating_ ? "OSCILLATING: max|dz| has not fallen over the last " + std::to_string(window_) +
                                  " steps and alternates step to step (a period-2 cycle: timestep too large, or a "
                                  "limit cycle in the nonlinear solve)"
                            : "STALLED: max|dz| has not fallen over the last " + std::to_string(window_) +
                                  " steps but is still above tolerance";
    return state_;
}

}  // namespace lsdtt_gpu
