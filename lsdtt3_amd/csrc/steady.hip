// steady.hip -- analytic fluvial steady state on the D8 flow object (see
// steady.hpp).
//
// Reference semantics (src/cpu/lsdtt_steady_state_cpu.cpp):
//   make_chi_weights :37-72      w_i = pow(K / K_i, 1/n) in float
//   compute_fluvial_steady_state_on_tree :76-170
//     A = accumulate() (dx^2 per valid cell in the flow object's order),
//     divided by dx when divide_area_by_dx, dx^2 if not positive;
//     amn = powf(float(A), -m/n) * w; chi = chi(receiver) + dx_eff * amn in
//     float, outlets 0; MaxChi = max over valid cells (1 if not positive);
//     f = (U/K)^(1/n) or target_relief / MaxChi (relief mode, K_used =
//     U (MaxChi / target_relief)^n); z = baselevel + f * chi
//   compute_d8_fluvial_steady_state :173-225  dx_eff = dx or dx * sqrt2f
//     (float) by the D8 code's parity
//   compute_dinf_fluvial_steady_state :227-436  the same amn on the
//     D-infinity graph's accumulate(); chi from one or two receivers
//     (FluvialTiles::dinf_chi)
// powf is evaluated as the correctly rounded value (pow in double, rounded
// to float); the reference's powf returns it or one ulp beside it,
// depending on the C library.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "steady.hpp"

namespace lsdtt_gpu {

namespace {

constexpr int kB = 256;

__global__ void k_area_weights(const uint8_t* __restrict__ mask, long long n, double w0, double* __restrict__ w,
                               double* __restrict__ acc)
{
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i < n) {
        const double v = mask[i] ? w0 : 0.0;
        w[i] = v;
        acc[i] = v;
    }
}

// amn per valid cell; chi 0 everywhere to start
__global__ void k_amn(const uint8_t* __restrict__ mask, const double* __restrict__ area, const float* __restrict__ w,
                      long long n, float dx, int divide, float expo, float* __restrict__ amn, float* __restrict__ chi)
{
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    chi[i] = 0.0f;
    if (!mask[i]) {
        amn[i] = 0.0f;
        return;
    }
    double A = area[i];
    if (divide) A /= static_cast<double>(dx);
    if (A <= 0.0) A = static_cast<double>(dx) * dx;
    const float fa = static_cast<float>(A);
    const float p = static_cast<float>(pow(static_cast<double>(fa), static_cast<double>(expo)));
    amn[i] = p * (w ? w[i] : 1.0f);
}

struct ChiSweep {
    const int* rcv;
    const uint8_t* dir;
    const float* amn;
    float dx, dxd;
    float* chi;
};

__device__ __forceinline__ void sweep_node(const ChiSweep& a, int v)
{
    const int r = a.rcv[v];
    if (r < 0) return;  // outlet: chi 0
    const float dxe = (a.dir[v] & 1u) ? a.dxd : a.dx;
    a.chi[v] = a.chi[r] + dxe * a.amn[v];
}

__global__ void k_sweep_level(const int* __restrict__ order, int begin, int end, ChiSweep a)
{
    const int i = begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (i < end) sweep_node(a, order[i]);
}

__global__ void __launch_bounds__(1024) k_sweep_tail(const int* __restrict__ order, const int* __restrict__ lstart,
                                                     int lo, int hi, ChiSweep a)
{
    for (int L = lo; L <= hi; ++L) {
        const int b = lstart[L], e = lstart[L + 1];
        for (int i = b + threadIdx.x; i < e; i += blockDim.x) sweep_node(a, order[i]);
        __syncthreads();
    }
}

// max chi over valid cells and over channel cells (chi >= 0: the float bit
// patterns order as integers), channel count
__global__ void __launch_bounds__(kB) k_chi_max(const uint8_t* __restrict__ mask, const float* __restrict__ chi,
                                                const double* __restrict__ area, long long n, double dx, int divide,
                                                double thr, unsigned* __restrict__ out, unsigned long long* nchan)
{
    unsigned mx = 0u, mc = 0u, cnt = 0u;
    for (long long i = blockIdx.x * (long long)kB + threadIdx.x; i < n; i += (long long)gridDim.x * kB) {
        if (!mask[i]) continue;
        const unsigned b = __float_as_uint(chi[i]);
        mx = max(mx, b);
        if (thr > 0.0) {
            const double q = divide ? area[i] / (dx > 1e-12 ? dx : 1e-12) : area[i];
            if (!(q < thr)) {
                mc = max(mc, b);
                ++cnt;
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        mx = max(mx, __shfl_xor(mx, o));
        mc = max(mc, __shfl_xor(mc, o));
        cnt += __shfl_xor(cnt, o);
    }
    if ((threadIdx.x & 63) == 0) {
        if (mx) atomicMax(&out[0], mx);
        if (mc) atomicMax(&out[1], mc);
        if (cnt) atomicAdd(nchan, static_cast<unsigned long long>(cnt));
    }
}

__global__ void k_steady_z(const uint8_t* __restrict__ mask, const float* __restrict__ chi, long long n, float base,
                           float f, float nodata, float* __restrict__ z)
{
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i < n) z[i] = mask[i] ? base + f * chi[i] : nodata;
}

// make_chi_weights (:37-72): w = powf(K / K_i, 1/n) on valid cells, 1 off
__global__ void k_chi_weights(const uint8_t* __restrict__ mask, const float* __restrict__ K, long long n, float Kref,
                              float inv_n, float* __restrict__ w)
{
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i < n)
        w[i] = mask[i] ? static_cast<float>(pow(static_cast<double>(Kref / K[i]), static_cast<double>(inv_n))) : 1.0f;
}

// steps 2-4 of both paths (:106-170, :358-436): max chi over valid cells
// (1 if not positive), f and K_used with the host's float pow, elevations,
// chi_c over the channel pixels
SteadyResult finish(const uint8_t* mask, const float* chi, const double* area, long long n, float dx,
                    const SteadyOptions& o, const float* w, float nodata, float* z_out, hipStream_t s)
{
    const int g = div_up(n, kB);
    dbuf<unsigned> mx;
    dbuf<unsigned long long> nch;
    mx.reserve(2);
    nch.reserve(1);
    LSDTT_HIP(hipMemsetAsync(mx.p, 0, sizeof(unsigned) * 2, s));
    LSDTT_HIP(hipMemsetAsync(nch.p, 0, sizeof(unsigned long long), s));
    k_chi_max<<<std::min(g, 2048), kB, 0, s>>>(mask, chi, area, n, static_cast<double>(dx),
                                               o.divide_area_by_dx ? 1 : 0,
                                               static_cast<double>(o.channel_threshold_area), mx.p, nch.p);
    LSDTT_LAUNCH_CHECK();
    unsigned hm[2];
    unsigned long long hn = 0;
    LSDTT_HIP(hipMemcpyAsync(hm, mx.p, sizeof hm, hipMemcpyDeviceToHost, s));
    LSDTT_HIP(hipMemcpyAsync(&hn, nch.p, sizeof hn, hipMemcpyDeviceToHost, s));
    LSDTT_HIP(hipStreamSynchronize(s));
    float max_chi, chi_c;
    std::memcpy(&max_chi, &hm[0], sizeof max_chi);
    std::memcpy(&chi_c, &hm[1], sizeof chi_c);
    if (max_chi <= 0.0f) max_chi = 1.0f;
    // step 3 (:112-123): scalars, with the host's float pow as the reference
    SteadyResult r;
    float f;
    if (o.relief_mode) {
        r.K_used = o.U * std::pow(max_chi / o.target_relief, o.n);
        f = o.target_relief / max_chi;
    } else {
        r.K_used = o.K;
        f = std::pow(o.U / o.K, 1.0f / o.n);
    }
    k_steady_z<<<g, kB, 0, s>>>(mask, chi, n, o.baselevel_elevation, f, nodata, z_out);
    LSDTT_LAUNCH_CHECK();
    LSDTT_HIP(hipStreamSynchronize(s));
    r.max_unit_height = max_chi;
    r.computed_relief = f * max_chi;
    if (w) r.erodibility_field_scale = r.K_used / o.K;
    if (o.channel_threshold_area > 0.0f) {
        r.n_channel_pixels = static_cast<long long>(hn);
        if (hn > 0) {
            r.max_unit_height_at_channel_head = chi_c;
            r.chi_c_reported = true;
        }
    }
    return r;
}

}  // namespace

void chi_weights(const uint8_t* mask, const float* K, long long n, float Kref, float nexp, float* w, hipStream_t s)
{
    k_chi_weights<<<div_up(n, kB), kB, 0, s>>>(mask, K, n, Kref, 1.0f / nexp, w);
    LSDTT_LAUNCH_CHECK();
}

SteadyResult steady_state_d8(FlowEngine& e, const SteadyOptions& o, const float* w, float nodata, float* z_out,
                             hipStream_t s)
{
    const long long n = e.n;
    const float dx = static_cast<float>(e.opt.dx);
    const float dxd = dx * static_cast<float>(kSqrt2);
    dbuf<double> wbuf, area;
    dbuf<float> amn, chi;
    wbuf.reserve(n);
    area.reserve(n);
    amn.reserve(n);
    chi.reserve(n);
    const int g = div_up(n, kB);
    k_area_weights<<<g, kB, 0, s>>>(e.mask, n, e.opt.dx * e.opt.dx, wbuf.p, area.p);
    LSDTT_LAUNCH_CHECK();
    e.upstream(true, wbuf.p, area.p, s);
    const float expo = -o.m / o.n;
    k_amn<<<g, kB, 0, s>>>(e.mask, area.p, w, n, dx, o.divide_area_by_dx ? 1 : 0, expo, amn.p, chi.p);
    LSDTT_LAUNCH_CHECK();
    ChiSweep a{e.rcv.p, e.dir.p, amn.p, dx, dxd, chi.p};
    const int nl = e.max_depth + 1;
    const int split = std::max(e.tail_from, 1);
    for (int L = 1; L < split && L < nl; ++L) {
        const int b = e.level_start_h[L], en = e.level_start_h[L + 1];
        if (en > b) {
            k_sweep_level<<<div_up(en - b, kB), kB, 0, s>>>(e.order.p, b, en, a);
            LSDTT_LAUNCH_CHECK();
        }
    }
    if (split < nl) {
        k_sweep_tail<<<1, 1024, 0, s>>>(e.order.p, e.level_start.p, split, nl - 1, a);
        LSDTT_LAUNCH_CHECK();
    }
    return finish(e.mask, chi.p, area.p, n, dx, o, w, nodata, z_out, s);
}

SteadyResult steady_state_dinf(FluvialTiles& t, const int* rec, const float* wr, const uint8_t* valid,
                               const double* area, const SteadyOptions& o, const float* w, float dx, float nodata,
                               float* z_out, hipStream_t s)
{
    const long long n = (long long)t.nr * t.nc;
    dbuf<float> amn, chi;
    amn.reserve(n);
    chi.reserve(n);
    const int g = div_up(n, kB);
    k_amn<<<g, kB, 0, s>>>(valid, area, w, n, dx, o.divide_area_by_dx ? 1 : 0, -o.m / o.n, amn.p, chi.p);
    LSDTT_LAUNCH_CHECK();
    t.dinf_up_order(rec, valid, s);
    t.dinf_chi(rec, wr, valid, amn.p, dx, chi.p, s);
    return finish(valid, chi.p, area, n, dx, o, w, nodata, z_out, s);
}

}  // namespace lsdtt_gpu
