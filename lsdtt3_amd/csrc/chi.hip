// chi.hip -- chi on the device flow object (see chi.hpp).
#include <hip/hip_runtime.h>

#include <cmath>
#include <rocprim/device/device_scan.hpp>

#include "chi.hpp"
#include "hostpool.hpp"

namespace lsdtt_gpu {

namespace {

constexpr int kB = 256;

__global__ void k_zero_u32(unsigned* __restrict__ a, long long n)
{
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0u;
}

// bit k of `bits` set when some valid cell has upstream count k.  Counts
// below kLdsBits are gathered in a workgroup bitmap first (most cells have
// small counts: one global atomic per word per workgroup instead of one per
// cell on the same few words); larger counts test the word before the
// atomic.
constexpr int kLdsBits = 8192;
__global__ void __launch_bounds__(kB) k_count_presence(const uint8_t* __restrict__ mask, const int* __restrict__ cnt,
                                                       long long n, unsigned* __restrict__ bits)
{
    __shared__ unsigned lb[kLdsBits / 32];
    for (int i = threadIdx.x; i < kLdsBits / 32; i += kB) lb[i] = 0u;
    __syncthreads();
    for (long long f = blockIdx.x * (long long)kB + threadIdx.x; f < n; f += (long long)gridDim.x * kB) {
        const uint8_t m = mask[f];
        const int k = cnt[f];
        if (!m) continue;
        const unsigned bit = 1u << (k & 31);
        if (k < kLdsBits) {
            if (!(lb[k >> 5] & bit)) atomicOr(&lb[k >> 5], bit);
        } else if (!(__hip_atomic_load(&bits[k >> 5], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & bit)) {
            atomicOr(&bits[k >> 5], bit);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kLdsBits / 32; i += kB)
        if (lb[i]) atomicOr(&bits[i], lb[i]);
}

__global__ void k_word_popc(const unsigned* __restrict__ bits, long long nw, int* __restrict__ pc)
{
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i < nw) pc[i] = __popc(bits[i]);
}

__global__ void k_chi_init(const uint8_t* __restrict__ mask, const int* __restrict__ rcv, long long n, float nodata,
                           float* __restrict__ chi)
{
    const long long f = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (f < n) chi[f] = mask[f] ? (rcv[f] < 0 ? 0.0f : __int_as_float(0x7fc00000)) : nodata;
}

struct ChiArgs {
    const int* rcv;
    const uint8_t* dir;
    const int* cnt;
    const double* vals;      // f(k) of the counts present, in increasing k
    const unsigned* bits;    // presence bitmap of the counts
    const int* wrank;        // set bits in the words before each word
    double dx;
    float* chi;
};

// f(k): the rank of k among the counts present indexes the compact table
__device__ __forceinline__ double f_of(const ChiArgs& a, int k)
{
    const int w = k >> 5;
    return a.vals[a.wrank[w] + __popc(a.bits[w] & ((1u << (k & 31)) - 1u))];
}

// lsdtt_chi_extraction_cpu.cpp:355-361: (step * 0.5) * (f_v + f_r) added to
// the receiver's float chi in double, rounded to float
__device__ __forceinline__ void chi_node(const ChiArgs& a, int v)
{
    const int r = a.rcv[v];
    const double step = (a.dir[v] & 1u) ? a.dx * kSqrt2 : a.dx;
    const double inc = step * 0.5 * (f_of(a, a.cnt[v]) + f_of(a, a.cnt[r]));
    a.chi[v] = static_cast<float>(static_cast<double>(a.chi[r]) + inc);
}

__global__ void k_chi_level(const int* __restrict__ order, int begin, int end, ChiArgs a)
{
    const int i = begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (i < end) chi_node(a, order[i]);
}

// one workgroup walks levels lo..hi: a receiver is one level shallower, so
// it is final when its donors read it
__global__ void __launch_bounds__(1024) k_chi_tail(const int* __restrict__ order, const int* __restrict__ lstart,
                                                   int lo, int hi, ChiArgs a)
{
    for (int L = lo; L <= hi; ++L) {
        const int b = lstart[L], e = lstart[L + 1];
        for (int i = b + threadIdx.x; i < e; i += blockDim.x) chi_node(a, order[i]);
        __syncthreads();
    }
}

__global__ void k_select_flag(const int* __restrict__ svec, const int* __restrict__ cnt, int m, int min_pixels,
                              int* __restrict__ flag)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < m) flag[p] = cnt[svec[p]] >= min_pixels ? 1 : 0;
}

__global__ void k_select_scatter(const int* __restrict__ svec, const int* __restrict__ flag,
                                 const int* __restrict__ pos, int m, const int* __restrict__ rcv,
                                 const float* __restrict__ chi, int* __restrict__ out_flat,
                                 int* __restrict__ out_rcv, float* __restrict__ out_chi)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < m && flag[p]) {
        const int v = svec[p], q = pos[p];
        out_flat[q] = v;
        out_rcv[q] = rcv[v];
        out_chi[q] = chi[v];
    }
}

// the counts present in increasing order: word w's set bits land at
// wrank[w] on; the last word also writes the total
__global__ void k_present_list(const unsigned* __restrict__ bits, const int* __restrict__ wrank, long long nw,
                               int* __restrict__ out, int* __restrict__ total)
{
    const long long w = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (w >= nw) return;
    unsigned b = bits[w];
    int o = wrank[w];
    if (w == nw - 1) *total = o + __popc(b);
    for (; b; b &= b - 1) out[o++] = static_cast<int>(w * 32 + __builtin_ctz(b));
}

}  // namespace

// room for the list of counts present: at most one per valid cell
long long ChiEngine::n_present_cap(long long n) { return n; }

void ChiEngine::raster(FlowEngine& e, double m_over_n, double A0, float nodata, float* chi, hipStream_t s)
{
    const long long n = e.n;
    const int g = div_up(n, kB);
    k_chi_init<<<g, kB, 0, s>>>(e.mask, e.rcv, n, nodata, chi);
    LSDTT_LAUNCH_CHECK();
    if (e.n_valid == 0) return;
    // counts present in the grid
    const long long nw = (n >> 5) + 1;
    bits.reserve(nw);
    k_zero_u32<<<div_up(nw, kB), kB, 0, s>>>(bits.p, nw);
    LSDTT_LAUNCH_CHECK();
    k_count_presence<<<std::min(g, 2048), kB, 0, s>>>(e.mask, e.cnt, n, bits.p);
    LSDTT_LAUNCH_CHECK();
    // rank table: set bits before each word
    wpc.reserve(nw + 1);
    wrank.reserve(nw + 1);
    present.reserve(n_present_cap(n) + 1);
    k_word_popc<<<div_up(nw, kB), kB, 0, s>>>(bits.p, nw, wpc.p);
    LSDTT_LAUNCH_CHECK();
    size_t tb = 0;
    LSDTT_HIP(rocprim::exclusive_scan(nullptr, tb, wpc.p, wrank.p, 0, (size_t)nw, rocprim::plus<int>(), s));
    LSDTT_HIP(rocprim::exclusive_scan(tmp.need(tb), tb, wpc.p, wrank.p, 0, (size_t)nw, rocprim::plus<int>(), s));
    // the counts present, in increasing order, then their number
    k_present_list<<<div_up(nw, kB), kB, 0, s>>>(bits.p, wrank.p, nw, present.p, present.p + n_present_cap(n));
    LSDTT_LAUNCH_CHECK();
    hread.reserve(1);
    LSDTT_HIP(hipMemcpyAsync(hread.p, present.p + n_present_cap(n), sizeof(int), hipMemcpyDeviceToHost, s));
    LSDTT_HIP(hipStreamSynchronize(s));
    const long long K = hread.p[0];
    kh.reserve(std::max<long long>(K, 1));
    LSDTT_HIP(hipMemcpyAsync(kh.p, present.p, sizeof(int) * K, hipMemcpyDeviceToHost, s));
    LSDTT_HIP(hipStreamSynchronize(s));
    // f(k) = pow(A0 / (k * dx^2), theta), the host libm's values
    // (lsdtt_chi_extraction_cpu.cpp:154-163 with A from :748-751), for the
    // counts present only, in increasing k
    const double pa = e.opt.dx * e.opt.dx;
    table_h.reserve(std::max<long long>(K, 1));
    const int* ks = kh.p;
    double* tv = table_h.p;
    host_parallel_for(K, 4096, [&](long long b, long long en) {
        for (long long i = b; i < en; ++i) {
            const double denom = static_cast<double>(ks[i]) * pa;
            tv[i] = std::pow(A0 / denom, m_over_n);
        }
    });
    table.reserve(std::max<long long>(K, 1));
    LSDTT_HIP(hipMemcpyAsync(table.p, table_h.p, sizeof(double) * K, hipMemcpyHostToDevice, s));
    ChiArgs a{e.rcv.p, e.dir.p, e.cnt.p, table.p, bits.p, wrank.p, e.opt.dx, chi};
    const int nl = e.max_depth + 1;
    const int split = std::max(e.tail_from, 1);
    for (int L = 1; L < split && L < nl; ++L) {
        const int b = e.level_start_h[L], en = e.level_start_h[L + 1];
        if (en > b) {
            k_chi_level<<<div_up(en - b, kB), kB, 0, s>>>(e.order.p, b, en, a);
            LSDTT_LAUNCH_CHECK();
        }
    }
    if (split < nl) {
        k_chi_tail<<<1, 1024, 0, s>>>(e.order.p, e.level_start.p, split, nl - 1, a);
        LSDTT_LAUNCH_CHECK();
    }
    // the pinned host table is the source of an asynchronous copy: it must
    // not change before the copy is made
    LSDTT_HIP(hipStreamSynchronize(s));
}

long long ChiEngine::threshold(FlowEngine& e, const float* chi, long long min_pixels, hipStream_t s)
{
    const int m = e.n_valid;
    if (m == 0) return 0;
    // counts are int32 on the device: a threshold above every count selects
    // nothing
    const int mp = static_cast<int>(std::min<long long>(std::max<long long>(min_pixels, 1), 0x7fffffffLL));
    flag.reserve(m + 1);
    pos.reserve(m + 1);
    sel_flat.reserve(m);
    sel_rcv.reserve(m);
    sel_chi.reserve(m);
    const int g = div_up(m, kB);
    k_select_flag<<<g, kB, 0, s>>>(e.svec.p, e.cnt.p, m, mp, flag.p);
    LSDTT_LAUNCH_CHECK();
    LSDTT_HIP(hipMemsetAsync(flag.p + m, 0, sizeof(int), s));
    size_t tb = 0;
    LSDTT_HIP(rocprim::exclusive_scan(nullptr, tb, flag.p, pos.p, 0, (size_t)(m + 1), rocprim::plus<int>(), s));
    LSDTT_HIP(
        rocprim::exclusive_scan(tmp.need(tb), tb, flag.p, pos.p, 0, (size_t)(m + 1), rocprim::plus<int>(), s));
    k_select_scatter<<<g, kB, 0, s>>>(e.svec.p, flag.p, pos.p, m, e.rcv.p, chi, sel_flat.p, sel_rcv.p, sel_chi.p);
    LSDTT_LAUNCH_CHECK();
    int total = 0;
    LSDTT_HIP(hipMemcpyAsync(&total, pos.p + m, sizeof(int), hipMemcpyDeviceToHost, s));
    LSDTT_HIP(hipStreamSynchronize(s));
    return total;
}

}  // namespace lsdtt_gpu
