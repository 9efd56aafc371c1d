// reduce.hpp -- deterministic grid reductions without a second launch.
//
// Every block reduces its values in a fixed order (wave shuffles, then LDS),
// writes one partial per block, and the last block to finish (atomic ticket)
// combines the partials in block order.  Results are therefore bit-identical
// from run to run, unlike floating-point atomics.
#pragma once

#include <hip/hip_runtime.h>

namespace lsdtt_gpu {

enum class RedOp { Sum, Max };

template <RedOp OP>
__device__ __forceinline__ double red_combine(double a, double b)
{
    if constexpr (OP == RedOp::Sum) return a + b;
    else return a > b ? a : b;
}

template <RedOp OP>
__device__ __forceinline__ double wave_reduce(double v)
{
    for (int o = 32; o > 0; o >>= 1) v = red_combine<OP>(v, __shfl_down(v, o));
    return v;
}

// Block reduction of NV values, each with its own op (ops[i] true = max).
// Result valid in thread 0.  blockDim.x must be a multiple of 64, <= 1024.
template <int NV>
__device__ __forceinline__ void block_reduce(double (&v)[NV], const bool (&is_max)[NV])
{
    __shared__ double sm[16][NV];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int i = 0; i < NV; ++i) v[i] = is_max[i] ? wave_reduce<RedOp::Max>(v[i]) : wave_reduce<RedOp::Sum>(v[i]);
    if (lane == 0)
#pragma unroll
        for (int i = 0; i < NV; ++i) sm[w][i] = v[i];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 0; i < NV; ++i) {
            double a = sm[0][i];
            for (int k = 1; k < nw; ++k) a = is_max[i] ? (a > sm[k][i] ? a : sm[k][i]) : a + sm[k][i];
            v[i] = a;
        }
    }
    __syncthreads();
}

// Block reduction whose result is stored as partial slot `slot` (for a
// companion kernel whose partials a later grid_reduce folds in).
template <int NV>
__device__ __forceinline__ void block_partial(double (&v)[NV], const bool (&is_max)[NV], double* partials, int slot)
{
    block_reduce<NV>(v, is_max);
    if (threadIdx.x == 0)
        for (int i = 0; i < NV; ++i) partials[(size_t)slot * NV + i] = v[i];
}

// Writes the block's partials, and in the last block combines all partials
// (in block order, then `extra` companion slots stored after them) into
// out[0..NV).  `ticket` must be zero on entry and is left zero.  Returns
// true in thread 0 of the last block.
template <int NV>
__device__ __forceinline__ bool grid_reduce(double (&v)[NV], const bool (&is_max)[NV], double* partials,
                                            unsigned* ticket, double* out, unsigned extra = 0)
{
    block_reduce<NV>(v, is_max);
    __shared__ bool last;
    if (threadIdx.x == 0) {
        for (int i = 0; i < NV; ++i) partials[(size_t)blockIdx.x * NV + i] = v[i];
        __threadfence();
        const unsigned t = atomicAdd(ticket, 1u);
        last = (t == gridDim.x - 1);
    }
    __syncthreads();
    if (!last) return false;
    __threadfence();
    // every thread folds a fixed stride of the partials, then the block
    // reduces in its fixed order: deterministic and parallel
    double w[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) w[i] = is_max[i] ? -1.7976931348623157e308 : 0.0;
    for (unsigned b = threadIdx.x; b < gridDim.x + extra; b += blockDim.x)
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            const double x = __hip_atomic_load(&partials[(size_t)b * NV + i], __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT);
            w[i] = is_max[i] ? (w[i] > x ? w[i] : x) : w[i] + x;
        }
    block_reduce<NV>(w, is_max);
    if (threadIdx.x == 0) {
        for (int i = 0; i < NV; ++i) out[i] = w[i];
        *ticket = 0u;
        return true;
    }
    return false;
}

// Folds `count` partial slots in slot order (thread-strided, then the block
// reduction's fixed order) into w[0..NV), valid in thread 0.  For a
// one-block finisher kernel launched after the kernel that wrote the
// partials: the kernel boundary makes them visible, which costs far less
// than grid_reduce's per-block device-scope fence (on a multi-XCD GPU each
// such fence writes back the XCD's L2: ~20 ns per block, 44 us for a
// 2048-block grid).
template <int NV>
__device__ __forceinline__ void fold_partials(const double* partials, unsigned count, const bool (&is_max)[NV],
                                              double (&w)[NV])
{
#pragma unroll
    for (int i = 0; i < NV; ++i) w[i] = is_max[i] ? -1.7976931348623157e308 : 0.0;
    for (unsigned b = threadIdx.x; b < count; b += blockDim.x)
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            const double x = partials[(size_t)b * NV + i];
            w[i] = is_max[i] ? (w[i] > x ? w[i] : x) : w[i] + x;
        }
    block_reduce<NV>(w, is_max);
}

}  // namespace lsdtt_gpu
