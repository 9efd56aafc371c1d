// common.hpp -- shared device helpers for the lsdtt3 MI355X LEM library.
//
// Grid conventions are the reference's (src/core/lsdtt_types.hpp:142-144,
// src/core/lsdtt_flow_base.hpp:265-271, src/utils/lsdtt_grid_neighbours.hpp):
// row-major flat index row*n_cols+col; D8 direction codes 0..7 clockwise from
// N (N, NE, E, SE, S, SW, W, NW); receiver -1 = nodata, -2 = outlet;
// direction 255 = none; boundary string N,E,S,W of o/n/p/b.
//
// On the device flat indices are int32 (every entry point rejects grids of
// 2^31 cells or more); values are widened to int64 at the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

namespace lsdtt_gpu {

constexpr int kNodataReceiver = -1;
constexpr int kOutletReceiver = -2;
constexpr uint8_t kNoDirection = 255;
constexpr double kSqrt2 = 1.4142135623730950488;

__constant__ static const int d8_row[8] = {-1, -1, 0, 1, 1, 1, 0, -1};
__constant__ static const int d8_col[8] = {0, 1, 1, 1, 0, -1, -1, -1};

struct BC {
    char n, e, s, w;
};

// BC as one 32-bit kernel argument: a struct of chars passed by value can
// be read from the argument segment with a vector load and a full wait;
// an integer argument stays in scalar registers
inline uint32_t bc_pack(BC b)
{
    return (uint32_t)(uint8_t)b.n | ((uint32_t)(uint8_t)b.e << 8) | ((uint32_t)(uint8_t)b.s << 16) |
           ((uint32_t)(uint8_t)b.w << 24);
}
__host__ __device__ __forceinline__ BC bc_unpack(uint32_t w)
{
    return BC{static_cast<char>(w & 0xff), static_cast<char>((w >> 8) & 0xff), static_cast<char>((w >> 16) & 0xff),
              static_cast<char>(w >> 24)};
}

// Result of stepping off a cell under the routing boundary rules
// (lsdtt_flowinfo_cpu.cpp:101-158): kind 0 = blocked (no-flux edge),
// 1 = lands inside the grid at (r, c), 2 = leaves through an 'o'/'b' edge.
struct Step {
    int kind, r, c;
};

__device__ __forceinline__ Step resolve_step(int row, int col, int dir, int nr, int nc, BC bc)
{
    int r = row + d8_row[dir], c = col + d8_col[dir];
    bool off = false;
    if (r < 0) {
        if (bc.n == 'p') r = nr - 1;
        else if (bc.n == 'o' || bc.n == 'b') off = true;
        else return {0, 0, 0};
    } else if (r >= nr) {
        if (bc.s == 'p') r = 0;
        else if (bc.s == 'o' || bc.s == 'b') off = true;
        else return {0, 0, 0};
    }
    if (c < 0) {
        if (bc.w == 'p') c = nc - 1;
        else if (bc.w == 'o' || bc.w == 'b') off = true;
        else return {0, 0, 0};
    } else if (c >= nc) {
        if (bc.e == 'p') c = 0;
        else if (bc.e == 'o' || bc.e == 'b') off = true;
        else return {0, 0, 0};
    }
    if (off) return {2, 0, 0};
    return {1, r, c};
}

// Distinct in-grid neighbours of `flat` (periodic wraps on 1- or 2-wide axes
// can reach one cell by two directions, and a 1-wide periodic axis can reach
// the cell itself, which is excluded).  Returns the count; `out` holds the
// neighbour flats in direction order.
__device__ __forceinline__ int distinct_neighbours(int flat, int nr, int nc, BC bc, int out[8])
{
    const int row = flat / nc, col = flat - (flat / nc) * nc;
    int k = 0;
    if (row > 0 && row < nr - 1 && col > 0 && col < nc - 1) {
#pragma unroll
        for (int d = 0; d < 8; ++d) out[k++] = flat + d8_row[d] * nc + d8_col[d];
        return k;
    }
    for (int d = 0; d < 8; ++d) {
        Step s = resolve_step(row, col, d, nr, nc, bc);
        if (s.kind != 1) continue;
        const int g = s.r * nc + s.c;
        if (g == flat) continue;
        bool dup = false;
        for (int i = 0; i < k; ++i) dup |= (out[i] == g);
        if (!dup) out[k++] = g;
    }
    return k;
}

inline int div_up(long long a, long long b) { return static_cast<int>((a + b - 1) / b); }

// a / b, correctly rounded, from yb = 1 / b (itself correctly rounded): one
// product and two FMAs instead of a division.  Markstein's correction: q0 =
// a*yb is within an ulp of a/b and r = a - b*q0 is exact, so q0 + r*yb
// rounds to the correctly rounded quotient (round to nearest, no underflow
// or overflow on the way; |a| outside [2^-900, 2^900] takes the division).
__device__ __forceinline__ double div_by_rcp(double a, double b, double yb)
{
    const double m = fabs(a);
    if (!(m >= 0x1p-900 && m <= 0x1p+900)) return a / b;
    const double q0 = a * yb;
    const double r = __builtin_fma(-q0, b, a);
    return __builtin_fma(r, yb, q0);
}

// XCD-aware tile numbering for a 1-D launch of `tiles` workgroups.  The
// hardware deals consecutive workgroups to its 8 XCDs in turn, and each XCD
// has its own L2; renumbered so, each XCD walks one contiguous band of tiles
// in order and its L2 holds the halo rows and columns that neighbouring
// tiles share (a bijection on [0, tiles)).
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_tile(int b, int tiles)
{
    const int per = tiles / kXcds, extra = tiles % kXcds;  // the first `extra` XCDs take one more
    const int x = b % kXcds, i = b / kXcds;
    return x < extra ? x * (per + 1) + i : extra * (per + 1) + (x - extra) * per + i;
}

struct hip_error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define LSDTT_HIP(expr)                                                                  \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            throw ::lsdtt_gpu::hip_error(std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Launch-configuration check: every kernel launch is followed by this.
#define LSDTT_LAUNCH_CHECK() LSDTT_HIP(hipGetLastError())

}  // namespace lsdtt_gpu
