// dinf.hip -- D-infinity receivers and weights (see dinf.hpp).
//
// Reference semantics restated (src/cpu/lsdtt_flowgraph_cpu.cpp):
//   facets (Tarboton 1997, Table 1) :62-67: facet a pairs a cardinal
//       neighbour e1 and a diagonal neighbour e2
//   validity :120-139: mask != 0, z finite and not the nodata value
//   resolve_neighbor :364-396: periodic edges wrap, other edges end the grid
//   build_dinf :423-663, per valid cell that is not on a 'b' edge and (with
//       interior_nodata_is_outlet) touches no invalid in-grid cell: over the
//       facets whose two cells exist, differ from the cell and are valid,
//       with e0 above one of them: s1 = (e0 - e1)/dx, s2 = (e1 - e2)/dx in
//       float; both lower: r = atan2(s2, s1), clamped to [0, pi/4] with
//       s = s1 or (e0 - e2)/(dx sqrt2) at the clamps, else
//       s = sqrt(s1^2 + s2^2); only e1 lower: r = 0, s = s1; only e2 lower:
//       r = pi/4, s = (e0 - e2)/(dx sqrt2).  The first facet of largest
//       s > 0 wins; r == 0 sends all flow to e1, r >= pi/4 to e2, otherwise
//       e2 gets r/(pi/4) and e1 the rest (one receiver if they coincide).
//
// Every operation is the reference's float operation (no contraction, IEEE
// division and square root).  atan2 is evaluated in double and rounded to
// float: the correctly rounded value, which the reference's atan2f returns
// or misses by one ulp depending on the C library.
#include <hip/hip_runtime.h>

#include <cmath>

#include "dinf.hpp"

namespace lsdtt_gpu {

namespace {

__constant__ const int kAC[8] = {0, 1, 1, 2, 2, 3, 3, 4};
__constant__ const int kAF[8] = {1, -1, 1, -1, 1, -1, 1, -1};

// resolve_neighbor: periodic wrap, else none (-1)
__device__ __forceinline__ int wrap_nb(int r, int c, int dr, int dc, int nr, int nc, BC bc)
{
    int tr = r + dr, tc = c + dc;
    if (tr < 0) {
        if (bc.n != 'p') return -1;
        tr = nr - 1;
    } else if (tr >= nr) {
        if (bc.s != 'p') return -1;
        tr = 0;
    }
    if (tc < 0) {
        if (bc.w != 'p') return -1;
        tc = nc - 1;
    } else if (tc >= nc) {
        if (bc.e != 'p') return -1;
        tc = 0;
    }
    return tr * nc + tc;
}


// facet a's cardinal (e1) and diagonal (e2) neighbours as slots of the 3x3
// neighbourhood, slot (dr + 1) * 3 + (dc + 1): e1 at (0,1) (-1,0) (-1,0)
// (0,-1) (0,-1) (1,0) (1,0) (0,1), e2 at (-1,1) (-1,1) (-1,-1) (-1,-1)
// (1,-1) (1,-1) (1,1) (1,1)
__constant__ const int kE1S[8] = {5, 1, 1, 3, 3, 7, 7, 5};
__constant__ const int kE2S[8] = {2, 2, 0, 0, 6, 6, 8, 8};

// one cell f = (r, c); INTERIOR: not on the grid's outer rows or columns,
// so no neighbour wraps or leaves the grid and no 'b' edge applies
template <class T, bool INTERIOR>
__device__ __forceinline__ void dinf_cell(const T* __restrict__ z, const uint8_t* __restrict__ mask, int nr, int nc,
                                          float cs, const BC& bc, int nodata_outlet, int use_nd, float nd, int f,
                                          int r, int c, int* __restrict__ rec, float* __restrict__ w,
                                          float* __restrict__ slope, float* __restrict__ bearing)
{
    // the 3x3 neighbourhood once: resolved (wrapped) cells, their float
    // elevations and validity; whether the raw neighbour is in the grid
    // (interior cells: every neighbour is in the grid, none wraps)
    int nb[9];
    float zz[9];
    bool vv[9], ing[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const int dr = k / 3 - 1, dc = k % 3 - 1;
        if (INTERIOR) {
            nb[k] = f + dr * nc + dc;
            ing[k] = true;
        } else {
            nb[k] = (k == 4) ? f : wrap_nb(r, c, dr, dc, nr, nc, bc);
            const int rr = r + dr, cc = c + dc;
            ing[k] = rr >= 0 && rr < nr && cc >= 0 && cc < nc;
        }
        const int g = nb[k] >= 0 ? nb[k] : f;
        const float v = static_cast<float>(z[g]);
        zz[k] = v;
        vv[k] = nb[k] >= 0 && mask[g] && isfinite(v) && !(use_nd && v == nd);
    }
    int r0 = -1, r1 = -1;
    float w0 = 0.0f, w1 = 0.0f, best_s = 0.0f, best_r = 0.0f;
    int best = -1;
    const float e0 = zz[4];
    bool ok = vv[4];
    if (!INTERIOR)
        ok = ok && !((bc.n == 'b' && r == 0) || (bc.s == 'b' && r == nr - 1) || (bc.e == 'b' && c == nc - 1) ||
                     (bc.w == 'b' && c == 0));
    if (ok && nodata_outlet) {
        // touches_interior_nodata: an invalid raw neighbour inside the grid
#pragma unroll
        for (int k = 0; k < 9; ++k)
            if (k != 4 && ing[k] && !vv[k]) ok = false;
    }
    if (ok) {
        const float pi4 = static_cast<float>(M_PI / 4.0);
        const float sqrt2f = static_cast<float>(1.4142135623730950488);
        // the winning facet's slopes, for its angle (one atan2 per cell)
        float bs1 = 0.0f, bs2 = 0.0f;
        bool interior = false;
        for (int a = 0; a < 8; ++a) {
            const int k1 = kE1S[a], k2 = kE2S[a];
            const int g1 = nb[k1], g2 = nb[k2];
            if (g1 < 0 || g2 < 0 || g1 == f || g2 == f) continue;
            if (!vv[k1] || !vv[k2]) continue;
            const float e1 = zz[k1], e2 = zz[k2];
            if (e0 <= e1 && e0 <= e2) continue;
            float rr, sl;
            bool in = false;
            const float s1 = (e0 - e1) / cs;
            const float s2 = (e1 - e2) / cs;
            if (e0 > e1 && e0 > e2) {
                // r = atan2(s2, s1) with s1 > 0 is < 0 exactly when s2 < 0;
                // it rounds above pi4 (the float above pi/4) when s2/s1 - 1
                // is above ~1.03e-7 and not when below; between the two
                // margins the rounded atan2 decides.  The interior
                // slope does not depend on r, so r is evaluated once,
                // for the winning facet.
                const double q = static_cast<double>(s2), p1 = static_cast<double>(s1);
                int cls;  // 0 below 0, 1 inside, 2 above pi/4
                if (s2 < 0.0f) cls = 0;
                else if (q < p1 * (1.0 + 0.9e-7)) cls = 1;
                else if (q > p1 * (1.0 + 1.2e-7)) cls = 2;
                else cls = static_cast<float>(atan2(q, p1)) > pi4 ? 2 : 1;
                if (cls == 0) {
                    rr = 0.0f;
                    sl = s1;
                } else if (cls == 2) {
                    rr = pi4;
                    sl = (e0 - e2) / (cs * sqrt2f);
                } else {
                    rr = 0.0f;  // set from atan2 below if this facet wins
                    in = true;
                    sl = sqrtf(s1 * s1 + s2 * s2);
                }
            } else if (e0 > e1) {
                rr = 0.0f;
                sl = s1;
            } else {
                rr = pi4;
                sl = (e0 - e2) / (cs * sqrt2f);
            }
            if (sl > best_s) {
                best_s = sl;
                best_r = rr;
                best = a;
                r0 = g1;
                r1 = g2;
                interior = in;
                bs1 = s1;
                bs2 = s2;
            }
        }
        if (best >= 0 && interior)
            best_r = static_cast<float>(atan2(static_cast<double>(bs2), static_cast<double>(bs1)));
        if (best >= 0) {
            if (best_r == 0.0f) {
                w0 = 1.0f;
                r1 = -1;
            } else if (best_r >= pi4) {
                r0 = r1;
                w0 = 1.0f;
                r1 = -1;
            } else if (r0 == r1) {
                w0 = 1.0f;
                r1 = -1;
            } else {
                w1 = best_r / pi4;
                w0 = 1.0f - w1;
            }
        } else {
            r0 = r1 = -1;
        }
    }
    rec[2 * (size_t)f] = r0;
    rec[2 * (size_t)f + 1] = r1;
    w[2 * (size_t)f] = w0;
    w[2 * (size_t)f + 1] = w1;
    if (slope) slope[f] = best >= 0 ? best_s : nd;
    if (bearing) {
        float br = nd;
        if (best >= 0) {
            const float rad2deg = static_cast<float>(180.0 / M_PI);
            const float ang =
                static_cast<float>(kAF[best]) * best_r + static_cast<float>(kAC[best]) * (static_cast<float>(M_PI) * 0.5f);
            br = 360.0f - (ang * rad2deg) + 90.0f;
            while (br >= 360.0f) br -= 360.0f;
            while (br < 0.0f) br += 360.0f;
        }
        bearing[f] = br;
    }
}

// interior cells, one per thread (grid-stride); keeping the wrap and edge
// logic out of this kernel keeps its code small
template <class T>
__global__ void __launch_bounds__(256) k_dinf_receivers(const T* __restrict__ z, const uint8_t* __restrict__ mask,
                                                        int nr, int nc, float cs, uint32_t bcw, int nodata_outlet,
                                                        int use_nd, float nd, int* __restrict__ rec,
                                                        float* __restrict__ w, float* __restrict__ slope,
                                                        float* __restrict__ bearing)
{
    const BC bc = bc_unpack(bcw);
    const int ic = nc - 2;
    const long long m = (long long)(nr - 2) * ic;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < m; i += (long long)gridDim.x * blockDim.x) {
        const int r = 1 + static_cast<int>(i / ic), c = 1 + static_cast<int>(i % ic);
        dinf_cell<T, true>(z, mask, nr, nc, cs, bc, nodata_outlet, use_nd, nd, r * nc + c, r, c, rec, w, slope,
                           bearing);
    }
}

// the outer ring: top row, bottom row, then the side columns of the rows
// between (fewer cells on grids of one or two rows or columns)
template <class T>
__global__ void __launch_bounds__(256) k_dinf_receivers_ring(const T* __restrict__ z,
                                                             const uint8_t* __restrict__ mask, int nr, int nc,
                                                             float cs, uint32_t bcw, int nodata_outlet, int use_nd,
                                                             float nd, int* __restrict__ rec, float* __restrict__ w,
                                                             float* __restrict__ slope, float* __restrict__ bearing)
{
    const BC bc = bc_unpack(bcw);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int nbottom = nr > 1 ? nc : 0;
    const int per = nc > 1 ? 2 : 1;
    const int nside = nr > 2 ? per * (nr - 2) : 0;
    int r, c;
    if (i < nc) {
        r = 0, c = i;
    } else if (i < nc + nbottom) {
        r = nr - 1, c = i - nc;
    } else if (i < nc + nbottom + nside) {
        const int k = i - nc - nbottom;
        r = 1 + k / per;
        c = (k % per) ? nc - 1 : 0;
    } else {
        return;
    }
    dinf_cell<T, false>(z, mask, nr, nc, cs, bc, nodata_outlet, use_nd, nd, r * nc + c, r, c, rec, w, slope, bearing);
}

}  // namespace

template <class T>
void dinf_receivers(const T* z, const uint8_t* mask, int nr, int nc, const DinfOptions& o, int* rec, float* w,
                    float* slope, float* bearing, hipStream_t s)
{
    const long long m = nr > 2 && nc > 2 ? (long long)(nr - 2) * (nc - 2) : 0;
    if (m > 0) {
        const int grid = static_cast<int>(std::max(1ll, std::min(8192ll, (long long)div_up(m, 256))));
        k_dinf_receivers<T><<<grid, 256, 0, s>>>(z, mask, nr, nc, static_cast<float>(o.dx), bc_pack(o.bc),
                                                 o.interior_nodata_is_outlet ? 1 : 0, o.use_nodata ? 1 : 0, o.nodata,
                                                 rec, w, slope, bearing);
        LSDTT_LAUNCH_CHECK();
    }
    const int nring = nc + (nr > 1 ? nc : 0) + (nr > 2 ? (nc > 1 ? 2 : 1) * (nr - 2) : 0);
    k_dinf_receivers_ring<T><<<div_up(nring, 256), 256, 0, s>>>(z, mask, nr, nc, static_cast<float>(o.dx),
                                                                bc_pack(o.bc), o.interior_nodata_is_outlet ? 1 : 0,
                                                                o.use_nodata ? 1 : 0, o.nodata, rec, w, slope,
                                                                bearing);
    LSDTT_LAUNCH_CHECK();
}

template void dinf_receivers<float>(const float*, const uint8_t*, int, int, const DinfOptions&, int*, float*, float*,
                                    float*, hipStream_t);
template void dinf_receivers<double>(const double*, const uint8_t*, int, int, const DinfOptions&, int*, float*,
                                     float*, float*, hipStream_t);

}  // namespace lsdtt_gpu
