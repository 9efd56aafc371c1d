/*
 * orc_fill.c -- epsilon-free depression filling restated.  TEST
 * INFRASTRUCTURE ONLY.
 *
 * Reference: src/cpu/lsdtt_raster_flow_preprocessing_cpu.cpp
 *   make_fill_inputs       :350-392   validity = mask != 0 && !is_nodata(z)
 *   seed_raster_boundary   :518-589   exits: non-periodic raster border cells,
 *                                     plus (only when some interior cell is
 *                                     invalid) cells 8-adjacent to an invalid
 *                                     cell; n/o/b codes are ignored
 *   priority_flood_core    :416-511   at min_slope <= 0 a lower neighbour is
 *                                     raised to exactly the spill level
 *   fill_pits (bc overload):727-782   periodic pairs must match; both pairs
 *                                     periodic is an error
 *   fill_pits_reconstruction :1203-1394 is pinned equal to the priority flood
 *                                     at min_slope <= 0 by the reference's own
 *                                     property P9 (tests/cpp/test_flow_properties.cpp)
 * src/core/lsdtt_raster_meta.hpp:124-131 (is_nodata, NaN aware).
 *
 * Parity: pinned by brute force (minimax path value over all paths to an
 * exit, computed by Bellman-Ford style relaxation in tests) and by the
 * reference's properties P1 (depression-free), P2 (never lowers), P6
 * (idempotent).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>

static const int D8R[8] = {-1, -1, 0, 1, 1, 1, 0, -1};
static const int D8C[8] = {0, 1, 1, 1, 0, -1, -1, -1};

typedef struct { float z; int64_t k; } node;

static void heap_push(node* h, int64_t* n, node v)
{
    int64_t i = (*n)++;
    while (i > 0) {
        int64_t p = (i - 1) / 2;
        if (h[p].z <= v.z) break;
        h[i] = h[p];
        i = p;
    }
    h[i] = v;
}

static node heap_pop(node* h, int64_t* n)
{
    node top = h[0];
    node last = h[--(*n)];
    int64_t i = 0;
    for (;;) {
        int64_t l = 2 * i + 1, r = l + 1, m = i;
        float mz = last.z;
        if (l < *n && h[l].z < mz) { m = l; mz = h[l].z; }
        if (r < *n && h[r].z < mz) { m = r; }
        if (m == i) break;
        h[i] = h[m];
        i = m;
    }
    if (*n > 0) h[i] = last;
    return top;
}

int orc_fill_priority_flood(int nr, int nc, const float* z,
                            const uint8_t* mask, double nodata,
                            const char* bc, float* filled)
{
    const int per_ns = bc[0] == 'p', per_ew = bc[1] == 'p';
    if (per_ns != (bc[2] == 'p') || per_ew != (bc[3] == 'p')) return ORC_ERR_BC;
    if (per_ns && per_ew) return ORC_ERR_BC;
    const int64_t n = (int64_t)nr * nc;
    const float ndv = (float)nodata;
    uint8_t* valid = (uint8_t*)malloc((size_t)(n ? n : 1));
    uint8_t* state = (uint8_t*)calloc((size_t)(n ? n : 1), 1); /* 0 open 1 queued 2 closed */
    node* heap = (node*)malloc(sizeof(node) * (size_t)(n ? n : 1) * 2);
    int64_t hn = 0;
    for (int64_t k = 0; k < n; ++k) {
        const double v = (double)z[k];
        const int nd = isnan(nodata) ? isnan(v) : (v == nodata);
        valid[k] = (mask[k] != 0 && !nd);
        filled[k] = valid[k] ? z[k] : ndv;
    }
    int interior_invalid = 0;
    for (int r = 1; r < nr - 1 && !interior_invalid; ++r)
        for (int c = 1; c < nc - 1; ++c)
            if (!valid[(int64_t)r * nc + c]) { interior_invalid = 1; break; }
    for (int r = 0; r < nr; ++r)
        for (int c = 0; c < nc; ++c) {
            const int64_t k = (int64_t)r * nc + c;
            if (!valid[k]) continue;
            int seed = (!per_ns && (r == 0 || r == nr - 1)) ||
                       (!per_ew && (c == 0 || c == nc - 1));
            if (!seed && interior_invalid)
                for (int d = 0; d < 8 && !seed; ++d) {
                    const int rr = r + D8R[d], cc = c + D8C[d];
                    if (rr < 0 || rr >= nr || cc < 0 || cc >= nc) continue;
                    if (!valid[(int64_t)rr * nc + cc]) seed = 1;
                }
            if (seed) {
                node v = {filled[k], k};
                heap_push(heap, &hn, v);
                state[k] = 1;
            }
        }
    /* epsilon-free flood: the FIFO of raised cells of the reference only
     * reorders pops of equal level, which cannot change the spill values */
    while (hn > 0) {
        node cn = heap_pop(heap, &hn);
        if (state[cn.k] == 2) continue;
        state[cn.k] = 2;
        const int r = (int)(cn.k / nc), c = (int)(cn.k % nc);
        for (int d = 0; d < 8; ++d) {
            int rr = r + D8R[d], cc = c + D8C[d];
            if (per_ns) { if (rr < 0) rr += nr; else if (rr >= nr) rr -= nr; }
            if (per_ew) { if (cc < 0) cc += nc; else if (cc >= nc) cc -= nc; }
            if (rr < 0 || rr >= nr || cc < 0 || cc >= nc) continue;
            const int64_t k = (int64_t)rr * nc + cc;
            if (!valid[k] || state[k] != 0) continue;
            if (filled[k] <= cn.z) filled[k] = cn.z;
            node v = {filled[k], k};
            heap_push(heap, &hn, v);
            state[k] = 1;
        }
    }
    free(valid);
    free(state);
    free(heap);
    return ORC_OK;
}
