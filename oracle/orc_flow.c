/*
 * orc_flow.c -- D8 flow object restated (receivers, donors, outlets, Kahn
 * topological order, s-vector, accumulation).  TEST INFRASTRUCTURE ONLY.
 *
 * Reference: src/cpu/lsdtt_flowinfo_cpu.cpp
 *   resolve_neighbour           :101-158
 *   is_forced_outlet_cell       :162-175
 *   build_receiver_graph_d8     :1077-1329
 *   build_donor_counts_and_outlets_d8 :1340-1408
 *   build_topological_order     :1428-1522
 *   build_s_vector              :1544-1731
 *   accumulate                  :1979-2046
 *   get_drainage_area_raster    :2095-2150
 * Direction tables: src/utils/lsdtt_grid_neighbours.hpp:50-51,75.
 *
 * Parity: pinned by the 17 fixtures x 4 option configurations of
 * tests/cpp/test_flow_d8_identity.cpp (FNV-1a hashes of every output); see
 * tests/test_oracle_flow.py.
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

static const int D8R[8] = {-1, -1, 0, 1, 1, 1, 0, -1};
static const int D8C[8] = {0, 1, 1, 1, 0, -1, -1, -1};
static const double SQRT2 = 1.4142135623730950488;

/* step from (r,c) in direction d under the boundary string.
 * returns 0 = blocked (no-flux), 1 = inside (writes *rr,*cc), 2 = off-grid
 * outlet path ('o'/'b').  lsdtt_flowinfo_cpu.cpp:101-158 */
static int resolve(int r, int c, int d, int nr, int nc, const char* bc,
                   int* rr, int* cc)
{
    int tr = r + D8R[d], tc = c + D8C[d];
    int off = 0;
    if (tr < 0) {
        if (bc[0] == 'p') tr = nr - 1;
        else if (bc[0] == 'o' || bc[0] == 'b') off = 1;
        else return 0;
    } else if (tr >= nr) {
        if (bc[2] == 'p') tr = 0;
        else if (bc[2] == 'o' || bc[2] == 'b') off = 1;
        else return 0;
    }
    if (tc < 0) {
        if (bc[3] == 'p') tc = nc - 1;
        else if (bc[3] == 'o' || bc[3] == 'b') off = 1;
        else return 0;
    } else if (tc >= nc) {
        if (bc[1] == 'p') tc = 0;
        else if (bc[1] == 'o' || bc[1] == 'b') off = 1;
        else return 0;
    }
    if (off) return 2;
    *rr = tr;
    *cc = tc;
    return 1;
}

int orc_d8_receivers(int nr, int nc, double dx, const double* z,
                     const uint8_t* mask, const char* bc,
                     int interior_nodata_is_outlet, int allow_pits_as_outlets,
                     const uint8_t* forced, int64_t* receiver, uint8_t* dir,
                     int64_t* fail_rc)
{
    const double card = dx, diag = dx * SQRT2;
    int failed = 0;
    for (int r = 0; r < nr; ++r)
        for (int c = 0; c < nc; ++c) {
            const int64_t f = (int64_t)r * nc + c;
            if (!mask[f]) { receiver[f] = -1; dir[f] = 255; continue; }
            if ((r == 0 && bc[0] == 'b') || (c == nc - 1 && bc[1] == 'b') ||
                (r == nr - 1 && bc[2] == 'b') || (c == 0 && bc[3] == 'b')) {
                receiver[f] = -2; dir[f] = 255; continue;
            }
            if (forced && forced[f]) { receiver[f] = -2; dir[f] = 255; continue; }
            double best = 0.0;
            int64_t brcv = -2;
            int bdir = 255, allowed = 0;
            for (int d = 0; d < 8; ++d) {
                int rr, cc;
                int k = resolve(r, c, d, nr, nc, bc, &rr, &cc);
                if (k == 2) { allowed = 1; continue; }
                if (k == 0) continue;
                const int64_t g = (int64_t)rr * nc + cc;
                if (!mask[g]) { if (interior_nodata_is_outlet) allowed = 1; continue; }
                if (g == f) continue;
                const double dz = z[f] - z[g];
                if (dz <= 0.0) continue;
                const double s = dz / ((d % 2 == 0) ? card : diag);
                if (s > best) { best = s; brcv = g; bdir = d; }
            }
            if (bdir != 255) { receiver[f] = brcv; dir[f] = (uint8_t)bdir; }
            else {
                receiver[f] = -2; dir[f] = 255;
                if (!allowed && !allow_pits_as_outlets && !failed) {
                    failed = 1;
                    if (fail_rc) *fail_rc = f;
                }
            }
        }
    return failed ? ORC_ERR_UNRESOLVED : ORC_OK;
}

int64_t orc_donors_outlets(int nr, int nc, const uint8_t* mask,
                           const int64_t* receiver, uint8_t* donors,
                           int64_t* outlets)
{
    const int64_t n = (int64_t)nr * nc;
    int64_t no = 0;
    memset(donors, 0, (size_t)n);
    for (int64_t f = 0; f < n; ++f) {
        if (!mask[f]) continue;
        const int64_t rc = receiver[f];
        if (rc >= 0) {
            if (rc >= n || rc == f || donors[rc] == 255) return -1;
            ++donors[rc];
        } else if (rc == -2) {
            outlets[no++] = f;
        }
    }
    return no;
}

int64_t orc_topo_kahn(int nr, int nc, const uint8_t* mask,
                      const int64_t* receiver, const uint8_t* donors,
                      int64_t* topo)
{
    const int64_t n = (int64_t)nr * nc;
    uint8_t* rem = (uint8_t*)malloc((size_t)(n ? n : 1));
    memcpy(rem, donors, (size_t)n);
    int64_t tail = 0, head = 0;
    for (int64_t f = 0; f < n; ++f)
        if (mask[f] && rem[f] == 0) topo[tail++] = f;
    while (head < tail) {
        const int64_t f = topo[head++];
        const int64_t rc = receiver[f];
        if (rc >= 0) {
            if (rem[rc] == 0) { free(rem); return -1; }
            if (--rem[rc] == 0) topo[tail++] = rc;
        }
    }
    free(rem);
    return tail;
}

/* donors of `flat` in ascending flat order, de-duplicated (periodic wraps
 * can reach one cell by two directions).  lsdtt_flowinfo_cpu.cpp:1583-1676 */
static int donors_of(int64_t flat, int nr, int nc, const uint8_t* mask,
                     const char* bc, const int64_t* receiver, int64_t* out)
{
    const int r = (int)(flat / nc), c = (int)(flat % nc);
    int k = 0;
    for (int d = 0; d < 8; ++d) {
        int rr, cc;
        if (resolve(r, c, d, nr, nc, bc, &rr, &cc) != 1) continue;
        const int64_t g = (int64_t)rr * nc + cc;
        if (!mask[g] || receiver[g] != flat) continue;
        int dup = 0;
        for (int i = 0; i < k; ++i) if (out[i] == g) dup = 1;
        if (!dup) out[k++] = g;
    }
    for (int i = 1; i < k; ++i) {          /* insertion sort, ascending */
        int64_t key = out[i];
        int j = i;
        while (j > 0 && out[j - 1] > key) { out[j] = out[j - 1]; --j; }
        out[j] = key;
    }
    return k;
}

int64_t orc_s_vector(int nr, int nc, const uint8_t* mask, const char* bc,
                     const int64_t* receiver, const int64_t* outlets,
                     int64_t n_outlets, int64_t* s_vector, int64_t* s_start,
                     uint64_t* s_size)
{
    const int64_t n = (int64_t)nr * nc;
    int64_t nvalid = 0;
    for (int64_t f = 0; f < n; ++f) {
        s_start[f] = -1;
        s_size[f] = 0;
        if (mask[f]) ++nvalid;
    }
    if (nvalid == 0) return 0;
    if (n_outlets == 0) return -ORC_ERR_NO_OUTLETS;

    typedef struct { int64_t flat; int64_t don[8]; int nd, next; } frame;
    int64_t cap = 1024;
    frame* stack = (frame*)malloc(sizeof(frame) * (size_t)cap);
    int64_t len = 0;
    for (int64_t o = 0; o < n_outlets; ++o) {
        const int64_t of = outlets[o];
        if (s_start[of] >= 0) continue;
        int64_t sp = 0;
        stack[sp].flat = of;
        s_start[of] = len; s_size[of] = 1; s_vector[len++] = of;
        stack[sp].nd = donors_of(of, nr, nc, mask, bc, receiver, stack[sp].don);
        stack[sp].next = 0;
        ++sp;
        while (sp > 0) {
            frame* fr = &stack[sp - 1];
            if (fr->next < fr->nd) {
                const int64_t ch = fr->don[fr->next++];
                if (s_start[ch] >= 0) { free(stack); return -1; }
                if (sp == cap) {
                    cap *= 2;
                    stack = (frame*)realloc(stack, sizeof(frame) * (size_t)cap);
                }
                frame* nf = &stack[sp++];
                nf->flat = ch;
                s_start[ch] = len; s_size[ch] = 1; s_vector[len++] = ch;
                nf->nd = donors_of(ch, nr, nc, mask, bc, receiver, nf->don);
                nf->next = 0;
                continue;
            }
            const uint64_t sz = s_size[fr->flat];
            --sp;
            if (sp > 0) s_size[stack[sp - 1].flat] += sz;
        }
    }
    free(stack);
    if (len != nvalid) return -1;
    return len;
}

void orc_accumulate(int64_t n, const uint8_t* mask, const int64_t* receiver,
                    const int64_t* topo, int64_t n_topo, const double* weights,
                    double dx, double* accum)
{
    const double w0 = dx * dx;
    for (int64_t f = 0; f < n; ++f)
        accum[f] = mask[f] ? (weights ? weights[f] : w0) : 0.0;
    for (int64_t i = 0; i < n_topo; ++i) {
        const int64_t f = topo[i];
        const int64_t rc = receiver[f];
        if (rc >= 0) accum[rc] += accum[f];
    }
}

void orc_upstream_counts(int64_t n, const uint8_t* mask,
                         const int64_t* receiver, const int64_t* topo,
                         int64_t n_topo, int64_t* counts)
{
    for (int64_t f = 0; f < n; ++f) counts[f] = mask[f] ? 1 : 0;
    for (int64_t i = 0; i < n_topo; ++i) {
        const int64_t f = topo[i];
        const int64_t rc = receiver[f];
        if (rc >= 0) counts[rc] += counts[f];
    }
}

/* 64-bit FNV-1a continuation (tests/cpp/test_flow_d8_identity.cpp:118-134) */
uint64_t orc_fnv1a(uint64_t h, const void* p, int64_t n)
{
    const unsigned char* c = (const unsigned char*)p;
    for (int64_t i = 0; i < n; ++i) { h ^= c[i]; h *= 1099511628211ULL; }
    return h;
}
