// flow.hpp -- D8 flow engine on the device.
//
// Replaces lsdtt_flowinfo::create (src/cpu/lsdtt_flowinfo_cpu.cpp:197-253)
// with a data-parallel pipeline whose outputs are byte-identical:
//   receivers            stencil, strict `>` first-wins over dirs 0..7
//   donor counts         gather of neighbours whose receiver is the cell
//                        (and a donor bit per neighbour)
//   outlets              flag + exclusive scan (row-major order kept)
//   depth from outlet    pointer jumping on packed (distance, next) words
//   level buckets        cells grouped by depth (counting sort)
//   upstream sweep       levels deep -> shallow: counts, Kahn keys, s-vector
//                        sibling offsets, weighted accumulation (donors
//                        summed in topological order)
//   topological order    sources (height 0) first in row-major order, then a
//                        radix sort of the other cells' (height-from-source,
//                        last-donor ancestor) keys == the reference's Kahn
//                        FIFO order
//   s-vector             outlet starts (scan of outlet sizes), then levels
//                        shallow -> deep: receiver's start + sibling offset
#pragma once

#include <vector>

#include "common.hpp"
#include "dbuf.hpp"
#include "ktimer.hpp"

namespace lsdtt_gpu {

struct FlowOptions {
    int nr = 0, nc = 0;
    double dx = 1.0;
    BC bc{'o', 'o', 'o', 'o'};
    bool interior_nodata_is_outlet = true;
    bool allow_pits_as_outlets = false;
};

// Routing tiles (32 columns by th rows, ntx per tile row) whose receiver
// structure changed: when dirty is set, a receiver kernel that writes a
// receiver different from the one already stored flags the tile of the cell
// and the tiles of its old and new receivers (their cross-tile donors).
struct TileMark {
    int* dirty = nullptr;
    int nc = 0, th = 0, ntx = 0;
};

// device ints read back with the pointer jumping's round flags
struct FlowReadInts {
    const int* src[6] = {};
    int count = 0;
};

class FlowEngine {
public:
    explicit FlowEngine(const FlowOptions& o);

    // receivers/dirs/donors/outlets, depths and level buckets.  z and mask
    // are device pointers; forced may be null.  Returns 0 or the first
    // unresolved flat + 1 as a negative number.
    template <class T>
    long long route(const T* z, const uint8_t* mask, const uint8_t* forced, hipStream_t s);

    // receivers and directions only (no failure check: for
    // allow_pits_as_outlets routing, where every cell resolves); zcopy: a
    // copy of z; mark: routing tiles to flag where a receiver changed
    template <class T>
    void receivers_only(const T* z, const uint8_t* mask, hipStream_t s, T* zcopy = nullptr,
                        const TileMark& mark = TileMark{});

    // Upstream sweep (requires route()).  keys: Kahn keys too; w/acc: optional
    // weighted double accumulation (w null -> no accumulation); sv_offsets:
    // the s-vector's sibling offsets too (s_vector() then skips its own pass)
    void upstream(bool keys, const double* w, double* acc, hipStream_t s, bool sv_offsets = false);
    // counts only, by atomics (order-free): used on the LEM path
    void upstream_counts_fast(hipStream_t s);
    void topological_order(hipStream_t s);  // requires upstream(keys=true)
    void s_vector(hipStream_t s);           // requires upstream()

    FlowOptions opt;
    long long n = 0;
    const uint8_t* mask = nullptr;  // device, borrowed
    int n_valid = 0, n_outlets = 0, max_depth = 0;
    int n_sources = 0;  // valid cells without donors
    std::vector<int> level_start_h;  // host copy, max_depth + 2 entries

    dbuf<int> rcv, outlets, depth, order, level_count, level_sub, level_start, cursor;
    dbuf<int> cnt, topo, topo_inner, start, svec, off;
    dbuf<uint8_t> dir, donors, dmask;  // dmask: donor bits (distinct_neighbours order)
    dbuf<unsigned long long> pj, key, sort_keys_a, sort_keys_b;
    dbuf<int> flags, scan_out, misc, pjflags, part_counts;
    hbuf<int> hread;  // pinned read-back of the round flags and single ints
    dscratch tmp;
    KernelTimer* timer = nullptr;  // optional, owned by the caller

    // per-level processing split: levels [0, tail_from) get one launch each
    // on the upstream/downstream sweeps, levels [tail_from, max_depth] run in
    // one single-workgroup launch
    int tail_from = 0;
    static constexpr int kTailThreshold = 4096;

    using ReadInts = FlowReadInts;
    static constexpr int kPjRounds = 34;  // enough for a path of 2^31 cells
    static constexpr int kReadSlot = 40;  // where the extra ints land in hread

private:
    // rounds the last pointer jumping needed plus a spare (the first batch
    // of the next one)
    int pj_hint = 12;
    bool off_ready = false;  // off holds the sibling offsets of the current routing
    // depth from the outlet by pointer jumping, into value_out; the extra
    // ints come back with the round flags in one read-back
    void pointer_jump(int* value_out, hipStream_t s, const ReadInts& extra = ReadInts{});
};

}  // namespace lsdtt_gpu
