// flow.hpp -- D8 flow engine on the device.
//
// Replaces lsdtt_flowinfo::create (src/cpu/lsdtt_flowinfo_cpu.cpp:197-253)
// with a data-parallel pipeline whose outputs are byte-identical:
//   receivers            stencil, strict `>` first-wins over dirs 0..7
//   donor counts         gather of neighbours whose receiver is the cell
//   outlets              flag + exclusive scan (row-major order kept)
//   depth from outlet    pointer jumping on packed (distance, next) words
//   level buckets        cells grouped by depth (counting sort)
//   upstream sweep       levels deep -> shallow: counts, Kahn keys, weighted
//                        accumulation (donors summed in topological order)
//   topological order    radix sort of (height-from-source, last-donor
//                        ancestor) keys == the reference's Kahn FIFO order
//   s-vector             sibling offsets + pointer jumping from outlet starts
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
    // weighted double accumulation (w null -> no accumulation).
    void upstream(bool keys, const double* w, double* acc, hipStream_t s);
    // counts only, by atomics (order-free): used on the LEM path
    void upstream_counts_fast(hipStream_t s);
    void topological_order(hipStream_t s);  // requires upstream(keys=true)
    void s_vector(hipStream_t s);           // requires upstream()

    FlowOptions opt;
    long long n = 0;
    const uint8_t* mask = nullptr;  // device, borrowed
    int n_valid = 0, n_outlets = 0, max_depth = 0;
    std::vector<int> level_start_h;  // host copy, max_depth + 2 entries

    dbuf<int> rcv, outlets, depth, root, order, level_count, level_start, cursor;
    dbuf<int> cnt, topo, start, svec, off;
    dbuf<uint8_t> dir, donors;
    dbuf<unsigned long long> pj, key, sort_keys_a, sort_keys_b;
    dbuf<int> flags, scan_out, misc, pjflags;
    dscratch tmp;
    KernelTimer* timer = nullptr;  // optional, owned by the caller

    // per-level processing split: levels [0, tail_from) get one launch each
    // on the upstream/downstream sweeps, levels [tail_from, max_depth] run in
    // one single-workgroup launch
    int tail_from = 0;
    static constexpr int kTailThreshold = 4096;

private:
    void pointer_jump(const int* weight_or_null, int* value_out, int* root_out, hipStream_t s);
};

}  // namespace lsdtt_gpu
