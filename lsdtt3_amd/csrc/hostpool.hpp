// hostpool.hpp -- a small pool of host threads for the library's host-side
// loops (values the device cannot reproduce bit for bit, such as the C
// library's pow).  Started on first use, kept until the process ends.
#pragma once

#include <functional>

namespace lsdtt_gpu {

// Calls fn(begin, end) on disjoint chunks covering [0, n), on the pool's
// threads and the caller's, and returns when all chunks are done.  Chunks
// hold at least `grain` items; a range below two grains runs on the caller
// alone.  Calls from different threads run one at a time.
void host_parallel_for(long long n, long long grain, const std::function<void(long long, long long)>& fn);

}  // namespace lsdtt_gpu
