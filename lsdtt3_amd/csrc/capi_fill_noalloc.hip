// capi_fill_noalloc.hip -- the epsilon-free fill with no device allocation
// (lsdtt_gpu_fill_device_noalloc in lsdtt3_gpu.h).
//
// lsdtt_gpu_fill_device keeps its working keys in a workspace the library
// allocates.  This variant takes them from `scratch`, a file-scope device
// array of fixed size, so a caller that may not allocate (graph capture,
// a serving loop, a harness that forbids hipMalloc) can still fill.  The
// array's host-side name is only a stand-in: its device address is looked
// up once per device with hipGetSymbolAddress and that address is what the
// kernels receive.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/lsdtt3_gpu.h"
#include "capi_util.hpp"
#include "common.hpp"

namespace lsdtt_gpu {
size_t fill_scratch_ints(int nr, int nc);
int fill_device(const float* z, const uint8_t* mask, int nr, int nc, double nodata, BC bc, float* out,
                int* scratch, hipStream_t s);
}  // namespace lsdtt_gpu

using namespace lsdtt_gpu;

namespace {

// room for the fill's working arrays of every grid of up to 2^22 cells
// (2048 x 2048, any shape: 10 ints per cell -- cheapest edges 2, keys 2,
// labels, parents, hooks 2, two boundary-cell lists -- a byte per cell, and
// 4096 for flags and compaction offsets): 164 MiB
constexpr size_t kScratchInts = 10ull * (1ull << 22) + (1ull << 20) + 4096ull;

static __device__ __align__(256) int scratch[kScratchInts];

constexpr int kMaxDevices = 64;

struct ScratchTable {
    std::mutex lock;  // serialises the calls of this process (one array per device)
    int* ptr[kMaxDevices] = {};
};

ScratchTable& table()
{
    static ScratchTable* t = new ScratchTable();  // never destroyed: the runtime may be torn down first
    return *t;
}

bool parse_fill_bc(const char* s, BC& bc)
{
    for (int i = 0; i < 4; ++i)
        if (s[i] != 'o' && s[i] != 'n' && s[i] != 'p' && s[i] != 'b') return false;
    bc = BC{s[0], s[1], s[2], s[3]};
    return true;
}

// the current device's copy of `scratch` (looked up on the device's first
// call); the table lock must be held
int* device_scratch()
{
    int dev = 0;
    LSDTT_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device index beyond the scratch table");
    ScratchTable& t = table();
    if (!t.ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw std::runtime_error(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
        t.ptr[dev] = static_cast<int*>(p);
    }
    return t.ptr[dev];
}

}  // namespace

extern "C" {

int64_t lsdtt_gpu_fill_noalloc_max_cells(void) { return int64_t(1) << 22; }

int lsdtt_gpu_fill_device_noalloc(const float* d_z, const uint8_t* d_mask, int32_t n_rows, int32_t n_cols,
                                  double nodata, const char* boundary_conditions, float* d_filled, void* stream)
{
    return guarded([&] {
        if (!d_z || !d_mask || !d_filled || !boundary_conditions) return fail_arg("null argument");
        if (n_rows <= 0 || n_cols <= 0) return fail_arg("grid must have at least one row and column");
        BC bc;
        if (!parse_fill_bc(boundary_conditions, bc)) return fail_arg("boundary conditions must be four of o/n/p/b");
        const bool pns = bc.n == 'p', pew = bc.e == 'p';
        if (pns != (bc.s == 'p')) return fail_arg("lsdtt::fill_pits: north and south periodic flags must match");
        if (pew != (bc.w == 'p')) return fail_arg("lsdtt::fill_pits: east and west periodic flags must match");
        if (pns && pew)
            return fail_arg("lsdtt::fill_pits: periodic in BOTH directions, so the grid has no drainage exit");
        const long long n = (long long)n_rows * n_cols;
        if (n > lsdtt_gpu_fill_noalloc_max_cells() || fill_scratch_ints(n_rows, n_cols) > kScratchInts)
            return fail_arg("lsdtt_gpu_fill_device_noalloc: grid of " + std::to_string(n) +
                            " cells exceeds the fixed scratch (at most 4194304 cells); use lsdtt_gpu_fill_device");
        ScratchTable& t = table();
        std::lock_guard<std::mutex> hold(t.lock);
        int* work = device_scratch();
        const hipStream_t s = static_cast<hipStream_t>(stream);
        // every scratch word the fill reads is written by it first (keys,
        // pointers, active tiles, flags): the last call's contents do not
        // matter
        fill_device(d_z, d_mask, n_rows, n_cols, nodata, bc, d_filled, work, s);
        LSDTT_HIP(hipStreamSynchronize(s));
        return LSDTT_GPU_OK;
    });
}

}  // extern "C"
