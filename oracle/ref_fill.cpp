// Test infrastructure only: a C-ABI shim over the reference's OWN depression
// fill, compiled from the reference sources where they lie (recipe: the `ref`
// target of oracle/Makefile, output oracle/_ref/libref_fill.so).  Nothing in
// the product links or loads this; the tests use it to pin the oracle's fill
// and the GPU fill against the reference itself, and the benchmark times it as
// the reference CPU baseline of the fill stage.
//
// Entry points bound here:
//   lsdtt::fill_pits(raster, min_slope, boundary_conditions, n_threads)
//     (src/cpu/lsdtt_raster_flow_preprocessing_cpu.hpp:191-198) -- the call
//     run_lem_step makes (src/cpu/lsdtt_lem_cpu.cpp:826);
//   lsdtt::fill_pits_priority_flood(raster, min_slope, n_threads)
//     (src/cpu/lsdtt_raster_flow_preprocessing_cpu.hpp:298-303).
// The raster is built with the (data, mask, meta) constructor
// (src/cpu/lsdtt_raster_cpu.hpp:71-75); invalid cells carry their input value.
#include "lsdtt_raster_flow_preprocessing_cpu.hpp"

#include <cstdint>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#define REF_API extern "C" __attribute__((visibility("default")))

namespace {

lsdtt::lsdtt_raster_cpu make_raster(int nr, int nc, double dx, const float* z,
                                    const std::uint8_t* mask)
{
    lsdtt::lsdtt_raster_meta meta;
    meta.n_rows = nr;
    meta.n_cols = nc;
    meta.cell_size = dx;
    const std::size_t n = std::size_t(nr) * std::size_t(nc);
    std::vector<float> data(z, z + n);
    std::vector<std::uint8_t> m(mask, mask + n);
    return lsdtt::lsdtt_raster_cpu(std::move(data), std::move(m), meta);
}

thread_local std::string g_err;

}  // namespace

// 0 on success, 1 on a reference exception (message via ref_fill_error()).
// mode 0: fill_pits(raster, min_slope, bc, n_threads)
// mode 1: fill_pits_priority_flood(raster, min_slope, n_threads)
REF_API int ref_fill(int nr, int nc, double dx, const float* z, const std::uint8_t* mask,
                     float min_slope, const char* bc, int n_threads, int mode, float* out)
{
    try {
        lsdtt::lsdtt_raster_cpu dem = make_raster(nr, nc, dx, z, mask);
        lsdtt::lsdtt_raster_cpu f = mode == 0
            ? lsdtt::fill_pits(dem, min_slope, std::string(bc), n_threads)
            : lsdtt::fill_pits_priority_flood(dem, min_slope, n_threads);
        const std::vector<float>& d = f.get_data();
        std::memcpy(out, d.data(), d.size() * sizeof(float));
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

REF_API const char* ref_fill_error() { return g_err.c_str(); }
