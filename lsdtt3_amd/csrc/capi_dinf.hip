// capi_dinf.hip -- D-infinity flow graph, accumulation and analytic steady
// state through the C ABI (lsdtt3_gpu.h): the lsdtt_flowgraph_cpu of method
// d_infinity (src/cpu/lsdtt_flowgraph_cpu.cpp:87-177, 423-663, 988-1033) and
// compute_dinf_fluvial_steady_state (src/cpu/lsdtt_steady_state_cpu.cpp:
// 227-436) on host arrays.
#include <hip/hip_runtime.h>

#include <cctype>
#include <cmath>
#include <vector>

#include "../../include/lsdtt3_gpu.h"
#include "capi_steady.hpp"
#include "capi_util.hpp"
#include "dbuf.hpp"
#include "dinf.hpp"
#include "fluvial.hpp"

using namespace lsdtt_gpu;

namespace {

// parse_boundary_conditions (lsdtt_flowgraph_cpu.cpp:252-300): four of
// o/n/p/b in any case, periodic edges in pairs
bool parse_dinf_bc(const char* s, BC& bc)
{
    char c[4];
    for (int i = 0; i < 4; ++i) {
        c[i] = static_cast<char>(std::tolower(static_cast<unsigned char>(s[i])));
        if (c[i] != 'o' && c[i] != 'n' && c[i] != 'p' && c[i] != 'b') return false;
    }
    if ((c[0] == 'p') != (c[2] == 'p') || (c[1] == 'p') != (c[3] == 'p')) return false;
    bc = BC{c[0], c[1], c[2], c[3]};
    return true;
}

__global__ void k_widen_rec(const int* __restrict__ a, long long n, long long* __restrict__ out)
{
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i < n) out[i] = a[i];
}

// the graph of the valid cells on the device: receivers and weights, the
// validity (mask != 0, z finite, z != nodata when nodata is finite) and,
// when asked, the accumulated area (by its tiles)
struct DinfDevice {
    dbuf<float> z, w, slope, bearing;
    dbuf<uint8_t> mask, valid;
    dbuf<int> rec;
    dbuf<double> area;
    std::vector<uint8_t> valid_h;
    FluvialTiles tiles;
};

void build_dinf(const float* z, const uint8_t* mask, int nr, int nc, double cell_size, const BC& bc,
                bool interior_nodata_is_outlet, double nodata, bool want_slope, bool want_bearing, bool want_area,
                DinfDevice& d)
{
    const long long n = (long long)nr * nc;
    d.z.reserve(n);
    d.mask.reserve(n);
    d.rec.reserve(2 * n);
    d.w.reserve(2 * n);
    if (want_slope) d.slope.reserve(n);
    if (want_bearing) d.bearing.reserve(n);
    LSDTT_HIP(hipMemcpy(d.z.p, z, sizeof(float) * n, hipMemcpyHostToDevice));
    LSDTT_HIP(hipMemcpy(d.mask.p, mask, n, hipMemcpyHostToDevice));
    DinfOptions o;
    o.dx = cell_size;
    o.bc = bc;
    o.interior_nodata_is_outlet = interior_nodata_is_outlet;
    o.use_nodata = std::isfinite(nodata);
    o.nodata = static_cast<float>(nodata);
    dinf_receivers<float>(d.z.p, d.mask.p, nr, nc, o, d.rec.p, d.w.p, want_slope ? d.slope.p : nullptr,
                          want_bearing ? d.bearing.p : nullptr, nullptr);
    d.valid_h.assign(static_cast<size_t>(n), 0);
    const float nd = static_cast<float>(nodata);
    for (long long i = 0; i < n; ++i)
        d.valid_h[static_cast<size_t>(i)] = mask[i] && std::isfinite(z[i]) && !(o.use_nodata && z[i] == nd);
    d.valid.reserve(n);
    LSDTT_HIP(hipMemcpy(d.valid.p, d.valid_h.data(), n, hipMemcpyHostToDevice));
    if (want_area) {
        d.tiles.resize(nr, nc, bc);
        d.tiles.dinf_order(d.rec.p, d.valid.p, nullptr);
        d.area.reserve(n);
        d.tiles.dinf_accumulate(d.rec.p, d.w.p, d.valid.p, cell_size * cell_size, d.area.p, false, nullptr);
    }
}

int check_grid(const float* z, const uint8_t* mask, int32_t n_rows, int32_t n_cols, double cell_size,
               const char* boundary_conditions, BC& bc)
{
    if (!z || !mask || !boundary_conditions) return fail_arg("null argument");
    if (n_rows <= 0 || n_cols <= 0) return fail_arg("grid must have at least one row and column");
    if (!(cell_size > 0.0) || !std::isfinite(cell_size)) return fail_arg("cell_size must be positive");
    if (!parse_dinf_bc(boundary_conditions, bc))
        return fail_arg("lsdtt_flowgraph_cpu: boundary_conditions must be four of o/n/p/b with periodic "
                        "edges in north-south and east-west pairs");
    if (2 * (long long)n_rows * n_cols >= (1ll << 31) - 1)
        return fail_arg("grids of 2^30 cells or more are not supported");
    return LSDTT_GPU_OK;
}

}  // namespace

extern "C" {

int lsdtt_gpu_dinf_graph(const float* z, const uint8_t* mask, int32_t n_rows, int32_t n_cols, double cell_size,
                         const char* boundary_conditions, int32_t interior_nodata_is_outlet, double nodata,
                         int64_t* receivers, float* weights, float* slope, float* bearing, double* area)
{
    return guarded([&] {
        BC bc;
        const int st = check_grid(z, mask, n_rows, n_cols, cell_size, boundary_conditions, bc);
        if (st != LSDTT_GPU_OK) return st;
        const long long n = (long long)n_rows * n_cols;
        DinfDevice d;
        build_dinf(z, mask, n_rows, n_cols, cell_size, bc, interior_nodata_is_outlet != 0, nodata, slope != nullptr,
                   bearing != nullptr, area != nullptr, d);
        if (area) LSDTT_HIP(hipMemcpy(area, d.area.p, sizeof(double) * n, hipMemcpyDeviceToHost));
        if (receivers) {
            dbuf<long long> r64;
            r64.reserve(2 * n);
            k_widen_rec<<<div_up(2 * n, 256), 256>>>(d.rec.p, 2 * n, r64.p);
            LSDTT_LAUNCH_CHECK();
            LSDTT_HIP(hipMemcpy(receivers, r64.p, sizeof(long long) * 2 * n, hipMemcpyDeviceToHost));
        }
        if (weights) LSDTT_HIP(hipMemcpy(weights, d.w.p, sizeof(float) * 2 * n, hipMemcpyDeviceToHost));
        if (slope) LSDTT_HIP(hipMemcpy(slope, d.slope.p, sizeof(float) * n, hipMemcpyDeviceToHost));
        if (bearing) LSDTT_HIP(hipMemcpy(bearing, d.bearing.p, sizeof(float) * n, hipMemcpyDeviceToHost));
        LSDTT_HIP(hipDeviceSynchronize());
        return LSDTT_GPU_OK;
    });
}

int lsdtt_gpu_dinf_steady_state(const float* z, const uint8_t* mask, int32_t n_rows, int32_t n_cols,
                                double cell_size, const char* boundary_conditions, int32_t interior_nodata_is_outlet,
                                double nodata, const lsdtt_gpu_steady_state_options* o,
                                const float* erodibility_field, float* z_out, double* area,
                                lsdtt_gpu_steady_state_result* result)
{
    return guarded([&] {
        if (!o || !z_out) return fail_arg("null argument");
        BC bc;
        const int st = check_grid(z, mask, n_rows, n_cols, cell_size, boundary_conditions, bc);
        if (st != LSDTT_GPU_OK) return st;
        const long long n = (long long)n_rows * n_cols;
        const SteadyOptions so = steady_options(o);
        DinfDevice d;
        build_dinf(z, mask, n_rows, n_cols, cell_size, bc, interior_nodata_is_outlet != 0, nodata, false, false,
                   true, d);
        hipStream_t s = nullptr;
        dbuf<float> w, Kd, zd;
        if (erodibility_field) {
            const std::string err = erodibility_field_error(erodibility_field, d.valid_h.data(), n, n_cols, so.K);
            if (!err.empty()) return fail_arg(err);
            Kd.reserve(n);
            w.reserve(n);
            LSDTT_HIP(hipMemcpyAsync(Kd.p, erodibility_field, sizeof(float) * n, hipMemcpyHostToDevice, s));
            chi_weights(d.valid.p, Kd.p, n, so.K, so.n, w.p, s);
        }
        zd.reserve(n);
        const SteadyResult r = steady_state_dinf(d.tiles, d.rec.p, d.w.p, d.valid.p, d.area.p, so,
                                                 erodibility_field ? w.p : nullptr, static_cast<float>(cell_size),
                                                 static_cast<float>(nodata), zd.p, s);
        LSDTT_HIP(hipMemcpy(z_out, zd.p, sizeof(float) * n, hipMemcpyDeviceToHost));
        if (area) LSDTT_HIP(hipMemcpy(area, d.area.p, sizeof(double) * n, hipMemcpyDeviceToHost));
        copy_steady_result(r, result);
        return LSDTT_GPU_OK;
    });
}

}  // extern "C"
