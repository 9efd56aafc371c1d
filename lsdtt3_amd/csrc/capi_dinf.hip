// capi_dinf.hip -- D-infinity flow graph and accumulation through the C ABI
// (lsdtt3_gpu.h): the lsdtt_flowgraph_cpu of method d_infinity
// (src/cpu/lsdtt_flowgraph_cpu.cpp:87-177, 423-663, 988-1033) on host arrays.
#include <hip/hip_runtime.h>

#include <cctype>
#include <cmath>
#include <vector>

#include "../../include/lsdtt3_gpu.h"
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

}  // namespace

extern "C" {

int lsdtt_gpu_dinf_graph(const float* z, const uint8_t* mask, int32_t n_rows, int32_t n_cols, double cell_size,
                         const char* boundary_conditions, int32_t interior_nodata_is_outlet, double nodata,
                         int64_t* receivers, float* weights, float* slope, float* bearing, double* area)
{
    return guarded([&] {
        if (!z || !mask || !boundary_conditions) return fail_arg("null argument");
        if (n_rows <= 0 || n_cols <= 0) return fail_arg("grid must have at least one row and column");
        if (!(cell_size > 0.0) || !std::isfinite(cell_size)) return fail_arg("cell_size must be positive");
        BC bc;
        if (!parse_dinf_bc(boundary_conditions, bc))
            return fail_arg("lsdtt_flowgraph_cpu: boundary_conditions must be four of o/n/p/b with periodic "
                            "edges in north-south and east-west pairs");
        const long long n = (long long)n_rows * n_cols;
        if (2 * n >= (1ll << 31) - 1) return fail_arg("grids of 2^30 cells or more are not supported");
        dbuf<float> dz, dw, dslope, dbear;
        dbuf<uint8_t> dm;
        dbuf<int> drec;
        dz.reserve(n);
        dm.reserve(n);
        drec.reserve(2 * n);
        dw.reserve(2 * n);
        if (slope) dslope.reserve(n);
        if (bearing) dbear.reserve(n);
        LSDTT_HIP(hipMemcpy(dz.p, z, sizeof(float) * n, hipMemcpyHostToDevice));
        LSDTT_HIP(hipMemcpy(dm.p, mask, n, hipMemcpyHostToDevice));
        DinfOptions o;
        o.dx = cell_size;
        o.bc = bc;
        o.interior_nodata_is_outlet = interior_nodata_is_outlet != 0;
        o.use_nodata = std::isfinite(nodata);
        o.nodata = static_cast<float>(nodata);
        dinf_receivers<float>(dz.p, dm.p, n_rows, n_cols, o, drec.p, dw.p, slope ? dslope.p : nullptr,
                              bearing ? dbear.p : nullptr, nullptr);
        if (area) {
            // accumulation counts valid cells only (the graph's validity)
            dbuf<uint8_t> dv;
            dv.reserve(n);
            std::vector<uint8_t> v(static_cast<size_t>(n));
            const float nd = static_cast<float>(nodata);
            for (long long i = 0; i < n; ++i)
                v[static_cast<size_t>(i)] = mask[i] && std::isfinite(z[i]) && !(o.use_nodata && z[i] == nd);
            LSDTT_HIP(hipMemcpy(dv.p, v.data(), n, hipMemcpyHostToDevice));
            FluvialTiles tiles;
            tiles.resize(n_rows, n_cols, bc);
            tiles.dinf_order(drec.p, dv.p, nullptr);
            dbuf<double> da;
            da.reserve(n);
            tiles.dinf_accumulate(drec.p, dw.p, dv.p, cell_size * cell_size, da.p, false, nullptr);
            LSDTT_HIP(hipMemcpy(area, da.p, sizeof(double) * n, hipMemcpyDeviceToHost));
        }
        if (receivers) {
            dbuf<long long> r64;
            r64.reserve(2 * n);
            k_widen_rec<<<div_up(2 * n, 256), 256>>>(drec.p, 2 * n, r64.p);
            LSDTT_LAUNCH_CHECK();
            LSDTT_HIP(hipMemcpy(receivers, r64.p, sizeof(long long) * 2 * n, hipMemcpyDeviceToHost));
        }
        if (weights) LSDTT_HIP(hipMemcpy(weights, dw.p, sizeof(float) * 2 * n, hipMemcpyDeviceToHost));
        if (slope) LSDTT_HIP(hipMemcpy(slope, dslope.p, sizeof(float) * n, hipMemcpyDeviceToHost));
        if (bearing) LSDTT_HIP(hipMemcpy(bearing, dbear.p, sizeof(float) * n, hipMemcpyDeviceToHost));
        LSDTT_HIP(hipDeviceSynchronize());
        return LSDTT_GPU_OK;
    });
}

}  // extern "C"
