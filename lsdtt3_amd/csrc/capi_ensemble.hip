// capi_ensemble.hip -- batched ensemble of independent LEM instances of
// uneven grid sizes and parameters on one GPU (lsdtt3_gpu.h), on the batch
// engine (batch.hpp): every kernel of a step runs once for all instances,
// the step's launch pattern is replayed from a captured hipGraph.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <memory>
#include <string>
#include <vector>

#include "../../include/lsdtt3_gpu.h"
#include "batch.hpp"
#include "capi_lem.hpp"
#include "capi_util.hpp"

using namespace lsdtt_gpu;

struct lsdtt_gpu_ensemble {
    std::unique_ptr<BatchEngine> eng;
    std::vector<bool> z_set;
    std::vector<LemStats> last;
    bool stepped = false;
};

namespace {

int check_ready(lsdtt_gpu_ensemble* e)
{
    for (bool b : e->z_set)
        if (!b) return fail_arg("elevation not set for every instance");
    return LSDTT_GPU_OK;
}

}  // namespace

extern "C" {

lsdtt_gpu_ensemble* lsdtt_gpu_ensemble_create(int32_t n_instances, const int32_t* n_rows, const int32_t* n_cols,
                                              const double* cell_size, const uint8_t* const* masks,
                                              const lsdtt_gpu_lem_options* options, int32_t n_workers)
{
    (void)n_workers;
    return guarded_ptr<lsdtt_gpu_ensemble>([&]() -> lsdtt_gpu_ensemble* {
        if (n_instances <= 0 || !n_rows || !n_cols || !cell_size || !masks || !options)
            throw std::invalid_argument("null argument or no instances");
        std::vector<BatchSpec> specs(static_cast<size_t>(n_instances));
        for (int i = 0; i < n_instances; ++i) {
            const lsdtt_gpu_lem_options& o = options[i];
            if (n_rows[i] <= 0 || n_cols[i] <= 0 || !masks[i])
                throw std::invalid_argument("instance " + std::to_string(i) + ": empty grid or null mask");
            if (!(cell_size[i] > 0.0) || !std::isfinite(cell_size[i]))
                throw std::invalid_argument("instance " + std::to_string(i) + ": cell_size must be positive");
            validate_options(o);
            BatchSpec& sp = specs[static_cast<size_t>(i)];
            sp.nr = n_rows[i];
            sp.nc = n_cols[i];
            sp.dx = cell_size[i];
            sp.mask = masks[i];
            sp.p = to_params(o, cell_size[i]);
        }
        auto E = std::make_unique<lsdtt_gpu_ensemble>();
        E->eng = std::make_unique<BatchEngine>(specs);
        E->z_set.assign(static_cast<size_t>(n_instances), false);
        // the reference's fixed-elevation set for each instance's boundary codes
        for (int i = 0; i < n_instances; ++i) {
            std::vector<int64_t> bf(static_cast<size_t>(n_rows[i]) * n_cols[i]);
            int64_t nb = 0;
            if (lsdtt_gpu_build_boundary_flats(n_rows[i], n_cols[i], masks[i], options[i].boundary_conditions,
                                               bf.data(), &nb) != LSDTT_GPU_OK)
                throw std::invalid_argument(lsdtt_gpu_last_error());
            std::vector<long long> v(bf.begin(), bf.begin() + nb);
            E->eng->set_boundary_flats(i, v.data(), nb);
        }
        return E.release();
    });
}

void lsdtt_gpu_ensemble_destroy(lsdtt_gpu_ensemble* e) { delete e; }

int lsdtt_gpu_ensemble_set_elevation(lsdtt_gpu_ensemble* e, int32_t i, const double* z)
{
    return guarded([&] {
        if (!e || !z || i < 0 || i >= e->eng->size()) return fail_arg("bad instance or null");
        e->eng->set_elevation(i, z);
        e->z_set[static_cast<size_t>(i)] = true;
        return LSDTT_GPU_OK;
    });
}

int lsdtt_gpu_ensemble_get_elevation(lsdtt_gpu_ensemble* e, int32_t i, double* z)
{
    return guarded([&] {
        if (!e || !z || i < 0 || i >= e->eng->size()) return fail_arg("bad instance or null");
        e->eng->get_elevation(i, z);
        return LSDTT_GPU_OK;
    });
}

int lsdtt_gpu_ensemble_get_erosion_rate(lsdtt_gpu_ensemble* e, int32_t i, float* rate)
{
    return guarded([&] {
        if (!e || !rate || i < 0 || i >= e->eng->size()) return fail_arg("bad instance or null");
        if (!e->stepped) return fail_arg("no step has run");
        e->eng->get_erosion_rate(i, rate);
        return LSDTT_GPU_OK;
    });
}

int lsdtt_gpu_ensemble_run(lsdtt_gpu_ensemble* e, int32_t n_steps, lsdtt_gpu_lem_stats* last_stats,
                           double* seconds)
{
    return guarded([&] {
        if (!e || n_steps < 0) return fail_arg("null ensemble or negative step count");
        const int rc = check_ready(e);
        if (rc != LSDTT_GPU_OK) return rc;
        LSDTT_HIP(hipDeviceSynchronize());
        const auto t0 = std::chrono::steady_clock::now();
        for (int k = 0; k < n_steps; ++k) e->eng->step(e->last);
        LSDTT_HIP(hipStreamSynchronize(e->eng->stream()));
        const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (n_steps > 0) e->stepped = true;
        if (seconds) *seconds = sec;
        if (last_stats)
            for (size_t i = 0; i < e->last.size(); ++i) to_c_stats(e->last[i], &last_stats[i]);
        return LSDTT_GPU_OK;
    });
}

int lsdtt_gpu_ensemble_set_graph(lsdtt_gpu_ensemble* e, int32_t enable)
{
    return guarded([&] {
        if (!e) return fail_arg("null ensemble");
        LSDTT_HIP(hipDeviceSynchronize());
        e->eng->set_graph(enable != 0);
        return LSDTT_GPU_OK;
    });
}

int lsdtt_gpu_ensemble_set_warm_start(lsdtt_gpu_ensemble* e, int32_t mode)
{
    return guarded([&] {
        if (!e) return fail_arg("null ensemble");
        if (mode != 0 && mode != 3) return fail_arg("ensemble warm start mode must be 0 or 3");
        e->eng->set_warm_start(mode);
        return LSDTT_GPU_OK;
    });
}

int lsdtt_gpu_ensemble_pattern(lsdtt_gpu_ensemble* e, int64_t* graph_launches, int64_t* graph_captures,
                               int64_t* extra_rounds, int64_t* extra_passes, int64_t* restarts,
                               int32_t* caps)
{
    return guarded([&] {
        if (!e) return fail_arg("null ensemble");
        const auto& c = e->eng->counters;
        if (graph_launches) *graph_launches = c.graph_launches;
        if (graph_captures) *graph_captures = c.graph_captures;
        if (extra_rounds) *extra_rounds = c.eager_rounds;
        if (extra_passes) *extra_passes = c.eager_passes;
        if (restarts) *restarts = c.refills;
        if (caps) {
            caps[0] = c.count_passes;
            caps[1] = c.sp_passes;
            caps[2] = c.rounds;
        }
        return LSDTT_GPU_OK;
    });
}

int lsdtt_gpu_ensemble_set_tile_height(lsdtt_gpu_ensemble* e, int32_t tile_height)
{
    return guarded([&] {
        if (!e) return fail_arg("null ensemble");
        if (!FluvialTiles::valid_tile_height(tile_height)) return fail_arg("tile height must be 8, 16 or 32");
        e->eng->set_tile_height(tile_height);
        return LSDTT_GPU_OK;
    });
}

int32_t lsdtt_gpu_ensemble_tile_height(lsdtt_gpu_ensemble* e) { return e ? e->eng->tile_height() : 0; }

int lsdtt_gpu_ensemble_autotune_tiles(lsdtt_gpu_ensemble* e, int32_t steps, int32_t* tile_height,
                                      double* ms_per_step)
{
    return guarded([&] {
        if (!e || steps < 1) return fail_arg("null ensemble or fewer than one step");
        const int rc = check_ready(e);
        if (rc != LSDTT_GPU_OK) return rc;
        const int th = e->eng->autotune_tiles(steps, ms_per_step);
        if (tile_height) *tile_height = th;
        return LSDTT_GPU_OK;
    });
}

}  // extern "C"
