// capi_ensemble.hip -- batched ensemble of independent LEM instances of
// uneven grid sizes and parameters on one GPU (lsdtt3_gpu.h).
//
// Each instance is a device-resident LEM context with its own non-blocking
// HIP stream.  A pool of host workers takes instances largest first and
// runs all of an instance's steps back to back; the step's host decisions
// (pass settling, Krylov and Newton convergence) only wait on that
// instance's stream, so the kernels of different instances overlap on the
// GPU and small grids that cannot fill it alone share it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "../../include/lsdtt3_gpu.h"
#include "capi_lem.hpp"
#include "capi_util.hpp"
#include "lem.hpp"

using namespace lsdtt_gpu;

namespace {

struct Instance {
    std::unique_ptr<LemEngine> eng;
    hipStream_t stream = nullptr;
    int nr = 0, nc = 0;
    bool z_set = false;
    LemStats last;
    int rc = 0;
    std::string err;
};

}  // namespace

struct lsdtt_gpu_ensemble {
    std::vector<Instance> inst;
    int workers = 1;
    double last_seconds = 0.0;
    ~lsdtt_gpu_ensemble()
    {
        (void)hipDeviceSynchronize();
        for (auto& i : inst) {
            i.eng.reset();
            if (i.stream) (void)hipStreamDestroy(i.stream);
        }
    }
};

extern "C" {

lsdtt_gpu_ensemble* lsdtt_gpu_ensemble_create(int32_t n_instances, const int32_t* n_rows, const int32_t* n_cols,
                                              const double* cell_size, const uint8_t* const* masks,
                                              const lsdtt_gpu_lem_options* options, int32_t n_workers)
{
    return guarded_ptr<lsdtt_gpu_ensemble>([&]() -> lsdtt_gpu_ensemble* {
        if (n_instances <= 0 || !n_rows || !n_cols || !cell_size || !masks || !options)
            throw std::invalid_argument("null argument or no instances");
        auto E = std::make_unique<lsdtt_gpu_ensemble>();
        E->inst.resize(static_cast<size_t>(n_instances));
        E->workers = std::max(1, std::min<int>(n_workers > 0 ? n_workers : 16, n_instances));
        for (int i = 0; i < n_instances; ++i) {
            const lsdtt_gpu_lem_options& o = options[i];
            if (n_rows[i] <= 0 || n_cols[i] <= 0 || !masks[i])
                throw std::invalid_argument("instance " + std::to_string(i) + ": empty grid or null mask");
            if (!(cell_size[i] > 0.0) || !std::isfinite(cell_size[i]))
                throw std::invalid_argument("instance " + std::to_string(i) + ": cell_size must be positive");
            validate_options(o);
            Instance& I = E->inst[static_cast<size_t>(i)];
            I.nr = n_rows[i];
            I.nc = n_cols[i];
            I.eng = std::make_unique<LemEngine>(n_rows[i], n_cols[i], cell_size[i], masks[i],
                                                to_params(o, cell_size[i]));
            LSDTT_HIP(hipStreamCreateWithFlags(&I.stream, hipStreamNonBlocking));
            // the reference's fixed-elevation set for the boundary codes
            std::vector<int64_t> bf(static_cast<size_t>(n_rows[i]) * n_cols[i]);
            int64_t nb = 0;
            if (lsdtt_gpu_build_boundary_flats(n_rows[i], n_cols[i], masks[i], o.boundary_conditions, bf.data(),
                                               &nb) != LSDTT_GPU_OK)
                throw std::invalid_argument(lsdtt_gpu_last_error());
            std::vector<long long> v(bf.begin(), bf.begin() + nb);
            I.eng->set_boundary_flats(v.data(), nb);
        }
        LSDTT_HIP(hipDeviceSynchronize());
        return E.release();
    });
}

void lsdtt_gpu_ensemble_destroy(lsdtt_gpu_ensemble* e) { delete e; }

int lsdtt_gpu_ensemble_set_elevation(lsdtt_gpu_ensemble* e, int32_t i, const double* z)
{
    return guarded([&] {
        if (!e || !z || i < 0 || i >= static_cast<int>(e->inst.size())) return fail_arg("bad instance or null");
        e->inst[static_cast<size_t>(i)].eng->set_elevation(z);
        e->inst[static_cast<size_t>(i)].z_set = true;
        return LSDTT_GPU_OK;
    });
}

int lsdtt_gpu_ensemble_get_elevation(lsdtt_gpu_ensemble* e, int32_t i, double* z)
{
    return guarded([&] {
        if (!e || !z || i < 0 || i >= static_cast<int>(e->inst.size())) return fail_arg("bad instance or null");
        e->inst[static_cast<size_t>(i)].eng->get_elevation(z);
        return LSDTT_GPU_OK;
    });
}

int lsdtt_gpu_ensemble_run(lsdtt_gpu_ensemble* e, int32_t n_steps, lsdtt_gpu_lem_stats* last_stats,
                           double* seconds)
{
    return guarded([&] {
        if (!e || n_steps < 0) return fail_arg("null ensemble or negative step count");
        for (auto& I : e->inst)
            if (!I.z_set) return fail_arg("elevation not set for every instance");
        const int m = static_cast<int>(e->inst.size());
        // largest first: the longest instances start at once
        std::vector<int> order(static_cast<size_t>(m));
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
            return (long long)e->inst[a].nr * e->inst[a].nc > (long long)e->inst[b].nr * e->inst[b].nc;
        });
        std::atomic<int> next{0};
        LSDTT_HIP(hipDeviceSynchronize());
        const auto t0 = std::chrono::steady_clock::now();
        auto work = [&] {
            for (;;) {
                const int k = next.fetch_add(1);
                if (k >= m) return;
                Instance& I = e->inst[static_cast<size_t>(order[static_cast<size_t>(k)])];
                try {
                    for (int s = 0; s < n_steps; ++s) I.rc |= I.eng->step(&I.last, I.stream);
                    LSDTT_HIP(hipStreamSynchronize(I.stream));
                } catch (const std::exception& ex) {
                    I.rc = -1;
                    I.err = ex.what();
                }
            }
        };
        std::vector<std::thread> pool;
        for (int w = 0; w < e->workers; ++w) pool.emplace_back(work);
        for (auto& t : pool) t.join();
        LSDTT_HIP(hipDeviceSynchronize());
        e->last_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (seconds) *seconds = e->last_seconds;
        for (int i = 0; i < m; ++i) {
            Instance& I = e->inst[static_cast<size_t>(i)];
            if (I.rc < 0) {
                set_error("instance " + std::to_string(i) + ": " + I.err);
                return LSDTT_GPU_ERR_RUNTIME;
            }
            if (last_stats) to_c_stats(I.last, &last_stats[i]);
        }
        return LSDTT_GPU_OK;
    });
}

}  // extern "C"
