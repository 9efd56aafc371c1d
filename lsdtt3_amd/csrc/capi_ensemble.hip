// capi_ensemble.hip -- batched ensemble of independent LEM instances of
// uneven grid sizes and parameters on one GPU (lsdtt3_gpu.h).
//
// Each instance is a device-resident LEM context with its own non-blocking
// HIP stream.  A pool of host workers takes instances largest first and
// runs all of an instance's steps back to back; the step's host decisions
// (pass settling, Krylov and Newton convergence) only wait on that
// instance's stream, so the kernels of different instances overlap on the
// GPU and small grids that cannot fill it alone share it.  The routing
// sweeps' tile height is chosen per instance by measurement
// (lsdtt_gpu_ensemble_autotune_tiles).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <map>
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

    // n_steps steps of every instance, concurrently; wall seconds.  A pool
    // of host workers takes instances largest first (the longest start at
    // once) and runs all of an instance's steps back to back.
    double run_all(int n_steps)
    {
        const int m = static_cast<int>(inst.size());
        std::vector<int> order(static_cast<size_t>(m));
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
            return (long long)inst[a].nr * inst[a].nc > (long long)inst[b].nr * inst[b].nc;
        });
        std::atomic<int> next{0};
        LSDTT_HIP(hipDeviceSynchronize());
        const auto t0 = std::chrono::steady_clock::now();
        auto work = [&] {
            for (;;) {
                const int k = next.fetch_add(1);
                if (k >= m) return;
                Instance& I = inst[static_cast<size_t>(order[static_cast<size_t>(k)])];
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
        for (int w = 0; w < workers; ++w) pool.emplace_back(work);
        for (auto& t : pool) t.join();
        LSDTT_HIP(hipDeviceSynchronize());
        const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        for (int i = 0; i < m; ++i)
            if (inst[static_cast<size_t>(i)].rc < 0)
                throw std::runtime_error("instance " + std::to_string(i) + ": " + inst[static_cast<size_t>(i)].err);
        return sec;
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
        try {
            e->last_seconds = e->run_all(n_steps);
        } catch (const std::runtime_error& ex) {
            set_error(ex.what());
            return LSDTT_GPU_ERR_RUNTIME;
        }
        if (seconds) *seconds = e->last_seconds;
        if (last_stats)
            for (size_t i = 0; i < e->inst.size(); ++i) to_c_stats(e->inst[i].last, &last_stats[i]);
        return LSDTT_GPU_OK;
    });
}

int lsdtt_gpu_ensemble_set_tile_height(lsdtt_gpu_ensemble* e, int32_t i, int32_t tile_height)
{
    return guarded([&] {
        if (!e || i < 0 || i >= static_cast<int>(e->inst.size())) return fail_arg("bad instance or null");
        if (!FluvialTiles::valid_tile_height(tile_height)) return fail_arg("tile height must be 8, 16 or 32");
        LSDTT_HIP(hipDeviceSynchronize());
        e->inst[static_cast<size_t>(i)].eng->tiles.set_tile_height(tile_height);
        return LSDTT_GPU_OK;
    });
}

int lsdtt_gpu_ensemble_autotune_tiles(lsdtt_gpu_ensemble* e, int32_t steps, int32_t* tile_heights,
                                      double* seconds)
{
    return guarded([&] {
        if (!e || steps < 1) return fail_arg("null ensemble or fewer than one step");
        for (auto& I : e->inst)
            if (!I.z_set) return fail_arg("elevation not set for every instance");
        const size_t m = e->inst.size();
        const auto t_begin = std::chrono::steady_clock::now();
        // 1. per size class (octave of the cell count), the height that runs
        //    the class's largest instance fastest on its own
        auto cells = [&](size_t i) { return (long long)e->inst[i].nr * e->inst[i].nc; };
        std::vector<int> cls(m), tuned(m, 16);
        std::map<int, size_t> rep;  // size class -> its largest instance
        for (size_t i = 0; i < m; ++i) {
            cls[i] = 63 - __builtin_clzll(static_cast<unsigned long long>(cells(i)));
            auto it = rep.find(cls[i]);
            if (it == rep.end()) rep.emplace(cls[i], i);
            else if (cells(i) > cells(it->second)) it->second = i;
        }
        for (const auto& [c, r] : rep) {
            Instance& I = e->inst[r];
            const int th = I.eng->autotune_tiles(steps, nullptr, I.stream);
            for (size_t i = 0; i < m; ++i)
                if (cls[i] == c) tuned[i] = th;
        }
        // 2. the instances share the GPU when they run together, which can
        //    favour other heights than running alone: time the ensemble with
        //    each uniform height and with the per-class heights, from the
        //    same surfaces, and keep the fastest assignment
        std::vector<dbuf<double>> saved(m);
        for (size_t i = 0; i < m; ++i) {
            Instance& I = e->inst[i];
            saved[i].reserve(static_cast<size_t>(I.eng->n));
            LSDTT_HIP(hipMemcpy(saved[i].p, I.eng->z.p, sizeof(double) * I.eng->n, hipMemcpyDeviceToDevice));
        }
        auto restore = [&] {
            for (size_t i = 0; i < m; ++i)
                LSDTT_HIP(hipMemcpy(e->inst[i].eng->z.p, saved[i].p, sizeof(double) * e->inst[i].eng->n,
                                    hipMemcpyDeviceToDevice));
        };
        // candidate 0 is the current assignment; another replaces it only
        // when it is faster by more than kMargin (trials of a few steps
        // carry a few per cent of noise)
        constexpr double kMargin = 0.03;
        std::vector<std::vector<int>> cands(1, std::vector<int>(m));
        for (size_t i = 0; i < m; ++i) cands[0][i] = e->inst[i].eng->tiles.tile_height();
        for (int h : FluvialTiles::kTileHeights) cands.emplace_back(m, h);
        cands.push_back(tuned);
        double cur_sec = 0.0, best_sec = 0.0;
        size_t best = 0;
        try {
            for (size_t c = 0; c < cands.size(); ++c) {
                if (c > 0 && std::find(cands.begin(), cands.begin() + static_cast<long>(c), cands[c]) !=
                                 cands.begin() + static_cast<long>(c))
                    continue;  // timed already
                for (size_t i = 0; i < m; ++i) e->inst[i].eng->tiles.set_tile_height(cands[c][i]);
                e->run_all(1);  // settles the pass and Krylov heuristics
                restore();
                const double sec = e->run_all(steps);
                restore();
                if (c == 0) cur_sec = best_sec = sec;
                else if (sec < best_sec && sec < (1.0 - kMargin) * cur_sec) {
                    best_sec = sec;
                    best = c;
                }
            }
        } catch (const std::runtime_error& ex) {
            restore();
            set_error(ex.what());
            return LSDTT_GPU_ERR_RUNTIME;
        }
        for (size_t i = 0; i < m; ++i) {
            e->inst[i].eng->tiles.set_tile_height(cands[best][i]);
            if (tile_heights) tile_heights[i] = cands[best][i];
        }
        LSDTT_HIP(hipDeviceSynchronize());
        if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_begin).count();
        return LSDTT_GPU_OK;
    });
}

}  // extern "C"
