// capi_timestep.hip -- LEM timestep control through the C ABI
// (lsdtt3_gpu.h): diagnosis, adaptive controller and convergence monitor of
// src/cpu/lsdtt_lem_timestep_cpu.hpp:62-292, with the median interior
// gradient computed on the device from the LEM context's resident surface.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>

#include "../../include/lsdtt3_gpu.h"
#include "capi_lem.hpp"
#include "capi_util.hpp"
#include "timestep.hpp"

using namespace lsdtt_gpu;

struct lsdtt_gpu_timestep_controller {
    std::unique_ptr<TimestepController> c;
};

struct lsdtt_gpu_convergence_monitor {
    std::unique_ptr<ConvergenceMonitor> m;
};

namespace {

double lem_gradient(lsdtt_gpu_lem* lem)
{
    LemEngine& e = *lem->eng;
    return median_interior_gradient(e.z.p, e.mask.p, e.nr, e.nc, e.dx, e.grad_keys, nullptr);
}

void copy_message(const std::string& s, char* out, int32_t len)
{
    if (!out || len <= 0) return;
    const size_t k = std::min(s.size(), static_cast<size_t>(len - 1));
    std::memcpy(out, s.data(), k);
    out[k] = '\0';
}

}  // namespace

extern "C" {

void lsdtt_gpu_timestep_options_init(lsdtt_gpu_timestep_options* o)
{
    if (!o) return;
    const TimestepOptions d;
    o->target_ratio = d.target_ratio;
    o->dt_min_yr = d.dt_min_yr;
    o->dt_max_yr = d.dt_max_yr;
    o->grow_factor = d.grow_factor;
    o->shrink_factor = d.shrink_factor;
    o->slope_refresh_interval = d.slope_refresh_interval;
    o->ladder_ratio = d.ladder_ratio;
    o->ratio_smoothing = d.ratio_smoothing;
}

void lsdtt_gpu_convergence_options_init(lsdtt_gpu_convergence_options* o)
{
    if (!o) return;
    const ConvergenceOptions d;
    o->dz_tolerance_m = d.dz_tolerance_m;
    o->dz_rate_tolerance_m_per_yr = d.dz_rate_tolerance_m_per_yr;
    o->patience_steps = d.patience_steps;
    o->min_dt_fraction = d.min_dt_fraction;
    o->stall_window = d.stall_window;
    o->stall_decay_threshold = d.stall_decay_threshold;
    o->oscillation_alternation_fraction = d.oscillation_alternation_fraction;
}

int lsdtt_gpu_lem_median_interior_gradient(lsdtt_gpu_lem* lem, double* out)
{
    return guarded([&] {
        if (!lem || !out) return fail_arg("null argument");
        if (!lem->z_set) return fail_arg("elevation not set");
        *out = lem_gradient(lem);
        return LSDTT_GPU_OK;
    });
}

int lsdtt_gpu_lem_diagnose_timestep(lsdtt_gpu_lem* lem, double uplift_rate, double timestep_yr,
                                    double target_ratio, lsdtt_gpu_timestep_diagnosis* out)
{
    return guarded([&] {
        if (!lem || !out) return fail_arg("null argument");
        if (!lem->z_set) return fail_arg("elevation not set");
        const TimestepDiagnosis d = diagnose_timestep(lem_gradient(lem), lem->eng->dx, uplift_rate, timestep_yr,
                                                      target_ratio > 0.0 ? target_ratio : 0.7);
        out->reference_slope = d.reference_slope;
        out->cell_relief_m = d.cell_relief_m;
        out->uplift_per_step_m = d.uplift_per_step_m;
        out->ratio = d.ratio;
        out->suggested_dt_yr = d.suggested_dt_yr;
        out->too_large = d.too_large ? 1 : 0;
        copy_message(d.warning, out->warning, static_cast<int32_t>(sizeof(out->warning)));
        return LSDTT_GPU_OK;
    });
}

int lsdtt_gpu_lem_set_timestep(lsdtt_gpu_lem* lem, double timestep_yr)
{
    return guarded([&] {
        if (!lem) return fail_arg("null argument");
        if (!(timestep_yr > 0.0) || !std::isfinite(timestep_yr)) return fail_arg("timestep must be positive");
        lem->eng->p.dt = timestep_yr;
        return LSDTT_GPU_OK;
    });
}

lsdtt_gpu_timestep_controller* lsdtt_gpu_timestep_controller_create(double dt_initial_yr,
                                                                    const lsdtt_gpu_timestep_options* o)
{
    return guarded_ptr<lsdtt_gpu_timestep_controller>([&]() -> lsdtt_gpu_timestep_controller* {
        if (!(dt_initial_yr > 0.0)) throw std::invalid_argument("initial timestep must be positive");
        TimestepOptions t;
        if (o) {
            t.target_ratio = o->target_ratio;
            t.dt_min_yr = o->dt_min_yr;
            t.dt_max_yr = o->dt_max_yr;
            t.grow_factor = o->grow_factor;
            t.shrink_factor = o->shrink_factor;
            t.slope_refresh_interval = o->slope_refresh_interval;
            t.ladder_ratio = o->ladder_ratio;
            t.ratio_smoothing = o->ratio_smoothing;
        }
        auto h = std::make_unique<lsdtt_gpu_timestep_controller>();
        h->c = std::make_unique<TimestepController>(dt_initial_yr, t);
        return h.release();
    });
}

void lsdtt_gpu_timestep_controller_destroy(lsdtt_gpu_timestep_controller* c) { delete c; }

int lsdtt_gpu_timestep_controller_update(lsdtt_gpu_timestep_controller* c, lsdtt_gpu_lem* lem,
                                         double max_local_dz_m, double* dt_next)
{
    return guarded([&] {
        if (!c || !lem) return fail_arg("null argument");
        if (!lem->z_set) return fail_arg("elevation not set");
        const double dt = c->c->update(max_local_dz_m, lem->eng->dx, [&] { return lem_gradient(lem); });
        if (dt_next) *dt_next = dt;
        return LSDTT_GPU_OK;
    });
}

int lsdtt_gpu_timestep_controller_get(const lsdtt_gpu_timestep_controller* c, double* dt, double* last_ratio,
                                      double* reference_slope, int64_t* n_shrinks, int64_t* n_grows,
                                      int64_t* n_changes)
{
    return guarded([&] {
        if (!c) return fail_arg("null argument");
        if (dt) *dt = c->c->dt();
        if (last_ratio) *last_ratio = c->c->last_ratio();
        if (reference_slope) *reference_slope = c->c->reference_slope();
        if (n_shrinks) *n_shrinks = c->c->n_shrinks();
        if (n_grows) *n_grows = c->c->n_grows();
        if (n_changes) *n_changes = c->c->n_changes();
        return LSDTT_GPU_OK;
    });
}

lsdtt_gpu_convergence_monitor* lsdtt_gpu_convergence_monitor_create(double dt_initial_yr,
                                                                    const lsdtt_gpu_convergence_options* o)
{
    return guarded_ptr<lsdtt_gpu_convergence_monitor>([&]() -> lsdtt_gpu_convergence_monitor* {
        ConvergenceOptions t;
        if (o) {
            t.dz_tolerance_m = o->dz_tolerance_m;
            t.dz_rate_tolerance_m_per_yr = o->dz_rate_tolerance_m_per_yr;
            t.patience_steps = o->patience_steps;
            t.min_dt_fraction = o->min_dt_fraction;
            t.stall_window = o->stall_window;
            t.stall_decay_threshold = o->stall_decay_threshold;
            t.oscillation_alternation_fraction = o->oscillation_alternation_fraction;
        }
        auto h = std::make_unique<lsdtt_gpu_convergence_monitor>();
        h->m = std::make_unique<ConvergenceMonitor>(dt_initial_yr, t);
        return h.release();
    });
}

void lsdtt_gpu_convergence_monitor_destroy(lsdtt_gpu_convergence_monitor* m) { delete m; }

int lsdtt_gpu_convergence_monitor_update(lsdtt_gpu_convergence_monitor* m, double max_local_dz_m,
                                         double dt_used_yr, int32_t* state)
{
    return guarded([&] {
        if (!m) return fail_arg("null argument");
        const int st = m->m->update(max_local_dz_m, dt_used_yr);
        if (state) *state = st;
        return LSDTT_GPU_OK;
    });
}

int lsdtt_gpu_convergence_monitor_get(const lsdtt_gpu_convergence_monitor* m, int32_t* state,
                                      int32_t* oscillating, int64_t* n_steps, double* last_dz,
                                      double* window_decay, char* message, int32_t message_len)
{
    return guarded([&] {
        if (!m) return fail_arg("null argument");
        if (state) *state = m->m->state();
        if (oscillating) *oscillating = m->m->oscillating() ? 1 : 0;
        if (n_steps) *n_steps = m->m->n_steps();
        if (last_dz) *last_dz = m->m->last_dz();
        if (window_decay) *window_decay = m->m->window_decay();
        copy_message(m->m->message(), message, message_len);
        return LSDTT_GPU_OK;
    });
}

}  // extern "C"
