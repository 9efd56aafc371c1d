// capi_lem.hip -- C ABI of the LEM timestep (lsdtt3_gpu.h).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "../../include/lsdtt3_gpu.h"
#include "capi_util.hpp"
#include "capi_lem.hpp"
#include "lem.hpp"

using namespace lsdtt_gpu;

namespace lsdtt_gpu {

bool bc_ok(const char* s)
{
    for (int i = 0; i < 4; ++i)
        if (s[i] != 'o' && s[i] != 'n' && s[i] != 'p' && s[i] != 'b') return false;
    return true;
}

LemParams to_params(const lsdtt_gpu_lem_options& o, double o_cell_size)
{
    LemParams p;
    p.dt = o.timestep;
    p.K = o.erodibility;
    p.m = o.m_exponent;
    p.n = o.n_exponent;
    p.threshold = o.incision_threshold_area;
    p.divide_area_by_dx = o.divide_area_by_dx;
    p.slope_clamp = o.slope_clamp;
    p.D = o.hillslope_diffusivity;
    p.nonlinear = o.enable_nonlinear_hillslopes;
    p.Sc = o.critical_hillslope_slope;
    p.U = o.uplift_rate;
    p.baselevel_fall = o.baselevel_fall_rate;
    // run_lem_step: D8 incision, else D-inf incision (lsdtt_lem_cpu.cpp:324-391)
    p.enable_fluvial = o.enable_d8_fluvial_erosion || o.enable_dinf_fluvial_erosion;
    p.fluvial_dinf = !o.enable_d8_fluvial_erosion && o.enable_dinf_fluvial_erosion;
    p.enable_hillslope = o.enable_hillslope_transport;
    p.enable_uplift = o.enable_uplift;
    p.bc = BC{o.boundary_conditions[0], o.boundary_conditions[1], o.boundary_conditions[2],
              o.boundary_conditions[3]};
    p.newton_abs_tol = o.newton_abs_tol;
    p.newton_rel_tol = o.newton_rel_tol;
    p.newton_max_iterations = o.newton_max_iterations;
    p.newton_max_line_search = o.newton_max_line_search;
    p.slope_ratio_cap = o.slope_ratio_cap;
    p.krylov_rel_tol = o.krylov_rel_tol;
    p.krylov_max_iterations = o.krylov_max_iterations;
    p.area_kernel = o.drainage_area_kernel_size;
    // lsdtt_stream_power_cpu.cpp:181-183
    p.width_factor = o.multiply_by_w_over_dx ? o.channel_width_w / o_cell_size : 1.0;
    p.regime = o.hillslope_coupling_regime;
    p.force_picard = o.hillslope_force_picard;
    p.erosion_rate_field = o.compute_erosion_rate_field;
    p.krylov_forcing = o.krylov_forcing;
    p.picard.picard_tol = o.picard_tol;
    p.picard.picard_max_iter = o.picard_max_iter;
    p.picard.omega = o.relaxation_omega;
    p.picard.r_max = o.picard_slope_ratio_cap;
    p.picard.dt_init = o.picard_dt_init;
    p.picard.dt_min = o.picard_dt_min;
    p.picard.dt_max = o.picard_dt_max;
    p.picard.grow = o.picard_grow_factor;
    p.picard.shrink = o.picard_shrink_factor;
    p.picard.dz_target = o.picard_dz_target;
    return p;
}

void validate_options(const lsdtt_gpu_lem_options& oo)
{
    const lsdtt_gpu_lem_options* o = &oo;
    if (!bc_ok(o->boundary_conditions)) throw std::invalid_argument("boundary conditions must be four of o/n/p/b");
    if (!(o->timestep > 0.0)) throw std::invalid_argument("timestep must be positive");
    if (o->enable_hillslope_transport && o->enable_nonlinear_hillslopes && !(o->critical_hillslope_slope > 0.0))
        throw std::invalid_argument("critical_hillslope_slope must be positive");
    if (o->drainage_area_kernel_size < 1 || o->drainage_area_kernel_size % 2 == 0)
        throw std::invalid_argument("drainage area moving-average kernel size must be a positive odd number");
    if (o->multiply_by_w_over_dx && (!(o->channel_width_w > 0.0) || !std::isfinite(o->channel_width_w)))
        throw std::invalid_argument(
            "channel_width_w must be finite and positive when the width correction is enabled");
    if (o->hillslope_coupling_regime < 0 || o->hillslope_coupling_regime > 2)
        throw std::invalid_argument("hillslope_coupling_regime must be 0 (auto), 1 (coupled) or 2 "
                                    "(channels_as_boundaries)");
}

void to_c_stats(const LemStats& st, lsdtt_gpu_lem_stats* out)
{
    out->delta_max_elevation = st.delta_max_elevation;
    out->delta_mean_elevation = st.delta_mean_elevation;
    out->max_local_elevation_change = st.max_local_elevation_change;
    out->p999_local_elevation_change = st.p999_local_elevation_change;
    out->mean_erosion_rate = st.mean_erosion_rate;
    out->max_erosion_rate = st.max_erosion_rate;
    out->n_valid = st.n_valid;
    out->cells_eroded = st.cells_eroded;
    out->newton_iterations = st.newton_iterations;
    out->krylov_iterations = st.krylov_iterations;
    out->hillslope_substeps = st.substeps;
    out->hillslope_rejected = st.rejected;
    out->hillslope_converged = st.hillslope_converged;
    out->hillslope_initial_residual = st.hillslope_initial_residual;
    out->hillslope_residual = st.hillslope_residual;
    out->hillslope_tolerance = st.hillslope_tolerance;
    out->max_slope_ratio = st.max_slope_ratio;
    out->count_passes = st.count_passes;
    out->count_tile_runs = st.count_tile_runs;
    out->hillslope_method = st.hillslope_method;
    out->picard_iterations = st.picard.picard_iters;
    out->picard_substeps = st.picard.substeps;
    out->picard_rejected = st.picard.rejected;
    out->picard_cg_iterations = st.picard.cg_iters;
    out->picard_limiter_fired = st.picard.limiter_fired;
    out->picard_max_slope_ratio = st.picard.max_slope_ratio;
    out->warm_start_mode = st.warm_start_mode;
    out->warm_start_accepted = st.warm_start_accepted;
    out->warm_start_residual = st.warm_start_residual;
    out->stream_power_tile_runs = st.stream_power_tile_runs;
    out->stream_power_passes = st.stream_power_passes;
    out->newton_restoration_iterations = st.restoration_iterations;
}

}  // namespace lsdtt_gpu

extern "C" {

void lsdtt_gpu_lem_options_init(lsdtt_gpu_lem_options* o)
{
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    // lsdtt_lem_options defaults (lsdtt_lem_cpu.hpp:97-182)
    o->timestep = 100.0;
    o->erodibility = 1e-3;
    o->m_exponent = 0.45;
    o->n_exponent = 1.0;
    o->incision_threshold_area = 0.0;
    o->divide_area_by_dx = 1;
    o->slope_clamp = 1e-5;
    o->hillslope_diffusivity = 1e-3;
    o->enable_nonlinear_hillslopes = 0;
    o->critical_hillslope_slope = 0.6;
    o->uplift_rate = 1e-4;
    o->baselevel_fall_rate = 0.0;
    o->enable_d8_fluvial_erosion = 1;
    o->enable_dinf_fluvial_erosion = 0;
    o->enable_hillslope_transport = 0;
    o->enable_uplift = 1;
    std::memcpy(o->boundary_conditions, "oooo", 4);
    // lsdtt_hillslope_newton_options defaults (newton .hpp:143-165)
    o->newton_abs_tol = 1e-12;
    o->newton_rel_tol = 1e-8;
    o->newton_max_iterations = 50;
    o->newton_max_line_search = 30;
    o->slope_ratio_cap = 0.9999;
    o->krylov_rel_tol = 1e-4;
    o->krylov_max_iterations = 500;
    o->krylov_forcing = 1;
    o->drainage_area_kernel_size = 1;
    o->multiply_by_w_over_dx = 0;
    o->channel_width_w = 1.0;
    o->hillslope_coupling_regime = 0;
    // lsdtt_roering_implicit_options defaults (lsdtt_hillslope_implicit_cpu.hpp:65-97)
    o->hillslope_force_picard = 0;
    o->picard_tol = 1e-4;
    o->picard_max_iter = 50;
    o->relaxation_omega = 1.0;
    o->picard_slope_ratio_cap = 0.99;
    o->picard_dt_init = 0.0;
    o->picard_dt_min = 0.0;
    o->picard_dt_max = 0.0;
    o->picard_grow_factor = 1.5;
    o->picard_shrink_factor = 0.5;
    o->picard_dz_target = 0.0;
}

lsdtt_gpu_lem* lsdtt_gpu_lem_create(int32_t n_rows, int32_t n_cols, double cell_size, const uint8_t* mask,
                                    const lsdtt_gpu_lem_options* o)
{
    return guarded_ptr<lsdtt_gpu_lem>([&]() -> lsdtt_gpu_lem* {
        if (!o || !mask) throw std::invalid_argument("null argument");
        if (n_rows <= 0 || n_cols <= 0) throw std::invalid_argument("grid must have at least one row and column");
        if (!(cell_size > 0.0) || !std::isfinite(cell_size))
            throw std::invalid_argument("cell_size must be finite and positive");
        validate_options(*o);
        auto L = std::make_unique<lsdtt_gpu_lem>();
        L->eng = std::make_unique<LemEngine>(n_rows, n_cols, cell_size, mask, to_params(*o, cell_size));
        return L.release();
    });
}

void lsdtt_gpu_lem_destroy(lsdtt_gpu_lem* lem) { delete lem; }

int lsdtt_gpu_lem_set_elevation(lsdtt_gpu_lem* lem, const double* z)
{
    return guarded([&] {
        if (!lem || !z) return fail_arg("null argument");
        lem->eng->set_elevation(z);
        lem->z_set = true;
        return LSDTT_GPU_OK;
    });
}

int lsdtt_gpu_lem_get_elevation(lsdtt_gpu_lem* lem, double* z)
{
    return guarded([&] {
        if (!lem || !z) return fail_arg("null argument");
        lem->eng->get_elevation(z);
        return LSDTT_GPU_OK;
    });
}

int lsdtt_gpu_lem_set_erodibility_field(lsdtt_gpu_lem* lem, const double* K)
{
    return guarded([&] {
        if (!lem) return fail_arg("null argument");
        if (K) {
            const long long n = lem->eng->n;
            for (long long i = 0; i < n; ++i)
                if (!std::isfinite(K[i]) || K[i] < 0.0)
                    return fail_arg("run_lem_step: erodibility field must be finite and non-negative");
        }
        lem->eng->set_erodibility(K);
        return LSDTT_GPU_OK;
    });
}

int lsdtt_gpu_lem_set_channel_mask(lsdtt_gpu_lem* lem, const uint8_t* channel)
{
    return guarded([&] {
        if (!lem) return fail_arg("null argument");
        lem->eng->set_channel_mask(channel);
        return LSDTT_GPU_OK;
    });
}

double* lsdtt_gpu_lem_elevation_device(lsdtt_gpu_lem* lem)
{
    if (!lem) return nullptr;
    lem->z_set = true;
    return lem->eng->elevation_device();
}

}  // extern "C"

namespace {

// uploads the fixed-elevation set when it differs from the last one
void use_boundary_flats(lsdtt_gpu_lem* lem, const int64_t* boundary_flats, int64_t n_boundary)
{
    if (!lem->bf_set || lem->bf_cache.size() != static_cast<size_t>(n_boundary) ||
        (n_boundary && std::memcmp(lem->bf_cache.data(), boundary_flats, sizeof(int64_t) * n_boundary) != 0)) {
        lem->bf_cache.assign(boundary_flats, boundary_flats + n_boundary);
        std::vector<long long> v(lem->bf_cache.begin(), lem->bf_cache.end());
        lem->eng->set_boundary_flats(v.data(), n_boundary);
        lem->bf_set = true;
    }
}

}  // namespace

extern "C" {

int lsdtt_gpu_lem_step(lsdtt_gpu_lem* lem, const int64_t* boundary_flats, int64_t n_boundary,
                       lsdtt_gpu_lem_stats* out)
{
    return guarded([&] {
        if (!lem || !out || (n_boundary > 0 && !boundary_flats)) return fail_arg("null argument");
        if (!lem->z_set) return fail_arg("elevation not set");
        use_boundary_flats(lem, boundary_flats, n_boundary);
        LemStats st;
        const int rc = lem->eng->step(&st, nullptr);
        to_c_stats(st, out);
        if (rc != 0) {
            set_error("hillslope Newton did not converge at the minimum sub-step");
            return LSDTT_GPU_ERR_NOT_CONVERGED;
        }
        return LSDTT_GPU_OK;
    });
}

int lsdtt_gpu_lem_get_erosion_rate(lsdtt_gpu_lem* lem, float* rate)
{
    return guarded([&] {
        if (!lem || !rate) return fail_arg("null argument");
        LemEngine& e = *lem->eng;
        if (!e.p.erosion_rate_field) return fail_arg("compute_erosion_rate_field is off");
        if (!e.ero_valid) return fail_arg("no step has run since the erosion rate field was enabled");
        LSDTT_HIP(hipDeviceSynchronize());
        LSDTT_HIP(hipMemcpy(rate, e.ero_rate.p, sizeof(float) * e.n, hipMemcpyDeviceToHost));
        return LSDTT_GPU_OK;
    });
}

int lsdtt_gpu_lem_get_channel_mask(lsdtt_gpu_lem* lem, uint8_t* channel)
{
    return guarded([&] {
        if (!lem || !channel) return fail_arg("null argument");
        LemEngine& e = *lem->eng;
        const size_t n = static_cast<size_t>(e.n);
        LSDTT_HIP(hipDeviceSynchronize());
        if (!e.channels_active()) {
            std::memset(channel, 0, n);
        } else if (!e.channel_h.empty()) {
            std::memcpy(channel, e.channel_h.data(), n);
        } else if (e.chan.p && e.darea_warm) {
            LSDTT_HIP(hipMemcpy(channel, e.chan.p, n, hipMemcpyDeviceToHost));
        } else {
            std::memset(channel, 0, n);
        }
        return LSDTT_GPU_OK;
    });
}

int lsdtt_gpu_lem_set_tile_height(lsdtt_gpu_lem* lem, int32_t tile_height)
{
    return guarded([&] {
        if (!lem) return fail_arg("null argument");
        if (!FluvialTiles::valid_tile_height(tile_height)) return fail_arg("tile height must be 8, 16 or 32");
        LSDTT_HIP(hipDeviceSynchronize());
        lem->eng->tiles.set_tile_height(tile_height);
        return LSDTT_GPU_OK;
    });
}

int32_t lsdtt_gpu_lem_tile_height(lsdtt_gpu_lem* lem) { return lem ? lem->eng->tiles.tile_height() : 0; }

int lsdtt_gpu_lem_autotune_tiles(lsdtt_gpu_lem* lem, const int64_t* boundary_flats, int64_t n_boundary,
                                 int32_t steps, int32_t* tile_height, double* ms_per_step)
{
    return guarded([&] {
        if (!lem || (n_boundary > 0 && !boundary_flats)) return fail_arg("null argument");
        if (!lem->z_set) return fail_arg("elevation not set");
        if (steps < 1) return fail_arg("autotune needs at least one step per tile height");
        use_boundary_flats(lem, boundary_flats, n_boundary);
        const int th = lem->eng->autotune_tiles(steps, ms_per_step, nullptr);
        if (tile_height) *tile_height = th;
        return LSDTT_GPU_OK;
    });
}

int lsdtt_gpu_lem_profile(lsdtt_gpu_lem* lem, int32_t enable)
{
    return guarded([&] {
        if (!lem) return fail_arg("null argument");
        LSDTT_HIP(hipDeviceSynchronize());
        lem->eng->timer.reset();
        lem->eng->timer.enabled = enable != 0;
        return LSDTT_GPU_OK;
    });
}

int32_t lsdtt_gpu_lem_profile_count(lsdtt_gpu_lem* lem)
{
    if (!lem) return 0;
    try {
        LSDTT_HIP(hipDeviceSynchronize());
        lem->eng->timer.collect();
        return static_cast<int32_t>(lem->eng->timer.totals().size());
    } catch (const std::exception& e) {
        set_error(e.what());
        return 0;
    }
}

int lsdtt_gpu_lem_profile_get(lsdtt_gpu_lem* lem, int32_t index, char* name, int32_t name_len, double* ms,
                              int64_t* launches, double* bytes)
{
    return guarded([&] {
        if (!lem) return fail_arg("null argument");
        lem->eng->timer.collect();
        const auto& t = lem->eng->timer.totals();
        if (index < 0 || index >= static_cast<int32_t>(t.size())) return fail_arg("profile index out of range");
        auto it = t.begin();
        std::advance(it, index);
        if (name && name_len > 0) {
            std::strncpy(name, it->first.c_str(), static_cast<size_t>(name_len - 1));
            name[name_len - 1] = '\0';
        }
        if (ms) *ms = it->second.ms;
        if (launches) *launches = it->second.launches;
        if (bytes) *bytes = it->second.bytes;
        return LSDTT_GPU_OK;
    });
}

int lsdtt_gpu_build_boundary_flats(int32_t n_rows, int32_t n_cols, const uint8_t* mask,
                                   const char* boundary_conditions, int64_t* out, int64_t* n_out)
{
    return guarded([&] {
        if (!mask || !boundary_conditions || !out || !n_out) return fail_arg("null argument");
        const char* bc = boundary_conditions;
        const bool fn = bc[0] == 'o' || bc[0] == 'b', fe = bc[1] == 'o' || bc[1] == 'b';
        const bool fs = bc[2] == 'o' || bc[2] == 'b', fw = bc[3] == 'o' || bc[3] == 'b';
        int64_t k = 0;
        for (int r = 0; r < n_rows; ++r)
            for (int c = 0; c < n_cols; ++c) {
                const int64_t f = static_cast<int64_t>(r) * n_cols + c;
                if (!mask[f]) continue;
                bool edge = false;
                for (int dr = -1; dr <= 1 && !edge; ++dr)
                    for (int dc = -1; dc <= 1; ++dc) {
                        if (dr == 0 && dc == 0) continue;
                        const int r2 = r + dr, c2 = c + dc;
                        if (r2 < 0) edge |= fn;
                        else if (r2 >= n_rows) edge |= fs;
                        else if (c2 < 0) edge |= fw;
                        else if (c2 >= n_cols) edge |= fe;
                        else if (!mask[static_cast<int64_t>(r2) * n_cols + c2])
                            edge |= (dr == -1 && fn) || (dr == 1 && fs) || (dc == -1 && fw) || (dc == 1 && fe);
                    }
                if (edge) out[k++] = f;
            }
        *n_out = k;
        return LSDTT_GPU_OK;
    });
}

}  // extern "C"
