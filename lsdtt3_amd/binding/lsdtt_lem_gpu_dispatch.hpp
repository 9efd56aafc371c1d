// lsdtt_lem_gpu_dispatch.hpp -- the reference-side binding of
// liblsdtt3_gpu.so for lsdtt::run_lem_step (src/cpu/lsdtt_lem_cpu.hpp:213-232),
// written against the reference's own headers: a maintainer includes it in
// src/cpu/lsdtt_lem_cpu.cpp and calls try_run_lem_step() at the top of
// run_lem_step, in the selective-dispatch style of
// src/gpu/lsdtt_surface_gpu.hpp:19-29 (the launcher returns a status; on any
// failure the CPU body runs unchanged).
//
//   #ifdef LSDTT_GPU_KERNELS
//   #include "lsdtt_lem_gpu_dispatch.hpp"
//   #endif
//   std::pair<lsdtt_raster_cpu, lsdtt_lem_step_stats>
//   run_lem_step(const lsdtt_raster_cpu& dem, const lsdtt_lem_options& opts,
//                const std::vector<std::int64_t>& boundary_flats, ...)
//   {
//   #ifdef LSDTT_GPU_KERNELS
//       static thread_local lsdtt_mi355x::gpu_lem_context gpu_ctx;
//       std::pair<lsdtt_raster_cpu, lsdtt_lem_step_stats> gpu_out;
//       if (lsdtt_mi355x::try_run_lem_step(gpu_ctx, dem, opts, boundary_flats, &gpu_out))
//           return gpu_out;
//   #endif
//       ... CPU body unchanged ...
//   }
//
// The header compiles against the reference tree as it is (a CPU test of
// this repository runs the compiler over it with the reference's include
// directories); it needs no reference change beyond that call.
#pragma once

#include <cstdint>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "lsdtt3_gpu.h"
#include "lsdtt_lem_cpu.hpp"

namespace lsdtt_mi355x {

// The run_lem_step configurations the device path computes; the others stay
// on the CPU: the explicit and cellular hillslope models, TR-BDF2, the D8
// drainage-area smoothing with D-infinity incision.
inline bool gpu_lem_supported(const lsdtt::lsdtt_lem_options& o)
{
    if (o.enable_hillslope_transport && o.hillslope_transport_model != "roering_implicit") return false;
    if (o.enable_hillslope_transport && o.roering_implicit_opts.time_integrator != 0) return false;
    if (o.boundary_conditions.size() != 4) return false;
    return true;
}

// lsdtt_lem_options (src/cpu/lsdtt_lem_cpu.hpp:97-163) -> lsdtt_gpu_lem_options
inline void to_gpu_options(const lsdtt::lsdtt_lem_options& o, lsdtt_gpu_lem_options* g)
{
    lsdtt_gpu_lem_options_init(g);
    g->timestep = o.timestep;
    g->erodibility = o.erodibility;
    g->m_exponent = o.m_exponent;
    g->n_exponent = o.n_exponent;
    g->incision_threshold_area = o.incision_threshold_area;
    g->divide_area_by_dx = o.divide_area_by_dx ? 1 : 0;
    g->slope_clamp = o.slope_clamp;
    g->hillslope_diffusivity = o.hillslope_diffusivity;
    g->enable_nonlinear_hillslopes = o.enable_nonlinear_hillslopes ? 1 : 0;
    g->critical_hillslope_slope = o.critical_hillslope_slope;
    g->uplift_rate = o.uplift_rate;
    g->baselevel_fall_rate = o.baselevel_fall_rate;
    g->enable_d8_fluvial_erosion = o.enable_d8_fluvial_erosion ? 1 : 0;
    g->enable_dinf_fluvial_erosion = o.enable_dinf_fluvial_erosion ? 1 : 0;
    g->enable_hillslope_transport = o.enable_hillslope_transport ? 1 : 0;
    g->enable_uplift = o.enable_uplift ? 1 : 0;
    std::memcpy(g->boundary_conditions, o.boundary_conditions.data(), 4);
    g->drainage_area_kernel_size = o.drainage_area_moving_average_kernel_size_pixels;
    g->multiply_by_w_over_dx = o.multiply_stream_power_term_by_w_over_dx ? 1 : 0;
    g->channel_width_w = o.channel_width_w;
    switch (o.coupling_regime) {
    case lsdtt::hillslope_coupling_regime::coupled: g->hillslope_coupling_regime = 1; break;
    case lsdtt::hillslope_coupling_regime::channels_as_boundaries: g->hillslope_coupling_regime = 2; break;
    default: g->hillslope_coupling_regime = 0; break;
    }
    g->hillslope_force_picard = o.hillslope_force_picard ? 1 : 0;
    const lsdtt::lsdtt_roering_implicit_options& r = o.roering_implicit_opts;
    g->picard_tol = r.picard_tol;
    g->picard_max_iter = r.picard_max_iter;
    g->relaxation_omega = r.relaxation_omega;
    g->picard_slope_ratio_cap = r.slope_ratio_cap;
    g->picard_dt_init = r.dt_init;
    g->picard_dt_min = r.dt_min;
    g->picard_dt_max = r.dt_max;
    g->picard_grow_factor = r.grow_factor;
    g->picard_shrink_factor = r.shrink_factor;
    g->picard_dz_target = r.dz_target;
    g->compute_erosion_rate_field = o.compute_erosion_rate_field ? 1 : 0;
}

// lsdtt_gpu_lem_stats -> lsdtt_lem_step_stats (src/cpu/lsdtt_lem_cpu.hpp:166-211),
// every field; the per-pixel erosion rate is read separately
inline void to_reference_stats(const lsdtt_gpu_lem_stats& st, lsdtt::lsdtt_lem_step_stats* s)
{
    s->delta_max_elevation = static_cast<float>(st.delta_max_elevation);
    s->delta_mean_elevation = static_cast<float>(st.delta_mean_elevation);
    s->max_local_elevation_change = static_cast<float>(st.max_local_elevation_change);
    s->p999_local_elevation_change = static_cast<float>(st.p999_local_elevation_change);
    s->mean_erosion_rate = static_cast<float>(st.mean_erosion_rate);
    s->max_erosion_rate = static_cast<float>(st.max_erosion_rate);
    s->erosion_rate.clear();
}

// One device context per grid shape, mask and option set, kept across steps
// (the CLI time loop calls run_lem_step once per step on the same grid).
struct gpu_lem_context {
    lsdtt_gpu_lem* h = nullptr;
    int nr = 0, nc = 0;
    double cell_size = 0.0;
    std::vector<std::uint8_t> mask;
    bool has_k_field = false, has_channel_mask = false;  // set on the device context
    lsdtt_gpu_lem_options opts;
    gpu_lem_context() { std::memset(&opts, 0, sizeof opts); }
    gpu_lem_context(const gpu_lem_context&) = delete;
    gpu_lem_context& operator=(const gpu_lem_context&) = delete;
    ~gpu_lem_context() { lsdtt_gpu_lem_destroy(h); }
};

// Runs the step on the GPU and returns true, or returns false (nothing
// written) when the configuration is not covered or any call fails.
inline bool try_run_lem_step(gpu_lem_context& ctx, const lsdtt::lsdtt_raster_cpu& dem,
                             const lsdtt::lsdtt_lem_options& opts,
                             const std::vector<std::int64_t>& boundary_flats,
                             std::pair<lsdtt::lsdtt_raster_cpu, lsdtt::lsdtt_lem_step_stats>* out)
{
    if (!gpu_lem_supported(opts)) return false;
    const lsdtt::lsdtt_raster_meta& m = dem.get_meta();
    const std::vector<std::uint8_t>& mask = dem.get_mask();
    lsdtt_gpu_lem_options g;
    to_gpu_options(opts, &g);
    // a new context when the grid, mask or any option but the timestep
    // changed (byte copies keep the zeroed padding of options_init, so the
    // byte comparison sees only the fields)
    lsdtt_gpu_lem_options g_cmp, c_cmp;
    std::memcpy(&g_cmp, &g, sizeof g);
    std::memcpy(&c_cmp, &ctx.opts, sizeof g);
    g_cmp.timestep = c_cmp.timestep = 0.0;
    const bool k_field = !opts.erodibility_field.empty(), channels = !opts.explicit_channel_mask.empty();
    const bool same = ctx.h && ctx.nr == m.n_rows && ctx.nc == m.n_cols && ctx.cell_size == m.cell_size &&
                      ctx.mask == mask && std::memcmp(&g_cmp, &c_cmp, sizeof g_cmp) == 0 &&
                      ctx.has_k_field == k_field && ctx.has_channel_mask == channels;
    if (!same) {
        lsdtt_gpu_lem_destroy(ctx.h);
        ctx.h = lsdtt_gpu_lem_create(m.n_rows, m.n_cols, m.cell_size, mask.data(), &g);
        if (!ctx.h) return false;
        ctx.nr = m.n_rows;
        ctx.nc = m.n_cols;
        ctx.cell_size = m.cell_size;
        ctx.mask = mask;
        ctx.has_k_field = k_field;
        ctx.has_channel_mask = channels;
        std::memcpy(&ctx.opts, &g, sizeof g);
    } else if (ctx.opts.timestep != g.timestep) {
        if (lsdtt_gpu_lem_set_timestep(ctx.h, g.timestep) != LSDTT_GPU_OK) return false;
        ctx.opts.timestep = g.timestep;
    }
    const std::vector<float>& zf = dem.get_data();
    std::vector<double> z(zf.begin(), zf.end());
    if (lsdtt_gpu_lem_set_elevation(ctx.h, z.data()) != LSDTT_GPU_OK) return false;
    if (k_field) {
        std::vector<double> K(opts.erodibility_field.begin(), opts.erodibility_field.end());
        if (lsdtt_gpu_lem_set_erodibility_field(ctx.h, K.data()) != LSDTT_GPU_OK) return false;
    }
    if (channels && lsdtt_gpu_lem_set_channel_mask(ctx.h, opts.explicit_channel_mask.data()) != LSDTT_GPU_OK)
        return false;
    lsdtt_gpu_lem_stats st;
    if (lsdtt_gpu_lem_step(ctx.h, boundary_flats.data(), static_cast<int64_t>(boundary_flats.size()), &st) !=
        LSDTT_GPU_OK)
        return false;
    if (lsdtt_gpu_lem_get_elevation(ctx.h, z.data()) != LSDTT_GPU_OK) return false;
    lsdtt::lsdtt_raster_cpu result = dem;
    std::vector<float>& zo = result.get_data();
    for (std::size_t i = 0; i < z.size(); ++i) zo[i] = static_cast<float>(z[i]);
    lsdtt::lsdtt_lem_step_stats s{};
    to_reference_stats(st, &s);
    if (opts.compute_erosion_rate_field) {
        s.erosion_rate.resize(z.size());
        if (lsdtt_gpu_lem_get_erosion_rate(ctx.h, s.erosion_rate.data()) != LSDTT_GPU_OK) return false;
    }
    *out = {std::move(result), std::move(s)};
    return true;
}

}  // namespace lsdtt_mi355x
