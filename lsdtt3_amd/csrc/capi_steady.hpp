// capi_steady.hpp -- the steady-state options, erodibility checks and result
// of the C ABI (lsdtt_gpu_steady_state_*), shared by the D8 flow object's
// and the D-infinity graph's entry points.
#pragma once

#include <cmath>
#include <string>

#include "../../include/lsdtt3_gpu.h"
#include "capi_util.hpp"
#include "steady.hpp"

namespace lsdtt_gpu {

// the reference's options are float (lsdtt_steady_state_cpu.hpp:62-89)
inline SteadyOptions steady_options(const lsdtt_gpu_steady_state_options* o)
{
    SteadyOptions so;
    so.U = static_cast<float>(o->U);
    so.K = static_cast<float>(o->K);
    so.m = static_cast<float>(o->m);
    so.n = static_cast<float>(o->n);
    so.relief_mode = o->relief_mode != 0;
    so.target_relief = static_cast<float>(o->target_relief);
    so.baselevel_elevation = static_cast<float>(o->baselevel_elevation);
    so.divide_area_by_dx = o->divide_area_by_dx != 0;
    so.channel_threshold_area = static_cast<float>(o->channel_threshold_area);
    return so;
}

// make_chi_weights' checks (lsdtt_steady_state_cpu.cpp:44-66) on the host
// copy of the valid cells; empty string when the field is usable
inline std::string erodibility_field_error(const float* K, const uint8_t* valid, long long n, int n_cols, float Kref)
{
    if (!(Kref > 0.0f) || !std::isfinite(Kref)) return "erodibility_field: the reference K must be > 0";
    for (long long i = 0; i < n; ++i)
        if (valid[i] && !(K[i] > 0.0f && std::isfinite(K[i])))
            return "erodibility_field: K must be > 0; pixel (" + std::to_string(i / n_cols) + ", " +
                   std::to_string(i % n_cols) + ") has K = " + std::to_string(K[i]);
    return std::string();
}

inline void copy_steady_result(const SteadyResult& r, lsdtt_gpu_steady_state_result* res)
{
    if (!res) return;
    res->K_used = r.K_used;
    res->max_unit_height = r.max_unit_height;
    res->computed_relief = r.computed_relief;
    res->erodibility_field_scale = r.erodibility_field_scale;
    res->max_unit_height_at_channel_head = r.max_unit_height_at_channel_head;
    res->n_channel_pixels = r.n_channel_pixels;
    res->chi_c_reported = r.chi_c_reported ? 1 : 0;
}

}  // namespace lsdtt_gpu
