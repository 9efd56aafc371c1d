// steady.hpp -- analytic fluvial steady state on the D8 flow object
// (compute_d8_fluvial_steady_state, src/cpu/lsdtt_steady_state_cpu.cpp:
// 76-225) and on the D-infinity graph (compute_dinf_fluvial_steady_state,
// :227-436): chi by a sweep from the outlets, z = baselevel + f chi.
#pragma once

#include "flow.hpp"
#include "fluvial.hpp"

namespace lsdtt_gpu {

// lsdtt_fluvial_steady_state_options (src/cpu/lsdtt_steady_state_cpu.hpp:
// 62-89), float fields as in the reference
struct SteadyOptions {
    float U = 1e-4f, K = 1e-4f, m = 0.45f, n = 1.0f;
    bool relief_mode = false;
    float target_relief = 1000.0f;
    float baselevel_elevation = 0.0f;
    bool divide_area_by_dx = true;
    float channel_threshold_area = 0.0f;
};

struct SteadyResult {
    float K_used = 0, max_unit_height = 0, computed_relief = 0, erodibility_field_scale = 1.0f;
    float max_unit_height_at_channel_head = 0;
    long long n_channel_pixels = 0;
    bool chi_c_reported = false;
};

// make_chi_weights: w = powf(Kref / K_i, 1/n) on valid cells (device arrays)
void chi_weights(const uint8_t* mask, const float* K, long long n, float Kref, float nexp, float* w, hipStream_t s);

// On a routed flow object: area (accumulate, the reference's order), chi,
// elevations (nodata off the valid cells) into z_out (device, n floats).
// w: per-cell chi weights (K / K_i)^(1/n) on the device, or null.
SteadyResult steady_state_d8(FlowEngine& e, const SteadyOptions& o, const float* w, float nodata, float* z_out,
                             hipStream_t s);

// On a D-infinity graph (rec / wr: two receivers and weights per cell, -1 =
// none) of the valid cells `valid` with its accumulated area (m^2 per
// cell): chi by the tile passes of t (sized to the grid and its boundary
// conditions; its order buffers are overwritten), then as steady_state_d8.
SteadyResult steady_state_dinf(FluvialTiles& t, const int* rec, const float* wr, const uint8_t* valid,
                               const double* area, const SteadyOptions& o, const float* w, float dx, float nodata,
                               float* z_out, hipStream_t s);

}  // namespace lsdtt_gpu
