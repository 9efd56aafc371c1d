// dinf.hpp -- D-infinity receivers and weights (Tarboton 1997) of a surface,
// restating build_dinf (src/cpu/lsdtt_flowgraph_cpu.cpp:423-663).
#pragma once

#include "common.hpp"

namespace lsdtt_gpu {

struct DinfOptions {
    double dx = 1.0;
    BC bc{'o', 'o', 'o', 'o'};
    bool interior_nodata_is_outlet = true;  // cells touching an invalid in-grid neighbour route nowhere
    bool use_nodata = false;                // a cell equal to `nodata` is invalid (besides the mask)
    float nodata = -9999.0f;
};

// Per cell f: rec[2f], rec[2f+1] (-1 = none) and weights w[2f], w[2f+1] in
// float; slope and bearing (degrees clockwise from north) may be null and
// are `nodata` on cells without receivers.  T is float or double; a double
// surface is rounded to float first (the reference's raster type).
template <class T>
void dinf_receivers(const T* z, const uint8_t* mask, int nr, int nc, const DinfOptions& o, int* rec, float* w,
                    float* slope, float* bearing, hipStream_t s);

}  // namespace lsdtt_gpu
