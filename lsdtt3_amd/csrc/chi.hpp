// chi.hpp -- chi on the device flow object.
//
// Reference: src/cpu/lsdtt_chi_extraction_cpu.cpp:154-179 (increment and
// step), :283-497 (full and thresholded basin recursion), :739-765
// (get_chi_raster, A = s_vector_size * dx^2).  chi(v) = float(double(chi(r))
// + step * 0.5 * (f(size v) + f(size r))) with f(k) = pow(A0 / (k dx^2),
// theta) depends only on the receiver's final value, so the depth-from-outlet
// levels of the flow object give the reference's values exactly.  f is
// evaluated on the host with the C library's pow for every count present in
// the grid (a compact table indexed by the count's rank among the counts
// present), which keeps chi bit-identical to a host evaluation with the same
// libm.
#pragma once

#include <vector>

#include "flow.hpp"

namespace lsdtt_gpu {

class ChiEngine {
public:
    // chi for every valid cell (outlets 0), nodata elsewhere; requires
    // FlowEngine::route and upstream
    void raster(FlowEngine& e, double m_over_n, double A0, float nodata, float* chi, hipStream_t s);

    // pixels with count >= min_pixels in s-vector order (requires
    // FlowEngine::s_vector and a chi raster): device flats, chi and
    // receivers; returns the number selected
    long long threshold(FlowEngine& e, const float* chi, long long min_pixels, hipStream_t s);

    dbuf<unsigned> bits;
    dbuf<int> wpc, wrank, present;
    dbuf<double> table;
    dbuf<int> flag, pos, sel_flat, sel_rcv;
    dbuf<float> sel_chi;
    dscratch tmp;
    hbuf<int> hread, kh;   // pinned: the number of counts present, the counts
    hbuf<double> table_h;  // pinned: f of each count present

private:
    static long long n_present_cap(long long n);
};

}  // namespace lsdtt_gpu
