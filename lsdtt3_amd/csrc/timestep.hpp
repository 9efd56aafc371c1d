// timestep.hpp -- LEM timestep control: the median interior gradient on the
// device, and the reference's host-side controller and convergence monitor.
//
// Reference: src/cpu/lsdtt_lem_timestep_cpu.hpp:62-292 (options, classes),
// src/cpu/lsdtt_lem_timestep_cpu.cpp:28-302 (median_interior_gradient,
// diagnose_lem_timestep, lsdtt_lem_timestep_controller::update,
// lsdtt_lem_convergence_monitor::update).  The controller and the monitor
// are a few scalar decisions per step (host code, as in the reference); the
// gradient median is a pass over the grid (device).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "common.hpp"
#include "dbuf.hpp"

namespace lsdtt_gpu {

// Median over the valid interior cells whose four cardinal neighbours are
// valid of sqrt(gx^2 + gy^2), gx = (z_e - z_w) / (2 dx), gy = (z_s - z_n) /
// (2 dx); 0 when there is none.  Exact: two radix selections of the FP64
// values' bit patterns.  keys: scratch of n 64-bit words (grown as needed).
double median_interior_gradient(const double* z, const uint8_t* mask, int nr, int nc, double dx,
                                dbuf<unsigned long long>& keys, hipStream_t s);

struct TimestepOptions {
    double target_ratio = 0.7;
    double dt_min_yr = 0.0;  // 0: dt_initial / 100
    double dt_max_yr = 0.0;  // 0: dt_initial * 100
    double grow_factor = 1.15;
    double shrink_factor = 0.5;
    int slope_refresh_interval = 25;
    double ladder_ratio = 2.0;  // 1: no ladder
    double ratio_smoothing = 0.15;
};

struct TimestepDiagnosis {
    double reference_slope = 0, cell_relief_m = 0, uplift_per_step_m = 0, ratio = 0, suggested_dt_yr = 0;
    bool too_large = false;
    std::string warning;
};

TimestepDiagnosis diagnose_timestep(double reference_slope, double dx, double uplift_rate, double timestep_yr,
                                    double target_ratio);

class TimestepController {
public:
    TimestepController(double dt_initial, const TimestepOptions& o);
    // the step just taken moved the surface by at most max_dz; slope() is
    // asked for the current median gradient on the refresh interval only
    template <class SlopeFn>
    double update(double max_dz, double dx, SlopeFn&& slope_now)
    {
        if (n_steps_ % (o_.slope_refresh_interval > 1 ? o_.slope_refresh_interval : 1) == 0) slope_ = slope_now();
        ++n_steps_;
        return decide(max_dz, dx);
    }
    double dt() const { return dt_; }
    double last_ratio() const { return last_ratio_; }
    double reference_slope() const { return slope_; }
    long n_shrinks() const { return n_shrinks_; }
    long n_grows() const { return n_grows_; }
    long n_changes() const { return n_changes_; }

private:
    double decide(double max_dz, double dx);
    TimestepOptions o_;
    double dt_, dt_raw_, dt_initial_, dt_min_, dt_max_;
    double slope_ = 0.0, last_ratio_ = 0.0, ratio_ema_ = 0.0;
    long n_steps_ = 0, n_shrinks_ = 0, n_grows_ = 0, n_changes_ = 0;
    int rung_ = 0, lowest_rung_ = 0;
};

struct ConvergenceOptions {
    double dz_tolerance_m = 1.0e-4;
    double dz_rate_tolerance_m_per_yr = 0.0;
    int patience_steps = 5;
    double min_dt_fraction = 0.25;
    int stall_window = 1000;
    double stall_decay_threshold = 0.99;
    double oscillation_alternation_fraction = 0.70;
};

class ConvergenceMonitor {
public:
    enum State { Running = 0, Converged = 1, Stalled = 2 };
    ConvergenceMonitor(double dt_initial, const ConvergenceOptions& o);
    State update(double max_dz, double dt_used);
    State state() const { return state_; }
    bool oscillating() const { return oscillating_; }
    long n_steps() const { return n_steps_; }
    double last_dz() const { return last_dz_; }
    double window_decay() const { return decay_; }
    const std::string& message() const { return message_; }

private:
    ConvergenceOptions o_;
    double dt_initial_, last_dz_ = 0.0, decay_ = 1.0;
    long n_steps_ = 0;
    int n_below_ = 0;
    bool oscillating_ = false;
    State state_ = Running;
    std::string message_;
    std::vector<double> hist_;  // the last max(8, stall_window) values, oldest first
    size_t window_;
};

}  // namespace lsdtt_gpu
