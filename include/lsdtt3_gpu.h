/*
 * lsdtt3_gpu.h -- C ABI of the MI355X LEM-timestep library (liblsdtt3_gpu.so).
 *
 * Plain pointers and sizes only.  Layout conventions are the reference's:
 * row-major flat index row*n_cols+col (int64 at this boundary), receiver -1 =
 * nodata and -2 = outlet, D8 direction codes 0..7 clockwise from N with 255 =
 * none, boundary string N,E,S,W of 'o' (open), 'n' (no flux), 'p'
 * (periodic), 'b' (base level) -- src/core/lsdtt_types.hpp:142-144,
 * src/core/lsdtt_flow_base.hpp:79-111 and :265-271.
 *
 * Error behaviour follows the reference's GPU dispatch rule
 * (docs/developer_notes/gpu_compute.md:51-62, src/gpu/lsdtt_surface_gpu.hpp:
 * 19-29): every call returns a status; on a non-zero status the outputs are
 * unspecified and the caller falls back to its CPU path.  No C++ exception
 * crosses this boundary; lsdtt_gpu_last_error() gives the message.
 *
 * Host-pointer entry points copy in and out.  *_device entry points take
 * device pointers and a hipStream_t (as void*) and leave results resident.
 */
#ifndef LSDTT3_GPU_H
#define LSDTT3_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LSDTT_GPU_OK 0
#define LSDTT_GPU_ERR_ARG 1          /* std::invalid_argument in the reference */
#define LSDTT_GPU_ERR_RUNTIME 2      /* HIP or allocation failure */
#define LSDTT_GPU_ERR_UNRESOLVED 3   /* "could not assign a receiver" */
#define LSDTT_GPU_ERR_NO_DEVICE 4
#define LSDTT_GPU_ERR_NOT_CONVERGED 5

/* 1 when a usable GPU is present.  Replaces lsdtt::gpu_available()
 * (src/core/lsdtt_backend_factory.cpp:33). */
int lsdtt_gpu_available(void);

/* Message of the last failing call on this thread ("" if none). */
const char* lsdtt_gpu_last_error(void);

/* Device memory for the *_device entry points (no counterpart in the
 * reference, whose arrays live on the host): plain hipMalloc / hipFree /
 * hipMemcpy on the current device.  kind: 1 host to device, 2 device to
 * host, 3 device to device.  _synchronize waits for all work on the device. */
void* lsdtt_gpu_device_alloc(int64_t bytes);
void lsdtt_gpu_device_free(void* p);
int lsdtt_gpu_memcpy(void* dst, const void* src, int64_t bytes, int32_t kind);
int lsdtt_gpu_synchronize(void);

/* ---------------------------------------------------------------------
 * D8 flow object.  Replaces constructing lsdtt_flowinfo
 * (src/cpu/lsdtt_flowinfo_cpu.cpp:183-253, options
 * src/core/lsdtt_flow_base.hpp:79-111) and its array queries
 * (receiver_of_unchecked, receiver_dir_of_unchecked,
 * donor_count_of_unchecked, outlets, topological_order, s-vector and
 * s_vector_start/size_unchecked, accumulate, get_drainage_area_raster;
 * src/core/lsdtt_flow_base.hpp:126-491).
 * ------------------------------------------------------------------- */
typedef struct lsdtt_gpu_flow lsdtt_gpu_flow;

typedef struct {
    int32_t n_rows;
    int32_t n_cols;
    double cell_size;
    char boundary_conditions[4]; /* N E S W */
    int32_t interior_nodata_is_outlet;  /* default 1 */
    int32_t allow_pits_as_outlets;      /* default 0; the LEM uses 1 */
    int32_t build_topological_order;    /* default 1 */
    int32_t build_s_vector;             /* default 1 */
} lsdtt_gpu_flow_options;

/* Sets the reference's defaults (oooo, 1, 0, 1, 1). */
void lsdtt_gpu_flow_options_init(lsdtt_gpu_flow_options* o, int32_t n_rows, int32_t n_cols,
                                 double cell_size);

lsdtt_gpu_flow* lsdtt_gpu_flow_create(const lsdtt_gpu_flow_options* o);
void lsdtt_gpu_flow_destroy(lsdtt_gpu_flow* f);

/* Build from host arrays.  z is float32 (the reference's raster type) or
 * float64 (the LEM's state); mask is uint8 (0 = nodata); forced outlets are
 * flat indices (may be NULL / 0).  Out-of-range or masked forced indices
 * give LSDTT_GPU_ERR_ARG, an unroutable cell LSDTT_GPU_ERR_UNRESOLVED. */
int lsdtt_gpu_flow_build_f32(lsdtt_gpu_flow* f, const float* z, const uint8_t* mask,
                             const int64_t* forced_outlets, int64_t n_forced);
int lsdtt_gpu_flow_build_f64(lsdtt_gpu_flow* f, const double* z, const uint8_t* mask,
                             const int64_t* forced_outlets, int64_t n_forced);

/* Build from device arrays already resident in HBM (no host copies); the
 * work is enqueued on `stream` (a hipStream_t, NULL = default stream). */
int lsdtt_gpu_flow_build_device(lsdtt_gpu_flow* f, const void* d_z, int32_t z_is_f64,
                                const uint8_t* d_mask, void* stream);

/* Sizes of the built object. */
int lsdtt_gpu_flow_sizes(const lsdtt_gpu_flow* f, int64_t* n_valid, int64_t* n_outlets,
                         int64_t* max_depth);

/* Copy results to host arrays (any pointer may be NULL to skip it):
 * receiver/start int64[n], dir/donors uint8[n], size uint64[n],
 * outlets int64[n_outlets], topo and s_vector int64[n_valid]. */
int lsdtt_gpu_flow_get(lsdtt_gpu_flow* f, int64_t* receiver, uint8_t* dir, uint8_t* donors,
                       int64_t* outlets, int64_t* topo, int64_t* s_vector, int64_t* s_start,
                       uint64_t* s_size);

/* Drainage area raster: float(n_contributing * cell_size^2) on valid cells,
 * nodata elsewhere (get_drainage_area_raster, lsdtt_flowinfo_cpu.cpp:
 * 2095-2150); int64 upstream counts may be requested as well (NULL skips). */
int lsdtt_gpu_flow_area(lsdtt_gpu_flow* f, double nodata, float* area, uint8_t* area_mask,
                        int64_t* counts);

/* accumulate(weights) in double, summed in topological order exactly as
 * lsdtt_flowinfo::accumulate (lsdtt_flowinfo_cpu.cpp:1979-2046); weights
 * NULL means cell_size^2 per valid cell. Output double[n]. */
int lsdtt_gpu_flow_accumulate(lsdtt_gpu_flow* f, const double* weights, double* accum);

/* Analytic fluvial steady state on the flow object:
 * compute_d8_fluvial_steady_state(flow, opts)
 * (src/cpu/lsdtt_steady_state_cpu.cpp:76-225; options :62-89, float in the
 * reference).  chi = chi(receiver) + dx_eff * powf(A, -m/n) * w (float,
 * outlets 0; A = accumulate() / dx when divide_area_by_dx); z = baselevel +
 * (U/K)^(1/n) chi, or in relief mode target_relief * chi / max(chi) with
 * K_used back-calculated; z_out float[n] (nodata off the valid cells).
 * erodibility_field (float[n], may be NULL) gives the K-weighted chi.  powf
 * is the correctly rounded value (the reference's C library returns it or
 * one ulp beside it).  For lsdtt-lem's snap
 * (snap_lem_surface_to_fluvial_steady_state, lsdtt_lem_cpu.cpp:813-845):
 * lsdtt_gpu_fill, then a flow object on the filled surface, then this. */
typedef struct {
    double U, K, m, n;           /* 1e-4, 1e-4, 0.45, 1 */
    int32_t relief_mode;         /* 0 */
    double target_relief;        /* 1000 m */
    double baselevel_elevation;  /* 0 */
    int32_t divide_area_by_dx;   /* 1 */
    double channel_threshold_area;  /* 0: chi_c not reported */
} lsdtt_gpu_steady_state_options;

typedef struct {
    double K_used, max_unit_height, computed_relief, erodibility_field_scale;
    double max_unit_height_at_channel_head;
    int64_t n_channel_pixels;
    int32_t chi_c_reported;
} lsdtt_gpu_steady_state_result;

void lsdtt_gpu_steady_state_options_init(lsdtt_gpu_steady_state_options* o);
int lsdtt_gpu_flow_steady_state(lsdtt_gpu_flow* f, const lsdtt_gpu_steady_state_options* o,
                                const float* erodibility_field, double nodata, float* z_out,
                                lsdtt_gpu_steady_state_result* result);

/* Chi raster: get_chi_raster(flow, {m_over_n, A_0})
 * (src/cpu/lsdtt_chi_extraction_cpu.cpp:739-765; recursion :283-365,
 * increment :154-179): every outlet starts at 0, A = contributing pixels x
 * cell_size^2, chi float on valid cells and `nodata` elsewhere.  Needs a flow
 * object built with the s-vector (LSDTT_GPU_ERR_RUNTIME otherwise, the
 * reference's runtime_error); non-finite m_over_n or A_0 <= 0 give
 * LSDTT_GPU_ERR_ARG.  chi float[n], chi_mask uint8[n] (either may be NULL). */
int lsdtt_gpu_flow_chi(lsdtt_gpu_flow* f, double m_over_n, double A_0, double nodata, float* chi,
                       uint8_t* chi_mask);

/* Same raster into a device buffer (float[n]) on `stream`. */
int lsdtt_gpu_flow_chi_device(lsdtt_gpu_flow* f, double m_over_n, double A_0, double nodata, float* d_chi,
                              void* stream);

/* Channel chi of every basin large enough: get_all_basin_outlet_flats(flow,
 * min_pixels) (src/cpu/lsdtt_flow_base.cpp:82-110) followed by
 * get_chi_from_outlets_zero_base_threshold(flow, outlets, min_pixels)
 * (lsdtt_chi_extraction_cpu.cpp:368-497, 954-990), concatenated: for each
 * qualifying outlet in row-major order, the pixels of its s-vector slice
 * with at least min_pixels contributing pixels, in s-vector order.
 * receivers[i] is the pixel's receiver (-2 for the basin outlet);
 * basin_start[b] is the index of basin b's outlet.  Pass NULL arrays to
 * obtain *n_pixels and *n_basins first (arrays sized n_valid suffice). */
int lsdtt_gpu_flow_chi_threshold(lsdtt_gpu_flow* f, double m_over_n, double A_0, int64_t min_pixels,
                                 int64_t* n_pixels, int64_t* n_basins, int64_t* flats, float* chi,
                                 int64_t* receivers, int64_t* basin_start);

/* ---------------------------------------------------------------------
 * Epsilon-free depression fill.  Replaces lsdtt::fill_pits(raster,
 * min_slope <= 0, boundary_conditions) (src/cpu/
 * lsdtt_raster_flow_preprocessing_cpu.cpp:707-782 -> reconstruction
 * :1203-1394 / priority flood :893-920).  valid = mask != 0 && z != nodata
 * (NaN-aware); invalid cells are written as nodata.  Periodic pairs must
 * match and may not both be periodic (LSDTT_GPU_ERR_ARG); grids of 2^29
 * cells or more are rejected (LSDTT_GPU_ERR_ARG).
 * lsdtt_gpu_fill_device works on device arrays on the caller's stream (a
 * hipStream_t, NULL = the null stream) and returns when the fill has
 * finished there.  Its working arrays (44 bytes per cell) live in one
 * library workspace: calls from several host threads or streams are safe
 * and run one at a time.
 * ------------------------------------------------------------------- */
int lsdtt_gpu_fill(const float* z, const uint8_t* mask, int32_t n_rows, int32_t n_cols,
                   double nodata, const char* boundary_conditions, float* filled);
int lsdtt_gpu_fill_device(const float* d_z, const uint8_t* d_mask, int32_t n_rows,
                          int32_t n_cols, double nodata, const char* boundary_conditions,
                          float* d_filled, void* stream);
/* No counterpart in the reference: lsdtt_gpu_fill_device for callers that
 * cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc).  Same arguments and results
 * (bit for bit); the working arrays live in a fixed device array of the
 * library (164 MiB per GPU, held until the process ends) instead of an
 * allocated workspace, so grids are limited to
 * lsdtt_gpu_fill_noalloc_max_cells() = 4194304 cells (2048 x 2048, any
 * shape); larger grids are rejected with LSDTT_GPU_ERR_ARG.  Calls share
 * the one array per GPU: within a process the library runs them one at a
 * time (each returns when its fill has finished on its stream); calls that
 * overlapped on one GPU would race on it. */
int lsdtt_gpu_fill_device_noalloc(const float* d_z, const uint8_t* d_mask, int32_t n_rows,
                                  int32_t n_cols, double nodata,
                                  const char* boundary_conditions, float* d_filled,
                                  void* stream);
int64_t lsdtt_gpu_fill_noalloc_max_cells(void);

/* ---------------------------------------------------------------------
 * LEM timestep.  Replaces lsdtt::run_lem_step (src/cpu/lsdtt_lem_cpu.hpp:
 * 220-232, .cpp:285-732) for the D8 fluvial + roering_implicit hillslope
 * configuration, with a device-resident FP64 elevation state.  Options
 * mirror lsdtt_lem_options (lsdtt_lem_cpu.hpp:97-182) and the Newton
 * options (lsdtt_hillslope_newton_cpu.hpp:143-193); statistics mirror
 * lsdtt_lem_step_stats (lsdtt_lem_cpu.hpp:185-211).
 * ------------------------------------------------------------------- */
typedef struct lsdtt_gpu_lem lsdtt_gpu_lem;

typedef struct {
    double timestep;                 /* dt (yr) */
    double erodibility;              /* K */
    double m_exponent, n_exponent;
    double incision_threshold_area;  /* on area/dx; 0 = every cell */
    int32_t divide_area_by_dx;       /* default 1 */
    double slope_clamp;              /* default 1e-5 */
    double hillslope_diffusivity;    /* D (m^2/yr) */
    int32_t enable_nonlinear_hillslopes;
    double critical_hillslope_slope; /* Sc */
    double uplift_rate;              /* U (m/yr) */
    double baselevel_fall_rate;
    int32_t enable_d8_fluvial_erosion;
    int32_t enable_hillslope_transport;  /* roering_implicit (Newton) */
    int32_t enable_uplift;
    char boundary_conditions[4];
    /* Newton (lsdtt_hillslope_newton_options) and its linear solves */
    double newton_abs_tol;           /* 1e-12 m/yr: the reference criterion */
    double newton_rel_tol;           /* on ||F(z^n)||inf, default 1e-8 */
    int32_t newton_max_iterations;   /* 50 */
    int32_t newton_max_line_search;  /* 30 */
    double slope_ratio_cap;          /* 0.9999 */
    double krylov_rel_tol;           /* BiCGSTAB (FP32), default 1e-4 */
    int32_t krylov_max_iterations;   /* 500 */
    /* drainage_area_moving_average_kernel_size_pixels: odd box average of
     * the drainage area before stream power (1 = none) */
    int32_t drainage_area_kernel_size;
    /* multiply_stream_power_term_by_w_over_dx, channel_width_w */
    int32_t multiply_by_w_over_dx;
    double channel_width_w;
    /* hillslope_coupling_regime: 0 automatic, 1 coupled, 2
     * channels_as_boundaries (lsdtt_lem_cpu.cpp:64-77).  The Dirichlet
     * channel set is the mask given to lsdtt_gpu_lem_set_channel_mask, or
     * else, each step, the cells whose D-infinity drainage area of the
     * post-incision surface (rounded to float, interior nodata as outlets,
     * lsdtt_lem_cpu.cpp:137-154) passes incision_threshold_area
     * (lsdtt_gpu_lem_get_channel_mask returns it). */
    int32_t hillslope_coupling_regime;
    /* hillslope_force_picard: skip Newton; otherwise Picard runs only when
     * Newton fails.  roering_implicit_opts of the damped Picard fallback
     * (lsdtt_hillslope_implicit_cpu.hpp:65-97; D, Sc and the law come from
     * the fields above, U is 0): */
    int32_t hillslope_force_picard;
    double picard_tol;               /* 1e-4 m */
    int32_t picard_max_iter;         /* 50 */
    double relaxation_omega;         /* 1 */
    double picard_slope_ratio_cap;   /* r_max, 0.99 */
    double picard_dt_init, picard_dt_min, picard_dt_max;  /* 0 = auto */
    double picard_grow_factor;       /* 1.5 */
    double picard_shrink_factor;     /* 0.5 */
    double picard_dz_target;         /* 0 = off */
    /* enable_dinf_fluvial_erosion (lsdtt_lem_options, lsdtt_lem_cpu.hpp:172;
     * run_lem_step :357-391): D-infinity incision
     * (compute_dinf_stream_power_incision, lsdtt_stream_power_cpu.cpp:
     * 329-602) when D8 incision is off; default 0 */
    int32_t enable_dinf_fluvial_erosion;
    /* compute_erosion_rate_field (lsdtt_lem_options, lsdtt_lem_cpu.hpp:170;
     * filled at :666-696): keep each step's per-pixel erosion rate for
     * lsdtt_gpu_lem_get_erosion_rate; default 0 */
    int32_t compute_erosion_rate_field;
    /* Krylov forcing of the Newton solve (no reference counterpart: the
     * reference factorises exactly).  0: every linear solve to
     * krylov_rel_tol; 1 (default): krylov_rel_tol for the first Newton
     * iteration, then Eisenstat-Walker choice 2 (gamma 0.9, exponent 2,
     * with its safeguard), never tighter than needed to reach the Newton
     * tolerance from the current residual, within [1e-6, 0.1].  The Newton
     * stop is unchanged, so either setting meets the same tolerance. */
    int32_t krylov_forcing;
} lsdtt_gpu_lem_options;

typedef struct {
    double delta_max_elevation;
    double delta_mean_elevation;
    double max_local_elevation_change;
    double p999_local_elevation_change;
    double mean_erosion_rate;
    double max_erosion_rate;
    int64_t n_valid;
    int64_t cells_eroded;             /* stream-power cells updated */
    int32_t newton_iterations;
    int32_t krylov_iterations;
    int32_t hillslope_substeps;
    int32_t hillslope_rejected;
    int32_t hillslope_converged;
    double hillslope_initial_residual; /* ||F(z^n)||inf, m/yr */
    double hillslope_residual;         /* final ||F||inf, m/yr */
    double hillslope_tolerance;
    double max_slope_ratio;
    int32_t count_passes;             /* tile passes of the upstream counts */
    int32_t stream_power_passes;      /* tile passes of the stream power */
    int64_t count_tile_runs;          /* 32 x 32 tiles processed by those passes */
    int64_t stream_power_tile_runs;
    int32_t hillslope_method;         /* 0 Newton, 1 damped Picard */
    int32_t picard_iterations;
    int32_t picard_substeps;
    int32_t picard_rejected;
    int32_t picard_cg_iterations;
    int32_t picard_limiter_fired;
    double picard_max_slope_ratio;
    /* warm start of the first Newton solve (lsdtt_gpu_lem_set_warm_start;
     * no counterpart in the reference): mode used, whether the predicted
     * start was kept (it must lower ||F||inf below ||F(z^n)||inf, and keep
     * the slope ratio feasible), and ||F||inf at the predicted start */
    int32_t warm_start_mode;
    int32_t warm_start_accepted;
    double warm_start_residual;
    /* Newton iterations whose line search ran in restoration mode: the
     * iterate's largest centred slope ratio above sqrt(slope_ratio_cap),
     * so a trial is accepted when it lowers that ratio
     * (lsdtt_hillslope_newton_cpu.cpp:553-633) */
    int32_t newton_restoration_iterations;
} lsdtt_gpu_lem_stats;

/* Defaults of lsdtt_lem_options, plus the Newton defaults above. */
void lsdtt_gpu_lem_options_init(lsdtt_gpu_lem_options* o);

lsdtt_gpu_lem* lsdtt_gpu_lem_create(int32_t n_rows, int32_t n_cols, double cell_size,
                                    const uint8_t* mask, const lsdtt_gpu_lem_options* o);
void lsdtt_gpu_lem_destroy(lsdtt_gpu_lem* lem);

/* State transfer (host arrays of n_rows*n_cols). */
int lsdtt_gpu_lem_set_elevation(lsdtt_gpu_lem* lem, const double* z);
int lsdtt_gpu_lem_get_elevation(lsdtt_gpu_lem* lem, double* z);
/* Per-pixel erodibility (check_lem_erodibility_field: finite, >= 0);
 * NULL returns to the uniform K. */
int lsdtt_gpu_lem_set_erodibility_field(lsdtt_gpu_lem* lem, const double* K);
/* explicit_channel_mask (lsdtt_lem_options, lsdtt_lem_cpu.cpp:93-103):
 * cells != 0 are fixed-elevation channel pixels for the hillslope step
 * (not unknowns), unless the regime is coupled.  NULL clears the mask. */
int lsdtt_gpu_lem_set_channel_mask(lsdtt_gpu_lem* lem, const uint8_t* channel);
/* Device address of the resident FP64 elevation (for device-side I/O).
 * The address is fixed for the life of the context: every step writes its
 * result there, so a surface written through it before a step is the
 * surface that step reads.  Fetching it counts as setting the surface. */
double* lsdtt_gpu_lem_elevation_device(lsdtt_gpu_lem* lem);

/* One step; boundary_flats as from build_boundary_flats (uploaded when it
 * changes).  When the hillslope Newton fails at its minimum sub-step the
 * step falls back to the damped Picard solver, as the reference does
 * (stats.hillslope_method = 1). */
int lsdtt_gpu_lem_step(lsdtt_gpu_lem* lem, const int64_t* boundary_flats, int64_t n_boundary,
                       lsdtt_gpu_lem_stats* stats);

/* lsdtt_lem_step_stats::erosion_rate (lsdtt_lem_cpu.hpp:205-210, filled at
 * lsdtt_lem_cpu.cpp:666-696) of the last step, float[n_rows*n_cols]: per
 * valid pixel (z_before - (z_after - U dt)) / dt in m/yr (positive =
 * erosion, the uplift backed out, deposition negative), 0 off the mask.
 * Needs compute_erosion_rate_field and at least one step. */
int lsdtt_gpu_lem_get_erosion_rate(lsdtt_gpu_lem* lem, float* rate);

/* The hillslope's Dirichlet channel set of the last step (uint8[n]): the
 * explicit mask, or the cells whose D-infinity drainage area passed
 * incision_threshold_area (build_channel_mask, lsdtt_lem_cpu.cpp:86-134,
 * on the post-incision surface), or none in the coupled regime. */
int lsdtt_gpu_lem_get_channel_mask(lsdtt_gpu_lem* lem, uint8_t* channel);

/* Tile shape of the upstream-count and stream-power sweeps: tiles of 32
 * columns by tile_height rows (8, 16 or 32; default 16).  No counterpart in
 * the reference, whose serial sweeps have no tiles
 * (lsdtt_stream_power_cpu.cpp:218-319); the results do not depend on it.
 * _autotune_tiles times `steps` steps (after one untimed step) with each
 * height from the current surface, restores the surface after each, keeps
 * the fastest (the current height unless another is more than 3% faster)
 * and returns it in *tile_height; ms_per_step (3 entries, may be NULL)
 * receives the times of heights 8, 16 and 32. */
int lsdtt_gpu_lem_set_tile_height(lsdtt_gpu_lem* lem, int32_t tile_height);
int32_t lsdtt_gpu_lem_tile_height(lsdtt_gpu_lem* lem);
int lsdtt_gpu_lem_autotune_tiles(lsdtt_gpu_lem* lem, const int64_t* boundary_flats, int64_t n_boundary,
                                 int32_t steps, int32_t* tile_height, double* ms_per_step);

/* ---------------------------------------------------------------------
 * LEM timestep control (src/cpu/lsdtt_lem_timestep_cpu.hpp:62-292,
 * lsdtt_lem_timestep_cpu.cpp:28-302).  The median interior gradient (valid
 * cells whose four cardinal neighbours are valid; sqrt(gx^2 + gy^2) of
 * central differences; the mean of the two middle values for an even count)
 * is computed on the device from the LEM context's resident surface, exactly
 * (radix selection of the FP64 values).  The controller and the monitor
 * are the reference's host-side decisions.
 * ------------------------------------------------------------------- */
typedef struct {
    double target_ratio;            /* 0.7 */
    double dt_min_yr, dt_max_yr;    /* 0: dt_initial / 100, dt_initial * 100 */
    double grow_factor;             /* 1.15 */
    double shrink_factor;           /* 0.5 */
    int32_t slope_refresh_interval; /* 25 */
    double ladder_ratio;            /* 2 (1: continuous dt) */
    double ratio_smoothing;         /* 0.15 */
} lsdtt_gpu_timestep_options;

typedef struct {
    double reference_slope, cell_relief_m, uplift_per_step_m, ratio, suggested_dt_yr;
    int32_t too_large;
    char warning[512];
} lsdtt_gpu_timestep_diagnosis;

typedef struct {
    double dz_tolerance_m;              /* 1e-4 */
    double dz_rate_tolerance_m_per_yr;  /* 0: off */
    int32_t patience_steps;             /* 5 */
    double min_dt_fraction;             /* 0.25 */
    int32_t stall_window;               /* 1000 */
    double stall_decay_threshold;       /* 0.99 */
    double oscillation_alternation_fraction;  /* 0.7 */
} lsdtt_gpu_convergence_options;

typedef struct lsdtt_gpu_timestep_controller lsdtt_gpu_timestep_controller;
typedef struct lsdtt_gpu_convergence_monitor lsdtt_gpu_convergence_monitor;

void lsdtt_gpu_timestep_options_init(lsdtt_gpu_timestep_options* o);
void lsdtt_gpu_convergence_options_init(lsdtt_gpu_convergence_options* o);
/* median_interior_gradient of the resident surface */
int lsdtt_gpu_lem_median_interior_gradient(lsdtt_gpu_lem* lem, double* gradient);
/* diagnose_lem_timestep (target_ratio <= 0: 0.7) */
int lsdtt_gpu_lem_diagnose_timestep(lsdtt_gpu_lem* lem, double uplift_rate, double timestep_yr,
                                    double target_ratio, lsdtt_gpu_timestep_diagnosis* out);
/* the timestep of the following steps (lsdtt_lem_options::timestep) */
int lsdtt_gpu_lem_set_timestep(lsdtt_gpu_lem* lem, double timestep_yr);
/* lsdtt_lem_timestep_controller: _update reports the step just taken
 * (max_local_elevation_change of its statistics) and returns the next dt;
 * the reference gradient is refreshed from `lem` on the options' interval */
lsdtt_gpu_timestep_controller* lsdtt_gpu_timestep_controller_create(double dt_initial_yr,
                                                                    const lsdtt_gpu_timestep_options* o);
void lsdtt_gpu_timestep_controller_destroy(lsdtt_gpu_timestep_controller* c);
int lsdtt_gpu_timestep_controller_update(lsdtt_gpu_timestep_controller* c, lsdtt_gpu_lem* lem,
                                         double max_local_dz_m, double* dt_next);
int lsdtt_gpu_timestep_controller_get(const lsdtt_gpu_timestep_controller* c, double* dt, double* last_ratio,
                                      double* reference_slope, int64_t* n_shrinks, int64_t* n_grows,
                                      int64_t* n_changes);
/* lsdtt_lem_convergence_monitor: state 0 running, 1 converged, 2 stalled */
lsdtt_gpu_convergence_monitor* lsdtt_gpu_convergence_monitor_create(double dt_initial_yr,
                                                                    const lsdtt_gpu_convergence_options* o);
void lsdtt_gpu_convergence_monitor_destroy(lsdtt_gpu_convergence_monitor* m);
int lsdtt_gpu_convergence_monitor_update(lsdtt_gpu_convergence_monitor* m, double max_local_dz_m,
                                         double dt_used_yr, int32_t* state);
int lsdtt_gpu_convergence_monitor_get(const lsdtt_gpu_convergence_monitor* m, int32_t* state,
                                      int32_t* oscillating, int64_t* n_steps, double* last_dz,
                                      double* window_decay, char* message, int32_t message_len);

/* ---------------------------------------------------------------------
 * D-infinity flow graph (lsdtt_flowgraph_cpu with method d_infinity,
 * src/cpu/lsdtt_flowgraph_cpu.cpp:87-177, 423-663) and its accumulation
 * (accumulate() with cell_size^2 per valid cell, :988-1033) from host
 * arrays.  Valid cells: mask != 0, z finite and (when nodata is finite)
 * z != nodata.  Per cell f: receivers int64[2f], [2f+1] (-1 = none; the
 * reference's CSR rows in slot order) and weights float[2f], [2f+1];
 * slope and bearing float rasters (nodata where there is no receiver);
 * area double raster (0 off the valid cells).  Any output may be NULL.
 * atan2 is the correctly rounded value; the reference's std::atan2(float,
 * float) returns it or one ulp beside it, depending on the C library.  The
 * accumulation adds a cell's donors in direction order (N, NE, ..., NW)
 * where the reference follows its Kahn order: the areas agree to rounding.
 * ------------------------------------------------------------------- */
int lsdtt_gpu_dinf_graph(const float* z, const uint8_t* mask, int32_t n_rows, int32_t n_cols, double cell_size,
                         const char* boundary_conditions, int32_t interior_nodata_is_outlet, double nodata,
                         int64_t* receivers, float* weights, float* slope, float* bearing, double* area);

/* Analytic fluvial steady state on the D-infinity graph of z:
 * compute_dinf_fluvial_steady_state(flow, opts)
 * (src/cpu/lsdtt_steady_state_cpu.cpp:227-436) on the graph and area of
 * lsdtt_gpu_dinf_graph.  Per valid cell with receivers, amn = powf(A, -m/n)
 * w as on D8; two receivers (neither weight >= 1 - 1e-6): dchi =
 * (chi(e2) - chi(e1)) / dx, disc = amn^2 - dchi^2, chi = chi(e1) +
 * dx sqrt(disc), or chi(e1) + dx amn when disc < 0; one receiver:
 * chi(r) + dx_eff amn with dx_eff = dx sqrt2 when r is one row and one
 * column away.  Scalars, z_out and chi_c as lsdtt_gpu_flow_steady_state;
 * nodata marks invalid input cells and fills z_out off them.  area (double
 * raster, may be NULL) receives the accumulated area used.  For lsdtt-lem's
 * snap with use_dinf (lsdtt_lem_cpu.cpp:830-836): lsdtt_gpu_fill, then
 * this on the filled surface. */
int lsdtt_gpu_dinf_steady_state(const float* z, const uint8_t* mask, int32_t n_rows, int32_t n_cols,
                                double cell_size, const char* boundary_conditions, int32_t interior_nodata_is_outlet,
                                double nodata, const lsdtt_gpu_steady_state_options* o,
                                const float* erodibility_field, float* z_out, double* area,
                                lsdtt_gpu_steady_state_result* result);

/* ---------------------------------------------------------------------
 * Batched ensemble: n independent LEM instances (uneven grids, per-instance
 * options) stepped together on one GPU.  No counterpart in the reference,
 * which runs one model per process (cli/lsdtt-lem.cpp:3553-3700); each
 * instance's step is run_lem_step as above (D8 incision, roering_implicit
 * Newton hillslope with its Picard fallback, uplift, boundary flats from
 * the instance's boundary codes, statistics).  Every kernel of a step runs
 * once for all instances, each instance's solver decisions are taken on
 * the device, and the step's launches are replayed from a captured
 * hipGraph.  An instance's results do not depend on the other instances.
 * Not supported here (LSDTT_GPU_ERR_ARG at creation): D-inf incision, area
 * smoothing, incision thresholds or channel sets, forced Picard.
 * n_workers is ignored (kept for source compatibility).
 * ------------------------------------------------------------------- */
typedef struct lsdtt_gpu_ensemble lsdtt_gpu_ensemble;

lsdtt_gpu_ensemble* lsdtt_gpu_ensemble_create(int32_t n_instances, const int32_t* n_rows, const int32_t* n_cols,
                                              const double* cell_size, const uint8_t* const* masks,
                                              const lsdtt_gpu_lem_options* options, int32_t n_workers);
void lsdtt_gpu_ensemble_destroy(lsdtt_gpu_ensemble* e);
int lsdtt_gpu_ensemble_set_elevation(lsdtt_gpu_ensemble* e, int32_t instance, const double* z);
int lsdtt_gpu_ensemble_get_elevation(lsdtt_gpu_ensemble* e, int32_t instance, double* z);
/* The last step's per-pixel erosion rate of an instance created with
 * compute_erosion_rate_field (as lsdtt_gpu_lem_get_erosion_rate). */
int lsdtt_gpu_ensemble_get_erosion_rate(lsdtt_gpu_ensemble* e, int32_t instance, float* rate);
/* n_steps steps of every instance; last_stats (n_instances entries, may be
 * NULL) receives each instance's last step; *seconds the wall time from
 * the first launch to the last step's completion. */
int lsdtt_gpu_ensemble_run(lsdtt_gpu_ensemble* e, int32_t n_steps, lsdtt_gpu_lem_stats* last_stats,
                           double* seconds);
/* enable != 0 (the default): each step replays the captured launch
 * pattern; 0: the same launches are enqueued one by one.  Results are the
 * same either way. */
int lsdtt_gpu_ensemble_set_graph(lsdtt_gpu_ensemble* e, int32_t enable);
/* Warm start of every instance's first Newton solve of a step (no
 * counterpart in the reference): mode 3 tries z^n plus the instance's
 * previous step's hillslope change as the first iterate, kept when its
 * residual is below the residual at z^n (and, nonlinear, its slope ratio
 * feasible); mode 0 (the default) starts from z^n.  The change is recorded
 * only by steps whose solve converged by Newton and is forgotten by
 * _set_elevation.  Results agree with mode 0 to the Newton tolerance;
 * the stats' warm_start_* fields report each instance's trial.  Needs
 * 4 bytes per cell more device memory. */
int lsdtt_gpu_ensemble_set_warm_start(lsdtt_gpu_ensemble* e, int32_t mode);
/* Launch-pattern counters: graph replays and captures, solver rounds and
 * tile passes launched outside the pattern (an instance that needed more
 * than its caps), steps restarted because a sweep did not settle within
 * the pattern, and caps[3] = the current pattern's count passes, stream
 * power passes and solver rounds.  Any pointer may be NULL. */
int lsdtt_gpu_ensemble_pattern(lsdtt_gpu_ensemble* e, int64_t* graph_launches, int64_t* graph_captures,
                               int64_t* extra_rounds, int64_t* extra_passes, int64_t* restarts,
                               int32_t* caps);
/* Tile height of the routing sweeps of every instance (8, 16 or 32; speed
 * only, results are the same). */
int lsdtt_gpu_ensemble_set_tile_height(lsdtt_gpu_ensemble* e, int32_t tile_height);
int32_t lsdtt_gpu_ensemble_tile_height(lsdtt_gpu_ensemble* e);
/* Times `steps` steps of the whole ensemble with each tile height from the
 * same surfaces (restored afterwards) and keeps the fastest (the current
 * height unless another is more than 3% faster); *tile_height receives it,
 * ms_per_step[3] (may be NULL) each height's time per step. */
int lsdtt_gpu_ensemble_autotune_tiles(lsdtt_gpu_ensemble* e, int32_t steps, int32_t* tile_height,
                                      double* ms_per_step);

/* Per-kernel timing with HIP events recorded on the step's stream (no
 * counterpart in the reference; for roofline measurement).  enable != 0
 * starts recording (and clears previous totals); _get returns, for kernel
 * group `index` < _count, its name, total milliseconds, launches and the
 * algorithmic bytes those launches move (each array element a kernel must
 * read or write counted once). */
int lsdtt_gpu_lem_profile(lsdtt_gpu_lem* lem, int32_t enable);
int32_t lsdtt_gpu_lem_profile_count(lsdtt_gpu_lem* lem);
int lsdtt_gpu_lem_profile_get(lsdtt_gpu_lem* lem, int32_t index, char* name, int32_t name_len,
                              double* ms, int64_t* launches, double* bytes);

/* ---------------------------------------------------------------------
 * Learned hillslope warm start.  No counterpart in the reference, which
 * always starts the hillslope Newton solve from z^n
 * (lsdtt_hillslope_newton_cpu.cpp:196-260).  A 2-D U-Net (`levels`
 * resolution levels, `base_channels` doubling per level, two 3x3
 * convolutions + ReLU per level each way, 2x2 average pooling, nearest
 * upsampling with skip concatenation, 1x1 head) with BF16 weights and
 * activations and FP32 accumulation on MFMA; FP32 master weights, Adam.
 * The LEM feeds it windows of tile + 2 halo cells with four features per
 * pixel (z minus the window mean over z_scale, |grad z| / Sc, the h^2
 * Laplacian over lap_scale, dt D / h^2 times kappa_scale) and takes its
 * output times dt D / h^2 * lap_scale as the predicted hillslope change.
 * ------------------------------------------------------------------- */
typedef struct lsdtt_gpu_cnn lsdtt_gpu_cnn;
typedef struct lsdtt_gpu_cnn_dataset lsdtt_gpu_cnn_dataset;

/* in_channels 1..32; base_channels a multiple of 32 with
 * base_channels * 2^(levels-1) <= 256; 2 <= levels <= 6.  Weights are
 * He-normal from `seed`, biases zero.  tile/halo: the LEM windows (256/16). */
lsdtt_gpu_cnn* lsdtt_gpu_cnn_create(int32_t in_channels, int32_t base_channels, int32_t levels, uint64_t seed,
                                    int32_t tile, int32_t halo);
void lsdtt_gpu_cnn_destroy(lsdtt_gpu_cnn* net);
int64_t lsdtt_gpu_cnn_num_params(const lsdtt_gpu_cnn* net);
/* FP32 parameter vector: per 3x3 layer W[cout][9][cin_pad] (tap = 3 (dy+1)
 * + dx+1, cin_pad = cin rounded up to 8, padded channels zero) then
 * bias[cout]; layers in the order: for each level from the finest, the
 * encoder's two convolutions; then for each level below the coarsest, from
 * the finest, the decoder's two convolutions (the first one's input is
 * [upsampled coarser output, encoder skip]); then the head's
 * weights[base_channels] and bias. */
int lsdtt_gpu_cnn_get_params(const lsdtt_gpu_cnn* net, float* params);
int lsdtt_gpu_cnn_set_params(lsdtt_gpu_cnn* net, const float* params);
int lsdtt_gpu_cnn_get_grads(const lsdtt_gpu_cnn* net, float* grads);
/* Host arrays: x FP32 NHWC [n][h][w][in_channels], y FP32 [n][h][w]; h and
 * w divisible by 2^(levels-1), coarsest level's pixel count by 16. */
int lsdtt_gpu_cnn_forward(lsdtt_gpu_cnn* net, const float* x, int32_t n, int32_t h, int32_t w, float* y);
/* After _forward: MSE over pixels with mask != 0 (NULL = all) against
 * target [n][h][w]; the parameter gradient is ADDED to the gradient buffer
 * (cleared by _zero_grad); *loss the MSE. */
int lsdtt_gpu_cnn_backward(lsdtt_gpu_cnn* net, const float* target, const uint8_t* mask, double* loss);
int lsdtt_gpu_cnn_zero_grad(lsdtt_gpu_cnn* net);
int lsdtt_gpu_cnn_adam_step(lsdtt_gpu_cnn* net, double lr, double beta1, double beta2, double eps);
/* Throughput of the network on random inputs of n x h x w: mean
 * milliseconds of a forward pass, of its 3x3 convolution stack alone, and
 * of a training iteration (forward, backward, Adam); *flops the 3x3
 * convolutions' FLOP of one forward pass (2 per multiply-add). */
int lsdtt_gpu_cnn_benchmark(lsdtt_gpu_cnn* net, int32_t n, int32_t h, int32_t w, int32_t reps, double* forward_ms,
                            double* conv_ms, double* train_ms, double* flops);

/* One 3x3 convolution layer of the network on host arrays, through the
 * network's own kernels (for checking them one layer at a time): x
 * [n][h][w][cin] and w [cout][9][cin] are rounded to BF16; y
 * [n][h][w][cout] receives the BF16 output (+ bias b, ReLU when relu != 0).
 * With gy != NULL (FP32 output gradient): gw [cout][9][cin], gb [cout] and,
 * if gx != NULL (cin a multiple of 32), gx [n][h][w][cin]; the masked
 * output gradient is rounded to BF16 first, as in training.  cout a
 * multiple of 32, n*h*w a multiple of 16. */
int lsdtt_gpu_cnn_conv3x3(const float* x, const float* w, const float* b, int32_t n, int32_t h, int32_t w_px,
                          int32_t cin, int32_t cout, int32_t relu, float* y, const float* gy, float* gx, float* gw,
                          float* gb);

/* Training pairs from LEM steps.  _capture makes every later step of `lem`
 * keep its post-incision surface and hillslope change; _add copies the
 * last captured pair into the dataset (device-resident; all pairs on the
 * dataset's grid, whose mask and hillslope DOFs are taken from the first
 * LEM added). */
int lsdtt_gpu_lem_capture_hillslope(lsdtt_gpu_lem* lem, int32_t enable);
lsdtt_gpu_cnn_dataset* lsdtt_gpu_cnn_dataset_create(void);
void lsdtt_gpu_cnn_dataset_destroy(lsdtt_gpu_cnn_dataset* ds);
int lsdtt_gpu_cnn_dataset_add(lsdtt_gpu_cnn_dataset* ds, lsdtt_gpu_lem* lem);
int32_t lsdtt_gpu_cnn_dataset_size(const lsdtt_gpu_cnn_dataset* ds);
/* n_iterations Adam steps (beta 0.9 / 0.999, eps 1e-8) on batches of
 * random windows; losses (n_iterations, may be NULL) the batch MSEs. */
int lsdtt_gpu_cnn_train(lsdtt_gpu_cnn* net, lsdtt_gpu_cnn_dataset* ds, int32_t n_iterations, int32_t batch,
                        double lr, uint64_t seed, float* losses);
/* Mean MSE of the network and of the zero prediction on n_batches random
 * batches (parameters unchanged). */
int lsdtt_gpu_cnn_evaluate(lsdtt_gpu_cnn* net, lsdtt_gpu_cnn_dataset* ds, int32_t n_batches, int32_t batch,
                           uint64_t seed, double* loss, double* zero_loss);
/* Warm start of each step's first Newton solve: mode 0 off (z^n), 1 the
 * explicit Euler predictor z^n + dt (-F(z^n)), 2 the network `net` (which
 * must outlive its use here), 3 the previous step's hillslope change
 * (z^n + (z_hs - z_in) of the last step, used when that step's Newton
 * solve converged over the same dt and the surface was not set since).
 * The predicted start is kept only when it lowers ||F||inf; the converged
 * residual is judged as without it. */
int lsdtt_gpu_lem_set_warm_start(lsdtt_gpu_lem* lem, lsdtt_gpu_cnn* net, int32_t mode);

/* Valid cells with an 8-neighbour off a fixed ('o'/'b') edge or masked in
 * a fixed direction (build_boundary_flats, lsdtt_lem_cpu.cpp:735-808);
 * writes up to n_rows*n_cols flats, count in *n_out. */
int lsdtt_gpu_build_boundary_flats(int32_t n_rows, int32_t n_cols, const uint8_t* mask,
                                   const char* boundary_conditions, int64_t* out, int64_t* n_out);

/* ------------------------------------------------------------------------
 * Raster files on either side of the LEM path (host code, no GPU needed):
 * single-band GeoTIFF (.tif/.tiff: strips or tiles, uncompressed, LZW or
 * DEFLATE, predictors 2 and 3, integer or float samples; written as float32
 * strips, DEFLATE + floating-point predictor when compress != 0) and ESRI
 * ASCII grids (.asc).  They replace the GDAL-backed lsdtt_data_loader's
 * read_full / write_full for these formats (src/core/lsdtt_data_loader.hpp,
 * src/core/lsdtt_data_loader.cpp:700-765): values as float32, mask 0 where a
 * value equals the no-data value or is not finite (those values become the
 * no-data value), north-up and square pixels required (geographic CRS
 * excepted), no-data -9999 when the file names none; the meta mirrors
 * lsdtt_raster_meta (src/core/lsdtt_raster_meta.hpp:44-84) with the CRS as
 * an EPSG code instead of WKT.
 * ------------------------------------------------------------------------ */
typedef struct lsdtt_gpu_raster_meta {
    int32_t n_rows, n_cols;
    double x_min;          /* west edge */
    double y_min;          /* south edge */
    double cell_size;      /* pixel width */
    double cell_size_y;    /* pixel height when it differs (geographic CRS), else 0 */
    double no_data_value;
    int32_t epsg;          /* CRS code, 0 when unknown */
    int32_t geographic;    /* 1: geographic CRS */
} lsdtt_gpu_raster_meta;

/* Header only: fills *meta. */
int lsdtt_gpu_raster_info(const char* path, lsdtt_gpu_raster_meta* meta);
/* data float[n_rows*n_cols] (row-major, north row first), mask uint8[n] or
 * NULL; meta must hold the shape lsdtt_gpu_raster_info returned. */
int lsdtt_gpu_raster_read(const char* path, float* data, uint8_t* mask, lsdtt_gpu_raster_meta* meta);
/* Cells with mask 0 (mask may be NULL) are written as the no-data value. */
int lsdtt_gpu_raster_write(const char* path, const float* data, const uint8_t* mask,
                           const lsdtt_gpu_raster_meta* meta, int32_t compress);

#ifdef __cplusplus
}
#endif
#endif
