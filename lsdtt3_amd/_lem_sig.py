"""ctypes signatures and structs of the LEM entry points of
include/lsdtt3_gpu.h."""
import ctypes as C


class TimestepOptions(C.Structure):
    _fields_ = [("target_ratio", C.c_double), ("dt_min_yr", C.c_double), ("dt_max_yr", C.c_double),
                ("grow_factor", C.c_double), ("shrink_factor", C.c_double), ("slope_refresh_interval", C.c_int32),
                ("ladder_ratio", C.c_double), ("ratio_smoothing", C.c_double)]


class TimestepDiagnosis(C.Structure):
    _fields_ = [("reference_slope", C.c_double), ("cell_relief_m", C.c_double), ("uplift_per_step_m", C.c_double),
                ("ratio", C.c_double), ("suggested_dt_yr", C.c_double), ("too_large", C.c_int32),
                ("warning", C.c_char * 512)]


class ConvergenceOptions(C.Structure):
    _fields_ = [("dz_tolerance_m", C.c_double), ("dz_rate_tolerance_m_per_yr", C.c_double),
                ("patience_steps", C.c_int32), ("min_dt_fraction", C.c_double), ("stall_window", C.c_int32),
                ("stall_decay_threshold", C.c_double), ("oscillation_alternation_fraction", C.c_double)]


class LemOptions(C.Structure):
    _fields_ = [
        ("timestep", C.c_double), ("erodibility", C.c_double),
        ("m_exponent", C.c_double), ("n_exponent", C.c_double),
        ("incision_threshold_area", C.c_double), ("divide_area_by_dx", C.c_int32),
        ("slope_clamp", C.c_double), ("hillslope_diffusivity", C.c_double),
        ("enable_nonlinear_hillslopes", C.c_int32), ("critical_hillslope_slope", C.c_double),
        ("uplift_rate", C.c_double), ("baselevel_fall_rate", C.c_double),
        ("enable_d8_fluvial_erosion", C.c_int32), ("enable_hillslope_transport", C.c_int32),
        ("enable_uplift", C.c_int32), ("boundary_conditions", C.c_char * 4),
        ("newton_abs_tol", C.c_double), ("newton_rel_tol", C.c_double),
        ("newton_max_iterations", C.c_int32), ("newton_max_line_search", C.c_int32),
        ("slope_ratio_cap", C.c_double), ("krylov_rel_tol", C.c_double),
        ("krylov_max_iterations", C.c_int32),
        ("drainage_area_kernel_size", C.c_int32), ("multiply_by_w_over_dx", C.c_int32),
        ("channel_width_w", C.c_double), ("hillslope_coupling_regime", C.c_int32),
        ("hillslope_force_picard", C.c_int32), ("picard_tol", C.c_double), ("picard_max_iter", C.c_int32),
        ("relaxation_omega", C.c_double), ("picard_slope_ratio_cap", C.c_double),
        ("picard_dt_init", C.c_double), ("picard_dt_min", C.c_double), ("picard_dt_max", C.c_double),
        ("picard_grow_factor", C.c_double), ("picard_shrink_factor", C.c_double),
        ("picard_dz_target", C.c_double), ("enable_dinf_fluvial_erosion", C.c_int32),
        ("compute_erosion_rate_field", C.c_int32), ("krylov_forcing", C.c_int32),
    ]


class LemStats(C.Structure):
    _fields_ = [
        ("delta_max_elevation", C.c_double), ("delta_mean_elevation", C.c_double),
        ("max_local_elevation_change", C.c_double), ("p999_local_elevation_change", C.c_double),
        ("mean_erosion_rate", C.c_double), ("max_erosion_rate", C.c_double),
        ("n_valid", C.c_int64), ("cells_eroded", C.c_int64),
        ("newton_iterations", C.c_int32), ("krylov_iterations", C.c_int32),
        ("hillslope_substeps", C.c_int32), ("hillslope_rejected", C.c_int32),
        ("hillslope_converged", C.c_int32),
        ("hillslope_initial_residual", C.c_double), ("hillslope_residual", C.c_double),
        ("hillslope_tolerance", C.c_double), ("max_slope_ratio", C.c_double),
        ("count_passes", C.c_int32), ("stream_power_passes", C.c_int32),
        ("count_tile_runs", C.c_int64), ("stream_power_tile_runs", C.c_int64),
        ("hillslope_method", C.c_int32), ("picard_iterations", C.c_int32), ("picard_substeps", C.c_int32),
        ("picard_rejected", C.c_int32), ("picard_cg_iterations", C.c_int32),
        ("picard_limiter_fired", C.c_int32), ("picard_max_slope_ratio", C.c_double),
        ("warm_start_mode", C.c_int32), ("warm_start_accepted", C.c_int32),
        ("warm_start_residual", C.c_double), ("newton_restoration_iterations", C.c_int32),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def declare(L, sig):
    vp, i32, i64, dbl = C.c_void_p, C.c_int32, C.c_int64, C.c_double
    sig(L, "lsdtt_gpu_lem_options_init", None, [C.POINTER(LemOptions)])
    sig(L, "lsdtt_gpu_lem_create", vp, [i32, i32, dbl, vp, C.POINTER(LemOptions)])
    sig(L, "lsdtt_gpu_lem_destroy", None, [vp])
    sig(L, "lsdtt_gpu_lem_set_elevation", C.c_int, [vp, vp])
    sig(L, "lsdtt_gpu_lem_get_elevation", C.c_int, [vp, vp])
    sig(L, "lsdtt_gpu_lem_set_erodibility_field", C.c_int, [vp, vp])
    sig(L, "lsdtt_gpu_lem_set_channel_mask", C.c_int, [vp, vp])
    sig(L, "lsdtt_gpu_ensemble_create", vp, [C.c_int32, vp, vp, vp, vp, vp, C.c_int32])
    sig(L, "lsdtt_gpu_ensemble_destroy", None, [vp])
    sig(L, "lsdtt_gpu_ensemble_set_elevation", C.c_int, [vp, C.c_int32, vp])
    sig(L, "lsdtt_gpu_ensemble_get_elevation", C.c_int, [vp, C.c_int32, vp])
    sig(L, "lsdtt_gpu_ensemble_run", C.c_int, [vp, C.c_int32, vp, vp])
    sig(L, "lsdtt_gpu_ensemble_get_erosion_rate", C.c_int, [vp, i32, vp])
    sig(L, "lsdtt_gpu_ensemble_set_graph", C.c_int, [vp, i32])
    sig(L, "lsdtt_gpu_ensemble_set_warm_start", C.c_int, [vp, i32])
    sig(L, "lsdtt_gpu_ensemble_pattern", C.c_int, [vp, vp, vp, vp, vp, vp, vp])
    sig(L, "lsdtt_gpu_ensemble_set_tile_height", C.c_int, [vp, i32])
    sig(L, "lsdtt_gpu_ensemble_tile_height", i32, [vp])
    sig(L, "lsdtt_gpu_ensemble_autotune_tiles", C.c_int, [vp, i32, vp, vp])
    sig(L, "lsdtt_gpu_lem_set_tile_height", C.c_int, [vp, i32])
    sig(L, "lsdtt_gpu_lem_tile_height", i32, [vp])
    sig(L, "lsdtt_gpu_lem_get_channel_mask", C.c_int, [vp, vp])
    sig(L, "lsdtt_gpu_lem_get_erosion_rate", C.c_int, [vp, vp])
    pd = C.POINTER(C.c_double)
    pi64 = C.POINTER(C.c_int64)
    sig(L, "lsdtt_gpu_timestep_options_init", None, [C.POINTER(TimestepOptions)])
    sig(L, "lsdtt_gpu_convergence_options_init", None, [C.POINTER(ConvergenceOptions)])
    sig(L, "lsdtt_gpu_lem_median_interior_gradient", C.c_int, [vp, pd])
    sig(L, "lsdtt_gpu_lem_diagnose_timestep", C.c_int, [vp, dbl, dbl, dbl, C.POINTER(TimestepDiagnosis)])
    sig(L, "lsdtt_gpu_lem_set_timestep", C.c_int, [vp, dbl])
    sig(L, "lsdtt_gpu_timestep_controller_create", vp, [dbl, C.POINTER(TimestepOptions)])
    sig(L, "lsdtt_gpu_timestep_controller_destroy", None, [vp])
    sig(L, "lsdtt_gpu_timestep_controller_update", C.c_int, [vp, vp, dbl, pd])
    sig(L, "lsdtt_gpu_timestep_controller_get", C.c_int, [vp, pd, pd, pd, pi64, pi64, pi64])
    sig(L, "lsdtt_gpu_convergence_monitor_create", vp, [dbl, C.POINTER(ConvergenceOptions)])
    sig(L, "lsdtt_gpu_convergence_monitor_destroy", None, [vp])
    sig(L, "lsdtt_gpu_convergence_monitor_update", C.c_int, [vp, dbl, dbl, C.POINTER(C.c_int32)])
    sig(L, "lsdtt_gpu_convergence_monitor_get", C.c_int, [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32), pi64,
                                                           pd, pd, C.c_char_p, i32])
    sig(L, "lsdtt_gpu_dinf_graph", C.c_int, [vp, vp, i32, i32, dbl, C.c_char_p, i32, dbl, vp, vp, vp, vp, vp])
    sig(L, "lsdtt_gpu_dinf_steady_state", C.c_int, [vp, vp, i32, i32, dbl, C.c_char_p, i32, dbl, vp, vp, vp, vp, vp])
    sig(L, "lsdtt_gpu_lem_autotune_tiles", C.c_int, [vp, vp, i64, i32, vp, vp])
    sig(L, "lsdtt_gpu_lem_elevation_device", C.c_void_p, [vp])
    sig(L, "lsdtt_gpu_lem_step", C.c_int, [vp, vp, i64, C.POINTER(LemStats)])
    sig(L, "lsdtt_gpu_build_boundary_flats", C.c_int, [i32, i32, vp, C.c_char_p, vp, vp])
    sig(L, "lsdtt_gpu_lem_profile", C.c_int, [vp, i32])
    sig(L, "lsdtt_gpu_lem_profile_count", C.c_int32, [vp])
    sig(L, "lsdtt_gpu_lem_profile_get", C.c_int, [vp, i32, C.c_char_p, i32, vp, vp, vp])
    u64 = C.c_uint64
    sig(L, "lsdtt_gpu_cnn_create", vp, [i32, i32, i32, u64, i32, i32])
    sig(L, "lsdtt_gpu_cnn_destroy", None, [vp])
    sig(L, "lsdtt_gpu_cnn_num_params", i64, [vp])
    sig(L, "lsdtt_gpu_cnn_get_params", C.c_int, [vp, vp])
    sig(L, "lsdtt_gpu_cnn_set_params", C.c_int, [vp, vp])
    sig(L, "lsdtt_gpu_cnn_get_grads", C.c_int, [vp, vp])
    sig(L, "lsdtt_gpu_cnn_forward", C.c_int, [vp, vp, i32, i32, i32, vp])
    sig(L, "lsdtt_gpu_cnn_backward", C.c_int, [vp, vp, vp, vp])
    sig(L, "lsdtt_gpu_cnn_zero_grad", C.c_int, [vp])
    sig(L, "lsdtt_gpu_cnn_adam_step", C.c_int, [vp, dbl, dbl, dbl, dbl])
    sig(L, "lsdtt_gpu_cnn_benchmark", C.c_int, [vp, i32, i32, i32, i32, vp, vp, vp, vp])
    sig(L, "lsdtt_gpu_cnn_conv3x3", C.c_int, [vp, vp, vp, i32, i32, i32, i32, i32, i32, vp, vp, vp, vp, vp])
    sig(L, "lsdtt_gpu_lem_capture_hillslope", C.c_int, [vp, i32])
    sig(L, "lsdtt_gpu_cnn_dataset_create", vp, [])
    sig(L, "lsdtt_gpu_cnn_dataset_destroy", None, [vp])
    sig(L, "lsdtt_gpu_cnn_dataset_add", C.c_int, [vp, vp])
    sig(L, "lsdtt_gpu_cnn_dataset_size", i32, [vp])
    sig(L, "lsdtt_gpu_cnn_train", C.c_int, [vp, vp, i32, i32, dbl, u64, vp])
    sig(L, "lsdtt_gpu_cnn_evaluate", C.c_int, [vp, vp, i32, i32, u64, vp, vp])
    sig(L, "lsdtt_gpu_lem_set_warm_start", C.c_int, [vp, vp, i32])
