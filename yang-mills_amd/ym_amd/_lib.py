"""Loader and ctypes signatures for libym_amd.so (include/ym_amd.h)."""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
_PKG = os.path.dirname(_HERE)


def library_path():
    return os.environ.get("YM_AMD_LIB", os.path.join(_PKG, "libym_amd.so"))


i32, i64, u64, dbl, flt = C.c_int32, C.c_int64, C.c_uint64, C.c_double, C.c_float
P = C.POINTER


class ChainDiagnostics(C.Structure):
    _fields_ = [("mean", dbl), ("variance", dbl), ("positive_window_lag", i64),
                ("tau_int", dbl), ("ess", dbl), ("standard_error", dbl)]


class Su3ChainConfig(C.Structure):
    _fields_ = [("extents", i64 * 4), ("beta", dbl), ("master_seed", u64), ("chain_id", u64),
                ("thermalization_sweeps", i64), ("measurements", i64),
                ("sweeps_between_measurements", i64), ("proposal_angle", dbl),
                ("autocorrelation_max_lag", i64), ("hits", i32), ("precision", i32),
                ("hot_start", i32), ("reunitarize", i32)]


class Su3ChainReport(C.Structure):
    _fields_ = [("derived_seed", u64), ("total_sweeps", i64), ("proposals", u64), ("accepts", u64),
                ("mean_acceptance", dbl), ("final_unitarity_residual", dbl),
                ("final_determinant_residual", dbl), ("diagnostics", ChainDiagnostics)]


class Su2ChainConfig(C.Structure):
    _fields_ = [("extents", i64 * 4), ("beta", dbl), ("master_seed", u64), ("chain_id", u64),
                ("thermalization_sweeps", i64), ("measurements", i64), ("sweeps_between_measurements", i64),
                ("overrelaxation_sweeps_per_heatbath", i64), ("max_heatbath_attempts_per_link", i64),
                ("autocorrelation_max_lag", i64), ("precision", i32), ("hot_start", i32)]


class Su2ChainReport(C.Structure):
    _fields_ = [("derived_seed", u64), ("total_heatbath_sweeps", i64), ("total_overrelaxation_sweeps", i64),
                ("heatbath_proposals", u64), ("heatbath_updates", u64), ("mean_heatbath_acceptance", dbl),
                ("final_link_unit_residual", dbl), ("diagnostics", ChainDiagnostics)]


class EfficiencyReport(C.Structure):
    _fields_ = [("diagnostics", ChainDiagnostics), ("reweighting_efficiency", dbl), ("weight_only_ess", dbl),
                ("combined_ess_proxy", dbl), ("total_cost", dbl), ("sampling_cost_per_ess", dbl),
                ("total_cost_per_ess", dbl)]


class StationarityReport(C.Structure):
    _fields_ = [("first_half_mean", dbl), ("second_half_mean", dbl), ("first_half_se", dbl),
                ("second_half_se", dbl), ("drift_z", dbl)]


class EffectiveMassPoint(C.Structure):
    _fields_ = [("tau", i64), ("full_estimate", dbl), ("jackknife_mean", dbl), ("standard_error", dbl)]


class ScalarFitResult(C.Structure):
    _fields_ = [("window_start", i64), ("window_end", i64), ("plateau_weighted_effective_mass", dbl),
                ("plateau_maximum_z", dbl), ("fitted_mass", dbl), ("fitted_amplitude", dbl), ("chi_square", dbl),
                ("degrees_of_freedom", i64), ("chi_square_per_dof", dbl), ("mass_jackknife_mean", dbl),
                ("mass_standard_error", dbl), ("amplitude_jackknife_mean", dbl),
                ("amplitude_standard_error", dbl), ("maximum_off_diagonal_correlation", dbl)]


class PoolSummary(C.Structure):
    _fields_ = [("pooled_mean", dbl), ("pooled_se", dbl), ("aggregate_ess", dbl), ("max_pairwise_z", dbl),
                ("max_stationarity_z", dbl), ("min_ess", dbl)]


# name -> (restype, argtypes); the exported symbol set of include/ym_amd.h
SIGNATURES = {
    "ym_lattice_create": (i32, [i32, P(i64), i32, P(C.c_void_p)]),
    "ym_lattice_destroy": (i32, [C.c_void_p]),
    "ym_lattice_bytes": (i32, [C.c_void_p, P(i64)]),
    "ym_lattice_set_identity": (i32, [C.c_void_p]),
    "ym_lattice_hot_start": (i32, [C.c_void_p, u64]),
    "ym_lattice_upload": (i32, [C.c_void_p, P(dbl), i64]),
    "ym_lattice_upload_unchecked": (i32, [C.c_void_p, P(dbl), i64]),
    "ym_lattice_download": (i32, [C.c_void_p, P(dbl), i64]),
    "ym_lattice_copy": (i32, [C.c_void_p, C.c_void_p]),
    "ym_su3_metropolis_sweeps": (i32, [C.c_void_p, dbl, dbl, i32, u64, u64, i32, i32, P(u64), P(u64)]),
    "ym_su3_metropolis_sweeps_profiled": (i32, [C.c_void_p, dbl, dbl, i32, u64, u64, i32, i32, P(u64), P(u64),
                                                P(flt), P(flt), P(i32)]),
    "ym_su3_autotune": (i32, [C.c_void_p, dbl, dbl, i32, i32, P(i32), P(i32), P(flt)]),
    "ym_su2_heatbath_sweeps": (i32, [C.c_void_p, dbl, u64, u64, i32, i32, P(u64), P(u64)]),
    "ym_su2_overrelax_sweeps": (i32, [C.c_void_p, i32]),
    "ym_su2_markov_sweeps": (i32, [C.c_void_p, dbl, u64, u64, i32, i32, i32, P(u64), P(u64)]),
    "ym_su2_run_chain": (i32, [P(Su2ChainConfig), P(dbl), P(Su2ChainReport)]),
    "ym_average_plaquette": (i32, [C.c_void_p, P(dbl)]),
    "ym_average_plaquette_noalloc": (i32, [C.c_void_p, P(dbl)]),
    "ym_wilson_action": (i32, [C.c_void_p, dbl, P(dbl)]),
    "ym_staple": (i32, [C.c_void_p, P(i64), i32, P(dbl)]),
    "ym_path_holonomy": (i32, [C.c_void_p, P(i64), P(i32), i64, P(dbl), P(i64)]),
    "ym_wilson_loops": (i32, [C.c_void_p, i32, i32, i32, P(dbl), P(dbl)]),
    "ym_polyakov_loop": (i32, [C.c_void_p, i32, P(dbl), P(dbl), P(dbl)]),
    "ym_su3_timeslice_a1pp": (i32, [C.c_void_p, P(dbl), P(dbl)]),
    "ym_su3_stout_spatial": (i32, [C.c_void_p, C.c_void_p, dbl, i32]),
    "ym_su3_blocked_extents": (i32, [C.c_void_p, P(i64)]),
    "ym_su3_spatial_block": (i32, [C.c_void_p, C.c_void_p]),
    "ym_link_residuals": (i32, [C.c_void_p, P(dbl), P(dbl)]),
    "ym_reunitarize": (i32, [C.c_void_p]),
    "ym_su3_run_chain": (i32, [P(Su3ChainConfig), P(dbl), P(Su3ChainReport)]),
    "ym_diagnostics": (i32, [P(dbl), i64, i64, P(ChainDiagnostics), P(dbl)]),
    "ym_ensemble_create": (i32, [i32, P(i64), P(u64), i32, P(C.c_void_p)]),
    "ym_ensemble_destroy": (i32, [C.c_void_p]),
    "ym_ensemble_set_identity": (i32, [C.c_void_p]),
    "ym_ensemble_hot_start": (i32, [C.c_void_p]),
    "ym_ensemble_chain_count": (i32, [C.c_void_p, P(i32), P(i64), P(i64)]),
    "ym_ensemble_upload_chain": (i32, [C.c_void_p, i32, P(dbl), i64]),
    "ym_ensemble_download_chain": (i32, [C.c_void_p, i32, P(dbl), i64]),
    "ym_ensemble_sweep_index": (i32, [C.c_void_p, P(u64)]),
    "ym_ensemble_ops_length": (i32, [C.c_void_p, i32, P(i64), P(i64)]),
    "ym_ensemble_run_cycles": (i32, [C.c_void_p, dbl, dbl, i32, i32, i32, i32, dbl, i32, i32, P(dbl), P(dbl),
                                     P(u64), P(flt)]),
    "ym_mlp_create": (i32, [u64, i64, P(C.c_void_p)]),
    "ym_mlp_destroy": (i32, [C.c_void_p]),
    "ym_mlp_get_params": (i32, [C.c_void_p, P(flt)]),
    "ym_mlp_set_params": (i32, [C.c_void_p, P(flt)]),
    "ym_mlp_dataset_size": (i32, [C.c_void_p, P(i64), P(i64)]),
    "ym_mlp_clear_pairs": (i32, [C.c_void_p]),
    "ym_mlp_collect_pairs": (i32, [C.c_void_p, C.c_void_p, P(i64)]),
    "ym_mlp_download_pairs": (i32, [C.c_void_p, i64, i64, P(flt), P(flt), P(i32)]),
    "ym_mlp_upload_pairs": (i32, [C.c_void_p, P(flt), P(flt), P(i32), i64]),
    "ym_mlp_forward": (i32, [C.c_void_p, P(flt), i64, i32, P(flt)]),
    "ym_mlp_loss_grad": (i32, [C.c_void_p, i64, i64, P(flt), P(flt)]),
    "ym_mlp_train": (i32, [C.c_void_p, i32, i32, flt, u64, P(flt), P(flt)]),
    "ym_su3_learned_sweeps": (i32, [C.c_void_p, C.c_void_p, dbl, u64, u64, i32, i32, dbl, i32, P(u64), P(u64),
                                    P(u64), P(flt)]),
    "ym_su3_link_chart": (i32, [C.c_void_p, P(i64), i32, P(dbl)]),
    "ym_su3_learned_probe": (i32, [C.c_void_p, C.c_void_p, dbl, u64, u64, P(i64), i32, P(dbl)]),
    "ym_evaluate": (i32, [P(dbl), i64, P(dbl), i64, dbl, dbl, dbl, P(EfficiencyReport)]),
    "ym_stationarity": (i32, [P(dbl), i64, i64, P(StationarityReport)]),
    "ym_pool_chains": (i32, [P(dbl), i64, i64, i64, i64, P(dbl), P(dbl), P(dbl), P(dbl), P(PoolSummary)]),
    "ym_periodic_connected": (i32, [P(dbl), i64, i64, P(dbl), P(dbl), P(dbl)]),
    "ym_connected_correlator": (i32, [P(dbl), i64, i64, i64, P(dbl)]),
    "ym_jackknife_correlator": (i32, [P(dbl), i64, i64, P(dbl), P(dbl), P(dbl), P(dbl)]),
    "ym_su2_config_encode": (i32, [P(i64), P(dbl), C.c_char_p, i64, P(i64)]),
    "ym_su2_config_decode": (i32, [C.c_char_p, i64, P(i64), P(dbl), i64]),
    "ym_checkpoint_bytes": (i32, [i32, i32, P(i64), P(i64)]),
    "ym_checkpoint_encode": (i32, [i32, i32, P(i64), P(dbl), u64, u64, C.c_void_p, i64, P(i64)]),
    "ym_checkpoint_header": (i32, [C.c_void_p, i64, P(i32), P(i32), P(i64), P(u64), P(u64)]),
    "ym_checkpoint_decode": (i32, [C.c_void_p, i64, P(dbl), i64]),
    "ym_cosh_effective_mass": (i32, [P(dbl), i64, P(P(dbl)), P(i64), i64, P(i64), i64, P(EffectiveMassPoint)]),
    "ym_analytic_periodic_single_state": (i32, [i64, dbl, dbl, P(dbl)]),
    "ym_correlated_scalar_fit": (i32, [P(dbl), i64, P(P(dbl)), P(i64), i64, P(i64), i64, P(ScalarFitResult)]),
    "ym_variational_connected_correlator": (i32, [P(dbl), i64, i64, P(dbl), P(dbl)]),
    "ym_principal_generalized_projection": (i32, [P(dbl), i64, i64, i64, P(dbl), P(dbl), P(dbl)]),
    "ym_derive_chain_seed": (u64, [u64, u64]),
    "ym_derive_su3_chain_seed": (u64, [u64, u64]),
    "ym_timer_begin": (i32, [C.c_void_p]),
    "ym_timer_end": (i32, [C.c_void_p, P(flt)]),
    "ym_status_string": (C.c_char_p, [i32]),
}


class YmError(RuntimeError):
    def __init__(self, status, where=""):
        self.status = int(status)
        super().__init__(f"{where}: status {self.status} ({status_string(self.status)})")


_lib = None


def lib():
    """Load libym_amd.so; raises if the HIP library has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    path = library_path()
    if not os.path.exists(path):
        raise ImportError(f"libym_amd.so not built at {path}; run __graft_entry__.build()")
    L = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def status_string(code):
    return lib().ym_status_string(int(code)).decode()


def check(status, where):
    if status != 0:
        raise YmError(status, where)
