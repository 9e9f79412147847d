"""ctypes loader for the CPU oracle (oracle/libym_oracle.so).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use this module.  The oracle restates QSOLKCB/YANG-MILLS
(reference/rust/src/*.rs) on the CPU; it is the checker, never the product.
"""
import ctypes as C
import os
import subprocess

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(REPO, "oracle")
LIB_PATH = os.path.join(ORACLE_DIR, "libym_oracle.so")

i64 = C.c_int64
u64 = C.c_uint64
dbl = C.c_double
i32 = C.c_int32
P = C.POINTER


class ChainDiagnostics(C.Structure):
    _fields_ = [("mean", dbl), ("variance", dbl), ("positive_window_lag", i64),
                ("tau_int", dbl), ("ess", dbl), ("standard_error", dbl)]


class Su3ChainConfig(C.Structure):
    _fields_ = [("extents", i64 * 4), ("beta", dbl), ("master_seed", u64), ("chain_id", u64),
                ("thermalization_sweeps", i64), ("measurements", i64),
                ("sweeps_between_measurements", i64), ("proposal_angle", dbl),
                ("autocorrelation_max_lag", i64), ("no_caps", i64)]


class Su3ChainReport(C.Structure):
    _fields_ = [("derived_seed", u64), ("total_sweeps", i64), ("proposals", u64), ("accepts", u64),
                ("mean_acceptance", dbl), ("final_unitarity_residual", dbl),
                ("final_determinant_residual", dbl)]


class Su2ChainConfig(C.Structure):
    _fields_ = [("extents", i64 * 4), ("beta", dbl), ("master_seed", u64), ("chain_id", u64),
                ("thermalization_sweeps", i64), ("measurements", i64),
                ("sweeps_between_measurements", i64), ("overrelaxation_sweeps_per_heatbath", i64),
                ("max_heatbath_attempts_per_link", i64), ("autocorrelation_max_lag", i64),
                ("no_caps", i64)]


class Su2ChainReport(C.Structure):
    _fields_ = [("derived_seed", u64), ("total_heatbath_sweeps", i64),
                ("total_overrelaxation_sweeps", i64), ("heatbath_proposals", u64),
                ("heatbath_updates", u64), ("mean_heatbath_acceptance", dbl),
                ("final_link_unit_residual", dbl)]


class EfficiencyReport(C.Structure):
    _fields_ = [("diag", ChainDiagnostics), ("reweighting_efficiency", dbl), ("weight_only_ess", dbl),
                ("combined_ess_proxy", dbl), ("total_cost", dbl), ("sampling_cost_per_ess", dbl),
                ("total_cost_per_ess", dbl)]


class StationarityReport(C.Structure):
    _fields_ = [("first_half_mean", dbl), ("second_half_mean", dbl), ("first_half_se", dbl),
                ("second_half_se", dbl), ("drift_z", dbl)]


class PoolReport(C.Structure):
    _fields_ = [("chain_mean", dbl * 16), ("chain_se", dbl * 16), ("chain_ess", dbl * 16),
                ("chain_stationarity_z", dbl * 16), ("pooled_mean", dbl), ("pooled_se", dbl),
                ("aggregate_ess", dbl), ("max_pairwise_z", dbl), ("max_stationarity_z", dbl),
                ("min_ess", dbl)]


class DevSu3Params(C.Structure):
    _fields_ = [("beta", dbl), ("angle", dbl), ("hits", i32), ("reunitarize", i32),
                ("key", u64), ("sweep", u64)]


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = C.CDLL(LIB_PATH)
    dp = P(dbl)
    ip = P(i64)
    sig = {
        "ymo_mix64": (u64, [u64]),
        "ymo_splitmix_next": (u64, [P(u64)]),
        "ymo_splitmix_unit": (dbl, [P(u64)]),
        "ymo_derive_chain_seed": (u64, [u64, u64]),
        "ymo_derive_su3_chain_seed": (u64, [u64, u64]),
        "ymo_proposal_angle_from_u64": (C.c_int, [u64, dbl, dp]),
        "ymo_su3_from_algebra": (C.c_int, [dp, dp]),
        "ymo_su3_gell_mann": (C.c_int, [C.c_int, dp]),
        "ymo_su3_checked_mul": (C.c_int, [dp, dp, dp]),
        "ymo_su3_matrix_residuals": (None, [dp, dp, dp]),
        "ymo_su3_structure_f": (C.c_int, [C.c_int, C.c_int, C.c_int, dp]),
        "ymo_su3_structure_d": (C.c_int, [C.c_int, C.c_int, C.c_int, dp]),
        "ymo_su3_casimirs": (None, [dp, dp]),
        "ymo_su3_identity": (C.c_int, [ip, dp]),
        "ymo_su3_wilson_action": (C.c_int, [ip, dp, dbl, dp]),
        "ymo_su3_average_plaquette": (C.c_int, [ip, dp, dp]),
        "ymo_su3_polyakov_loop": (C.c_int, [ip, dp, ip, C.c_int, dp, dp]),
        "ymo_su3_gauge_transform": (C.c_int, [ip, dp, dp, dp]),
        "ymo_su3_max_residuals": (None, [ip, dp, dp, dp]),
        "ymo_su3_traceless_cycle_link": (None, [dp]),
        "ymo_su3_deterministic_site_gauge": (C.c_int, [ip, dp]),
        "ymo_su3_local_link_action": (C.c_int, [ip, dp, ip, C.c_int, dbl, dp]),
        "ymo_su3_local_action_difference": (C.c_int, [ip, dp, ip, C.c_int, dp, dbl, dp]),
        "ymo_su3_metropolis_sweep": (C.c_int, [ip, dp, dbl, dbl, P(u64), P(u64), P(u64)]),
        "ymo_su3_run_chain": (C.c_int, [P(Su3ChainConfig), dp, P(ChainDiagnostics), P(Su3ChainReport), dp]),
        "ymo_su3_path_holonomy": (C.c_int, [ip, dp, ip, P(i32), i64, dp, ip]),
        "ymo_su3_rect_wilson_loop": (C.c_int, [ip, dp, ip, C.c_int, C.c_int, i64, i64, dp, dp]),
        "ymo_su3_a1pp_basis": (C.c_int, [ip, dp, i64, dp, dp, ip, ip]),
        "ymo_su3_stout_smear": (C.c_int, [ip, dp, dp, dbl, i64, i64]),
        "ymo_su3_max_link_residual": (dbl, [ip, dp, dp]),
        "ymo_su3_blocked_extents": (C.c_int, [ip, ip]),
        "ymo_su3_spatial_block": (C.c_int, [ip, dp, dp]),
        "ymo_su2_from_axis_angle": (C.c_int, [dp, dbl, dp]),
        "ymo_su2_mul": (None, [dp, dp, dp]),
        "ymo_su2_identity": (C.c_int, [ip, dp]),
        "ymo_su2_plaquette": (C.c_int, [ip, dp, ip, C.c_int, C.c_int, dp]),
        "ymo_su2_wilson_action": (C.c_int, [ip, dp, dbl, dp]),
        "ymo_su2_path_holonomy": (C.c_int, [ip, dp, ip, P(i32), i64, dp, ip]),
        "ymo_su2_rect_wilson_loop": (C.c_int, [ip, dp, ip, C.c_int, C.c_int, i64, i64, dp]),
        "ymo_su2_polyakov_loop": (C.c_int, [ip, dp, ip, C.c_int, dp]),
        "ymo_su2_staple": (C.c_int, [ip, dp, ip, C.c_int, dp]),
        "ymo_su2_local_action_difference": (C.c_int, [ip, dp, ip, C.c_int, dp, dbl, dp]),
        "ymo_su2_max_unit_residual": (dbl, [ip, dp]),
        "ymo_su2_deterministic_site_gauge": (C.c_int, [ip, dp]),
        "ymo_su2_heatbath_sweep": (C.c_int, [ip, dp, dbl, P(u64), i64, P(u64), P(u64)]),
        "ymo_su2_overrelaxation_sweep": (C.c_int, [ip, dp]),
        "ymo_su2_average_plaquette": (C.c_int, [ip, dp, dp]),
        "ymo_su2_run_chain": (C.c_int, [P(Su2ChainConfig), dp, P(ChainDiagnostics), P(Su2ChainReport), dp]),
        "ymo_su2_config_encode": (i64, [ip, dp, C.c_char_p, i64]),
        "ymo_diagnostics": (C.c_int, [dp, i64, i64, P(ChainDiagnostics), dp]),
        "ymo_evaluate": (C.c_int, [dp, i64, dp, i64, dbl, dbl, dbl, P(EfficiencyReport)]),
        "ymo_stationarity": (C.c_int, [dp, i64, i64, P(StationarityReport)]),
        "ymo_pool_chains": (C.c_int, [dp, i64, i64, i64, i64, P(PoolReport)]),
        "ymo_periodic_connected": (C.c_int, [dp, i64, i64, dp, dp, dp]),
        "ymo_connected_from_samples": (C.c_int, [dp, i64, i64, i64, dp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _dev_signatures(L)
    _lib = L
    return L


def _dev_signatures(L):
    dp = P(dbl)
    ip = P(i64)
    sig = {
        "ymd_philox4x32_10": (None, [P(C.c_uint32), P(C.c_uint32), P(C.c_uint32)]),
        "ymd_exp_neg": (dbl, [dbl]),
        "ymd_sincos_small": (None, [dbl, dp, dp]),
        "ymd_sincos_short": (None, [dbl, dp, dp]),
        "ymd_log": (dbl, [dbl]),
        "ymd_log1p": (dbl, [dbl]),
        "ymd_expm1": (dbl, [dbl]),
        "ymd_sincos_turns": (None, [dbl, dp, dp]),
        "ymd_slots_per_parity": (i64, [ip]),
        "ymd_su3_to_device": (C.c_int, [ip, dp, C.c_void_p, C.c_int]),
        "ymd_su3_from_device": (C.c_int, [ip, C.c_void_p, C.c_int, dp]),
        "ymd_su2_to_device": (C.c_int, [ip, dp, C.c_void_p, C.c_int]),
        "ymd_su2_from_device": (C.c_int, [ip, C.c_void_p, C.c_int, dp]),
        "ymd_su3_hot_start": (C.c_int, [ip, C.c_void_p, C.c_int, u64]),
        "ymd_su2_hot_start": (C.c_int, [ip, C.c_void_p, C.c_int, u64]),
        "ymd_su3_metropolis_sweep": (C.c_int, [ip, C.c_void_p, C.c_int, P(DevSu3Params), P(u64), P(u64)]),
        "ymd_su3_staple": (C.c_int, [ip, C.c_void_p, C.c_int, ip, C.c_int, dp]),
        "ymd_su3_average_plaquette": (C.c_int, [ip, C.c_void_p, C.c_int, dp]),
        "ymd_su2_heatbath_sweep": (C.c_int, [ip, C.c_void_p, C.c_int, dbl, u64, u64, i64, P(u64), P(u64)]),
        "ymd_su2_overrelax_sweep": (C.c_int, [ip, C.c_void_p, C.c_int]),
        "ymd_su2_heatbath_link": (C.c_int, [dp, dbl, u64, C.c_uint32, C.c_int, u64, i64, dp, ip]),
        "ymd_su2_heatbath_samples": (C.c_int, [dp, dbl, u64, C.c_int, u64, i64, i64, dp, ip]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name, None)
        if f is None:
            continue
        f.restype = res
        f.argtypes = args


def dptr(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(P(dbl))


def iptr(a):
    a = np.ascontiguousarray(a, dtype=np.int64)
    return a, a.ctypes.data_as(P(i64))


def ext(n):
    a = (i64 * 4)(*[int(v) for v in n])
    return a


def su3_identity(n):
    links = np.zeros(int(np.prod(n)) * 4 * 18, dtype=np.float64)
    st = lib().ymo_su3_identity(ext(n), dptr(links))
    assert st == 0
    return links


def su2_identity(n):
    links = np.zeros(int(np.prod(n)) * 4 * 4, dtype=np.float64)
    st = lib().ymo_su2_identity(ext(n), dptr(links))
    assert st == 0
    return links


def link_offset(n, x, mu, ncomp):
    site = ((x[0] * n[1] + x[1]) * n[2] + x[2]) * n[3] + x[3]
    return (site * 4 + mu) * ncomp


def su3_from_algebra(coef):
    c = np.ascontiguousarray(coef, dtype=np.float64)
    out = np.zeros(18)
    st = lib().ymo_su3_from_algebra(dptr(c), dptr(out))
    assert st == 0, st
    return out


def mat_from18(a):
    a = np.asarray(a, dtype=np.float64).reshape(3, 3, 2)
    return a[..., 0] + 1j * a[..., 1]


def mat_to18(m):
    m = np.asarray(m, dtype=np.complex128)
    out = np.empty((3, 3, 2))
    out[..., 0] = m.real
    out[..., 1] = m.imag
    return out.reshape(18)
