"""FP32-storage error against FP64 on identical configurations.

Links stored in FP32 (SU(2): 4 reals, SU(3): 18 reals) are multiplied in
FP32; the accept test, the new link, reunitarisation and every reduction are
FP64.  This module measures what that costs on the observables of the
reference (plaquette su2_ensemble.rs:348 / su3_ensemble.rs:228, rectangular
Wilson loops lib.rs:438-481, Polyakov loop lib.rs:483-505 /
su3_lattice.rs:185-212, stout smearing glueball_smearing.rs:257): an FP64
chain is thermalised, its configuration is copied into an FP32 lattice, and
every observable is evaluated on both.  Chain-level agreement (FP32 and FP64
chains continued with the same key) is reported as the difference of the
plaquette means in units of their combined standard error.
"""
import numpy as np

from .lattice import (Lattice4, Su3Lattice4, derive_chain_seed, derive_su3_chain_seed, diagnostics,
                      spatial_stout_smear)


def _errors(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b))), float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


def su2_precision(extent, beta, therm=50, chain=100, loops=8, or_per_hb=3, seed=0):
    n = (extent,) * 4
    key = derive_chain_seed(0x5EED_0002, seed)
    l64 = Lattice4(n, 64)
    l64.hot_start(key)
    for s in range(therm):
        l64.heatbath_sweeps(beta, key, s, 1)
        l64.overrelax_sweeps(or_per_hb)
    l32 = Lattice4(n, 32)
    l32.upload(l64.download())
    out = {"workload": f"SU(2) {extent}^4 beta={beta}", "thermalisation": f"{therm} x (1 HB + {or_per_hb} OR), FP64"}
    p64, p32 = l64.average_plaquette(), l32.average_plaquette()
    out["plaquette_fp64"] = p64
    out["plaquette_abs_err"] = abs(p32 - p64)
    w64, _ = l64.wilson_loops(loops, loops, 0)
    w32, _ = l32.wilson_loops(loops, loops, 0)
    out["wilson_loops_max_abs_err"], out["wilson_loops_max_rel_err"] = _errors(w32, w64)
    out["wilson_8x8_fp64"] = float(w64[loops - 1, loops - 1])
    pl64, a64 = l64.polyakov_loop(3)
    pl32, a32 = l32.polyakov_loop(3)
    out["polyakov_abs_err"] = abs(pl32 - pl64)
    out["polyakov_modulus_abs_err"] = abs(a32 - a64)
    out["unit_residual_fp32_storage"] = l32.link_residuals()[0]
    # chains continued from the same configuration with the same key
    series = {32: [], 64: []}
    t32 = 0.0
    for prec, lat in ((32, l32), (64, l64)):
        lat.timer_begin()
        for s in range(chain):
            lat.heatbath_sweeps(beta, key, therm + s, 1)
            lat.overrelax_sweeps(or_per_hb)
            series[prec].append(lat.average_plaquette())
        ms = lat.timer_end()
        if prec == 32:
            t32 = ms
    d32 = diagnostics(series[32], max(1, chain // 5))[0]
    d64 = diagnostics(series[64], max(1, chain // 5))[0]
    out["chain_plaquette_fp32"] = d32.mean
    out["chain_plaquette_fp64"] = d64.mean
    out["chain_plaquette_z"] = abs(d32.mean - d64.mean) / max(1e-300, np.hypot(d32.standard_error,
                                                                               d64.standard_error))
    out["chain_unit_residual_fp32"] = l32.link_residuals()[0]
    out["fp32_device_ms_per_update_cycle"] = t32 / chain
    return out


def su3_precision(extent, beta=6.0, therm=100, stout_levels=10, rho=0.1, seed=0, hits=8):
    n = (extent,) * 4
    key = derive_su3_chain_seed(0x5EED_0003, seed)
    l64 = Su3Lattice4(n, 64)
    l64.metropolis_sweeps(beta, 0.35, key, 0, therm, hits)
    l32 = Su3Lattice4(n, 32)
    l32.upload(l64.download())
    out = {"workload": f"SU(3) {extent}^4 beta={beta}", "thermalisation": f"{therm} Metropolis sweeps, FP64"}
    p64, p32 = l64.average_plaquette(), l32.average_plaquette()
    out["plaquette_fp64"] = p64
    out["plaquette_abs_err"] = abs(p32 - p64)
    w64, _ = l64.wilson_loops(4, 4, 0)
    w32, _ = l32.wilson_loops(4, 4, 0)
    out["wilson_loops_4x4_max_abs_err"], _ = _errors(w32, w64)
    s64 = spatial_stout_smear(l64, rho, stout_levels)
    s32 = spatial_stout_smear(l32, rho, stout_levels)
    a64 = s64.timeslice_a1pp()
    a32 = s32.timeslice_a1pp()
    out[f"stout{stout_levels}_plaquette_fp64"] = s64.average_plaquette()
    out[f"stout{stout_levels}_plaquette_abs_err"] = abs(s32.average_plaquette() - s64.average_plaquette())
    out[f"stout{stout_levels}_a1pp_max_abs_err"] = max(_errors(a32[0], a64[0])[0], _errors(a32[1], a64[1])[0])
    u, d = l32.link_residuals()
    out["unit_residual_fp32_storage"], out["det_residual_fp32_storage"] = u, d
    u, d = s32.link_residuals()
    out["stout_unit_residual_fp32"] = u
    return out
