"""Statistical agreement with the reference and FP32-storage error bounds.

Level (c): GPU chains with the reference's V2 schedule (4^4, beta = 6, cold
start, 192 + 128 x 2 sweeps, 8 proposals per link, angle 0.35;
su3_confirmation_reference.rs) against the reference's frozen pooled
plaquette 0.5992920042685975 +- 0.0008268453534905086
(su3-plaquette-baseline-v2.json), with FP64 link residuals <= 1e-12.
FP32 storage: observables of an FP32 copy of an FP64 configuration against
FP64 (SU(2) heat-bath/over-relaxation lattices and SU(3) with stout).
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

V2_MEAN, V2_SE = 0.5992920042685975, 0.0008268453534905086


@pytest.mark.parametrize("prec", [64, 32])
def test_level_c_chains_agree_with_reference_v2(ym, prec):
    n = (4, 4, 4, 4)
    series, acc = [], []
    for chain in range(4):
        key = ym.derive_su3_chain_seed(0x5EED, chain)
        lat = ym.Su3Lattice4(n, prec)
        meas, tp, ta = [], 0, 0
        for s in range(192 + 256):
            p, a = lat.metropolis_sweeps(6.0, 0.35, key, s, 1, 8)
            tp += p
            ta += a
            if s >= 192 and (s - 192) % 2 == 1:
                meas.append(lat.average_plaquette())
        series.append(meas)
        acc.append(ta / tp)
        if prec == 64:
            u, d = lat.link_residuals()
            assert u <= 1e-12 and d <= 1e-12
    pool = ym.stats.pool_chains(np.array(series), 24)
    z = abs(pool.pooled_mean - V2_MEAN) / math.hypot(pool.pooled_se, V2_SE)
    assert z < 2.0, (pool.pooled_mean, pool.pooled_se, z)
    assert abs(np.mean(acc) - 0.8675) < 0.005, acc


def test_su2_fp32_storage_errors(ym):
    r = ym.precision.su2_precision(8, 2.3, therm=10, chain=20, loops=4)
    assert r["plaquette_abs_err"] < 2e-6
    assert r["wilson_loops_max_abs_err"] < 2e-6
    assert r["polyakov_abs_err"] < 2e-6
    assert r["unit_residual_fp32_storage"] < 1e-6 and r["chain_unit_residual_fp32"] < 1e-6


def test_su3_fp32_storage_errors(ym):
    r = ym.precision.su3_precision(8, therm=20, stout_levels=3)
    assert r["plaquette_abs_err"] < 2e-6
    assert r["wilson_loops_4x4_max_abs_err"] < 2e-6
    assert r["stout3_plaquette_abs_err"] < 5e-6 and r["stout3_a1pp_max_abs_err"] < 5e-6
    assert r["unit_residual_fp32_storage"] < 1e-6 and r["stout_unit_residual_fp32"] < 1e-6
