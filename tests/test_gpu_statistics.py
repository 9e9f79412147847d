"""Statistical agreement with the reference and FP32-storage error bounds.

Level (c): GPU chains with the reference's V2 schedule (4^4, beta = 6, cold
start, 192 + 128 x 2 sweeps, 8 proposals per link, angle 0.35;
su3_confirmation_reference.rs) against the reference's frozen pooled
plaquette 0.5992920042685975 +- 0.0008268453534905086
(su3-plaquette-baseline-v2.json), with FP64 link residuals <= 1e-12.
SU(2): device heat-bath + over-relaxation chains against reference-order
CPU chains of the same job and the published 10^4 beta = 2.3 pooled mean.
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


def _jackknife_over_chains(means):
    m = np.asarray(means, dtype=np.float64)
    n = m.size
    reps = (m.sum() - m) / (n - 1)
    return float(m.mean()), float(math.sqrt((n - 1) / n * np.sum((reps - reps.mean()) ** 2)))


def test_level_c_against_reference_chains_in_the_same_job(ym):
    """Level (c) as stated: fresh reference-order CPU chains (the oracle's
    restatement of run_chain, su3_ensemble.rs:408-480: lexicographic order,
    SplitMix64 stream, checked FP64 products) run in this job beside FP32 GPU
    checkerboard chains, same schedule (4^4, beta = 6, cold start, 128
    thermalisation sweeps + 32 measurements 2 sweeps apart; the CPU chains
    cost about 0.2 s per sweep, so 8 of them run in parallel threads),
    independent seeds.  Pooled plaquettes agree within 2 sigma of the
    jackknife over chains; GPU acceptance matches the CPU's."""
    import ctypes as C
    import threading

    import oracle_lib as O

    n, therm, meas, stride = (4, 4, 4, 4), 128, 32, 2
    master = 0x3C0FFEE000000001
    ncpu = 8
    cpu = [None] * ncpu

    def run(i):
        cfg = O.Su3ChainConfig()
        cfg.extents[:] = list(n)
        cfg.beta, cfg.master_seed, cfg.chain_id = 6.0, master, 100 + i
        cfg.thermalization_sweeps, cfg.measurements, cfg.sweeps_between_measurements = therm, meas, stride
        cfg.proposal_angle, cfg.autocorrelation_max_lag = 0.35, 24
        m = np.zeros(meas)
        d, r = O.ChainDiagnostics(), O.Su3ChainReport()
        assert O.lib().ymo_su3_run_chain(C.byref(cfg), O.dptr(m), C.byref(d), C.byref(r), None) == 0
        cpu[i] = (m, r.accepts / r.proposals)

    threads = [threading.Thread(target=run, args=(i,)) for i in range(ncpu)]
    for t in threads:
        t.start()
    gpu, acc = [], []
    for chain in range(16):
        key = ym.derive_su3_chain_seed(master, 200 + chain)
        lat = ym.Su3Lattice4(n, 32)
        meas_s, tp, ta = [], 0, 0
        for s in range(therm + meas * stride):
            p, a = lat.metropolis_sweeps(6.0, 0.35, key, s, 1, 8)
            tp, ta = tp + p, ta + a
            if s >= therm and (s - therm) % stride == stride - 1:
                meas_s.append(lat.average_plaquette())
        gpu.append(np.mean(meas_s))
        acc.append(ta / tp)
    for t in threads:
        t.join()
    gm, gse = _jackknife_over_chains(gpu)
    cm, cse = _jackknife_over_chains([c[0].mean() for c in cpu])
    z = abs(gm - cm) / math.hypot(gse, cse)
    assert z < 2.0, (gm, gse, cm, cse, z)
    assert abs(np.mean(acc) - np.mean([c[1] for c in cpu])) < 0.004
    assert 0.56 <= cm <= 0.63  # the reference's external envelope (su3_benchmark.rs:23-28)


def _su3_fp32_gpu_vs_fp64_cpu(ym, L, equil, chains, sweeps, hits=10):
    """Level (c) at the benchmark lattice: one configuration equilibrated on
    the GPU (FP32, `equil` sweeps from cold), then `chains` FP32 GPU chains
    and `chains` FP64 CPU chains (the device-order replay with FP64 link
    storage, threaded over each sub-pass) continue from it under their own
    keys for `sweeps` sweeps, the plaquette after every sweep.  Returns the
    per-chain mean plaquettes of both sides."""
    import ctypes as C

    import oracle_lib as O

    n = (L,) * 4
    master = 0x3C0FFEE000003200 + L
    lat = ym.Su3Lattice4(n, 32)
    lat.metropolis_sweeps(6.0, 0.35, ym.derive_su3_chain_seed(master, 0), 0, equil, hits)
    gpu = []
    for c in range(chains):
        g = lat.clone()
        key = ym.derive_su3_chain_seed(master, 100 + c)
        pl = []
        for s in range(sweeps):
            g.metropolis_sweeps(6.0, 0.35, key, s, 1, hits)
            pl.append(g.average_plaquette())
        gpu.append(np.mean(pl))
    # the FP64 chains start from the FP64 Gram-Schmidt projection of the same links
    r64 = ym.Su3Lattice4(n, 64)
    r64.upload(lat.download(), checked=False)
    r64.reunitarize()
    vh = O.lib().ymd_slots_per_parity(O.ext(n))
    start = np.zeros(8 * 18 * vh)
    assert O.lib().ymd_su3_to_device(O.ext(n), O.dptr(r64.download()), start.ctypes.data_as(C.c_void_p), 64) == 0
    cpu = []
    for c in range(chains):
        buf = start.copy()
        key = ym.derive_su3_chain_seed(master, 200 + c)
        pl = []
        for s in range(sweeps):
            prm = O.DevSu3Params(beta=6.0, angle=0.35, hits=hits, reunitarize=1, key=key, sweep=s)
            a, b = C.c_uint64(), C.c_uint64()
            assert O.lib().ymd_su3_metropolis_sweep(O.ext(n), buf.ctypes.data_as(C.c_void_p), 64, C.byref(prm),
                                                   C.byref(a), C.byref(b)) == 0
            v = C.c_double()
            assert O.lib().ymd_su3_average_plaquette(O.ext(n), buf.ctypes.data_as(C.c_void_p), 64, C.byref(v)) == 0
            pl.append(v.value)
        cpu.append(np.mean(pl))
    return gpu, cpu


@pytest.mark.parametrize("L,equil,sweeps", [(16, 2000, 40), (32, 2000, 8)])
def test_su3_level_c_fp32_gpu_against_fp64_cpu_chains(ym, L, equil, sweeps):
    """Level (c) at 16^4 and at the benchmark's 32^4, beta = 6.0, 10 hits:
    8 FP32 GPU chains and 8 FP64 CPU chains continuing from one
    GPU-equilibrated configuration agree in the pooled plaquette within 2
    sigma of the jackknife over chains."""
    print(f"[level-c] SU(3) {L}^4: {equil} equilibration sweeps, 8 + 8 chains x {sweeps} sweeps", flush=True)
    gpu, cpu = _su3_fp32_gpu_vs_fp64_cpu(ym, L, equil, 8, sweeps)
    gm, gse = _jackknife_over_chains(gpu)
    cm, cse = _jackknife_over_chains(cpu)
    z = abs(gm - cm) / math.hypot(gse, cse)
    print(f"[level-c] {L}^4 GPU FP32 {gm:.7f} +- {gse:.7f}, CPU FP64 {cm:.7f} +- {cse:.7f}, z = {z:.2f}", flush=True)
    assert z < 2.0, (gm, gse, cm, cse, z)
    assert 0.58 < gm < 0.61


# docs/PUBLISHED_PLAQUETTE_BASELINE.md:109-126: two 10^4 chains at beta = 2.3,
# 40 + 48 Markov sweeps (1 heat-bath + 1 over-relaxation), pooled mean and SE
PUBLISHED_SU2_MEAN, PUBLISHED_SU2_SE = 0.6022919679191184, 0.0004173286094490


@pytest.mark.parametrize("prec", [64, 32])
def test_su2_level_c_against_reference_chains_and_published_mean(ym, prec):
    """Level (c) for SU(2): device heat-bath + over-relaxation chains
    (ym_su2_run_chain: checkerboard order, Philox) with the reference's
    published schedule (su2_benchmark.rs:21-35: 10^4, beta = 2.3, cold start,
    40 thermalisation + 48 measurements, one over-relaxation sweep per
    heat-bath sweep) against reference-order CPU chains run in this job (the
    oracle's restatement of su2_ensemble.rs:488-552: lexicographic order,
    SplitMix64, normalised FP64 quaternion products; about 7 s per chain, 8
    in parallel threads), and against the published pooled mean.  Means
    agree within 2 sigma of the jackknife over chains; heat-bath
    acceptances agree."""
    import ctypes as C
    import threading

    import oracle_lib as O

    n, therm, meas = (10, 10, 10, 10), 40, 48
    master = 0x2B00201211160101
    ncpu = 8
    cpu = [None] * ncpu

    def run(i):
        cfg = O.Su2ChainConfig()
        cfg.extents[:] = list(n)
        cfg.beta, cfg.master_seed, cfg.chain_id = 2.3, master, 1000 + i
        cfg.thermalization_sweeps, cfg.measurements, cfg.sweeps_between_measurements = therm, meas, 1
        cfg.overrelaxation_sweeps_per_heatbath, cfg.max_heatbath_attempts_per_link = 1, 10_000
        cfg.autocorrelation_max_lag, cfg.no_caps = 12, 0
        m = np.zeros(meas)
        d, r = O.ChainDiagnostics(), O.Su2ChainReport()
        assert O.lib().ymo_su2_run_chain(C.byref(cfg), O.dptr(m), C.byref(d), C.byref(r), None) == 0
        cpu[i] = (m, r.mean_heatbath_acceptance)

    threads = [threading.Thread(target=run, args=(i,)) for i in range(ncpu)]
    for t in threads:
        t.start()
    gpu, acc = [], []
    for chain in range(16):
        rep = ym.su2_run_chain(n, 2.3, master, 2000 + chain, therm, meas, 1, 1, 10_000, 12, precision=prec)
        assert rep.total_heatbath_sweeps == therm + meas and rep.total_overrelaxation_sweeps == therm + meas
        assert rep.heatbath_updates == (therm + meas) * 4 * 10**4
        assert rep.final_link_unit_residual <= (1e-12 if prec == 64 else 1e-6)
        gpu.append(rep.measurements.mean())
        acc.append(rep.mean_heatbath_acceptance)
    for t in threads:
        t.join()
    gm, gse = _jackknife_over_chains(gpu)
    cm, cse = _jackknife_over_chains([c[0].mean() for c in cpu])
    z = abs(gm - cm) / math.hypot(gse, cse)
    assert z < 2.0, (gm, gse, cm, cse, z)
    zp = abs(gm - PUBLISHED_SU2_MEAN) / math.hypot(gse, PUBLISHED_SU2_SE)
    print(f"[level-c] SU(2) 10^4 FP{prec} GPU {gm:.6f} +- {gse:.6f}, CPU reference-order {cm:.6f} +- {cse:.6f}, "
          f"z = {z:.2f}; published {PUBLISHED_SU2_MEAN:.6f}, z = {zp:.2f}; acceptance {np.mean(acc):.4f} vs "
          f"{np.mean([c[1] for c in cpu]):.4f}", flush=True)
    assert zp < 2.0, (gm, gse, zp)
    assert abs(np.mean(acc) - np.mean([c[1] for c in cpu])) < 0.003, (np.mean(acc), [c[1] for c in cpu])
    assert 0.59 <= gm <= 0.62  # the published source-resolution envelope (su2_benchmark.rs:23-24)


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
