"""The benchmark workloads at their full sizes.

* SU(3) 24^4 and 32^4, beta = 6.0, 10 proposals per link (generators
  0..7, 0, 1), FP32 storage, with the sub-pass mapping the autotuner picks:
  one sweep bit for bit against the device-order replay (oracle/dev_exact.c,
  threaded over the links of each sub-pass), and 10 spatial stout levels at
  rho = 0.1 in FP32 against an FP64 twin holding the same links.
* SU(2) 48^4 at beta = 2.3 and 2.5: one heat-bath + over-relaxation Markov
  sweep bit for bit against the replay; at 32^4 the averaged Wilson loops up
  to 8 x 8 and the Polyakov loops in FP32 against FP64.
* The 256-chain production set (64 each of 4^3x10, 6^3x12, 8^3x16,
  12^3x24, beta = 5.7, keys derive_su3_chain_seed(master, id)): one
  hipGraph-captured cycle, then two chains of every geometry against single
  lattices swept alone (links, accepts, plaquette, A1++ on raw and 4 stout
  levels).
* The learned proposal at 16^3x32, beta = 6.0: trained on pairs of a plain
  chain at that volume, BF16 inference; deterministic sweeps, acceptance, and
  the plaquette of the learned chain within the plain chain's range.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu


def _dev_buffer(n, nc, prec):
    vh = O.lib().ymd_slots_per_parity(O.ext(n))
    return np.zeros(8 * nc * vh, dtype=np.float32 if prec == 32 else np.float64)


def _dev_to_aos(n, buf, prec, nc):
    aos = np.zeros(int(np.prod(n)) * 4 * nc)
    f = O.lib().ymd_su3_from_device if nc == 18 else O.lib().ymd_su2_from_device
    assert f(O.ext(n), buf.ctypes.data_as(C.c_void_p), prec, O.dptr(aos)) == 0
    return aos


@pytest.mark.parametrize("L", [24, 32])
def test_su3_ten_hit_sweep_bitwise_at_full_size(ym, L):
    n = (L, L, L, L)
    key = ym.derive_su3_chain_seed(0x2C02198105010002, 0)
    print(f"[fullsize] SU(3) {n} FP32 autotune + 1 sweep, 10 hits", flush=True)
    lat = ym.Su3Lattice4(n, 32)
    lat.hot_start(key ^ 0xF00D)
    best, cand = lat.autotune(6.0, 0.35, hits=10, reps=1)
    assert cand
    gp, ga = lat.metropolis_sweeps(6.0, 0.35, key, 3, 1, 10)
    buf = _dev_buffer(n, 18, 32)
    assert O.lib().ymd_su3_hot_start(O.ext(n), buf.ctypes.data_as(C.c_void_p), 32, key ^ 0xF00D) == 0
    params = O.DevSu3Params(beta=6.0, angle=0.35, hits=10, reunitarize=1, key=key, sweep=3)
    cp, ca = C.c_uint64(), C.c_uint64()
    assert O.lib().ymd_su3_metropolis_sweep(O.ext(n), buf.ctypes.data_as(C.c_void_p), 32, C.byref(params),
                                           C.byref(cp), C.byref(ca)) == 0
    assert (gp, ga) == (cp.value, ca.value), (best, gp, ga, cp.value, ca.value)
    assert gp == 10 * 4 * L ** 4
    assert np.array_equal(lat.download(), _dev_to_aos(n, buf, 32, 18)), best


@pytest.mark.parametrize("L", [24, 32])
def test_su3_ten_stout_levels_fp32_vs_fp64(ym, L):
    n = (L, L, L, L)
    print(f"[fullsize] SU(3) {n} 10 stout levels FP32 vs FP64", flush=True)
    lat = ym.Su3Lattice4(n, 32)
    key = ym.derive_su3_chain_seed(0x2C02198105010002, 7)
    lat.metropolis_sweeps(6.0, 0.35, key, 0, 20, 10)
    twin = ym.Su3Lattice4(n, 64)
    twin.upload(lat.download(), checked=False)  # the FP32 values as they are
    s32 = ym.spatial_stout_smear(lat, 0.1, 10)
    s64 = ym.spatial_stout_smear(twin, 0.1, 10)
    p32, p64 = s32.average_plaquette(), s64.average_plaquette()
    assert abs(p32 - p64) < 2e-6, (p32, p64)
    assert p64 > lat.average_plaquette() + 0.2  # smearing raises the plaquette
    a32, r32 = s32.timeslice_a1pp()
    a64, r64 = s64.timeslice_a1pp()
    assert np.abs(a32 - a64).max() < 5e-6 and np.abs(r32 - r64).max() < 5e-6
    u, d = s32.link_residuals()
    assert u < 2e-6 and d < 2e-6


@pytest.mark.parametrize("beta", [2.3, 2.5])
def test_su2_markov_sweep_bitwise_at_48(ym, beta):
    n = (48, 48, 48, 48)
    key = ym.derive_chain_seed(0x2B00201211160101, int(beta * 10))
    print(f"[fullsize] SU(2) {n} FP32 beta={beta} HB + OR against the replay", flush=True)
    lat = ym.Lattice4(n, 32)
    lat.hot_start(key ^ 0xABC)
    gp, gu = lat.markov_sweeps(beta, key, 0, 1, 1)
    buf = _dev_buffer(n, 4, 32)
    assert O.lib().ymd_su2_hot_start(O.ext(n), buf.ctypes.data_as(C.c_void_p), 32, key ^ 0xABC) == 0
    cp, cu = C.c_uint64(), C.c_uint64()
    assert O.lib().ymd_su2_heatbath_sweep(O.ext(n), buf.ctypes.data_as(C.c_void_p), 32, beta, key, 0, 10_000,
                                          C.byref(cp), C.byref(cu)) == 0
    assert O.lib().ymd_su2_overrelax_sweep(O.ext(n), buf.ctypes.data_as(C.c_void_p), 32) == 0
    assert (gp, gu) == (cp.value, cu.value)
    assert gu == 4 * 48 ** 4
    assert np.array_equal(lat.download(), _dev_to_aos(n, buf, 32, 4))


def test_su2_wilson_8x8_and_polyakov_fp32_vs_fp64_at_32(ym):
    n = (32, 32, 32, 32)
    print("[fullsize] SU(2) 32^4 W(R,T <= 8) and Polyakov FP32 vs FP64", flush=True)
    l64 = ym.Lattice4(n, 64)
    key = ym.derive_chain_seed(0x2B00201211160101, 77)
    l64.hot_start(key)
    l64.markov_sweeps(2.3, key, 0, 10, 3)
    l32 = ym.Lattice4(n, 32)
    l32.upload(l64.download())
    re64, _ = l64.wilson_loops(8, 8, 0)
    re32, _ = l32.wilson_loops(8, 8, 0)
    assert np.abs(re32 - re64).max() < 2e-6
    assert re64[0, 0] > 0.55 and re64[3, 3] < re64[1, 1] < re64[0, 0]  # loops decay with their area
    for d in range(4):
        p64, a64 = l64.polyakov_loop(d)
        p32, a32 = l32.polyakov_loop(d)
        assert abs(p32 - p64) < 2e-6 and abs(a32 - a64) < 2e-6


def test_production_256_chain_cycle_against_single_lattices(ym):
    geoms = ym.ensemble.PRODUCTION_GEOMETRIES
    master = ym.ensemble.PRODUCTION_MASTER_SEED
    ids = list(range(256))
    extents = [geoms[i // 64] for i in ids]
    print("[fullsize] 256-chain production set, one graph cycle", flush=True)
    ens = ym.Ensemble.from_chain_ids(extents, ids, master, 32)
    res = ens.run_cycles(5.7, 0.35, hits=8, sweeps_per_cycle=2, n_cycles=1, stout_levels=4, rho=0.1,
                         use_graph=True)
    assert ens.sweep_index() == 2
    for c in (0, 63, 64, 127, 128, 191, 192, 255):
        lat = ym.Su3Lattice4(extents[c], 32)
        _, acc = lat.metropolis_sweeps(5.7, 0.35, ym.derive_su3_chain_seed(master, c), 0, 2, 8)
        assert int(res.accepts[c]) == acc, c
        assert np.array_equal(ens.download_chain(c), lat.download()), c
        assert res.plaquette[0, c] == lat.average_plaquette(), c
        ops = res.ops[0][c]
        pl, rc = lat.timeslice_a1pp()
        assert np.array_equal(ops[0, :, 0], pl) and np.array_equal(ops[0, :, 1], rc), c
        for lev in (1, 4):
            sm = ym.spatial_stout_smear(lat, 0.1, lev)
            pl, rc = sm.timeslice_a1pp()
            assert np.array_equal(ops[lev, :, 0], pl) and np.array_equal(ops[lev, :, 1], rc), (c, lev)


def test_learned_proposal_at_16x3x32(ym):
    n = (16, 16, 16, 32)
    key = 0x5EED_1EA2_1632
    print("[fullsize] learned proposal 16^3x32, BF16 inference", flush=True)
    plain = ym.Su3Lattice4(n, 32)
    plain.metropolis_sweeps(6.0, 0.35, key, 0, 300, 8)
    m = ym.learned.Mlp(seed=21, dataset_capacity=4 * 4 * 16 ** 3 * 32)
    work = plain.clone()
    for k in range(4):
        work.metropolis_sweeps(6.0, 0.35, key, 300 + 5 * k, 5, 8)
        m.collect_pairs(work)
    losses, _ = m.train(epochs=3, batch=8192, lr=3e-3)
    assert losses[-1] < losses[0]
    a, b = plain.clone(), plain.clone()
    ca = ym.learned.learned_sweeps(a, m, 6.0, key, 1000, 20, metro_hits=0, bf16=True)
    cb = ym.learned.learned_sweeps(b, m, 6.0, key, 1000, 20, metro_hits=0, bf16=True)
    assert np.array_equal(a.download(), b.download())
    assert (ca.learned_proposals, ca.learned_accepts) == (cb.learned_proposals, cb.learned_accepts)
    assert ca.learned_proposals >= 0.99 * 20 * 4 * 16 ** 3 * 32
    assert ca.learned_accepts / ca.learned_proposals > 0.5
    # the learned chain keeps the plaquette in the physical range (the
    # statistical agreement of the two samplers is tested on 8^4 chains)
    pl = []
    for k in range(10):
        ym.learned.learned_sweeps(a, m, 6.0, key, 2000 + k, 1, metro_hits=0, bf16=True)
        pl.append(a.average_plaquette())
    assert 0.585 < np.mean(pl) < 0.61, np.mean(pl)
    u, d = a.link_residuals()
    assert u < 1e-6 and d < 1e-6
