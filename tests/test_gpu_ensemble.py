"""Batched ensembles (ym_ensemble_*) on the GPU.

Chain c of an ensemble must reproduce, bit for bit, a single lattice of the
same extents and precision swept with the same key (the single-lattice path
is itself pinned decision for decision by the device-order replay in the
oracle): links, accept counts, plaquette, A1++ timeslice operators on raw and
stout-smeared links.  The hipGraph-captured cycle must give the same results
as the eager launches on every replay.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GEOMS = [(4, 4, 4, 6), (6, 6, 6, 4), (4, 4, 4, 8), (6, 6, 6, 6)]
# 6 x 81 + 4 x 16 = 550 tiles of 256 links: the wave-specialised batched sweep
BIG = [(12, 12, 12, 24)] * 6 + [(8, 8, 8, 16)] * 4


def keys(ym, n):
    return [ym.derive_su3_chain_seed(0x3F0020260915_0001, c) for c in range(n)]


@pytest.mark.parametrize("prec", [32, 64])
def test_chains_match_single_lattices_bitwise(ym, prec):
    ks = keys(ym, len(GEOMS))
    ens = ym.Ensemble(GEOMS, ks, prec)
    res = ens.run_cycles(5.7, 0.35, hits=8, sweeps_per_cycle=3, n_cycles=1, stout_levels=2, rho=0.1,
                         use_graph=False)
    assert ens.sweep_index() == 3
    for c, (g, k) in enumerate(zip(GEOMS, ks)):
        lat = ym.Su3Lattice4(g, prec)
        _, acc = lat.metropolis_sweeps(5.7, 0.35, k, 0, 3, 8)
        assert int(res.accepts[c]) == acc
        assert np.array_equal(ens.download_chain(c), lat.download())
        assert res.plaquette[0, c] == lat.average_plaquette()
        ops = res.ops[0][c]
        pl, rc = lat.timeslice_a1pp()
        assert np.array_equal(ops[0, :, 0], pl) and np.array_equal(ops[0, :, 1], rc)
        for lev in (1, 2):
            sm = ym.spatial_stout_smear(lat, 0.1, lev)
            pl, rc = sm.timeslice_a1pp()
            assert np.array_equal(ops[lev, :, 0], pl) and np.array_equal(ops[lev, :, 1], rc)


@pytest.mark.parametrize("prec,hits", [(32, 8), (64, 8), (32, 10), (64, 10), (32, 3)])
def test_wave_specialised_batch_matches_single_lattices(ym, prec, hits):
    """From 512 tiles of 256 links per sub-pass the batched sweep runs the
    wave-specialised kernel (staples staged in LDS, tiles of all chains in one
    list; 8 and 10 hits have unrolled instances, other counts loop); every
    chain still equals a single lattice swept alone."""
    geoms = BIG
    ks = keys(ym, len(geoms))
    ens = ym.Ensemble(geoms, ks, prec)
    res = ens.run_cycles(6.0, 0.35, hits=hits, sweeps_per_cycle=2, n_cycles=1, stout_levels=0, rho=0.1,
                         use_graph=False)
    for c, (g, k) in enumerate(zip(geoms, ks)):
        lat = ym.Su3Lattice4(g, prec)
        _, acc = lat.metropolis_sweeps(6.0, 0.35, k, 0, 2, hits)
        assert int(res.accepts[c]) == acc, c
        assert np.array_equal(ens.download_chain(c), lat.download()), c
        assert res.plaquette[0, c] == lat.average_plaquette()


@pytest.mark.parametrize("geoms", [GEOMS, BIG], ids=["small", "550-tiles"])
def test_graph_replays_match_eager(ym, geoms):
    ks = keys(ym, len(geoms))
    runs = []
    for use_graph in (True, False):
        ens = ym.Ensemble(geoms, ks, 32)
        a = ens.run_cycles(5.7, 0.35, sweeps_per_cycle=2, n_cycles=3, stout_levels=2, use_graph=use_graph)
        # eager work and synchronisation between two graph calls
        mid = ens.sweeps(5.7, 0.35, 1)
        b = ens.run_cycles(5.7, 0.35, sweeps_per_cycle=1, n_cycles=2, stout_levels=1, use_graph=use_graph)
        links = [ens.download_chain(c) for c in range(len(geoms))]
        runs.append((a, mid, b, links, ens.sweep_index()))
        ens.close()
    (a1, m1, b1, l1, s1), (a2, m2, b2, l2, s2) = runs
    assert s1 == s2 == 2 * 3 + 1 + 2
    assert np.array_equal(a1.plaquette, a2.plaquette) and np.array_equal(b1.plaquette, b2.plaquette)
    for k in range(3):
        for c in range(len(geoms)):
            assert np.array_equal(a1.ops[k][c], a2.ops[k][c])
    for k in range(2):
        for c in range(len(geoms)):
            assert np.array_equal(b1.ops[k][c], b2.ops[k][c])
    assert np.array_equal(a1.accepts, a2.accepts) and np.array_equal(m1, m2) and np.array_equal(b1.accepts, b2.accepts)
    for x, y in zip(l1, l2):
        assert np.array_equal(x, y)
    # measurements of successive cycles differ (the chain moved)
    assert not np.array_equal(a1.plaquette[0], a1.plaquette[1])


def test_identity_ensemble_observables_closed_form(ym):
    ens = ym.Ensemble(GEOMS, keys(ym, len(GEOMS)), 64)
    res = ens.run_cycles(5.7, 0.35, sweeps_per_cycle=0, n_cycles=1, stout_levels=3)
    assert np.all(res.plaquette == 1.0)
    for c in range(len(GEOMS)):
        assert np.all(res.ops[0][c] == 1.0)
    assert np.all(res.accepts == 0)


def test_production_pipeline_small(ym):
    rep = ym.ensemble.production(chains_per_geometry=4, geometries=((4, 4, 4, 10), (6, 6, 6, 12)),
                                 stout_levels=2, sweeps=6)
    assert rep.plaquette.shape == (8,)
    assert np.all((rep.plaquette > 0.2) & (rep.plaquette < 1.0))
    assert np.all((rep.acceptance > 0.1) & (rep.acceptance < 0.99))
    for (g, lev, kind), ev in rep.correlators.items():
        assert ev.full.shape == (g[3],) and ev.jackknife.shape == (4, g[3])
        assert np.all(np.isfinite(ev.jackknife_error))
        # C(tau) is symmetric under tau -> T - tau (periodic estimator)
        assert np.allclose(ev.full[1:], ev.full[1:][::-1], rtol=0, atol=1e-14)
    # host analysis of each (geometry, level): results or the reference's error status,
    # the same as the analysis of the same samples run on their own
    assert set(rep.analysis) == {(g, lev) for g in rep.geometries for lev in range(3)}
    for (g, lev), an in rep.analysis.items():
        var = an["variational"]
        assert isinstance(var, ym.stats.VariationalCorrelator) and var.matrices.shape == (g[3], 2, 2)
        assert np.allclose(var.matrices[:, 0, 0], rep.correlators[(g, lev, 0)].full, rtol=0, atol=2e-12)
        for key in ("effective_mass", "fit", "projection"):
            r = an[key]
            assert not isinstance(r, dict) or r["status"] >= 100 or r["status"] in (54, 74, 75), (key, r)


def test_correlator_analysis_on_device_samples(ym):
    """A1++ samples of 64 small chains measured on the device feed the host
    analysis: the variational diagonal replays the jackknife correlators and
    the GEVP, when admitted, returns a normalised projection."""
    ens = ym.Ensemble.from_chain_ids([(4, 4, 4, 8)] * 64, list(range(64)), 0x5EED)
    try:
        res = ens.run_cycles(5.7, 0.35, sweeps_per_cycle=20, n_cycles=1, stout_levels=1)
    finally:
        ens.close()
    pl = np.array([res.ops[0][c][1, :, 0] for c in range(64)])
    rc = np.array([res.ops[0][c][1, :, 1] for c in range(64)])
    an = ym.ensemble.correlator_analysis(pl, rc)
    var = an["variational"]
    for k, ev in ((0, an["plaquette"]), (1, an["rectangle"])):
        assert np.max(np.abs(var.matrices[:, k, k] - ev.full)) <= 2e-12
    pr = an["projection"]
    if isinstance(pr, ym.stats.PrincipalProjection):
        assert pr.projected_correlator[1] == 1.0 and pr.principal_eigenvalue > 0
        c = pr.coefficients
        m1 = var.matrices[1]
        assert abs(c @ (0.5 * (m1 + m1.T)) @ c - 1.0) <= 1e-10
    else:
        assert pr["status"] in (144, 145, 146, 147)


def test_ensemble_rejects_bad_arguments(ym):
    with pytest.raises(ym.YmError) as e:
        ym.Ensemble([(4, 4, 4, 5)], [1], 32)
    assert e.value.status == 99
    ens = ym.Ensemble([(4, 4, 6, 6)], [1], 32)
    with pytest.raises(ym.YmError) as e:
        ens.run_cycles(5.7, 0.35, sweeps_per_cycle=1, stout_levels=1)
    assert e.value.status == 52
    with pytest.raises(ym.YmError) as e:
        ens.run_cycles(13.0, 0.35, sweeps_per_cycle=1, measure=False)
    assert e.value.status == 30
