"""Product host statistics (libym_amd.so, no GPU needed) against the
reference's golden vectors, closed forms, and the oracle bit for bit.

Reference: sampling_efficiency.rs:57-135 (test-vectors/sampling-efficiency-v1.json),
glueball_correlators.rs:111-172 (glueball-connected-correlator-phase3c-v1.json),
glueball_scalar_production.rs:383-470, glueball_effective_mass.rs:161-215,
su3_benchmark.rs:186-237, su3_confirmation.rs:105-263.
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

import oracle_lib as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_vectors")


def vec(name, vid=None):
    with open(os.path.join(GOLD, name + ".json")) as f:
        d = json.load(f)
    if vid is not None and "vectors" in d:
        return next(v for v in d["vectors"] if v["id"] == vid)
    return d


@pytest.fixture(scope="module")
def st(ym):
    return ym.stats


def test_evaluate_matches_reference_fixture(st):
    v = vec("sampling-efficiency-v1")
    c = v["cost"]
    r = st.evaluate(v["values"], v["max_lag"], c["setup"], c["thermalization"], c["sampling"])
    u = v["unit_weights"]
    assert abs(r.diagnostics.tau_int - u["tau_int"]) <= 1e-12
    assert abs(r.diagnostics.ess - u["autocorrelation_ess"]) <= 1e-12
    assert abs(r.combined_ess_proxy - u["combined_ess_proxy"]) <= 1e-12
    assert abs(r.sampling_cost_per_ess - u["sampling_cost_per_proxy"]) <= 1e-12
    assert abs(r.total_cost_per_ess - u["total_cost_per_proxy"]) <= 1e-12
    r = st.evaluate(v["values"], v["max_lag"], c["setup"], c["thermalization"], c["sampling"],
                    log_weights=np.log([1.0, 1.0, 1.0, 3.0]))
    w = v["weights_1_1_1_3"]
    assert abs(r.reweighting_efficiency - w["weight_efficiency"]) <= 1e-12
    assert abs(r.weight_only_ess - w["weight_only_ess"]) <= 1e-12
    assert abs(r.total_cost_per_ess - w["total_cost_per_proxy"]) <= 1e-12
    r = st.evaluate(v["correlated_values"], v["correlated_max_lag"])
    assert abs(r.diagnostics.tau_int - v["correlated_tau_int"]) <= 1e-12


def test_evaluate_rejections(st, ym):
    with pytest.raises(ym.YmError) as e:
        st.evaluate([1.0] * 5, 2)
    assert e.value.status == 85  # constant series
    with pytest.raises(ym.YmError) as e:
        st.evaluate([1.0, 2.0, 3.0], 3)
    assert e.value.status == 81
    with pytest.raises(ym.YmError) as e:
        st.evaluate([1.0, 2.0, 3.0], 1, sampling=0.0)
    assert e.value.status == 83
    with pytest.raises(ym.YmError) as e:
        st.evaluate(np.arange(4097.0), 1)
    assert e.value.status == 80


def shifted_defect_samples():
    """The phase-3C fixture: one cycle link at timeslice c of configuration c
    (glueball-connected-correlator-phase3c-v1.json), operator values from the
    oracle's A1++ basis (itself pinned by glueball-operators-phase3a-v1.json)."""
    L = O.lib()
    n = (3, 3, 3, 4)
    cyc = np.zeros(18)
    L.ymo_su3_traceless_cycle_link(O.dptr(cyc))
    pl, rc = np.zeros((4, 4)), np.zeros((4, 4))
    for c in range(4):
        links = O.su3_identity(n)
        off = O.link_offset(n, (0, 0, 0, c), 0, 18)
        links[off:off + 18] = cyc
        for t in range(4):
            p, r, pc, rcn = C.c_double(), C.c_double(), C.c_int64(), C.c_int64()
            assert L.ymo_su3_a1pp_basis(O.ext(n), O.dptr(links), t, C.byref(p), C.byref(r), C.byref(pc),
                                        C.byref(rcn)) == 0
            pl[c, t], rc[c, t] = p.value, r.value
    return pl, rc


def test_connected_correlators_match_reference_fixture(st):
    e = vec("glueball-connected-correlator-phase3c-v1", "shifted-defect-analytic-correlators")["expected"]
    pl, rc = shifted_defect_samples()
    for samples, key in ((pl, "plaquette"), (rc, "rectangle")):
        mean, raw, conn = st.periodic_connected(samples)
        assert abs(mean - e[key + "_vacuum_mean"]) <= 2e-12
        assert np.max(np.abs(raw - e[key + "_raw"])) <= 2e-12
        assert np.max(np.abs(conn - e[key + "_connected"])) <= 2e-12
        assert np.array_equal(st.connected_from_samples(samples), conn)


def test_jackknife_replicas_and_error_closed_form(st):
    rng = np.random.default_rng(7)
    x = rng.normal(size=(9, 6))
    ev = st.correlator_evidence(x)
    assert np.array_equal(ev.full, st.connected_from_samples(x))
    for c in range(9):
        # a delete-one replica is the estimator on the remaining rows
        _, _, conn = st.periodic_connected(np.delete(x, c, axis=0))
        assert np.max(np.abs(ev.jackknife[c] - conn)) <= 1e-15
    m = ev.jackknife.mean(axis=0)
    err = np.sqrt(8.0 / 9.0 * ((ev.jackknife - m) ** 2).sum(axis=0))
    assert np.allclose(ev.jackknife_mean, m, rtol=0, atol=1e-15)
    assert np.allclose(ev.jackknife_error, err, rtol=1e-13, atol=0)
    # for a linear statistic (the mean) the jackknife error is the textbook
    # standard error of the mean: sample std / sqrt(N)
    y = rng.normal(size=(12, 1))
    ev = st.correlator_evidence(y)
    mean_reps = np.array([np.delete(y, c).mean() for c in range(12)])
    jk = np.sqrt(11.0 / 12.0 * ((mean_reps - mean_reps.mean()) ** 2).sum())
    assert abs(jk - y.std(ddof=1) / np.sqrt(12)) <= 1e-14


def test_connected_correlator_rejections(st, ym):
    with pytest.raises(ym.YmError) as e:
        st.connected_from_samples(np.zeros((2, 3)), omitted=2)
    assert e.value.status == 75
    with pytest.raises(ym.YmError) as e:
        st.connected_from_samples(np.zeros((1, 3)), omitted=0)
    assert e.value.status == 70
    with pytest.raises(ym.YmError) as e:
        st.periodic_connected(np.array([[1.0, np.nan]]))
    assert e.value.status == 54


def test_statistics_bitwise_equal_to_oracle(st):
    L = O.lib()
    rng = np.random.default_rng(11)
    for n, lag in ((64, 8), (400, 40), (1000, 100)):
        x = np.cumsum(rng.normal(size=n)) * 0.01 + rng.normal(size=n)
        w = rng.normal(size=n) * 0.3
        r = st.evaluate(x, lag, 0.5, 1.5, 2.0, log_weights=w)
        ro = O.EfficiencyReport()
        assert L.ymo_evaluate(O.dptr(x), n, O.dptr(np.ascontiguousarray(w)), lag, 0.5, 1.5, 2.0, C.byref(ro)) == 0
        for a, b in ((r.diagnostics.tau_int, ro.diag.tau_int), (r.diagnostics.ess, ro.diag.ess),
                     (r.reweighting_efficiency, ro.reweighting_efficiency),
                     (r.total_cost_per_ess, ro.total_cost_per_ess)):
            assert a == b
        s = st.stationarity(x, lag // 2)
        so = O.StationarityReport()
        assert L.ymo_stationarity(O.dptr(x), n, lag // 2, C.byref(so)) == 0
        assert s.drift_z == so.drift_z and s.first_half_se == so.first_half_se
    series = rng.normal(size=(4, 200)) + np.arange(4)[:, None] * 0.01
    p = st.pool_chains(series, 20, 10)
    po = O.PoolReport()
    flat = np.ascontiguousarray(series.reshape(-1))
    assert L.ymo_pool_chains(O.dptr(flat), 4, 200, 20, 10, C.byref(po)) == 0
    assert p.pooled_mean == po.pooled_mean and p.pooled_se == po.pooled_se
    assert p.max_pairwise_z == po.max_pairwise_z and p.max_stationarity_z == po.max_stationarity_z
    assert list(p.chain_ess) == list(po.chain_ess)[:4]
    samples = rng.normal(size=(16, 10))
    _, raw, conn = st.periodic_connected(samples)
    m, raw2, conn2 = C.c_double(), np.zeros(10), np.zeros(10)
    flat = np.ascontiguousarray(samples.reshape(-1))
    assert L.ymo_periodic_connected(O.dptr(flat), 16, 10, C.byref(m), O.dptr(raw2), O.dptr(conn2)) == 0
    assert np.array_equal(raw, raw2) and np.array_equal(conn, conn2)
    for om in (-1, 0, 7, 15):
        c2 = np.zeros(10)
        assert L.ymo_connected_from_samples(O.dptr(flat), 16, 10, om, O.dptr(c2)) == 0
        assert np.array_equal(st.connected_from_samples(samples, None if om < 0 else om), c2)


def test_pooled_summary_closed_forms(st):
    # two independent chains: pooled SE = sqrt(se1^2 + se2^2) / 2, z = |dm| / sqrt(se1^2 + se2^2)
    rng = np.random.default_rng(3)
    s = rng.normal(size=(2, 300))
    p = st.pool_chains(s, 5)
    assert abs(p.pooled_mean - s.mean(axis=1).mean()) <= 1e-15
    assert abs(p.pooled_se - np.sqrt((p.chain_se ** 2).sum()) / 2) <= 1e-16
    assert abs(p.max_pairwise_z - abs(p.chain_mean[0] - p.chain_mean[1]) / np.sqrt((p.chain_se ** 2).sum())) <= 1e-14
    assert p.aggregate_ess == p.chain_ess.sum() and p.min_ess == p.chain_ess.min()
