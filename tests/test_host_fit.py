"""Host FP64 correlator analysis: effective mass, correlated fit, GEVP.

The product entry points (libym_amd.so: ym_cosh_effective_mass,
ym_correlated_scalar_fit, ym_variational_connected_correlator,
ym_principal_generalized_projection; no GPU needed) and the oracle's
restatement (oracle/ref_fit.py) are both checked against the reference's
golden vectors (glueball-effective-mass-phase3d-v1.json,
glueball-scalar-fit-phase3e-v1.json, glueball-scalar-variational-phase3j-v1.json)
and against closed forms the mathematics fixes; then against each other on
random inputs.

Reference: glueball_effective_mass.rs:161-268, glueball_scalar_fit.rs:156-699,
glueball_scalar_variational.rs:628-951.
"""
import json
import math
import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "oracle"))
import ref_fit as R  # noqa: E402

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_vectors")


def vec(name, vid):
    with open(os.path.join(GOLD, name + ".json")) as f:
        d = json.load(f)
    return next(v for v in d["vectors"] if v["id"] == vid)


@pytest.fixture(scope="module")
def st(ym):
    return ym.stats


def status_of(fn, *a):
    import ym_amd
    try:
        fn(*a)
    except ym_amd.YmError as e:
        return e.status
    except R.OracleError as e:
        return e.status
    return 0


# ------------------------------------------------------------ effective mass
def em_fixture():
    full = R.analytic_periodic_single_state(16, 0.5, 1.0)
    jk = [R.analytic_periodic_single_state(16, m, 1.0) for m in (0.48, 0.49, 0.50, 0.51, 0.52)]
    return full, jk


def test_effective_mass_golden_vectors(st):
    v = vec("glueball-effective-mass-phase3d-v1", "periodic-single-state-cosh-mass")
    j = vec("glueball-effective-mass-phase3d-v1", "delete-one-jackknife-propagation")
    full, jk = em_fixture()
    assert np.allclose(st.analytic_periodic_single_state(16, 0.5, 1.0), full, rtol=0, atol=0)
    s = st.cosh_effective_mass_with_jackknife(full, jk, v["analysis_lags"])
    o = R.cosh_effective_mass_with_jackknife(full, jk, v["analysis_lags"])
    tol = v["absolute_tolerance"]
    assert s.time_extent == 16 and s.jackknife_replicates == 5
    for i in range(len(v["analysis_lags"])):
        for got in (s.full_estimate[i], o[i][1]):
            assert abs(got - v["expected_effective_mass_each_lag"]) <= tol
        for got in (s.jackknife_mean[i], o[i][2]):
            assert abs(got - j["expected_jackknife_mean_each_lag"]) <= tol
        for got in (s.standard_error[i], o[i][3]):
            assert abs(got - j["expected_standard_error_each_lag"]) <= tol
    # closed form: sqrt((N-1)/N sum (m_i - 0.5)^2) = sqrt(0.8 * 0.001) for these five masses
    assert abs(s.standard_error[0] - math.sqrt(0.0008)) <= 2e-12
    a = vec("glueball-effective-mass-phase3d-v1", "amplitude-cancellation")
    f2 = st.analytic_periodic_single_state(16, a["full_mass"], a["full_amplitude"])
    jk2 = [st.analytic_periodic_single_state(16, m, a["replica_amplitude"]) for m in (0.48, 0.49, 0.5, 0.51, 0.52)]
    r = st.cosh_effective_mass_with_jackknife(f2, jk2, [a["analysis_lag"]])
    assert abs(r.full_estimate[0] - a["expected_effective_mass"]) <= a["absolute_tolerance"]


def test_effective_mass_golden_rejections(st):
    full, jk = em_fixture()
    for fn in (st.cosh_effective_mass_with_jackknife, R.cosh_effective_mass_with_jackknife):
        f = list(full)
        f[2] = 0.0
        assert status_of(fn, f, jk, [2]) == 108          # NonPositiveCentralCorrelator
        f = list(full)
        f[2] = 10.0
        assert status_of(fn, f, jk, [2]) == 109          # AcoshDomainViolation
        bad = [list(r) for r in jk]
        bad[0].pop()
        assert status_of(fn, full, bad, [2]) == 106      # ShapeMismatch
        assert status_of(fn, full, [full, full], [2]) == 102
        assert status_of(fn, full, [full] * 65, [2]) == 103
        for lags in ([0], [8], [2, 2], [3, 2]):
            assert status_of(fn, full, jk, lags) == 105  # InvalidAnalysisLag
        assert status_of(fn, full, jk, []) == 104
        f = list(full)
        f[1] = float("nan")
        assert status_of(fn, f, jk, [2]) == 107
        bad = [list(r) for r in jk]
        bad[0][1] = float("inf")
        assert status_of(fn, full, bad, [2]) == 107
        assert status_of(fn, full[:3], [full[:3]] * 3, [1]) == 100
        assert status_of(fn, [1.0] * 65, [[1.0] * 65] * 3, [1]) == 101
    for fn in (st.analytic_periodic_single_state, R.analytic_periodic_single_state):
        assert status_of(fn, 3, 0.5, 1.0) == 111 and status_of(fn, 16, -0.5, 1.0) == 111


def test_effective_mass_product_vs_oracle(st):
    rng = np.random.default_rng(4)
    for T in (8, 16, 24, 64):
        m = rng.uniform(0.2, 1.2)
        base = np.array(R.analytic_periodic_single_state(T, m, 2.0))
        full = base * (1 + 1e-3 * rng.standard_normal(T))
        jk = [base * (1 + 1e-3 * rng.standard_normal(T)) for _ in range(int(rng.integers(3, 65)))]
        lags = list(range(1, (T - 1) // 2 + (1 if T % 2 else 0)))
        lags = [t for t in lags if 2 * t < T and t < T - 1]
        s = st.cosh_effective_mass_with_jackknife(full, jk, lags)
        o = R.cosh_effective_mass_with_jackknife(full, jk, lags)
        for i in range(len(lags)):
            assert s.full_estimate[i] == o[i][1] and s.jackknife_mean[i] == o[i][2]
            assert s.standard_error[i] == o[i][3]


# ------------------------------------------------------------ correlated fit
def test_scalar_fit_golden_canonical(st):
    v = vec("glueball-scalar-fit-phase3e-v1", "canonical-correlated-single-state-fit")
    c = vec("glueball-scalar-fit-phase3e-v1", "correlated-covariance-not-diagonal")
    j = vec("glueball-scalar-fit-phase3e-v1", "jackknife-parameter-refit")
    full, jk, cands = R.canonical_scalar_fit_fixture()
    assert cands == [tuple(w) for w in v["candidate_windows"]]
    tol = v["absolute_tolerance"]
    o = R.correlated_scalar_fit(full, jk, cands)
    p = st.correlated_scalar_fit(full, jk, cands)
    assert o["window"] == tuple(v["expected_selected_window"]) == (p.window_start, p.window_end)
    assert all(abs(a - b) <= 1e-12 for a, b in zip(o["replica_masses"], j["replica_mass_grid_results"]))
    for name, got in (("oracle", o), ("product", {k: getattr(p, k) for k, _ in p._fields_})):
        assert abs(got["plateau_weighted_effective_mass"] - v["expected_plateau_weighted_effective_mass"]) <= tol, name
        assert got["plateau_maximum_z"] <= 1e-10
        assert abs(got["fitted_mass"] - v["expected_fitted_mass"]) <= 1e-15
        assert abs(got["fitted_amplitude"] - v["expected_fitted_amplitude"]) <= tol
        assert abs(got["chi_square"] - v["expected_chi_square"]) <= tol
        assert got["degrees_of_freedom"] == v["expected_degrees_of_freedom"]
        assert abs(got["mass_jackknife_mean"] - j["expected_mass_jackknife_mean"]) <= 1e-15
        assert abs(got["mass_standard_error"] - j["expected_mass_standard_error"]) <= tol
        assert abs(got["maximum_off_diagonal_correlation"] - c["expected_maximum_off_diagonal_correlation"]) <= tol
        assert got["amplitude_standard_error"] > 0.0
    # closed forms: the Hadamard replicas' correlation is 1/17 and their grid masses
    # give sqrt((N-1)/N sum (m_i - 0.5)^2) = sqrt(7/8 * 4e-5)
    assert abs(p.maximum_off_diagonal_correlation - 1.0 / 17.0) <= 2e-10
    assert abs(p.mass_standard_error - math.sqrt(7.0 / 8.0 * 4.0e-5)) <= 2e-10


def test_scalar_fit_golden_rejections(st):
    full, jk, _ = R.canonical_scalar_fit_fixture()
    for fn in (st.correlated_scalar_fit, R.correlated_scalar_fit):
        assert status_of(fn, full, jk, [(2, 5), (2, 5)]) == 123          # DuplicateCandidateWindow
        assert status_of(fn, full, jk, [(3, 8)]) == 122                  # midpoint window
        assert status_of(fn, full, jk, []) == 120
        assert status_of(fn, full, jk, [(0, 3)]) == 122
        assert status_of(fn, full, jk, [(2, 3)]) == 122
        assert status_of(fn, full, jk, [(1, 3)] * 17) == 121
        assert status_of(fn, full, jk[:4], [(3, 6)]) == 126              # 4 replicas, 4 points
        # eight identical replicas: the plateau errors vanish (no window passes)
        assert status_of(fn, full, [full] * 8, [(3, 6)]) == 125
        # rank-one replica spread: plateau passes, covariance is singular
        g = np.exp(-0.3 * np.arange(16))
        rk1 = [np.array(full) * (1 + 1e-3 * s * g) for s in np.linspace(-1, 1, 8)]
        assert status_of(fn, full, rk1, [(3, 6)]) == 127
    # the reference's unit test form of that vector: invert the zero covariance
    cov = R._covariance([full] * 8, 3, 6)
    assert status_of(R._invert, cov) == 127


def test_scalar_fit_window_tie_prefers_later_start(st):
    v = vec("glueball-scalar-fit-phase3e-v1", "window-tie-prefers-later-start")
    full, jk, _ = R.canonical_scalar_fit_fixture()
    w = [tuple(x) for x in v["passing_equal_length_windows"]]
    for order in (w, w[::-1]):
        p = st.correlated_scalar_fit(full, jk, order)
        assert (p.window_start, p.window_end) == tuple(v["expected_selected_window"])
        assert R.correlated_scalar_fit(full, jk, order)["window"] == tuple(v["expected_selected_window"])


def test_scalar_fit_product_vs_oracle_random(st):
    """Random correlated replica sets around single periodic states (T = 12,
    16, 24; 9..40 replicas): both implementations take the same window and
    the same grid masses; continuous outputs agree to 1e-10."""
    rng = np.random.default_rng(12)
    checked = 0
    for trial in range(12):
        T = int(rng.choice([12, 16, 24]))
        nrep = int(rng.integers(9, 41))
        m = float(rng.uniform(0.2, 1.2))
        full = np.array(R.analytic_periodic_single_state(T, m, float(rng.uniform(0.5, 3.0))))
        t = np.arange(T) / T
        jk = [full * (1 + 0.004 * (rng.standard_normal() + rng.standard_normal() * t +
                                   0.3 * rng.standard_normal(T))) for _ in range(nrep)]
        cands = [(1, 3), (2, 5), (3, 5), (2, 4), (1, 4)]
        cands = [w for w in cands if 2 * w[1] < T]
        ps = status_of(st.correlated_scalar_fit, full, jk, cands)
        os_ = status_of(R.correlated_scalar_fit, full, jk, cands)
        assert ps == os_
        if ps:
            continue
        p = st.correlated_scalar_fit(full, jk, cands)
        o = R.correlated_scalar_fit(full, jk, cands)
        assert (p.window_start, p.window_end) == o["window"]
        assert p.fitted_mass == o["fitted_mass"] and p.mass_jackknife_mean == o["mass_jackknife_mean"]
        assert p.plateau_weighted_effective_mass == o["plateau_weighted_effective_mass"]
        for k in ("fitted_amplitude", "chi_square", "mass_standard_error", "amplitude_standard_error",
                  "maximum_off_diagonal_correlation"):
            assert abs(getattr(p, k) - o[k]) <= 1e-10 * max(1.0, abs(o[k])), k
        checked += 1
    assert checked >= 6


# --------------------------------------------------------------------- GEVP
def two_state_closed_form(t):
    return R.periodic_weight(0.4, t, 16) / R.periodic_weight(0.4, 1, 16)


def test_gevp_golden_and_closed_form(st):
    v = vec("glueball-scalar-variational-phase3j-v1", "canonical-two-state-gevp")
    ms = np.array(R.canonical_two_state_matrices())
    lam_closed = R.periodic_weight(0.4, 2, 16) / R.periodic_weight(0.4, 1, 16)
    p = st.principal_generalized_projection(ms)
    lo, co, po = R.principal_generalized_projection(ms.tolist())
    for lam in (p.principal_eigenvalue, lo):
        assert abs(lam - v["expected_principal_eigenvalue"]) <= 2e-15
        assert abs(lam - lam_closed) <= 2e-15
    for proj in (p.projected_correlator, po):
        assert proj[1] == v["projected_reference_value"]
        for t in range(16):
            assert abs(proj[t] - two_state_closed_form(t)) <= 2e-11
    # the principal vector annihilates the excited coupling (0.25, 1)
    for c in (p.coefficients, co):
        assert abs(0.25 * c[0] + 1.0 * c[1]) <= 1e-12 and c[0] > 0
    assert np.allclose(p.coefficients, co, rtol=1e-12, atol=1e-15)


def test_gevp_invariants(st):
    ms = np.array(R.canonical_two_state_matrices())
    base = st.principal_generalized_projection(ms)
    # independent rescaling of the two operators
    sc = np.array([3.0, 0.25])
    scaled = ms * sc[None, :, None] * sc[None, None, :]
    r = st.principal_generalized_projection(scaled)
    assert abs(r.principal_eigenvalue - base.principal_eigenvalue) <= 1e-14
    assert np.max(np.abs(r.projected_correlator - base.projected_correlator)) <= 1e-12
    assert np.allclose(r.coefficients * sc, base.coefficients, rtol=1e-12)
    # operator order
    sw = ms[:, ::-1, ::-1].copy()
    r = st.principal_generalized_projection(sw)
    assert abs(r.principal_eigenvalue - base.principal_eigenvalue) <= 1e-14
    assert np.max(np.abs(r.projected_correlator - base.projected_correlator)) <= 1e-12
    # same direction; the sign convention (first component positive) applies in the new order
    assert np.allclose(-r.coefficients[::-1], base.coefficients, rtol=1e-12) and r.coefficients[0] > 0
    # rejections (both implementations)
    for fn in (st.principal_generalized_projection, lambda m: R.principal_generalized_projection(m.tolist())):
        bad = ms.copy()
        bad[4, 0, 1] += 1e-6
        assert status_of(fn, bad) == 142                       # MatrixNotSymmetric
        bad = ms.copy()
        bad[3, 1, 1] = float("nan")
        assert status_of(fn, bad) == 74                        # NonFiniteAccumulator
        bad = ms.copy()
        bad[1] = [[1.0, 1.0], [1.0, 1.0]]
        assert status_of(fn, bad) == 144                       # singular reference
        bad = ms.copy()
        bad[1, 0, 0] = -1.0
        assert status_of(fn, bad) == 144
        bad = ms.copy()
        bad[2] = bad[1] * 0.5
        assert status_of(fn, bad) == 145                       # C(t1) proportional to C(t0)
        assert status_of(fn, ms[:2]) == 75
    assert status_of(st.principal_generalized_projection, ms, 0, 2) == 143
    assert status_of(st.principal_generalized_projection, ms[:, :, :], 1, 3) == 143


def test_gevp_product_vs_oracle_random(st):
    rng = np.random.default_rng(21)
    for _ in range(50):
        T = int(rng.integers(3, 20))
        m = rng.uniform(0.1, 1.5, 2)
        amp = rng.uniform(0.2, 2.0, 2)
        cp = rng.standard_normal((2, 2))
        mats = np.zeros((T, 2, 2))
        for t in range(T):
            for n in range(2):
                mats[t] += amp[n] * R.periodic_weight(m[n], t, T) * np.outer(cp[n], cp[n])
        mats += 1e-4 * rng.standard_normal((T, 1, 1)) * np.eye(2)
        ps = status_of(st.principal_generalized_projection, mats)
        os_ = status_of(lambda x: R.principal_generalized_projection(x.tolist()), mats)
        assert ps == os_
        if ps:
            continue
        p = st.principal_generalized_projection(mats)
        lo, co, po = R.principal_generalized_projection(mats.tolist())
        assert abs(p.principal_eigenvalue - lo) <= 1e-12 * abs(lo)
        assert np.allclose(p.coefficients, co, rtol=1e-10, atol=1e-13)
        assert np.allclose(p.projected_correlator, po, rtol=1e-10, atol=1e-13)


# -------------------------------------------------- variational correlator
def test_variational_correlator(st):
    rng = np.random.default_rng(5)
    for nconf, T in ((1, 3), (7, 10), (64, 24)):
        pl = 0.55 + 0.01 * rng.standard_normal((nconf, T))
        rc = 0.30 + 0.01 * rng.standard_normal((nconf, T)) + 0.5 * (pl - 0.55)
        v = st.scalar_variational_connected_correlator(pl, rc)
        vm, om = R.scalar_variational_connected_correlator(pl, rc)
        assert v.configuration_count == nconf and v.time_extent == T
        assert np.allclose(v.vacuum_means, vm, rtol=0, atol=0)
        assert np.max(np.abs(v.matrices - np.array(om))) == 0.0      # same summation order
        for k, op in ((0, pl), (1, rc)):                            # diagonal replays the phase-3C estimator
            _, _, conn = st.periodic_connected(op)
            assert np.max(np.abs(v.matrices[:, k, k] - conn)) <= 2e-12
        assert np.array_equal(v.matrices[:, 0, 1], v.matrices[:, 1, 0])
    pl = np.ones((2, 4))
    assert status_of(st.scalar_variational_connected_correlator, np.zeros((0, 4)), np.zeros((0, 4))) in (70, 75)
    assert status_of(st.scalar_variational_connected_correlator, pl[:, :2], pl[:, :2]) == 140
    assert status_of(st.scalar_variational_connected_correlator, np.ones((65, 4)), np.ones((65, 4))) == 71
    assert status_of(st.scalar_variational_connected_correlator, np.ones((2, 65)), np.ones((2, 65))) == 72
    bad = pl.copy()
    bad[1, 2] = float("inf")
    assert status_of(st.scalar_variational_connected_correlator, bad, pl) == 54


def test_variational_correlator_shifted_defect_fixture(st):
    """glueball-connected-correlator-phase3c-v1.json: four configurations of
    3^3 x 4, configuration c with one traceless-cycle link at [0,0,0,c],
    direction 0.  A1++ samples from the oracle; the matrix diagonals are the
    frozen connected correlators, and since both operators move only on the
    defect slice, C_pr(tau) = sqrt(C_pp(tau) C_rr(tau)) at tau = 0 and
    -sqrt(C_pp C_rr) elsewhere (closed form of a one-slice defect)."""
    import ctypes as C
    import oracle_lib as O
    L = O.lib()
    v = vec("glueball-connected-correlator-phase3c-v1", "shifted-defect-analytic-correlators")["expected"]
    n, T = (3, 3, 3, 4), 4
    cyc = np.zeros(18)
    L.ymo_su3_traceless_cycle_link(O.dptr(cyc))
    pl, rc = np.zeros((4, T)), np.zeros((4, T))
    for c in range(4):
        links = O.su3_identity(n)
        off = O.link_offset(n, (0, 0, 0, c), 0, 18)
        links[off:off + 18] = cyc
        for t in range(T):
            p, r, pc, rcn = C.c_double(), C.c_double(), C.c_int64(), C.c_int64()
            assert L.ymo_su3_a1pp_basis(O.ext(n), O.dptr(links), t, C.byref(p), C.byref(r), C.byref(pc),
                                        C.byref(rcn)) == 0
            pl[c, t], rc[c, t] = p.value, r.value
    res = st.scalar_variational_connected_correlator(pl, rc)
    assert abs(res.vacuum_means[0] - v["plaquette_vacuum_mean"]) <= 2e-12
    assert abs(res.vacuum_means[1] - v["rectangle_vacuum_mean"]) <= 2e-12
    for tau in range(T):
        cpp, crr, cpr = res.matrices[tau, 0, 0], res.matrices[tau, 1, 1], res.matrices[tau, 0, 1]
        assert abs(cpp - v["plaquette_connected"][tau]) <= 2e-12
        assert abs(crr - v["rectangle_connected"][tau]) <= 2e-12
        sign = 1.0 if tau == 0 else -1.0
        assert abs(cpr - sign * math.sqrt(cpp * crr)) <= 2e-12
