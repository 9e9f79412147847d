"""Pin the CPU oracle against every golden vector the reference holds.

The fixtures under tests/golden/reference_vectors/ are the reference's own
test vectors (test-vectors/*.json); tests/golden/su3_confirmation_v2_frozen.json
holds the frozen V2 constants of bin/su3_confirmation_reference.rs:32-54.
Closed forms and invariants cover what the vectors leave open.
"""
import ctypes as C
import json
import os
import threading

import numpy as np
import pytest

import oracle_lib as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def vec(name, vid=None):
    with open(os.path.join(GOLD, "reference_vectors", name + ".json")) as f:
        d = json.load(f)
    if vid is None:
        return d
    if "vectors" in d:
        return next(v for v in d["vectors"] if v["id"] == vid)
    return d


L = O.lib()


# ------------------------------------------------------------------ RNG
def test_splitmix64_first_five():
    v = vec("su2-ensemble-v1", "splitmix64-seed-zero-first-five")
    s = C.c_uint64(int(v["seed"], 16))
    assert [L.ymo_splitmix_next(C.byref(s)) for _ in range(5)] == [int(h, 16) for h in v["outputs_hex"]]


def test_chain_seed_derivations():
    v = vec("su3-ensemble-v1", "su3-stream-golden-seeds")
    m = int(v["master_seed"], 16)
    assert [L.ymo_derive_su3_chain_seed(m, c) for c in v["chain_ids"]] == [int(h, 16) for h in v["derived_seeds_hex"]]
    v1 = vec("su3-plaquette-baseline-v1", "two-independent-chain-predeclaration")
    assert [L.ymo_derive_su3_chain_seed(int(v1["master_seed"], 16), c) for c in v1["chain_ids"]] == \
        [int(h, 16) for h in v1["derived_seeds"]]
    v2 = vec("su3-plaquette-baseline-v2", "four-fresh-chain-confirmation")
    assert [L.ymo_derive_su3_chain_seed(int(v2["master_seed"], 16), c) for c in v2["chain_ids"]] == \
        [int(h, 16) for h in v2["derived_seeds"]]
    # su2_benchmark.rs:29 frozen SU(2) baseline seeds
    assert [L.ymo_derive_chain_seed(0x2B00201211160101, c) for c in (0, 1)] == [0x044DF781F86A2ADE, 0x4EBD94783D508D2E]


def test_paired_angle_grid():
    out, neg = C.c_double(), C.c_double()
    rng = np.random.default_rng(3)
    for raw in list(rng.integers(0, 2 ** 63, 200, dtype=np.uint64)) + [0, 2 ** 52 - 1]:
        raw = int(raw)
        assert L.ymo_proposal_angle_from_u64(raw, 0.35, C.byref(out)) == 0
        assert L.ymo_proposal_angle_from_u64(raw | (1 << 63), 0.35, C.byref(neg)) == 0
        assert out.value < 0 and neg.value == -out.value and abs(out.value) < 0.35
        m = raw & ((1 << 52) - 1)
        assert -out.value == 0.35 * (2 * m + 1) * 2.0 ** -53
    # angle bounds (su3_ensemble.rs:27-28): below 1e-8 and the smallest double are rejected
    assert L.ymo_proposal_angle_from_u64(5, 0.99e-8, C.byref(out)) == 31
    assert L.ymo_proposal_angle_from_u64(5, 5e-324, C.byref(out)) == 31


# ------------------------------------------------------------- SU(3) algebra
def gm(a):
    out = np.zeros(18)
    assert L.ymo_su3_gell_mann(a, O.dptr(out)) == 0
    return O.mat_from18(out)


def test_gell_mann_trace_normalisation_and_structure_constants():
    for a in range(8):
        for b in range(8):
            t = np.trace(gm(a) @ gm(b))
            assert abs(t - (2.0 if a == b else 0.0)) <= 2e-15
    f, d = C.c_double(), C.c_double()
    L.ymo_su3_structure_f(0, 1, 2, C.byref(f))
    assert abs(f.value - 1.0) <= 2e-15
    L.ymo_su3_structure_f(0, 4, 5, C.byref(f))
    assert abs(f.value + 0.5) <= 2e-15
    L.ymo_su3_structure_d(0, 0, 7, C.byref(d))
    assert abs(d.value - 0.5773502691896258) <= 2e-15
    L.ymo_su3_structure_d(7, 7, 7, C.byref(d))
    assert abs(d.value + 0.5773502691896258) <= 2e-15
    q, c = np.zeros(18), np.zeros(18)
    L.ymo_su3_casimirs(O.dptr(q), O.dptr(c))
    assert np.max(np.abs(O.mat_from18(q) - 4.0 / 3.0 * np.eye(3))) <= 3e-15
    assert np.max(np.abs(O.mat_from18(c) - 10.0 / 9.0 * np.eye(3))) <= 8e-15


def test_exp_theta3_closed_form_and_single_generator_exp():
    coef = np.zeros(8)
    coef[2] = 0.7
    U = O.mat_from18(O.su3_from_algebra(coef))
    assert np.max(np.abs(U - np.diag([np.exp(0.35j), np.exp(-0.35j), 1.0]))) <= 2e-14
    # every single-generator exponential equals the closed form exp(i theta lambda/2)
    from scipy.linalg import expm
    for a in range(8):
        for th in (-0.9, 0.35, 0.999):
            coef = np.zeros(8)
            coef[a] = th
            U = O.mat_from18(O.su3_from_algebra(coef))
            assert np.max(np.abs(U - expm(1j * th * gm(a) / 2))) <= 1e-14


# ------------------------------------------------------------- lattices
def cycle():
    out = np.zeros(18)
    L.ymo_su3_traceless_cycle_link(O.dptr(out))
    return out


def su3_gauge_field(n):
    g = np.zeros(int(np.prod(n)) * 18)
    for i, x in enumerate(np.ndindex(*n)):
        assert L.ymo_su3_deterministic_site_gauge((C.c_int64 * 4)(*x), O.dptr(g[i * 18:(i + 1) * 18])) == 0
    return g


def test_su3_lattice_vectors():
    n = (3, 3, 3, 3)
    a = C.c_double()
    links = O.su3_identity(n)
    assert L.ymo_su3_wilson_action(O.ext(n), O.dptr(links), 2.0, C.byref(a)) == 0 and a.value == 0.0
    links[0:18] = cycle()
    assert L.ymo_su3_wilson_action(O.ext(n), O.dptr(links), 2.0, C.byref(a)) == 0
    assert abs(a.value - 12.0) <= 1e-12
    # gauge invariance (su3-lattice-v1.json)
    g = su3_gauge_field(n)
    t = np.zeros_like(links)
    assert L.ymo_su3_gauge_transform(O.ext(n), O.dptr(links), O.dptr(g), O.dptr(t)) == 0
    b, c = C.c_double(), C.c_double()
    L.ymo_su3_wilson_action(O.ext(n), O.dptr(links), 1.7, C.byref(b))
    L.ymo_su3_wilson_action(O.ext(n), O.dptr(t), 1.7, C.byref(c))
    assert abs(b.value - c.value) <= 1e-10
    u, d = C.c_double(), C.c_double()
    L.ymo_su3_max_residuals(O.ext(n), O.dptr(t), C.byref(u), C.byref(d))
    assert u.value <= 5e-12 and d.value <= 5e-12


def test_su2_wilson_and_observable_vectors():
    n = (3, 3, 3, 3)
    links = O.su2_identity(n)
    a = C.c_double()
    assert L.ymo_su2_wilson_action(O.ext(n), O.dptr(links), 2.0, C.byref(a)) == 0 and a.value == 0.0
    r = C.c_double()
    zero = (C.c_int64 * 4)(0, 0, 0, 0)
    assert L.ymo_su2_rect_wilson_loop(O.ext(n), O.dptr(links), zero, 0, 1, 2, 1, C.byref(r)) == 0 and r.value == 1.0
    links[0:4] = [0.0, 1.0, 0.0, 0.0]
    assert L.ymo_su2_wilson_action(O.ext(n), O.dptr(links), 2.0, C.byref(a)) == 0
    assert abs(a.value - 12.0) <= 1e-12
    assert L.ymo_su2_rect_wilson_loop(O.ext(n), O.dptr(links), zero, 0, 1, 2, 1, C.byref(r)) == 0
    assert abs(r.value) <= 1e-12
    assert L.ymo_su2_polyakov_loop(O.ext(n), O.dptr(links), zero, 0, C.byref(r)) == 0
    assert abs(r.value) <= 1e-12
    ident = np.array([1.0, 0.0, 0.0, 0.0])
    assert L.ymo_su2_local_action_difference(O.ext(n), O.dptr(links), zero, 0, O.dptr(ident), 2.0, C.byref(r)) == 0
    assert abs(r.value + 12.0) <= 1e-12


def test_su2_staple_identity_and_local_delta():
    """1/2 Re Tr(U V) equals the sum of the six plaquettes (lib.rs:507-546)."""
    n = (3, 4, 3, 4)
    links = O.su2_identity(n)
    rng = np.random.default_rng(1)
    q = rng.standard_normal((len(links) // 4, 4))
    links[:] = (q / np.linalg.norm(q, axis=1, keepdims=True)).reshape(-1)
    for x, mu in [((0, 0, 0, 0), 0), ((2, 3, 1, 2), 3)]:
        V = np.zeros(4)
        assert L.ymo_su2_staple(O.ext(n), O.dptr(links), (C.c_int64 * 4)(*x), mu, O.dptr(V)) == 0
        off = O.link_offset(n, x, mu, 4)
        UV = np.zeros(4)
        L.ymo_su2_mul(O.dptr(np.ascontiguousarray(links[off:off + 4])), O.dptr(V / np.linalg.norm(V)), O.dptr(UV))
        plaq_sum = 0.0
        p = np.zeros(4)
        for nu in range(4):
            if nu == mu:
                continue
            L.ymo_su2_plaquette(O.ext(n), O.dptr(links), (C.c_int64 * 4)(*x), mu, nu, O.dptr(p))
            plaq_sum += p[0]
            xm = list(x)
            xm[nu] = (xm[nu] - 1) % n[nu]
            L.ymo_su2_plaquette(O.ext(n), O.dptr(links), (C.c_int64 * 4)(*xm), mu, nu, O.dptr(p))
            plaq_sum += p[0]
        assert abs(UV[0] * np.linalg.norm(V) - plaq_sum) <= 1e-13


def test_su3_local_delta_matches_full_action():
    v = vec("su3-ensemble-v1", "local-action-difference-vs-full-action")
    n = tuple(v["extents"])
    links = O.su3_identity(n)
    coef = np.zeros(8)
    coef[v["generator"]] = v["angle"]
    cand = O.su3_from_algebra(coef)
    x = (C.c_int64 * 4)(*v["link"]["x"])
    local = C.c_double()
    assert L.ymo_su3_local_action_difference(O.ext(n), O.dptr(links), x, v["link"]["mu"], O.dptr(cand),
                                             v["beta"], C.byref(local)) == 0
    prop = links.copy()
    off = O.link_offset(n, v["link"]["x"], v["link"]["mu"], 18)
    prop[off:off + 18] = cand
    a0, a1 = C.c_double(), C.c_double()
    L.ymo_su3_wilson_action(O.ext(n), O.dptr(links), v["beta"], C.byref(a0))
    L.ymo_su3_wilson_action(O.ext(n), O.dptr(prop), v["beta"], C.byref(a1))
    assert abs(local.value - (a1.value - a0.value)) <= v["absolute_residual_max"]


# ----------------------------------------------------------- SU(3) chains
def su3_config(extents, beta, master, chain, therm, meas, stride, angle, lag):
    c = O.Su3ChainConfig()
    c.extents[:] = list(extents)
    c.beta, c.master_seed, c.chain_id = beta, master, chain
    c.thermalization_sweeps, c.measurements, c.sweeps_between_measurements = therm, meas, stride
    c.proposal_angle, c.autocorrelation_max_lag = angle, lag
    return c


def run_su3(cfg):
    m = np.zeros(cfg.measurements)
    d, r = O.ChainDiagnostics(), O.Su3ChainReport()
    st = L.ymo_su3_run_chain(C.byref(cfg), O.dptr(m), C.byref(d), C.byref(r), None)
    return st, m, d, r


def test_su3_golden_beta0_chain():
    v = vec("su3-ensemble-v1", "frozen-reference-chain-golden-evidence")
    c = v["config"]
    st, m, d, r = run_su3(su3_config(c["extents"], c["beta"], int(c["master_seed"], 16), c["chain_id"],
                                     c["thermalization_sweeps"], c["measurements"],
                                     c["sweeps_between_measurements"], c["proposal_angle"],
                                     c["autocorrelation_max_lag"]))
    tol = v["floating_absolute_tolerance"]
    assert st == 0
    assert r.derived_seed == int(v["derived_seed_hex"], 16) and r.total_sweeps == v["total_sweeps"]
    assert r.proposals == v["metropolis_proposals"] and r.accepts == v["metropolis_accepts"]
    assert np.max(np.abs(m - np.array(v["raw_measurements"]))) <= tol
    g = v["diagnostics"]
    ac = np.zeros(c["autocorrelation_max_lag"] + 1)
    d2 = O.ChainDiagnostics()
    L.ymo_diagnostics(O.dptr(m), len(m), c["autocorrelation_max_lag"], C.byref(d2), O.dptr(ac))
    assert np.max(np.abs(ac - np.array(g["autocorrelation"]))) <= tol
    for k in ("mean", "variance", "standard_error"):
        assert abs(getattr(d, k) - g[k]) <= tol
    assert d.positive_window_lag == g["positive_window_lag"]
    assert abs(d.tau_int - g["integrated_autocorrelation_time"]) <= tol
    assert abs(d.ess - g["effective_sample_size"]) <= tol
    assert abs(r.final_unitarity_residual - v["final_link_unitarity_residual"]) <= tol
    assert abs(r.final_determinant_residual - v["final_link_determinant_residual"]) <= tol


def test_su3_beta0_all_accept_two_draws_per_proposal():
    n = (2, 2, 2, 2)
    links = O.su3_identity(n)
    seed = 0x3C00198212230101
    rng = C.c_uint64(seed)
    p, a = C.c_uint64(), C.c_uint64()
    assert L.ymo_su3_metropolis_sweep(O.ext(n), O.dptr(links), 0.0, 0.35, C.byref(rng), C.byref(p), C.byref(a)) == 0
    assert p.value == a.value == 512
    ref = C.c_uint64(seed)
    for _ in range(2 * 512):
        L.ymo_splitmix_next(C.byref(ref))
    assert rng.value == ref.value


def test_su3_maximum_schedule_group_stress():
    v = vec("su3-ensemble-v1", "maximum-schedule-group-stress")
    st, m, d, r = run_su3(su3_config(v["extents"], v["beta"], int(v["master_seed"], 16), v["chain_id"],
                                     v["maximum_total_sweeps"] - 2, 2, 1, v["proposal_angle"], 1))
    assert st == 0
    assert r.final_unitarity_residual <= 5e-12 and r.final_determinant_residual <= 5e-12


@pytest.mark.slow
def test_su3_beta6_v1_v2_frozen_summaries():
    """22M Metropolis decisions: V1 (2 chains) and V2 (4 chains) at 4^4, beta=6."""
    v1 = vec("su3-plaquette-baseline-v1", "observed-v1-failure")
    with open(os.path.join(GOLD, "su3_confirmation_v2_frozen.json")) as f:
        v2 = json.load(f)
    jobs = [(0x2C02198105010001, c) for c in (0, 1)] + [(0x2C02198105010002, c) for c in range(4)]
    res = {}

    def work(master, chain):
        res[(master, chain)] = run_su3(su3_config((4, 4, 4, 4), 6.0, master, chain, 192, 128, 2, 0.35, 24))

    th = [threading.Thread(target=work, args=j) for j in jobs]
    [t.start() for t in th]
    [t.join() for t in th]
    for j in jobs:
        assert res[j][0] == 0
        assert res[j][3].proposals == 3670016

    def close(a, b):
        return abs(a - b) <= 1e-12 + 1e-10 * abs(b)

    # V1 (bin/su3_benchmark_reference.rs:28-34)
    s1 = np.concatenate([res[(0x2C02198105010001, c)][1] for c in (0, 1)])
    p1 = O.PoolReport()
    assert L.ymo_pool_chains(O.dptr(s1), 2, 128, 24, 12, C.byref(p1)) == 0
    assert close(p1.pooled_mean, v1["pooled_mean"])
    assert close(p1.min_ess, v1["minimum_effective_sample_size_observed"])
    assert close(p1.max_pairwise_z, v1["chain_consistency_z"])
    assert close(p1.max_stationarity_z, v1["maximum_stationarity_z_observed"])
    for c in (0, 1):
        assert close(res[(0x2C02198105010001, c)][3].mean_acceptance, v1["chain_acceptances"][c])
    # V2 (bin/su3_confirmation_reference.rs:32-54)
    s2 = np.concatenate([res[(0x2C02198105010002, c)][1] for c in range(4)])
    p2 = O.PoolReport()
    assert L.ymo_pool_chains(O.dptr(s2), 4, 128, 24, 12, C.byref(p2)) == 0
    for c in range(4):
        assert close(p2.chain_mean[c], v2["chain_means"][c])
        assert close(p2.chain_se[c], v2["chain_standard_errors"][c])
        assert close(p2.chain_ess[c], v2["chain_effective_sample_sizes"][c])
        assert close(p2.chain_stationarity_z[c], v2["stationarity_z"][c])
        assert close(res[(0x2C02198105010002, c)][3].mean_acceptance, v2["chain_acceptances"][c])
    assert close(p2.pooled_mean, v2["pooled_mean"])
    assert close(p2.pooled_se, v2["pooled_standard_error"])
    assert close(p2.aggregate_ess, v2["aggregate_ess"])
    assert close(p2.max_pairwise_z, v2["max_pairwise_z"])
    assert close(p2.max_stationarity_z, v2["max_stationarity_z"])
    mu = max(res[(0x2C02198105010002, c)][3].final_unitarity_residual for c in range(4))
    md = max(res[(0x2C02198105010002, c)][3].final_determinant_residual for c in range(4))
    assert abs(mu - v2["max_unitarity_residual"]) <= 1e-13
    assert abs(md - v2["max_determinant_residual"]) <= 1e-13


# -------------------------------------------------------------- SU(2) chains
def su2_config(extents, beta, master, chain, therm, meas, stride, n_or, attempts, lag):
    c = O.Su2ChainConfig()
    c.extents[:] = list(extents)
    c.beta, c.master_seed, c.chain_id = beta, master, chain
    c.thermalization_sweeps, c.measurements, c.sweeps_between_measurements = therm, meas, stride
    c.overrelaxation_sweeps_per_heatbath, c.max_heatbath_attempts_per_link = n_or, attempts
    c.autocorrelation_max_lag = lag
    return c


def test_su2_beta0_haar_band():
    """tests/su2_ensemble.rs:75-81 and su2-ensemble-v1.json beta-zero band."""
    c = su2_config((2, 2, 2, 2), 0.0, 0x0123456789ABCDEF, 0, 1, 64, 1, 1, 10_000, 8)
    m = np.zeros(64)
    d, r = O.ChainDiagnostics(), O.Su2ChainReport()
    assert L.ymo_su2_run_chain(C.byref(c), O.dptr(m), C.byref(d), C.byref(r), None) == 0
    assert abs(d.mean) <= 0.12 and d.ess >= 8.0
    assert 0.45 <= r.mean_heatbath_acceptance <= 1.0 and r.final_link_unit_residual <= 1e-12


def test_su2_overrelaxation_is_microcanonical():
    """su2-ensemble-v1.json: |dS| <= 2e-10 after one OR sweep (beta 2, seed feedfacecafebeef)."""
    n = (2, 2, 2, 2)
    links = O.su2_identity(n)
    rng = C.c_uint64(0xFEEDFACECAFEBEEF)
    p, u = C.c_uint64(), C.c_uint64()
    assert L.ymo_su2_heatbath_sweep(O.ext(n), O.dptr(links), 2.0, C.byref(rng), 100_000, C.byref(p), C.byref(u)) == 0
    a0, a1 = C.c_double(), C.c_double()
    L.ymo_su2_wilson_action(O.ext(n), O.dptr(links), 2.0, C.byref(a0))
    assert L.ymo_su2_overrelaxation_sweep(O.ext(n), O.dptr(links)) == 0
    L.ymo_su2_wilson_action(O.ext(n), O.dptr(links), 2.0, C.byref(a1))
    assert abs(a1.value - a0.value) <= 2e-10


def test_su2_heatbath_draws_the_exact_conditional_density():
    """At a fixed staple the heat-bath r0 = (1/2)Re Tr(U V)/|V| follows
    p(r0) ~ sqrt(1 - r0^2) exp(alpha r0); Kolmogorov-Smirnov against the
    numerically integrated CDF (a dropped Jacobian or sign fails it)."""
    from scipy import integrate, stats
    n = (2, 2, 2, 2)
    links = O.su2_identity(n)
    rng = C.c_uint64(12345)
    p, u = C.c_uint64(), C.c_uint64()
    L.ymo_su2_heatbath_sweep(O.ext(n), O.dptr(links), 1.0, C.byref(rng), 100_000, C.byref(p), C.byref(u))
    samples, alphas = [], []
    x = (C.c_int64 * 4)(0, 0, 0, 0)
    V = np.zeros(4)
    L.ymo_su2_staple(O.ext(n), O.dptr(links), x, 0, O.dptr(V))
    alpha = 0.5 * np.linalg.norm(V)
    base = links.copy()
    for k in range(4000):
        links[:] = base
        rng2 = C.c_uint64(1000 + k)
        # a heat-bath sweep visits link 0 first: its value is one draw at this staple
        L.ymo_su2_heatbath_sweep(O.ext(n), O.dptr(links), 0.5, C.byref(rng2), 100_000, C.byref(p), C.byref(u))
        Ulink = links[0:4]
        Vh = V / np.linalg.norm(V)
        samples.append(Ulink[0] * Vh[0] - Ulink[1] * Vh[1] - Ulink[2] * Vh[2] - Ulink[3] * Vh[3])
    grid = np.linspace(-1, 1, 4001)
    dens = np.sqrt(np.clip(1 - grid ** 2, 0, None)) * np.exp(alpha * grid)
    cdf = integrate.cumulative_trapezoid(dens, grid, initial=0.0)
    cdf /= cdf[-1]
    ks = stats.kstest(samples, lambda t: np.interp(t, grid, cdf))
    assert ks.pvalue > 1e-3, ks


def test_su2_documented_chain_consistent_with_published_pooled_mean():
    """docs/PUBLISHED_PLAQUETTE_BASELINE.md:111-126 quotes the 10^4 beta=2.3
    pooled mean 0.6022919679191184 (SE 4.17e-4).  No reference test asserts the
    per-chain values, and they do not replay under this restatement (see the
    oracle notes); the restated chains must agree with the pooled mean
    statistically and satisfy the frozen gates."""
    res = {}

    def work(cid):
        c = su2_config((10, 10, 10, 10), 2.3, 0x2B00201211160101, cid, 40, 48, 1, 1, 10_000, 12)
        m = np.zeros(48)
        d, r = O.ChainDiagnostics(), O.Su2ChainReport()
        assert L.ymo_su2_run_chain(C.byref(c), O.dptr(m), C.byref(d), C.byref(r), None) == 0
        res[cid] = (m, d, r)

    th = [threading.Thread(target=work, args=(i,)) for i in (0, 1)]
    [t.start() for t in th]
    [t.join() for t in th]
    s = np.concatenate([res[0][0], res[1][0]])
    p = O.PoolReport()
    assert L.ymo_pool_chains(O.dptr(s), 2, 48, 12, 6, C.byref(p)) == 0
    z = abs(p.pooled_mean - 0.6022919679191184) / np.hypot(p.pooled_se, 4.173286094490e-4)
    assert z < 3.0, z
    assert 0.59 <= p.pooled_mean <= 0.62 and p.min_ess >= 6.0 and p.max_pairwise_z <= 5 and p.max_stationarity_z <= 5
    assert max(res[c][2].final_link_unit_residual for c in (0, 1)) <= 1e-12


# ------------------------------------------------------------- observables
def test_a1pp_identity_counts_and_gauge_invariance():
    v = vec("glueball-operators-phase3a-v1", "identity-3cubed-spatial-slice")
    n = tuple(v["extents"])
    links = O.su3_identity(n)
    p, r, pc, rc = C.c_double(), C.c_double(), C.c_int64(), C.c_int64()
    assert L.ymo_su3_a1pp_basis(O.ext(n), O.dptr(links), 0, C.byref(p), C.byref(r), C.byref(pc), C.byref(rc)) == 0
    e = v["expected"]
    assert (p.value, r.value, pc.value, rc.value) == (e["spatial_plaquette_1x1"], e["spatial_rectangle_2x1"],
                                                      e["plaquette_loop_count"], e["rectangle_loop_count"])
    links[0:18] = cycle()
    g = su3_gauge_field(n)
    t = np.zeros_like(links)
    L.ymo_su3_gauge_transform(O.ext(n), O.dptr(links), O.dptr(g), O.dptr(t))
    p2, r2 = C.c_double(), C.c_double()
    L.ymo_su3_a1pp_basis(O.ext(n), O.dptr(links), 0, C.byref(p), C.byref(r), C.byref(pc), C.byref(rc))
    L.ymo_su3_a1pp_basis(O.ext(n), O.dptr(t), 0, C.byref(p2), C.byref(r2), C.byref(pc), C.byref(rc))
    assert abs(p.value - p2.value) <= 2e-11 and abs(r.value - r2.value) <= 2e-11
    # rejected geometries
    assert L.ymo_su3_a1pp_basis(O.ext(n), O.dptr(links), 3, C.byref(p), C.byref(r), C.byref(pc), C.byref(rc)) == 50
    n2 = (2, 3, 3, 3)
    l2 = O.su3_identity(n2)
    assert L.ymo_su3_a1pp_basis(O.ext(n2), O.dptr(l2), 0, C.byref(p), C.byref(r), C.byref(pc), C.byref(rc)) == 51
    n3 = (3, 4, 5, 3)
    l3 = O.su3_identity(n3)
    assert L.ymo_su3_a1pp_basis(O.ext(n3), O.dptr(l3), 0, C.byref(p), C.byref(r), C.byref(pc), C.byref(rc)) == 52


def test_stout_frozen_vectors():
    v = vec("glueball-spatial-stout-phase3b1-v1", "canonical-nontrivial-fixture-changes-spatial-links")
    n = (3, 3, 3, 3)
    links = O.su3_identity(n)
    out = np.zeros_like(links)
    assert L.ymo_su3_stout_smear(O.ext(n), O.dptr(links), O.dptr(out), 0.1, 3, 0) == 0
    assert L.ymo_su3_max_link_residual(O.ext(n), O.dptr(links), O.dptr(out)) == 0.0
    links[0:18] = cycle()
    assert L.ymo_su3_stout_smear(O.ext(n), O.dptr(links), O.dptr(out), 0.1, 0, 0) == 0
    assert np.array_equal(out, links)
    assert L.ymo_su3_stout_smear(O.ext(n), O.dptr(links), O.dptr(out), 0.1, 1, 0) == 0
    e = v["expected"]
    assert abs(L.ymo_su3_max_link_residual(O.ext(n), O.dptr(links), O.dptr(out)) - e["maximum_all_link_residual"]) <= 1e-12
    p, r, pc, rc = C.c_double(), C.c_double(), C.c_int64(), C.c_int64()
    L.ymo_su3_a1pp_basis(O.ext(n), O.dptr(out), 0, C.byref(p), C.byref(r), C.byref(pc), C.byref(rc))
    assert abs(p.value - e["spatial_plaquette_1x1"]) <= 1e-12 and abs(r.value - e["spatial_rectangle_2x1"]) <= 1e-12
    assert np.array_equal(out.reshape(-1, 4, 18)[:, 3], links.reshape(-1, 4, 18)[:, 3])
    # gauge covariance 5e-11 and 8-level group residuals
    g = su3_gauge_field(n)
    t, ts, st = np.zeros_like(links), np.zeros_like(links), np.zeros_like(links)
    L.ymo_su3_gauge_transform(O.ext(n), O.dptr(links), O.dptr(g), O.dptr(t))
    L.ymo_su3_stout_smear(O.ext(n), O.dptr(t), O.dptr(ts), 0.1, 1, 0)
    L.ymo_su3_gauge_transform(O.ext(n), O.dptr(out), O.dptr(g), O.dptr(st))
    assert L.ymo_su3_max_link_residual(O.ext(n), O.dptr(ts), O.dptr(st)) <= 5e-11
    assert L.ymo_su3_stout_smear(O.ext(n), O.dptr(links), O.dptr(out), 0.1, 8, 0) == 0
    u, d = C.c_double(), C.c_double()
    L.ymo_su3_max_residuals(O.ext(n), O.dptr(out), C.byref(u), C.byref(d))
    assert u.value <= 5e-12 and d.value <= 5e-12
    # rejections (glueball_smearing.rs:87-113)
    assert L.ymo_su3_stout_smear(O.ext(n), O.dptr(links), O.dptr(out), 0.0, 1, 0) == 61
    assert L.ymo_su3_stout_smear(O.ext(n), O.dptr(links), O.dptr(out), 0.250001, 1, 0) == 61
    assert L.ymo_su3_stout_smear(O.ext(n), O.dptr(links), O.dptr(out), 0.1, 9, 0) == 62


def _block(n, links):
    c = (C.c_int64 * 4)()
    assert L.ymo_su3_blocked_extents(O.ext(n), c) == 0
    nc = tuple(c)
    out = np.zeros(int(np.prod(nc)) * 4 * 18)
    assert L.ymo_su3_spatial_block(O.ext(n), O.dptr(links), O.dptr(out)) == 0
    return nc, out


def test_spatial_block_frozen_vectors():
    """glueball-spatial-block-phase3b2-v1.json (glueball_blocking.rs:105-136)."""
    n = (6, 6, 6, 3)
    nc, out = _block(n, O.su3_identity(n))
    assert nc == (3, 3, 3, 3)
    ident = O.su3_identity(nc)
    assert L.ymo_su3_max_link_residual(O.ext(nc), O.dptr(out), O.dptr(ident)) == 0.0
    # canonical output: a cycle link at the fine origin, direction 0
    v = vec("glueball-spatial-block-phase3b2-v1", "canonical-coarse-operator-output")
    fine = O.su3_identity(n)
    fine[0:18] = cycle()
    nc, out = _block(n, fine)
    p, r, pc, rc = C.c_double(), C.c_double(), C.c_int64(), C.c_int64()
    assert L.ymo_su3_a1pp_basis(O.ext(nc), O.dptr(out), 0, C.byref(p), C.byref(r), C.byref(pc), C.byref(rc)) == 0
    e = v["expected"]
    assert abs(p.value - e["spatial_plaquette_1x1"]) <= v["absolute_tolerance"]
    assert abs(r.value - e["spatial_rectangle_2x1"]) <= v["absolute_tolerance"]
    # the coarse link at the origin is the cycle times the identity
    assert np.array_equal(out[0:18], cycle())
    u, d = C.c_double(), C.c_double()
    L.ymo_su3_max_residuals(O.ext(nc), O.dptr(out), C.byref(u), C.byref(d))
    assert u.value <= 5e-12 and d.value <= 5e-12
    # temporal link sampled exactly
    ft = O.su3_identity(n)
    off = O.link_offset(n, (0, 0, 0, 0), 3, 18)
    ft[off:off + 18] = cycle()
    nc, out = _block(n, ft)
    assert np.array_equal(out[O.link_offset(nc, (0, 0, 0, 0), 3, 18):][:18], cycle())
    # gauge covariance under the coarse restriction y -> (2y0, 2y1, 2y2, y3)
    g = su3_gauge_field(n)
    t = np.zeros_like(fine)
    L.ymo_su3_gauge_transform(O.ext(n), O.dptr(fine), O.dptr(g), O.dptr(t))
    _, bt = _block(n, t)
    gc = np.zeros(int(np.prod(nc)) * 18)
    for i, y in enumerate(np.ndindex(*nc)):
        x = (2 * y[0], 2 * y[1], 2 * y[2], y[3])
        assert L.ymo_su3_deterministic_site_gauge((C.c_int64 * 4)(*x), O.dptr(gc[i * 18:(i + 1) * 18])) == 0
    _, b = _block(n, fine)
    tb = np.zeros_like(b)
    L.ymo_su3_gauge_transform(O.ext(nc), O.dptr(b), O.dptr(gc), O.dptr(tb))
    assert L.ymo_su3_max_link_residual(O.ext(nc), O.dptr(bt), O.dptr(tb)) <= 5e-11
    # rejected geometries, in the reference's order
    c = (C.c_int64 * 4)()
    for bad, code in (((4, 4, 4, 3), 51), ((7, 7, 7, 3), 55), ((6, 8, 6, 3), 52)):
        assert L.ymo_su3_blocked_extents(O.ext(bad), c) == code
        lb = O.su3_identity(bad)
        assert L.ymo_su3_spatial_block(O.ext(bad), O.dptr(lb), O.dptr(np.zeros(18))) == code


def test_spatial_block_is_two_step_holonomy():
    """A coarse spatial link equals the fine two-step path holonomy from the
    embedded site (path_holonomy.rs:132-170), on a random SU(3) field."""
    n = (6, 6, 6, 4)
    rng = np.random.default_rng(11)
    links = O.su3_identity(n)
    for i in range(links.size // 18):
        links[i * 18:(i + 1) * 18] = O.su3_from_algebra(rng.uniform(-0.4, 0.4, 8))
    nc, out = _block(n, links)
    h = np.zeros(18)
    endp = (C.c_int64 * 4)()
    for y in [(0, 0, 0, 0), (1, 2, 0, 3), (2, 2, 2, 1)]:
        x = (2 * y[0], 2 * y[1], 2 * y[2], y[3])
        for mu in range(3):
            steps = (C.c_int32 * 2)(mu + 1, mu + 1)
            assert L.ymo_su3_path_holonomy(O.ext(n), O.dptr(links), (C.c_int64 * 4)(*x), steps, 2, O.dptr(h), endp) == 0
            off = O.link_offset(nc, y, mu, 18)
            assert np.max(np.abs(out[off:off + 18] - h)) <= 1e-15
        off = O.link_offset(nc, y, 3, 18)
        assert np.array_equal(out[off:off + 18], links[O.link_offset(n, x, 3, 18):][:18])


def test_connected_correlator_vectors():
    v = vec("glueball-connected-correlator-phase3c-v1", "shifted-defect-analytic-correlators")
    n = (3, 3, 3, 4)
    pl, rc = [], []
    for c in range(4):
        links = O.su3_identity(n)
        off = O.link_offset(n, (0, 0, 0, c), 0, 18)
        links[off:off + 18] = cycle()
        for t in range(4):
            p, r, pc, rcn = C.c_double(), C.c_double(), C.c_int64(), C.c_int64()
            L.ymo_su3_a1pp_basis(O.ext(n), O.dptr(links), t, C.byref(p), C.byref(r), C.byref(pc), C.byref(rcn))
            pl.append(p.value)
            rc.append(r.value)
    e = v["expected"]
    for samples, key in ((np.array(pl), "plaquette"), (np.array(rc), "rectangle")):
        mean = C.c_double()
        raw, conn = np.zeros(4), np.zeros(4)
        assert L.ymo_periodic_connected(O.dptr(samples), 4, 4, C.byref(mean), O.dptr(raw), O.dptr(conn)) == 0
        assert abs(mean.value - e[key + "_vacuum_mean"]) <= 2e-12
        assert np.max(np.abs(raw - e[key + "_raw"])) <= 2e-12
        assert np.max(np.abs(conn - e[key + "_connected"])) <= 2e-12
        full = np.zeros(4)
        assert L.ymo_connected_from_samples(O.dptr(samples), 4, 4, -1, O.dptr(full)) == 0
        assert np.max(np.abs(full - conn)) <= 2e-12
        # delete-one jackknife replica = estimator on the remaining configurations
        jk = np.zeros(4)
        assert L.ymo_connected_from_samples(O.dptr(samples), 4, 4, 2, O.dptr(jk)) == 0
        rest = np.ascontiguousarray(np.delete(samples.reshape(4, 4), 2, axis=0).reshape(-1))
        m2, raw2, conn2 = C.c_double(), np.zeros(4), np.zeros(4)
        L.ymo_periodic_connected(O.dptr(rest), 3, 4, C.byref(m2), O.dptr(raw2), O.dptr(conn2))
        assert np.max(np.abs(jk - conn2)) <= 1e-15


def test_sampling_efficiency_fixture():
    v = vec("sampling-efficiency-v1")
    vals = np.array(v["values"], dtype=np.float64)
    r = O.EfficiencyReport()
    c = v["cost"]
    assert L.ymo_evaluate(O.dptr(vals), len(vals), None, v["max_lag"], c["setup"], c["thermalization"], c["sampling"], C.byref(r)) == 0
    u = v["unit_weights"]
    assert abs(r.diag.tau_int - u["tau_int"]) <= 1e-12 and abs(r.diag.ess - u["autocorrelation_ess"]) <= 1e-12
    assert abs(r.combined_ess_proxy - u["combined_ess_proxy"]) <= 1e-12
    assert abs(r.sampling_cost_per_ess - u["sampling_cost_per_proxy"]) <= 1e-12
    assert abs(r.total_cost_per_ess - u["total_cost_per_proxy"]) <= 1e-12
    lw = np.log(np.array([1.0, 1.0, 1.0, 3.0]))
    assert L.ymo_evaluate(O.dptr(vals), len(vals), O.dptr(lw), v["max_lag"], c["setup"], c["thermalization"], c["sampling"], C.byref(r)) == 0
    w = v["weights_1_1_1_3"]
    assert abs(r.reweighting_efficiency - w["weight_efficiency"]) <= 1e-12
    assert abs(r.weight_only_ess - w["weight_only_ess"]) <= 1e-12
    assert abs(r.total_cost_per_ess - w["total_cost_per_proxy"]) <= 1e-12
    cv = np.array(v["correlated_values"], dtype=np.float64)
    assert L.ymo_evaluate(O.dptr(cv), len(cv), None, v["correlated_max_lag"], 0.0, 0.0, 1.0, C.byref(r)) == 0
    assert abs(r.diag.tau_int - v["correlated_tau_int"]) <= 1e-12


def test_path_holonomy_invariants():
    """path-holonomy-v1.json on tests/path_holonomy.rs:33-40's two-link fixture."""
    n = (3, 3, 3, 3)
    links = O.su3_identity(n)
    links[O.link_offset(n, (0, 0, 0, 0), 0, 18):][:18] = O.su3_from_algebra([0.4, 0, 0.2, 0, 0, 0, 0, 0.1])
    links[O.link_offset(n, (1, 0, 0, 0), 1, 18):][:18] = O.su3_from_algebra([0, -0.3, 0, 0.25, 0, 0, 0.1, -0.05])
    z = (C.c_int64 * 4)(0, 0, 0, 0)

    def hol(start, steps):
        h, end = np.zeros(18), (C.c_int64 * 4)()
        s = (C.c_int32 * len(steps))(*steps)
        assert L.ymo_su3_path_holonomy(O.ext(n), O.dptr(links), start, s, len(steps), O.dptr(h), end) == 0
        return O.mat_from18(h), tuple(end)

    fwd, end = hol(z, [1, 2])
    back, ret = hol((C.c_int64 * 4)(*end), [-2, -1])
    assert ret == (0, 0, 0, 0) and np.max(np.abs(back - fwd.conj().T)) <= 5e-12
    u1, mid = hol(z, [1])
    u2, _ = hol((C.c_int64 * 4)(*mid), [2])
    assert np.max(np.abs(fwd - u1 @ u2)) <= 5e-12
    g = su3_gauge_field(n)
    t = np.zeros_like(links)
    L.ymo_su3_gauge_transform(O.ext(n), O.dptr(links), O.dptr(g), O.dptr(t))
    re, im, re2, im2 = C.c_double(), C.c_double(), C.c_double(), C.c_double()
    L.ymo_su3_rect_wilson_loop(O.ext(n), O.dptr(links), z, 0, 1, 1, 1, C.byref(re), C.byref(im))
    L.ymo_su3_rect_wilson_loop(O.ext(n), O.dptr(t), z, 0, 1, 1, 1, C.byref(re2), C.byref(im2))
    assert abs(complex(re.value, im.value) - complex(re2.value, im2.value)) <= 2e-11
