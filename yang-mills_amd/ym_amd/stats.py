"""Host FP64 statistics over libym_amd.so (include/ym_amd.h).

Mirrors the reference's analysis functions: ``evaluate``
(sampling_efficiency.rs:57-135), the half-split ``stationarity`` gate
(su3_benchmark.rs:186-237), multi-chain pooling (su3_confirmation.rs:105-263),
``periodic_connected`` (glueball_correlators.rs:111-172),
``connected_from_samples`` (glueball_scalar_production.rs:383-443) and the
delete-one jackknife of ``correlator_evidence``
(glueball_scalar_production.rs:445-470) with the error convention of
glueball_effective_mass.rs:161-215; the correlator analysis that consumes it:
``cosh_effective_mass_with_jackknife`` (glueball_effective_mass.rs:161-232),
``correlated_scalar_fit`` (glueball_scalar_fit.rs:595-650),
``scalar_variational_connected_correlator`` and
``principal_generalized_projection`` (glueball_scalar_variational.rs:628-925).
"""
import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._lib import EffectiveMassPoint, EfficiencyReport, PoolSummary, ScalarFitResult, StationarityReport, check, lib


def _arr(x):
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
    return a, a.ctypes.data_as(C.POINTER(C.c_double))


def evaluate(values, max_lag, setup=0.0, thermalization=0.0, sampling=1.0, log_weights=None):
    v, vp = _arr(values)
    lwp = None
    if log_weights is not None:
        lw, lwp = _arr(log_weights)
        if lw.shape != v.shape:
            raise ValueError("log_weights must match values")
    r = EfficiencyReport()
    check(lib().ym_evaluate(vp, v.size, lwp, int(max_lag), float(setup), float(thermalization), float(sampling),
                            C.byref(r)), "ym_evaluate")
    return r


def stationarity(values, half_lag):
    v, vp = _arr(values)
    r = StationarityReport()
    check(lib().ym_stationarity(vp, v.size, int(half_lag), C.byref(r)), "ym_stationarity")
    return r


@dataclass
class PoolReport:
    chain_mean: np.ndarray
    chain_se: np.ndarray
    chain_ess: np.ndarray
    chain_drift_z: np.ndarray
    pooled_mean: float
    pooled_se: float
    aggregate_ess: float
    max_pairwise_z: float
    max_stationarity_z: float
    min_ess: float


def pool_chains(series, max_lag, half_lag=0):
    """series: (k, n) array, one row per chain."""
    s, sp = _arr(series)
    if s.ndim != 2:
        raise ValueError("series must be (chains, samples)")
    k, n = s.shape
    m, se, ess, dz = (np.zeros(k) for _ in range(4))
    out = PoolSummary()
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    check(lib().ym_pool_chains(sp, k, n, int(max_lag), int(half_lag), dp(m), dp(se), dp(ess), dp(dz), C.byref(out)),
          "ym_pool_chains")
    return PoolReport(m, se, ess, dz, out.pooled_mean, out.pooled_se, out.aggregate_ess, out.max_pairwise_z,
                      out.max_stationarity_z, out.min_ess)


def periodic_connected(samples):
    """samples: (nconf, T).  Returns (mean, raw[T], connected[T])."""
    s, sp = _arr(samples)
    nconf, T = s.shape
    mean, raw, conn = C.c_double(), np.zeros(T), np.zeros(T)
    check(lib().ym_periodic_connected(sp, nconf, T, C.byref(mean), raw.ctypes.data_as(C.POINTER(C.c_double)),
                                      conn.ctypes.data_as(C.POINTER(C.c_double))), "ym_periodic_connected")
    return mean.value, raw, conn


def connected_from_samples(samples, omitted=None):
    s, sp = _arr(samples)
    nconf, T = s.shape
    conn = np.zeros(T)
    check(lib().ym_connected_correlator(sp, nconf, T, -1 if omitted is None else int(omitted),
                                        conn.ctypes.data_as(C.POINTER(C.c_double))), "ym_connected_correlator")
    return conn


@dataclass
class CorrelatorEvidence:
    full: np.ndarray
    jackknife: np.ndarray       # (nconf, T) delete-one replicas
    jackknife_mean: np.ndarray
    jackknife_error: np.ndarray


def correlator_evidence(samples):
    s, sp = _arr(samples)
    nconf, T = s.shape
    full, rep, jm, je = np.zeros(T), np.zeros((nconf, T)), np.zeros(T), np.zeros(T)
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    check(lib().ym_jackknife_correlator(sp, nconf, T, dp(full), dp(rep), dp(jm), dp(je)), "ym_jackknife_correlator")
    return CorrelatorEvidence(full, rep, jm, je)


# ------------------------------------------------------------ F1 analysis
def _rows(replicas):
    """Replica correlators as row pointers + lengths (the reference's &[Vec<f64>])."""
    rows = [np.ascontiguousarray(np.asarray(r, dtype=np.float64)) for r in replicas]
    ptrs = (C.POINTER(C.c_double) * max(1, len(rows)))(*[r.ctypes.data_as(C.POINTER(C.c_double)) for r in rows])
    lens = (C.c_int64 * max(1, len(rows)))(*[r.size for r in rows])
    return rows, ptrs, lens


@dataclass
class EffectiveMassSeries:
    time_extent: int
    jackknife_replicates: int
    tau: np.ndarray
    full_estimate: np.ndarray
    jackknife_mean: np.ndarray
    standard_error: np.ndarray


def cosh_effective_mass_with_jackknife(full, jackknife, lags):
    """glueball_effective_mass.rs:161-232."""
    f, fp = _arr(full)
    rows, ptrs, lens = _rows(jackknife)
    lg = np.ascontiguousarray(np.asarray(lags, dtype=np.int64))
    out = (EffectiveMassPoint * max(1, lg.size))()
    check(lib().ym_cosh_effective_mass(fp, f.size, ptrs, lens, len(rows), lg.ctypes.data_as(C.POINTER(C.c_int64)),
                                       lg.size, out), "ym_cosh_effective_mass")
    pts = out[:lg.size]
    return EffectiveMassSeries(f.size, len(rows), np.array([p.tau for p in pts], dtype=np.int64),
                               np.array([p.full_estimate for p in pts]), np.array([p.jackknife_mean for p in pts]),
                               np.array([p.standard_error for p in pts]))


def analytic_periodic_single_state(time_extent, mass, amplitude):
    """glueball_effective_mass.rs:240-268."""
    out = np.zeros(max(0, int(time_extent)))
    check(lib().ym_analytic_periodic_single_state(int(time_extent), float(mass), float(amplitude),
                                                  out.ctypes.data_as(C.POINTER(C.c_double))),
          "ym_analytic_periodic_single_state")
    return out


def correlated_scalar_fit(full, jackknife, candidate_windows):
    """glueball_scalar_fit.rs:595-650.  candidate_windows: [(start, end), ...]."""
    f, fp = _arr(full)
    rows, ptrs, lens = _rows(jackknife)
    w = np.ascontiguousarray(np.asarray(candidate_windows, dtype=np.int64).reshape(-1))
    r = ScalarFitResult()
    check(lib().ym_correlated_scalar_fit(fp, f.size, ptrs, lens, len(rows), w.ctypes.data_as(C.POINTER(C.c_int64)),
                                         w.size // 2, C.byref(r)), "ym_correlated_scalar_fit")
    return r


@dataclass
class VariationalCorrelator:
    configuration_count: int
    time_extent: int
    vacuum_means: np.ndarray
    matrices: np.ndarray  # (T, 2, 2)


def scalar_variational_connected_correlator(plaquette, rectangle):
    """glueball_scalar_variational.rs:628-744 on A1++ samples: plaquette and
    rectangle are (nconf, T) arrays of per-timeslice operator values."""
    p = np.asarray(plaquette, dtype=np.float64)
    q = np.asarray(rectangle, dtype=np.float64)
    if p.shape != q.shape or p.ndim != 2:
        raise ValueError("plaquette and rectangle samples must be two (nconf, T) arrays")
    s = np.ascontiguousarray(np.stack([p, q], axis=-1))
    nconf, T = p.shape
    vm, mats = np.zeros(2), np.zeros((T, 2, 2))
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    check(lib().ym_variational_connected_correlator(dp(s), nconf, T, dp(vm), dp(mats)),
          "ym_variational_connected_correlator")
    return VariationalCorrelator(nconf, T, vm, mats)


@dataclass
class PrincipalProjection:
    reference_time: int
    diagonalization_time: int
    principal_eigenvalue: float
    coefficients: np.ndarray
    projected_correlator: np.ndarray


def principal_generalized_projection(matrices, reference_time=1, diagonalization_time=2):
    """glueball_scalar_variational.rs:746-925.  matrices: (T, 2, 2)."""
    m = np.ascontiguousarray(np.asarray(matrices, dtype=np.float64))
    if m.ndim != 3 or m.shape[1:] != (2, 2):
        raise ValueError("matrices must be (T, 2, 2)")
    T = m.shape[0]
    lam, coef, proj = C.c_double(), np.zeros(2), np.zeros(T)
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    check(lib().ym_principal_generalized_projection(dp(m), T, int(reference_time), int(diagonalization_time),
                                                    C.byref(lam), dp(coef), dp(proj)),
          "ym_principal_generalized_projection")
    return PrincipalProjection(int(reference_time), int(diagonalization_time), lam.value, coef, proj)
