"""Host FP64 statistics over libym_amd.so (include/ym_amd.h).

Mirrors the reference's analysis functions: ``evaluate``
(sampling_efficiency.rs:57-135), the half-split ``stationarity`` gate
(su3_benchmark.rs:186-237), multi-chain pooling (su3_confirmation.rs:105-263),
``periodic_connected`` (glueball_correlators.rs:111-172),
``connected_from_samples`` (glueball_scalar_production.rs:383-443) and the
delete-one jackknife of ``correlator_evidence``
(glueball_scalar_production.rs:445-470) with the error convention of
glueball_effective_mass.rs:161-215.
"""
import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._lib import EfficiencyReport, PoolSummary, StationarityReport, check, lib


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
