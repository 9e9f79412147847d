"""Batched ensembles of independent SU(3) chains (include/ym_amd.h,
ym_ensemble_*).

The reference generates independent configurations one chain after another
(generate_independent_configurations, glueball_scalar_production.rs:257-297):
key derive_su3_chain_seed(MASTER_SEED, chain_id), cold start, 68 Metropolis
sweeps at beta = 5.7 with proposal angle 0.35, then the A1++ operator basis
on raw and stout-smeared links and the connected correlators with a
delete-one jackknife.  ``Ensemble`` runs all chains together on the device,
each sweep and measurement being one batched launch over every chain, with the
sweep/measure cycle captured as a hipGraph.
"""
import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import stats
from ._lib import check, lib
from .lattice import derive_su3_chain_seed

# glueball_scalar_production.rs:33-40 and glueball_smearing.rs:23
PRODUCTION_MASTER_SEED = 0x3F00_2026_0915_0001
PRODUCTION_BETA = 5.7
PRODUCTION_ANGLE = 0.35
PRODUCTION_SWEEPS = 64 + 4
CANONICAL_RHO = 0.10

# the batched production set: 64 chains each of four uneven geometries
PRODUCTION_GEOMETRIES = ((4, 4, 4, 10), (6, 6, 6, 12), (8, 8, 8, 16), (12, 12, 12, 24))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


@dataclass
class CycleResult:
    plaquette: np.ndarray                 # (n_cycles, nchains)
    ops: list = field(default_factory=list)   # ops[cycle][chain] -> (levels + 1, T_c, 2)
    accepts: np.ndarray = None            # (nchains,) accepted proposals in this call
    device_ms: float = 0.0


class Ensemble:
    def __init__(self, extents, keys, precision=32):
        ext = np.ascontiguousarray(np.asarray(extents, dtype=np.int64).reshape(-1, 4))
        k = np.ascontiguousarray(np.asarray(keys, dtype=np.uint64).reshape(-1))
        if ext.shape[0] != k.size:
            raise ValueError("one key per chain")
        self.extents = [tuple(int(v) for v in e) for e in ext]
        self.keys = [int(v) for v in k]
        self.precision = precision
        h = C.c_void_p()
        check(lib().ym_ensemble_create(len(self.keys), ext.ctypes.data_as(C.POINTER(C.c_int64)),
                                       k.ctypes.data_as(C.POINTER(C.c_uint64)), precision, C.byref(h)),
              "ym_ensemble_create")
        self._h = h

    @classmethod
    def from_chain_ids(cls, extents, chain_ids, master_seed=PRODUCTION_MASTER_SEED, precision=32):
        return cls(extents, [derive_su3_chain_seed(master_seed, int(c)) for c in chain_ids], precision)

    def close(self):
        if getattr(self, "_h", None):
            lib().ym_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def nchains(self):
        return len(self.keys)

    def set_identity(self):
        check(lib().ym_ensemble_set_identity(self._h), "ym_ensemble_set_identity")

    def hot_start(self):
        check(lib().ym_ensemble_hot_start(self._h), "ym_ensemble_hot_start")

    def sweep_index(self):
        v = C.c_uint64()
        check(lib().ym_ensemble_sweep_index(self._h, C.byref(v)), "ym_ensemble_sweep_index")
        return v.value

    def download_chain(self, c):
        n = int(np.prod(self.extents[c])) * 4 * 18
        out = np.zeros(n)
        check(lib().ym_ensemble_download_chain(self._h, int(c), _dp(out), n), "ym_ensemble_download_chain")
        return out

    def upload_chain(self, c, aos):
        a = np.ascontiguousarray(np.asarray(aos, dtype=np.float64).reshape(-1))
        check(lib().ym_ensemble_upload_chain(self._h, int(c), _dp(a), a.size), "ym_ensemble_upload_chain")

    def run_cycles(self, beta, angle, hits=8, sweeps_per_cycle=1, n_cycles=1, stout_levels=0, rho=CANONICAL_RHO,
                   measure=True, use_graph=True, reunitarize=True):
        nch = self.nchains
        ops_len, offs = C.c_int64(), np.zeros(nch, dtype=np.int64)
        check(lib().ym_ensemble_ops_length(self._h, int(stout_levels), C.byref(ops_len),
                                           offs.ctypes.data_as(C.POINTER(C.c_int64))), "ym_ensemble_ops_length")
        plaq = np.zeros((n_cycles, nch)) if measure else None
        ops = np.zeros((n_cycles, ops_len.value)) if measure else None
        acc = np.zeros(nch, dtype=np.uint64)
        ms = C.c_float()
        check(lib().ym_ensemble_run_cycles(self._h, float(beta), float(angle), int(hits), 1 if reunitarize else 0,
                                           int(sweeps_per_cycle), int(stout_levels), float(rho), int(n_cycles),
                                           1 if use_graph else 0, _dp(plaq) if measure else None,
                                           _dp(ops) if measure else None,
                                           acc.ctypes.data_as(C.POINTER(C.c_uint64)), C.byref(ms)),
              "ym_ensemble_run_cycles")
        res = CycleResult(plaq, [], acc, ms.value)
        if measure:
            tsum = ops_len.value // (2 * (stout_levels + 1))
            for k in range(n_cycles):
                blk = ops[k].reshape(stout_levels + 1, tsum, 2)
                res.ops.append([blk[:, offs[c]:offs[c] + self.extents[c][3], :] for c in range(nch)])
        return res

    def sweeps(self, beta, angle, n_sweeps, hits=8, reunitarize=True):
        """Sweeps only (no measurement); returns accepted proposals per chain."""
        return self.run_cycles(beta, angle, hits, sweeps_per_cycle=n_sweeps, n_cycles=1, measure=False,
                               use_graph=False, reunitarize=reunitarize).accepts


@dataclass
class ProductionReport:
    geometries: tuple
    chains_per_geometry: int
    plaquette: np.ndarray          # (nchains,)
    acceptance: np.ndarray         # (nchains,)
    correlators: dict              # (geometry, level, kind) -> stats.CorrelatorEvidence
    device_ms: float
    sweeps: int
    analysis: dict = None          # (geometry, level) -> correlator_analysis(...)


def _attempt(fn, *a):
    """Result of an analysis step, or the status of the reference error it fails with."""
    from ._lib import YmError
    try:
        return fn(*a)
    except YmError as e:
        return {"status": e.status, "error": str(e)}


def correlator_analysis(plaquette, rectangle, candidate_windows=None):
    """The host FP64 analysis of one ensemble's A1++ samples ((nconf, T)
    arrays of per-timeslice plaquette and rectangle operators): connected
    correlators with delete-one jackknife (glueball_scalar_production.rs:383-470),
    cosh effective masses of the plaquette channel (glueball_effective_mass.rs:161-232),
    the correlated one-state fit over predeclared windows
    (glueball_scalar_fit.rs:595-650), and the 2 x 2 variational correlator with
    its principal generalized projection (glueball_scalar_variational.rs:628-925).
    Steps that the data do not admit report the reference's error status."""
    plaquette = np.asarray(plaquette, dtype=np.float64)
    rectangle = np.asarray(rectangle, dtype=np.float64)
    T = plaquette.shape[1]
    ev = stats.correlator_evidence(plaquette)
    lags = [t for t in range(1, T) if 2 * t < T and t < T - 1]
    if candidate_windows is None:
        candidate_windows = [(a, a + n - 1) for n in (3, 4) for a in (1, 2) if 2 * (a + n - 1) < T]
    out = {"plaquette": ev, "rectangle": stats.correlator_evidence(rectangle)}
    out["effective_mass"] = _attempt(stats.cosh_effective_mass_with_jackknife, ev.full, ev.jackknife, lags)
    out["fit"] = _attempt(stats.correlated_scalar_fit, ev.full, ev.jackknife, candidate_windows)
    var = _attempt(stats.scalar_variational_connected_correlator, plaquette, rectangle)
    out["variational"] = var
    out["projection"] = _attempt(stats.principal_generalized_projection, var.matrices) \
        if isinstance(var, stats.VariationalCorrelator) else var
    return out


def production(chains_per_geometry=64, geometries=PRODUCTION_GEOMETRIES, stout_levels=4, rho=CANONICAL_RHO,
               sweeps=PRODUCTION_SWEEPS, beta=PRODUCTION_BETA, angle=PRODUCTION_ANGLE,
               master_seed=PRODUCTION_MASTER_SEED, hits=8, use_graph=True):
    """The batched glueball production: chain ids 0.. over the geometries
    (chains_per_geometry each), `sweeps` sweeps from a cold start in one
    graph cycle, then plaquette and the A1++ basis on raw + stout levels, and
    per geometry the connected correlators C(t) with delete-one jackknife over
    its chains (glueball_scalar_production.rs:383-470)."""
    extents, ids = [], []
    for gi, g in enumerate(geometries):
        for j in range(chains_per_geometry):
            extents.append(g)
            ids.append(gi * chains_per_geometry + j)
    ens = Ensemble.from_chain_ids(extents, ids, master_seed)
    try:
        r = ens.run_cycles(beta, angle, hits, sweeps_per_cycle=sweeps, n_cycles=1, stout_levels=stout_levels,
                           rho=rho, use_graph=use_graph)
    finally:
        ens.close()
    corr, analysis = {}, {}
    for gi, g in enumerate(geometries):
        rows = range(gi * chains_per_geometry, (gi + 1) * chains_per_geometry)
        for lev in range(stout_levels + 1):
            samples = [np.array([r.ops[0][c][lev, :, kind] for c in rows]) for kind in range(2)]
            for kind in range(2):
                corr[(g, lev, kind)] = stats.correlator_evidence(samples[kind])
            if chains_per_geometry >= 3 and g[3] >= 4:
                analysis[(g, lev)] = correlator_analysis(samples[0], samples[1])
    links_per_sweep = np.array([4 * np.prod(e) for e in extents], dtype=np.float64)
    acceptance = r.accepts.astype(np.float64) / (links_per_sweep * hits * sweeps)
    return ProductionReport(tuple(geometries), chains_per_geometry, r.plaquette[0], acceptance, corr, r.device_ms,
                            sweeps, analysis)
