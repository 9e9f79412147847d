"""Lattice handles and reference-named entry points over the C ABI."""
import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from ._lib import ChainDiagnostics, Su2ChainConfig, Su3ChainConfig, Su3ChainReport, check, lib
from ._lib import Su2ChainReport as _Su2ChainReportC


def _ext(extents):
    return (C.c_int64 * 4)(*[int(v) for v in extents])


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class _Lattice:
    """Device-resident lattice (lib.rs:235-756 Lattice4 / su3_lattice.rs:20-313)."""

    GROUP = 0
    NCOMP = 0

    def __init__(self, extents, precision=32):
        self.extents = tuple(int(v) for v in extents)
        self.precision = int(precision)
        h = C.c_void_p()
        check(lib().ym_lattice_create(self.GROUP, _ext(self.extents), self.precision, C.byref(h)),
              "ym_lattice_create")
        self._h = h

    # lifetime -----------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().ym_lattice_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    @classmethod
    def new_identity(cls, extents, precision=32):
        return cls(extents, precision)

    # sizes ----------------------------------------------------------------
    def site_count(self):
        return int(np.prod(self.extents))

    def link_count(self):
        return 4 * self.site_count()

    def device_bytes(self):
        b = C.c_int64()
        check(lib().ym_lattice_bytes(self._h, C.byref(b)), "ym_lattice_bytes")
        return b.value

    # data -----------------------------------------------------------------
    def set_identity(self):
        check(lib().ym_lattice_set_identity(self._h), "ym_lattice_set_identity")

    def hot_start(self, seed):
        check(lib().ym_lattice_hot_start(self._h, int(seed)), "ym_lattice_hot_start")

    def upload(self, aos, checked=True):
        """Set every link from a host AoS image.  checked=True applies the
        reference's link checks (finite, unitary, unit determinant / unit
        norm) and fails closed; checked=False stores the values as given."""
        a = np.ascontiguousarray(aos, dtype=np.float64).reshape(-1)
        if checked:
            check(lib().ym_lattice_upload(self._h, _dp(a), a.size), "ym_lattice_upload")
        else:
            check(lib().ym_lattice_upload_unchecked(self._h, _dp(a), a.size), "ym_lattice_upload_unchecked")

    def download(self):
        a = np.empty(self.link_count() * self.NCOMP, dtype=np.float64)
        check(lib().ym_lattice_download(self._h, _dp(a), a.size), "ym_lattice_download")
        return a

    def copy_from(self, other):
        check(lib().ym_lattice_copy(self._h, other._h), "ym_lattice_copy")

    def clone(self):
        out = type(self)(self.extents, self.precision)
        out.copy_from(self)
        return out

    # observables ----------------------------------------------------------
    def average_plaquette(self):
        v = C.c_double()
        check(lib().ym_average_plaquette(self._h, C.byref(v)), "ym_average_plaquette")
        return v.value

    def wilson_action(self, beta):
        v = C.c_double()
        check(lib().ym_wilson_action(self._h, float(beta), C.byref(v)), "ym_wilson_action")
        return v.value

    def staple(self, x, mu):
        out = np.zeros(self.NCOMP)
        check(lib().ym_staple(self._h, _ext(x), int(mu), _dp(out)), "ym_staple")
        return out

    def path_holonomy(self, start, steps):
        """Ordered product along `steps` (+-(d+1)) from `start`; returns
        (matrix as NCOMP doubles, end point)."""
        st = np.ascontiguousarray(np.asarray(steps, dtype=np.int32))
        out = np.zeros(self.NCOMP)
        end = (C.c_int64 * 4)()
        check(lib().ym_path_holonomy(self._h, _ext(start), st.ctypes.data_as(C.POINTER(C.c_int32)), st.size,
                                     _dp(out), end), "ym_path_holonomy")
        return out, tuple(end)

    def wilson_loops(self, rmax, tmax, plane_set=0):
        re = np.zeros(rmax * tmax)
        im = np.zeros(rmax * tmax)
        check(lib().ym_wilson_loops(self._h, int(rmax), int(tmax), int(plane_set), _dp(re), _dp(im)),
              "ym_wilson_loops")
        return re.reshape(rmax, tmax), im.reshape(rmax, tmax)

    def polyakov_loop(self, direction):
        re, im, ab = C.c_double(), C.c_double(), C.c_double()
        check(lib().ym_polyakov_loop(self._h, int(direction), C.byref(re), C.byref(im), C.byref(ab)),
              "ym_polyakov_loop")
        return complex(re.value, im.value), ab.value

    def link_residuals(self):
        u, d = C.c_double(), C.c_double()
        check(lib().ym_link_residuals(self._h, C.byref(u), C.byref(d)), "ym_link_residuals")
        return u.value, d.value

    def reunitarize(self):
        check(lib().ym_reunitarize(self._h), "ym_reunitarize")

    # configuration formats -------------------------------------------------
    def save_checkpoint(self, path, key, next_sweep):
        """Binary checkpoint: links in storage precision + chain key + next sweep."""
        from . import codec
        return codec.save_checkpoint(path, self, key, next_sweep)

    @staticmethod
    def load_checkpoint(path):
        """(lattice, key, next_sweep) from a checkpoint file."""
        from . import codec
        return codec.load_checkpoint(path)

    # timing ---------------------------------------------------------------
    def timer_begin(self):
        check(lib().ym_timer_begin(self._h), "ym_timer_begin")

    def timer_end(self):
        ms = C.c_float()
        check(lib().ym_timer_end(self._h, C.byref(ms)), "ym_timer_end")
        return ms.value


class Lattice4(_Lattice):
    """SU(2) lattice (lib.rs:235)."""

    GROUP = 2
    NCOMP = 4

    def max_link_unit_residual(self):
        return self.link_residuals()[0]

    def heatbath_sweeps(self, beta, key, first_sweep, n_sweeps=1, max_attempts=10_000):
        p, u = C.c_uint64(0), C.c_uint64(0)
        st = lib().ym_su2_heatbath_sweeps(self._h, float(beta), int(key), int(first_sweep), int(n_sweeps),
                                          int(max_attempts), C.byref(p), C.byref(u))
        check(st, "ym_su2_heatbath_sweeps")
        return p.value, u.value

    def overrelax_sweeps(self, n_sweeps=1):
        check(lib().ym_su2_overrelax_sweeps(self._h, int(n_sweeps)), "ym_su2_overrelax_sweeps")

    def markov_sweeps(self, beta, key, first_sweep, n_sweeps=1, overrelaxation_sweeps_per_heatbath=1,
                      max_attempts=10_000):
        """su2_ensemble.rs:465-486 markov_sweep x n: (heat-bath proposals, updates)."""
        p, u = C.c_uint64(0), C.c_uint64(0)
        st = lib().ym_su2_markov_sweeps(self._h, float(beta), int(key), int(first_sweep), int(n_sweeps),
                                        int(overrelaxation_sweeps_per_heatbath), int(max_attempts), C.byref(p),
                                        C.byref(u))
        check(st, "ym_su2_markov_sweeps")
        return p.value, u.value

    def to_canonical_config(self):
        """lib.rs:587-614: the YM-SU2-CONFIG-V1 text of this configuration."""
        from . import codec
        return codec.to_canonical_config(self)

    @classmethod
    def from_canonical_config(cls, text, precision=64):
        """lib.rs:616-694: a new lattice from canonical text."""
        from . import codec
        return codec.from_canonical_config(text, precision)


class Su3Lattice4(_Lattice):
    """SU(3) lattice (su3_lattice.rs:20)."""

    GROUP = 3
    NCOMP = 18

    def max_link_unitarity_residual(self):
        return self.link_residuals()[0]

    def max_link_determinant_residual(self):
        return self.link_residuals()[1]

    def metropolis_sweeps(self, beta, angle, key, first_sweep, n_sweeps=1, hits=8, reunitarize=True):
        p, a = C.c_uint64(0), C.c_uint64(0)
        st = lib().ym_su3_metropolis_sweeps(self._h, float(beta), float(angle), int(hits), int(key),
                                            int(first_sweep), int(n_sweeps), 1 if reunitarize else 0,
                                            C.byref(p), C.byref(a))
        check(st, "ym_su3_metropolis_sweeps")
        return p.value, a.value

    AUTOTUNE_CANDIDATES = (("row", 0), ("row", 2), ("row", 4), ("row", 8), ("row", 16),
                           ("compact", 1), ("compact", 2), ("compact", 4), ("compact", 8))

    def autotune(self, beta=6.0, angle=0.35, hits=8, reps=3):
        """Pick the fastest thread-to-link mapping of the Metropolis sub-pass
        for this lattice.  Returns ((mode, strip), {(mode, strip): us per
        sub-pass}).  On lattices large enough for the wave-specialised kernel
        the one-link-per-thread kernel is timed with the best mapping too,
        under the mode "<mode>/per-link"; it is the result when faster."""
        mode, strip = C.c_int32(), C.c_int32()
        us = (C.c_float * 10)()
        check(lib().ym_su3_autotune(self._h, float(beta), float(angle), int(hits), int(reps), C.byref(mode),
                                    C.byref(strip), us), "ym_su3_autotune")
        names = ("row", "compact")
        best = (names[mode.value % 2] + ("/per-link" if mode.value >= 2 else ""), strip.value)
        cand = {c: us[i] for i, c in enumerate(self.AUTOTUNE_CANDIDATES) if us[i] >= 0}
        if us[9] >= 0:
            cand[(names[mode.value % 2] + "/per-link", strip.value)] = us[9]
        return best, cand

    def metropolis_sweeps_profiled(self, beta, angle, key, first_sweep, n_sweeps, hits=8, reunitarize=True):
        """Same sweeps with HIP events around every sub-pass launch.

        Returns (proposals, accepts, total_ms, kernel_ms_sum, launches)."""
        p, a = C.c_uint64(0), C.c_uint64(0)
        tot, ksum, nl = C.c_float(), C.c_float(), C.c_int32()
        st = lib().ym_su3_metropolis_sweeps_profiled(self._h, float(beta), float(angle), int(hits), int(key),
                                                     int(first_sweep), int(n_sweeps), 1 if reunitarize else 0,
                                                     C.byref(p), C.byref(a), C.byref(tot), C.byref(ksum),
                                                     C.byref(nl))
        check(st, "ym_su3_metropolis_sweeps_profiled")
        return p.value, a.value, tot.value, ksum.value, nl.value

    def metropolis_sweeps_timed(self, beta, angle, key, first_sweep, n_sweeps, hits=8, reunitarize=True):
        """Same sweeps with one HIP event pair around all launches.

        Returns (proposals, accepts, device_ms)."""
        p, a = C.c_uint64(0), C.c_uint64(0)
        tot = C.c_float()
        st = lib().ym_su3_metropolis_sweeps_profiled(self._h, float(beta), float(angle), int(hits), int(key),
                                                     int(first_sweep), int(n_sweeps), 1 if reunitarize else 0,
                                                     C.byref(p), C.byref(a), C.byref(tot), None, None)
        check(st, "ym_su3_metropolis_sweeps_profiled")
        return p.value, a.value, tot.value

    def timeslice_a1pp(self):
        T = self.extents[3]
        pl, rc = np.zeros(T), np.zeros(T)
        check(lib().ym_su3_timeslice_a1pp(self._h, _dp(pl), _dp(rc)), "ym_su3_timeslice_a1pp")
        return pl, rc


# ------------------------------------------------------------ free functions
def metropolis_sweep(lattice, beta, proposal_angle, key, sweep, hits=8):
    """su3_ensemble.rs:342-391 (checkerboard + Philox order): (proposals, accepts)."""
    return lattice.metropolis_sweeps(beta, proposal_angle, key, sweep, 1, hits)


def heatbath_sweep(lattice, beta, key, sweep, max_attempts_per_link=10_000):
    """su2_ensemble.rs:321-342: (proposals, updates)."""
    return lattice.heatbath_sweeps(beta, key, sweep, 1, max_attempts_per_link)


def overrelaxation_sweep(lattice):
    """su2_ensemble.rs:344-346."""
    lattice.overrelax_sweeps(1)


def average_plaquette(lattice):
    """su2_ensemble.rs:348-373 / su3_ensemble.rs:228-259."""
    return lattice.average_plaquette()


def spatial_stout_smear(lattice, rho, levels):
    """glueball_smearing.rs:257-268: returns a new lattice."""
    out = Su3Lattice4(lattice.extents, lattice.precision)
    check(lib().ym_su3_stout_spatial(out.handle, lattice.handle, float(rho), int(levels)), "ym_su3_stout_spatial")
    return out


def blocked_extents(lattice):
    """glueball_blocking.rs:85-94: coarse extents (N/2, N/2, N/2, N3)."""
    c = (C.c_int64 * 4)()
    check(lib().ym_su3_blocked_extents(lattice.handle, c), "ym_su3_blocked_extents")
    return tuple(int(v) for v in c)


def spatial_stride2_block(lattice):
    """glueball_blocking.rs:105-136: returns a new coarse lattice with
    V_mu(y) = U_mu(x) U_mu(x+mu) (spatial mu), V_3(y) = U_3(x), x = (2y0, 2y1, 2y2, y3)."""
    out = Su3Lattice4(blocked_extents(lattice), lattice.precision)
    check(lib().ym_su3_spatial_block(out.handle, lattice.handle), "ym_su3_spatial_block")
    return out


def scalar_a1pp_operator_basis(lattice, time_slice):
    """glueball_operators.rs:143-160 for one slice: (plaquette_1x1, rectangle_2x1)."""
    pl, rc = lattice.timeslice_a1pp()
    return pl[time_slice], rc[time_slice]


def derive_chain_seed(master_seed, chain_id):
    return lib().ym_derive_chain_seed(int(master_seed), int(chain_id))


def derive_su3_chain_seed(master_seed, chain_id):
    return lib().ym_derive_su3_chain_seed(int(master_seed), int(chain_id))


def diagnostics(values, max_lag):
    """su2_ensemble.rs:375-463: (ChainDiagnostics, autocorrelation array)."""
    v = np.ascontiguousarray(values, dtype=np.float64)
    ac = np.zeros(int(max_lag) + 1)
    d = ChainDiagnostics()
    check(lib().ym_diagnostics(_dp(v), v.size, int(max_lag), C.byref(d), _dp(ac)), "ym_diagnostics")
    return d, ac


@dataclass
class ChainReport:
    derived_seed: int
    total_sweeps: int
    proposals: int
    accepts: int
    mean_acceptance: float
    measurements: np.ndarray
    diagnostics: ChainDiagnostics
    final_unitarity_residual: float
    final_determinant_residual: float
    extra: dict = field(default_factory=dict)


def run_chain(extents, beta, master_seed, chain_id, thermalization_sweeps, measurements,
              sweeps_between_measurements, proposal_angle, autocorrelation_max_lag, hits=8,
              precision=32, hot_start=False, reunitarize=True):
    """su3_ensemble.rs:408-480 on the device."""
    cfg = Su3ChainConfig()
    cfg.extents[:] = [int(v) for v in extents]
    cfg.beta = float(beta)
    cfg.master_seed = int(master_seed)
    cfg.chain_id = int(chain_id)
    cfg.thermalization_sweeps = int(thermalization_sweeps)
    cfg.measurements = int(measurements)
    cfg.sweeps_between_measurements = int(sweeps_between_measurements)
    cfg.proposal_angle = float(proposal_angle)
    cfg.autocorrelation_max_lag = int(autocorrelation_max_lag)
    cfg.hits = int(hits)
    cfg.precision = int(precision)
    cfg.hot_start = 1 if hot_start else 0
    cfg.reunitarize = 1 if reunitarize else 0
    meas = np.zeros(int(measurements))
    rep = Su3ChainReport()
    check(lib().ym_su3_run_chain(C.byref(cfg), _dp(meas), C.byref(rep)), "ym_su3_run_chain")
    return ChainReport(rep.derived_seed, rep.total_sweeps, rep.proposals, rep.accepts, rep.mean_acceptance,
                       meas, rep.diagnostics, rep.final_unitarity_residual, rep.final_determinant_residual)


@dataclass
class Su2ChainReport:
    """su2_ensemble.rs:173-186 ChainReport."""
    derived_seed: int
    total_heatbath_sweeps: int
    total_overrelaxation_sweeps: int
    heatbath_proposals: int
    heatbath_updates: int
    mean_heatbath_acceptance: float
    measurements: np.ndarray
    diagnostics: ChainDiagnostics
    final_link_unit_residual: float


def su2_run_chain(extents, beta, master_seed, chain_id, thermalization_sweeps, measurements,
                  sweeps_between_measurements, overrelaxation_sweeps_per_heatbath=1,
                  max_heatbath_attempts_per_link=10_000, autocorrelation_max_lag=12, precision=64,
                  hot_start=False):
    """su2_ensemble.rs:488-552 run_chain on the device."""
    cfg = Su2ChainConfig()
    cfg.extents[:] = [int(v) for v in extents]
    cfg.beta = float(beta)
    cfg.master_seed = int(master_seed)
    cfg.chain_id = int(chain_id)
    cfg.thermalization_sweeps = int(thermalization_sweeps)
    cfg.measurements = int(measurements)
    cfg.sweeps_between_measurements = int(sweeps_between_measurements)
    cfg.overrelaxation_sweeps_per_heatbath = int(overrelaxation_sweeps_per_heatbath)
    cfg.max_heatbath_attempts_per_link = int(max_heatbath_attempts_per_link)
    cfg.autocorrelation_max_lag = int(autocorrelation_max_lag)
    cfg.precision = int(precision)
    cfg.hot_start = 1 if hot_start else 0
    meas = np.zeros(max(int(measurements), 0))
    rep = _Su2ChainReportC()
    check(lib().ym_su2_run_chain(C.byref(cfg), _dp(meas), C.byref(rep)), "ym_su2_run_chain")
    return Su2ChainReport(rep.derived_seed, rep.total_heatbath_sweeps, rep.total_overrelaxation_sweeps,
                          rep.heatbath_proposals, rep.heatbath_updates, rep.mean_heatbath_acceptance, meas,
                          rep.diagnostics, rep.final_link_unit_residual)
