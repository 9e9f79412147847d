"""The C-ABI boundary where it must behave as the reference's types and
drivers do.

* Uploads fail closed, before anything is stored, on links the reference's
  types cannot hold: Su3::from_checked_matrix (su3.rs:242-253) rejects
  non-finite entries, unitarity residuals and determinant residuals above
  5e-12; Su2 values (lib.rs:155-180) are finite, non-zero and unit.  FP32
  storage applies a 2^-16 tolerance.  The unchecked upload stores anything.
* ym_su2_run_chain (su2_ensemble.rs:488-552) equals the device-order replay
  of its Markov sweeps (one heat-bath sweep then the over-relaxation sweeps,
  su2_ensemble.rs:465-486) and validates its configuration in the
  reference's order (su2_ensemble.rs:70-103).
"""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu


def su3_phase(m18, eps):
    """Multiply a link by e^{i eps}: unitary, determinant e^{3 i eps}."""
    z = np.exp(1j * eps)
    m = m18.reshape(9, 2)
    c = (m[:, 0] + 1j * m[:, 1]) * z
    return np.stack([c.real, c.imag], axis=1).reshape(18)


@pytest.mark.parametrize("prec", [64, 32])
def test_su3_upload_fails_closed(ym, prec):
    n = (4, 4, 4, 4)
    lat = ym.Su3Lattice4(n, prec)
    lat.hot_start(ym.derive_su3_chain_seed(0xB0, prec))
    good = ym.Su3Lattice4(n, 64)
    good.hot_start(ym.derive_su3_chain_seed(0xB1, prec))
    aos = good.download()
    before = lat.download()
    k = O.link_offset(n, (1, 2, 3, 0), 2, 18)
    tol_ok, tol_bad = (1e-13, 1e-9) if prec == 64 else (1e-7, 1e-3)
    cases = []
    a = aos.copy(); a[k + 5] = np.nan; cases.append((a, 23))
    a = aos.copy(); a[k + 7] = np.inf; cases.append((a, 23))
    a = aos.copy(); a[k:k + 18] *= 1.0 + tol_bad; cases.append((a, 25))
    a = aos.copy(); a[k:k + 18] = su3_phase(aos[k:k + 18], tol_bad); cases.append((a, 26))
    for bad, code in cases:
        with pytest.raises(ym.YmError) as e:
            lat.upload(bad)
        assert e.value.status == code
        assert np.array_equal(lat.download(), before), "a rejected upload changed the lattice"
    # the first failing link in site order decides
    a = aos.copy()
    a[O.link_offset(n, (3, 3, 3, 3), 0, 18) + 1] = np.nan
    a[k:k + 18] *= 1.0 + tol_bad
    with pytest.raises(ym.YmError) as e:
        lat.upload(a)
    assert e.value.status == 25
    # within tolerance passes; the unchecked upload stores anything
    a = aos.copy(); a[k:k + 18] *= 1.0 + tol_ok
    lat.upload(a)
    a = aos.copy(); a[k + 5] = np.nan
    lat.upload(a, checked=False)
    assert np.isnan(lat.download()[k + 5])
    lat.upload(aos)
    got = lat.download()
    assert np.array_equal(got, aos) if prec == 64 else np.abs(got - aos).max() < 1e-7


@pytest.mark.parametrize("prec", [64, 32])
def test_su2_upload_fails_closed(ym, prec):
    n = (4, 4, 4, 4)
    lat = ym.Lattice4(n, prec)
    aos = O.su2_identity(n)
    before = lat.download()
    k = O.link_offset(n, (0, 1, 2, 3), 1, 4)
    big = 1e-9 if prec == 64 else 1e-3
    cases = []
    a = aos.copy(); a[k + 2] = np.nan; cases.append((a, 13))
    a = aos.copy(); a[k:k + 4] = 0.0; cases.append((a, 14))
    a = aos.copy(); a[k:k + 4] = [1.0 + big, 0.0, 0.0, 0.0]; cases.append((a, 15))
    a = aos.copy(); a[k:k + 4] = [0.6, 0.6, 0.6, 0.0]; cases.append((a, 15))
    for bad, code in cases:
        with pytest.raises(ym.YmError) as e:
            lat.upload(bad)
        assert e.value.status == code
        assert np.array_equal(lat.download(), before)
    a = aos.copy(); a[k:k + 4] = [0.0, 0.6, 0.0, 0.8]
    lat.upload(a)
    assert np.abs(lat.download() - a).max() < (1e-15 if prec == 64 else 1e-7)


def test_ensemble_upload_fails_closed(ym):
    ens = ym.Ensemble([(4, 4, 4, 4), (4, 4, 4, 6)], [11, 12], 32)
    lat = ym.Su3Lattice4((4, 4, 4, 6), 64)
    lat.hot_start(5)
    aos = lat.download()
    before = ens.download_chain(1)
    bad = aos.copy()
    bad[100] = np.nan
    with pytest.raises(ym.YmError) as e:
        ens.upload_chain(1, bad)
    assert e.value.status == 23
    assert np.array_equal(ens.download_chain(1), before)
    ens.upload_chain(1, aos)
    assert np.abs(ens.download_chain(1) - aos).max() < 1e-7


def _replay_su2_chain(n, beta, key, therm, meas, nor, prec):
    """The device-order replay of run_chain's Markov sweeps; plaquettes by the
    reference-exact oracle on the replayed links."""
    vh = O.lib().ymd_slots_per_parity(O.ext(n))
    buf = np.zeros(8 * 4 * vh, dtype=np.float32 if prec == 32 else np.float64)
    assert O.lib().ymd_su2_to_device(O.ext(n), O.dptr(O.su2_identity(n)), buf.ctypes.data_as(C.c_void_p), prec) == 0
    prop = upd = 0
    out = []
    for s in range(therm + meas):
        p, u = C.c_uint64(), C.c_uint64()
        assert O.lib().ymd_su2_heatbath_sweep(O.ext(n), buf.ctypes.data_as(C.c_void_p), prec, beta, key, s, 10_000,
                                              C.byref(p), C.byref(u)) == 0
        prop, upd = prop + p.value, upd + u.value
        for _ in range(nor):
            assert O.lib().ymd_su2_overrelax_sweep(O.ext(n), buf.ctypes.data_as(C.c_void_p), prec) == 0
        if s >= therm:
            aos = np.zeros(int(np.prod(n)) * 16)
            assert O.lib().ymd_su2_from_device(O.ext(n), buf.ctypes.data_as(C.c_void_p), prec, O.dptr(aos)) == 0
            v = C.c_double()
            assert O.lib().ymo_su2_average_plaquette(O.ext(n), O.dptr(aos), C.byref(v)) == 0
            out.append(v.value)
    return np.array(out), prop, upd


@pytest.mark.parametrize("prec,nor", [(64, 1), (64, 3), (32, 2)])
def test_su2_run_chain_matches_replay(ym, prec, nor):
    n, beta, master, chain = (4, 4, 6, 4), 2.3, 0x0123456789ABCDEF, 5
    therm, meas = 3, 5
    rep = ym.su2_run_chain(n, beta, master, chain, therm, meas, 1, nor, 10_000, 2, precision=prec)
    key = ym.derive_chain_seed(master, chain)
    assert rep.derived_seed == key
    want, prop, upd = _replay_su2_chain(n, beta, key, therm, meas, nor, prec)
    assert rep.heatbath_proposals == prop and rep.heatbath_updates == upd
    assert rep.total_heatbath_sweeps == therm + meas and rep.total_overrelaxation_sweeps == (therm + meas) * nor
    assert np.abs(rep.measurements - want).max() <= (1e-13 if prec == 64 else 1e-6)
    assert rep.mean_heatbath_acceptance == upd / prop
    d, _ = ym.diagnostics(rep.measurements, 2)
    assert d.mean == rep.diagnostics.mean and d.standard_error == rep.diagnostics.standard_error


def test_su2_markov_sweeps_equal_heatbath_then_overrelaxation(ym):
    n = (6, 4, 4, 8)
    a, b = ym.Lattice4(n, 32), ym.Lattice4(n, 32)
    a.hot_start(9)
    b.hot_start(9)
    pa, ua = a.markov_sweeps(2.5, 77, 10, 3, 2)
    pb = ub = 0
    for k in range(3):
        p, u = b.heatbath_sweeps(2.5, 77, 10 + k, 1)
        pb, ub = pb + p, ub + u
        b.overrelax_sweeps(2)
    assert (pa, ua) == (pb, ub)
    assert np.array_equal(a.download(), b.download())


def test_su2_run_chain_validation_order(ym):
    base = dict(extents=(4, 4, 4, 4), beta=2.3, master_seed=1, chain_id=0, thermalization_sweeps=2,
                measurements=4, sweeps_between_measurements=1, overrelaxation_sweeps_per_heatbath=1,
                max_heatbath_attempts_per_link=100, autocorrelation_max_lag=2)
    cases = [(dict(beta=8.5), 30), (dict(beta=float("nan")), 30), (dict(measurements=(1 << 20) + 1), 35),
             (dict(thermalization_sweeps=(1 << 20)), 35), (dict(measurements=1, autocorrelation_max_lag=0), 32),
             (dict(sweeps_between_measurements=0), 33), (dict(overrelaxation_sweeps_per_heatbath=17), 40),
             (dict(max_heatbath_attempts_per_link=0), 41), (dict(autocorrelation_max_lag=4), 34),
             # the first failing check wins: beta before the schedule
             (dict(beta=9.0, sweeps_between_measurements=0), 30)]
    for over, code in cases:
        kw = dict(base)
        kw.update(over)
        with pytest.raises(ym.YmError) as e:
            ym.su2_run_chain(**kw)
        assert e.value.status == code, (over, e.value.status)
