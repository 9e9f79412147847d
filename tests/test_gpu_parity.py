"""GPU parity tests: libym_amd.so (HIP, gfx950) against the CPU oracle.

Two bars:
  * bit-exact against the device-order replay (oracle/dev_exact.c): hot start,
    SU(3) Metropolis sweeps (FP32 and FP64 storage) and SU(2) heat-bath /
    over-relaxation sweeps reproduce every link bit and every accept count;
  * FP64 storage against the reference-exact restatement (oracle/ref_*.c) on
    identical configurations: plaquette, action, staples, Delta S, Wilson
    loops, Polyakov loops, A1++ operators and stout smearing agree to
    1e-13 relative (absolute where the value can vanish).
"""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

REL = 1e-13


def dev_buffer(n, nc, prec):
    vh = O.lib().ymd_slots_per_parity(O.ext(n))
    return np.zeros(8 * nc * vh, dtype=np.float32 if prec == 32 else np.float64)


def dev_to_aos(n, buf, prec, nc):
    aos = np.zeros(int(np.prod(n)) * 4 * nc)
    f = O.lib().ymd_su3_from_device if nc == 18 else O.lib().ymd_su2_from_device
    assert f(O.ext(n), buf.ctypes.data_as(C.c_void_p), prec, O.dptr(aos)) == 0
    return aos


def aos_to_dev(n, aos, prec, nc):
    buf = dev_buffer(n, nc, prec)
    f = O.lib().ymd_su3_to_device if nc == 18 else O.lib().ymd_su2_to_device
    assert f(O.ext(n), O.dptr(np.ascontiguousarray(aos)), buf.ctypes.data_as(C.c_void_p), prec) == 0
    return buf


def assert_rel(a, b, tol=REL, scale=None):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    s = scale if scale is not None else max(1.0, float(np.max(np.abs(b))))
    err = float(np.max(np.abs(a - b))) / s
    assert err <= tol, f"max error {err:.3e} > {tol:.1e}"


# ------------------------------------------------------------- round trips
@pytest.mark.parametrize("group,prec", [(3, 32), (3, 64), (2, 32), (2, 64)])
def test_upload_download_round_trip(ym, group, prec):
    n = (4, 3, 5, 6)
    cls = ym.Su3Lattice4 if group == 3 else ym.Lattice4
    lat = cls(n, prec)
    rng = np.random.default_rng(7)
    a = rng.standard_normal(lat.link_count() * cls.NCOMP)
    lat.upload(a, checked=False)
    b = lat.download()
    expect = a if prec == 64 else a.astype(np.float32).astype(np.float64)
    assert np.array_equal(b, expect)


@pytest.mark.parametrize("prec", [32, 64])
def test_su3_hot_start_matches_replay(ym, prec):
    n = (4, 4, 2, 6)
    key = ym.derive_su3_chain_seed(0x1234, 5)
    lat = ym.Su3Lattice4(n, prec)
    lat.hot_start(key)
    buf = dev_buffer(n, 18, prec)
    assert O.lib().ymd_su3_hot_start(O.ext(n), buf.ctypes.data_as(C.c_void_p), prec, key) == 0
    assert np.array_equal(lat.download(), dev_to_aos(n, buf, prec, 18))
    unit, det = lat.link_residuals()
    assert unit < (1e-6 if prec == 32 else 1e-14) and det < (1e-6 if prec == 32 else 1e-14)


@pytest.mark.parametrize("prec", [32, 64])
def test_su2_hot_start_matches_replay(ym, prec):
    n = (4, 2, 4, 4)
    lat = ym.Lattice4(n, prec)
    lat.hot_start(99)
    buf = dev_buffer(n, 4, prec)
    assert O.lib().ymd_su2_hot_start(O.ext(n), buf.ctypes.data_as(C.c_void_p), prec, 99) == 0
    assert np.array_equal(lat.download(), dev_to_aos(n, buf, prec, 4))


# ------------------------------------------------ decision-for-decision chains
# angle <= 0.5 runs the short sincos series, larger angles the full one
@pytest.mark.parametrize("prec,n,beta,hits,angle", [
    (32, (4, 4, 4, 4), 6.0, 8, 0.35),
    (32, (2, 4, 6, 4), 5.7, 10, 0.35),
    (64, (4, 4, 4, 4), 6.0, 8, 0.35),
    (64, (4, 2, 2, 8), 0.0, 3, 0.35),
    (64, (4, 4, 4, 4), 5.7, 10, 0.5),
    (32, (4, 4, 4, 4), 6.0, 10, 0.8),
    (64, (2, 4, 4, 4), 6.0, 8, 1.0),
])
def test_su3_metropolis_matches_replay_bitwise(ym, prec, n, beta, hits, angle):
    key = ym.derive_su3_chain_seed(0x2c02198105010002, 1)
    lat = ym.Su3Lattice4(n, prec)
    lat.hot_start(key ^ 0x55)
    buf = dev_buffer(n, 18, prec)
    assert O.lib().ymd_su3_hot_start(O.ext(n), buf.ctypes.data_as(C.c_void_p), prec, key ^ 0x55) == 0
    params = O.DevSu3Params(beta=beta, angle=angle, hits=hits, reunitarize=1, key=key, sweep=0)
    for sweep in range(3):
        gp, ga = lat.metropolis_sweeps(beta, angle, key, sweep, 1, hits)
        params.sweep = sweep
        cp, ca = C.c_uint64(), C.c_uint64()
        st = O.lib().ymd_su3_metropolis_sweep(O.ext(n), buf.ctypes.data_as(C.c_void_p), prec, C.byref(params),
                                             C.byref(cp), C.byref(ca))
        assert st == 0
        assert gp == cp.value and ga == ca.value, f"sweep {sweep}: gpu {ga}/{gp} replay {ca.value}/{cp.value}"
        assert np.array_equal(lat.download(), dev_to_aos(n, buf, prec, 18)), f"links differ after sweep {sweep}"
    if beta == 0.0:
        assert ga == gp  # every proposal accepted at beta = 0


@pytest.mark.parametrize("prec,hits,angle", [(32, 8, 0.35), (64, 8, 0.35), (32, 10, 0.35), (64, 10, 0.35),
                                              (32, 3, 0.8)])
def test_su3_metropolis_wave_specialised_path_bitwise(ym, prec, hits, angle):
    """Lattices of >= 256 tiles take the wave-specialised sub-pass kernel
    (staples of the next tile overlapping the hits of the current one; 8 and
    10 hits have unrolled instances, other counts loop); one sweep of
    16^3x32 against the device-order replay."""
    n = (16, 16, 16, 32)
    key = ym.derive_su3_chain_seed(0x2c02198105010002, 2)
    lat = ym.Su3Lattice4(n, prec)
    lat.hot_start(key ^ 0x77)
    buf = dev_buffer(n, 18, prec)
    assert O.lib().ymd_su3_hot_start(O.ext(n), buf.ctypes.data_as(C.c_void_p), prec, key ^ 0x77) == 0
    gp, ga = lat.metropolis_sweeps(6.0, angle, key, 5, 1, hits)
    params = O.DevSu3Params(beta=6.0, angle=angle, hits=hits, reunitarize=1, key=key, sweep=5)
    cp, ca = C.c_uint64(), C.c_uint64()
    assert O.lib().ymd_su3_metropolis_sweep(O.ext(n), buf.ctypes.data_as(C.c_void_p), prec, C.byref(params),
                                           C.byref(cp), C.byref(ca)) == 0
    assert gp == cp.value and ga == ca.value
    assert np.array_equal(lat.download(), dev_to_aos(n, buf, prec, 18))


@pytest.mark.parametrize("prec", [32, 64])
def test_su2_heatbath_overrelax_match_replay_bitwise(ym, prec):
    n = (4, 4, 2, 6)
    key = ym.derive_chain_seed(0x2b00201211160101, 0)
    lat = ym.Lattice4(n, prec)
    buf = dev_buffer(n, 4, prec)
    aos = O.su2_identity(n)
    buf = aos_to_dev(n, aos, prec, 4)
    for sweep in range(3):
        gp, gu = lat.heatbath_sweeps(2.3, key, sweep, 1, 10_000)
        cp, cu = C.c_uint64(), C.c_uint64()
        st = O.lib().ymd_su2_heatbath_sweep(O.ext(n), buf.ctypes.data_as(C.c_void_p), prec, 2.3, key, sweep,
                                           10_000, C.byref(cp), C.byref(cu))
        assert st == 0
        assert (gp, gu) == (cp.value, cu.value)
        assert np.array_equal(lat.download(), dev_to_aos(n, buf, prec, 4))
        lat.overrelax_sweeps(1)
        assert O.lib().ymd_su2_overrelax_sweep(O.ext(n), buf.ctypes.data_as(C.c_void_p), prec) == 0
        assert np.array_equal(lat.download(), dev_to_aos(n, buf, prec, 4))


@pytest.mark.parametrize("prec,beta", [(32, 2.3), (64, 2.5), (32, 0.5)])
def test_su2_heatbath_hot_start_matches_replay_bitwise(ym, prec, beta):
    """Full waves with varied staples (hot start, 8^3 x 6): every link's
    accepted attempt, hence the proposal count and the new link, equals the
    sequential rejection loop of the replay, whichever lanes evaluated it."""
    n = (8, 8, 8, 6)
    key = ym.derive_chain_seed(0x2b00201211160101, 9)
    lat = ym.Lattice4(n, prec)
    lat.hot_start(key)
    buf = dev_buffer(n, 4, prec)
    assert O.lib().ymd_su2_hot_start(O.ext(n), buf.ctypes.data_as(C.c_void_p), prec, key) == 0
    for sweep in range(2):
        gp, gu = lat.heatbath_sweeps(beta, key, sweep, 1, 10_000)
        cp, cu = C.c_uint64(), C.c_uint64()
        assert O.lib().ymd_su2_heatbath_sweep(O.ext(n), buf.ctypes.data_as(C.c_void_p), prec, beta, key, sweep,
                                              10_000, C.byref(cp), C.byref(cu)) == 0
        assert (gp, gu) == (cp.value, cu.value) and gp > gu
        assert np.array_equal(lat.download(), dev_to_aos(n, buf, prec, 4))


def test_su2_heatbath_attempt_limit_reported(ym):
    # A cap of one attempt at large alpha rejects most first attempts: the
    # sweep must return HeatbathAttemptLimitExceeded (status 42), not hang.
    lat = ym.Lattice4((4, 4, 4, 4), 32)
    with pytest.raises(ym.YmError) as e:
        lat.heatbath_sweeps(8.0, 1, 0, 1, 1)
    assert e.value.status == 42


@pytest.mark.parametrize("prec", [32, 64])
def test_su3_staple_matches_replay_bitwise(ym, prec):
    n = (4, 4, 4, 6)
    lat = ym.Su3Lattice4(n, prec)
    lat.hot_start(3)
    buf = dev_buffer(n, 18, prec)
    O.lib().ymd_su3_hot_start(O.ext(n), buf.ctypes.data_as(C.c_void_p), prec, 3)
    for x, mu in [((0, 0, 0, 0), 0), ((1, 2, 3, 5), 3), ((3, 3, 0, 1), 2)]:
        g = lat.staple(x, mu)
        c = np.zeros(18)
        xs = (C.c_int64 * 4)(*x)
        assert O.lib().ymd_su3_staple(O.ext(n), buf.ctypes.data_as(C.c_void_p), prec, xs, mu, O.dptr(c)) == 0
        assert np.array_equal(g, c)


def test_su3_fp32_staple_from_two_rows_vs_fp64_full_products(ym):
    """The device's FP32 staple (rows 0 and 1 of each neighbour, third rows
    rebuilt) against the staple of the whole stored links formed in FP64,
    on the same FP32 hot-start links: equal to FP32 rounding."""
    n = (4, 4, 4, 6)
    lat = ym.Su3Lattice4(n, 32)
    lat.hot_start(9)
    aos = lat.download()

    def link(y, d):
        o = O.link_offset(n, tuple(v % n[k] for k, v in enumerate(y)), d, 18)
        return O.mat_from18(aos[o:o + 18])

    def sh(y, d, s):
        z = list(y)
        z[d] += s
        return z

    for x, mu in [((0, 0, 0, 0), 0), ((1, 2, 3, 5), 3), ((3, 3, 0, 1), 2), ((2, 1, 3, 4), 1)]:
        ref = np.zeros((3, 3), dtype=complex)
        for nu in range(4):
            if nu == mu:
                continue
            xpm = sh(x, mu, 1)
            ref += link(xpm, nu) @ link(sh(x, nu, 1), mu).conj().T @ link(x, nu).conj().T
            ref += link(sh(xpm, nu, -1), nu).conj().T @ link(sh(x, nu, -1), mu).conj().T @ link(sh(x, nu, -1), nu)
        got = O.mat_from18(lat.staple(x, mu))
        assert np.abs(got - ref).max() <= 2e-6 * np.abs(ref).max(), (x, mu)


def test_fp32_sweep_never_reads_a_stored_third_row(ym):
    """FP32 sweeps read rows 0 and 1 of every link (staple neighbours and the
    updated link) and rebuild third rows: a field whose stored third rows are
    overwritten sweeps to the same links, bit for bit, once every link has
    been updated (one full sweep), and so do the accept counts."""
    n = (4, 4, 4, 6)
    clean = ym.Su3Lattice4(n, 32)
    clean.hot_start(21)
    aos = clean.download()
    bad = aos.reshape(-1, 18).copy()
    bad[:, 12:18] = 0.5
    dirty = ym.Su3Lattice4(n, 32)
    dirty.upload(bad.reshape(-1), checked=False)
    for lat in (clean, dirty):
        lat.metropolis_sweeps(6.0, 0.35, 5, 0, 1, 10)
    a, b = clean.download(), dirty.download()
    assert np.array_equal(a, b)
    pa, ca = clean.metropolis_sweeps(6.0, 0.35, 5, 1, 2, 8)
    pb, cb = dirty.metropolis_sweeps(6.0, 0.35, 5, 1, 2, 8)
    assert (pa, ca) == (pb, cb)
    assert np.array_equal(clean.download(), dirty.download())


def test_determinism_same_key_same_links(ym):
    n = (4, 4, 4, 4)
    a, b = ym.Su3Lattice4(n, 32), ym.Su3Lattice4(n, 32)
    for lat in (a, b):
        lat.hot_start(11)
        lat.metropolis_sweeps(6.0, 0.35, 77, 0, 2, 8)
    assert np.array_equal(a.download(), b.download())


# -------------------------------------------- FP64 vs the reference restatement
def hot_fp64(ym, n, seed, group=3):
    cls = ym.Su3Lattice4 if group == 3 else ym.Lattice4
    lat = cls(n, 64)
    lat.hot_start(seed)
    return lat, lat.download()


def test_su3_plaquette_and_action_fp64(ym):
    n = (4, 4, 4, 6)
    lat, aos = hot_fp64(ym, n, 21)
    lat.metropolis_sweeps(6.0, 0.35, 5, 0, 2, 8, True)
    aos = lat.download()
    ref = C.c_double()
    assert O.lib().ymo_su3_average_plaquette(O.ext(n), O.dptr(aos), C.byref(ref)) == 0
    assert_rel(lat.average_plaquette(), ref.value)
    assert O.lib().ymo_su3_wilson_action(O.ext(n), O.dptr(aos), 6.0, C.byref(ref)) == 0
    assert_rel(lat.wilson_action(6.0), ref.value, scale=abs(ref.value))


@pytest.mark.parametrize("group,n", [(3, (16, 16, 16, 32)), (3, (8, 8, 32, 16)), (2, (32, 32, 32, 32))])
def test_fp32_plaquette_tile_order_vs_fp64(ym, group, n):
    """FP32 links on lattices whose x0-slices are a multiple of 8 tiles, so the
    plaquette kernel deals x1-bands to the XCDs (and A1++ runs its slot-order
    site pass): the results equal those of an FP64 lattice holding the same
    values (the FP64 kernels are pinned against the reference restatement)
    to FP32 product rounding; a missing or doubled tile would move the
    plaquette by ~1/tiles.  (The reference restatement itself rejects FP32
    links: its products check unitarity to 5e-12.)"""
    cls = ym.Su3Lattice4 if group == 3 else ym.Lattice4
    lat = cls(n, 32)
    lat.hot_start(77)
    if group == 3:
        lat.metropolis_sweeps(6.0, 0.35, 78, 0, 1, 8)
    else:
        lat.heatbath_sweeps(2.3, 78, 0, 1)
    twin = cls(n, 64)
    twin.upload(lat.download(), checked=False)  # the FP32 values as they are
    assert abs(lat.average_plaquette() - twin.average_plaquette()) <= 1e-6
    if group == 3 and n[0] == n[1] == n[2]:
        pl, rc = lat.timeslice_a1pp()
        pl64, rc64 = twin.timeslice_a1pp()
        assert np.max(np.abs(pl - pl64)) <= 1e-6 and np.max(np.abs(rc - rc64)) <= 1e-6


def test_su2_plaquette_and_staple_fp64(ym):
    n = (4, 4, 2, 6)
    lat, aos = hot_fp64(ym, n, 4, group=2)
    ref = C.c_double()
    assert O.lib().ymo_su2_average_plaquette(O.ext(n), O.dptr(aos), C.byref(ref)) == 0
    assert_rel(lat.average_plaquette(), ref.value)
    for x, mu in [((0, 0, 0, 0), 0), ((3, 1, 1, 5), 2)]:
        out = np.zeros(4)
        assert O.lib().ymo_su2_staple(O.ext(n), O.dptr(aos), (C.c_int64 * 4)(*x), mu, O.dptr(out)) == 0
        assert_rel(lat.staple(x, mu), out)


def test_su3_staple_reproduces_reference_local_action_and_delta(ym):
    """Re Tr(U A)/3 summed with the six plaquettes (su3_ensemble.rs:261-334)."""
    n = (4, 4, 4, 4)
    lat, aos = hot_fp64(ym, n, 8)
    beta = 5.7
    for x, mu in [((0, 0, 0, 0), 2), ((1, 3, 2, 0), 0), ((3, 3, 3, 3), 3)]:
        A = O.mat_from18(lat.staple(x, mu))
        off = O.link_offset(n, x, mu, 18)
        U = O.mat_from18(aos[off:off + 18])
        local = beta * (6.0 - np.trace(U @ A).real / 3.0)
        ref = C.c_double()
        assert O.lib().ymo_su3_local_link_action(O.ext(n), O.dptr(aos), (C.c_int64 * 4)(*x), mu, beta, C.byref(ref)) == 0
        assert_rel(local, ref.value, scale=abs(ref.value))
        coef = np.zeros(8)
        coef[4] = 0.21
        R = O.mat_from18(O.su3_from_algebra(coef))
        cand = R @ U
        dS = -(beta / 3.0) * np.trace((cand - U) @ A).real
        assert O.lib().ymo_su3_local_action_difference(O.ext(n), O.dptr(aos), (C.c_int64 * 4)(*x), mu,
                                                       O.dptr(O.mat_to18(cand)), beta, C.byref(ref)) == 0
        assert abs(dS - ref.value) <= 1e-12 * max(1.0, abs(ref.value))


@pytest.mark.parametrize("group", [3, 2])
def test_wilson_loops_fp64_vs_reference(ym, group):
    n = (4, 4, 4, 4)
    lat, aos = hot_fp64(ym, n, 31, group)
    rmax, tmax = 3, 2
    for plane_set, planes in [(0, [(m, k) for m in range(4) for k in range(4) if m != k]),
                              (1, [(0, 3), (1, 3), (2, 3)])]:
        re, im = lat.wilson_loops(rmax, tmax, plane_set)
        for r in range(1, rmax + 1):
            for t in range(1, tmax + 1):
                acc_re = acc_im = 0.0
                for (m, k) in planes:
                    for site in np.ndindex(*n):
                        s = (C.c_int64 * 4)(*site)
                        if group == 3:
                            a, b = C.c_double(), C.c_double()
                            assert O.lib().ymo_su3_rect_wilson_loop(O.ext(n), O.dptr(aos), s, m, k, r, t, C.byref(a), C.byref(b)) == 0
                            acc_re += a.value
                            acc_im += b.value
                        else:
                            a = C.c_double()
                            assert O.lib().ymo_su2_rect_wilson_loop(O.ext(n), O.dptr(aos), s, m, k, r, t, C.byref(a)) == 0
                            acc_re += a.value
                cnt = len(planes) * int(np.prod(n))
                assert abs(re[r - 1, t - 1] - acc_re / cnt) <= REL
                assert abs(im[r - 1, t - 1] - acc_im / cnt) <= REL


@pytest.mark.parametrize("group", [3, 2])
def test_path_holonomy_fp64_vs_reference(ym, group):
    """Open and closed paths from several starts against the oracle's
    ordered product (path_holonomy.rs:132-157, lib.rs:384-416); the closed
    r x t rectangle must reproduce the Wilson loop of the same plane, and a
    path that walks out and back must give the identity."""
    n = (4, 3, 5, 6)
    lat, aos = hot_fp64(ym, n, 47, group)
    rng = np.random.default_rng(5)
    ncomp = 18 if group == 3 else 4
    fn = O.lib().ymo_su3_path_holonomy if group == 3 else O.lib().ymo_su2_path_holonomy
    cases = [(0, 0, 0, 0), (3, 2, 4, 5), (1, 0, 2, 3)]
    for start in cases:
        for length in (1, 7, 40):
            steps = rng.choice([1, 2, 3, 4, -1, -2, -3, -4], size=length).astype(np.int32)
            got, end = lat.path_holonomy(start, steps)
            ref = np.zeros(ncomp)
            rend = (C.c_int64 * 4)()
            assert fn(O.ext(n), O.dptr(aos), (C.c_int64 * 4)(*start),
                      steps.ctypes.data_as(C.POINTER(C.c_int32)), length, O.dptr(ref), rend) == 0
            assert_rel(got, ref)
            assert end == tuple(rend)
            # closed-form end point: start + net displacement, periodic
            disp = np.zeros(4, np.int64)
            for s in steps:
                disp[abs(s) - 1] += 1 if s > 0 else -1
            assert end == tuple(int(v) for v in (np.asarray(start) + disp) % np.asarray(n))
    # out and back: P followed by reverse(-P) is the identity
    steps = np.array([1, 2, 2, 4, -3, 1], np.int32)
    back = np.concatenate([steps, -steps[::-1]])
    got, end = lat.path_holonomy((2, 1, 3, 0), back)
    ident = np.zeros(ncomp)
    if group == 3:
        ident[[0, 8, 16]] = 1.0
    else:
        ident[0] = 1.0
    assert_rel(got, ident, tol=1e-12)
    assert end == (2, 1, 3, 0)
    # closed 2x1 rectangle in plane (0, 3) from one site = that site's Wilson loop
    s = (1, 2, 0, 4)
    rect = np.array([1, 1, 4, -1, -1, -4], np.int32)
    got, _ = lat.path_holonomy(s, rect)
    if group == 3:
        a, b = C.c_double(), C.c_double()
        assert O.lib().ymo_su3_rect_wilson_loop(O.ext(n), O.dptr(aos), (C.c_int64 * 4)(*s), 0, 3, 2, 1,
                                                C.byref(a), C.byref(b)) == 0
        tr = complex(got[0] + got[8] + got[16], got[1] + got[9] + got[17]) / 3
        assert abs(tr - complex(a.value, b.value)) <= REL
    else:
        a = C.c_double()
        assert O.lib().ymo_su2_rect_wilson_loop(O.ext(n), O.dptr(aos), (C.c_int64 * 4)(*s), 0, 3, 2, 1,
                                                C.byref(a)) == 0
        assert abs(got[0] - a.value) <= REL
    # the reference's error variants
    with pytest.raises(ym.YmError) as e:
        lat.path_holonomy((0, 0, 0, 0), np.zeros(0, np.int32))
    assert e.value.status == 9
    with pytest.raises(ym.YmError) as e:
        lat.path_holonomy((0, 0, 0, 0), np.array([1, 5], np.int32))
    assert e.value.status == 4
    with pytest.raises(ym.YmError) as e:
        lat.path_holonomy((4, 0, 0, 0), np.array([1], np.int32))
    assert e.value.status == 17
    with pytest.raises(ym.YmError) as e:
        lat.path_holonomy((0, 0, 0, 0), np.ones((1 << 20) + 1, np.int32))
    assert e.value.status == 10


@pytest.mark.parametrize("group", [3, 2])
def test_polyakov_fp64_vs_reference(ym, group):
    n = (4, 2, 4, 6)
    lat, aos = hot_fp64(ym, n, 12, group)
    for d in range(4):
        val, ab = lat.polyakov_loop(d)
        acc = 0j
        accabs = 0.0
        cnt = 0
        for site in np.ndindex(*n):
            if site[d] != 0:
                continue
            s = (C.c_int64 * 4)(*site)
            if group == 3:
                a, b = C.c_double(), C.c_double()
                assert O.lib().ymo_su3_polyakov_loop(O.ext(n), O.dptr(aos), s, d, C.byref(a), C.byref(b)) == 0
                z = complex(a.value, b.value)
            else:
                a = C.c_double()
                assert O.lib().ymo_su2_polyakov_loop(O.ext(n), O.dptr(aos), s, d, C.byref(a)) == 0
                z = complex(a.value, 0.0)
            acc += z
            accabs += abs(z)
            cnt += 1
        assert abs(val - acc / cnt) <= REL
        assert abs(ab - accabs / cnt) <= REL


@pytest.mark.parametrize("prec", [64, 32])
def test_polyakov_centre_slice_and_closed_form(ym, prec):
    """ym_polyakov_loop (mean over the base) under the centre-slice transform
    (P -> z P, tests/su3_center.rs:97-120) and on diagonal fields whose loop
    has a closed form (test_polyakov_pin.py pins the oracle the same way)."""
    import test_polyakov_pin as PP
    n = (4, 4, 6, 8)
    tol = 1e-13 if prec == 64 else 5e-6
    links = PP.hot_su3(n, 0xABC)
    lat = ym.Su3Lattice4(n, prec)
    lat.upload(links)
    other = ym.Su3Lattice4(n, prec)
    p0 = lat.average_plaquette()
    for d in range(4):
        before, ab = lat.polyakov_loop(d)
        for element in (1, 2):
            other.upload(PP.centre_transform(n, links, d, n[d] - 1, element))
            after, ab2 = other.polyakov_loop(d)
            assert abs(after - PP.Z[element] * before) <= tol, (d, element)
            assert abs(ab2 - ab) <= tol
            assert abs(other.average_plaquette() - p0) <= tol
    for d in range(4):
        diag, ang = PP.diagonal_field(n, d, 70 + d)
        lat.upload(diag)
        val, ab = lat.polyakov_loop(d)
        base = [x for x in np.ndindex(*n) if x[d] == 0]
        want = np.mean([PP.closed_form(n, ang, d, x) for x in base])
        want_abs = np.mean([abs(PP.closed_form(n, ang, d, x)) for x in base])
        assert abs(val - want) <= tol, d
        assert abs(ab - want_abs) <= tol, d


def test_a1pp_fp64_vs_reference(ym):
    n = (4, 4, 4, 5)
    lat, aos = hot_fp64(ym, n, 17)
    pl, rc = lat.timeslice_a1pp()
    for t in range(n[3]):
        p, r = C.c_double(), C.c_double()
        pc, rc_ = C.c_int64(), C.c_int64()
        assert O.lib().ymo_su3_a1pp_basis(O.ext(n), O.dptr(aos), t, C.byref(p), C.byref(r), C.byref(pc), C.byref(rc_)) == 0
        assert abs(pl[t] - p.value) <= REL and abs(rc[t] - r.value) <= REL


@pytest.mark.parametrize("sweeps,levels", [(0, 1), (0, 3), (20, 1), (20, 4)])
def test_stout_fp32_closed_form_against_fp64_series(ym, sweeps, levels):
    """FP32 storage smears with exp(iQ) by Cayley-Hamilton (Morningstar-Peardon),
    FP64 storage with the reference's scaled Taylor series (pinned to the
    reference restatement at 1e-13).  On the same links -- a hot start, where
    Q ranges widely and Tr Q^3 takes both signs, and a configuration after 20
    sweeps -- the two agree link by link to FP32 rounding; temporal links
    are copied unchanged."""
    n = (8, 8, 8, 6)
    lat = ym.Su3Lattice4(n, 32)
    lat.hot_start(31)
    if sweeps:
        lat.metropolis_sweeps(6.0, 0.35, 5, 0, sweeps, 8)
    links = lat.download()
    twin = ym.Su3Lattice4(n, 64)
    twin.upload(links, checked=False)
    a = ym.spatial_stout_smear(lat, 0.1, levels).download().reshape(-1, 4, 18)
    b = ym.spatial_stout_smear(twin, 0.1, levels).download().reshape(-1, 4, 18)
    assert np.abs(a - b).max() < 1e-6 * (1 + levels), np.abs(a - b).max()
    assert np.array_equal(a[:, 3], links.reshape(-1, 4, 18)[:, 3])
    assert np.abs(a[:, :3] - links.reshape(-1, 4, 18)[:, :3]).max() > 0.01  # the spatial links moved


def test_stout_frozen_cycle_fixture(ym):
    """glueball-spatial-stout-phase3b1-v1.json: one level at rho = 0.1."""
    n = (3, 3, 3, 3)
    aos = O.su3_identity(n)
    cyc = np.zeros(18)
    O.lib().ymo_su3_traceless_cycle_link(O.dptr(cyc))
    aos[0:18] = cyc
    src = ym.Su3Lattice4(n, 64)
    src.upload(aos)
    out = ym.spatial_stout_smear(src, 0.1, 1)
    sm = out.download()
    d = (sm - aos).reshape(-1, 9, 2)
    change = float(np.max(np.hypot(d[..., 0], d[..., 1])))  # max_link_residual, glueball_smearing.rs:270
    assert abs(change - 0.2158247298307305) <= 1e-12
    pl, rc = out.timeslice_a1pp()
    assert abs(pl[0] - 0.9693015447000209) <= 1e-12
    assert abs(rc[0] - 0.9511047606560571) <= 1e-12
    # temporal links unchanged, bit for bit
    assert np.array_equal(sm.reshape(-1, 4, 18)[:, 3], aos.reshape(-1, 4, 18)[:, 3])


@pytest.mark.parametrize("levels", [1, 3])
def test_stout_fp64_vs_reference_hot(ym, levels):
    n = (4, 4, 4, 4)
    lat, aos = hot_fp64(ym, n, 23)
    lat.metropolis_sweeps(6.0, 0.35, 3, 0, 2, 8, True)
    aos = lat.download()
    out = ym.spatial_stout_smear(lat, 0.1, levels).download()
    ref = np.zeros_like(aos)
    assert O.lib().ymo_su3_stout_smear(O.ext(n), O.dptr(aos), O.dptr(ref), 0.1, levels, 0) == 0
    assert float(np.max(np.abs(out - ref))) <= 1e-13


def test_spatial_block_frozen_fixture(ym):
    """glueball-spatial-block-phase3b2-v1.json on the device: identity blocks
    to identity, the cycle fixture gives the frozen A1++ outputs (1e-12), the
    temporal link is sampled exactly, geometries are rejected in order."""
    n = (6, 6, 6, 3)
    fine = ym.Su3Lattice4(n, 64)
    coarse = ym.spatial_stride2_block(fine)
    assert coarse.extents == (3, 3, 3, 3)
    assert np.array_equal(coarse.download(), O.su3_identity((3, 3, 3, 3)))
    cyc = np.zeros(18)
    O.lib().ymo_su3_traceless_cycle_link(O.dptr(cyc))
    aos = O.su3_identity(n)
    aos[0:18] = cyc
    off3 = O.link_offset(n, (2, 4, 0, 1), 3, 18)
    aos[off3:off3 + 18] = cyc
    fine.upload(aos)
    coarse = ym.spatial_stride2_block(fine)
    pl, rc = coarse.timeslice_a1pp()
    assert abs(pl[0] - 0.9506172839506173) <= 1e-12 and abs(rc[0] - 0.9259259259259259) <= 1e-12
    got = coarse.download()
    assert np.array_equal(got[O.link_offset((3, 3, 3, 3), (1, 2, 0, 1), 3, 18):][:18], cyc)
    for bad, code in (((4, 4, 4, 3), 51), ((7, 7, 7, 3), 55), ((6, 8, 6, 3), 52)):
        with pytest.raises(ym.YmError) as e:
            ym.spatial_stride2_block(ym.Su3Lattice4(bad, 64))
        assert e.value.status == code


@pytest.mark.parametrize("prec,n", [(64, (8, 8, 8, 6)), (64, (6, 6, 6, 5)), (32, (12, 12, 12, 8))])
def test_spatial_block_vs_reference(ym, prec, n):
    """Hot-start configurations (odd N3 included): FP64 storage matches the
    reference-exact blocking to 1e-13; FP32 storage to its rounding."""
    lat = ym.Su3Lattice4(n, prec)
    lat.hot_start(ym.derive_su3_chain_seed(0xB10C, prec))
    aos = lat.download()
    ref_aos = aos.copy()
    if prec == 32:  # the reference's 5e-12 group check needs FP64-unitary links
        r64 = ym.Su3Lattice4(n, 64)
        r64.upload(aos, checked=False)
        r64.reunitarize()
        ref_aos = r64.download()
    nc = (n[0] // 2, n[1] // 2, n[2] // 2, n[3])
    ref = np.zeros(int(np.prod(nc)) * 4 * 18)
    assert O.lib().ymo_su3_spatial_block(O.ext(n), O.dptr(ref_aos), O.dptr(ref)) == 0
    got = ym.spatial_stride2_block(lat).download()
    assert float(np.max(np.abs(got - ref))) <= (1e-13 if prec == 64 else 2e-6)
    if prec == 64:
        u, d = ym.spatial_stride2_block(lat).link_residuals()
        assert u <= 1e-13 and d <= 1e-13


def test_residuals_fp64_vs_reference(ym):
    n = (4, 4, 2, 4)
    lat, aos = hot_fp64(ym, n, 2)
    u, d = lat.link_residuals()
    ru, rd = C.c_double(), C.c_double()
    O.lib().ymo_su3_max_residuals(O.ext(n), O.dptr(aos), C.byref(ru), C.byref(rd))
    assert abs(u - ru.value) <= 1e-15 and abs(d - rd.value) <= 2e-15
    assert u <= 1e-14


def test_fixtures_closed_forms(ym):
    """su3-lattice-v1.json: identity action 0; single cycle link action 12 at beta 2."""
    n = (3, 3, 3, 3)
    lat = ym.Su3Lattice4(n, 64)
    assert lat.wilson_action(2.0) == 0.0
    assert lat.average_plaquette() == 1.0
    aos = O.su3_identity(n)
    O.lib().ymo_su3_traceless_cycle_link(O.dptr(aos[0:18]))
    lat.upload(aos)
    assert abs(lat.wilson_action(2.0) - 12.0) <= 1e-12
    re, im = lat.wilson_loops(1, 1, 0)
    assert abs(re[0, 0] - lat.average_plaquette()) <= 1e-15
    su2 = ym.Lattice4(n, 64)
    a2 = O.su2_identity(n)
    a2[0:4] = [0.0, 1.0, 0.0, 0.0]
    su2.upload(a2)
    assert abs(su2.wilson_action(2.0) - 12.0) <= 1e-12


def test_autotune_keeps_results_identical(ym):
    """The tuned tile visiting order permutes independent tiles: the chain is
    bit-identical to an untuned lattice with the same key."""
    n = (8, 8, 32, 16)  # N2 * N3 / 2 = 256: whole tiles per (x0, x1), strips 0, 2, 4, 8
    a = ym.Su3Lattice4(n, 32)
    b = ym.Su3Lattice4(n, 32)
    key = ym.derive_su3_chain_seed(7, 3)
    a.hot_start(key)
    b.hot_start(key)
    best, cand = a.autotune(6.0, 0.35, 8, reps=2)
    assert best in cand
    assert sorted(cand) == [("compact", 1), ("compact", 2), ("compact", 4), ("row", 0), ("row", 2), ("row", 4),
                            ("row", 8)]
    pa = a.metropolis_sweeps(6.0, 0.35, key, 0, 3, 8)
    pb = b.metropolis_sweeps(6.0, 0.35, key, 0, 3, 8)
    assert pa == pb
    assert np.array_equal(a.download(), b.download())
    # the timed variants run the same sweeps
    c = a.clone()
    pa = a.metropolis_sweeps(6.0, 0.35, key, 3, 2, 8)
    p1, a1, tot, ksum, nl = b.metropolis_sweeps_profiled(6.0, 0.35, key, 3, 2, 8)
    p2, a2, ms = c.metropolis_sweeps_timed(6.0, 0.35, key, 3, 2, 8)
    assert pa == (p1, a1) == (p2, a2)
    assert nl == 16 and 0 < ksum <= tot and ms > 0
    assert np.array_equal(a.download(), b.download()) and np.array_equal(a.download(), c.download())


def test_autotune_times_per_link_kernel_on_large_lattices(ym):
    """From 256 tiles up the sub-pass is the wave-specialised kernel and the
    tuner also times the one-link-per-thread kernel with the best mapping;
    whichever it keeps, the chain is bit-identical to an untuned lattice."""
    n = (16, 16, 16, 32)  # 65536 links per (mu, parity): 256 tiles
    a = ym.Su3Lattice4(n, 32)
    b = ym.Su3Lattice4(n, 32)
    key = ym.derive_su3_chain_seed(11, 2)
    a.hot_start(key)
    b.hot_start(key)
    best, cand = a.autotune(6.0, 0.35, 8, reps=2)
    assert best in cand
    per_link = [c for c in cand if c[0].endswith("/per-link")]
    assert len(per_link) == 1 and all(t > 0 for t in cand.values())
    pa = a.metropolis_sweeps(6.0, 0.35, key, 0, 2, 8)
    pb = b.metropolis_sweeps(6.0, 0.35, key, 0, 2, 8)
    assert pa == pb
    assert np.array_equal(a.download(), b.download())
