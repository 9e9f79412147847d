"""CPU checks of the device-order specification (oracle/dev_exact.c).

The device specification has no reference counterpart (Philox counter RNG,
checkerboard order, closed-form single-generator rotations, staple-based
Delta S, own FP64 transcendentals, hot start).  It is pinned here by
published known answers (Random123 Philox KATs), libm agreement of the
transcendental kernels, the identities that make the staple formula equal
to the reference's six-plaquette local action, Haar moments of the hot start,
and statistical agreement of the replayed chain with the reference's frozen
beta = 6 plaquette.
"""
import ctypes as C
import math
import os

import numpy as np
import pytest
from scipy.linalg import expm

import oracle_lib as O

L = O.lib()


def philox(ctr, key):
    o = (C.c_uint32 * 4)()
    L.ymd_philox4x32_10((C.c_uint32 * 4)(*ctr), (C.c_uint32 * 2)(*key), o)
    return list(o)


def test_philox4x32_10_random123_known_answers():
    assert philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def ulps(a, b):
    return abs(a - b) / max(math.ulp(abs(b)), 5e-324)


def test_fp64_kernels_against_libm():
    worst = {}
    for d in np.concatenate([np.linspace(0, 50, 5001), np.logspace(-12, 2.8, 2000)]):
        worst["exp"] = max(worst.get("exp", 0), ulps(L.ymd_exp_neg(d), math.exp(-d)))
    s, c = C.c_double(), C.c_double()
    for x in np.linspace(-0.785, 0.785, 4001):
        L.ymd_sincos_small(x, C.byref(s), C.byref(c))
        worst["sin"] = max(worst.get("sin", 0), abs(s.value - math.sin(x)) / math.ulp(1.0))
        worst["cos"] = max(worst.get("cos", 0), abs(1.0 + c.value - math.cos(x)) / math.ulp(1.0))
    for x in np.linspace(-0.2887, 0.2887, 4001):  # the short series' range (proposal angles <= 0.5)
        if x == 0.0:
            continue
        # relative to sin x and to cos x - 1 = -2 sin^2(x/2) (no cancellation), in ulps of each
        L.ymd_sincos_short(x, C.byref(s), C.byref(c))
        worst["sin_short"] = max(worst.get("sin_short", 0), ulps(s.value, math.sin(x)))
        worst["cm1_short"] = max(worst.get("cm1_short", 0), ulps(c.value, -2.0 * math.sin(x / 2) ** 2))
        L.ymd_sincos_small(x, C.byref(s), C.byref(c))
        worst["cm1"] = max(worst.get("cm1", 0), ulps(c.value, -2.0 * math.sin(x / 2) ** 2))
    for t in np.linspace(0, 1, 4001, endpoint=False):
        L.ymd_sincos_turns(t, C.byref(s), C.byref(c))
        worst["turns"] = max(worst.get("turns", 0), abs(s.value - math.sin(2 * math.pi * t)) / math.ulp(1.0),
                             abs(c.value - math.cos(2 * math.pi * t)) / math.ulp(1.0))
    for w in np.concatenate([np.logspace(-300, 0, 3001), np.linspace(0.01, 4, 3001)]):
        if w != 1.0:
            worst["log"] = max(worst.get("log", 0), ulps(L.ymd_log(w), math.log(w)))
    for y in np.concatenate([-np.logspace(-300, -1e-12, 2001), np.linspace(-0.999, 0.9, 2001)]):
        if y != 0:
            worst["log1p"] = max(worst.get("log1p", 0), ulps(L.ymd_log1p(y), math.log1p(y)))
    for x in -np.logspace(-300, 2.5, 3001):
        worst["expm1"] = max(worst.get("expm1", 0), ulps(L.ymd_expm1(x), math.expm1(x)))
    assert all(v <= 4.0 for v in worst.values()), worst


def dev_hot(n, prec, key, nc=18):
    vh = L.ymd_slots_per_parity(O.ext(n))
    buf = np.zeros(8 * nc * vh, dtype=np.float32 if prec == 32 else np.float64)
    f = L.ymd_su3_hot_start if nc == 18 else L.ymd_su2_hot_start
    assert f(O.ext(n), buf.ctypes.data_as(C.c_void_p), prec, key) == 0
    return buf


def dev_aos(n, buf, prec, nc=18):
    aos = np.zeros(int(np.prod(n)) * 4 * nc)
    f = L.ymd_su3_from_device if nc == 18 else L.ymd_su2_from_device
    assert f(O.ext(n), buf.ctypes.data_as(C.c_void_p), prec, O.dptr(aos)) == 0
    return aos


def test_su3_hot_start_haar_moments():
    """Haar SU(3): E Tr U = 0, E|Tr U|^2 = 1, E|Tr U|^4 = 2, E (Tr U)^3 = 1
    (the last is zero for U(3): it pins the determinant-one completion)."""
    n = (8, 8, 8, 8)
    aos = dev_aos(n, dev_hot(n, 64, 0xABCDEF), 64).reshape(-1, 9, 2)
    m = (aos[..., 0] + 1j * aos[..., 1]).reshape(-1, 3, 3)
    assert np.max(np.abs(np.einsum("kij,kil->kjl", m.conj(), m) - np.eye(3))) <= 1e-14
    assert np.max(np.abs(np.linalg.det(m) - 1.0)) <= 1e-14
    t = np.trace(m, axis1=1, axis2=2)
    N = len(t)
    assert abs(t.mean()) <= 5 / math.sqrt(N)
    assert abs(np.mean(np.abs(t) ** 2) - 1.0) <= 5 * math.sqrt(1.0 / N)
    assert abs(np.mean(np.abs(t) ** 4) - 2.0) <= 5 * math.sqrt(24.0 / N)
    assert abs(np.mean(t ** 3) - 1.0) <= 5 * math.sqrt(6.0 / N)


def test_su2_hot_start_haar_moments():
    n = (8, 8, 8, 8)
    q = dev_aos(n, dev_hot(n, 64, 7, nc=4), 64, nc=4).reshape(-1, 4)
    assert np.max(np.abs(np.sum(q * q, axis=1) - 1)) <= 1e-15
    N = len(q)
    assert abs(q[:, 0].mean()) <= 5 * 0.5 / math.sqrt(N)
    assert abs(np.mean(q[:, 0] ** 2) - 0.25) <= 5 * math.sqrt(1.0 / 8 / N)


def test_staple_formula_reproduces_reference_local_action():
    """Delta S = -(beta/3) Re Tr((R U - U) A) with the device staple A equals
    the reference's six-plaquette local_action_difference for every generator
    (su3_ensemble.rs:261-334)."""
    n = (4, 4, 4, 4)
    buf = dev_hot(n, 64, 42)
    aos = dev_aos(n, buf, 64)
    beta = 6.0
    for x, mu in [((0, 0, 0, 0), 0), ((1, 2, 3, 0), 1), ((3, 0, 2, 1), 3)]:
        A = np.zeros(18)
        assert L.ymd_su3_staple(O.ext(n), buf.ctypes.data_as(C.c_void_p), 64, (C.c_int64 * 4)(*x), mu, O.dptr(A)) == 0
        A = O.mat_from18(A)
        off = O.link_offset(n, x, mu, 18)
        U = O.mat_from18(aos[off:off + 18])
        ref = C.c_double()
        L.ymo_su3_local_link_action(O.ext(n), O.dptr(aos), (C.c_int64 * 4)(*x), mu, beta, C.byref(ref))
        assert abs(beta * (6 - np.trace(U @ A).real / 3) - ref.value) <= 1e-12 * abs(ref.value)
        for a in range(8):
            coef = np.zeros(8)
            coef[a] = 0.3
            R = O.mat_from18(O.su3_from_algebra(coef))
            cand = R @ U
            L.ymo_su3_local_action_difference(O.ext(n), O.dptr(aos), (C.c_int64 * 4)(*x), mu,
                                              O.dptr(O.mat_to18(cand)), beta, C.byref(ref))
            assert abs(-(beta / 3) * np.trace((cand - U) @ A).real - ref.value) <= 1e-12


def _staple_fp64_full(aos, n, x, mu):
    """The staple from whole stored links, in FP64 (numpy)."""
    def link(y, d):
        o = O.link_offset(n, tuple(v % n[k] for k, v in enumerate(y)), d, 18)
        return O.mat_from18(aos[o:o + 18])

    def sh(y, d, s):
        z = list(y)
        z[d] += s
        return z

    A = np.zeros((3, 3), dtype=complex)
    for nu in range(4):
        if nu == mu:
            continue
        xpm = sh(x, mu, 1)
        A += link(xpm, nu) @ link(sh(x, nu, 1), mu).conj().T @ link(x, nu).conj().T
        A += link(sh(xpm, nu, -1), nu).conj().T @ link(sh(x, nu, -1), mu).conj().T @ link(sh(x, nu, -1), nu)
    return A


def test_fp32_staple_from_two_rows_matches_full_products():
    """The FP32 staple reads rows 0 and 1 of each neighbour and rebuilds third
    rows as conj(row 0 x row 1) (forward X (Z Y)^+, backward (Y X)^+ Z).  On
    FP32 hot-start links it equals the staple of the whole stored links,
    formed in FP64, to FP32 rounding; and it does not change when every
    stored third row is overwritten, so no neighbour's third row is read.  A
    wrong conjugation, sign or operand order in the rebuilt rows moves the
    result by O(1)."""
    n = (4, 4, 4, 6)
    buf = dev_hot(n, 32, 9)
    aos = dev_aos(n, buf, 32)
    sites = [((0, 0, 0, 0), 0), ((1, 2, 3, 5), 3), ((3, 3, 0, 1), 2), ((2, 1, 3, 4), 1)]
    got = {}
    for x, mu in sites:
        A = np.zeros(18)
        assert L.ymd_su3_staple(O.ext(n), buf.ctypes.data_as(C.c_void_p), 32, (C.c_int64 * 4)(*x), mu, O.dptr(A)) == 0
        got[(x, mu)] = A.copy()
        ref = _staple_fp64_full(aos, n, x, mu)
        assert np.abs(O.mat_from18(A) - ref).max() <= 2e-6 * np.abs(ref).max(), (x, mu)
    # overwrite the third row of every stored link and repeat
    bad = aos.reshape(-1, 18).copy()
    bad[:, 12:18] = 7.0
    buf2 = np.zeros_like(buf)
    assert L.ymd_su3_to_device(O.ext(n), O.dptr(bad.reshape(-1)), buf2.ctypes.data_as(C.c_void_p), 32) == 0
    for x, mu in sites:
        A = np.zeros(18)
        assert L.ymd_su3_staple(O.ext(n), buf2.ctypes.data_as(C.c_void_p), 32, (C.c_int64 * 4)(*x), mu, O.dptr(A)) == 0
        assert np.array_equal(A, got[(x, mu)]), (x, mu)
        # the FP64 full-product staple of the overwritten links is far off
        assert np.abs(O.mat_from18(A) - _staple_fp64_full(bad.reshape(-1), n, x, mu)).max() > 0.1


# Gell-Mann matrices lambda_1 .. lambda_8; generator h mod 8 of a hit is
# T = lambda_{h mod 8 + 1} / 2 (the reference's generator(), su3.rs:325-352).
_GELL_MANN = [np.array(m, dtype=complex) for m in (
    [[0, 1, 0], [1, 0, 0], [0, 0, 0]],
    [[0, -1j, 0], [1j, 0, 0], [0, 0, 0]],
    [[1, 0, 0], [0, -1, 0], [0, 0, 0]],
    [[0, 0, 1], [0, 0, 0], [1, 0, 0]],
    [[0, 0, -1j], [0, 0, 0], [1j, 0, 0]],
    [[0, 0, 0], [0, 0, 1], [0, 1, 0]],
    [[0, 0, 0], [0, 0, -1j], [0, 1j, 0]],
    np.diag([1, 1, -2]) / math.sqrt(3.0),
)]


def _numpy_hit_sweep(n, aos, beta, angle, hits, key, sweep):
    """One checkerboard sweep written from the definitions: for each (mu,
    parity) sub-pass and link, hit h draws Philox at (site, mu | h << 5,
    sweep) under the key, takes theta on the reference's grid from words 0-1
    and the uniform from words 2-3, proposes U' = expm(i theta T_{h mod 8}) U,
    and accepts iff dS = -(beta/3) Re Tr((U' - U) A) <= 0 or u < exp(-dS),
    A the FP64 staple of the current field.  Returns the links and the
    accept count."""
    aos = aos.copy()
    acc = 0
    for mu in range(4):
        for par in range(2):
            for x in np.ndindex(*n):
                if sum(x) % 2 != par:
                    continue
                A = _staple_fp64_full(aos, n, x, mu)
                off = O.link_offset(n, x, mu, 18)
                u = O.mat_from18(aos[off:off + 18])
                site = ((x[0] * n[1] + x[1]) * n[2] + x[2]) * n[3] + x[3]
                for h in range(hits):
                    o = philox((site, mu | (h << 5), sweep & 0xFFFFFFFF, sweep >> 32),
                               (key & 0xFFFFFFFF, key >> 32))
                    raw = (o[1] << 32) | o[0]
                    theta = angle * float(2 * (raw & ((1 << 52) - 1)) + 1) * 2.0 ** -53
                    theta = theta if raw >> 63 else -theta
                    uacc = float(((o[3] << 32) | o[2]) >> 11) * 2.0 ** -53
                    cand = expm(1j * theta * _GELL_MANN[h % 8] / 2) @ u
                    dS = -(beta / 3) * np.trace((cand - u) @ A).real
                    if dS <= 0 or uacc < math.exp(-dS):
                        u = cand
                        acc += 1
                aos[off:off + 18] = O.mat_to18(u)
    return aos, acc


def test_replay_hits_follow_generator_exponentials():
    """The replay's closed-form hits (rotations of rows by cos/sin of theta/2,
    phases for lambda_3 and lambda_8, Delta S from the 2 x 2 block of U A)
    against the definitions in numpy: one FP64 sweep of 10 hits per link
    (every generator, 0 and 1 twice) on a hot 4x4x2x4 lattice at beta = 6,
    angle 0.8, without reunitarisation.  The accept count must agree exactly
    and every link to 1e-12; the same comparison with theta of the opposite
    sign or a generator off by one fails by far."""
    n = (4, 4, 2, 4)
    beta, angle, hits, key, sweep = 6.0, 0.8, 10, 0x1234ABCD5678EF01, 3
    buf = dev_hot(n, 64, 11)
    start = dev_aos(n, buf, 64)
    p = O.DevSu3Params(beta=beta, angle=angle, hits=hits, reunitarize=0, key=key, sweep=sweep)
    pr, ac = C.c_uint64(), C.c_uint64()
    assert L.ymd_su3_metropolis_sweep(O.ext(n), buf.ctypes.data_as(C.c_void_p), 64, C.byref(p),
                                      C.byref(pr), C.byref(ac)) == 0
    got = dev_aos(n, buf, 64)
    want, acc = _numpy_hit_sweep(n, start, beta, angle, hits, key, sweep)
    assert pr.value == hits * 4 * int(np.prod(n))
    assert 0 < ac.value < pr.value
    assert ac.value == acc, (ac.value, acc)
    assert np.max(np.abs(got - want)) <= 1e-12, np.max(np.abs(got - want))


def test_beta0_replay_accepts_everything_and_stays_unitary():
    n = (4, 4, 2, 4)
    buf = dev_hot(n, 64, 5)
    p = O.DevSu3Params(beta=0.0, angle=0.35, hits=8, reunitarize=1, key=9, sweep=0)
    pr, ac = C.c_uint64(), C.c_uint64()
    assert L.ymd_su3_metropolis_sweep(O.ext(n), buf.ctypes.data_as(C.c_void_p), 64, C.byref(p), C.byref(pr), C.byref(ac)) == 0
    assert pr.value == ac.value == 8 * 4 * 128
    u, d = C.c_double(), C.c_double()
    L.ymo_su3_max_residuals(O.ext(n), O.dptr(dev_aos(n, buf, 64)), C.byref(u), C.byref(d))
    assert u.value <= 1e-14 and d.value <= 1e-14


def test_checkerboard_rejects_odd_extents():
    n = (4, 4, 4, 3)
    buf = dev_hot(n, 64, 1)
    p = O.DevSu3Params(beta=6.0, angle=0.35, hits=8, reunitarize=1, key=9, sweep=0)
    pr, ac = C.c_uint64(), C.c_uint64()
    assert L.ymd_su3_metropolis_sweep(O.ext(n), buf.ctypes.data_as(C.c_void_p), 64, C.byref(p), C.byref(pr), C.byref(ac)) == 99


@pytest.mark.slow
def test_replayed_device_chain_matches_reference_plaquette():
    """The checkerboard/Philox sampler targets the same distribution as the
    reference: the V2 schedule (4^4, beta = 6, cold start, 192 + 128 x 2
    sweeps, 8 proposals per link per sweep, angle 0.35), four chains; the
    pooled plaquette must lie within 2 sigma of the frozen V2 pooled value
    0.5992920042685975 +- 0.0008268453534905086 (su3-plaquette-baseline-v2.json)
    and the acceptance near the frozen 0.8675."""
    n = (4, 4, 4, 4)
    series, acc = [], []
    for chain in range(4):
        key = L.ymo_derive_su3_chain_seed(0x5EED, chain)
        vh = L.ymd_slots_per_parity(O.ext(n))
        buf = np.zeros(8 * 18 * vh)
        assert L.ymd_su3_to_device(O.ext(n), O.dptr(O.su3_identity(n)), buf.ctypes.data_as(C.c_void_p), 64) == 0
        p = O.DevSu3Params(beta=6.0, angle=0.35, hits=8, reunitarize=1, key=key, sweep=0)
        pr, ac = C.c_uint64(), C.c_uint64()
        meas, tp, ta = [], 0, 0
        for s in range(192 + 256):
            p.sweep = s
            L.ymd_su3_metropolis_sweep(O.ext(n), buf.ctypes.data_as(C.c_void_p), 64, C.byref(p), C.byref(pr), C.byref(ac))
            tp += pr.value
            ta += ac.value
            if s >= 192 and (s - 192) % 2 == 1:
                v = C.c_double()
                L.ymo_su3_average_plaquette(O.ext(n), O.dptr(dev_aos(n, buf, 64)), C.byref(v))
                meas.append(v.value)
        series.append(meas)
        acc.append(ta / tp)
    s = np.ascontiguousarray(np.concatenate(series))
    pool = O.PoolReport()
    assert L.ymo_pool_chains(O.dptr(s), 4, 128, 24, 0, C.byref(pool)) == 0
    z = abs(pool.pooled_mean - 0.5992920042685975) / math.hypot(pool.pooled_se, 0.0008268453534905086)
    assert z < 2.0, (pool.pooled_mean, pool.pooled_se, z)
    assert abs(np.mean(acc) - 0.8675) < 0.005, acc


def test_replay_plaquette_matches_reference_exact():
    """ymd_su3_average_plaquette (the CPU chains' measurement in the level-(c)
    comparisons) against the reference-exact average_plaquette: identity
    gives 1, a hot-start FP64 configuration agrees to 1e-13."""
    import ctypes as C
    L = O.lib()
    n = (4, 6, 4, 6)
    vh = L.ymd_slots_per_parity(O.ext(n))
    buf = np.zeros(8 * 18 * vh)
    aos = O.su3_identity(n)
    assert L.ymd_su3_to_device(O.ext(n), O.dptr(aos), buf.ctypes.data_as(C.c_void_p), 64) == 0
    v = C.c_double()
    assert L.ymd_su3_average_plaquette(O.ext(n), buf.ctypes.data_as(C.c_void_p), 64, C.byref(v)) == 0
    assert v.value == 1.0
    assert L.ymd_su3_hot_start(O.ext(n), buf.ctypes.data_as(C.c_void_p), 64, 4242) == 0
    p = O.DevSu3Params(beta=6.0, angle=0.35, hits=8, reunitarize=1, key=9, sweep=0)
    a, b = C.c_uint64(), C.c_uint64()
    assert L.ymd_su3_metropolis_sweep(O.ext(n), buf.ctypes.data_as(C.c_void_p), 64, C.byref(p), C.byref(a),
                                      C.byref(b)) == 0
    assert L.ymd_su3_from_device(O.ext(n), buf.ctypes.data_as(C.c_void_p), 64, O.dptr(aos)) == 0
    ref = C.c_double()
    assert L.ymo_su3_average_plaquette(O.ext(n), O.dptr(aos), C.byref(ref)) == 0
    assert L.ymd_su3_average_plaquette(O.ext(n), buf.ctypes.data_as(C.c_void_p), 64, C.byref(v)) == 0
    assert abs(v.value - ref.value) <= 1e-13


def test_replay_is_independent_of_the_thread_count():
    """The replay runs the links of a sub-pass on OpenMP threads: 1 and 3
    threads give the same links, counts and plaquette, bit for bit."""
    import subprocess
    import sys
    code = r'''
import ctypes as C, hashlib, sys
import numpy as np
sys.path.insert(0, "TESTS")
import oracle_lib as O
L = O.lib()
n = (4, 4, 6, 8)
vh = L.ymd_slots_per_parity(O.ext(n))
b3 = np.zeros(8 * 18 * vh, dtype=np.float32)
assert L.ymd_su3_hot_start(O.ext(n), b3.ctypes.data_as(C.c_void_p), 32, 5) == 0
p = O.DevSu3Params(beta=6.0, angle=0.35, hits=10, reunitarize=1, key=3, sweep=1)
a, c = C.c_uint64(), C.c_uint64()
assert L.ymd_su3_metropolis_sweep(O.ext(n), b3.ctypes.data_as(C.c_void_p), 32, C.byref(p), C.byref(a), C.byref(c)) == 0
b2 = np.zeros(8 * 4 * vh, dtype=np.float32)
assert L.ymd_su2_hot_start(O.ext(n), b2.ctypes.data_as(C.c_void_p), 32, 5) == 0
u, w = C.c_uint64(), C.c_uint64()
assert L.ymd_su2_heatbath_sweep(O.ext(n), b2.ctypes.data_as(C.c_void_p), 32, 2.3, 3, 1, 10000, C.byref(u), C.byref(w)) == 0
assert L.ymd_su2_overrelax_sweep(O.ext(n), b2.ctypes.data_as(C.c_void_p), 32) == 0
v = C.c_double()
assert L.ymd_su3_average_plaquette(O.ext(n), b3.ctypes.data_as(C.c_void_p), 32, C.byref(v)) == 0
print(hashlib.sha1(b3.tobytes() + b2.tobytes()).hexdigest(), a.value, c.value, u.value, w.value, v.value.hex())
'''.replace("TESTS", os.path.dirname(os.path.abspath(__file__)))
    outs = []
    for t in ("1", "3"):
        env = dict(os.environ, OMP_NUM_THREADS=t)
        r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr
        outs.append(r.stdout.strip())
    assert outs[0] == outs[1], outs
