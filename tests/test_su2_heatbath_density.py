"""Pin of the device-order SU(2) heat-bath (oracle/dev_exact.c,
ymd_su2_heatbath_link, the per-link step of the replay that the GPU sweeps
reproduce bit for bit) against the distribution it must sample.

For a staple V with |V| = alpha / beta the local conditional density of the
new link U is exp(beta (U V)_0) dHaar(U) (su2_ensemble.rs:210-218), (U V)_0
= (1/2) Re Tr(U V).  Writing U V^ = (x, sqrt(1 - x^2) n) with V^ = V / |V|:
x has density sqrt(1 - x^2) e^{alpha x} on [-1, 1] and n is uniform on S^2,
independent of x.  The tests draw 100,000 links per alpha from the replay and
compare x with that density (Kolmogorov-Smirnov against the numerically
integrated CDF), every component of n with the uniform law on [-1, 1]
(Archimedes), and the mean number of attempts with the geometric law of the
sqrt(1 - r0^2) rejection step.  Small alpha exercises the log1p branch of
the inverse CDF, large alpha the plain-log branch.  The same samples read
with the staple conjugated the other way, or against the density of a
mis-normalised alpha, are rejected: the tests would catch those mistakes.
"""
import ctypes as C

import numpy as np
import pytest
from scipy import stats

import oracle_lib as O

BETA = 2.3
KEY = 0x2B00201211160101
N = 100_000


def _qmul(a, b):
    """Quaternion product in the reference convention (lib.rs:212-224), rows of a with b."""
    a0, a1, a2, a3 = a[..., 0], a[..., 1], a[..., 2], a[..., 3]
    b0, b1, b2, b3 = b[..., 0], b[..., 1], b[..., 2], b[..., 3]
    return np.stack([a0 * b0 - a1 * b1 - a2 * b2 - a3 * b3,
                     a0 * b1 + a1 * b0 + a2 * b3 - a3 * b2,
                     a0 * b2 - a1 * b3 + a2 * b0 + a3 * b1,
                     a0 * b3 + a1 * b2 - a2 * b1 + a3 * b0], axis=-1)


def _density_cdf(alpha):
    t = np.linspace(-1.0, 1.0, 200_001)
    w = np.sqrt(np.clip(1.0 - t * t, 0.0, None)) * np.exp(alpha * (t - 1.0))
    cum = np.concatenate([[0.0], np.cumsum(0.5 * (w[1:] + w[:-1]) * np.diff(t))])
    return lambda v: np.interp(v, t, cum / cum[-1])


def _acceptance(alpha):
    """Probability that one attempt is accepted: E[sqrt(1 - r^2)] under the
    truncated exponential proposal e^{alpha r} on [-1, 1]."""
    t = np.linspace(-1.0, 1.0, 200_001)
    e = np.exp(alpha * (t - 1.0))
    w = np.sqrt(np.clip(1.0 - t * t, 0.0, None)) * e
    return np.trapezoid(w, t) / np.trapezoid(e, t)


def _draw(alpha, seed=0):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=4)
    vhat = v / np.linalg.norm(v)
    a = np.ascontiguousarray(vhat * (alpha / BETA))
    out = np.zeros((N, 4))
    att = np.zeros(N, dtype=np.int64)
    st = O.lib().ymd_su2_heatbath_samples(O.dptr(a), BETA, KEY, 1, 7, 10_000, N, O.dptr(out),
                                          att.ctypes.data_as(C.POINTER(C.c_int64)))
    assert st == 0
    return vhat, out, att


@pytest.mark.parametrize("alpha", [0.05, 0.4, 2.0, 8.0, 30.0])
def test_heatbath_link_samples_the_local_conditional_density(alpha):
    vhat, U, att = _draw(alpha)
    assert np.abs(np.einsum("ij,ij->i", U, U) - 1.0).max() < 1e-14
    W = _qmul(U, vhat)
    x = W[:, 0]
    ks = stats.kstest(x, _density_cdf(alpha))
    assert ks.pvalue > 1e-3, (alpha, ks)
    n = W[:, 1:] / np.sqrt(np.clip(1.0 - x * x, 1e-300, None))[:, None]
    assert np.abs(np.linalg.norm(n, axis=1) - 1.0).max() < 1e-6
    for c in range(3):
        kc = stats.kstest(n[:, c], stats.uniform(loc=-1.0, scale=2.0).cdf)
        assert kc.pvalue > 1e-3, (alpha, c, kc)
        # independent of x: no correlation beyond sampling noise
        assert abs(np.corrcoef(x, n[:, c])[0, 1]) < 5.0 / np.sqrt(N)
    # attempts: geometric with success probability E[sqrt(1 - r0^2)]
    p = _acceptance(alpha)
    mean, sd = 1.0 / p, np.sqrt((1.0 - p) / p**2 / N)
    assert abs(att.mean() - mean) < 5.0 * sd, (alpha, att.mean(), mean, sd)


@pytest.mark.parametrize("alpha", [2.0, 8.0, 30.0])
def test_density_test_rejects_planted_mistakes(alpha):
    vhat, U, _ = _draw(alpha)
    cdf = _density_cdf(alpha)
    # U = R V^ instead of R V^dagger: read the samples against the conjugate staple
    vconj = vhat * np.array([1.0, -1.0, -1.0, -1.0])
    wrong = _qmul(U, vconj)[:, 0]
    assert stats.kstest(wrong, cdf).pvalue < 1e-6
    # alpha mis-normalised by 10%
    x = _qmul(U, vhat)[:, 0]
    assert stats.kstest(x, _density_cdf(0.9 * alpha)).pvalue < 1e-6


def test_heatbath_link_zero_staple_and_attempt_limit():
    """|V| = 0: the new link is the Haar draw R itself (heatbath_link's
    norm == 0 branch); an attempt limit of 1 reports exhaustion without
    touching the output."""
    a = np.zeros(4)
    out = np.zeros(4)
    att = C.c_int64()
    assert O.lib().ymd_su2_heatbath_link(O.dptr(a), BETA, KEY, 3, 0, 0, 10_000, O.dptr(out), C.byref(att)) == 1
    assert abs(np.dot(out, out) - 1.0) < 1e-15 and att.value >= 1
    # large alpha accepts rarely on the first attempt: find a draw that fails
    a = np.array([30.0 / BETA, 0.0, 0.0, 0.0])
    sentinel = np.full(4, 7.0)
    fails = 0
    for site in range(64):
        out = sentinel.copy()
        r = O.lib().ymd_su2_heatbath_link(O.dptr(a), BETA, KEY, site, 0, 0, 1, O.dptr(out), C.byref(att))
        assert att.value == 1
        if r == 0:
            fails += 1
            assert np.array_equal(out, sentinel)
    assert fails > 0
