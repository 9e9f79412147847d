"""CPU oracle (test infrastructure only) for the host correlator analysis.

Step-by-step restatement of the reference's FP64 analysis of scalar
glueball correlators, used by tests/ as the checker of libym_amd.so's
ym_cosh_effective_mass, ym_correlated_scalar_fit,
ym_variational_connected_correlator and ym_principal_generalized_projection.
Nothing in the product imports this file.

Follows (reference/rust/src/):
  glueball_effective_mass.rs:96-232    validation, cosh mass, jackknife
  glueball_effective_mass.rs:240-268   analytic single-state fixture
  glueball_scalar_fit.rs:156-650       windows, plateau, covariance, inverse,
                                       profile grid fit, jackknife refits
  glueball_scalar_fit.rs:652-699       canonical Hadamard fixture
  glueball_scalar_variational.rs:132-925  symmetrisation, stable 2x2
                                       generalized eigenvalue, whitening,
                                       projection, cross-correlator
  glueball_scalar_variational.rs:927-951  canonical two-state matrices

Rounding follows the reference: sums in index order, no contraction except
where the reference calls mul_add (done exactly here through rationals);
acosh is the reference's f64::acosh, ln(x + sqrt(x - 1) sqrt(x + 1)); hypot is
the C library's (numpy.hypot).  Errors are raised as OracleError carrying the
status number the product library uses for the same reference variant.

Pinned by the reference's golden vectors
(glueball-effective-mass-phase3d-v1.json, glueball-scalar-fit-phase3e-v1.json,
glueball-scalar-variational-phase3j-v1.json) and by closed forms: a single
periodic state gives its mass exactly at every lag, and for two states in a
2 x 2 basis the principal generalized eigenvalue is w0(t1)/w0(t0) with
projection w0(t)/w0(t0) (tests/test_host_fit.py).
"""
import math
from fractions import Fraction

import numpy as np


class OracleError(Exception):
    def __init__(self, status, name):
        super().__init__(f"{name} ({status})")
        self.status = status
        self.name = name


def _err(name):
    return OracleError(STATUS[name], name)


STATUS = {
    "TimeExtentTooSmall": 100, "TimeExtentTooLarge": 101, "TooFewJackknifeReplicates": 102,
    "TooManyJackknifeReplicates": 103, "EmptyAnalysisWindow": 104, "InvalidAnalysisLag": 105,
    "ShapeMismatch": 106, "NonFiniteCorrelator": 107, "NonPositiveCentralCorrelator": 108,
    "AcoshDomainViolation": 109, "NonFiniteStatisticEM": 110, "InvalidFixtureParameter": 111,
    "EmptyCandidateWindows": 120, "TooManyCandidateWindows": 121, "InvalidCandidateWindow": 122,
    "DuplicateCandidateWindow": 123, "DegeneratePlateauUncertainty": 124, "NoAcceptedPlateauWindow": 125,
    "InsufficientReplicatesForCovariance": 126, "CovarianceNotPositiveDefinite": 127, "SingularCovariance": 128,
    "NonFiniteStatisticFit": 129, "NoFiniteGridFit": 130,
    "EmptyConfigurationSet": 70, "TooManyConfigurations": 71, "VarTimeExtentTooSmall": 140,
    "VarTimeExtentTooLarge": 72, "NonFiniteObservable": 54, "NonFiniteAccumulator": 74,
    "CorrelatorShapeMismatch": 75, "MatrixNotSymmetric": 142, "InvalidReferenceTimes": 143,
    "ReferenceMatrixNotPositiveDefinite": 144, "DegeneratePrincipalEigenvector": 145,
    "NonPositivePrincipalEigenvalue": 146, "ProjectionNormalizationFailure": 147,
}

fin = math.isfinite


def fma(a, b, c):
    """Correctly rounded a*b + c (f64::mul_add)."""
    if not (fin(a) and fin(b) and fin(c)):
        return a * b + c
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def acosh_ref(x):
    return math.log(x + math.sqrt(x - 1.0) * math.sqrt(x + 1.0))


def hypot(a, b):
    return float(np.hypot(a, b))


# --------------------------------------------------- glueball_effective_mass.rs
def _validate_em(full, jk, lags):  # :96-139
    T = len(full)
    if T < 4:
        raise _err("TimeExtentTooSmall")
    if T > 64:
        raise _err("TimeExtentTooLarge")
    if len(jk) < 3:
        raise _err("TooFewJackknifeReplicates")
    if len(jk) > 64:
        raise _err("TooManyJackknifeReplicates")
    if len(lags) == 0:
        raise _err("EmptyAnalysisWindow")
    if any(not fin(v) for v in full):
        raise _err("NonFiniteCorrelator")
    if any(len(r) != T for r in jk):
        raise _err("ShapeMismatch")
    if any(not fin(v) for r in jk for v in r):
        raise _err("NonFiniteCorrelator")
    prev = 0
    for i, tau in enumerate(lags):
        if tau == 0 or tau >= T - 1 or 2 * tau >= T or (i != 0 and tau <= prev):
            raise _err("InvalidAnalysisLag")
        prev = tau


def _cosh_mass(c, tau):  # :141-159
    centre = c[tau]
    if centre <= 0.0:
        raise _err("NonPositiveCentralCorrelator")
    num = c[tau - 1] + c[tau + 1]
    den = 2.0 * centre
    arg = num / den
    if not (fin(num) and fin(den) and fin(arg)):
        raise _err("NonFiniteStatisticEM")
    if arg < 1.0:
        raise _err("AcoshDomainViolation")
    m = acosh_ref(arg)
    if not fin(m):
        raise _err("NonFiniteStatisticEM")
    return m


def cosh_effective_mass_with_jackknife(full, jk, lags):  # :161-232
    full = [float(v) for v in full]
    jk = [[float(v) for v in r] for r in jk]
    lags = [int(t) for t in lags]
    _validate_em(full, jk, lags)
    pts = []
    for tau in lags:
        f = _cosh_mass(full, tau)
        ms, s = [], 0.0
        for r in jk:
            m = _cosh_mass(r, tau)
            s += m
            if not fin(s):
                raise _err("NonFiniteStatisticEM")
            ms.append(m)
        n = float(len(ms))
        mean = s / n
        q = 0.0
        for m in ms:
            d = m - mean
            q += d * d
        var = ((n - 1.0) / n) * q
        se = math.sqrt(var)
        pts.append((tau, f, mean, se))
    return pts


def analytic_periodic_single_state(T, mass, amplitude):  # :240-268
    if not (4 <= T <= 64) or not fin(mass) or mass <= 0.0 or not fin(amplitude) or amplitude <= 0.0:
        raise _err("InvalidFixtureParameter")
    return [amplitude * (math.exp(-mass * float(tau)) + math.exp(-mass * float(T - tau))) for tau in range(T)]


# ------------------------------------------------------- glueball_scalar_fit.rs
def _validate_windows(T, cands):  # :156-182
    if not cands:
        raise _err("EmptyCandidateWindows")
    if len(cands) > 16:
        raise _err("TooManyCandidateWindows")
    for i, (a, b) in enumerate(cands):
        if b < a:
            raise _err("InvalidCandidateWindow")
        n = b - a + 1
        if a == 0 or n < 3 or n > 8 or 2 * b >= T:
            raise _err("InvalidCandidateWindow")
        if (a, b) in cands[:i]:
            raise _err("DuplicateCandidateWindow")


def _plateau(full, jk, a, b):  # :197-237
    pts = cosh_effective_mass_with_jackknife(full, jk, list(range(a, b + 1)))
    ws, w = 0.0, 0.0
    for _, f, _, se in pts:
        if not fin(se) or se <= 0.0:
            raise _err("DegeneratePlateauUncertainty")
        wi = 1.0 / (se * se)
        ws += wi * f
        w += wi
    mw = ws / w
    zmax = 0.0
    for _, f, _, se in pts:
        zmax = max(zmax, abs((f - mw) / se))
    return (a, b), mw, zmax


def _choose(full, jk, cands):  # :239-274
    _validate_windows(len(full), cands)
    best = None
    for a, b in cands:
        try:
            sel = _plateau(full, jk, a, b)
        except OracleError as e:
            if e.name == "DegeneratePlateauUncertainty":
                continue
            raise
        if sel[2] > 2.5:
            continue
        n = b - a + 1
        if best is None:
            best = sel
        else:
            nb = best[0][1] - best[0][0] + 1
            if n > nb or (n == nb and a > best[0][0]):
                best = sel
    if best is None:
        raise _err("NoAcceptedPlateauWindow")
    return best


def _covariance(jk, a, b):  # :276-322
    n = b - a + 1
    if len(jk) <= n:
        raise _err("InsufficientReplicatesForCovariance")
    means = [0.0] * n
    for r in jk:
        for c in range(n):
            means[c] += r[a + c]
    N = float(len(jk))
    means = [m / N for m in means]
    cov = [[0.0] * n for _ in range(n)]
    for r in jk:
        for i in range(n):
            left = r[a + i] - means[i]
            for j in range(n):
                cov[i][j] += left * (r[a + j] - means[j])
    f = (N - 1.0) / N
    return [[v * f for v in row] for row in cov]


def _max_abs(m):
    s = 0.0
    for row in m:
        for v in row:
            s = max(s, abs(v))
    return s


def _check_pd(m):  # :324-359
    n = len(m)
    scale = _max_abs(m)
    if n == 0 or not fin(scale) or scale <= 0.0:
        raise _err("CovarianceNotPositiveDefinite")
    low = [[0.0] * n for _ in range(n)]
    for i in range(n):
        for j in range(i + 1):
            s = m[i][j]
            for k in range(j):
                s -= low[i][k] * low[j][k]
            if i == j:
                if not fin(s) or s <= scale * 1.0e-12:
                    raise _err("CovarianceNotPositiveDefinite")
                low[i][j] = math.sqrt(s)
            else:
                low[i][j] = s / low[j][j]


def _invert(m):  # :361-421
    _check_pd(m)
    n = len(m)
    scale = _max_abs(m)
    aug = [list(m[i]) + [1.0 if k == i else 0.0 for k in range(n)] for i in range(n)]
    for col in range(n):
        piv, pa = col, abs(aug[col][col])
        for r in range(col + 1, n):
            if abs(aug[r][col]) > pa:
                pa, piv = abs(aug[r][col]), r
        if not fin(pa) or pa <= scale * 1.0e-12:
            raise _err("SingularCovariance")
        if piv != col:
            aug[piv], aug[col] = aug[col], aug[piv]
        p = aug[col][col]
        aug[col] = [v / p for v in aug[col]]
        for r in range(n):
            if r == col:
                continue
            f = aug[r][col]
            aug[r] = [aug[r][k] - f * aug[col][k] for k in range(2 * n)]
    return [row[n:] for row in aug]


def _qf(x, m, y):  # :423-438
    s = 0.0
    for i in range(len(m)):
        inner = 0.0
        for j in range(len(m)):
            inner += m[i][j] * y[j]
        s += x[i] * inner
    return s


def _grid_fit(T, a, b, vals, inv):  # :441-526
    best = None
    for g in range(0, 1951):
        mass = 0.05 + 0.001 * float(g)
        if mass > 2.0 + 1.0e-15:
            break
        basis = [math.exp(-mass * float(t)) + math.exp(-mass * float(T - t)) for t in range(a, b + 1)]
        den = _qf(basis, inv, basis)
        if den <= 0.0 or not fin(den):
            continue
        amp = _qf(basis, inv, vals) / den
        if amp <= 0.0 or not fin(amp):
            continue
        res = [v - amp * f for v, f in zip(vals, basis)]
        chi2 = _qf(res, inv, res)
        if chi2 < 0.0 or not fin(chi2):
            continue
        if best is None or chi2 < best[2] or (chi2 == best[2] and mass < best[0]):
            best = (mass, amp, chi2)
    if best is None:
        raise _err("NoFiniteGridFit")
    return best


def _jk_se(vals):  # :548-565
    n = float(len(vals))
    mean = sum(vals) / n
    q = 0.0
    for v in vals:
        d = v - mean
        q += d * d
    return mean, math.sqrt(((n - 1.0) / n) * q)


def _max_corr(cov):  # :567-593
    m = 0.0
    for i in range(len(cov)):
        si = math.sqrt(cov[i][i])
        for j in range(i):
            sj = math.sqrt(cov[j][j])
            m = max(m, abs((cov[i][j] / si) / sj))
    return m


def correlated_scalar_fit(full, jk, cands):  # :595-650
    full = [float(v) for v in full]
    jk = [[float(v) for v in r] for r in jk]
    cands = [(int(a), int(b)) for a, b in cands]
    (a, b), mw, zmax = _choose(full, jk, cands)
    cov = _covariance(jk, a, b)
    inv = _invert(cov)
    cmax = _max_corr(cov)
    T = len(full)
    mass, amp, chi2 = _grid_fit(T, a, b, full[a:b + 1], inv)
    ms, amps = [], []
    for r in jk:
        fm, fa, _ = _grid_fit(T, a, b, r[a:b + 1], inv)
        ms.append(fm)
        amps.append(fa)
    mm, mse = _jk_se(ms)
    am, ase = _jk_se(amps)
    dof = (b - a + 1) - 2
    return dict(window=(a, b), plateau_weighted_effective_mass=mw, plateau_maximum_z=zmax, fitted_mass=mass,
                fitted_amplitude=amp, chi_square=chi2, degrees_of_freedom=dof, chi_square_per_dof=chi2 / dof,
                mass_jackknife_mean=mm, mass_standard_error=mse, amplitude_jackknife_mean=am,
                amplitude_standard_error=ase, maximum_off_diagonal_correlation=cmax, replica_masses=ms)


HADAMARD_8 = [
    [1, 1, 1, 1, 1, 1, 1, 1], [1, -1, 1, -1, 1, -1, 1, -1], [1, 1, -1, -1, 1, 1, -1, -1],
    [1, -1, -1, 1, 1, -1, -1, 1], [1, 1, 1, 1, -1, -1, -1, -1], [1, -1, 1, -1, -1, 1, -1, 1],
    [1, 1, -1, -1, -1, -1, 1, 1], [1, -1, -1, 1, -1, 1, 1, -1],
]


def canonical_scalar_fit_fixture():  # :652-699
    full = analytic_periodic_single_state(16, 0.5, 1.0)
    jk = []
    for row in HADAMARD_8:
        rep = list(full)
        for tau in range(1, 8):
            pattern = float(row[1]) if tau == 1 else float(row[tau]) + 0.25 * float(row[1])
            rep[tau] = full[tau] * (1.0 + 0.005 * pattern)
        jk.append(rep)
    return full, jk, [(1, 3), (2, 5), (3, 6)]


# ----------------------------------------------- glueball_scalar_variational.rs
def _sym_ok(m):  # :136-147
    r = abs(m[0][1] - m[1][0])
    if not fin(r):
        return False
    scale = max(abs(m[0][1]), abs(m[1][0]), 1.0)
    return fin(scale) and r <= 2.0e-12 * scale


def _stable_average(x, y):  # :149-165
    if not fin(x) or not fin(y):
        raise _err("NonFiniteAccumulator")
    if np.float64(x).tobytes() == np.float64(y).tobytes():
        return x
    d = y - x
    a = x + 0.5 * d
    if not fin(d) or not fin(a):
        raise _err("NonFiniteAccumulator")
    return a


def _symmetrized(m):  # :167-170
    c = _stable_average(m[0][1], m[1][0])
    return [[m[0][0], c], [c, m[1][1]]]


def _validate_series(ms):  # :172-185
    if len(ms) < 3 or len(ms) > 64:
        raise _err("CorrelatorShapeMismatch")
    for m in ms:
        if not all(fin(v) for row in m for v in row):
            raise _err("NonFiniteAccumulator")
        if not _sym_ok(m):
            raise _err("MatrixNotSymmetric")


def _cholesky_metric_norm(v, l00, l10, l11):  # :192-205
    first = l00 * v[0] + l10 * v[1]
    second = l11 * v[1]
    n = hypot(first, second)
    if not fin(first) or not fin(second) or not fin(n) or n <= 0.0:
        raise _err("ProjectionNormalizationFailure")
    return n


def _divide_by_product(num, d1, d2):  # :207-248
    if not fin(num) or not fin(d1) or not fin(d2) or d1 == 0.0 or d2 == 0.0:
        raise _err("NonFiniteAccumulator")
    large, small = (d1, d2) if abs(d1) >= abs(d2) else (d2, d1)
    i1 = num / large
    v1 = i1 / small
    if fin(i1) and fin(v1) and (i1 != 0.0 or num == 0.0):
        return v1
    i2 = num / small
    v2 = i2 / large
    if fin(i2) and fin(v2):
        return v2
    raise _err("NonFiniteAccumulator")


def _compensated_det(a, d, b):  # :250-282
    p = a * d
    q = b * b
    pe = fma(a, d, -p)
    qe = fma(b, b, -q)
    return (p - q) + (pe - qe)


def _solve_characteristic(rdet, roff, d1, d2, off):  # :284-378
    if not fin(rdet) or rdet <= 0.0 or not fin(roff) or not fin(d1) or not fin(d2) or not fin(off):
        raise _err("NonFiniteAccumulator")
    pd, md = 1.0 + roff, 1.0 - roff
    srd = math.sqrt(rdet)
    if pd <= 0.0 or md <= 0.0 or srd <= 0.0:
        raise _err("ReferenceMatrixNotPositiveDefinite")
    avg = 0.5 * d1 + 0.5 * d2
    plus, minus = avg + off, avg - off
    cross = 0.5 * d1 - 0.5 * d2
    if not (fin(avg) and fin(plus) and fin(minus) and fin(cross)):
        raise _err("NonFiniteAccumulator")
    t1, t2, to = plus / pd, minus / md, cross / srd
    if not (fin(t1) and fin(t2) and fin(to)):
        raise _err("NonFiniteAccumulator")
    if to == 0.0 and t1 == t2:
        raise _err("DegeneratePrincipalEigenvector")
    mid = 0.5 * t1 + 0.5 * t2
    half = 0.5 * t1 - 0.5 * t2
    rad = hypot(half, to)
    if not (fin(mid) and fin(half) and fin(rad)):
        raise _err("NonFiniteAccumulator")
    if rad == 0.0:
        raise _err("DegeneratePrincipalEigenvector")
    if mid >= 0.0:
        pr = mid + rad
        if not fin(pr):
            raise _err("NonFiniteAccumulator")
        return pr
    smaller = mid - rad
    if not fin(smaller) or smaller == 0.0:
        raise _err("NonFiniteAccumulator")
    det = _compensated_det(d1, d2, off)
    return _divide_by_product(det, rdet, smaller)


def _principal_eigenvalue(ref, diag):  # :380-464
    ref = _symmetrized(ref)
    diag = _symmetrized(diag)
    r1, r2 = ref[0][0], ref[1][1]
    if not fin(r1) or not fin(r2) or r1 <= 0.0 or r2 <= 0.0:
        raise _err("ReferenceMatrixNotPositiveDefinite")
    rs = math.sqrt(r1) * math.sqrt(r2)
    roff = ref[0][1] / rs
    rdet = 1.0 - roff * roff
    if not fin(roff) or not fin(rdet) or rdet <= 1.0e-14:
        raise _err("ReferenceMatrixNotPositiveDefinite")
    d1, d2, off = diag[0][0] / r1, diag[1][1] / r2, diag[0][1] / rs
    if not (fin(d1) and fin(d2) and fin(off)):
        raise _err("NonFiniteAccumulator")
    if roff == 0.0 and off == 0.0:
        if d1 == d2:
            raise _err("DegeneratePrincipalEigenvector")
        return max(d1, d2)
    try:
        return _solve_characteristic(rdet, roff, d1, d2, off)
    except OracleError as e:
        if e.name != "NonFiniteAccumulator":
            raise
    scale = max(abs(d1), abs(d2), abs(off))
    if not fin(scale) or scale <= 0.0:
        raise _err("NonFiniteAccumulator")
    root = scale * _solve_characteristic(rdet, roff, d1 / scale, d2 / scale, off / scale)
    if not fin(root):
        raise _err("NonFiniteAccumulator")
    return root


def _whitened(m, r1, r2, rs, roff, rdet):  # :481-566
    m = _symmetrized(m)
    n1, n2, no = m[0][0] / r1, m[1][1] / r2, m[0][1] / rs
    if not (fin(n1) and fin(n2) and fin(no)):
        raise _err("NonFiniteAccumulator")
    if roff == 0.0:
        return [[n1, no], [no, n2]]
    scale = max(abs(n1), abs(n2), abs(no))
    if scale == 0.0:
        return [[0.0, 0.0], [0.0, 0.0]]
    f, s, o = n1 / scale, n2 / scale, no / scale
    sd = math.sqrt(rdet)
    wf = f
    wo = fma(-roff, f, o) / sd
    ln = fma(roff, roff * f, fma(-2.0 * roff, o, s))
    ws = ln / rdet
    return [[wf * scale, wo * scale], [wo * scale, ws * scale]]


def _whitened_projection(m, c0, r1, r2, rs, roff, rdet, mv):  # :568-603
    m = _symmetrized(m)
    if m == c0:
        return 1.0
    w = _whitened(m, r1, r2, rs, roff, rdet)
    row0 = fma(w[0][0], mv[0], w[0][1] * mv[1])
    row1 = fma(w[1][0], mv[0], w[1][1] * mv[1])
    v = fma(mv[0], row0, mv[1] * row1)
    if not fin(v):
        raise _err("NonFiniteAccumulator")
    return v


def principal_generalized_projection(ms, t0=1, t1=2):  # :746-925
    ms = [[[float(v) for v in row] for row in m] for m in ms]
    _validate_series(ms)
    if t0 != 1 or t1 != 2 or t1 >= len(ms):
        raise _err("InvalidReferenceTimes")
    c0, c1 = _symmetrized(ms[t0]), _symmetrized(ms[t1])
    a, d = c0[0][0], c0[1][1]
    if not fin(a) or not fin(d) or a <= 0.0 or d <= 0.0:
        raise _err("ReferenceMatrixNotPositiveDefinite")
    ds = math.sqrt(a) * math.sqrt(d)
    noff = c0[0][1] / ds
    ndet = 1.0 - noff * noff
    if not fin(noff) or not fin(ndet) or ndet <= 1.0e-14:
        raise _err("ReferenceMatrixNotPositiveDefinite")
    l00 = math.sqrt(a)
    l10 = c0[1][0] / l00
    l11 = math.sqrt(d) * math.sqrt(ndet)
    inv_lower = -_divide_by_product(l10, l00, l11)
    inv_l = [[1.0 / l00, 0.0], [inv_lower, 1.0 / l11]]
    w = _whitened(c1, a, d, ds, noff, ndet)
    wa, wd, wo = w[0][0], w[1][1], w[0][1]
    lam = _principal_eigenvalue(c0, c1)
    if not fin(lam) or lam <= 0.0:
        raise _err("NonPositivePrincipalEigenvalue")
    if wo != 0.0:
        half = 0.5 * wa - 0.5 * wd
        rad = hypot(half, wo)
        if rad == 0.0:
            raise _err("DegeneratePrincipalEigenvector")
        cs = max(abs(rad), abs(half), abs(wo))
        if half >= 0.0:
            ev = [rad / cs + half / cs, wo / cs]
        else:
            ev = [wo / cs, rad / cs - half / cs]
    elif wa > wd:
        ev = [1.0, 0.0]
    elif wd > wa:
        ev = [0.0, 1.0]
    else:
        raise _err("DegeneratePrincipalEigenvector")
    en = hypot(ev[0], ev[1])
    ev = [ev[0] / en, ev[1] / en]
    inv_lt = [[inv_l[0][0], inv_l[1][0]], [inv_l[0][1], inv_l[1][1]]]
    coef = [inv_lt[0][0] * ev[0] + inv_lt[0][1] * ev[1], inv_lt[1][0] * ev[0] + inv_lt[1][1] * ev[1]]
    mn = _cholesky_metric_norm(coef, l00, l10, l11)
    coef = [coef[0] / mn, coef[1] / mn]
    if abs(coef[0]) > 1.0e-14:
        if coef[0] < 0.0:
            coef = [-coef[0], -coef[1]]
    elif coef[1] < 0.0:
        coef = [-coef[0], -coef[1]]
    mv = [l00 * coef[0] + l10 * coef[1], l11 * coef[1]]
    nn = hypot(mv[0], mv[1])
    if not fin(nn * nn) or abs(nn * nn - 1.0) > 2.0e-12:
        raise _err("ProjectionNormalizationFailure")
    proj = []
    for t, m in enumerate(ms):
        if t == t0:
            proj.append(1.0)
        elif t == t1:
            proj.append(lam)
        else:
            proj.append(_whitened_projection(m, c0, a, d, ds, noff, ndet, mv))
    return lam, coef, proj


def scalar_variational_connected_correlator(plaq, rect):  # :628-744 on operator samples
    plaq = [[float(v) for v in r] for r in plaq]
    rect = [[float(v) for v in r] for r in rect]
    nconf = len(plaq)
    if nconf == 0:
        raise _err("EmptyConfigurationSet")
    if nconf > 64:
        raise _err("TooManyConfigurations")
    T = len(plaq[0])
    if T < 3:
        raise _err("VarTimeExtentTooSmall")
    if T > 64:
        raise _err("VarTimeExtentTooLarge")
    s0 = s1 = 0.0
    samples = []
    for c in range(nconf):
        for t in range(T):
            v = (plaq[c][t], rect[c][t])
            if not (fin(v[0]) and fin(v[1])):
                raise _err("NonFiniteObservable")
            s0 += v[0]
            s1 += v[1]
            samples.append(v)
    cnt = float(nconf * T)
    m0, m1 = s0 / cnt, s1 / cnt
    mats = []
    for tau in range(T):
        pp = rr = pr = rp = 0.0
        for c in range(nconf):
            off = c * T
            for t in range(T):
                x, y = samples[off + t], samples[off + (t + tau) % T]
                pp += x[0] * y[0]
                rr += x[1] * y[1]
                pr += x[0] * y[1]
                rp += x[1] * y[0]
        cross = 0.5 * (pr / cnt + rp / cnt) - m0 * m1
        mats.append([[pp / cnt - m0 * m0, cross], [cross, rr / cnt - m1 * m1]])
    return (m0, m1), mats


def periodic_weight(mass, t, T):  # :927-929
    return math.exp(-mass * float(t)) + math.exp(-mass * float(T - t))


def canonical_two_state_matrices():  # :931-951
    gc, ec = (1.0, 0.5), (0.25, 1.0)
    out = []
    for t in range(16):
        g = 1.2 * periodic_weight(0.4, t, 16)
        e = 0.8 * periodic_weight(0.9, t, 16)
        out.append([[g * gc[i] * gc[j] + e * ec[i] * ec[j] for j in range(2)] for i in range(2)])
    return out
