// Host FP64 analysis of glueball correlators: the periodic cosh effective
// mass with delete-one jackknife errors, the correlated one-state fit over a
// predeclared set of plateau windows, and the two-operator (plaquette,
// rectangle) variational correlator matrix with its 2 x 2 generalized
// eigenvalue projection.  These consume the per-timeslice operator sums the
// device kernels produce; they run in microseconds and stay in FP64 on the
// host.  Every entry point fails closed with a status from include/ym_amd.h.
#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/ym_amd.h"

namespace {

// glueball_effective_mass.rs:17-22
constexpr int64_t kMinReplicas = 3, kMaxReplicas = 64;
constexpr int64_t kMinTime = 4, kMaxTime = 64;
// glueball_scalar_fit.rs:15-23
constexpr int64_t kMinWindow = 3, kMaxWindow = 8, kMaxCandidates = 16;
constexpr double kPlateauMaxZ = 2.5;
constexpr double kGridMin = 0.05, kGridMax = 2.0, kGridStep = 0.001;
constexpr int64_t kGridSteps = 1950;
constexpr double kPivotRel = 1.0e-12;
// glueball_scalar_variational.rs:24-29, glueball_correlators.rs:19-20
constexpr int64_t kT0 = 1, kT1 = 2;
constexpr double kSymmetryTol = 2.0e-12, kNormTol = 2.0e-12, kPdTol = 1.0e-14;
constexpr int64_t kMaxConfigurations = 64, kMaxCorrelatorTime = 64;
constexpr double kSqrtHalf = 0.70710678118654752440;

bool is_fin(double v) { return std::isfinite(v); }

// ------------------------------------------------------------ effective mass
// m(tau) = acosh((C(tau-1) + C(tau+1)) / (2 C(tau))).
int32_t cosh_mass(const double *c, int64_t tau, double *m) {
    const double centre = c[tau];
    if (centre <= 0.0) return YM_EM_NON_POSITIVE_CENTRE;
    const double num = c[tau - 1] + c[tau + 1];
    const double den = 2.0 * centre;
    const double arg = num / den;
    if (!is_fin(num) || !is_fin(den) || !is_fin(arg)) return YM_EM_NON_FINITE_STATISTIC;
    if (arg < 1.0) return YM_EM_ACOSH_DOMAIN;
    // acosh in the reference's form (Rust f64::acosh), so effective masses
    // match it to the last bit and rounding-level decisions agree
    const double v = std::log(arg + std::sqrt(arg - 1.0) * std::sqrt(arg + 1.0));
    if (!is_fin(v)) return YM_EM_NON_FINITE_STATISTIC;
    *m = v;
    return YM_OK;
}

int32_t validate_em(const double *full, int64_t T, const double *const *rep, const int64_t *rep_len, int64_t nrep,
                    const int64_t *lags, int64_t nlags) {
    if (T < kMinTime) return YM_EM_TIME_EXTENT_TOO_SMALL;
    if (T > kMaxTime) return YM_EM_TIME_EXTENT_TOO_LARGE;
    if (nrep < kMinReplicas) return YM_EM_TOO_FEW_REPLICAS;
    if (nrep > kMaxReplicas) return YM_EM_TOO_MANY_REPLICAS;
    if (nlags <= 0) return YM_EM_EMPTY_WINDOW;
    for (int64_t t = 0; t < T; ++t)
        if (!is_fin(full[t])) return YM_EM_NON_FINITE_CORRELATOR;
    for (int64_t r = 0; r < nrep; ++r)
        if (rep_len[r] != T) return YM_EM_SHAPE_MISMATCH;
    for (int64_t r = 0; r < nrep; ++r)
        for (int64_t t = 0; t < T; ++t)
            if (!is_fin(rep[r][t])) return YM_EM_NON_FINITE_CORRELATOR;
    int64_t prev = 0;
    for (int64_t i = 0; i < nlags; ++i) {
        const int64_t tau = lags[i];
        if (tau <= 0 || tau >= T - 1 || 2 * tau >= T || (i > 0 && tau <= prev)) return YM_EM_INVALID_LAG;
        prev = tau;
    }
    return YM_OK;
}

// Delete-one jackknife of a statistic: mean and sqrt((N-1)/N sum (v - mean)^2).
bool jackknife_se(const std::vector<double> &v, double *mean, double *se) {
    const double n = (double)v.size();
    double s = 0.0;
    for (double x : v) s += x;
    const double mb = s / n;
    double q = 0.0;
    for (double x : v) {
        const double d = x - mb;
        q += d * d;
    }
    const double var = ((n - 1.0) / n) * q;
    const double e = std::sqrt(var);
    if (!is_fin(mb) || !is_fin(var) || var < 0.0 || !is_fin(e)) return false;
    *mean = mb;
    *se = e;
    return true;
}

int32_t effective_mass_points(const double *full, int64_t T, const double *const *rep, const int64_t *rep_len,
                              int64_t nrep, const int64_t *lags, int64_t nlags, ym_effective_mass_point *out) {
    int32_t st = validate_em(full, T, rep, rep_len, nrep, lags, nlags);
    if (st) return st;
    std::vector<double> m((size_t)nrep);
    for (int64_t i = 0; i < nlags; ++i) {
        const int64_t tau = lags[i];
        double f;
        if ((st = cosh_mass(full, tau, &f))) return st;
        double acc = 0.0;
        for (int64_t r = 0; r < nrep; ++r) {
            if ((st = cosh_mass(rep[r], tau, &m[r]))) return st;
            acc += m[r];
            if (!is_fin(acc)) return YM_EM_NON_FINITE_STATISTIC;
        }
        double mb, se;
        if (!jackknife_se(m, &mb, &se)) return YM_EM_NON_FINITE_STATISTIC;
        out[i].tau = tau;
        out[i].full_estimate = f;
        out[i].jackknife_mean = mb;
        out[i].standard_error = se;
    }
    return YM_OK;
}

// ------------------------------------------------------------ correlated fit
using Mat = std::vector<std::vector<double>>;

struct Plateau {
    int64_t start = 0, end = 0;
    double weighted_mass = 0.0, max_z = 0.0;
};

// Inverse-variance weighted mean of the effective masses over the window and
// the largest standardized deviation from it.  Zero errors disqualify the
// window (returned as YM_FIT_DEGENERATE_PLATEAU).
int32_t plateau(const double *full, int64_t T, const double *const *rep, const int64_t *rep_len, int64_t nrep,
                int64_t a, int64_t b, Plateau *out) {
    const int64_t n = b - a + 1;
    std::vector<int64_t> lags((size_t)n);
    for (int64_t i = 0; i < n; ++i) lags[i] = a + i;
    std::vector<ym_effective_mass_point> pts((size_t)n);
    int32_t st = effective_mass_points(full, T, rep, rep_len, nrep, lags.data(), n, pts.data());
    if (st) return st;
    double ws = 0.0, w = 0.0;
    for (const auto &p : pts) {
        if (!is_fin(p.standard_error) || p.standard_error <= 0.0) return YM_FIT_DEGENERATE_PLATEAU;
        const double wi = 1.0 / (p.standard_error * p.standard_error);
        ws += wi * p.full_estimate;
        w += wi;
        if (!is_fin(ws) || !is_fin(w)) return YM_FIT_NON_FINITE_STATISTIC;
    }
    if (w <= 0.0) return YM_FIT_NON_FINITE_STATISTIC;
    const double mw = ws / w;
    if (!is_fin(mw)) return YM_FIT_NON_FINITE_STATISTIC;
    double zmax = 0.0;
    for (const auto &p : pts) {
        const double z = std::fabs((p.full_estimate - mw) / p.standard_error);
        if (!is_fin(z)) return YM_FIT_NON_FINITE_STATISTIC;
        zmax = std::fmax(zmax, z);
    }
    out->start = a;
    out->end = b;
    out->weighted_mass = mw;
    out->max_z = zmax;
    return YM_OK;
}

// Jackknife covariance of the replica correlators over the window:
// (N-1)/N sum_r (C_r - Cbar)(C_r - Cbar)^T.
int32_t window_covariance(const double *const *rep, int64_t nrep, int64_t a, int64_t n, Mat &cov) {
    if (nrep <= n) return YM_FIT_INSUFFICIENT_REPLICAS;
    std::vector<double> mean((size_t)n, 0.0);
    for (int64_t r = 0; r < nrep; ++r)
        for (int64_t i = 0; i < n; ++i) {
            const double v = rep[r][a + i];
            if (!is_fin(v)) return YM_FIT_NON_FINITE_STATISTIC;
            mean[i] += v;
        }
    const double nd = (double)nrep;
    for (double &m : mean) m /= nd;
    cov.assign((size_t)n, std::vector<double>((size_t)n, 0.0));
    for (int64_t r = 0; r < nrep; ++r)
        for (int64_t i = 0; i < n; ++i) {
            const double di = rep[r][a + i] - mean[i];
            for (int64_t j = 0; j < n; ++j) cov[i][j] += di * (rep[r][a + j] - mean[j]);
        }
    const double f = (nd - 1.0) / nd;
    for (auto &row : cov)
        for (double &v : row) {
            v *= f;
            if (!is_fin(v)) return YM_FIT_NON_FINITE_STATISTIC;
        }
    return YM_OK;
}

double max_abs(const Mat &m) {
    double s = 0.0;
    for (const auto &row : m)
        for (double v : row) s = std::fmax(s, std::fabs(v));
    return s;
}

// Positive definiteness by a Cholesky factorisation whose pivots must exceed
// 1e-12 of the largest entry.
int32_t check_positive_definite(const Mat &c) {
    const size_t n = c.size();
    const double scale = max_abs(c);
    if (n == 0 || !is_fin(scale) || scale <= 0.0) return YM_FIT_COVARIANCE_NOT_PD;
    Mat l(n, std::vector<double>(n, 0.0));
    for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j <= i; ++j) {
            double s = c[i][j];
            for (size_t k = 0; k < j; ++k) s -= l[i][k] * l[j][k];
            if (i == j) {
                if (!is_fin(s) || s <= scale * kPivotRel) return YM_FIT_COVARIANCE_NOT_PD;
                l[i][i] = std::sqrt(s);
            } else {
                l[i][j] = s / l[j][j];
            }
        }
    return YM_OK;
}

// Gauss-Jordan inverse with partial pivoting (pivots below 1e-12 of the
// largest entry are singular).
int32_t invert(const Mat &c, Mat &inv) {
    int32_t st = check_positive_definite(c);
    if (st) return st;
    const size_t n = c.size();
    const double scale = max_abs(c);
    Mat a(n, std::vector<double>(2 * n, 0.0));
    for (size_t i = 0; i < n; ++i) {
        for (size_t j = 0; j < n; ++j) a[i][j] = c[i][j];
        a[i][n + i] = 1.0;
    }
    for (size_t col = 0; col < n; ++col) {
        size_t piv = col;
        double best = std::fabs(a[col][col]);
        for (size_t r = col + 1; r < n; ++r)
            if (std::fabs(a[r][col]) > best) {
                best = std::fabs(a[r][col]);
                piv = r;
            }
        if (!is_fin(best) || best <= scale * kPivotRel) return YM_FIT_SINGULAR_COVARIANCE;
        if (piv != col) std::swap(a[piv], a[col]);
        const double p = a[col][col];
        for (double &v : a[col]) v /= p;
        for (size_t r = 0; r < n; ++r) {
            if (r == col) continue;
            const double f = a[r][col];
            for (size_t k = 0; k < 2 * n; ++k) a[r][k] -= f * a[col][k];
        }
    }
    inv.assign(n, std::vector<double>(n, 0.0));
    for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j < n; ++j) {
            if (!is_fin(a[i][n + j])) return YM_FIT_SINGULAR_COVARIANCE;
            inv[i][j] = a[i][n + j];
        }
    return YM_OK;
}

// x^T M y, accumulated row by row.
bool bilinear(const std::vector<double> &x, const Mat &m, const std::vector<double> &y, double *out) {
    double s = 0.0;
    for (size_t i = 0; i < m.size(); ++i) {
        double inner = 0.0;
        for (size_t j = 0; j < m.size(); ++j) inner += m[i][j] * y[j];
        s += x[i] * inner;
    }
    *out = s;
    return is_fin(s);
}

struct GridFit {
    double mass, amplitude, chi2;
};

// Profile likelihood on the fixed mass grid 0.05, 0.051, ..., 2.0: for each
// mass the amplitude minimising chi^2 = r^T C^-1 r (r = data - A f(m),
// f(m)_tau = exp(-m tau) + exp(-m (T - tau))) is b^T C^-1 d / b^T C^-1 b;
// the smallest chi^2 wins, ties to the smaller mass.
int32_t grid_fit(int64_t T, int64_t a, const std::vector<double> &d, const Mat &inv, GridFit *out) {
    const size_t n = d.size();
    for (double v : d)
        if (!is_fin(v)) return YM_FIT_NON_FINITE_STATISTIC;
    bool have = false;
    GridFit best{0.0, 0.0, 0.0};
    std::vector<double> f(n), r(n);
    for (int64_t g = 0; g <= kGridSteps; ++g) {
        const double m = kGridMin + kGridStep * (double)g;
        if (m > kGridMax + 1.0e-15) break;
        for (size_t i = 0; i < n; ++i) {
            const int64_t tau = a + (int64_t)i;
            f[i] = std::exp(-m * (double)tau) + std::exp(-m * (double)(T - tau));
            if (!is_fin(f[i]) || f[i] <= 0.0) return YM_FIT_NON_FINITE_STATISTIC;
        }
        double den, num, chi2;
        if (!bilinear(f, inv, f, &den)) return YM_FIT_NON_FINITE_STATISTIC;
        if (den <= 0.0) continue;
        if (!bilinear(f, inv, d, &num)) return YM_FIT_NON_FINITE_STATISTIC;
        const double amp = num / den;
        if (amp <= 0.0 || !is_fin(amp)) continue;
        for (size_t i = 0; i < n; ++i) r[i] = d[i] - amp * f[i];
        if (!bilinear(r, inv, r, &chi2)) return YM_FIT_NON_FINITE_STATISTIC;
        if (chi2 < 0.0) continue;
        if (!have || chi2 < best.chi2 || (chi2 == best.chi2 && m < best.mass)) {
            best = {m, amp, chi2};
            have = true;
        }
    }
    if (!have) return YM_FIT_NO_GRID_FIT;
    *out = best;
    return YM_OK;
}

// --------------------------------------------------------- 2 x 2 GEVP
struct Sym2 {
    double a, b, d;  // [[a, b], [b, d]]
};

// Symmetrise by the midpoint of the two off-diagonal entries.
bool symmetrised(const double *m, Sym2 *s) {
    for (int k = 0; k < 4; ++k)
        if (!is_fin(m[k])) return false;
    const double mid = (m[1] == m[2]) ? m[1] : m[1] + 0.5 * (m[2] - m[1]);
    if (!is_fin(mid)) return false;
    *s = {m[0], mid, m[3]};
    return true;
}

bool symmetric_enough(const double *m) {
    const double r = std::fabs(m[1] - m[2]);
    const double scale = std::fmax(std::fmax(std::fabs(m[1]), std::fabs(m[2])), 1.0);
    return is_fin(r) && r <= kSymmetryTol * scale;
}

}  // namespace

extern "C" {

// glueball_effective_mass.rs:161-232
int32_t ym_cosh_effective_mass(const double *full, int64_t T, const double *const *replicas,
                               const int64_t *replica_lengths, int64_t nrep, const int64_t *lags, int64_t nlags,
                               ym_effective_mass_point *out) {
    if (!full || (nrep > 0 && (!replicas || !replica_lengths)) || (nlags > 0 && (!lags || !out)) || nrep < 0)
        return YM_INVALID_ARGUMENT;
    for (int64_t r = 0; r < nrep; ++r)
        if (!replicas[r]) return YM_INVALID_ARGUMENT;
    return effective_mass_points(full, T, replicas, replica_lengths, nrep, lags, nlags, out);
}

// glueball_effective_mass.rs:240-268
int32_t ym_analytic_periodic_single_state(int64_t T, double mass, double amplitude, double *out) {
    if (!out) return YM_INVALID_ARGUMENT;
    if (T < kMinTime || T > kMaxTime || !is_fin(mass) || mass <= 0.0 || !is_fin(amplitude) || amplitude <= 0.0)
        return YM_EM_INVALID_FIXTURE;
    for (int64_t tau = 0; tau < T; ++tau) {
        const double v = amplitude * (std::exp(-mass * (double)tau) + std::exp(-mass * (double)(T - tau)));
        if (!is_fin(v)) return YM_EM_NON_FINITE_STATISTIC;
        out[tau] = v;
    }
    return YM_OK;
}

// glueball_scalar_fit.rs:156-650
int32_t ym_correlated_scalar_fit(const double *full, int64_t T, const double *const *replicas,
                                 const int64_t *replica_lengths, int64_t nrep, const int64_t *windows, int64_t nwin,
                                 ym_scalar_fit_result *out) {
    if (!full || !out || nrep < 0 || nwin < 0 || (nwin > 0 && !windows) ||
        (nrep > 0 && (!replicas || !replica_lengths)))
        return YM_INVALID_ARGUMENT;
    for (int64_t r = 0; r < nrep; ++r)
        if (!replicas[r]) return YM_INVALID_ARGUMENT;
    // candidate windows: non-empty, at most 16, 3..8 points, start >= 1,
    // ending before the temporal midpoint, no duplicates
    if (nwin == 0) return YM_FIT_EMPTY_CANDIDATES;
    if (nwin > kMaxCandidates) return YM_FIT_TOO_MANY_CANDIDATES;
    for (int64_t w = 0; w < nwin; ++w) {
        const int64_t a = windows[2 * w], b = windows[2 * w + 1];
        if (b < a || a < 0) return YM_FIT_INVALID_CANDIDATE;
        const int64_t n = b - a + 1;
        if (a == 0 || n < kMinWindow || n > kMaxWindow || 2 * b >= T) return YM_FIT_INVALID_CANDIDATE;
        for (int64_t v = 0; v < w; ++v)
            if (windows[2 * v] == a && windows[2 * v + 1] == b) return YM_FIT_DUPLICATE_CANDIDATE;
    }
    // plateau selection: longest passing window (max |z| <= 2.5), ties to the later start
    bool have = false;
    Plateau best;
    for (int64_t w = 0; w < nwin; ++w) {
        Plateau p;
        const int32_t st = plateau(full, T, replicas, replica_lengths, nrep, windows[2 * w], windows[2 * w + 1], &p);
        if (st == YM_FIT_DEGENERATE_PLATEAU) continue;
        if (st) return st;
        if (p.max_z > kPlateauMaxZ) continue;
        const int64_t n = p.end - p.start + 1;
        const int64_t nb = best.end - best.start + 1;
        if (!have || n > nb || (n == nb && p.start > best.start)) {
            best = p;
            have = true;
        }
    }
    if (!have) return YM_FIT_NO_PLATEAU;
    const int64_t a = best.start, n = best.end - best.start + 1;
    Mat cov, inv;
    int32_t st = window_covariance(replicas, nrep, a, n, cov);
    if (st) return st;
    if ((st = invert(cov, inv))) return st;
    // largest |correlation| between distinct window points
    double cmax = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        if (!(cov[i][i] > 0.0) || !is_fin(cov[i][i])) return YM_FIT_COVARIANCE_NOT_PD;
        const double si = std::sqrt(cov[i][i]);
        for (int64_t j = 0; j < i; ++j) {
            const double sj = std::sqrt(cov[j][j]);
            const double c = std::fabs((cov[i][j] / si) / sj);
            if (!is_fin(c)) return YM_FIT_NON_FINITE_STATISTIC;
            cmax = std::fmax(cmax, c);
        }
    }
    std::vector<double> d((size_t)n);
    for (int64_t i = 0; i < n; ++i) d[i] = full[a + i];
    GridFit ff;
    if ((st = grid_fit(T, a, d, inv, &ff))) return st;
    std::vector<double> ms((size_t)nrep), as((size_t)nrep);
    for (int64_t r = 0; r < nrep; ++r) {
        for (int64_t i = 0; i < n; ++i) d[i] = replicas[r][a + i];
        GridFit rf;
        if ((st = grid_fit(T, a, d, inv, &rf))) return st;
        ms[r] = rf.mass;
        as[r] = rf.amplitude;
    }
    double mm, mse, am, ase;
    if (!jackknife_se(ms, &mm, &mse) || !jackknife_se(as, &am, &ase)) return YM_FIT_NON_FINITE_STATISTIC;
    const int64_t dof = n - 2;
    const double chi2dof = ff.chi2 / (double)dof;
    if (!is_fin(chi2dof)) return YM_FIT_NON_FINITE_STATISTIC;
    out->window_start = a;
    out->window_end = best.end;
    out->plateau_weighted_effective_mass = best.weighted_mass;
    out->plateau_maximum_z = best.max_z;
    out->fitted_mass = ff.mass;
    out->fitted_amplitude = ff.amplitude;
    out->chi_square = ff.chi2;
    out->degrees_of_freedom = dof;
    out->chi_square_per_dof = chi2dof;
    out->mass_jackknife_mean = mm;
    out->mass_standard_error = mse;
    out->amplitude_jackknife_mean = am;
    out->amplitude_standard_error = ase;
    out->maximum_off_diagonal_correlation = cmax;
    return YM_OK;
}

// glueball_scalar_variational.rs:628-744 on operator samples: samples holds
// (plaquette, rectangle) pairs, row-major [configuration][time][2].
// C_ij(tau) = 1/2 [<O_i(t) O_j(t+tau)> + <O_j(t) O_i(t+tau)>] - <O_i><O_j>,
// every configuration and time origin weighted equally; the diagonal must
// replay the single-operator periodic estimator to 2e-12.
int32_t ym_variational_connected_correlator(const double *samples, int64_t nconf, int64_t T, double *vacuum_means,
                                            double *matrices) {
    if (!samples || !vacuum_means || !matrices || nconf < 0 || T < 0) return YM_INVALID_ARGUMENT;
    if (nconf == 0) return YM_EMPTY_CONFIGURATION_SET;
    if (nconf > kMaxConfigurations) return YM_TOO_MANY_CONFIGURATIONS;
    if (T < 3) return YM_VAR_TIME_EXTENT_TOO_SMALL;
    if (T > kMaxCorrelatorTime) return YM_TIME_EXTENT_TOO_LARGE;
    const int64_t count = nconf * T;
    double s0 = 0.0, s1 = 0.0;
    for (int64_t i = 0; i < count; ++i) {
        const double p = samples[2 * i], r = samples[2 * i + 1];
        if (!is_fin(p) || !is_fin(r)) return YM_NON_FINITE_OBSERVABLE;
        s0 += p;
        s1 += r;
        if (!is_fin(s0) || !is_fin(s1)) return YM_NON_FINITE_ACCUMULATOR;
    }
    const double cnt = (double)count;
    const double m0 = s0 / cnt, m1 = s1 / cnt;
    if (!is_fin(m0) || !is_fin(m1)) return YM_NON_FINITE_ACCUMULATOR;
    for (int64_t tau = 0; tau < T; ++tau) {
        double pp = 0.0, rr = 0.0, pr = 0.0, rp = 0.0;
        for (int64_t c = 0; c < nconf; ++c) {
            const double *row = samples + 2 * c * T;
            for (int64_t t = 0; t < T; ++t) {
                const double *x = row + 2 * t, *y = row + 2 * ((t + tau) % T);
                pp += x[0] * y[0];
                rr += x[1] * y[1];
                pr += x[0] * y[1];
                rp += x[1] * y[0];
            }
            if (!is_fin(pp) || !is_fin(rr) || !is_fin(pr) || !is_fin(rp)) return YM_NON_FINITE_ACCUMULATOR;
        }
        const double cross = 0.5 * (pr / cnt + rp / cnt) - m0 * m1;
        double *m = matrices + 4 * tau;
        m[0] = pp / cnt - m0 * m0;
        m[1] = cross;
        m[2] = cross;
        m[3] = rr / cnt - m1 * m1;
        for (int k = 0; k < 4; ++k)
            if (!is_fin(m[k])) return YM_NON_FINITE_ACCUMULATOR;
    }
    // the diagonal replays the single-operator estimator
    std::vector<double> op((size_t)count), raw((size_t)T), conn((size_t)T);
    for (int k = 0; k < 2; ++k) {
        for (int64_t i = 0; i < count; ++i) op[i] = samples[2 * i + k];
        double mean;
        int32_t st = ym_periodic_connected(op.data(), nconf, T, &mean, raw.data(), conn.data());
        if (st) return st;
        for (int64_t tau = 0; tau < T; ++tau)
            if (std::fabs(matrices[4 * tau + 3 * k] - conn[tau]) > kSymmetryTol) return YM_CORRELATOR_SHAPE;
    }
    vacuum_means[0] = m0;
    vacuum_means[1] = m1;
    return YM_OK;
}

// glueball_scalar_variational.rs:746-925: largest lambda of
// C(t1) v = lambda C(t0) v (t0 = 1, t1 = 2) for symmetric 2 x 2 matrices
// (row-major, 4 per time), v normalised to v^T C(t0) v = 1 with its first
// (else second) component positive, and the projection v^T C(t) v at every
// time (exactly 1 at t0 and lambda at t1).
//
// Method: D = diag(sqrt C0_00, sqrt C0_11) makes C0 unit-diagonal with
// off-diagonal r; the basis (e0 + e1)/sqrt2, (e0 - e1)/sqrt2 diagonalises it
// to diag(1 + r, 1 - r); whitening by those square roots turns the problem
// into an ordinary symmetric 2 x 2 eigenproblem H u = lambda u with
// H = [[P / (1 + r), X / s], [X / s, M / (1 - r)]], P, M the +/- sums of the
// normalised C1, X half the difference of its diagonal, s = sqrt(1 - r^2).
// The results do not depend on the units of either operator (D) or on
// their order.
int32_t ym_principal_generalized_projection(const double *matrices, int64_t T, int64_t t0, int64_t t1,
                                            double *eigenvalue, double *coefficients, double *projected) {
    if (!matrices || !eigenvalue || !coefficients || !projected || T < 0) return YM_INVALID_ARGUMENT;
    if (T < 3 || T > kMaxCorrelatorTime) return YM_CORRELATOR_SHAPE;
    for (int64_t t = 0; t < T; ++t) {
        for (int k = 0; k < 4; ++k)
            if (!is_fin(matrices[4 * t + k])) return YM_NON_FINITE_ACCUMULATOR;
        if (!symmetric_enough(matrices + 4 * t)) return YM_VAR_MATRIX_NOT_SYMMETRIC;
    }
    if (t0 != kT0 || t1 != kT1 || t1 >= T) return YM_VAR_INVALID_REFERENCE_TIMES;
    Sym2 c0, c1;
    if (!symmetrised(matrices + 4 * t0, &c0) || !symmetrised(matrices + 4 * t1, &c1)) return YM_NON_FINITE_ACCUMULATOR;
    if (!(c0.a > 0.0) || !(c0.d > 0.0) || !is_fin(c0.a) || !is_fin(c0.d)) return YM_VAR_REFERENCE_NOT_PD;
    const double sa = std::sqrt(c0.a), sd = std::sqrt(c0.d);
    const double r = (c0.b / sa) / sd;
    const double det0 = (1.0 - r) * (1.0 + r);
    if (!is_fin(r) || !is_fin(det0) || det0 <= kPdTol) return YM_VAR_REFERENCE_NOT_PD;
    // normalised C1
    const double p = (c1.a / sa) / sa, q = (c1.b / sa) / sd, s = (c1.d / sd) / sd;
    if (!is_fin(p) || !is_fin(q) || !is_fin(s)) return YM_NON_FINITE_ACCUMULATOR;
    // whitened H in the +/- basis
    const double sq = std::sqrt(det0);
    const double hp = (0.5 * (p + s) + q) / (1.0 + r);
    const double hm = (0.5 * (p + s) - q) / (1.0 - r);
    const double hx = (0.5 * (p - s)) / sq;
    if (!is_fin(hp) || !is_fin(hm) || !is_fin(hx)) return YM_NON_FINITE_ACCUMULATOR;
    const double mid = 0.5 * (hp + hm), half = 0.5 * (hp - hm);
    const double rad = std::hypot(half, hx);
    if (!is_fin(mid) || !is_fin(rad)) return YM_NON_FINITE_ACCUMULATOR;
    if (rad == 0.0) return YM_VAR_DEGENERATE_EIGENVECTOR;
    double lam;
    if (mid >= 0.0) {
        lam = mid + rad;
    } else {
        // mid + rad cancels: lambda_max = det(H) / lambda_min
        const double lo = mid - rad;
        lam = (hp * hm - hx * hx) / lo;
    }
    if (!is_fin(lam)) return YM_NON_FINITE_ACCUMULATOR;
    if (lam <= 0.0) return YM_VAR_NON_POSITIVE_EIGENVALUE;
    // principal eigenvector of H without forming lambda - diagonal
    double u0, u1;
    if (half >= 0.0) {
        u0 = rad + half;
        u1 = hx;
    } else {
        u0 = hx;
        u1 = rad - half;
    }
    const double un = std::hypot(u0, u1);
    if (!is_fin(un) || un <= 0.0) return YM_VAR_DEGENERATE_EIGENVECTOR;
    u0 /= un;
    u1 /= un;
    // back to the operator basis: w = (u+ / sqrt(1+r) (1, 1) + u- / sqrt(1-r) (1, -1)) / sqrt2, v = D^-1 w
    const double kp = u0 / std::sqrt(1.0 + r), km = u1 / std::sqrt(1.0 - r);
    double v0 = (kp + km) * kSqrtHalf / sa;
    double v1 = (kp - km) * kSqrtHalf / sd;
    // normalise in the C0 metric
    const double g = v0 * (c0.a * v0 + c0.b * v1) + v1 * (c0.b * v0 + c0.d * v1);
    if (!is_fin(g) || g <= 0.0) return YM_VAR_PROJECTION_NORMALIZATION;
    const double gs = std::sqrt(g);
    v0 /= gs;
    v1 /= gs;
    if ((std::fabs(v0) > kPdTol) ? v0 < 0.0 : v1 < 0.0) {
        v0 = -v0;
        v1 = -v1;
    }
    const double check = v0 * (c0.a * v0 + c0.b * v1) + v1 * (c0.b * v0 + c0.d * v1);
    if (!is_fin(check) || std::fabs(check - 1.0) > kNormTol) return YM_VAR_PROJECTION_NORMALIZATION;
    for (int64_t t = 0; t < T; ++t) {
        if (t == t0) {
            projected[t] = 1.0;
        } else if (t == t1) {
            projected[t] = lam;
        } else {
            Sym2 c;
            if (!symmetrised(matrices + 4 * t, &c)) return YM_NON_FINITE_ACCUMULATOR;
            const double val = v0 * (c.a * v0 + c.b * v1) + v1 * (c.b * v0 + c.d * v1);
            if (!is_fin(val)) return YM_NON_FINITE_ACCUMULATOR;
            projected[t] = val;
        }
    }
    *eigenvalue = lam;
    coefficients[0] = v0;
    coefficients[1] = v1;
    return YM_OK;
}

}  // extern "C"
