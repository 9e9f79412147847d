// Host FP64 statistics of the measurement pipeline: sampling-efficiency
// report, half-split stationarity gate, pooled multi-chain summaries, the
// periodic connected correlator and its delete-one jackknife.  Sums run in
// index order, as the reference's iterators do, so results agree with the
// reference to the last bit on the same inputs.
#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/ym_amd.h"

namespace {

bool all_finite(const double *v, int64_t n) {
    for (int64_t i = 0; i < n; ++i)
        if (!std::isfinite(v[i])) return false;
    return true;
}

}  // namespace

extern "C" {

// sampling_efficiency.rs:57-135
int32_t ym_evaluate(const double *v, int64_t n, const double *log_weights, int64_t max_lag, double setup,
                    double thermalization, double sampling, ym_efficiency_report *out) {
    if (!v || !out) return YM_INVALID_ARGUMENT;
    if (n < 2 || n > 4096) return YM_EFF_INVALID_LENGTH;
    if (max_lag == 0 || max_lag >= n || max_lag > 256) return YM_EFF_INVALID_LAG;
    for (double c : {setup, thermalization, sampling})
        if (!std::isfinite(c) || c < 0.0) return YM_EFF_INVALID_COST;
    if (sampling == 0.0) return YM_EFF_INVALID_COST;
    const double total = setup + thermalization + sampling;
    if (!std::isfinite(total)) return YM_EFF_INVALID_COST;
    // Kish efficiency of the importance weights, max-shifted before exp
    double kish = 1.0;
    if (log_weights) {
        if (!all_finite(log_weights, n)) return YM_EFF_INVALID_WEIGHT;
        double top = -INFINITY;
        for (int64_t i = 0; i < n; ++i) top = std::fmax(top, log_weights[i]);
        double s1 = 0.0, s2 = 0.0;
        for (int64_t i = 0; i < n; ++i) {
            const double w = std::exp(log_weights[i] - top);
            s1 += w;
            s2 += w * w;
        }
        kish = std::fmin((s1 / (double)n) * (s1 / s2), 1.0);
    }
    bool constant = true;
    for (int64_t i = 1; i < n && constant; ++i) constant = (v[i] == v[0]);
    if (constant) return YM_EFF_DEGENERATE;
    ym_chain_diagnostics d;
    if (ym_diagnostics(v, n, max_lag, &d, nullptr) != YM_OK) return YM_EFF_INVALID_DIAGNOSTIC;
    if (d.variance <= 0.0) return YM_EFF_DEGENERATE;
    if (!std::isfinite(d.standard_error) || d.standard_error <= 0.0) return YM_EFF_NON_FINITE;
    ym_efficiency_report r;
    r.diagnostics = d;
    r.reweighting_efficiency = kish;
    r.weight_only_ess = (double)n * kish;
    r.combined_ess_proxy = d.ess * kish;
    r.total_cost = total;
    r.sampling_cost_per_ess = sampling / r.combined_ess_proxy;
    r.total_cost_per_ess = total / r.combined_ess_proxy;
    for (double x : {r.reweighting_efficiency, r.weight_only_ess, r.combined_ess_proxy, r.sampling_cost_per_ess,
                     r.total_cost_per_ess})
        if (!std::isfinite(x) || x <= 0.0) return YM_EFF_NON_FINITE;
    *out = r;
    return YM_OK;
}

// su3_benchmark.rs:186-237, su2_benchmark.rs:188-245
int32_t ym_stationarity(const double *v, int64_t n, int64_t half_lag, ym_stationarity_report *out) {
    if (!v || !out) return YM_INVALID_ARGUMENT;
    if (n % 2 != 0 || half_lag < 0 || half_lag >= n / 2) return YM_WRONG_CONFIGURATION;
    const int64_t h = n / 2;
    bool flat1 = true, flat2 = true;
    for (int64_t i = 1; i < h; ++i) {
        flat1 = flat1 && v[i] == v[0];
        flat2 = flat2 && v[h + i] == v[h];
    }
    if (flat1 || flat2) return YM_STATIONARITY_FAILED;
    ym_chain_diagnostics a, b;
    if (ym_diagnostics(v, h, half_lag, &a, nullptr) != YM_OK || ym_diagnostics(v + h, h, half_lag, &b, nullptr) != YM_OK)
        return YM_STATIONARITY_FAILED;
    if (a.variance == 0.0 || b.variance == 0.0) return YM_STATIONARITY_FAILED;
    const double spread = std::sqrt(a.standard_error * a.standard_error + b.standard_error * b.standard_error);
    if (spread == 0.0) return YM_STATIONARITY_FAILED;
    const double z = std::fabs(a.mean - b.mean) / spread;
    if (!std::isfinite(z)) return YM_NON_FINITE_STATISTIC;
    out->first_half_mean = a.mean;
    out->second_half_mean = b.mean;
    out->first_half_se = a.standard_error;
    out->second_half_se = b.standard_error;
    out->drift_z = z;
    return YM_OK;
}

// su3_confirmation.rs:105-263; su3_benchmark.rs:239-362; su2_benchmark.rs:247-362
int32_t ym_pool_chains(const double *series, int64_t k, int64_t n, int64_t max_lag, int64_t half_lag,
                       double *chain_mean, double *chain_se, double *chain_ess, double *chain_drift_z,
                       ym_pool_summary *out) {
    if (!series || !out || k < 1 || n < 2) return YM_INVALID_ARGUMENT;
    std::vector<double> m(k), se(k), ess(k), dz(k, 0.0);
    double drift_max = 0.0, ess_min = INFINITY;
    for (int64_t c = 0; c < k; ++c) {
        ym_chain_diagnostics d;
        int32_t st = ym_diagnostics(series + c * n, n, max_lag, &d, nullptr);
        if (st != YM_OK) return st;
        m[c] = d.mean;
        se[c] = d.standard_error;
        ess[c] = d.ess;
        if (half_lag > 0) {
            ym_stationarity_report s;
            if ((st = ym_stationarity(series + c * n, n, half_lag, &s)) != YM_OK) return st;
            dz[c] = s.drift_z;
            drift_max = std::fmax(drift_max, s.drift_z);
        }
        ess_min = std::fmin(ess_min, d.ess);
    }
    double sum_m = -0.0, sum_se2 = -0.0, sum_ess = -0.0, zmax = 0.0;
    for (int64_t c = 0; c < k; ++c) sum_m += m[c];
    for (int64_t c = 0; c < k; ++c) sum_se2 += se[c] * se[c];
    for (int64_t c = 0; c < k; ++c) sum_ess += ess[c];
    for (int64_t a = 0; a < k; ++a)
        for (int64_t b = a + 1; b < k; ++b) {
            const double gap = std::fabs(m[a] - m[b]);
            const double scale = std::sqrt(se[a] * se[a] + se[b] * se[b]);
            if (!std::isfinite(gap) || !std::isfinite(scale) || scale <= 0.0) return YM_NON_FINITE_STATISTIC;
            zmax = std::fmax(zmax, gap / scale);
        }
    out->pooled_mean = sum_m / (double)k;
    out->pooled_se = std::sqrt(sum_se2) / (double)k;
    out->aggregate_ess = sum_ess;
    out->max_pairwise_z = zmax;
    out->max_stationarity_z = drift_max;
    out->min_ess = ess_min;
    for (int64_t c = 0; c < k; ++c) {
        if (chain_mean) chain_mean[c] = m[c];
        if (chain_se) chain_se[c] = se[c];
        if (chain_ess) chain_ess[c] = ess[c];
        if (chain_drift_z) chain_drift_z[c] = dz[c];
    }
    return YM_OK;
}

// glueball_correlators.rs:111-172
int32_t ym_periodic_connected(const double *x, int64_t nconf, int64_t T, double *mean, double *raw, double *conn) {
    if (!x || !mean || !raw || !conn || nconf < 0 || T < 0) return YM_INVALID_ARGUMENT;
    const int64_t count = nconf * T;
    if (count == 0) return YM_CORRELATOR_SHAPE;
    double acc = 0.0;
    for (int64_t i = 0; i < count; ++i) {
        if (!std::isfinite(x[i])) return YM_NON_FINITE_OBSERVABLE;
        acc += x[i];
        if (!std::isfinite(acc)) return YM_NON_FINITE_ACCUMULATOR;
    }
    const double mu = acc / (double)count;
    const double mu2 = mu * mu;
    for (int64_t tau = 0; tau < T; ++tau) {
        double s = 0.0;
        for (int64_t c = 0; c < nconf; ++c) {
            const double *row = x + c * T;
            for (int64_t t = 0; t < T; ++t) {
                s += row[t] * row[(t + tau) % T];
                if (!std::isfinite(s)) return YM_NON_FINITE_ACCUMULATOR;
            }
        }
        raw[tau] = s / (double)count;
        conn[tau] = raw[tau] - mu2;
    }
    *mean = mu;
    return YM_OK;
}

// glueball_scalar_production.rs:383-443
int32_t ym_connected_correlator(const double *x, int64_t nconf, int64_t T, int64_t omitted, double *conn) {
    if (!x || !conn) return YM_INVALID_ARGUMENT;
    if (nconf <= 0 || T <= 0) return YM_EMPTY_CONFIGURATION_SET;
    if (omitted >= nconf) return YM_CORRELATOR_SHAPE;
    const int64_t kept = omitted >= 0 ? nconf - 1 : nconf;
    if (kept == 0) return YM_EMPTY_CONFIGURATION_SET;
    const double count = (double)(kept * T);
    double acc = 0.0;
    for (int64_t c = 0; c < nconf; ++c) {
        if (c == omitted) continue;
        for (int64_t t = 0; t < T; ++t) {
            if (!std::isfinite(x[c * T + t])) return YM_NON_FINITE_STATISTIC;
            acc += x[c * T + t];
        }
    }
    const double mu = acc / count;
    const double mu2 = mu * mu;
    if (!std::isfinite(mu) || !std::isfinite(mu2)) return YM_NON_FINITE_STATISTIC;
    for (int64_t tau = 0; tau < T; ++tau) {
        double s = 0.0;
        for (int64_t c = 0; c < nconf; ++c) {
            if (c == omitted) continue;
            const double *row = x + c * T;
            for (int64_t t = 0; t < T; ++t) s += row[t] * row[(t + tau) % T];
        }
        const double value = s / count - mu2;
        if (!std::isfinite(value)) return YM_NON_FINITE_STATISTIC;
        conn[tau] = value;
    }
    return YM_OK;
}

// glueball_scalar_production.rs:445-470 and glueball_effective_mass.rs:161-215
int32_t ym_jackknife_correlator(const double *x, int64_t nconf, int64_t T, double *full, double *replicas,
                                double *jk_mean, double *jk_error) {
    if (!x || !full || !jk_mean || !jk_error) return YM_INVALID_ARGUMENT;
    int32_t st = ym_connected_correlator(x, nconf, T, -1, full);
    if (st != YM_OK) return st;
    if (nconf < 2) return YM_EMPTY_CONFIGURATION_SET;
    std::vector<double> rep((size_t)(nconf * T));
    for (int64_t c = 0; c < nconf; ++c)
        if ((st = ym_connected_correlator(x, nconf, T, c, rep.data() + c * T)) != YM_OK) return st;
    const double nd = (double)nconf;
    for (int64_t tau = 0; tau < T; ++tau) {
        double s = 0.0;
        for (int64_t c = 0; c < nconf; ++c) s += rep[c * T + tau];
        const double mb = s / nd;
        double q = 0.0;
        for (int64_t c = 0; c < nconf; ++c) {
            const double d = rep[c * T + tau] - mb;
            q += d * d;
        }
        jk_mean[tau] = mb;
        jk_error[tau] = std::sqrt(((nd - 1.0) / nd) * q);
        if (!std::isfinite(jk_mean[tau]) || !std::isfinite(jk_error[tau])) return YM_NON_FINITE_STATISTIC;
    }
    if (replicas)
        for (size_t i = 0; i < rep.size(); ++i) replicas[i] = rep[i];
    return YM_OK;
}

}  // extern "C"
