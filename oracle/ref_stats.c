/*
 * ORACLE -- TEST INFRASTRUCTURE ONLY.
 *
 * Reference-exact restatement of the host statistics of QSOLKCB/YANG-MILLS:
 * chain diagnostics (positive-window tau_int / ESS / SE), the sampling-
 * efficiency report, the half-split stationarity gate, pooled multi-chain
 * summaries, and the periodic connected-correlator estimator with its
 * delete-one jackknife.  Same summation order as the Rust iterators.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "ym_oracle.h"

/* su2_ensemble.rs:375-463.  `autocorr_out` (optional) receives max_lag+1
 * values. */
int ymo_diagnostics(const double *v, int64_t n, int64_t max_lag, ymo_chain_diagnostics *out, double *autocorr_out) {
    if (n < 2) return YMO_MEASUREMENT_COUNT;
    if (max_lag >= n || max_lag < 0) return YMO_LAG_OUT_OF_RANGE;
    for (int64_t i = 0; i < n; ++i) if (!isfinite(v[i])) return YMO_NON_FINITE_STATISTIC;
    double nd = (double)n;
    double s = -0.0;
    for (int64_t i = 0; i < n; ++i) s += v[i];
    double mean = s / nd;
    double vs = -0.0;
    for (int64_t i = 0; i < n; ++i) { double d = v[i] - mean; vs += d * d; }
    double var = vs / nd;
    if (!isfinite(mean) || !isfinite(var)) return YMO_NON_FINITE_STATISTIC;
    if (autocorr_out) autocorr_out[0] = 1.0;
    out->mean = mean;
    out->variance = var;
    if (var == 0.0) {
        if (autocorr_out) for (int64_t k = 1; k <= max_lag; ++k) autocorr_out[k] = 0.0;
        out->positive_window_lag = 0;
        out->tau_int = 0.5;
        out->ess = nd;
        out->standard_error = 0.0;
        return YMO_OK;
    }
    double tau = 0.5;
    int64_t wlag = 0;
    int open = 1;
    for (int64_t lag = 1; lag <= max_lag; ++lag) {
        int64_t pairs = n - lag;
        double c = -0.0;
        for (int64_t i = 0; i < pairs; ++i) c += (v[i] - mean) * (v[i + lag] - mean);
        double cov = c / (double)pairs;
        double rho = cov / var;
        if (!isfinite(rho)) return YMO_NON_FINITE_STATISTIC;
        if (autocorr_out) autocorr_out[lag] = rho;
        if (open && rho > 0.0) { tau += rho; wlag = lag; }
        else open = 0;
    }
    if (!isfinite(tau) || tau < 0.5) return YMO_NON_FINITE_STATISTIC;
    double ess = fmin(nd / (2.0 * tau), nd);
    double se = sqrt(var / ess);
    if (!isfinite(ess) || !isfinite(se)) return YMO_NON_FINITE_STATISTIC;
    out->positive_window_lag = wlag;
    out->tau_int = tau;
    out->ess = ess;
    out->standard_error = se;
    return YMO_OK;
}

/* sampling_efficiency.rs:57-135.  log_weights may be NULL (unit weights). */
int ymo_evaluate(const double *v, int64_t n, const double *log_weights, int64_t max_lag, double setup,
                 double thermalization, double sampling, ymo_efficiency_report *out) {
    if (n < 2 || n > 4096) return YMO_EFF_INVALID_LENGTH;
    if (max_lag == 0 || max_lag >= n || max_lag > 256) return YMO_EFF_INVALID_LAG;
    double costs[3] = {setup, thermalization, sampling};
    for (int i = 0; i < 3; ++i) if (!isfinite(costs[i]) || costs[i] < 0.0) return YMO_EFF_INVALID_COST;
    if (sampling == 0.0) return YMO_EFF_INVALID_COST;
    double total = setup + thermalization + sampling;
    if (!isfinite(total)) return YMO_EFF_INVALID_COST;
    double eff = 1.0;
    if (log_weights) {
        double mx = -INFINITY;
        for (int64_t i = 0; i < n; ++i) if (!isfinite(log_weights[i])) return YMO_EFF_INVALID_WEIGHT;
        for (int64_t i = 0; i < n; ++i) mx = fmax(mx, log_weights[i]);
        double sum = 0.0, sq = 0.0;
        for (int64_t i = 0; i < n; ++i) { double w = exp(log_weights[i] - mx); sum += w; sq += w * w; }
        eff = fmin((sum / (double)n) * (sum / sq), 1.0);
    }
    int constant = 1;
    for (int64_t i = 1; i < n; ++i) if (v[i] != v[0]) { constant = 0; break; }
    if (constant) return YMO_EFF_DEGENERATE;
    ymo_chain_diagnostics d;
    if (ymo_diagnostics(v, n, max_lag, &d, NULL)) return YMO_EFF_INVALID_DIAGNOSTIC;
    if (d.variance <= 0.0) return YMO_EFF_DEGENERATE;
    if (!isfinite(d.standard_error) || d.standard_error <= 0.0) return YMO_EFF_NON_FINITE;
    out->diag = d;
    out->reweighting_efficiency = eff;
    out->weight_only_ess = (double)n * eff;
    out->combined_ess_proxy = d.ess * eff;
    out->total_cost = total;
    out->sampling_cost_per_ess = sampling / out->combined_ess_proxy;
    out->total_cost_per_ess = total / out->combined_ess_proxy;
    double chk[5] = {eff, out->weight_only_ess, out->combined_ess_proxy, out->sampling_cost_per_ess, out->total_cost_per_ess};
    for (int i = 0; i < 5; ++i) if (!isfinite(chk[i]) || chk[i] <= 0.0) return YMO_EFF_NON_FINITE;
    return YMO_OK;
}

/* su3_benchmark.rs:186-237 / su2_benchmark.rs:188-245 -- half-split drift z. */
int ymo_stationarity(const double *v, int64_t n, int64_t half_lag, ymo_stationarity_report *out) {
    if (n % 2 != 0 || half_lag >= n / 2) return YMO_WRONG_CONFIGURATION;
    int64_t h = n / 2;
    int c1 = 1, c2 = 1;
    for (int64_t i = 1; i < h; ++i) { if (v[i] != v[0]) c1 = 0; if (v[h + i] != v[h]) c2 = 0; }
    if (c1 || c2) return YMO_STATIONARITY_FAILED;
    ymo_chain_diagnostics a, b;
    if (ymo_diagnostics(v, h, half_lag, &a, NULL) || ymo_diagnostics(v + h, h, half_lag, &b, NULL)) return YMO_STATIONARITY_FAILED;
    if (a.variance == 0.0 || b.variance == 0.0) return YMO_STATIONARITY_FAILED;
    double diff = fabs(a.mean - b.mean);
    double scale = sqrt(a.standard_error * a.standard_error + b.standard_error * b.standard_error);
    if (scale == 0.0) return YMO_STATIONARITY_FAILED;
    double z = diff / scale;
    if (!isfinite(z)) return YMO_NON_FINITE_STATISTIC;
    out->first_half_mean = a.mean;
    out->second_half_mean = b.mean;
    out->first_half_se = a.standard_error;
    out->second_half_se = b.standard_error;
    out->drift_z = z;
    return YMO_OK;
}

/* su3_confirmation.rs:105-125 (pairwise) and 127-263 (pooling): pooled mean
 * = mean of chain means, pooled SE = sqrt(sum SE^2)/k, aggregate ESS = sum,
 * max pairwise z = max |m_i-m_j| / sqrt(SE_i^2+SE_j^2).  `series` holds k
 * chains of n values each.  The two-chain benchmark summaries
 * (su2_benchmark.rs:247-362, su3_benchmark.rs:239-362) are the k=2 case. */
int ymo_pool_chains(const double *series, int64_t k, int64_t n, int64_t max_lag, int64_t half_lag, ymo_pool_report *out) {
    if (k < 1 || k > 16) return YMO_WRONG_CONFIGURATION;
    double sm = -0.0, sse = -0.0, sess = -0.0, maxz = 0.0, maxs = 0.0, mine = INFINITY;
    for (int64_t c = 0; c < k; ++c) {
        ymo_chain_diagnostics d;
        int st = ymo_diagnostics(series + c * n, n, max_lag, &d, NULL);
        if (st) return st;
        out->chain_mean[c] = d.mean;
        out->chain_se[c] = d.standard_error;
        out->chain_ess[c] = d.ess;
        ymo_stationarity_report s;
        if (half_lag > 0) {
            if ((st = ymo_stationarity(series + c * n, n, half_lag, &s))) return st;
            out->chain_stationarity_z[c] = s.drift_z;
            maxs = fmax(maxs, s.drift_z);
        } else {
            out->chain_stationarity_z[c] = 0.0;
        }
        mine = fmin(mine, d.ess);
    }
    for (int64_t c = 0; c < k; ++c) sm += out->chain_mean[c];
    for (int64_t c = 0; c < k; ++c) sse += out->chain_se[c] * out->chain_se[c];
    for (int64_t c = 0; c < k; ++c) sess += out->chain_ess[c];
    for (int64_t a = 0; a < k; ++a)
        for (int64_t b = a + 1; b < k; ++b) {
            double diff = fabs(out->chain_mean[a] - out->chain_mean[b]);
            double sc = sqrt(out->chain_se[a] * out->chain_se[a] + out->chain_se[b] * out->chain_se[b]);
            if (!isfinite(diff) || !isfinite(sc) || sc <= 0.0) return YMO_NON_FINITE_STATISTIC;
            maxz = fmax(maxz, diff / sc);
        }
    out->pooled_mean = sm / (double)k;
    out->pooled_se = sqrt(sse) / (double)k;
    out->aggregate_ess = sess;
    out->max_pairwise_z = maxz;
    out->max_stationarity_z = maxs;
    out->min_ess = mine;
    return YMO_OK;
}

/* glueball_correlators.rs:111-172 */
int ymo_periodic_connected(const double *samples, int64_t nconf, int64_t T, double *mean_out, double *raw, double *conn) {
    int64_t len = nconf * T;
    if (len == 0) return YMO_CORRELATOR_SHAPE;
    double ms = 0.0;
    for (int64_t i = 0; i < len; ++i) {
        if (!isfinite(samples[i])) return YMO_NON_FINITE_OBSERVABLE;
        ms += samples[i];
        if (!isfinite(ms)) return YMO_NON_FINITE_ACCUMULATOR;
    }
    double mean = ms / (double)len;
    double sq = mean * mean;
    for (int64_t tau = 0; tau < T; ++tau) {
        double s = 0.0;
        for (int64_t c = 0; c < nconf; ++c)
            for (int64_t t = 0; t < T; ++t) {
                double p = samples[c * T + t] * samples[c * T + (t + tau) % T];
                s += p;
                if (!isfinite(s)) return YMO_NON_FINITE_ACCUMULATOR;
            }
        double r = s / (double)len;
        raw[tau] = r;
        conn[tau] = r - sq;
    }
    *mean_out = mean;
    return YMO_OK;
}

/* glueball_scalar_production.rs:383-443 -- omitted < 0 means "keep all". */
int ymo_connected_from_samples(const double *samples, int64_t nconf, int64_t T, int64_t omitted, double *conn) {
    if (nconf <= 0 || T <= 0) return YMO_EMPTY_CONFIGURATION_SET;
    if (omitted >= nconf) return YMO_CORRELATOR_SHAPE;
    int64_t kept = omitted >= 0 ? nconf - 1 : nconf;
    if (kept == 0) return YMO_EMPTY_CONFIGURATION_SET;
    double cnt = (double)(kept * T);
    double ms = 0.0;
    for (int64_t c = 0; c < nconf; ++c) {
        if (c == omitted) continue;
        for (int64_t t = 0; t < T; ++t) {
            if (!isfinite(samples[c * T + t])) return YMO_NON_FINITE_STATISTIC;
            ms += samples[c * T + t];
        }
    }
    double mean = ms / cnt;
    double msq = mean * mean;
    for (int64_t tau = 0; tau < T; ++tau) {
        double s = 0.0;
        for (int64_t c = 0; c < nconf; ++c) {
            if (c == omitted) continue;
            for (int64_t t = 0; t < T; ++t) s += samples[c * T + t] * samples[c * T + (t + tau) % T];
        }
        double v = s / cnt - msq;
        if (!isfinite(v)) return YMO_NON_FINITE_STATISTIC;
        conn[tau] = v;
    }
    return YMO_OK;
}
