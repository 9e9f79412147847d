/*
 * ORACLE -- TEST INFRASTRUCTURE ONLY.
 *
 * Reference-exact CPU restatement of the SU(2) lattice core (lib.rs) and the
 * SU(2) heat-bath / over-relaxation ensemble (su2_ensemble.rs) of
 * QSOLKCB/YANG-MILLS.  Same operation order, no FMA contraction, glibc libm
 * (expm1, log1p, sin, cos, sqrt) as Rust std uses on linux-gnu.  Only tests/,
 * smoke() and bench.py's cpu_baseline leg may load it.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>

#include "ym_oracle.h"

#define SU2_MAX_LINKS (1u << 20)        /* lib.rs:21 */
#define SU2_MAX_BETA 8.0                /* su2_ensemble.rs:18 */
#define SU2_MAX_SWEEPS (1 << 20)        /* su2_ensemble.rs:19 */
#define SU2_MAX_OR 16                   /* su2_ensemble.rs:21 */
#define ALPHA_ZERO_LIMIT 1.0e-300       /* su2_ensemble.rs:22 */
#define SERIALIZATION_TOLERANCE 2.0e-12 /* lib.rs:20 */

typedef struct { double c[4]; } su2;

/* lib.rs:106-130 -- max-abs scaling, then division by the Euclidean norm. */
static int normalize_scaled4(const double v[4], double out[4]) {
    double scale = 0.0;
    for (int i = 0; i < 4; ++i) scale = fmax(scale, fabs(v[i]));
    if (scale == 0.0 || !isfinite(scale)) return 0;
    double s[4];
    for (int i = 0; i < 4; ++i) s[i] = v[i] / scale;
    double sq = -0.0;
    for (int i = 0; i < 4; ++i) sq += s[i] * s[i];
    double norm = sqrt(sq);
    if (norm == 0.0 || !isfinite(norm)) return 0;
    for (int i = 0; i < 4; ++i) out[i] = s[i] / norm;
    return 1;
}
static int normalize_scaled3(const double v[3], double out[3]) {
    double scale = 0.0;
    for (int i = 0; i < 3; ++i) scale = fmax(scale, fabs(v[i]));
    if (scale == 0.0 || !isfinite(scale)) return 0;
    double s[3];
    for (int i = 0; i < 3; ++i) s[i] = v[i] / scale;
    double sq = -0.0;
    for (int i = 0; i < 3; ++i) sq += s[i] * s[i];
    double norm = sqrt(sq);
    if (norm == 0.0 || !isfinite(norm)) return 0;
    for (int i = 0; i < 3; ++i) out[i] = s[i] / norm;
    return 1;
}

/* lib.rs:153-159 */
static int su2_from_normalized(const double c[4], su2 *out) {
    for (int i = 0; i < 4; ++i) if (!isfinite(c[i])) return YMO_SU2_NON_FINITE;
    if (!normalize_scaled4(c, out->c)) return YMO_SU2_ZERO_NORM;
    return YMO_OK;
}
static su2 su2_dagger(su2 a) { su2 r = {{a.c[0], -a.c[1], -a.c[2], -a.c[3]}}; return r; }
/* lib.rs:212-224 -- raw quaternion product, renormalised. */
static su2 su2_mul(su2 a, su2 b) {
    double a0 = a.c[0], a1 = a.c[1], a2 = a.c[2], a3 = a.c[3];
    double b0 = b.c[0], b1 = b.c[1], b2 = b.c[2], b3 = b.c[3];
    double raw[4] = {
        a0 * b0 - a1 * b1 - a2 * b2 - a3 * b3,
        a0 * b1 + a1 * b0 + a2 * b3 - a3 * b2,
        a0 * b2 - a1 * b3 + a2 * b0 + a3 * b1,
        a0 * b3 + a1 * b2 - a2 * b1 + a3 * b0,
    };
    su2 r;
    if (!normalize_scaled4(raw, r.c)) { r.c[0] = NAN; r.c[1] = r.c[2] = r.c[3] = NAN; }
    return r;
}
static double su2_norm_sq(su2 a) { double s = -0.0; for (int i = 0; i < 4; ++i) s += a.c[i] * a.c[i]; return s; }

/* lib.rs:175-186 */
int ymo_su2_from_axis_angle(const double axis[3], double angle, double out[4]) {
    if (!isfinite(angle) || !isfinite(axis[0]) || !isfinite(axis[1]) || !isfinite(axis[2])) return YMO_SU2_INVALID_AXIS;
    double ax[3];
    if (!normalize_scaled3(axis, ax)) return YMO_SU2_INVALID_AXIS;
    double half = 0.5 * angle;
    double s = sin(half);
    out[0] = cos(half); out[1] = s * ax[0]; out[2] = s * ax[1]; out[3] = s * ax[2];
    return YMO_OK;
}
void ymo_su2_mul(const double a[4], const double b[4], double out[4]) {
    su2 x, y; memcpy(x.c, a, sizeof x.c); memcpy(y.c, b, sizeof y.c);
    su2 r = su2_mul(x, y); memcpy(out, r.c, sizeof r.c);
}

/* ------------------------------------------------------------------ lattice */
typedef struct { int64_t n[4]; double *links; } lat2; /* ((site*4+mu)*4 + c) */

static inline int64_t s2_site(const lat2 *l, const int64_t x[4]) { return ((x[0] * l->n[1] + x[1]) * l->n[2] + x[2]) * l->n[3] + x[3]; }
static inline su2 s2_get(const lat2 *l, const int64_t x[4], int mu) { su2 r; memcpy(r.c, l->links + (s2_site(l, x) * 4 + mu) * 4, sizeof r.c); return r; }
static inline void s2_set(lat2 *l, const int64_t x[4], int mu, su2 v) { memcpy(l->links + (s2_site(l, x) * 4 + mu) * 4, v.c, sizeof v.c); }
static inline void s2_plus(const lat2 *l, const int64_t x[4], int mu, int64_t y[4]) { memcpy(y, x, 4 * sizeof(int64_t)); y[mu] = (y[mu] + 1) % l->n[mu]; }
static inline void s2_minus(const lat2 *l, const int64_t x[4], int mu, int64_t y[4]) { memcpy(y, x, 4 * sizeof(int64_t)); y[mu] = y[mu] == 0 ? l->n[mu] - 1 : y[mu] - 1; }
static int64_t s2_sites(const int64_t n[4]) { return n[0] * n[1] * n[2] * n[3]; }

#define FOR_SITES2(l, x)                                \
    for (x[0] = 0; x[0] < (l)->n[0]; ++x[0])            \
        for (x[1] = 0; x[1] < (l)->n[1]; ++x[1])        \
            for (x[2] = 0; x[2] < (l)->n[2]; ++x[2])    \
                for (x[3] = 0; x[3] < (l)->n[3]; ++x[3])

/* lib.rs:237-263 */
int ymo_su2_identity(const int64_t n[4], double *links) {
    for (int d = 0; d < 4; ++d) if (n[d] < 2) return YMO_INVALID_EXTENT;
    int64_t count = s2_sites(n) * 4;
    for (int64_t i = 0; i < count; ++i) { links[4 * i] = 1.0; links[4 * i + 1] = links[4 * i + 2] = links[4 * i + 3] = 0.0; }
    return YMO_OK;
}

/* lib.rs:348-361 */
static su2 s2_plaquette(const lat2 *l, const int64_t x[4], int mu, int nu) {
    int64_t xm[4], xn[4];
    s2_plus(l, x, mu, xm);
    s2_plus(l, x, nu, xn);
    return su2_mul(su2_mul(su2_mul(s2_get(l, x, mu), s2_get(l, xm, nu)), su2_dagger(s2_get(l, xn, mu))), su2_dagger(s2_get(l, x, nu)));
}
int ymo_su2_plaquette(const int64_t n[4], const double *links, const int64_t x[4], int mu, int nu, double out[4]) {
    if (mu < 0 || mu > 3 || nu < 0 || nu > 3) return YMO_DIRECTION_OUT_OF_RANGE;
    if (mu == nu) return YMO_DIRECTIONS_MUST_DIFFER;
    lat2 l = {{n[0], n[1], n[2], n[3]}, (double *)links};
    su2 p = s2_plaquette(&l, x, mu, nu);
    memcpy(out, p.c, sizeof p.c);
    return YMO_OK;
}

/* lib.rs:363-382 */
int ymo_su2_wilson_action(const int64_t n[4], const double *links, double beta, double *out) {
    if (!isfinite(beta)) return YMO_NON_FINITE_BETA;
    if (beta < 0.0) return YMO_NEGATIVE_BETA;
    lat2 l = {{n[0], n[1], n[2], n[3]}, (double *)links};
    double sum = 0.0;
    int64_t x[4];
    FOR_SITES2(&l, x) for (int mu = 0; mu < 4; ++mu) for (int nu = mu + 1; nu < 4; ++nu) sum += 1.0 - s2_plaquette(&l, x, mu, nu).c[0];
    if (!isfinite(sum)) return YMO_NON_FINITE_ACTION;
    double a = beta * sum;
    if (!isfinite(a)) return YMO_NON_FINITE_ACTION;
    *out = a;
    return YMO_OK;
}

/* lib.rs:384-416 -- steps encoded +(d+1) forward / -(d+1) backward. */
int ymo_su2_path_holonomy(const int64_t n[4], const double *links, const int64_t start[4], const int32_t *steps,
                          int64_t nsteps, double out[4], int64_t end[4]) {
    if (nsteps <= 0) return YMO_EMPTY_PATH;
    if (nsteps > (1 << 20)) return YMO_PATH_TOO_LONG;
    lat2 l = {{n[0], n[1], n[2], n[3]}, (double *)links};
    int64_t cur[4], y[4];
    memcpy(cur, start, sizeof cur);
    su2 h = {{1.0, 0.0, 0.0, 0.0}};
    for (int64_t s = 0; s < nsteps; ++s) {
        int code = steps[s];
        int d = (code > 0 ? code : -code) - 1;
        if (code == 0 || d > 3) return YMO_DIRECTION_OUT_OF_RANGE;
        su2 seg;
        if (code > 0) { seg = s2_get(&l, cur, d); s2_plus(&l, cur, d, y); }
        else { s2_minus(&l, cur, d, y); seg = su2_dagger(s2_get(&l, y, d)); }
        memcpy(cur, y, sizeof cur);
        h = su2_mul(h, seg);
    }
    memcpy(out, h.c, sizeof h.c);
    if (end) memcpy(end, cur, sizeof cur);
    return YMO_OK;
}

/* lib.rs:438-481 */
int ymo_su2_rect_wilson_loop(const int64_t n[4], const double *links, const int64_t start[4], int mu, int nu,
                             int64_t r, int64_t t, double *out) {
    if (mu < 0 || mu > 3 || nu < 0 || nu > 3) return YMO_DIRECTION_OUT_OF_RANGE;
    if (mu == nu) return YMO_DIRECTIONS_MUST_DIFFER;
    if (r <= 0 || t <= 0) return YMO_ZERO_RECTANGLE_EXTENT;
    int64_t k = 0, np = 2 * (r + t);
    if (np > (1 << 20)) return YMO_PATH_TOO_LONG;
    int32_t *buf = (int32_t *)malloc(sizeof(int32_t) * np);
    if (!buf) return YMO_ALLOCATION_FAILED;
    for (int64_t i = 0; i < r; ++i) buf[k++] = mu + 1;
    for (int64_t i = 0; i < t; ++i) buf[k++] = nu + 1;
    for (int64_t i = 0; i < r; ++i) buf[k++] = -(mu + 1);
    for (int64_t i = 0; i < t; ++i) buf[k++] = -(nu + 1);
    double h[4];
    int64_t end[4];
    int st = ymo_su2_path_holonomy(n, links, start, buf, k, h, end);
    free(buf);
    if (st) return st;
    for (int d = 0; d < 4; ++d) if (end[d] != start[d]) return YMO_PATH_NOT_CLOSED;
    *out = h[0];
    return YMO_OK;
}

/* lib.rs:483-505 */
int ymo_su2_polyakov_loop(const int64_t n[4], const double *links, const int64_t start[4], int dir, double *out) {
    if (dir < 0 || dir > 3) return YMO_DIRECTION_OUT_OF_RANGE;
    lat2 l = {{n[0], n[1], n[2], n[3]}, (double *)links};
    int64_t cur[4], y[4];
    memcpy(cur, start, sizeof cur);
    su2 h = {{1.0, 0.0, 0.0, 0.0}};
    for (int64_t s = 0; s < n[dir]; ++s) { h = su2_mul(h, s2_get(&l, cur, dir)); s2_plus(&l, cur, dir, y); memcpy(cur, y, sizeof cur); }
    *out = h.c[0];
    return YMO_OK;
}

/* lib.rs:507-546 -- 6-term staple, forward then backward per nu. */
static void s2_staple(const lat2 *l, const int64_t x[4], int mu, double out[4]) {
    double comp[4] = {0.0, 0.0, 0.0, 0.0};
    for (int nu = 0; nu < 4; ++nu) {
        if (nu == mu) continue;
        int64_t xpm[4], xpn[4], xmn[4], xmnpm[4];
        s2_plus(l, x, mu, xpm);
        s2_plus(l, x, nu, xpn);
        su2 f = su2_mul(su2_mul(s2_get(l, xpm, nu), su2_dagger(s2_get(l, xpn, mu))), su2_dagger(s2_get(l, x, nu)));
        s2_minus(l, x, nu, xmn);
        s2_plus(l, xmn, mu, xmnpm);
        su2 b = su2_mul(su2_mul(su2_dagger(s2_get(l, xmnpm, nu)), su2_dagger(s2_get(l, xmn, mu))), s2_get(l, xmn, nu));
        for (int i = 0; i < 4; ++i) comp[i] += f.c[i];
        for (int i = 0; i < 4; ++i) comp[i] += b.c[i];
    }
    memcpy(out, comp, sizeof comp);
}
int ymo_su2_staple(const int64_t n[4], const double *links, const int64_t x[4], int mu, double out[4]) {
    if (mu < 0 || mu > 3) return YMO_DIRECTION_OUT_OF_RANGE;
    lat2 l = {{n[0], n[1], n[2], n[3]}, (double *)links};
    s2_staple(&l, x, mu, out);
    return YMO_OK;
}

/* lib.rs:548-585 -- local action from the 6 plaquettes; the difference
 * evaluates the candidate on a copy of the lattice as the reference does. */
static double s2_local_action(const lat2 *l, const int64_t x[4], int mu, double beta) {
    double sum = 0.0;
    for (int nu = 0; nu < 4; ++nu) {
        if (nu == mu) continue;
        su2 f = s2_plaquette(l, x, mu, nu);
        int64_t xmn[4]; s2_minus(l, x, nu, xmn);
        su2 b = s2_plaquette(l, xmn, mu, nu);
        sum += 1.0 - f.c[0];
        sum += 1.0 - b.c[0];
    }
    return beta * sum;
}
int ymo_su2_local_action_difference(const int64_t n[4], const double *links, const int64_t x[4], int mu,
                                    const double cand[4], double beta, double *out) {
    if (!isfinite(beta)) return YMO_NON_FINITE_BETA;
    if (beta < 0.0) return YMO_NEGATIVE_BETA;
    lat2 l = {{n[0], n[1], n[2], n[3]}, (double *)links};
    double old_a = s2_local_action(&l, x, mu, beta);
    double *copy = (double *)malloc(sizeof(double) * 16 * s2_sites(n));
    if (!copy) return YMO_ALLOCATION_FAILED;
    memcpy(copy, links, sizeof(double) * 16 * s2_sites(n));
    lat2 c = {{n[0], n[1], n[2], n[3]}, copy};
    su2 v; memcpy(v.c, cand, sizeof v.c);
    s2_set(&c, x, mu, v);
    double new_a = s2_local_action(&c, x, mu, beta);
    free(copy);
    double d = new_a - old_a;
    if (!isfinite(d)) return YMO_NON_FINITE_ACTION;
    *out = d;
    return YMO_OK;
}

/* lib.rs:696-712 */
double ymo_su2_max_unit_residual(const int64_t n[4], const double *links) {
    int64_t count = s2_sites(n) * 4;
    double m = 0.0;
    for (int64_t i = 0; i < count; ++i) {
        su2 v; memcpy(v.c, links + 4 * i, sizeof v.c);
        double sq = su2_norm_sq(v);
        if (!isfinite(sq)) return INFINITY;
        m = fmax(m, fabs(sq - 1.0));
    }
    return m;
}

/* lib.rs:714-... deterministic fixtures */
int ymo_su2_deterministic_site_gauge(const int64_t x[4], double out[4]) {
    int64_t code = x[0] + 3 * x[1] + 7 * x[2] + 11 * x[3];
    double angle = 0.137 * ((double)code + 1.0);
    double axis[3] = {1.0, -0.5, 0.25};
    return ymo_su2_from_axis_angle(axis, angle, out);
}

/* ---------------------------------------------------------------- ensemble */
/* su2_ensemble.rs:195-202 */
static int staple_norm(const double c[4], double *out) {
    double sq = -0.0;
    for (int i = 0; i < 4; ++i) sq += c[i] * c[i];
    double n = sqrt(sq);
    if (!isfinite(n)) return YMO_NON_FINITE_STATISTIC;
    *out = n;
    return YMO_OK;
}
/* su2_ensemble.rs:204-209 */
static void uniform_sphere_direction(uint64_t *rng, double d[3]) {
    double z = 2.0 * ymo_splitmix_unit(rng) - 1.0;
    double phi = 6.283185307179586 * ymo_splitmix_unit(rng);
    double planar = sqrt(fmax(1.0 - z * z, 0.0));
    d[0] = planar * cos(phi);
    d[1] = planar * sin(phi);
    d[2] = z;
}
/* su2_ensemble.rs:220-254 */
static int sample_local_heatbath(double alpha, uint64_t *rng, int64_t max_attempts, su2 *out, uint64_t *attempts) {
    if (!isfinite(alpha) || alpha < 0.0) return YMO_NON_FINITE_STATISTIC;
    for (int64_t attempt = 1; attempt <= max_attempts; ++attempt) {
        double u = ymo_splitmix_unit(rng);
        double r0;
        if (alpha <= ALPHA_ZERO_LIMIT) {
            r0 = 2.0 * u - 1.0;
        } else {
            double tm = -expm1(-2.0 * alpha);
            double y = -log1p(-u * tm) / alpha;
            r0 = 1.0 - y;
            if (r0 < -1.0) r0 = -1.0;
            if (r0 > 1.0) r0 = 1.0;
        }
        double radial = sqrt(fmax(1.0 - r0 * r0, 0.0));
        if (ymo_splitmix_unit(rng) <= radial) {
            double d[3];
            uniform_sphere_direction(rng, d);
            double c[4] = {r0, radial * d[0], radial * d[1], radial * d[2]};
            int st = su2_from_normalized(c, out);
            if (st) return st;
            *attempts = (uint64_t)attempt;
            return YMO_OK;
        }
    }
    return YMO_HEATBATH_ATTEMPT_LIMIT;
}
/* su2_ensemble.rs:256-281 */
static int heatbath_link(lat2 *l, const int64_t x[4], int mu, double beta, uint64_t *rng, int64_t max_attempts, uint64_t *att) {
    double comp[4], norm;
    s2_staple(l, x, mu, comp);
    int st = staple_norm(comp, &norm);
    if (st) return st;
    double alpha = beta * norm;
    if (!isfinite(alpha)) return YMO_NON_FINITE_STATISTIC;
    su2 r;
    if ((st = sample_local_heatbath(alpha, rng, max_attempts, &r, att))) return st;
    su2 next = r;
    if (norm != 0.0) {
        su2 ns;
        if ((st = su2_from_normalized(comp, &ns))) return st;
        next = su2_mul(r, su2_dagger(ns));
    }
    s2_set(l, x, mu, next);
    return YMO_OK;
}
/* su2_ensemble.rs:283-299 */
static int overrelax_link(lat2 *l, const int64_t x[4], int mu) {
    double comp[4], norm;
    s2_staple(l, x, mu, comp);
    int st = staple_norm(comp, &norm);
    if (st) return st;
    if (norm == 0.0) return YMO_OK;
    su2 ns;
    if ((st = su2_from_normalized(comp, &ns))) return st;
    su2 old = s2_get(l, x, mu);
    su2 refl = su2_mul(su2_mul(su2_dagger(ns), su2_dagger(old)), su2_dagger(ns));
    s2_set(l, x, mu, refl);
    return YMO_OK;
}
/* su2_ensemble.rs:321-342 */
int ymo_su2_heatbath_sweep(const int64_t n[4], double *links, double beta, uint64_t *rng, int64_t max_attempts,
                           uint64_t *proposals, uint64_t *updates) {
    if (!isfinite(beta) || beta < 0.0 || beta > SU2_MAX_BETA) return YMO_BETA_OUT_OF_RANGE;
    if (max_attempts == 0) return YMO_ZERO_ATTEMPT_LIMIT;
    lat2 l = {{n[0], n[1], n[2], n[3]}, links};
    uint64_t prop = 0, upd = 0;
    int64_t x[4];
    FOR_SITES2(&l, x) for (int mu = 0; mu < 4; ++mu) {
        uint64_t att;
        int st = heatbath_link(&l, x, mu, beta, rng, max_attempts, &att);
        if (st) return st;
        prop += att;
        upd += 1;
    }
    *proposals = prop;
    *updates = upd;
    return YMO_OK;
}
/* su2_ensemble.rs:344-346 */
int ymo_su2_overrelaxation_sweep(const int64_t n[4], double *links) {
    lat2 l = {{n[0], n[1], n[2], n[3]}, links};
    int64_t x[4];
    FOR_SITES2(&l, x) for (int mu = 0; mu < 4; ++mu) {
        int st = overrelax_link(&l, x, mu);
        if (st) return st;
    }
    return YMO_OK;
}
/* su2_ensemble.rs:348-373 */
int ymo_su2_average_plaquette(const int64_t n[4], const double *links, double *out) {
    lat2 l = {{n[0], n[1], n[2], n[3]}, (double *)links};
    double sum = 0.0;
    int64_t count = 0, x[4];
    FOR_SITES2(&l, x) for (int mu = 0; mu < 4; ++mu) for (int nu = mu + 1; nu < 4; ++nu) {
        sum += s2_plaquette(&l, x, mu, nu).c[0];
        count += 1;
    }
    double v = sum / (double)count;
    if (!isfinite(v)) return YMO_NON_FINITE_STATISTIC;
    *out = v;
    return YMO_OK;
}

/* su2_ensemble.rs:57-107 */
static int su2_validate(const ymo_su2_chain_config *c) {
    if (!isfinite(c->beta) || c->beta < 0.0 || c->beta > SU2_MAX_BETA) return YMO_BETA_OUT_OF_RANGE;
    if (c->measurements > SU2_MAX_SWEEPS) return YMO_SCHEDULE_OUT_OF_RANGE;
    int64_t total = c->thermalization_sweeps + c->measurements * c->sweeps_between_measurements;
    if (total > SU2_MAX_SWEEPS) return YMO_SCHEDULE_OUT_OF_RANGE;
    if (c->measurements < 2) return YMO_MEASUREMENT_COUNT;
    if (c->sweeps_between_measurements == 0) return YMO_ZERO_STRIDE;
    if (c->overrelaxation_sweeps_per_heatbath > SU2_MAX_OR) return YMO_TOO_MANY_OR;
    if (c->max_heatbath_attempts_per_link == 0) return YMO_ZERO_ATTEMPT_LIMIT;
    if (c->autocorrelation_max_lag >= c->measurements) return YMO_LAG_OUT_OF_RANGE;
    for (int d = 0; d < 4; ++d) if (c->extents[d] < 2) return YMO_INVALID_EXTENT;
    if (!c->no_caps && s2_sites(c->extents) * 4 > (int64_t)SU2_MAX_LINKS) return YMO_LATTICE_TOO_LARGE;
    return YMO_OK;
}

/* su2_ensemble.rs:465-552 -- cold start, (HB + n_or OR) Markov sweeps. */
int ymo_su2_run_chain(const ymo_su2_chain_config *c, double *measurements_out, ymo_chain_diagnostics *diag,
                      ymo_su2_chain_report *rep, double *links_out) {
    int st = su2_validate(c);
    if (st) return st;
    int64_t count = s2_sites(c->extents) * 4;
    double *links = links_out ? links_out : (double *)malloc(sizeof(double) * 4 * count);
    if (!links) return YMO_ALLOCATION_FAILED;
    ymo_su2_identity(c->extents, links);
    uint64_t seed = ymo_derive_chain_seed(c->master_seed, c->chain_id), rng = seed;
    uint64_t hp = 0, hu = 0, p, u;
    int64_t nor = 0;
    int64_t total_meas_sweeps = c->measurements * c->sweeps_between_measurements;
    for (int64_t s = 0; s < c->thermalization_sweeps + total_meas_sweeps; ++s) {
        if ((st = ymo_su2_heatbath_sweep(c->extents, links, c->beta, &rng, c->max_heatbath_attempts_per_link, &p, &u))) goto done;
        hp += p; hu += u;
        for (int64_t k = 0; k < c->overrelaxation_sweeps_per_heatbath; ++k) {
            if ((st = ymo_su2_overrelaxation_sweep(c->extents, links))) goto done;
            nor += 1;
        }
        int64_t after = s + 1 - c->thermalization_sweeps;
        if (after > 0 && after % c->sweeps_between_measurements == 0) {
            int64_t m = after / c->sweeps_between_measurements - 1;
            if ((st = ymo_su2_average_plaquette(c->extents, links, &measurements_out[m]))) goto done;
        }
    }
    if ((st = ymo_diagnostics(measurements_out, c->measurements, c->autocorrelation_max_lag, diag, NULL))) goto done;
    rep->derived_seed = seed;
    rep->heatbath_proposals = hp;
    rep->heatbath_updates = hu;
    rep->total_heatbath_sweeps = c->thermalization_sweeps + total_meas_sweeps;
    rep->total_overrelaxation_sweeps = nor;
    rep->mean_heatbath_acceptance = (double)hu / (double)hp;
    rep->final_link_unit_residual = ymo_su2_max_unit_residual(c->extents, links);
    if (!isfinite(rep->mean_heatbath_acceptance) || !isfinite(rep->final_link_unit_residual)) st = YMO_NON_FINITE_STATISTIC;
done:
    if (!links_out) free(links);
    return st;
}

/* ------------------------------------------------ canonical config codec */
/* lib.rs:587-694 -- YM-SU2-CONFIG-V1: header, extents, link_count, then one
 * line per link with the four components as 16-digit lower-case hex bits.
 * Returns the number of bytes needed (excluding NUL) or a negative status. */
int64_t ymo_su2_config_encode(const int64_t n[4], const double *links, char *buf, int64_t cap) {
    if (ymo_su2_max_unit_residual(n, links) > SERIALIZATION_TOLERANCE) return -YMO_SU2_NON_UNIT;
    int64_t count = s2_sites(n) * 4;
    int64_t need = 0;
    char line[128];
    int k = snprintf(line, sizeof line, "YM-SU2-CONFIG-V1\nextents=%lld,%lld,%lld,%lld\nlink_count=%lld\n",
                     (long long)n[0], (long long)n[1], (long long)n[2], (long long)n[3], (long long)count);
    if (buf && need + k <= cap) memcpy(buf + need, line, k);
    need += k;
    for (int64_t i = 0; i < count; ++i) {
        uint64_t b[4];
        memcpy(b, links + 4 * i, sizeof b);
        k = snprintf(line, sizeof line, "%016llx,%016llx,%016llx,%016llx\n", (unsigned long long)b[0],
                     (unsigned long long)b[1], (unsigned long long)b[2], (unsigned long long)b[3]);
        if (buf && need + k <= cap) memcpy(buf + need, line, k);
        need += k;
    }
    return need;
}
