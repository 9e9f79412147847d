/*
 * ORACLE -- TEST INFRASTRUCTURE ONLY.
 *
 * Reference-exact CPU restatement of the SU(3) algebra, lattice, Metropolis
 * ensemble, path-holonomy, A1++ operator, spatial-stout and correlator code of
 * QSOLKCB/YANG-MILLS (Rust, MPL-2.0).  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load this library, and only as the checker
 * or the timed CPU baseline.  The product library never links it.
 *
 * Arithmetic follows the Rust source operation by operation: no FMA
 * contraction (built with -ffp-contract=off), the same left-to-right
 * summation order, glibc libm for exp/hypot (which Rust std calls on
 * linux-gnu), and the same fail-closed checks (every checked product verifies
 * unitarity and determinant residuals).  Each function cites the reference
 * file:line it restates.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "ym_oracle.h"

/* ---------------------------------------------------------------- algebra */
/* reference/rust/src/su3.rs:18-25 */
#define SU3_ACCEPTANCE_TOLERANCE 5.0e-12
#define SU3_EXP_MAX_COEFFICIENT 64.0
#define INV_SQRT_3 0.5773502691896258
#define EXP_TAYLOR_TERMS 64
#define EXP_SCALE_TARGET 0.5
#define EXP_TERM_CUTOFF 1.0e-18

typedef struct { double re, im; } cplx;
typedef struct { cplx m[3][3]; } mat3;

/* su3.rs:50-66 -- Complex64 add/sub/mul/scale (no contraction). */
static inline cplx c_add(cplx a, cplx b) { cplx r = {a.re + b.re, a.im + b.im}; return r; }
static inline cplx c_sub(cplx a, cplx b) { cplx r = {a.re - b.re, a.im - b.im}; return r; }
static inline cplx c_mul(cplx a, cplx b) {
    cplx r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
    return r;
}
static inline cplx c_scale(cplx a, double f) { cplx r = {a.re * f, a.im * f}; return r; }
static inline cplx c_conj(cplx a) { cplx r = {a.re, -a.im}; return r; }
static inline double c_abs(cplx a) { return hypot(a.re, a.im); } /* su3.rs:68-70 */
static inline int c_finite(cplx a) { return isfinite(a.re) && isfinite(a.im); }

static const cplx C_ZERO = {0.0, 0.0};
static const cplx C_ONE = {1.0, 0.0};
static const cplx C_I = {0.0, 1.0};

static mat3 m_zero(void) { mat3 r; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = C_ZERO; return r; }
static mat3 m_identity(void) { mat3 r = m_zero(); r.m[0][0] = r.m[1][1] = r.m[2][2] = C_ONE; return r; }

static int m_finite(const mat3 *a) {
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) if (!c_finite(a->m[i][j])) return 0;
    return 1;
}
static mat3 m_add(mat3 a, mat3 b) { mat3 r; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = c_add(a.m[i][j], b.m[i][j]); return r; }
static mat3 m_sub(mat3 a, mat3 b) { mat3 r; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = c_sub(a.m[i][j], b.m[i][j]); return r; }
static mat3 m_scale_real(mat3 a, double f) { mat3 r; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = c_scale(a.m[i][j], f); return r; }
static mat3 m_scale_complex(mat3 a, cplx f) { mat3 r; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = c_mul(a.m[i][j], f); return r; }

/* su3.rs:147-159 -- value starts at ZERO and adds the three products in order. */
static mat3 m_mul(mat3 a, mat3 b) {
    mat3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            cplx v = C_ZERO;
            for (int k = 0; k < 3; ++k) v = c_add(v, c_mul(a.m[i][k], b.m[k][j]));
            r.m[i][j] = v;
        }
    return r;
}
static mat3 m_dagger(mat3 a) { mat3 r; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = c_conj(a.m[j][i]); return r; }
static cplx m_trace(mat3 a) { return c_add(c_add(a.m[0][0], a.m[1][1]), a.m[2][2]); }

/* su3.rs:177-191 */
static cplx m_det(mat3 x) {
    cplx a = x.m[0][0], b = x.m[0][1], c = x.m[0][2];
    cplx d = x.m[1][0], e = x.m[1][1], f = x.m[1][2];
    cplx g = x.m[2][0], h = x.m[2][1], i = x.m[2][2];
    cplx t1 = c_mul(a, c_sub(c_mul(e, i), c_mul(f, h)));
    cplx t2 = c_mul(b, c_sub(c_mul(d, i), c_mul(f, g)));
    cplx t3 = c_mul(c, c_sub(c_mul(d, h), c_mul(e, g)));
    return c_add(c_sub(t1, t2), t3);
}

/* su3.rs:193-199 -- fold(0.0, f64::max) over row-major |entries|. */
static double m_max_abs_entry(mat3 a) {
    double m = 0.0;
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) m = fmax(m, c_abs(a.m[i][j]));
    return m;
}
/* su3.rs:201-211 */
static double m_one_norm(mat3 a) {
    double mc = 0.0;
    for (int j = 0; j < 3; ++j) {
        double s = 0.0;
        for (int i = 0; i < 3; ++i) s += c_abs(a.m[i][j]);
        mc = fmax(mc, s);
    }
    return mc;
}
static double m_max_residual(mat3 a, mat3 b) { return m_max_abs_entry(m_sub(a, b)); }
static double m_hermiticity_residual(mat3 a) { return m_max_residual(a, m_dagger(a)); }
/* su3.rs:219-223 */
static double m_unitarity_residual(mat3 a) {
    double left = m_max_residual(m_mul(m_dagger(a), a), m_identity());
    double right = m_max_residual(m_mul(a, m_dagger(a)), m_identity());
    return fmax(left, right);
}
/* su3.rs:225-229 */
static double m_determinant_residual(mat3 a) {
    double direct = c_abs(c_sub(m_det(a), C_ONE));
    double dag = c_abs(c_sub(m_det(m_dagger(a)), C_ONE));
    return fmax(direct, dag);
}

/* su3.rs:242-253 -- Su3::from_checked_matrix */
static int su3_checked(mat3 a, mat3 *out) {
    if (!m_finite(&a)) return YMO_SU3_NON_FINITE_MATRIX;
    if (m_unitarity_residual(a) > SU3_ACCEPTANCE_TOLERANCE) return YMO_SU3_NON_UNITARY;
    if (m_determinant_residual(a) > SU3_ACCEPTANCE_TOLERANCE) return YMO_SU3_DET_NOT_ONE;
    *out = a;
    return YMO_OK;
}
/* su3.rs:286-288 -- Su3::mul is a checked product. */
static int su3_mul(mat3 a, mat3 b, mat3 *out) { return su3_checked(m_mul(a, b), out); }

/* su3.rs:325-348 */
static int gell_mann(int index, mat3 *out) {
    mat3 r = m_zero();
    const cplx one = C_ONE, mone = {-1.0, 0.0}, i = C_I, mi = {0.0, -1.0};
    switch (index) {
    case 0: r.m[0][1] = one; r.m[1][0] = one; break;
    case 1: r.m[0][1] = mi; r.m[1][0] = i; break;
    case 2: r.m[0][0] = one; r.m[1][1] = mone; break;
    case 3: r.m[0][2] = one; r.m[2][0] = one; break;
    case 4: r.m[0][2] = mi; r.m[2][0] = i; break;
    case 5: r.m[1][2] = one; r.m[2][1] = one; break;
    case 6: r.m[1][2] = mi; r.m[2][1] = i; break;
    case 7: {
        cplx d = {INV_SQRT_3, 0.0}, d2 = {-2.0 * INV_SQRT_3, 0.0};
        r.m[0][0] = d; r.m[1][1] = d; r.m[2][2] = d2; break;
    }
    default: return YMO_SU3_GENERATOR_INDEX;
    }
    *out = r;
    return YMO_OK;
}
/* su3.rs:350-352 */
static int generator(int index, mat3 *out) {
    mat3 g; int st = gell_mann(index, &g);
    if (st) return st;
    *out = m_scale_real(g, 0.5);
    return YMO_OK;
}

/* su3.rs:395-407 */
static int algebra_matrix(const double coef[8], mat3 *out) {
    mat3 m = m_zero();
    for (int a = 0; a < 8; ++a) {
        if (!isfinite(coef[a])) return YMO_SU3_NON_FINITE_COEFFICIENT;
        if (fabs(coef[a]) > SU3_EXP_MAX_COEFFICIENT) return YMO_SU3_COEFFICIENT_TOO_LARGE;
        mat3 g; generator(a, &g);
        m = m_add(m, m_scale_real(g, coef[a]));
    }
    *out = m;
    return YMO_OK;
}

/* su3.rs:409-450 -- scaling and squaring with a bounded Taylor series. */
static int matrix_exponential(mat3 a, mat3 *out) {
    if (!m_finite(&a)) return YMO_SU3_NON_FINITE_MATRIX;
    double norm = m_one_norm(a);
    if (!isfinite(norm)) return YMO_SU3_NON_FINITE_MATRIX;
    int squarings = 0;
    double scaled_norm = norm;
    while (scaled_norm > EXP_SCALE_TARGET) {
        squarings += 1;
        if (squarings > 32) return YMO_SU3_EXP_SCALING_OVERFLOW;
        scaled_norm *= 0.5;
    }
    double scale = ldexp(1.0, -squarings);
    mat3 scaled = m_scale_real(a, scale);
    mat3 sum = m_identity(), term = m_identity();
    for (int order = 1; order <= EXP_TAYLOR_TERMS; ++order) {
        term = m_scale_real(m_mul(term, scaled), 1.0 / (double)order);
        sum = m_add(sum, term);
        if (m_max_abs_entry(term) <= EXP_TERM_CUTOFF) break;
    }
    for (int s = 0; s < squarings; ++s) sum = m_mul(sum, sum);
    if (!m_finite(&sum)) return YMO_SU3_NON_FINITE_MATRIX;
    *out = sum;
    return YMO_OK;
}

/* su3.rs:255-262 -- Su3::from_algebra_coefficients */
static int su3_from_algebra(const double coef[8], mat3 *out) {
    mat3 h; int st = algebra_matrix(coef, &h);
    if (st) return st;
    mat3 ah = m_scale_complex(h, C_I), e;
    st = matrix_exponential(ah, &e);
    if (st) return st;
    return su3_checked(e, out);
}

/* ------------------------------------------------------- marshalling helpers */
static mat3 load_m(const double *p) {
    mat3 r;
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) { r.m[i][j].re = p[2 * (3 * i + j)]; r.m[i][j].im = p[2 * (3 * i + j) + 1]; }
    return r;
}
static void store_m(double *p, mat3 a) {
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) { p[2 * (3 * i + j)] = a.m[i][j].re; p[2 * (3 * i + j) + 1] = a.m[i][j].im; }
}

int ymo_su3_from_algebra(const double coef[8], double *out18) {
    mat3 r; int st = su3_from_algebra(coef, &r);
    if (st) return st;
    store_m(out18, r);
    return YMO_OK;
}
int ymo_su3_gell_mann(int index, double *out18) {
    mat3 r; int st = gell_mann(index, &r);
    if (st) return st;
    store_m(out18, r);
    return YMO_OK;
}
int ymo_su3_checked_mul(const double *a18, const double *b18, double *out18) {
    mat3 r; int st = su3_mul(load_m(a18), load_m(b18), &r);
    if (st) return st;
    store_m(out18, r);
    return YMO_OK;
}
void ymo_su3_matrix_residuals(const double *a18, double *unit, double *det) {
    mat3 a = load_m(a18);
    *unit = m_unitarity_residual(a);
    *det = m_determinant_residual(a);
}
/* su3.rs:354-368 -- structure constants via commutators / anticommutators. */
int ymo_su3_structure_f(int a, int b, int c, double *out) {
    mat3 ta, tb, tc;
    if (generator(a, &ta) || generator(b, &tb) || generator(c, &tc)) return YMO_SU3_GENERATOR_INDEX;
    mat3 comm = m_sub(m_mul(ta, tb), m_mul(tb, ta));
    *out = 2.0 * m_trace(m_mul(comm, tc)).im;
    return YMO_OK;
}
int ymo_su3_structure_d(int a, int b, int c, double *out) {
    mat3 ta, tb, tc;
    if (generator(a, &ta) || generator(b, &tb) || generator(c, &tc)) return YMO_SU3_GENERATOR_INDEX;
    mat3 anti = m_add(m_mul(ta, tb), m_mul(tb, ta));
    *out = 2.0 * m_trace(m_mul(anti, tc)).re;
    return YMO_OK;
}
/* su3.rs:370-393 */
void ymo_su3_casimirs(double *quad18, double *cubic18) {
    mat3 q = m_zero();
    for (int a = 0; a < 8; ++a) { mat3 ta; generator(a, &ta); q = m_add(q, m_mul(ta, ta)); }
    store_m(quad18, q);
    mat3 s = m_zero();
    for (int a = 0; a < 8; ++a) {
        mat3 ta; generator(a, &ta);
        for (int b = 0; b < 8; ++b) {
            mat3 tb; generator(b, &tb);
            for (int c = 0; c < 8; ++c) {
                mat3 tc; generator(c, &tc);
                double d; ymo_su3_structure_d(a, b, c, &d);
                s = m_add(s, m_scale_real(m_mul(m_mul(ta, tb), tc), d));
            }
        }
    }
    store_m(cubic18, s);
}

/* ------------------------------------------------------------- SU(3) lattice */
/* su3_lattice.rs:16-17: the reference caps lattices at 2^18 links; the oracle
 * lifts the cap to a caller flag (YMO_NO_CAP) so larger synthetic lattices
 * can be checked, and reports LatticeTooLarge otherwise. */
#define SU3_MAX_LINKS (1u << 18)

typedef struct {
    int64_t n[4];
    double *links; /* ((site*4+mu)*18 + c), reference AoS, x3 fastest */
} lat3;

static inline int64_t site_index(const lat3 *l, const int64_t x[4]) {
    return ((x[0] * l->n[1] + x[1]) * l->n[2] + x[2]) * l->n[3] + x[3];
}
static inline mat3 link_get(const lat3 *l, const int64_t x[4], int mu) {
    return load_m(l->links + (site_index(l, x) * 4 + mu) * 18);
}
static inline void link_set(lat3 *l, const int64_t x[4], int mu, mat3 v) {
    store_m(l->links + (site_index(l, x) * 4 + mu) * 18, v);
}
static inline void nb_plus(const lat3 *l, const int64_t x[4], int mu, int64_t y[4]) {
    memcpy(y, x, 4 * sizeof(int64_t));
    y[mu] = (y[mu] + 1) % l->n[mu];
}
static inline void nb_minus(const lat3 *l, const int64_t x[4], int mu, int64_t y[4]) {
    memcpy(y, x, 4 * sizeof(int64_t));
    y[mu] = y[mu] == 0 ? l->n[mu] - 1 : y[mu] - 1;
}
static int64_t lat_sites(const int64_t n[4]) { return n[0] * n[1] * n[2] * n[3]; }

static int validate_extents(const int64_t n[4], int cap) {
    for (int d = 0; d < 4; ++d) if (n[d] < 2) return YMO_INVALID_EXTENT;
    if (cap && lat_sites(n) * 4 > (int64_t)SU3_MAX_LINKS) return YMO_LATTICE_TOO_LARGE;
    return YMO_OK;
}

/* su3_lattice.rs:85-106 */
int ymo_su3_identity(const int64_t n[4], double *links) {
    int st = validate_extents(n, 0);
    if (st) return st;
    int64_t count = lat_sites(n) * 4;
    mat3 id = m_identity();
    for (int64_t i = 0; i < count; ++i) store_m(links + i * 18, id);
    return YMO_OK;
}

/* su3_lattice.rs:168-181 -- checked left-to-right product. */
static int su3_plaquette(const lat3 *l, const int64_t x[4], int mu, int nu, mat3 *out) {
    int64_t xm[4], xn[4];
    nb_plus(l, x, mu, xm);
    nb_plus(l, x, nu, xn);
    mat3 p; int st;
    if ((st = su3_mul(link_get(l, x, mu), link_get(l, xm, nu), &p))) return st;
    if ((st = su3_mul(p, m_dagger(link_get(l, xn, mu)), &p))) return st;
    if ((st = su3_mul(p, m_dagger(link_get(l, x, nu)), &p))) return st;
    *out = p;
    return YMO_OK;
}
static double normalized_real_trace(mat3 a) { return m_trace(a).re / 3.0; }

#define FOR_SITES(l, x)                                                   \
    for (x[0] = 0; x[0] < (l)->n[0]; ++x[0])                              \
        for (x[1] = 0; x[1] < (l)->n[1]; ++x[1])                          \
            for (x[2] = 0; x[2] < (l)->n[2]; ++x[2])                      \
                for (x[3] = 0; x[3] < (l)->n[3]; ++x[3])

/* su3_lattice.rs:214-242 */
int ymo_su3_wilson_action(const int64_t n[4], const double *links, double beta, double *out) {
    if (!isfinite(beta)) return YMO_NON_FINITE_BETA;
    if (beta < 0.0) return YMO_NEGATIVE_BETA;
    lat3 l = {{n[0], n[1], n[2], n[3]}, (double *)links};
    double sum = 0.0;
    int64_t x[4];
    FOR_SITES(&l, x) {
        for (int mu = 0; mu < 4; ++mu)
            for (int nu = mu + 1; nu < 4; ++nu) {
                mat3 p; int st = su3_plaquette(&l, x, mu, nu, &p);
                if (st) return st;
                sum += 1.0 - normalized_real_trace(p);
                if (!isfinite(sum)) return YMO_NON_FINITE_ACTION;
            }
    }
    double action = beta * sum;
    if (!isfinite(action)) return YMO_NON_FINITE_ACTION;
    *out = action;
    return YMO_OK;
}

/* su3_ensemble.rs:228-259 */
int ymo_su3_average_plaquette(const int64_t n[4], const double *links, double *out) {
    lat3 l = {{n[0], n[1], n[2], n[3]}, (double *)links};
    double sum = 0.0;
    int64_t count = 0, x[4];
    FOR_SITES(&l, x) {
        for (int mu = 0; mu < 4; ++mu)
            for (int nu = mu + 1; nu < 4; ++nu) {
                mat3 p; int st = su3_plaquette(&l, x, mu, nu, &p);
                if (st) return st;
                double v = normalized_real_trace(p);
                if (!isfinite(v)) return YMO_NON_FINITE_STATISTIC;
                sum += v;
                count += 1;
            }
    }
    if (count == 0) return YMO_NON_FINITE_STATISTIC;
    double mean = sum / (double)count;
    if (!isfinite(mean)) return YMO_NON_FINITE_STATISTIC;
    *out = mean;
    return YMO_OK;
}

/* su3_lattice.rs:185-212 -- ordered product along one extent, (1/3) Tr. */
int ymo_su3_polyakov_loop(const int64_t n[4], const double *links, const int64_t start[4], int dir,
                          double *re, double *im) {
    if (dir < 0 || dir > 3) return YMO_DIRECTION_OUT_OF_RANGE;
    lat3 l = {{n[0], n[1], n[2], n[3]}, (double *)links};
    int64_t x[4];
    memcpy(x, start, sizeof x);
    mat3 h = m_identity();
    for (int64_t s = 0; s < n[dir]; ++s) {
        int st = su3_mul(h, link_get(&l, x, dir), &h);
        if (st) return st;
        int64_t y[4]; nb_plus(&l, x, dir, y); memcpy(x, y, sizeof x);
    }
    cplx t = c_scale(m_trace(h), 1.0 / 3.0);
    *re = t.re; *im = t.im;
    return YMO_OK;
}

/* su3_lattice.rs:244-283 -- U'_mu(x) = g(x) U_mu(x) g(x+mu)^dagger (checked). */
int ymo_su3_gauge_transform(const int64_t n[4], const double *links, const double *gauge_sites, double *out) {
    lat3 l = {{n[0], n[1], n[2], n[3]}, (double *)links};
    lat3 o = {{n[0], n[1], n[2], n[3]}, out};
    int64_t x[4];
    FOR_SITES(&l, x) {
        mat3 gx = load_m(gauge_sites + site_index(&l, x) * 18);
        for (int mu = 0; mu < 4; ++mu) {
            int64_t xp[4]; nb_plus(&l, x, mu, xp);
            mat3 gxp = load_m(gauge_sites + site_index(&l, xp) * 18), v;
            int st;
            if ((st = su3_mul(gx, link_get(&l, x, mu), &v))) return st;
            if ((st = su3_mul(v, m_dagger(gxp), &v))) return st;
            link_set(&o, x, mu, v);
        }
    }
    return YMO_OK;
}

/* su3_lattice.rs:285-297 */
void ymo_su3_max_residuals(const int64_t n[4], const double *links, double *unit, double *det) {
    int64_t count = lat_sites(n) * 4;
    double u = 0.0, d = 0.0;
    for (int64_t i = 0; i < count; ++i) u = fmax(u, m_unitarity_residual(load_m(links + i * 18)));
    for (int64_t i = 0; i < count; ++i) d = fmax(d, m_determinant_residual(load_m(links + i * 18)));
    *unit = u;
    *det = d;
}

/* su3_lattice.rs:318-340 -- fixtures */
void ymo_su3_traceless_cycle_link(double *out18) {
    mat3 r = m_zero();
    r.m[0][1] = C_ONE; r.m[1][2] = C_ONE; r.m[2][0] = C_ONE;
    store_m(out18, r);
}
int ymo_su3_deterministic_site_gauge(const int64_t x[4], double *out18) {
    int64_t red[4];
    for (int d = 0; d < 4; ++d) red[d] = x[d] % 1024;
    int64_t code = (red[0] + 3 * red[1] + 7 * red[2] + 11 * red[3]) % 1024;
    double coef[8] = {0};
    coef[2] = 0.031 * ((double)code + 1.0);
    return ymo_su3_from_algebra(coef, out18);
}

/* ------------------------------------------------------------ SU(3) ensemble */
/* su3_ensemble.rs:19-33 */
#define SU3_STREAM_DOMAIN 0x535533454e535631ull
#define SU3_MAX_BETA 12.0
#define SU3_MAX_SWEEPS 512
#define SU3_MAX_MEASUREMENTS 512
#define SU3_MIN_ANGLE 1.0e-8
#define SU3_MAX_ANGLE 1.0
#define SU3_PLAQUETTE_TOLERANCE 1.0e-12
#define SU3_LINK_RESIDUAL_LIMIT 5.0e-12
#define ANGLE_GRID_MASK ((1ull << 52) - 1ull)
#define ANGLE_GRID_SCALE (1.0 / 9007199254740992.0)

/* su2_ensemble.rs:46-50 / su3_ensemble.rs:183-187 */
uint64_t ymo_mix64(uint64_t v) {
    v = (v ^ (v >> 30)) * 0xBF58476D1CE4E5B9ull;
    v = (v ^ (v >> 27)) * 0x94D049BB133111EBull;
    return v ^ (v >> 31);
}
/* su2_ensemble.rs:34-43 */
uint64_t ymo_splitmix_next(uint64_t *state) {
    *state += 0x9E3779B97F4A7C15ull;
    return ymo_mix64(*state);
}
double ymo_splitmix_unit(uint64_t *state) {
    return (double)(ymo_splitmix_next(state) >> 11) * (1.0 / 9007199254740992.0);
}
/* su2_ensemble.rs:52-55 */
uint64_t ymo_derive_chain_seed(uint64_t master, uint64_t chain) {
    uint64_t lane = (chain + 1ull) * 0x9E3779B97F4A7C15ull;
    return ymo_mix64(master ^ lane);
}
/* su3_ensemble.rs:189-192 */
uint64_t ymo_derive_su3_chain_seed(uint64_t master, uint64_t chain) {
    uint64_t lane = (chain + 1ull) * 0x9E3779B97F4A7C15ull;
    return ymo_mix64(master ^ SU3_STREAM_DOMAIN ^ lane);
}

static int validate_angle(double a) {
    if (!isfinite(a) || a < SU3_MIN_ANGLE || a > SU3_MAX_ANGLE) return YMO_ANGLE_OUT_OF_RANGE;
    return YMO_OK;
}

/* su3_ensemble.rs:199-213 */
int ymo_proposal_angle_from_u64(uint64_t raw, double proposal_angle, double *out) {
    int st = validate_angle(proposal_angle);
    if (st) return st;
    uint64_t m = raw & ANGLE_GRID_MASK;
    uint64_t odd = m * 2ull + 1ull;
    double magnitude = proposal_angle * (double)odd * ANGLE_GRID_SCALE;
    double sign = (raw >> 63) == 0 ? -1.0 : 1.0;
    double angle = sign * magnitude;
    if (!isfinite(angle) || angle == 0.0 || fabs(angle) >= proposal_angle) return YMO_NON_FINITE_STATISTIC;
    *out = angle;
    return YMO_OK;
}

/* su3_ensemble.rs:261-307 -- six plaquettes with the link replaced. */
static int local_link_action_with_value(const lat3 *l, const int64_t x[4], int mu, mat3 w, double beta, double *out) {
    if (!isfinite(beta) || beta < 0.0 || beta > SU3_MAX_BETA) return YMO_BETA_OUT_OF_RANGE;
    double sum = 0.0;
    int st;
    for (int nu = 0; nu < 4; ++nu) {
        if (nu == mu) continue;
        int64_t xmu[4], xnu[4], bx[4], bxmu[4];
        nb_plus(l, x, mu, xmu);
        nb_plus(l, x, nu, xnu);
        mat3 f;
        if ((st = su3_mul(w, link_get(l, xmu, nu), &f))) return st;
        if ((st = su3_mul(f, m_dagger(link_get(l, xnu, mu)), &f))) return st;
        if ((st = su3_mul(f, m_dagger(link_get(l, x, nu)), &f))) return st;
        nb_minus(l, x, nu, bx);
        nb_plus(l, bx, mu, bxmu);
        mat3 b;
        if ((st = su3_mul(link_get(l, bx, mu), link_get(l, bxmu, nu), &b))) return st;
        if ((st = su3_mul(b, m_dagger(w), &b))) return st;
        if ((st = su3_mul(b, m_dagger(link_get(l, bx, nu)), &b))) return st;
        sum += 1.0 - normalized_real_trace(f);
        sum += 1.0 - normalized_real_trace(b);
        if (!isfinite(sum)) return YMO_NON_FINITE_STATISTIC;
    }
    double action = beta * sum;
    if (!isfinite(action)) return YMO_NON_FINITE_STATISTIC;
    *out = action;
    return YMO_OK;
}

/* su3_ensemble.rs:319-334 */
static int local_action_difference(const lat3 *l, const int64_t x[4], int mu, mat3 cand, double beta, double *out) {
    double old_a, new_a;
    int st;
    if ((st = local_link_action_with_value(l, x, mu, link_get(l, x, mu), beta, &old_a))) return st;
    if ((st = local_link_action_with_value(l, x, mu, cand, beta, &new_a))) return st;
    double d = new_a - old_a;
    if (!isfinite(d)) return YMO_NON_FINITE_STATISTIC;
    *out = d;
    return YMO_OK;
}

int ymo_su3_local_link_action(const int64_t n[4], const double *links, const int64_t x[4], int mu, double beta, double *out) {
    lat3 l = {{n[0], n[1], n[2], n[3]}, (double *)links};
    return local_link_action_with_value(&l, x, mu, link_get(&l, x, mu), beta, out);
}
int ymo_su3_local_action_difference(const int64_t n[4], const double *links, const int64_t x[4], int mu,
                                    const double *cand18, double beta, double *out) {
    lat3 l = {{n[0], n[1], n[2], n[3]}, (double *)links};
    return local_action_difference(&l, x, mu, load_m(cand18), beta, out);
}

/* su3_ensemble.rs:336-340 */
static int generator_proposal(int a, double angle, mat3 *out) {
    double coef[8] = {0};
    coef[a] = angle;
    return su3_from_algebra(coef, out);
}

/* su3_ensemble.rs:342-391 -- lexicographic sites, mu, then the 8 generators;
 * one next_u64 for the angle and one unit_f64 for the acceptance draw. */
int ymo_su3_metropolis_sweep(const int64_t n[4], double *links, double beta, double angle, uint64_t *rng,
                             uint64_t *proposals, uint64_t *accepts) {
    if (!isfinite(beta) || beta < 0.0 || beta > SU3_MAX_BETA) return YMO_BETA_OUT_OF_RANGE;
    int st = validate_angle(angle);
    if (st) return st;
    lat3 l = {{n[0], n[1], n[2], n[3]}, links};
    uint64_t prop = 0, acc = 0;
    int64_t x[4];
    FOR_SITES(&l, x) {
        for (int mu = 0; mu < 4; ++mu)
            for (int a = 0; a < 8; ++a) {
                uint64_t raw = ymo_splitmix_next(rng);
                double th;
                if ((st = ymo_proposal_angle_from_u64(raw, angle, &th))) return st;
                mat3 rot, cand;
                if ((st = generator_proposal(a, th, &rot))) return st;
                mat3 old = link_get(&l, x, mu);
                if ((st = su3_mul(rot, old, &cand))) return st;
                double delta;
                if ((st = local_action_difference(&l, x, mu, cand, beta, &delta))) return st;
                double acceptance = exp(-fmax(delta, 0.0));
                if (!isfinite(acceptance) || acceptance < 0.0 || acceptance > 1.0) return YMO_NON_FINITE_STATISTIC;
                double u = ymo_splitmix_unit(rng);
                prop += 1;
                if (u < acceptance) {
                    link_set(&l, x, mu, cand);
                    acc += 1;
                }
            }
    }
    *proposals = prop;
    *accepts = acc;
    return YMO_OK;
}

/* su3_ensemble.rs:58-97 -- Su3EnsembleConfig::validate (caps optional). */
static int su3_validate_config(const ymo_su3_chain_config *c) {
    if (!isfinite(c->beta) || c->beta < 0.0 || c->beta > SU3_MAX_BETA) return YMO_BETA_OUT_OF_RANGE;
    int st = validate_angle(c->proposal_angle);
    if (st) return st;
    if (c->measurements < 2 || (!c->no_caps && c->measurements > SU3_MAX_MEASUREMENTS)) return YMO_MEASUREMENT_COUNT;
    if (c->sweeps_between_measurements == 0) return YMO_ZERO_STRIDE;
    if (c->autocorrelation_max_lag >= c->measurements) return YMO_LAG_OUT_OF_RANGE;
    int64_t total = c->thermalization_sweeps + c->measurements * c->sweeps_between_measurements;
    if (!c->no_caps && total > SU3_MAX_SWEEPS) return YMO_SCHEDULE_OUT_OF_RANGE;
    for (int d = 0; d < 4; ++d) if (c->extents[d] < 2) return YMO_LATTICE_OUT_OF_RANGE;
    if (!c->no_caps && lat_sites(c->extents) * 4 > (int64_t)SU3_MAX_LINKS) return YMO_LATTICE_OUT_OF_RANGE;
    return YMO_OK;
}

/* su3_ensemble.rs:408-480 -- cold start, therm + measurements x stride,
 * per-measurement plaquette with range check, diagnostics, residual gates.
 * `measurements_out` holds c->measurements doubles; `links_out` (optional)
 * receives the final configuration. */
int ymo_su3_run_chain(const ymo_su3_chain_config *c, double *measurements_out, ymo_chain_diagnostics *diag,
                      ymo_su3_chain_report *rep, double *links_out) {
    int st = su3_validate_config(c);
    if (st) return st;
    uint64_t seed = ymo_derive_su3_chain_seed(c->master_seed, c->chain_id);
    uint64_t rng = seed;
    int64_t count = lat_sites(c->extents) * 4;
    double *links = links_out ? links_out : (double *)malloc(sizeof(double) * 18 * count);
    if (!links) return YMO_ALLOCATION_FAILED;
    ymo_su3_identity(c->extents, links);
    uint64_t prop = 0, acc = 0, p, a;
    for (int64_t s = 0; s < c->thermalization_sweeps; ++s) {
        if ((st = ymo_su3_metropolis_sweep(c->extents, links, c->beta, c->proposal_angle, &rng, &p, &a))) goto done;
        prop += p; acc += a;
    }
    for (int64_t m = 0; m < c->measurements; ++m) {
        for (int64_t s = 0; s < c->sweeps_between_measurements; ++s) {
            if ((st = ymo_su3_metropolis_sweep(c->extents, links, c->beta, c->proposal_angle, &rng, &p, &a))) goto done;
            prop += p; acc += a;
        }
        double v;
        if ((st = ymo_su3_average_plaquette(c->extents, links, &v))) goto done;
        if (v < -0.5 - SU3_PLAQUETTE_TOLERANCE || v > 1.0 + SU3_PLAQUETTE_TOLERANCE) { st = YMO_NON_FINITE_STATISTIC; goto done; }
        measurements_out[m] = v;
    }
    if ((st = ymo_diagnostics(measurements_out, c->measurements, c->autocorrelation_max_lag, diag, NULL))) { st = YMO_DIAGNOSTIC; goto done; }
    if (prop == 0) { st = YMO_NON_FINITE_STATISTIC; goto done; }
    rep->derived_seed = seed;
    rep->total_sweeps = c->thermalization_sweeps + c->measurements * c->sweeps_between_measurements;
    rep->proposals = prop;
    rep->accepts = acc;
    rep->mean_acceptance = (double)acc / (double)prop;
    ymo_su3_max_residuals(c->extents, links, &rep->final_unitarity_residual, &rep->final_determinant_residual);
    if (!isfinite(rep->final_unitarity_residual) || !isfinite(rep->final_determinant_residual) ||
        rep->final_unitarity_residual > SU3_LINK_RESIDUAL_LIMIT || rep->final_determinant_residual > SU3_LINK_RESIDUAL_LIMIT)
        st = YMO_NON_FINITE_STATISTIC;
done:
    if (!links_out) free(links);
    return st;
}

/* ------------------------------------------------------------ path holonomy */
/* path_holonomy.rs:132-157 -- steps encoded +(d+1) forward, -(d+1) backward. */
int ymo_su3_path_holonomy(const int64_t n[4], const double *links, const int64_t start[4], const int32_t *steps,
                          int64_t nsteps, double *out18, int64_t end[4]) {
    if (nsteps <= 0) return YMO_EMPTY_PATH;
    if (nsteps > (1 << 20)) return YMO_PATH_TOO_LONG;
    lat3 l = {{n[0], n[1], n[2], n[3]}, (double *)links};
    int64_t cur[4];
    memcpy(cur, start, sizeof cur);
    mat3 h = m_identity();
    for (int64_t s = 0; s < nsteps; ++s) {
        int code = steps[s];
        int d = (code > 0 ? code : -code) - 1;
        if (code == 0 || d > 3) return YMO_DIRECTION_OUT_OF_RANGE;
        mat3 seg;
        int64_t y[4];
        if (code > 0) {
            seg = link_get(&l, cur, d);
            nb_plus(&l, cur, d, y);
        } else {
            nb_minus(&l, cur, d, y);
            seg = m_dagger(link_get(&l, y, d));
        }
        memcpy(cur, y, sizeof cur);
        int st = su3_mul(h, seg, &h);
        if (st) return st;
    }
    store_m(out18, h);
    if (end) memcpy(end, cur, sizeof cur);
    return YMO_OK;
}

/* path_holonomy.rs:219-260 followed by wilson_loop (172-181). */
static int rect_steps(int mu, int nu, int64_t r, int64_t t, int32_t *buf) {
    int64_t k = 0;
    for (int64_t i = 0; i < r; ++i) buf[k++] = mu + 1;
    for (int64_t i = 0; i < t; ++i) buf[k++] = nu + 1;
    for (int64_t i = 0; i < r; ++i) buf[k++] = -(mu + 1);
    for (int64_t i = 0; i < t; ++i) buf[k++] = -(nu + 1);
    return (int)k;
}
int ymo_su3_rect_wilson_loop(const int64_t n[4], const double *links, const int64_t start[4], int mu, int nu,
                             int64_t r, int64_t t, double *re, double *im) {
    if (mu < 0 || mu > 3 || nu < 0 || nu > 3) return YMO_DIRECTION_OUT_OF_RANGE;
    if (mu == nu) return YMO_DIRECTIONS_MUST_DIFFER;
    if (r <= 0 || t <= 0) return YMO_ZERO_RECTANGLE_EXTENT;
    if (2 * (r + t) > (1 << 20)) return YMO_PATH_TOO_LONG;
    int32_t *buf = (int32_t *)malloc(sizeof(int32_t) * 2 * (r + t));
    if (!buf) return YMO_ALLOCATION_FAILED;
    int k = rect_steps(mu, nu, r, t, buf);
    double h[18];
    int64_t end[4];
    int st = ymo_su3_path_holonomy(n, links, start, buf, k, h, end);
    free(buf);
    if (st) return st;
    for (int d = 0; d < 4; ++d) if (end[d] != start[d]) return YMO_PATH_NOT_CLOSED;
    cplx tr = c_scale(m_trace(load_m(h)), 1.0 / 3.0);
    *re = tr.re; *im = tr.im;
    return YMO_OK;
}

/* -------------------------------------------------------- A1++ operators */
/* glueball_operators.rs:76-91 */
static int a1pp_validate(const int64_t n[4], int64_t t) {
    if (t < 0 || t >= n[3]) return YMO_TIME_SLICE_OUT_OF_RANGE;
    for (int d = 0; d < 3; ++d) if (n[d] < 3) return YMO_SPATIAL_EXTENT_TOO_SMALL;
    if (n[0] != n[1] || n[1] != n[2]) return YMO_SPATIAL_NOT_CUBIC;
    return YMO_OK;
}
/* glueball_operators.rs:93-139 */
static int average_orbit(const int64_t n[4], const double *links, int64_t t, const int (*orient)[2], int norient,
                         int64_t e1, int64_t e2, double *mean, int64_t *count_out) {
    double sum = 0.0;
    int64_t count = 0;
    for (int o = 0; o < norient; ++o) {
        for (int64_t x0 = 0; x0 < n[0]; ++x0)
            for (int64_t x1 = 0; x1 < n[1]; ++x1)
                for (int64_t x2 = 0; x2 < n[2]; ++x2) {
                    int64_t s[4] = {x0, x1, x2, t};
                    double re, im;
                    int st = ymo_su3_rect_wilson_loop(n, links, s, orient[o][0], orient[o][1], e1, e2, &re, &im);
                    if (st) return st;
                    if (!isfinite(re)) return YMO_NON_FINITE_OBSERVABLE;
                    sum += re;
                    if (!isfinite(sum)) return YMO_NON_FINITE_OBSERVABLE;
                    count += 1;
                }
    }
    if (count == 0) return YMO_COUNT_OVERFLOW;
    *mean = sum / (double)count;
    if (!isfinite(*mean)) return YMO_NON_FINITE_OBSERVABLE;
    *count_out = count;
    return YMO_OK;
}
/* glueball_operators.rs:24-26, 143-160 */
int ymo_su3_a1pp_basis(const int64_t n[4], const double *links, int64_t t, double *plaq, double *rect,
                       int64_t *plaq_count, int64_t *rect_count) {
    static const int po[3][2] = {{0, 1}, {0, 2}, {1, 2}};
    static const int ro[6][2] = {{0, 1}, {0, 2}, {1, 0}, {1, 2}, {2, 0}, {2, 1}};
    int st = a1pp_validate(n, t);
    if (st) return st;
    if ((st = average_orbit(n, links, t, po, 3, 1, 1, plaq, plaq_count))) return st;
    return average_orbit(n, links, t, ro, 6, 2, 1, rect, rect_count);
}

/* ------------------------------------------------------- spatial stout smear */
/* glueball_smearing.rs:21-31 */
#define STOUT_MAX_RHO 0.25
#define STOUT_MAX_LEVELS 8
#define ALGEBRA_RESIDUAL_TOLERANCE 2.0e-12

/* glueball_smearing.rs:137-172 */
static int spatial_staple_sum(const lat3 *l, const int64_t x[4], int mu, double rho, mat3 *out) {
    mat3 sum = m_zero();
    int st;
    for (int nu = 0; nu < 3; ++nu) {
        if (nu == mu) continue;
        int64_t xpm[4], xpn[4], xmn[4], xmnpm[4];
        nb_plus(l, x, mu, xpm);
        nb_plus(l, x, nu, xpn);
        mat3 f;
        if ((st = su3_mul(link_get(l, x, nu), link_get(l, xpn, mu), &f))) return st;
        if ((st = su3_mul(f, m_dagger(link_get(l, xpm, nu)), &f))) return st;
        sum = m_add(sum, f);
        nb_minus(l, x, nu, xmn);
        nb_plus(l, xmn, mu, xmnpm);
        mat3 b;
        if ((st = su3_mul(m_dagger(link_get(l, xmn, nu)), link_get(l, xmn, mu), &b))) return st;
        if ((st = su3_mul(b, link_get(l, xmnpm, nu), &b))) return st;
        sum = m_add(sum, b);
    }
    mat3 w = m_scale_real(sum, rho);
    if (!m_finite(&w)) return YMO_NON_FINITE_STAPLE;
    *out = w;
    return YMO_OK;
}
/* glueball_smearing.rs:174-193 */
static int stout_q(mat3 staple, mat3 link, mat3 *out) {
    mat3 omega = m_mul(staple, m_dagger(link));
    if (!m_finite(&omega)) return YMO_NON_FINITE_STAPLE;
    mat3 anti = m_sub(m_dagger(omega), omega);
    cplx half_i = {0.0, 0.5};
    mat3 first = m_scale_complex(anti, half_i);
    cplx tc = c_scale(c_mul(m_trace(anti), C_I), 1.0 / 6.0);
    mat3 q = m_sub(first, m_scale_complex(m_identity(), tc));
    if (!m_finite(&q)) return YMO_NON_FINITE_STAPLE;
    if (m_hermiticity_residual(q) > ALGEBRA_RESIDUAL_TOLERANCE || c_abs(m_trace(q)) > ALGEBRA_RESIDUAL_TOLERANCE)
        return YMO_ALGEBRA_PROJECTION;
    *out = q;
    return YMO_OK;
}
/* glueball_smearing.rs:195-205 */
static int q_coefficients(mat3 q, double coef[8]) {
    for (int a = 0; a < 8; ++a) {
        mat3 g; gell_mann(a, &g);
        cplx p = m_trace(m_mul(q, g));
        if (!c_finite(p) || fabs(p.im) > ALGEBRA_RESIDUAL_TOLERANCE) return YMO_ALGEBRA_PROJECTION;
        coef[a] = p.re;
    }
    return YMO_OK;
}
/* glueball_smearing.rs:225-251 -- synchronous: reads level n, writes n+1. */
static int smear_one_level(const lat3 *in, lat3 *out, double rho) {
    int64_t count = lat_sites(in->n) * 4;
    memcpy(out->links, in->links, sizeof(double) * 18 * count);
    int64_t x[4];
    int st;
    FOR_SITES(in, x) {
        for (int mu = 0; mu < 3; ++mu) {
            mat3 old = link_get(in, x, mu), staple, q, rot, upd;
            double coef[8];
            if ((st = spatial_staple_sum(in, x, mu, rho, &staple))) return st;
            if ((st = stout_q(staple, old, &q))) return st;
            if ((st = q_coefficients(q, coef))) return st;
            if ((st = su3_from_algebra(coef, &rot))) return st;
            if ((st = su3_mul(rot, old, &upd))) return st;
            if (m_unitarity_residual(upd) > SU3_ACCEPTANCE_TOLERANCE || m_determinant_residual(upd) > SU3_ACCEPTANCE_TOLERANCE)
                return YMO_SU3_NON_UNITARY;
            link_set(out, x, mu, upd);
        }
    }
    return YMO_OK;
}
/* glueball_smearing.rs:87-113, 257-268.  `max_levels` is the level cap
 * (the reference's is 8; callers may lift it for longer smearing). */
int ymo_su3_stout_smear(const int64_t n[4], const double *in, double *out, double rho, int64_t levels, int64_t max_levels) {
    if (!isfinite(rho)) return YMO_NON_FINITE_RHO;
    if (rho <= 0.0 || rho > STOUT_MAX_RHO) return YMO_RHO_OUT_OF_RANGE;
    if (levels < 0 || levels > (max_levels > 0 ? max_levels : STOUT_MAX_LEVELS)) return YMO_TOO_MANY_LEVELS;
    for (int d = 0; d < 3; ++d) if (n[d] < 3) return YMO_SPATIAL_EXTENT_TOO_SMALL;
    if (n[0] != n[1] || n[1] != n[2]) return YMO_SPATIAL_NOT_CUBIC;
    int64_t count = lat_sites(n) * 4;
    memcpy(out, in, sizeof(double) * 18 * count);
    if (levels == 0) return YMO_OK;
    double *tmp = (double *)malloc(sizeof(double) * 18 * count);
    if (!tmp) return YMO_ALLOCATION_FAILED;
    int st = YMO_OK;
    for (int64_t lv = 0; lv < levels; ++lv) {
        memcpy(tmp, out, sizeof(double) * 18 * count);
        lat3 a = {{n[0], n[1], n[2], n[3]}, tmp};
        lat3 b = {{n[0], n[1], n[2], n[3]}, out};
        if ((st = smear_one_level(&a, &b, rho))) break;
    }
    free(tmp);
    return st;
}
/* ----------------------------------------------- spatial stride-2 blocking */
/* glueball_blocking.rs:71-94 -- coarse extents; checks in the reference's order. */
int ymo_su3_blocked_extents(const int64_t n[4], int64_t coarse[4]) {
    if (n[0] != n[1] || n[1] != n[2]) return YMO_SPATIAL_NOT_CUBIC;
    if (n[0] < 6) return YMO_SPATIAL_EXTENT_TOO_SMALL;
    if (n[0] % 2 != 0) return YMO_SPATIAL_EXTENT_NOT_EVEN;
    for (int d = 0; d < 3; ++d) coarse[d] = n[d] / 2;
    coarse[3] = n[3];
    return YMO_OK;
}
/* glueball_blocking.rs:105-136 -- V_mu(y) = U_mu(x) U_mu(x+mu) (checked
 * product, then the 5e-12 group check), V_3(y) = U_3(x), x = (2y0,2y1,2y2,y3). */
int ymo_su3_spatial_block(const int64_t n[4], const double *in, double *out) {
    int64_t c[4];
    int st = ymo_su3_blocked_extents(n, c);
    if (st) return st;
    lat3 f = {{n[0], n[1], n[2], n[3]}, (double *)in};
    lat3 g = {{c[0], c[1], c[2], c[3]}, out};
    int64_t y[4];
    FOR_SITES(&g, y) {
        const int64_t x[4] = {2 * y[0], 2 * y[1], 2 * y[2], y[3]};
        for (int mu = 0; mu < 3; ++mu) {
            int64_t xp[4];
            nb_plus(&f, x, mu, xp);
            mat3 v;
            if ((st = su3_mul(link_get(&f, x, mu), link_get(&f, xp, mu), &v))) return st;
            if (m_unitarity_residual(v) > SU3_ACCEPTANCE_TOLERANCE || m_determinant_residual(v) > SU3_ACCEPTANCE_TOLERANCE)
                return YMO_SU3_NON_UNITARY;
            link_set(&g, y, mu, v);
        }
        link_set(&g, y, 3, link_get(&f, x, 3));
    }
    return YMO_OK;
}
/* glueball_smearing.rs:270-300 */
double ymo_su3_max_link_residual(const int64_t n[4], const double *a, const double *b) {
    int64_t count = lat_sites(n) * 4;
    double m = 0.0;
    for (int64_t i = 0; i < count; ++i) m = fmax(m, m_max_residual(load_m(a + i * 18), load_m(b + i * 18)));
    return m;
}
