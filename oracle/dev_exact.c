/*
 * ORACLE -- TEST INFRASTRUCTURE ONLY.
 *
 * Precision-independent part of the device-order CPU replay: Philox4x32-10,
 * the FP64 transcendental kernels, slot indexing, and the precision
 * dispatch.  The per-precision sweep code is instantiated twice from
 * dev_tmpl.h (float and double storage).  See dev_exact.h for the
 * specification summary.  Built with -ffp-contract=off: every FMA below is
 * an explicit fma()/fmaf() call.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "dev_exact.h"
#include "ym_oracle.h"

/* ---------------------------------------------------------------- Philox */
/* Salmon et al., SC'11: Philox4x32 with 10 rounds, multipliers
 * 0xD2511F53 / 0xCD9E8D57 and Weyl key increments 0x9E3779B9 / 0xBB67AE85. */
void ymd_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static inline void philox_draw(uint64_t key, uint32_t site, uint32_t w1, uint64_t sweep, uint32_t out[4]) {
    uint32_t ctr[4] = {site, w1, (uint32_t)sweep, (uint32_t)(sweep >> 32)};
    uint32_t k[2] = {(uint32_t)key, (uint32_t)(key >> 32)};
    ymd_philox4x32_10(ctr, k, out);
}
static inline double unit53(uint32_t lo, uint32_t hi) {
    uint64_t v = ((uint64_t)hi << 32) | lo;
    return (double)(v >> 11) * 0x1p-53;
}
static inline uint64_t raw64(uint32_t lo, uint32_t hi) { return ((uint64_t)hi << 32) | lo; }

enum { KIND_SU3_HIT = 0, KIND_HB_ATTEMPT = 1, KIND_HB_DIRECTION = 2, KIND_HOT_START = 3 };
static inline uint32_t word1(int mu, int kind, uint32_t index) { return (uint32_t)mu | ((uint32_t)kind << 2) | (index << 5); }

/* ------------------------------------------------------ FP64 kernels */
/* exp(-d), d >= 0: Cody-Waite reduction k = rint(x/ln2), r = x - k ln2 with
 * a two-constant ln2 and FMA, then the degree-13 Taylor series of e^r in
 * Horner form, scaled by 2^k. */
double ymd_exp_neg(double d) {
    if (!(d <= 745.0)) return 0.0;
    double x = -d;
    double k = rint(x * 0x1.71547652b82fep+0);
    double r = fma(-k, 0x1.62e42fefa39efp-1, x);
    r = fma(-k, 0x1.abc9e3b39803fp-56, r);
    double p = 0x1.6124613a86d09p-33;          /* 1/13! */
    p = fma(p, r, 0x1.1eed8eff8d898p-29);      /* 1/12! */
    p = fma(p, r, 0x1.ae64567f544e4p-26);      /* 1/11! */
    p = fma(p, r, 0x1.27e4fb7789f5cp-22);      /* 1/10! */
    p = fma(p, r, 0x1.71de3a556c734p-19);      /* 1/9!  */
    p = fma(p, r, 0x1.a01a01a01a01ap-16);      /* 1/8!  */
    p = fma(p, r, 0x1.a01a01a01a01ap-13);      /* 1/7!  */
    p = fma(p, r, 0x1.6c16c16c16c17p-10);      /* 1/6!  */
    p = fma(p, r, 0x1.1111111111111p-7);       /* 1/5!  */
    p = fma(p, r, 0x1.5555555555555p-5);       /* 1/4!  */
    p = fma(p, r, 0x1.5555555555555p-3);       /* 1/3!  */
    p = fma(p, r, 0x1.0000000000000p-1);       /* 1/2!  */
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    return ldexp(p, (int)k);
}

/* sin x and cos x - 1 for |x| <= pi/4: Taylor series to x^19 / x^20 in
 * Horner form over x^2. */
void ymd_sincos_small(double x, double *s, double *cm1) {
    double x2 = x * x;
    double ps = -0x1.2f49b46814157p-57;        /* -1/19! */
    ps = fma(ps, x2, 0x1.952c77030ad4ap-49);   /*  1/17! */
    ps = fma(ps, x2, -0x1.ae7f3e733b81fp-41);  /* -1/15! */
    ps = fma(ps, x2, 0x1.6124613a86d09p-33);   /*  1/13! */
    ps = fma(ps, x2, -0x1.ae64567f544e4p-26);  /* -1/11! */
    ps = fma(ps, x2, 0x1.71de3a556c734p-19);   /*  1/9!  */
    ps = fma(ps, x2, -0x1.a01a01a01a01ap-13);  /* -1/7!  */
    ps = fma(ps, x2, 0x1.1111111111111p-7);    /*  1/5!  */
    ps = fma(ps, x2, -0x1.5555555555555p-3);   /* -1/3!  */
    *s = fma(x * x2, ps, x);
    double pc = 0x1.e542ba4020225p-62;         /*  1/20! */
    pc = fma(pc, x2, -0x1.6827863b97d97p-53);  /* -1/18! */
    pc = fma(pc, x2, 0x1.ae7f3e733b81fp-45);   /*  1/16! */
    pc = fma(pc, x2, -0x1.93974a8c07c9dp-37);  /* -1/14! */
    pc = fma(pc, x2, 0x1.1eed8eff8d898p-29);   /*  1/12! */
    pc = fma(pc, x2, -0x1.27e4fb7789f5cp-22);  /* -1/10! */
    pc = fma(pc, x2, 0x1.a01a01a01a01ap-16);   /*  1/8!  */
    pc = fma(pc, x2, -0x1.6c16c16c16c17p-10);  /* -1/6!  */
    pc = fma(pc, x2, 0x1.5555555555555p-5);    /*  1/4!  */
    pc = fma(pc, x2, -0x1p-1);                 /* -1/2!  */
    *cm1 = x2 * pc;
}

/* sin x and cos x - 1 for |x| <= 0.29, the device's sincos_short: the series
 * in y = -x^2 with positive 1/n! coefficients, stopped at x^13 / x^12.  The
 * SU(3) hit uses it for proposal angles <= 0.5. */
void ymd_sincos_short(double x, double *s, double *cm1) {
    double y = -(x * x);
    double ps = 0x1.6124613a86d09p-33;         /* 1/13! */
    ps = fma(ps, y, 0x1.ae64567f544e4p-26);    /* 1/11! */
    ps = fma(ps, y, 0x1.71de3a556c734p-19);    /* 1/9!  */
    ps = fma(ps, y, 0x1.a01a01a01a01ap-13);    /* 1/7!  */
    ps = fma(ps, y, 0x1.1111111111111p-7);     /* 1/5!  */
    ps = fma(ps, y, 0x1.5555555555555p-3);     /* 1/3!  */
    *s = fma(x * y, ps, x);
    double pc = 0x1.1eed8eff8d898p-29;         /* 1/12! */
    pc = fma(pc, y, 0x1.27e4fb7789f5cp-22);    /* 1/10! */
    pc = fma(pc, y, 0x1.a01a01a01a01ap-16);    /* 1/8!  */
    pc = fma(pc, y, 0x1.6c16c16c16c17p-10);    /* 1/6!  */
    pc = fma(pc, y, 0x1.5555555555555p-5);     /* 1/4!  */
    pc = fma(pc, y, 0x1p-1);                   /* 1/2!  */
    *cm1 = y * pc;
}

/* sin(2 pi t), cos(2 pi t): exact quarter-turn reduction t' = t - q/4 with
 * q = rint(4t), then the Taylor kernel on x = 2 pi t' (|x| <= pi/4). */
void ymd_sincos_turns(double t, double *s, double *c) {
    double q = rint(4.0 * t);
    double tr = t - q * 0.25;
    double x = tr * 0x1.921fb54442d18p+2;
    double ss, cm1;
    ymd_sincos_small(x, &ss, &cm1);
    double cc = 1.0 + cm1;
    switch (((int64_t)q) & 3) {
    case 0: *s = ss; *c = cc; break;
    case 1: *s = cc; *c = -ss; break;
    case 2: *s = -ss; *c = -cc; break;
    default: *s = -cc; *c = ss; break;
    }
}

/* log w, w > 0: w = m 2^e with m in [sqrt(1/2), sqrt(2)), f = (m-1)/(m+1),
 * log m = 2 atanh f = 2f (1 + f^2/3 + ... + f^22/23), log w = e ln2 + log m. */
double ymd_log(double w) {
    int e;
    double m = frexp(w, &e);
    if (m < 0x1.6a09e667f3bcdp-1) { m = m * 2.0; e -= 1; }
    double f = (m - 1.0) / (m + 1.0);
    double f2 = f * f;
    double q = 0x1.642c8590b2164p-5;           /* 1/23 */
    q = fma(q, f2, 0x1.8618618618618p-5);      /* 1/21 */
    q = fma(q, f2, 0x1.af286bca1af28p-5);      /* 1/19 */
    q = fma(q, f2, 0x1.e1e1e1e1e1e1ep-5);      /* 1/17 */
    q = fma(q, f2, 0x1.1111111111111p-4);      /* 1/15 */
    q = fma(q, f2, 0x1.3b13b13b13b14p-4);      /* 1/13 */
    q = fma(q, f2, 0x1.745d1745d1746p-4);      /* 1/11 */
    q = fma(q, f2, 0x1.c71c71c71c71cp-4);      /* 1/9  */
    q = fma(q, f2, 0x1.2492492492492p-3);      /* 1/7  */
    q = fma(q, f2, 0x1.999999999999ap-3);      /* 1/5  */
    q = fma(q, f2, 0x1.5555555555555p-2);      /* 1/3  */
    double tf = 2.0 * f;
    double lm = fma(tf * f2, q, tf);
    double de = (double)e;
    return fma(de, 0x1.62e42fefa39efp-1, fma(de, 0x1.abc9e3b39803fp-56, lm));
}

/* log1p(y), y > -1: w = 1 + y, log w corrected by (y - (w - 1)) / w. */
double ymd_log1p(double y) {
    double w = 1.0 + y;
    if (w == 1.0) return y;
    return ymd_log(w) + (y - (w - 1.0)) / w;
}

/* expm1(x), x <= 0: Taylor to x^14/14! for x > -ln2/2, else exp(x) - 1. */
double ymd_expm1(double x) {
    if (x > -0x1.62e42fefa39efp-2) {
        double p = 0x1.93974a8c07c9dp-37;      /* 1/14! */
        p = fma(p, x, 0x1.6124613a86d09p-33);  /* 1/13! */
        p = fma(p, x, 0x1.1eed8eff8d898p-29);
        p = fma(p, x, 0x1.ae64567f544e4p-26);
        p = fma(p, x, 0x1.27e4fb7789f5cp-22);
        p = fma(p, x, 0x1.71de3a556c734p-19);
        p = fma(p, x, 0x1.a01a01a01a01ap-16);
        p = fma(p, x, 0x1.a01a01a01a01ap-13);
        p = fma(p, x, 0x1.6c16c16c16c17p-10);
        p = fma(p, x, 0x1.1111111111111p-7);
        p = fma(p, x, 0x1.5555555555555p-5);
        p = fma(p, x, 0x1.5555555555555p-3);
        p = fma(p, x, 0x1p-1);                 /* 1/2! */
        return fma(p, x * x, x);
    }
    return ymd_exp_neg(-x) - 1.0;
}

/* ----------------------------------------------------------- slot layout */
int64_t ymd_slots_per_parity(const int64_t n[4]) {
    int64_t n3p = n[3] + (n[3] & 1);
    return n[0] * n[1] * n[2] * n3p / 2;
}
typedef struct {
    int64_t n[4], n3p, vh;
} geom;
static geom make_geom(const int64_t n[4]) {
    geom g;
    for (int d = 0; d < 4; ++d) g.n[d] = n[d];
    g.n3p = n[3] + (n[3] & 1);
    g.vh = n[0] * n[1] * n[2] * g.n3p / 2;
    return g;
}
static inline int64_t slot_of(const geom *g, const int64_t x[4]) {
    int64_t row = (x[0] * g->n[1] + x[1]) * g->n[2] + x[2];
    return (row * g->n3p + x[3]) >> 1;
}
static inline int parity_of(const int64_t x[4]) { return (int)((x[0] + x[1] + x[2] + x[3]) & 1); }
static inline uint32_t lex_of(const geom *g, const int64_t x[4]) {
    return (uint32_t)(((x[0] * g->n[1] + x[1]) * g->n[2] + x[2]) * g->n[3] + x[3]);
}
static inline void shift(const geom *g, const int64_t x[4], int d, int s, int64_t y[4]) {
    memcpy(y, x, 4 * sizeof(int64_t));
    y[d] = (y[d] + s + g->n[d]) % g->n[d];
}
/* slot -> site of parity p; returns 0 for the padding slot of odd N3 */
static inline int site_of(const geom *g, int p, int64_t s, int64_t x[4]) {
    int64_t half = g->n3p / 2;
    int64_t row = s / half, j = s % half;
    x[2] = row % g->n[2];
    int64_t r2 = row / g->n[2];
    x[1] = r2 % g->n[1];
    x[0] = r2 / g->n[1];
    x[3] = 2 * j + ((p + x[0] + x[1] + x[2]) & 1);
    return x[3] < g->n[3];
}

/* ------------------------------------------------ SU(2) quaternions */
static inline double qnorm2(const double *a) {
    double n = a[0] * a[0];
    n = fma(a[1], a[1], n); n = fma(a[2], a[2], n); n = fma(a[3], a[3], n);
    return n;
}
/* product in the reference convention (lib.rs:212-224) as an FMA chain */
static inline void qmul64(const double *a, const double *b, double *r) {
    double r0 = a[0] * b[0];
    r0 = fma(-a[1], b[1], r0); r0 = fma(-a[2], b[2], r0); r0 = fma(-a[3], b[3], r0);
    double r1 = a[0] * b[1];
    r1 = fma(a[1], b[0], r1); r1 = fma(a[2], b[3], r1); r1 = fma(-a[3], b[2], r1);
    double r2 = a[0] * b[2];
    r2 = fma(-a[1], b[3], r2); r2 = fma(a[2], b[0], r2); r2 = fma(a[3], b[1], r2);
    double r3 = a[0] * b[3];
    r3 = fma(a[1], b[2], r3); r3 = fma(-a[2], b[1], r3); r3 = fma(a[3], b[0], r3);
    r[0] = r0; r[1] = r1; r[2] = r2; r[3] = r3;
}
static inline void qnormalize(double *a) {
    double inv = 1.0 / sqrt(qnorm2(a));
    for (int c = 0; c < 4; ++c) a[c] *= inv;
}

/* Heat-bath update of one SU(2) link from its FP64 staple a: the device form
 * of sample_local_heatbath + heatbath_link (su2_ensemble.rs:220-281).
 *   alpha = beta |a|; per link 1/alpha and tm = 1 - e^{-2 alpha};
 *   attempt k draws Philox(site, mu | HB_ATTEMPT << 2 | k << 5, sweep) -> u1, u2;
 *   r0 = 2 u1 - 1 when alpha <= 1e-300 (the reference's zero limit), else
 *   r0 = 1 + L / alpha with L = ln(1 - u1 tm): log1p(-u1 tm) when
 *   1/alpha > 2 (small alpha, where 1 - u1 tm is close to 1), log of the
 *   fused 1 - u1 tm otherwise; r0 clamped to [-1, 1];
 *   accepted iff u2^2 <= 1 - r0^2 (the reference's u2 <= sqrt(1 - r0^2));
 *   direction z = 2 u3 - 1, phi = 2 pi u4 from Philox(site, mu | HB_DIRECTION << 2, sweep);
 *   R = (r0, sqrt(1 - r0^2) n), U = R (a / |a|)^dagger (U = R when |a| = 0), normalised.
 * Returns 1 with the link in out, or 0 (out untouched) when all max_attempts
 * attempts were rejected; *attempts = attempts used. */
int ymd_su2_heatbath_link(const double a[4], double beta, uint64_t key, uint32_t site, int mu, uint64_t sweep,
                          int64_t max_attempts, double out[4], int64_t *attempts) {
    double norm = sqrt(qnorm2(a));
    double alpha = beta * norm;
    double inv_alpha = 0.0, tm = 0.0;
    if (alpha > 1.0e-300) {
        inv_alpha = 1.0 / alpha;
        tm = -ymd_expm1(-2.0 * alpha);
    }
    double r0 = 0.0;
    int64_t att = 0;
    int ok = 0;
    while (att < max_attempts) {
        uint32_t o[4];
        philox_draw(key, site, word1(mu, KIND_HB_ATTEMPT, (uint32_t)att), sweep, o);
        att += 1;
        double u1 = unit53(o[0], o[1]), u2 = unit53(o[2], o[3]);
        if (inv_alpha == 0.0) {
            r0 = 2.0 * u1 - 1.0;
        } else {
            double L = inv_alpha > 2.0 ? ymd_log1p(-u1 * tm) : ymd_log(fma(-u1, tm, 1.0));
            r0 = fma(L, inv_alpha, 1.0);
            if (r0 < -1.0) r0 = -1.0;
            if (r0 > 1.0) r0 = 1.0;
        }
        if (u2 * u2 <= fma(-r0, r0, 1.0)) { ok = 1; break; }
    }
    *attempts = att;
    if (!ok) return 0;
    double radial = sqrt(fmax(fma(-r0, r0, 1.0), 0.0));
    uint32_t o[4];
    philox_draw(key, site, word1(mu, KIND_HB_DIRECTION, 0), sweep, o);
    double z = 2.0 * unit53(o[0], o[1]) - 1.0;
    double sn, cs;
    ymd_sincos_turns(unit53(o[2], o[3]), &sn, &cs);
    double planar = sqrt(fmax(fma(-z, z, 1.0), 0.0));
    double rp = radial * planar;
    double R[4] = {r0, rp * cs, rp * sn, radial * z};
    if (norm == 0.0) {
        memcpy(out, R, 4 * sizeof(double));
    } else {
        double inv_norm = 1.0 / norm;
        double ad[4] = {a[0] * inv_norm, -(a[1] * inv_norm), -(a[2] * inv_norm), -(a[3] * inv_norm)};
        qmul64(R, ad, out);
    }
    qnormalize(out);
    return 1;
}

/* n draws of ymd_su2_heatbath_link for one staple, draw i at site i (density
 * tests); out: n x 4, attempts: n (0 links of exhausted draws). */
int ymd_su2_heatbath_samples(const double a[4], double beta, uint64_t key, int mu, uint64_t sweep,
                             int64_t max_attempts, int64_t n, double *out, int64_t *attempts) {
    int status = YMO_OK;
    for (int64_t i = 0; i < n; ++i) {
        double *u = out + 4 * i;
        if (!ymd_su2_heatbath_link(a, beta, key, (uint32_t)i, mu, sweep, max_attempts, u, &attempts[i])) {
            u[0] = u[1] = u[2] = u[3] = 0.0;
            status = YMO_HEATBATH_ATTEMPT_LIMIT;
        }
    }
    return status;
}

/* Loops over the independent links of one checkerboard sub-pass (and over
 * sites in the layout changes and hot starts) run on OpenMP threads when the
 * oracle is built with -fopenmp: each link's arithmetic is unchanged and
 * counts are integer sums, so results do not depend on the thread count. */
#define YMD_PRAGMA(x) _Pragma(#x)
#ifdef _OPENMP
#define YMD_PARALLEL_FOR YMD_PRAGMA(omp parallel for schedule(static))
#define YMD_PARALLEL_FOR_REDUCE(...) YMD_PRAGMA(omp parallel for schedule(static) reduction(__VA_ARGS__))
#else
#define YMD_PARALLEL_FOR
#define YMD_PARALLEL_FOR_REDUCE(...)
#endif

/* The SU(3) hit's sincos, as the device's sincos_hit: the short series for
 * proposal angles <= 0.5, else the full one. */
static void sincos_hit_angle(double angle, double x, double *s, double *cm1) {
    if (angle <= 0.5) ymd_sincos_short(x, s, cm1);
    else ymd_sincos_small(x, s, cm1);
}

#define REAL float
#define RSUF f32
#define RFMA fmaf
#include "dev_tmpl.h"
#undef REAL
#undef RSUF
#undef RFMA

#define REAL double
#define RSUF f64
#define RFMA fma
#include "dev_tmpl.h"
#undef REAL
#undef RSUF
#undef RFMA

/* ------------------------------------------------------------- dispatch */
#define DISPATCH(prec, call32, call64) ((prec) == 32 ? (call32) : (prec) == 64 ? (call64) : YMO_UNSUPPORTED)

int ymd_su3_to_device(const int64_t n[4], const double *aos, void *dev, int precision) {
    return DISPATCH(precision, to_device_f32(n, aos, dev, 18), to_device_f64(n, aos, dev, 18));
}
int ymd_su3_from_device(const int64_t n[4], const void *dev, int precision, double *aos) {
    return DISPATCH(precision, from_device_f32(n, dev, aos, 18), from_device_f64(n, dev, aos, 18));
}
int ymd_su2_to_device(const int64_t n[4], const double *aos, void *dev, int precision) {
    return DISPATCH(precision, to_device_f32(n, aos, dev, 4), to_device_f64(n, aos, dev, 4));
}
int ymd_su2_from_device(const int64_t n[4], const void *dev, int precision, double *aos) {
    return DISPATCH(precision, from_device_f32(n, dev, aos, 4), from_device_f64(n, dev, aos, 4));
}
int ymd_su3_hot_start(const int64_t n[4], void *dev, int precision, uint64_t key) {
    return DISPATCH(precision, su3_hot_f32(n, dev, key), su3_hot_f64(n, dev, key));
}
int ymd_su2_hot_start(const int64_t n[4], void *dev, int precision, uint64_t key) {
    return DISPATCH(precision, su2_hot_f32(n, dev, key), su2_hot_f64(n, dev, key));
}
int ymd_su3_metropolis_sweep(const int64_t n[4], void *dev, int precision, const ymd_su3_params *p,
                             uint64_t *proposals, uint64_t *accepts) {
    return DISPATCH(precision, su3_sweep_f32(n, dev, p, proposals, accepts), su3_sweep_f64(n, dev, p, proposals, accepts));
}
int ymd_su3_staple(const int64_t n[4], const void *dev, int precision, const int64_t x[4], int mu, double *out18) {
    return DISPATCH(precision, su3_staple_out_f32(n, dev, x, mu, out18), su3_staple_out_f64(n, dev, x, mu, out18));
}
int ymd_su2_heatbath_sweep(const int64_t n[4], void *dev, int precision, double beta, uint64_t key, uint64_t sweep,
                           int64_t max_attempts, uint64_t *proposals, uint64_t *updates) {
    return DISPATCH(precision, su2_hb_f32(n, dev, beta, key, sweep, max_attempts, proposals, updates),
                    su2_hb_f64(n, dev, beta, key, sweep, max_attempts, proposals, updates));
}
int ymd_su3_average_plaquette(const int64_t n[4], const void *dev, int precision, double *out) {
    return DISPATCH(precision, su3_plaq_f32(n, dev, out), su3_plaq_f64(n, dev, out));
}
int ymd_su2_overrelax_sweep(const int64_t n[4], void *dev, int precision) {
    return DISPATCH(precision, su2_or_f32(n, dev), su2_or_f64(n, dev));
}
