/*
 * ORACLE -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU replay of the device-order specification that the HIP kernels
 * implement (the reference has no counterpart: it is single-threaded,
 * lexicographic and SplitMix64-driven).  The replay is written independently
 * of the product sources and is used to show, decision for decision and bit
 * for bit, that the GPU chain follows the specification:
 *
 *  - Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) keyed by a chain seed
 *    derived with the reference's derive_su3_chain_seed / derive_chain_seed;
 *    counter = (lexicographic site, mu | kind<<2 | index<<5, sweep lo, hi);
 *  - even/odd checkerboard: for mu = 0..3, parity 0 then 1;
 *  - device layout: per (mu, parity) NC planes of Vh slots, slot of a site =
 *    (row * N3p + x3) >> 1 with N3p = N3 rounded up to even;
 *  - SU(3): FP32 or FP64 staple A with a fixed FMA chain, held in FP64;
 *    closed-form single-generator rotations, Delta S = -(beta/3) Re Tr((R-I)M)
 *    with the entries of M = U*A it needs formed from the current U as one
 *    FMA chain each, accept iff Delta S <= 0 or u < exp(-Delta S), U <- R U
 *    with R the identity rotation (c, s) = (1, 0) when rejected, FP64
 *    Gram-Schmidt reunitarisation before the store;
 *  - SU(2): the reference heat-bath algorithm (truncated-exponential inverse
 *    CDF + sqrt(1-r0^2) rejection, written with 1/alpha and 1/|V| formed once
 *    per link and no division or square root per attempt) and
 *    over-relaxation on FP32/FP64 storage;
 *  - transcendental functions from named published methods (Cody-Waite
 *    reduction + Taylor for exp; atanh series for log; Taylor sin/cos on a
 *    quarter-turn reduced argument), evaluated with explicit FMA only.
 */
#ifndef YM_DEV_EXACT_H
#define YM_DEV_EXACT_H
#include <stdint.h>

typedef struct {
    double beta, angle;
    int32_t hits, reunitarize;
    uint64_t key, sweep;
} ymd_su3_params;

void ymd_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double ymd_exp_neg(double d);
void ymd_sincos_small(double x, double *s, double *cm1);
void ymd_sincos_short(double x, double *s, double *cm1);
double ymd_log(double w);
double ymd_log1p(double y);
double ymd_expm1(double x);
void ymd_sincos_turns(double t, double *s, double *c);
int64_t ymd_slots_per_parity(const int64_t n[4]);

/* precision: 32 (float storage) or 64 (double storage) */
int ymd_su3_to_device(const int64_t n[4], const double *aos, void *dev, int precision);
int ymd_su3_from_device(const int64_t n[4], const void *dev, int precision, double *aos);
int ymd_su2_to_device(const int64_t n[4], const double *aos, void *dev, int precision);
int ymd_su2_from_device(const int64_t n[4], const void *dev, int precision, double *aos);
int ymd_su3_hot_start(const int64_t n[4], void *dev, int precision, uint64_t key);
int ymd_su2_hot_start(const int64_t n[4], void *dev, int precision, uint64_t key);
int ymd_su3_metropolis_sweep(const int64_t n[4], void *dev, int precision, const ymd_su3_params *p,
                             uint64_t *proposals, uint64_t *accepts);
int ymd_su3_staple(const int64_t n[4], const void *dev, int precision, const int64_t x[4], int mu, double *out18);
/* average plaquette of device-layout links, FP64 products, deterministic for any thread count */
int ymd_su3_average_plaquette(const int64_t n[4], const void *dev, int precision, double *out);
int ymd_su2_heatbath_sweep(const int64_t n[4], void *dev, int precision, double beta, uint64_t key, uint64_t sweep,
                           int64_t max_attempts, uint64_t *proposals, uint64_t *updates);
int ymd_su2_overrelax_sweep(const int64_t n[4], void *dev, int precision);
/* one heat-bath draw from an FP64 staple (the per-link step of the sweep) */
int ymd_su2_heatbath_link(const double a[4], double beta, uint64_t key, uint32_t site, int mu, uint64_t sweep,
                          int64_t max_attempts, double out[4], int64_t *attempts);
int ymd_su2_heatbath_samples(const double a[4], double beta, uint64_t key, int mu, uint64_t sweep,
                             int64_t max_attempts, int64_t n, double *out, int64_t *attempts);
#endif
