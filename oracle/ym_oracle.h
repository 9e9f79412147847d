/*
 * ORACLE -- TEST INFRASTRUCTURE ONLY.
 *
 * C interface of the CPU oracle: (1) the reference-exact restatement of
 * QSOLKCB/YANG-MILLS (ref_su3.c, ref_su2.c, ref_stats.c) and (2) the
 * device-order replay (dev_exact.c) that restates, on the CPU, the even/odd
 * checkerboard + Philox4x32-10 arithmetic specification the HIP kernels
 * implement.  Loaded by tests/ (ctypes), __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg only.  The product library never links it.
 *
 * Lattices are caller-owned FP64 arrays in the reference AoS layout:
 * link (site*4 + mu), site lexicographic with x3 fastest; SU(3) links are
 * 18 doubles (row-major 3x3, re/im interleaved), SU(2) links 4 doubles
 * (quaternion a0..a3 with U = a0 I - i a.sigma).
 */
#ifndef YM_ORACLE_H
#define YM_ORACLE_H
#include <stdint.h>

/* Status codes shared with the product C ABI (include/ym_amd.h) so that
 * tests can compare error behaviour directly. */
enum {
    YMO_OK = 0,
    YMO_INVALID_EXTENT = 1,
    YMO_LATTICE_TOO_LARGE = 2,
    YMO_ALLOCATION_FAILED = 3,
    YMO_DIRECTION_OUT_OF_RANGE = 4,
    YMO_DIRECTIONS_MUST_DIFFER = 5,
    YMO_NON_FINITE_BETA = 6,
    YMO_NEGATIVE_BETA = 7,
    YMO_NON_FINITE_ACTION = 8,
    YMO_EMPTY_PATH = 9,
    YMO_PATH_TOO_LONG = 10,
    YMO_PATH_NOT_CLOSED = 11,
    YMO_ZERO_RECTANGLE_EXTENT = 12,
    YMO_SU2_NON_FINITE = 13,
    YMO_SU2_ZERO_NORM = 14,
    YMO_SU2_NON_UNIT = 15,
    YMO_SU2_INVALID_AXIS = 16,
    YMO_SU3_GENERATOR_INDEX = 20,
    YMO_SU3_NON_FINITE_COEFFICIENT = 21,
    YMO_SU3_COEFFICIENT_TOO_LARGE = 22,
    YMO_SU3_NON_FINITE_MATRIX = 23,
    YMO_SU3_EXP_SCALING_OVERFLOW = 24,
    YMO_SU3_NON_UNITARY = 25,
    YMO_SU3_DET_NOT_ONE = 26,
    YMO_BETA_OUT_OF_RANGE = 30,
    YMO_ANGLE_OUT_OF_RANGE = 31,
    YMO_MEASUREMENT_COUNT = 32,
    YMO_ZERO_STRIDE = 33,
    YMO_LAG_OUT_OF_RANGE = 34,
    YMO_SCHEDULE_OUT_OF_RANGE = 35,
    YMO_LATTICE_OUT_OF_RANGE = 36,
    YMO_NON_FINITE_STATISTIC = 37,
    YMO_ACCOUNTING_OVERFLOW = 38,
    YMO_DIAGNOSTIC = 39,
    YMO_TOO_MANY_OR = 40,
    YMO_ZERO_ATTEMPT_LIMIT = 41,
    YMO_HEATBATH_ATTEMPT_LIMIT = 42,
    YMO_TIME_SLICE_OUT_OF_RANGE = 50,
    YMO_SPATIAL_EXTENT_TOO_SMALL = 51,
    YMO_SPATIAL_NOT_CUBIC = 52,
    YMO_COUNT_OVERFLOW = 53,
    YMO_NON_FINITE_OBSERVABLE = 54,
    YMO_SPATIAL_EXTENT_NOT_EVEN = 55,
    YMO_NON_FINITE_RHO = 60,
    YMO_RHO_OUT_OF_RANGE = 61,
    YMO_TOO_MANY_LEVELS = 62,
    YMO_NON_FINITE_STAPLE = 63,
    YMO_ALGEBRA_PROJECTION = 64,
    YMO_EMPTY_CONFIGURATION_SET = 70,
    YMO_TOO_MANY_CONFIGURATIONS = 71,
    YMO_TIME_EXTENT_TOO_LARGE = 72,
    YMO_EXTENTS_MISMATCH = 73,
    YMO_NON_FINITE_ACCUMULATOR = 74,
    YMO_CORRELATOR_SHAPE = 75,
    YMO_EFF_INVALID_LENGTH = 80,
    YMO_EFF_INVALID_LAG = 81,
    YMO_EFF_INVALID_WEIGHT = 82,
    YMO_EFF_INVALID_COST = 83,
    YMO_EFF_INVALID_DIAGNOSTIC = 84,
    YMO_EFF_DEGENERATE = 85,
    YMO_EFF_NON_FINITE = 86,
    YMO_WRONG_CONFIGURATION = 90,
    YMO_STATIONARITY_FAILED = 91,
    YMO_UNSUPPORTED = 99,
};

typedef struct {
    double mean, variance;
    int64_t positive_window_lag;
    double tau_int, ess, standard_error;
} ymo_chain_diagnostics;

typedef struct {
    int64_t extents[4];
    double beta;
    uint64_t master_seed, chain_id;
    int64_t thermalization_sweeps, measurements, sweeps_between_measurements;
    double proposal_angle;
    int64_t autocorrelation_max_lag;
    int64_t no_caps; /* non-zero lifts the reference's size/sweep caps */
} ymo_su3_chain_config;

typedef struct {
    uint64_t derived_seed;
    int64_t total_sweeps;
    uint64_t proposals, accepts;
    double mean_acceptance;
    double final_unitarity_residual, final_determinant_residual;
} ymo_su3_chain_report;

typedef struct {
    int64_t extents[4];
    double beta;
    uint64_t master_seed, chain_id;
    int64_t thermalization_sweeps, measurements, sweeps_between_measurements;
    int64_t overrelaxation_sweeps_per_heatbath, max_heatbath_attempts_per_link;
    int64_t autocorrelation_max_lag;
    int64_t no_caps;
} ymo_su2_chain_config;

typedef struct {
    uint64_t derived_seed;
    int64_t total_heatbath_sweeps, total_overrelaxation_sweeps;
    uint64_t heatbath_proposals, heatbath_updates;
    double mean_heatbath_acceptance, final_link_unit_residual;
} ymo_su2_chain_report;

typedef struct {
    ymo_chain_diagnostics diag;
    double reweighting_efficiency, weight_only_ess, combined_ess_proxy;
    double total_cost, sampling_cost_per_ess, total_cost_per_ess;
} ymo_efficiency_report;

typedef struct {
    double first_half_mean, second_half_mean, first_half_se, second_half_se, drift_z;
} ymo_stationarity_report;

typedef struct {
    double chain_mean[16], chain_se[16], chain_ess[16], chain_stationarity_z[16];
    double pooled_mean, pooled_se, aggregate_ess, max_pairwise_z, max_stationarity_z, min_ess;
} ymo_pool_report;

/* ---- reference-exact: RNG and seeds ---- */
uint64_t ymo_mix64(uint64_t v);
uint64_t ymo_splitmix_next(uint64_t *state);
double ymo_splitmix_unit(uint64_t *state);
uint64_t ymo_derive_chain_seed(uint64_t master, uint64_t chain);
uint64_t ymo_derive_su3_chain_seed(uint64_t master, uint64_t chain);
int ymo_proposal_angle_from_u64(uint64_t raw, double proposal_angle, double *out);

/* ---- reference-exact: SU(3) ---- */
int ymo_su3_from_algebra(const double coef[8], double *out18);
int ymo_su3_gell_mann(int index, double *out18);
int ymo_su3_checked_mul(const double *a18, const double *b18, double *out18);
void ymo_su3_matrix_residuals(const double *a18, double *unit, double *det);
int ymo_su3_structure_f(int a, int b, int c, double *out);
int ymo_su3_structure_d(int a, int b, int c, double *out);
void ymo_su3_casimirs(double *quad18, double *cubic18);
int ymo_su3_identity(const int64_t n[4], double *links);
int ymo_su3_wilson_action(const int64_t n[4], const double *links, double beta, double *out);
int ymo_su3_average_plaquette(const int64_t n[4], const double *links, double *out);
int ymo_su3_polyakov_loop(const int64_t n[4], const double *links, const int64_t start[4], int dir, double *re, double *im);
int ymo_su3_gauge_transform(const int64_t n[4], const double *links, const double *gauge_sites, double *out);
void ymo_su3_max_residuals(const int64_t n[4], const double *links, double *unit, double *det);
void ymo_su3_traceless_cycle_link(double *out18);
int ymo_su3_deterministic_site_gauge(const int64_t x[4], double *out18);
int ymo_su3_local_link_action(const int64_t n[4], const double *links, const int64_t x[4], int mu, double beta, double *out);
int ymo_su3_local_action_difference(const int64_t n[4], const double *links, const int64_t x[4], int mu, const double *cand18, double beta, double *out);
int ymo_su3_metropolis_sweep(const int64_t n[4], double *links, double beta, double angle, uint64_t *rng, uint64_t *proposals, uint64_t *accepts);
int ymo_su3_run_chain(const ymo_su3_chain_config *c, double *measurements_out, ymo_chain_diagnostics *diag, ymo_su3_chain_report *rep, double *links_out);
int ymo_su3_path_holonomy(const int64_t n[4], const double *links, const int64_t start[4], const int32_t *steps, int64_t nsteps, double *out18, int64_t end[4]);
int ymo_su3_rect_wilson_loop(const int64_t n[4], const double *links, const int64_t start[4], int mu, int nu, int64_t r, int64_t t, double *re, double *im);
int ymo_su3_a1pp_basis(const int64_t n[4], const double *links, int64_t t, double *plaq, double *rect, int64_t *plaq_count, int64_t *rect_count);
int ymo_su3_stout_smear(const int64_t n[4], const double *in, double *out, double rho, int64_t levels, int64_t max_levels);
double ymo_su3_max_link_residual(const int64_t n[4], const double *a, const double *b);
int ymo_su3_blocked_extents(const int64_t n[4], int64_t coarse[4]);
int ymo_su3_spatial_block(const int64_t n[4], const double *in, double *out);

/* ---- reference-exact: SU(2) ---- */
int ymo_su2_from_axis_angle(const double axis[3], double angle, double out[4]);
void ymo_su2_mul(const double a[4], const double b[4], double out[4]);
int ymo_su2_identity(const int64_t n[4], double *links);
int ymo_su2_plaquette(const int64_t n[4], const double *links, const int64_t x[4], int mu, int nu, double out[4]);
int ymo_su2_wilson_action(const int64_t n[4], const double *links, double beta, double *out);
int ymo_su2_path_holonomy(const int64_t n[4], const double *links, const int64_t start[4], const int32_t *steps, int64_t nsteps, double out[4], int64_t end[4]);
int ymo_su2_rect_wilson_loop(const int64_t n[4], const double *links, const int64_t start[4], int mu, int nu, int64_t r, int64_t t, double *out);
int ymo_su2_polyakov_loop(const int64_t n[4], const double *links, const int64_t start[4], int dir, double *out);
int ymo_su2_staple(const int64_t n[4], const double *links, const int64_t x[4], int mu, double out[4]);
int ymo_su2_local_action_difference(const int64_t n[4], const double *links, const int64_t x[4], int mu, const double cand[4], double beta, double *out);
double ymo_su2_max_unit_residual(const int64_t n[4], const double *links);
int ymo_su2_deterministic_site_gauge(const int64_t x[4], double out[4]);
int ymo_su2_heatbath_sweep(const int64_t n[4], double *links, double beta, uint64_t *rng, int64_t max_attempts, uint64_t *proposals, uint64_t *updates);
int ymo_su2_overrelaxation_sweep(const int64_t n[4], double *links);
int ymo_su2_average_plaquette(const int64_t n[4], const double *links, double *out);
int ymo_su2_run_chain(const ymo_su2_chain_config *c, double *measurements_out, ymo_chain_diagnostics *diag, ymo_su2_chain_report *rep, double *links_out);
int64_t ymo_su2_config_encode(const int64_t n[4], const double *links, char *buf, int64_t cap);

/* ---- reference-exact: statistics ---- */
int ymo_diagnostics(const double *v, int64_t n, int64_t max_lag, ymo_chain_diagnostics *out, double *autocorr_out);
int ymo_evaluate(const double *v, int64_t n, const double *log_weights, int64_t max_lag, double setup, double thermalization, double sampling, ymo_efficiency_report *out);
int ymo_stationarity(const double *v, int64_t n, int64_t half_lag, ymo_stationarity_report *out);
int ymo_pool_chains(const double *series, int64_t k, int64_t n, int64_t max_lag, int64_t half_lag, ymo_pool_report *out);
int ymo_periodic_connected(const double *samples, int64_t nconf, int64_t T, double *mean_out, double *raw, double *conn);
int ymo_connected_from_samples(const double *samples, int64_t nconf, int64_t T, int64_t omitted, double *conn);

#endif
