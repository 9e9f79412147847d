/*
 * ym_amd.h -- C ABI of the MI355X (gfx950) lattice link-update and
 * gauge-observable library (libym_amd.so).
 *
 * Drop-in boundary for the hot path of QSOLKCB/YANG-MILLS.  The reference
 * exposes Rust module functions compiled into each binary; each entry point
 * below names the reference function it replaces (file:line under
 * reference/rust/src/).  Conventions:
 *   - every function returns an int32_t status: YM_OK (0) or a code whose
 *     value enumerates the reference's error variants (lib.rs:53-73,
 *     su2_ensemble.rs:110-121, su3.rs:292-300, su3_lattice.rs:27-39,
 *     su3_ensemble.rs:124-137, glueball_*.rs); no C++ exception crosses
 *     the ABI and no call aborts the process;
 *   - outputs are written through caller-owned pointers; the library owns
 *     device memory, the caller owns host buffers;
 *   - host lattice buffers use the reference AoS FP64 layout: link index
 *     site*4 + mu, site lexicographic with x3 fastest (lib.rs:284-289);
 *     SU(3) links are 18 doubles (row-major 3x3, re/im interleaved,
 *     su3.rs:82-85), SU(2) links 4 doubles (a0..a3, U = a0 I - i a.sigma,
 *     lib.rs:1-12);
 *   - each lattice handle owns one HIP stream; calls are synchronous and a
 *     handle must not be used from two host threads at once.
 */
#ifndef YM_AMD_H
#define YM_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ status codes */
enum {
    YM_OK = 0,
    YM_INVALID_EXTENT = 1,            /* ReferenceError::InvalidLatticeExtent */
    YM_LATTICE_TOO_LARGE = 2,         /* ReferenceError::LatticeTooLarge (device memory) */
    YM_ALLOCATION_FAILED = 3,         /* ReferenceError::AllocationFailed */
    YM_DIRECTION_OUT_OF_RANGE = 4,    /* ReferenceError::DirectionOutOfRange */
    YM_DIRECTIONS_MUST_DIFFER = 5,
    YM_NON_FINITE_BETA = 6,
    YM_NEGATIVE_BETA = 7,
    YM_NON_FINITE_ACTION = 8,
    YM_EMPTY_PATH = 9,                /* PathHolonomyError::EmptyPath */
    YM_PATH_TOO_LONG = 10,            /* PathHolonomyError::PathTooLong */
    YM_ZERO_RECTANGLE_EXTENT = 12,
    YM_SU2_NON_FINITE = 13,           /* ReferenceError::NonFiniteSu2Component */
    YM_SU2_ZERO_NORM = 14,            /* ReferenceError::ZeroSu2Norm */
    YM_SU2_NON_UNIT = 15,             /* ReferenceError::NonUnitSu2 */
    YM_COORDINATE_OUT_OF_RANGE = 17,  /* ReferenceError::CoordinateOutOfRange */
    YM_INVALID_CONFIGURATION_ENCODING = 18, /* ReferenceError::InvalidConfigurationEncoding */
    YM_LATTICE_SIZE_OVERFLOW = 19,    /* ReferenceError::LatticeSizeOverflow */
    YM_SU3_NON_FINITE_MATRIX = 23,    /* Su3Error::NonFiniteMatrix */
    YM_SU3_NON_UNITARY = 25,          /* Su3Error::NonUnitaryMatrix */
    YM_SU3_DETERMINANT_NOT_ONE = 26,  /* Su3Error::DeterminantNotOne */
    YM_BETA_OUT_OF_RANGE = 30,        /* Su3EnsembleError / EnsembleError::BetaOutsideReferenceRange */
    YM_ANGLE_OUT_OF_RANGE = 31,       /* Su3EnsembleError::ProposalAngleOutsideReferenceRange */
    YM_MEASUREMENT_COUNT = 32,
    YM_ZERO_STRIDE = 33,
    YM_LAG_OUT_OF_RANGE = 34,
    YM_SCHEDULE_OUT_OF_RANGE = 35,
    YM_NON_FINITE_STATISTIC = 37,
    YM_TOO_MANY_OR = 40,
    YM_ZERO_ATTEMPT_LIMIT = 41,       /* EnsembleError::ZeroHeatbathAttemptLimit */
    YM_HEATBATH_ATTEMPT_LIMIT = 42,   /* EnsembleError::HeatbathAttemptLimitExceeded */
    YM_TIME_SLICE_OUT_OF_RANGE = 50,
    YM_SPATIAL_EXTENT_TOO_SMALL = 51,
    YM_SPATIAL_NOT_CUBIC = 52,
    YM_SPATIAL_EXTENT_NOT_EVEN = 55,  /* GlueballBlockingError::SpatialExtentNotEven */
    YM_NON_FINITE_RHO = 60,
    YM_RHO_OUT_OF_RANGE = 61,
    YM_TOO_MANY_LEVELS = 62,
    YM_NON_FINITE_OBSERVABLE = 54,    /* GlueballCorrelatorError::NonFiniteObservable */
    YM_EMPTY_CONFIGURATION_SET = 70,
    YM_TOO_MANY_CONFIGURATIONS = 71,
    YM_TIME_EXTENT_TOO_LARGE = 72,
    YM_NON_FINITE_ACCUMULATOR = 74,   /* GlueballCorrelatorError::NonFiniteAccumulator */
    YM_CORRELATOR_SHAPE = 75,
    YM_EFF_INVALID_LENGTH = 80,
    YM_EFF_INVALID_LAG = 81,
    YM_EFF_INVALID_WEIGHT = 82,
    YM_EFF_INVALID_COST = 83,
    YM_EFF_INVALID_DIAGNOSTIC = 84,
    YM_EFF_DEGENERATE = 85,
    YM_EFF_NON_FINITE = 86,
    YM_EM_TIME_EXTENT_TOO_SMALL = 100, /* EffectiveMassError::TimeExtentTooSmall (T < 4) */
    YM_EM_TIME_EXTENT_TOO_LARGE = 101, /* EffectiveMassError::TimeExtentTooLarge (T > 64) */
    YM_EM_TOO_FEW_REPLICAS = 102,      /* EffectiveMassError::TooFewJackknifeReplicates (< 3) */
    YM_EM_TOO_MANY_REPLICAS = 103,     /* EffectiveMassError::TooManyJackknifeReplicates (> 64) */
    YM_EM_EMPTY_WINDOW = 104,          /* EffectiveMassError::EmptyAnalysisWindow */
    YM_EM_INVALID_LAG = 105,           /* EffectiveMassError::InvalidAnalysisLag */
    YM_EM_SHAPE_MISMATCH = 106,        /* EffectiveMassError::ShapeMismatch */
    YM_EM_NON_FINITE_CORRELATOR = 107, /* EffectiveMassError::NonFiniteCorrelator */
    YM_EM_NON_POSITIVE_CENTRE = 108,   /* EffectiveMassError::NonPositiveCentralCorrelator */
    YM_EM_ACOSH_DOMAIN = 109,          /* EffectiveMassError::AcoshDomainViolation */
    YM_EM_NON_FINITE_STATISTIC = 110,  /* EffectiveMassError::NonFiniteStatistic */
    YM_EM_INVALID_FIXTURE = 111,       /* EffectiveMassError::InvalidFixtureParameter */
    YM_FIT_EMPTY_CANDIDATES = 120,     /* ScalarFitError::EmptyCandidateWindows */
    YM_FIT_TOO_MANY_CANDIDATES = 121,  /* ScalarFitError::TooManyCandidateWindows */
    YM_FIT_INVALID_CANDIDATE = 122,    /* ScalarFitError::InvalidCandidateWindow */
    YM_FIT_DUPLICATE_CANDIDATE = 123,  /* ScalarFitError::DuplicateCandidateWindow */
    YM_FIT_DEGENERATE_PLATEAU = 124,   /* ScalarFitError::DegeneratePlateauUncertainty */
    YM_FIT_NO_PLATEAU = 125,           /* ScalarFitError::NoAcceptedPlateauWindow */
    YM_FIT_INSUFFICIENT_REPLICAS = 126, /* ScalarFitError::InsufficientReplicatesForCovariance */
    YM_FIT_COVARIANCE_NOT_PD = 127,    /* ScalarFitError::CovarianceNotPositiveDefinite */
    YM_FIT_SINGULAR_COVARIANCE = 128,  /* ScalarFitError::SingularCovariance */
    YM_FIT_NON_FINITE_STATISTIC = 129, /* ScalarFitError::NonFiniteStatistic */
    YM_FIT_NO_GRID_FIT = 130,          /* ScalarFitError::NoFiniteGridFit */
    YM_VAR_TIME_EXTENT_TOO_SMALL = 140, /* VariationalError::TimeExtentTooSmall (T < 3) */
    YM_VAR_MATRIX_NOT_SYMMETRIC = 142, /* VariationalError::MatrixNotSymmetric */
    YM_VAR_INVALID_REFERENCE_TIMES = 143, /* VariationalError::InvalidReferenceTimes */
    YM_VAR_REFERENCE_NOT_PD = 144,     /* VariationalError::ReferenceMatrixNotPositiveDefinite */
    YM_VAR_DEGENERATE_EIGENVECTOR = 145, /* VariationalError::DegeneratePrincipalEigenvector */
    YM_VAR_NON_POSITIVE_EIGENVALUE = 146, /* VariationalError::NonPositivePrincipalEigenvalue */
    YM_VAR_PROJECTION_NORMALIZATION = 147, /* VariationalError::ProjectionNormalizationFailure */
    YM_WRONG_CONFIGURATION = 90,      /* benchmark gate: schedule does not allow the split */
    YM_STATIONARITY_FAILED = 91,      /* benchmark gate: constant half or zero spread */
    YM_WORKSPACE_EXCEEDED = 95,       /* input larger than an allocation-free entry point's fixed workspace */
    YM_KERNEL_HANDOFF_TIMEOUT = 94,   /* a sweep kernel's bounded wave-to-wave wait expired (no reference counterpart) */
    YM_INVALID_ARGUMENT = 96,         /* null pointer, wrong group/precision, size mismatch */
    YM_NO_DEVICE = 97,
    YM_HIP_ERROR = 98,
    YM_UNSUPPORTED = 99               /* e.g. a checkerboard sweep on an odd extent */
};

typedef struct ym_lattice ym_lattice;

/* ------------------------------------------------------- lattice handles */
/* Lattice4::new_identity (lib.rs:237-263) / Su3Lattice4::new_identity
 * (su3_lattice.rs:85-106): group 2 or 3, precision 32 (FP32 link storage)
 * or 64 (FP64 link storage), every extent >= 2.  Links start at identity.
 * The reference's 2^20 / 2^18 link caps do not apply (device memory only). */
int32_t ym_lattice_create(int32_t group, const int64_t extents[4], int32_t precision, ym_lattice **out);
int32_t ym_lattice_destroy(ym_lattice *lat);
/* Device bytes held by the link field. */
int32_t ym_lattice_bytes(const ym_lattice *lat, int64_t *bytes);
int32_t ym_lattice_set_identity(ym_lattice *lat);
/* Hot start (no reference counterpart; the reference is cold-start only,
 * su2_ensemble.rs:490, su3_ensemble.rs:412): Haar-random links from
 * Philox4x32-10 keyed by `seed`, counter (site, mu, draw). */
int32_t ym_lattice_hot_start(ym_lattice *lat, uint64_t seed);
/* Lattice4::set_link / Su3Lattice4::set_link for every link at once
 * (lib.rs:323-331, su3_lattice.rs:154-162): `count` = links * (4 or 18).
 * Fails closed, before anything is stored, on a link the reference's types
 * cannot hold (the first such link in site order decides the status):
 * SU(3) as Su3::from_checked_matrix (su3.rs:242-253) -- YM_SU3_NON_FINITE_MATRIX,
 * YM_SU3_NON_UNITARY (max entry of |U^+U - 1|, |UU^+ - 1|), then
 * YM_SU3_DETERMINANT_NOT_ONE; SU(2) as Su2 (lib.rs:155-180) --
 * YM_SU2_NON_FINITE, YM_SU2_ZERO_NORM, YM_SU2_NON_UNIT.  Tolerance: the
 * reference's 5e-12 (SU(3)) / 2e-12 (SU(2)) for FP64 storage, 2^-16 for FP32
 * storage.  The image moves through a staging buffer of at most 2^16 sites
 * that the handle allocates on first use and keeps. */
int32_t ym_lattice_upload(ym_lattice *lat, const double *aos, int64_t count);
/* ym_lattice_upload without the link checks (no reference counterpart: the
 * reference's link types cannot carry unchecked values).  For tests and for
 * callers that have checked the configuration themselves. */
int32_t ym_lattice_upload_unchecked(ym_lattice *lat, const double *aos, int64_t count);
/* Lattice4::link / Su3Lattice4::link for every link (lib.rs:319-321). */
int32_t ym_lattice_download(const ym_lattice *lat, double *aos, int64_t count);
/* Clone (same group, precision and extents). */
int32_t ym_lattice_copy(ym_lattice *dst, const ym_lattice *src);

/* ------------------------------------------------------------- samplers */
/* su3_ensemble.rs:342-391 metropolis_sweep, as `n_sweeps` even/odd
 * checkerboard sweeps (mu = 0..3, parity 0 then 1).  Per link: `hits`
 * single-generator proposals, generator h mod 8 (hits = 8 is one reference
 * sweep's proposal set), angle from the reference's paired grid
 * (su3_ensemble.rs:199-213), random numbers from Philox4x32-10 keyed by
 * `key` (derive_su3_chain_seed, su3_ensemble.rs:189) with counter
 * (site, mu, hit, first_sweep + k).  Accept iff dS <= 0 or u < exp(-dS) in
 * FP64.  reunitarize != 0 applies FP64 Gram-Schmidt to every updated link
 * before it is stored (always on for FP32 storage).  Bounds as the
 * reference: 0 <= beta <= 12, 1e-8 <= angle <= 1.  Counts are accumulated
 * into *proposals / *accepts (not reset).  Requires even extents. */
int32_t ym_su3_metropolis_sweeps(ym_lattice *lat, double beta, double proposal_angle, int32_t hits, uint64_t key,
                                 uint64_t first_sweep, int32_t n_sweeps, int32_t reunitarize, uint64_t *proposals,
                                 uint64_t *accepts);
/* Autotuner (no reference counterpart): times the Metropolis sub-pass for
 * nine thread-to-link mappings -- row tiles (256 consecutive slots) visited
 * in x1-strips of width {0 = slot order, 2, 4, 8, 16}, and compact tiles
 * (2 x 2 (x0, x1) rows x 64 slots) in X1-strips of {1, 2, 4, 8} -- on a
 * scratch copy of the links (reps timed sweeps each, after one untimed),
 * keeps the fastest for this lattice and returns it in *best_mode (0 row, 1
 * compact) and *best_strip; us_per_candidate (10 entries, may be NULL) gets
 * the mean time per sub-pass, -1 where a mapping does not fit the extents.
 * From 256 tiles of 256 links up the sub-pass uses the wave-specialised
 * kernel; there the one-link-per-thread kernel is also timed with the best
 * mapping (entry 9, else -1) and kept if faster, reported as *best_mode + 2. 
 * Every mapping permutes independent links, so results do not depend on it. */
int32_t ym_su3_autotune(ym_lattice *lat, double beta, double proposal_angle, int32_t hits, int32_t reps,
                        int32_t *best_mode, int32_t *best_strip, float *us_per_candidate);
/* Profiling variant of ym_su3_metropolis_sweeps (same arithmetic and
 * results): HIP events on the handle's stream around every sub-pass launch.
 * *total_ms = first start to last stop; *kernel_ms_sum = sum of the
 * per-launch durations; *launches = 8 * n_sweeps.  With kernel_ms_sum and
 * launches both NULL only the span is timed (two events, none between
 * launches).  No reference counterpart (the reference has no timers). */
int32_t ym_su3_metropolis_sweeps_profiled(ym_lattice *lat, double beta, double proposal_angle, int32_t hits,
                                          uint64_t key, uint64_t first_sweep, int32_t n_sweeps, int32_t reunitarize,
                                          uint64_t *proposals, uint64_t *accepts, float *total_ms,
                                          float *kernel_ms_sum, int32_t *launches);
/* su2_ensemble.rs:321-342 heatbath_sweep, checkerboard order; attempts of a
 * link draw Philox(site, mu, attempt, sweep).  0 <= beta <= 8,
 * max_attempts > 0.  A link that exhausts its attempts keeps its value and
 * the call returns YM_HEATBATH_ATTEMPT_LIMIT after the sweeps finish.
 * *proposals (attempts) and *updates are accumulated. */
int32_t ym_su2_heatbath_sweeps(ym_lattice *lat, double beta, uint64_t key, uint64_t first_sweep, int32_t n_sweeps,
                               int32_t max_attempts, uint64_t *proposals, uint64_t *updates);
/* su2_ensemble.rs:283-299, 344-346 overrelaxation_sweep, checkerboard order. */
int32_t ym_su2_overrelax_sweeps(ym_lattice *lat, int32_t n_sweeps);
/* su2_ensemble.rs:465-486 markov_sweep, n_sweeps times: one heat-bath sweep
 * (sweep index first_sweep + k) then `overrelaxation_sweeps_per_heatbath`
 * (0..16) over-relaxation sweeps.  Counts accumulate as in
 * ym_su2_heatbath_sweeps; YM_HEATBATH_ATTEMPT_LIMIT after the sweeps if a
 * link exhausted its attempts. */
int32_t ym_su2_markov_sweeps(ym_lattice *lat, double beta, uint64_t key, uint64_t first_sweep, int32_t n_sweeps,
                             int32_t overrelaxation_sweeps_per_heatbath, int32_t max_attempts, uint64_t *proposals,
                             uint64_t *updates);

/* ----------------------------------------------------------- observables */
/* su2_ensemble.rs:348-373 / su3_ensemble.rs:228-259 average_plaquette:
 * (1/6V) sum_x sum_{mu<nu} (1/N) Re Tr P_{mu nu}(x), FP64 accumulation. */
int32_t ym_average_plaquette(const ym_lattice *lat, double *out);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops or harnesses that
 * forbid hipMalloc): ym_average_plaquette with the same kernels and the same
 * result, bit for bit, but its per-block partials in a fixed library-wide
 * device array of 65,536 doubles (512 KiB per GPU, held until the process
 * ends) instead of the handle's lazily grown buffer; it calls no hipMalloc,
 * hipFree or allocating library routine.  Limit: 2 * ceil(Vh / 256) <= 65,536
 * partials with Vh the slots per parity (V/2 for even extents), i.e. at most
 * 2^23 slots per parity (a 64^4 lattice); larger lattices return
 * YM_WORKSPACE_EXCEEDED.  Every call on a GPU shares that
 * GPU's array: calls that overlap on one GPU (two host threads, two
 * streams) race. */
int32_t ym_average_plaquette_noalloc(const ym_lattice *lat, double *out);
/* lib.rs:363-382 / su3_lattice.rs:214-242 wilson_action: beta * sum (1 - P). */
int32_t ym_wilson_action(const ym_lattice *lat, double beta, double *out);
/* lib.rs:507-546 staple (SU(2)); for SU(3) the same six-term sum A with
 * Re Tr(U A) = sum of the six plaquettes touching U (su3_ensemble.rs:261-307).
 * out: 4 (SU(2)) or 18 (SU(3)) doubles, computed as in the sweep kernels. */
int32_t ym_staple(const ym_lattice *lat, const int64_t x[4], int32_t mu, double *out);
/* Averaged rectangular Wilson loops: for 1 <= R <= rmax, 1 <= T <= tmax,
 * the mean over all sites and the planes of `plane_set` of
 * rectangular_wilson_loop(x, mu, nu, R, T) (lib.rs:471-481,
 * path_holonomy.rs:219-260 + wilson_loop :172-181): R steps along mu, T along
 * nu.  plane_set 0: all 12 ordered (mu, nu); 1: mu spatial, nu = 3 (time);
 * 2: the 6 ordered spatial pairs.  re/im: rmax*tmax doubles, index
 * (R-1)*tmax + (T-1); im is 0 for SU(2). */
/* path_holonomy.rs:132-170 (SU(3)), lib.rs:384-436 (SU(2)): ordered product
 * along a path from `start`; steps[k] = +(d+1) multiplies by U_d(x) and
 * moves to x + d, -(d+1) moves to x - d and multiplies by U_d(x-d)^dagger.
 * FP64 product of the stored links; out gets 18 (SU(3)) or 4 (SU(2))
 * doubles in the host layout, end (may be NULL) the end point.
 * 1 <= nsteps <= 2^20. */
int32_t ym_path_holonomy(const ym_lattice *lat, const int64_t start[4], const int32_t *steps, int64_t nsteps,
                         double *out, int64_t end[4]);
int32_t ym_wilson_loops(const ym_lattice *lat, int32_t rmax, int32_t tmax, int32_t plane_set, double *re, double *im);
/* Polyakov loop averaged over all starting sites of the hyperplane
 * x_dir = 0 (lib.rs:483-505 SU(2) (1/2)Re Tr, su3_lattice.rs:185-212 SU(3)
 * (1/3)Tr): *re, *im of the mean, *abs = mean of |P(x)|. */
int32_t ym_polyakov_loop(const ym_lattice *lat, int32_t dir, double *re, double *im, double *abs_mean);
/* glueball_operators.rs:143-160 scalar_a1pp_operator_basis for every time
 * slice t = 0..N3-1 at once: plaq_t[t] (1x1, 3 orientations) and rect_t[t]
 * (2x1, 6 orientations), each the mean of (1/3) Re Tr over L^3 sites.
 * SU(3) only; spatial extents equal and >= 3. */
int32_t ym_su3_timeslice_a1pp(const ym_lattice *lat, double *plaq_t, double *rect_t);
/* glueball_smearing.rs:257-268 spatial_stout_smear: `levels` synchronous
 * spatial stout steps with parameter rho (0 < rho <= 0.25) from src into
 * dst (same shape; src is not modified; temporal links copied).  The
 * reference's 8-level cap is lifted to 64. */
int32_t ym_su3_stout_spatial(ym_lattice *dst, const ym_lattice *src, double rho, int32_t levels);

/* glueball_blocking.rs:71-94 blocked_extents: coarse extents (N/2, N/2, N/2,
 * N3) of src; fails with YM_SPATIAL_NOT_CUBIC, YM_SPATIAL_EXTENT_TOO_SMALL
 * (N < 6) or YM_SPATIAL_EXTENT_NOT_EVEN, in the reference's order. */
int32_t ym_su3_blocked_extents(const ym_lattice *src, int64_t coarse[4]);
/* glueball_blocking.rs:105-136 spatial_stride2_block: dst (created with the
 * blocked extents and src's precision) receives V_mu(y) = U_mu(x) U_mu(x+mu)
 * for spatial mu and V_3(y) = U_3(x), x = (2y0, 2y1, 2y2, y3).  The product
 * is formed in the storage precision; the reference's per-link group check
 * is left to ym_link_residuals. */
int32_t ym_su3_spatial_block(ym_lattice *dst, const ym_lattice *src);
/* su3_lattice.rs:285-297 max_link_{unitarity,determinant}_residual (SU(3));
 * lib.rs:696-712 max_link_unit_residual in *unit for SU(2) (*det = 0). */
int32_t ym_link_residuals(const ym_lattice *lat, double *unit, double *det);
/* FP64 Gram-Schmidt (SU(3)) / normalisation (SU(2)) of every link; no
 * reference counterpart (the reference fails closed on residuals). */
int32_t ym_reunitarize(ym_lattice *lat);

/* ------------------------------------------------------ chain drivers */
typedef struct {
    int64_t extents[4];
    double beta;
    uint64_t master_seed, chain_id;
    int64_t thermalization_sweeps, measurements, sweeps_between_measurements;
    double proposal_angle;
    int64_t autocorrelation_max_lag;
    int32_t hits;        /* proposals per link per sweep (8 = reference) */
    int32_t precision;   /* 32 or 64 */
    int32_t hot_start;   /* 0 = cold (reference), 1 = Haar-random from the chain key */
    int32_t reunitarize; /* FP64 reunitarisation per update (forced for FP32) */
} ym_su3_chain_config;

typedef struct {
    double mean, variance;
    int64_t positive_window_lag;
    double tau_int, ess, standard_error;
} ym_chain_diagnostics;

typedef struct {
    uint64_t derived_seed;
    int64_t total_sweeps;
    uint64_t proposals, accepts;
    double mean_acceptance;
    double final_unitarity_residual, final_determinant_residual;
    ym_chain_diagnostics diagnostics;
} ym_su3_chain_report;

/* su3_ensemble.rs:408-480 run_chain on the device: validate (beta, angle,
 * schedule; no size cap), key = derive_su3_chain_seed(master, chain)
 * (su3_ensemble.rs:189-192), thermalisation, then `measurements` x
 * (`sweeps_between_measurements` sweeps + average plaquette), diagnostics
 * (su2_ensemble.rs:375-463), final residuals.  measurements_out holds
 * config->measurements doubles. */
int32_t ym_su3_run_chain(const ym_su3_chain_config *config, double *measurements_out, ym_su3_chain_report *report);

/* su2_ensemble.rs:56-68 EnsembleConfig, plus the storage precision and the
 * start (the reference is cold-start only). */
typedef struct {
    int64_t extents[4];
    double beta;
    uint64_t master_seed, chain_id;
    int64_t thermalization_sweeps, measurements, sweeps_between_measurements;
    int64_t overrelaxation_sweeps_per_heatbath;
    int64_t max_heatbath_attempts_per_link;
    int64_t autocorrelation_max_lag;
    int32_t precision;   /* 32 or 64 */
    int32_t hot_start;   /* 0 = cold (reference), 1 = Haar-random from the chain key */
} ym_su2_chain_config;

/* su2_ensemble.rs:173-186 ChainReport (measurements returned separately). */
typedef struct {
    uint64_t derived_seed;
    int64_t total_heatbath_sweeps, total_overrelaxation_sweeps;
    uint64_t heatbath_proposals, heatbath_updates;
    double mean_heatbath_acceptance;
    double final_link_unit_residual;
    ym_chain_diagnostics diagnostics;
} ym_su2_chain_report;

/* su2_ensemble.rs:488-552 run_chain on the device: EnsembleConfig::validate
 * (su2_ensemble.rs:70-103, same bounds and order: beta in [0, 8], at most
 * 2^20 sweeps, >= 2 measurements, stride > 0, <= 16 over-relaxation sweeps,
 * attempt limit > 0, lag < measurements), key = derive_chain_seed(master,
 * chain) (su2_ensemble.rs:46-55), thermalisation Markov sweeps, then
 * `measurements` x (`sweeps_between_measurements` Markov sweeps + average
 * plaquette), diagnostics, acceptance and the final unit residual.
 * measurements_out holds config->measurements doubles. */
int32_t ym_su2_run_chain(const ym_su2_chain_config *config, double *measurements_out, ym_su2_chain_report *report);

/* ------------------------------------------------------ batched ensembles */
/* Independent SU(3) chains of uneven geometries in one device buffer; every
 * sweep and measurement is one batched launch over all chains.  Reference
 * counterpart: generate_independent_configurations
 * (glueball_scalar_production.rs:257-297), which runs the chains one after
 * another with key derive_su3_chain_seed(MASTER_SEED, chain_id), a cold
 * start (su3_ensemble.rs:412) and metropolis_sweep (su3_ensemble.rs:342);
 * chain c of an ensemble reproduces bit for bit a single lattice created
 * with the same extents and precision and swept with keys[c]. */
typedef struct ym_ensemble ym_ensemble;

/* extents: nchains x 4 (all even, >= 2); keys: nchains Philox keys
 * (normally derive_su3_chain_seed(master, chain_id)).  Cold start, sweep
 * index 0. */
int32_t ym_ensemble_create(int32_t nchains, const int64_t *extents, const uint64_t *keys, int32_t precision,
                           ym_ensemble **out);
int32_t ym_ensemble_destroy(ym_ensemble *ens);
/* Cold start of every chain (su3_ensemble.rs:412); resets the sweep index. */
int32_t ym_ensemble_set_identity(ym_ensemble *ens);
/* Haar-random start of every chain from its key (no reference counterpart). */
int32_t ym_ensemble_hot_start(ym_ensemble *ens);
int32_t ym_ensemble_chain_count(const ym_ensemble *ens, int32_t *nchains, int64_t *total_sites, int64_t *time_sum);
/* Reference AoS FP64 layout of one chain (volume x 4 x 18 doubles); the
 * upload applies ym_lattice_upload's link checks and fails closed. */
int32_t ym_ensemble_upload_chain(ym_ensemble *ens, int32_t chain, const double *aos, int64_t count);
int32_t ym_ensemble_download_chain(const ym_ensemble *ens, int32_t chain, double *aos, int64_t count);
int32_t ym_ensemble_sweep_index(const ym_ensemble *ens, uint64_t *sweep);
/* Length of one cycle's operator block for `stout_levels` and each chain's
 * timeslice offset t_off[c]: chain c, level l, timeslice t, kind k (0 1x1
 * plaquettes, 1 2x1 rectangles) sits at ((l * time_sum + t_off[c] + t) * 2 + k). */
int32_t ym_ensemble_ops_length(const ym_ensemble *ens, int32_t stout_levels, int64_t *ops_len, int64_t *chain_offsets);
/* n_cycles x [sweeps_per_cycle checkerboard Metropolis sweeps of every chain
 * (sweep indices continue from the ensemble's counter), then, when plaquette
 * is not NULL, one measurement: average_plaquette per chain
 * (su3_ensemble.rs:228-259) into plaquette[cycle * nchains + c] and the A1++
 * timeslice basis (glueball_operators.rs:93-160) of the raw links and of
 * stout levels 1..stout_levels (glueball_smearing.rs:137-268, rho) into
 * ops[cycle * ops_len + ...]].  use_graph != 0 captures one cycle as a
 * hipGraph and replays it; results are identical either way.  accepts
 * (nchains, may be NULL) is accumulated; device_ms (may be NULL) is the
 * device time of all cycles.  Bounds as ym_su3_metropolis_sweeps and
 * ym_su3_stout_spatial (cubic spatial extents >= 3 when measuring). */
int32_t ym_ensemble_run_cycles(ym_ensemble *ens, double beta, double proposal_angle, int32_t hits,
                               int32_t reunitarize, int32_t sweeps_per_cycle, int32_t stout_levels, double rho,
                               int32_t n_cycles, int32_t use_graph, double *plaquette, double *ops,
                               uint64_t *accepts, float *device_ms);

/* ------------------------------------------------- learned link proposal */
/* No reference counterpart (the reference samples with plain
 * single-generator Metropolis, su3_ensemble.rs:336-391).  An MLP
 * (12 gauge-invariant staple features -> 64 tanh -> 64 tanh -> 16) gives the
 * mean and log-width of a Gaussian in the exponential chart around the
 * staple-aligned link W(A); each link gets one independence proposal with
 * the exact Metropolis-Hastings ratio (Haar Jacobian included), followed by
 * `metro_hits` plain Metropolis hits.  The handle owns the FP32 parameters
 * (6032: W1[64][12], b1[64], W2[64][64], b2[64], W3[16][64], b3[16]), Adam
 * moments and a device dataset of (features, chart coordinates) pairs. */
typedef struct ym_mlp ym_mlp;
int32_t ym_mlp_create(uint64_t seed, int64_t dataset_capacity, ym_mlp **out);
int32_t ym_mlp_destroy(ym_mlp *mlp);
int32_t ym_mlp_get_params(const ym_mlp *mlp, float *params);
/* Also resets the Adam moments and step count. */
int32_t ym_mlp_set_params(ym_mlp *mlp, const float *params);
int32_t ym_mlp_dataset_size(const ym_mlp *mlp, int64_t *n, int64_t *capacity);
int32_t ym_mlp_clear_pairs(ym_mlp *mlp);
/* Appends one (staple features, chart coordinates, valid) row per link of an
 * SU(3) FP32 lattice (8 Vh rows, group-major). */
int32_t ym_mlp_collect_pairs(ym_mlp *mlp, const ym_lattice *lat, int64_t *total);
int32_t ym_mlp_download_pairs(const ym_mlp *mlp, int64_t first, int64_t count, float *features, float *targets,
                              int32_t *valid);
/* Replaces the dataset with `count` host rows. */
int32_t ym_mlp_upload_pairs(ym_mlp *mlp, const float *features, const float *targets, const int32_t *valid,
                            int64_t count);
/* out[n][16] = MLP(features[n][12]); bf16 != 0 rounds weights and layer
 * inputs to bfloat16 (FP32 accumulation). */
int32_t ym_mlp_forward(const ym_mlp *mlp, const float *features, int64_t n, int32_t bf16, float *out);
/* Mean Gaussian negative log-likelihood of dataset rows [first, first+count)
 * and its gradient (6032). */
int32_t ym_mlp_loss_grad(ym_mlp *mlp, int64_t first, int64_t count, float *loss, float *grad);
/* Adam (lr, beta1 0.9, beta2 0.999) over shuffled minibatches of the
 * dataset; epoch_loss (epochs, may be NULL) gets the mean loss per epoch. */
int32_t ym_mlp_train(ym_mlp *mlp, int32_t epochs, int32_t batch, float lr, uint64_t seed, float *epoch_loss,
                     float *device_ms);
/* n_sweeps checkerboard sweeps (SU(3), FP32 storage): per link one learned
 * independence proposal with MH correction, then metro_hits Metropolis hits
 * (generator h mod 8, proposal_angle), FP64 reunitarisation.  Counts are
 * accumulated.  The network is only read (its BF16 weight fragments are
 * repacked by ym_mlp_set_params / ym_mlp_train), and counters and events
 * belong to the lattice handle, so calls on different lattices may share one
 * MLP concurrently (from different host threads) while no call changes it. */
int32_t ym_su3_learned_sweeps(ym_lattice *lat, const ym_mlp *mlp, double beta, uint64_t key, uint64_t first_sweep,
                              int32_t n_sweeps, int32_t metro_hits, double proposal_angle, int32_t bf16,
                              uint64_t *learned_proposals, uint64_t *learned_accepts, uint64_t *metro_accepts,
                              float *device_ms);
/* Probe of the link (x, mu): out[0..17] staple A, [18..35] frame W(A),
 * [36..43] chart coordinates of the link, [44..46] chart phases,
 * [47..58] features, [59] frame ok, [60] chart ok, [61] log Haar Jacobian. */
int32_t ym_su3_link_chart(const ym_lattice *lat, const int64_t x[4], int32_t mu, double *out62);
/* Probe of one learned MH step (no reference counterpart; for tests): for
 * the link (x, mu) of the lattice as it stands, the values k_su3_learned
 * computes for it in sub-pass (mu, parity of x) of sweep `sweep` with the
 * FP32 network, by the same device function: out[0..7] chart coordinates of
 * the link, [8..15] mu, [16..23] sigma, [24..31] proposal eps', [32] log q(U),
 * [33] log q(U'), [34] (beta/3) Re Tr((U' - U) A), [35] log MH ratio,
 * [36] accept uniform, [37] proposal inside the chart, [38] accepted,
 * [39] chart of the link ok, [40..57] U' (18 doubles, host layout). */
int32_t ym_su3_learned_probe(const ym_lattice *lat, const ym_mlp *mlp, double beta, uint64_t key, uint64_t sweep,
                             const int64_t x[4], int32_t mu, double *out58);

/* ------------------------------------------------- host FP64 statistics */
/* su2_ensemble.rs:375-463 diagnostics.  autocorr (optional) gets max_lag+1. */
int32_t ym_diagnostics(const double *values, int64_t n, int64_t max_lag, ym_chain_diagnostics *out, double *autocorr);

typedef struct {
    ym_chain_diagnostics diagnostics;
    double reweighting_efficiency, weight_only_ess, combined_ess_proxy;
    double total_cost, sampling_cost_per_ess, total_cost_per_ess;
} ym_efficiency_report;

/* sampling_efficiency.rs:57-135 evaluate: 2 <= n <= 4096, 1 <= max_lag <
 * min(n, 257); log_weights NULL = unit weights (Kish efficiency 1); costs
 * finite and >= 0 with sampling > 0 (seconds, or any cost unit). */
int32_t ym_evaluate(const double *values, int64_t n, const double *log_weights, int64_t max_lag, double setup_cost,
                    double thermalization_cost, double sampling_cost, ym_efficiency_report *out);

typedef struct {
    double first_half_mean, second_half_mean, first_half_se, second_half_se, drift_z;
} ym_stationarity_report;

/* su3_benchmark.rs:186-237 / su2_benchmark.rs:188-245 stationarity gate:
 * n even, half_lag < n/2; z = |mean_1 - mean_2| / sqrt(se_1^2 + se_2^2). */
int32_t ym_stationarity(const double *values, int64_t n, int64_t half_lag, ym_stationarity_report *out);

typedef struct {
    double pooled_mean, pooled_se, aggregate_ess;
    double max_pairwise_z, max_stationarity_z, min_ess;
} ym_pool_summary;

/* su3_confirmation.rs:105-263 (k chains) and su3_benchmark.rs:239-362 /
 * su2_benchmark.rs:247-362 (k = 2): `series` holds k chains of n values
 * each.  Pooled mean = mean of chain means, pooled SE = sqrt(sum SE^2) / k,
 * aggregate ESS = sum of ESS, max pairwise z.  half_lag > 0 also applies the
 * stationarity gate per chain.  Per-chain arrays (k entries each) may be
 * NULL. */
int32_t ym_pool_chains(const double *series, int64_t k, int64_t n, int64_t max_lag, int64_t half_lag,
                       double *chain_mean, double *chain_se, double *chain_ess, double *chain_drift_z,
                       ym_pool_summary *out);

/* glueball_correlators.rs:111-172 periodic connected estimator over
 * nconf x T samples (row-major, one row per configuration):
 * raw(tau) = sum_c sum_t O_c(t) O_c(t+tau) / (nconf T), conn = raw - mean^2. */
int32_t ym_periodic_connected(const double *samples, int64_t nconf, int64_t T, double *mean, double *raw,
                              double *conn);
/* glueball_scalar_production.rs:383-443 connected_from_samples; omitted < 0
 * keeps every configuration, otherwise that configuration is left out. */
int32_t ym_connected_correlator(const double *samples, int64_t nconf, int64_t T, int64_t omitted, double *conn);
/* glueball_scalar_production.rs:445-470 correlator_evidence (full estimate
 * plus the nconf delete-one replicas) and the jackknife error of
 * glueball_effective_mass.rs:161-215, sqrt((N-1)/N sum_i (C_i - C_bar)^2),
 * per lag.  replicas (nconf x T) may be NULL. */
int32_t ym_jackknife_correlator(const double *samples, int64_t nconf, int64_t T, double *full, double *replicas,
                                double *jackknife_mean, double *jackknife_error);


/* ------------------------------------------------ configuration formats
 * Host-only (no device): they work on the AoS FP64 arrays that
 * ym_lattice_download returns and ym_lattice_upload takes. */
/* lib.rs:587-614 to_canonical_config: the canonical YM-SU2-CONFIG-V1 text of
 * an SU(2) configuration (aos: 4 doubles per link).  Fails with
 * YM_SU2_NON_UNIT when a link's |norm^2 - 1| exceeds 2e-12 and with
 * YM_LATTICE_TOO_LARGE above 2^20 links (the format's bound).  Writes
 * *length bytes (no terminating NUL) when buf holds them; buf == NULL only
 * reports *length. */
int32_t ym_su2_config_encode(const int64_t extents[4], const double *aos, char *buf, int64_t capacity,
                             int64_t *length);
/* lib.rs:616-694 from_canonical_config: extents from the header lines, then
 * (aos != NULL, capacity >= 4 * links) the links, each kept bit for bit;
 * errors in the reference's order (YM_INVALID_CONFIGURATION_ENCODING,
 * YM_LATTICE_TOO_LARGE, YM_INVALID_EXTENT, YM_SU2_NON_FINITE,
 * YM_SU2_NON_UNIT); text that is not byte-identical to the canonical
 * encoding of its links is rejected. */
int32_t ym_su2_config_decode(const char *text, int64_t length, int64_t extents[4], double *aos, int64_t capacity);
/* Binary checkpoint of a chain (no reference counterpart; it serves the
 * chain drivers su3_ensemble.rs:408-480 / su2_ensemble.rs:488-560 and the
 * production generator glueball_scalar_production.rs:257-297 for resume):
 * group (2, 3), storage precision (32, 64), extents, the chain key, the next
 * sweep index and the links in the storage precision, AoS order, with an
 * FNV-1a 64 checksum.  The sweeps draw from a counter-based stream keyed by
 * (key, sweep, site, direction, hit), so links + key + next sweep restore a
 * chain exactly.  ym_checkpoint_bytes gives the encoded size. */
int32_t ym_checkpoint_bytes(int32_t group, int32_t precision, const int64_t extents[4], int64_t *bytes);
int32_t ym_checkpoint_encode(int32_t group, int32_t precision, const int64_t extents[4], const double *aos,
                             uint64_t key, uint64_t next_sweep, void *buf, int64_t capacity, int64_t *length);
/* Validates magic, sizes and checksum (YM_INVALID_CONFIGURATION_ENCODING). */
int32_t ym_checkpoint_header(const void *buf, int64_t length, int32_t *group, int32_t *precision, int64_t extents[4],
                             uint64_t *key, uint64_t *next_sweep);
/* Links back into AoS FP64 (count = links x 4 or 18 doubles), exactly. */
int32_t ym_checkpoint_decode(const void *buf, int64_t length, double *aos, int64_t count);

/* ------------------------------------------- host FP64 correlator analysis
 * Replica correlators are passed as nrep row pointers with their lengths
 * (the reference's &[Vec<f64>]); every row must hold T values. */
typedef struct {
    int64_t tau;
    double full_estimate, jackknife_mean, standard_error;
} ym_effective_mass_point;
/* glueball_effective_mass.rs:161-232 cosh_effective_mass_with_jackknife:
 * m(tau) = acosh((C(tau-1) + C(tau+1)) / (2 C(tau))) at strictly increasing
 * lags 0 < tau < T/2 on the full correlator and on each delete-one replica;
 * standard error sqrt((N-1)/N sum_i (m_i - m_bar)^2).  4 <= T <= 64,
 * 3 <= nrep <= 64.  out holds nlags points. */
int32_t ym_cosh_effective_mass(const double *full, int64_t T, const double *const *replicas,
                               const int64_t *replica_lengths, int64_t nrep, const int64_t *lags, int64_t nlags,
                               ym_effective_mass_point *out);
/* glueball_effective_mass.rs:240-268 analytic_periodic_single_state:
 * out[tau] = A (exp(-m tau) + exp(-m (T - tau))), tau = 0..T-1. */
int32_t ym_analytic_periodic_single_state(int64_t T, double mass, double amplitude, double *out);

typedef struct {
    int64_t window_start, window_end;
    double plateau_weighted_effective_mass, plateau_maximum_z;
    double fitted_mass, fitted_amplitude, chi_square;
    int64_t degrees_of_freedom;
    double chi_square_per_dof;
    double mass_jackknife_mean, mass_standard_error;
    double amplitude_jackknife_mean, amplitude_standard_error;
    double maximum_off_diagonal_correlation;
} ym_scalar_fit_result;
/* glueball_scalar_fit.rs:156-650 correlated_scalar_fit: among nwin
 * predeclared windows (pairs start, end; 3..8 points, start >= 1,
 * 2 end < T, unique, at most 16) the longest whose effective masses stay
 * within 2.5 jackknife errors of their inverse-variance mean (ties: later
 * start) is fitted with C(tau) = A (exp(-m tau) + exp(-m (T - tau))) by
 * correlated chi^2 (jackknife covariance of the window, Gauss-Jordan
 * inverse) profiled over the mass grid 0.05:0.001:2.0; each replica is
 * refitted with the same inverse for the jackknife errors. */
int32_t ym_correlated_scalar_fit(const double *full, int64_t T, const double *const *replicas,
                                 const int64_t *replica_lengths, int64_t nrep, const int64_t *windows, int64_t nwin,
                                 ym_scalar_fit_result *out);
/* glueball_scalar_variational.rs:628-744 scalar_variational_connected_correlator
 * from the A1++ operator samples (as ym_su3_timeslice_a1pp returns them):
 * samples is [nconf][T][2] (plaquette, rectangle).  matrices receives T
 * symmetric 2 x 2 matrices (row-major, 4 doubles each),
 * C_ij(tau) = 1/2 [<O_i(t) O_j(t+tau)> + <O_j(t) O_i(t+tau)>] - <O_i><O_j>;
 * vacuum_means receives <O_0>, <O_1>.  1 <= nconf <= 64, 3 <= T <= 64. */
int32_t ym_variational_connected_correlator(const double *samples, int64_t nconf, int64_t T, double *vacuum_means,
                                            double *matrices);
/* glueball_scalar_variational.rs:746-925 principal_generalized_projection:
 * largest lambda of C(t1) v = lambda C(t0) v with (t0, t1) = (1, 2), v
 * normalised to v^T C(t0) v = 1 (first, else second, component positive);
 * projected[t] = v^T C(t) v for t = 0..T-1. */
int32_t ym_principal_generalized_projection(const double *matrices, int64_t T, int64_t t0, int64_t t1,
                                            double *eigenvalue, double *coefficients, double *projected);

/* su2_ensemble.rs:46-55 mix64 / derive_chain_seed and su3_ensemble.rs:189-192. */
uint64_t ym_derive_chain_seed(uint64_t master_seed, uint64_t chain_id);
uint64_t ym_derive_su3_chain_seed(uint64_t master_seed, uint64_t chain_id);

/* --------------------------------------------------------------- timing */
/* HIP events on the handle's stream around the next calls: begin records
 * the start event; end records the stop event, synchronises and returns the
 * elapsed milliseconds. */
int32_t ym_timer_begin(ym_lattice *lat);
int32_t ym_timer_end(ym_lattice *lat, float *ms);

/* Human-readable name of a status code (static storage). */
const char *ym_status_string(int32_t status);

#ifdef __cplusplus
}
#endif
#endif
