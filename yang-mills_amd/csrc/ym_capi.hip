// C ABI of libym_amd.so: lattice handles, sampler and observable entry
// points, the device chain driver and host FP64 statistics.  See
// include/ym_amd.h for the contract and the reference function each entry
// point replaces.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <algorithm>
#include <complex>
#include <cstring>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "../../include/ym_amd.h"
#include "ym_device.hpp"
#include "ym_internal.hpp"

namespace ym {

__global__ void k_fill_zero(double *p, size_t n) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0.0;
}
__global__ void k_fill_zero_u64(unsigned long long *p, size_t n) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0ull;
}
__global__ void k_fill_zero_i32(int *p, size_t n) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0;
}
void launch_fill_zero(double *p, size_t n, hipStream_t st) {
    if (n) hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n);
}
void launch_fill_zero_u64(unsigned long long *p, size_t n, hipStream_t st) {
    if (n) hipLaunchKernelGGL(k_fill_zero_u64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n);
}
void launch_fill_zero_i32(int *p, size_t n, hipStream_t st) {
    if (n) hipLaunchKernelGGL(k_fill_zero_i32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n);
}

// Device copy of a link field, 16 bytes per lane (a wavefront moves one
// contiguous 1 KB block per instruction, the access width of the link loads
// of the sweep kernels), grid-stride over n16 16-byte words.  Link fields are
// multiples of 16 bytes.  It also serves as the known-size reference that
// calibrates the L2 <-> fabric byte counters.
__global__ __launch_bounds__(256) void k_copy16(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n16) {
    const size_t stride = (size_t)gridDim.x * 256;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}
void launch_copy16(const void *src, void *dst, size_t bytes, hipStream_t st) {
    const size_t n16 = bytes / 16;
    if (!n16) return;
    const size_t blocks = std::min<size_t>((n16 + 255) / 256, 8192);
    hipLaunchKernelGGL(k_copy16, dim3((unsigned)blocks), dim3(256), 0, st, (const uint4 *)src, (uint4 *)dst, n16);
}

// >= one slot per wave of 256-thread blocks over Vh slots (the finest grid of
// the sweep kernels) and >= the 256 blocks of the wave-specialised kernel
int64_t tally_slots(const Geom &g) { return std::max<int64_t>(((int64_t)g.vh + 255) / 256 * 4, 1024); }

// One block: out[0] = sum of the n slots, eight independent loads in flight
// per thread.
__global__ __launch_bounds__(1024) void k_tally_sum(const unsigned long long *__restrict__ t, int64_t n,
                                                    unsigned long long *__restrict__ out) {
    __shared__ unsigned long long red[16];
    unsigned long long v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int64_t i = threadIdx.x;
    for (; i + 7 * 1024 < n; i += 8 * 1024) {
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] += t[i + j * 1024];
    }
    for (; i < n; i += 1024) v[0] += t[i];
    unsigned long long w = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) w += v[j];
    w = block_sum_u64<1024>(w, red);
    if (threadIdx.x == 0) out[0] = w;
}
void launch_tally_sum(const unsigned long long *tally, int64_t n, unsigned long long *out, hipStream_t st) {
    hipLaunchKernelGGL(k_tally_sum, dim3(1), dim3(1024), 0, st, tally, n, out);
}

int64_t stage_sites_for(int64_t volume) { return std::min<int64_t>(volume, (int64_t)1 << 16); }

int32_t upload_aos_staged(int group, int prec, void *links, const Geom &g, const double *aos, double *stage,
                          int64_t stage_sites, hipStream_t st) {
    const int nc = group == 3 ? 18 : 4;
    for (int64_t s0 = 0; s0 < g.volume; s0 += stage_sites) {
        const int64_t ns = std::min<int64_t>(stage_sites, g.volume - s0);
        // the copy into the staging buffer is ordered after the previous
        // chunk's kernel on the same stream
        if (hipMemcpyAsync(stage, aos + s0 * 4 * nc, sizeof(double) * (size_t)ns * 4 * nc, hipMemcpyHostToDevice,
                           st) != hipSuccess)
            return YM_HIP_ERROR;
        if (group == 3) launch_su3_from_aos_range(prec, links, g, stage, (int)s0, (int)ns, st);
        else launch_su2_from_aos_range(prec, links, g, stage, (int)s0, (int)ns, st);
        if (hipGetLastError() != hipSuccess) return YM_HIP_ERROR;
    }
    return hipStreamSynchronize(st) == hipSuccess ? YM_OK : YM_HIP_ERROR;
}

int32_t download_aos_staged(int group, int prec, const void *links, const Geom &g, double *aos, double *stage,
                            int64_t stage_sites, hipStream_t st) {
    const int nc = group == 3 ? 18 : 4;
    for (int64_t s0 = 0; s0 < g.volume; s0 += stage_sites) {
        const int64_t ns = std::min<int64_t>(stage_sites, g.volume - s0);
        if (group == 3) launch_su3_to_aos_range(prec, links, g, stage, (int)s0, (int)ns, st);
        else launch_su2_to_aos_range(prec, links, g, stage, (int)s0, (int)ns, st);
        if (hipGetLastError() != hipSuccess) return YM_HIP_ERROR;
        if (hipMemcpyAsync(aos + s0 * 4 * nc, stage, sizeof(double) * (size_t)ns * 4 * nc, hipMemcpyDeviceToHost,
                           st) != hipSuccess)
            return YM_HIP_ERROR;
    }
    return hipStreamSynchronize(st) == hipSuccess ? YM_OK : YM_HIP_ERROR;
}

double link_check_tolerance(int group, int prec) {
    if (prec == 32) return 0x1p-16;
    return group == 3 ? 5.0e-12 : 2.0e-12;
}

namespace {
// Su3::from_checked_matrix (su3.rs:242-253): finite entries, then
// max |U^+U - 1| and |UU^+ - 1| entrywise (Mat3::unitarity_residual,
// su3.rs:219-223), then |det U - 1| (determinant_residual, su3.rs:225-229).
int32_t check_su3_link(const double *m, double tol) {
    for (int c = 0; c < 18; ++c)
        if (!std::isfinite(m[c])) return YM_SU3_NON_FINITE_MATRIX;
    typedef std::complex<double> cd;
    cd u[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) u[i][j] = cd(m[(3 * i + j) * 2], m[(3 * i + j) * 2 + 1]);
    double r = 0.0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            cd a = 0.0, b = 0.0;
            for (int k = 0; k < 3; ++k) {
                a += std::conj(u[k][i]) * u[k][j];
                b += u[i][k] * std::conj(u[j][k]);
            }
            const double d = (i == j) ? 1.0 : 0.0;
            r = std::max(r, std::max(std::abs(a - d), std::abs(b - d)));
        }
    if (!(r <= tol)) return YM_SU3_NON_UNITARY;
    const cd det = u[0][0] * (u[1][1] * u[2][2] - u[1][2] * u[2][1]) -
                   u[0][1] * (u[1][0] * u[2][2] - u[1][2] * u[2][0]) +
                   u[0][2] * (u[1][0] * u[2][1] - u[1][1] * u[2][0]);
    if (!(std::abs(det - 1.0) <= tol)) return YM_SU3_DETERMINANT_NOT_ONE;
    return YM_OK;
}
// Su2 values (lib.rs:155-180): finite components, non-zero norm, unit norm
// and every component within 1 + tol.
int32_t check_su2_link(const double *q, double tol) {
    for (int c = 0; c < 4; ++c)
        if (!std::isfinite(q[c])) return YM_SU2_NON_FINITE;
    const double n2 = q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
    if (n2 == 0.0) return YM_SU2_ZERO_NORM;
    for (int c = 0; c < 4; ++c)
        if (std::fabs(q[c]) > 1.0 + tol) return YM_SU2_NON_UNIT;
    if (!(std::fabs(n2 - 1.0) <= tol)) return YM_SU2_NON_UNIT;
    return YM_OK;
}
}  // namespace

int32_t validate_aos(int group, const double *aos, int64_t nlinks, double tol) {
    const int nc = group == 3 ? 18 : 4;
    unsigned nt = std::thread::hardware_concurrency();
    nt = std::max(1u, std::min(nt, 16u));
    if (nlinks < (1 << 16)) nt = 1;
    std::vector<int64_t> first(nt, INT64_MAX);
    std::vector<int32_t> code(nt, YM_OK);
    auto work = [&](unsigned t) {
        const int64_t lo = nlinks * (int64_t)t / nt, hi = nlinks * (int64_t)(t + 1) / nt;
        for (int64_t i = lo; i < hi; ++i) {
            const int32_t c = group == 3 ? check_su3_link(aos + i * nc, tol) : check_su2_link(aos + i * nc, tol);
            if (c) {
                first[t] = i;
                code[t] = c;
                return;
            }
        }
    };
    std::vector<std::thread> pool;
    try {
        for (unsigned t = 1; t < nt; ++t) pool.emplace_back(work, t);
    } catch (...) {
        for (auto &th : pool) th.join();
        return YM_ALLOCATION_FAILED;
    }
    work(0);
    for (auto &th : pool) th.join();
    int64_t best = INT64_MAX;
    int32_t st = YM_OK;
    for (unsigned t = 0; t < nt; ++t)
        if (code[t] && first[t] < best) {
            best = first[t];
            st = code[t];
        }
    return st;
}

}  // namespace ym

struct ym_lattice {
    int group = 0, prec = 0;
    int device = 0;          // HIP device current when the handle was created
    int64_t n[4] = {0, 0, 0, 0};
    ym::Geom g{};
    int ncomp = 0;           // reals per link
    size_t link_bytes = 0;   // bytes of the link field
    void *links = nullptr;
    hipStream_t stream = nullptr;
    double *part = nullptr;  // per-block partial results
    size_t part_cap = 0;     // doubles
    double *dres = nullptr;  // small result buffer (64 doubles)
    double *hres = nullptr;  // pinned host copy of dres (64 doubles)
    unsigned long long *dctr = nullptr;  // 4 counters
    unsigned long long *tally = nullptr; // tally_slots(g) per-block counts of the sweep kernels
    int *dstatus = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t aux0 = nullptr, aux1 = nullptr;  // timing of the learned sweeps
    // Device memory created on first use and kept until destroy, so no
    // call after the first allocates: the host-layout staging buffer
    // (stage_sites sites of AoS links), a second link field (stout
    // ping-pong, autotuning) and a byte buffer for path steps and result
    // vectors longer than dres.
    double *stage = nullptr;
    int64_t stage_sites = 0;
    void *spare = nullptr;
    void *aux = nullptr;
    size_t aux_cap = 0;
};

namespace ym {
int32_t lattice_info(const ym_lattice *l, int32_t *group, int32_t *prec, int64_t ext[4]) {
    if (!l) return YM_INVALID_ARGUMENT;
    *group = l->group;
    *prec = l->prec;
    for (int d = 0; d < 4; ++d) ext[d] = l->n[d];
    return YM_OK;
}
void *lattice_links(const ym_lattice *l) { return l->links; }
unsigned long long *lattice_counters(const ym_lattice *l) { return l->dctr; }
unsigned long long *lattice_tally(const ym_lattice *l) { return l->tally; }
void lattice_aux_events(const ym_lattice *l, hipEvent_t *e0, hipEvent_t *e1) {
    *e0 = l->aux0;
    *e1 = l->aux1;
}
hipStream_t lattice_stream(const ym_lattice *l) { return l->stream; }
}  // namespace ym

namespace {

constexpr int64_t kMaxSites = (int64_t)1 << 30;

int32_t hip_status(hipError_t e) { return e == hipSuccess ? YM_OK : YM_HIP_ERROR; }

#define YM_HIP(call)                              \
    do {                                          \
        hipError_t e__ = (call);                  \
        if (e__ != hipSuccess) return YM_HIP_ERROR; \
    } while (0)

int32_t ensure_part(ym_lattice *l, size_t n) {
    if (n <= l->part_cap) return YM_OK;
    if (l->part) (void)hipFree(l->part);
    l->part = nullptr;
    l->part_cap = 0;
    if (hipMalloc((void **)&l->part, n * sizeof(double)) != hipSuccess) return YM_ALLOCATION_FAILED;
    l->part_cap = n;
    return YM_OK;
}

int32_t ensure_stage(ym_lattice *l) {
    if (l->stage) return YM_OK;
    const int64_t sites = ym::stage_sites_for(l->g.volume);
    if (hipMalloc((void **)&l->stage, sizeof(double) * (size_t)sites * 4 * l->ncomp) != hipSuccess) {
        l->stage = nullptr;
        return YM_ALLOCATION_FAILED;
    }
    l->stage_sites = sites;
    return YM_OK;
}

int32_t ensure_spare(ym_lattice *l) {
    if (l->spare) return YM_OK;
    if (hipMalloc(&l->spare, l->link_bytes) != hipSuccess) {
        l->spare = nullptr;
        return YM_ALLOCATION_FAILED;
    }
    return YM_OK;
}

int32_t ensure_aux(ym_lattice *l, size_t bytes) {
    if (bytes <= l->aux_cap) return YM_OK;
    if (l->aux) {
        (void)hipStreamSynchronize(l->stream);
        (void)hipFree(l->aux);
    }
    l->aux = nullptr;
    l->aux_cap = 0;
    const size_t cap = std::max(bytes, (size_t)4096);
    if (hipMalloc(&l->aux, cap) != hipSuccess) return YM_ALLOCATION_FAILED;
    l->aux_cap = cap;
    return YM_OK;
}

int32_t finish(ym_lattice *l) {
    YM_HIP(hipGetLastError());
    YM_HIP(hipStreamSynchronize(l->stream));
    return YM_OK;
}

bool all_even(const ym_lattice *l) {
    for (int d = 0; d < 4; ++d)
        if (l->n[d] % 2) return false;
    return true;
}

// ncols <= 64 results through the pinned buffer: a device-to-host copy into
// pageable memory is staged by the runtime and costs several microseconds
// more per measurement call.
int32_t sum_partials_from(ym_lattice *l, const double *part, int nblk, int ncols, double *host_out) {
    if (ncols < 1 || ncols > 64) return YM_INVALID_ARGUMENT;
    ym::launch_sum_partials(part, nblk, ncols, nblk, l->dres, l->stream);
    YM_HIP(hipGetLastError());
    YM_HIP(hipMemcpyAsync(l->hres, l->dres, sizeof(double) * ncols, hipMemcpyDeviceToHost, l->stream));
    const int32_t st = finish(l);
    if (st) return st;
    std::memcpy(host_out, l->hres, sizeof(double) * ncols);
    return YM_OK;
}
int32_t sum_partials(ym_lattice *l, int nblk, int ncols, double *host_out) {
    return sum_partials_from(l, l->part, nblk, ncols, host_out);
}

}  // namespace

namespace ym {
void *lattice_aux(const ym_lattice *lc, size_t bytes) {
    ym_lattice *l = const_cast<ym_lattice *>(lc);
    return ensure_aux(l, bytes) == YM_OK ? l->aux : nullptr;
}
}  // namespace ym

extern "C" {

const char *ym_status_string(int32_t s) {
    switch (s) {
    case YM_OK: return "ok";
    case YM_INVALID_EXTENT: return "every lattice extent must be >= 2";
    case YM_LATTICE_TOO_LARGE: return "lattice exceeds the device limit";
    case YM_ALLOCATION_FAILED: return "device allocation failed";
    case YM_DIRECTION_OUT_OF_RANGE: return "lattice direction must be in 0..4";
    case YM_DIRECTIONS_MUST_DIFFER: return "directions must differ";
    case YM_NON_FINITE_BETA: return "beta must be finite";
    case YM_NEGATIVE_BETA: return "beta must be non-negative";
    case YM_NON_FINITE_ACTION: return "Wilson action evaluation produced a non-finite result";
    case YM_ZERO_RECTANGLE_EXTENT: return "rectangular Wilson-loop extents must be positive";
    case YM_EMPTY_PATH: return "path is empty";
    case YM_PATH_TOO_LONG: return "path exceeds the bounded length";
    case YM_SU2_NON_FINITE: return "SU(2) components must be finite";
    case YM_SU2_NON_UNIT: return "serialized SU(2) components must satisfy the unit invariant";
    case YM_INVALID_CONFIGURATION_ENCODING: return "invalid configuration encoding";
    case YM_LATTICE_SIZE_OVERFLOW: return "lattice size arithmetic overflowed";
    case YM_COORDINATE_OUT_OF_RANGE: return "lattice coordinate is outside its periodic fundamental domain";
    case YM_BETA_OUT_OF_RANGE: return "beta is outside the bounded reference range";
    case YM_ANGLE_OUT_OF_RANGE: return "Metropolis proposal angle is outside the bounded reference range";
    case YM_MEASUREMENT_COUNT: return "measurement count is outside the accepted range";
    case YM_ZERO_STRIDE: return "measurement stride must be positive";
    case YM_LAG_OUT_OF_RANGE: return "autocorrelation lag must be smaller than the measurement count";
    case YM_SCHEDULE_OUT_OF_RANGE: return "schedule exceeds the accepted range";
    case YM_NON_FINITE_STATISTIC: return "calculation produced a non-finite value";
    case YM_TOO_MANY_OR: return "over-relaxation sweep count exceeds the bound";
    case YM_ZERO_ATTEMPT_LIMIT: return "heat-bath rejection attempt limit must be positive";
    case YM_HEATBATH_ATTEMPT_LIMIT: return "SU(2) heat-bath rejection sampler exhausted its bounded attempt limit";
    case YM_TIME_SLICE_OUT_OF_RANGE: return "Euclidean time slice is outside the lattice";
    case YM_SPATIAL_EXTENT_TOO_SMALL: return "every spatial extent must be at least 3";
    case YM_SPATIAL_NOT_CUBIC: return "spatial extents must be equal";
    case YM_SPATIAL_EXTENT_NOT_EVEN: return "stride-2 blocking requires even spatial extents";
    case YM_NON_FINITE_RHO: return "stout rho must be finite";
    case YM_RHO_OUT_OF_RANGE: return "stout rho must satisfy 0 < rho <= 0.25";
    case YM_TOO_MANY_LEVELS: return "stout level count exceeds the limit";
    case YM_NON_FINITE_OBSERVABLE: return "operator sample is non-finite";
    case YM_EMPTY_CONFIGURATION_SET: return "configuration set is empty";
    case YM_TOO_MANY_CONFIGURATIONS: return "too many configurations";
    case YM_TIME_EXTENT_TOO_LARGE: return "time extent exceeds the limit";
    case YM_NON_FINITE_ACCUMULATOR: return "correlator accumulator became non-finite";
    case YM_CORRELATOR_SHAPE: return "configuration rows must be non-empty and of one length";
    case YM_EFF_INVALID_LENGTH: return "series length must be in 2..=4096";
    case YM_EFF_INVALID_LAG: return "maximum lag must be in 1..min(n, 257)";
    case YM_EFF_INVALID_WEIGHT: return "log weights must be finite";
    case YM_EFF_INVALID_COST: return "costs must be finite, non-negative, with positive sampling cost";
    case YM_EFF_INVALID_DIAGNOSTIC: return "chain diagnostics rejected the series";
    case YM_EFF_DEGENERATE: return "series is constant";
    case YM_EFF_NON_FINITE: return "efficiency report is non-finite or non-positive";
    case YM_EM_TIME_EXTENT_TOO_SMALL: return "effective mass requires time extent >= 4";
    case YM_EM_TIME_EXTENT_TOO_LARGE: return "effective-mass time extent exceeds 64";
    case YM_EM_TOO_FEW_REPLICAS: return "at least three delete-one jackknife replicas are required";
    case YM_EM_TOO_MANY_REPLICAS: return "jackknife replica count exceeds 64";
    case YM_EM_EMPTY_WINDOW: return "analysis lag set must not be empty";
    case YM_EM_INVALID_LAG: return "lags must be strictly increasing and lie before the temporal midpoint";
    case YM_EM_SHAPE_MISMATCH: return "full and jackknife correlators must share one time extent";
    case YM_EM_NON_FINITE_CORRELATOR: return "correlator input must be finite";
    case YM_EM_NON_POSITIVE_CENTRE: return "cosh effective mass requires C(tau) > 0";
    case YM_EM_ACOSH_DOMAIN: return "cosh effective-mass acosh argument must be >= 1";
    case YM_EM_NON_FINITE_STATISTIC: return "effective-mass statistic became non-finite";
    case YM_EM_INVALID_FIXTURE: return "analytic fixture parameters are invalid";
    case YM_FIT_EMPTY_CANDIDATES: return "at least one candidate window is required";
    case YM_FIT_TOO_MANY_CANDIDATES: return "candidate-window count exceeds 16";
    case YM_FIT_INVALID_CANDIDATE: return "candidate windows must hold 3..8 points, start >= 1 and end before the midpoint";
    case YM_FIT_DUPLICATE_CANDIDATE: return "candidate windows must be unique";
    case YM_FIT_DEGENERATE_PLATEAU: return "plateau selection requires positive finite jackknife errors";
    case YM_FIT_NO_PLATEAU: return "no candidate window passed the plateau criterion";
    case YM_FIT_INSUFFICIENT_REPLICAS: return "more jackknife replicas than fitted time points are required";
    case YM_FIT_COVARIANCE_NOT_PD: return "jackknife covariance must be positive definite";
    case YM_FIT_SINGULAR_COVARIANCE: return "jackknife covariance inversion failed";
    case YM_FIT_NON_FINITE_STATISTIC: return "fit statistic became non-finite";
    case YM_FIT_NO_GRID_FIT: return "mass grid produced no admissible fit";
    case YM_VAR_TIME_EXTENT_TOO_SMALL: return "variational correlator requires at least three time slices";
    case YM_VAR_MATRIX_NOT_SYMMETRIC: return "correlator matrix violates the symmetry tolerance";
    case YM_VAR_INVALID_REFERENCE_TIMES: return "reference and diagonalization times must be 1 and 2, inside the series";
    case YM_VAR_REFERENCE_NOT_PD: return "reference correlator matrix is not positive definite";
    case YM_VAR_DEGENERATE_EIGENVECTOR: return "principal generalized eigenvector is not unique";
    case YM_VAR_NON_POSITIVE_EIGENVALUE: return "principal generalized eigenvalue is not positive";
    case YM_VAR_PROJECTION_NORMALIZATION: return "projected reference correlator is not normalized";
    case YM_WRONG_CONFIGURATION: return "schedule does not admit the half split";
    case YM_STATIONARITY_FAILED: return "stationarity gate failed (constant half or zero spread)";
    case YM_SU2_ZERO_NORM: return "SU(2) link has zero norm";
    case YM_SU3_NON_FINITE_MATRIX: return "SU(3) link has a non-finite entry";
    case YM_SU3_NON_UNITARY: return "SU(3) link violates the unitarity tolerance";
    case YM_SU3_DETERMINANT_NOT_ONE: return "SU(3) link determinant differs from one beyond the tolerance";
    case YM_KERNEL_HANDOFF_TIMEOUT: return "a wave of a sweep kernel waited past its bound for its partner wave";
    case YM_WORKSPACE_EXCEEDED: return "input exceeds the fixed device workspace of an allocation-free entry point";
    case YM_INVALID_ARGUMENT: return "invalid argument";
    case YM_NO_DEVICE: return "no HIP device";
    case YM_HIP_ERROR: return "HIP runtime error";
    case YM_UNSUPPORTED: return "operation not supported for this lattice";
    default: return "unknown status";
    }
}

uint64_t ym_derive_chain_seed(uint64_t master, uint64_t chain) {
    uint64_t v = master ^ ((chain + 1ull) * 0x9E3779B97F4A7C15ull);
    v = (v ^ (v >> 30)) * 0xBF58476D1CE4E5B9ull;
    v = (v ^ (v >> 27)) * 0x94D049BB133111EBull;
    return v ^ (v >> 31);
}
uint64_t ym_derive_su3_chain_seed(uint64_t master, uint64_t chain) {
    return ym_derive_chain_seed(master ^ 0x535533454e535631ull, chain);
}

int32_t ym_lattice_create(int32_t group, const int64_t extents[4], int32_t precision, ym_lattice **out) {
    if (!extents || !out) return YM_INVALID_ARGUMENT;
    *out = nullptr;
    if ((group != 2 && group != 3) || (precision != 32 && precision != 64)) return YM_INVALID_ARGUMENT;
    for (int d = 0; d < 4; ++d)
        if (extents[d] < 2) return YM_INVALID_EXTENT;
    int64_t vol = 1;
    for (int d = 0; d < 4; ++d) {
        if (extents[d] > kMaxSites) return YM_LATTICE_TOO_LARGE;
        vol *= extents[d];
        if (vol > kMaxSites) return YM_LATTICE_TOO_LARGE;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return YM_NO_DEVICE;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return YM_NO_DEVICE;
    ym_lattice *l = new (std::nothrow) ym_lattice();
    if (!l) return YM_ALLOCATION_FAILED;
    l->device = dev;
    l->group = group;
    l->prec = precision;
    for (int d = 0; d < 4; ++d) l->n[d] = extents[d];
    l->g = ym::make_geom(extents);
    l->ncomp = group == 3 ? 18 : 4;
    l->link_bytes = group == 3 ? ym::su3_pairs(l->g.vh) * 2 * (precision / 8)
                               : (size_t)8 * 4 * l->g.vh * (precision / 8);
    int32_t st = YM_OK;
    if (hipStreamCreateWithFlags(&l->stream, hipStreamNonBlocking) != hipSuccess) st = YM_HIP_ERROR;
    if (!st && hipMalloc(&l->links, l->link_bytes) != hipSuccess) st = YM_LATTICE_TOO_LARGE;
    if (!st && hipMalloc((void **)&l->dres, 64 * sizeof(double)) != hipSuccess) st = YM_ALLOCATION_FAILED;
    if (!st && hipHostMalloc((void **)&l->hres, 64 * sizeof(double), hipHostMallocDefault) != hipSuccess)
        st = YM_ALLOCATION_FAILED;
    if (!st && hipMalloc((void **)&l->dctr, 4 * sizeof(unsigned long long)) != hipSuccess) st = YM_ALLOCATION_FAILED;
    if (!st && hipMalloc((void **)&l->tally, (size_t)ym::tally_slots(l->g) * sizeof(unsigned long long)) != hipSuccess)
        st = YM_ALLOCATION_FAILED;
    if (!st && hipMalloc((void **)&l->dstatus, sizeof(int)) != hipSuccess) st = YM_ALLOCATION_FAILED;
    if (!st && (hipEventCreate(&l->ev0) != hipSuccess || hipEventCreate(&l->ev1) != hipSuccess)) st = YM_HIP_ERROR;
    if (!st && (hipEventCreate(&l->aux0) != hipSuccess || hipEventCreate(&l->aux1) != hipSuccess)) st = YM_HIP_ERROR;
    if (!st) st = ym_lattice_set_identity(l);
    if (st) {
        ym_lattice_destroy(l);
        return st;
    }
    *out = l;
    return YM_OK;
}

int32_t ym_lattice_destroy(ym_lattice *l) {
    if (!l) return YM_INVALID_ARGUMENT;
    if (l->stream) (void)hipStreamSynchronize(l->stream);
    if (l->links) (void)hipFree(l->links);
    if (l->part) (void)hipFree(l->part);
    if (l->dres) (void)hipFree(l->dres);
    if (l->hres) (void)hipHostFree(l->hres);
    if (l->dctr) (void)hipFree(l->dctr);
    if (l->tally) (void)hipFree(l->tally);
    if (l->dstatus) (void)hipFree(l->dstatus);
    if (l->ev0) (void)hipEventDestroy(l->ev0);
    if (l->ev1) (void)hipEventDestroy(l->ev1);
    if (l->aux0) (void)hipEventDestroy(l->aux0);
    if (l->aux1) (void)hipEventDestroy(l->aux1);
    if (l->stage) (void)hipFree(l->stage);
    if (l->spare) (void)hipFree(l->spare);
    if (l->aux) (void)hipFree(l->aux);
    if (l->stream) (void)hipStreamDestroy(l->stream);
    delete l;
    return YM_OK;
}

int32_t ym_lattice_bytes(const ym_lattice *l, int64_t *bytes) {
    if (!l || !bytes) return YM_INVALID_ARGUMENT;
    *bytes = (int64_t)l->link_bytes;
    return YM_OK;
}

int32_t ym_lattice_set_identity(ym_lattice *l) {
    if (!l) return YM_INVALID_ARGUMENT;
    if (l->group == 3) ym::launch_su3_identity(l->prec, l->links, l->g, l->stream);
    else ym::launch_su2_identity(l->prec, l->links, l->g, l->stream);
    return finish(l);
}

int32_t ym_lattice_hot_start(ym_lattice *l, uint64_t seed) {
    if (!l) return YM_INVALID_ARGUMENT;
    if (l->group == 3) ym::launch_su3_hot(l->prec, l->links, l->g, seed, l->stream);
    else ym::launch_su2_hot(l->prec, l->links, l->g, seed, l->stream);
    return finish(l);
}

int32_t ym_lattice_upload_unchecked(ym_lattice *l, const double *aos, int64_t count) {
    if (!l || !aos) return YM_INVALID_ARGUMENT;
    if (count != (int64_t)l->g.volume * 4 * l->ncomp) return YM_INVALID_ARGUMENT;
    int32_t st = ensure_stage(l);
    if (st) return st;
    return ym::upload_aos_staged(l->group, l->prec, l->links, l->g, aos, l->stage, l->stage_sites, l->stream);
}

int32_t ym_lattice_upload(ym_lattice *l, const double *aos, int64_t count) {
    if (!l || !aos) return YM_INVALID_ARGUMENT;
    if (count != (int64_t)l->g.volume * 4 * l->ncomp) return YM_INVALID_ARGUMENT;
    const int32_t st = ym::validate_aos(l->group, aos, count / l->ncomp, ym::link_check_tolerance(l->group, l->prec));
    if (st) return st;
    return ym_lattice_upload_unchecked(l, aos, count);
}

int32_t ym_lattice_download(const ym_lattice *lc, double *aos, int64_t count) {
    ym_lattice *l = const_cast<ym_lattice *>(lc);
    if (!l || !aos) return YM_INVALID_ARGUMENT;
    if (count != (int64_t)l->g.volume * 4 * l->ncomp) return YM_INVALID_ARGUMENT;
    int32_t st = ensure_stage(l);
    if (st) return st;
    return ym::download_aos_staged(l->group, l->prec, l->links, l->g, aos, l->stage, l->stage_sites, l->stream);
}

int32_t ym_lattice_copy(ym_lattice *dst, const ym_lattice *src) {
    if (!dst || !src) return YM_INVALID_ARGUMENT;
    if (dst->group != src->group || dst->prec != src->prec) return YM_INVALID_ARGUMENT;
    for (int d = 0; d < 4; ++d)
        if (dst->n[d] != src->n[d]) return YM_INVALID_ARGUMENT;
    YM_HIP(hipStreamSynchronize(src->stream));
    ym::launch_copy16(src->links, dst->links, src->link_bytes, dst->stream);
    YM_HIP(hipGetLastError());
    return finish(dst);
}

int32_t ym_su3_metropolis_sweeps(ym_lattice *l, double beta, double angle, int32_t hits, uint64_t key,
                                 uint64_t first_sweep, int32_t n_sweeps, int32_t reunitarize, uint64_t *proposals,
                                 uint64_t *accepts) {
    if (!l || l->group != 3 || hits < 1 || n_sweeps < 0) return YM_INVALID_ARGUMENT;
    if (!std::isfinite(beta) || beta < 0.0 || beta > 12.0) return YM_BETA_OUT_OF_RANGE;
    if (!std::isfinite(angle) || angle < 1.0e-8 || angle > 1.0) return YM_ANGLE_OUT_OF_RANGE;
    if (!all_even(l)) return YM_UNSUPPORTED;
    const int reunit = (l->prec == 32) ? 1 : (reunitarize != 0);
    const int64_t nt = ym::tally_slots(l->g);
    ym::launch_fill_zero_u64(l->tally, (size_t)nt, l->stream);
    ym::launch_fill_zero_i32(l->dstatus, 1, l->stream);
    for (int k = 0; k < n_sweeps; ++k)
        for (int mu = 0; mu < 4; ++mu)
            for (int p = 0; p < 2; ++p)
                ym::launch_su3_metropolis(l->prec, l->links, l->g, mu, p, beta, angle, hits, key,
                                          first_sweep + (uint64_t)k, reunit, l->tally, l->dstatus, l->stream);
    ym::launch_tally_sum(l->tally, nt, l->dctr, l->stream);
    YM_HIP(hipGetLastError());
    unsigned long long acc = 0;
    int status = 0;
    YM_HIP(hipMemcpyAsync(&acc, l->dctr, sizeof(acc), hipMemcpyDeviceToHost, l->stream));
    YM_HIP(hipMemcpyAsync(&status, l->dstatus, sizeof(int), hipMemcpyDeviceToHost, l->stream));
    int32_t st = finish(l);
    if (st) return st;
    if (status) return YM_KERNEL_HANDOFF_TIMEOUT;
    if (proposals) *proposals += (uint64_t)n_sweeps * (uint64_t)hits * 4ull * (uint64_t)l->g.volume;
    if (accepts) *accepts += acc;
    return YM_OK;
}

int32_t ym_su3_autotune(ym_lattice *l, double beta, double angle, int32_t hits, int32_t reps, int32_t *best_mode,
                        int32_t *best_strip, float *us_per_candidate) {
    if (!l || l->group != 3 || hits < 1 || reps < 1 || !best_mode || !best_strip) return YM_INVALID_ARGUMENT;
    if (!std::isfinite(beta) || beta < 0.0 || beta > 12.0) return YM_BETA_OUT_OF_RANGE;
    if (!std::isfinite(angle) || angle < 1.0e-8 || angle > 1.0) return YM_ANGLE_OUT_OF_RANGE;
    if (!all_even(l)) return YM_UNSUPPORTED;
    // Candidates: row tiles with x1-strip widths {0 (slot order), 2, 4, 8,
    // 16}, compact 2 x 2 tiles with X1-strip widths {1, 2, 4, 8}.  Every
    // candidate permutes independent links, so all produce the same links;
    // the sweeps run on a scratch copy of the field.
    const int mode[9] = {0, 0, 0, 0, 0, 1, 1, 1, 1};
    const int strip[9] = {0, 2, 4, 8, 16, 1, 2, 4, 8};
    int32_t st = ensure_spare(l);
    if (st) return st;
    void *scratch = l->spare;
    st = hip_status(hipMemcpyAsync(scratch, l->links, l->link_bytes, hipMemcpyDeviceToDevice, l->stream));
    float best = 0.0f;
    int bm = l->g.tmode, bs = l->g.t1;
    for (int c = 0; c < 9 && !st; ++c) {
        ym::Geom g = l->g;
        g.tmode = mode[c];
        g.t1 = strip[c];
        ym::set_divisors(g);
        const bool valid = mode[c] == 0 ? (strip[c] == 0 || (g.tk > 0 && g.n[1] % strip[c] == 0))
                                        : (g.tq > 0 && (g.n[1] / 2) % strip[c] == 0);
        if (!valid) {
            if (us_per_candidate) us_per_candidate[c] = -1.0f;
            continue;
        }
        // one untimed sweep, then `reps` timed sweeps
        for (int k = 0; k <= reps && !st; ++k) {
            if (k == 1) st = hip_status(hipEventRecord(l->ev0, l->stream));
            for (int mu = 0; mu < 4; ++mu)
                for (int p = 0; p < 2; ++p)
                    ym::launch_su3_metropolis(l->prec, scratch, g, mu, p, beta, angle, hits, 0x5EEDull, (uint64_t)k,
                                              1, l->tally, l->dstatus, l->stream);
        }
        float ms = 0.0f;
        if (!st) st = hip_status(hipEventRecord(l->ev1, l->stream));
        if (!st) st = hip_status(hipEventSynchronize(l->ev1));
        if (!st) st = hip_status(hipEventElapsedTime(&ms, l->ev0, l->ev1));
        const float us = ms * 1e3f / (8.0f * (float)reps);
        if (us_per_candidate) us_per_candidate[c] = us;
        if (!st && (best == 0.0f || us < best)) {
            best = us;
            bm = mode[c];
            bs = strip[c];
        }
    }
    // With at least 256 tiles the wave-specialised kernel runs; when a block
    // gets only one or two tiles its producer and consumer waves barely
    // overlap, so the one-link-per-thread kernel is timed too, with the best
    // mapping, and kept when it is faster.
    int kper = 0;
    if (us_per_candidate) us_per_candidate[9] = -1.0f;
    if (!st && (l->g.vh + 255) / 256 >= 256) {
        ym::Geom g = l->g;
        g.tmode = bm;
        g.t1 = bs;
        g.kper = 1;
        ym::set_divisors(g);
        for (int k = 0; k <= reps && !st; ++k) {
            if (k == 1) st = hip_status(hipEventRecord(l->ev0, l->stream));
            for (int mu = 0; mu < 4; ++mu)
                for (int p = 0; p < 2; ++p)
                    ym::launch_su3_metropolis(l->prec, scratch, g, mu, p, beta, angle, hits, 0x5EEDull, (uint64_t)k,
                                              1, l->tally, l->dstatus, l->stream);
        }
        float ms = 0.0f;
        if (!st) st = hip_status(hipEventRecord(l->ev1, l->stream));
        if (!st) st = hip_status(hipEventSynchronize(l->ev1));
        if (!st) st = hip_status(hipEventElapsedTime(&ms, l->ev0, l->ev1));
        const float us = ms * 1e3f / (8.0f * (float)reps);
        if (us_per_candidate) us_per_candidate[9] = us;
        if (!st && us < best) kper = 1;
    }
    if (!st) st = finish(l);
    if (!st) {
        l->g.tmode = bm;
        l->g.t1 = bs;
        l->g.kper = kper;
        ym::set_divisors(l->g);
        *best_mode = bm + 2 * kper;
        *best_strip = bs;
    }
    return st;
}

int32_t ym_su3_metropolis_sweeps_profiled(ym_lattice *l, double beta, double angle, int32_t hits, uint64_t key,
                                          uint64_t first_sweep, int32_t n_sweeps, int32_t reunitarize,
                                          uint64_t *proposals, uint64_t *accepts, float *total_ms,
                                          float *kernel_ms_sum, int32_t *launches) {
    if (!l || l->group != 3 || hits < 1 || n_sweeps < 1 || !total_ms || (!kernel_ms_sum != !launches))
        return YM_INVALID_ARGUMENT;
    if (!std::isfinite(beta) || beta < 0.0 || beta > 12.0) return YM_BETA_OUT_OF_RANGE;
    if (!std::isfinite(angle) || angle < 1.0e-8 || angle > 1.0) return YM_ANGLE_OUT_OF_RANGE;
    if (!all_even(l)) return YM_UNSUPPORTED;
    const int reunit = (l->prec == 32) ? 1 : (reunitarize != 0);
    const int nl = 8 * n_sweeps;
    // per-launch events only when per-launch times are asked for: an event
    // between two launches delays the second one
    const bool per_launch = kernel_ms_sum != nullptr;
    std::vector<hipEvent_t> ev(per_launch ? 2 * (size_t)nl : 2, nullptr);
    int32_t st = YM_OK;
    for (auto &e : ev)
        if (hipEventCreate(&e) != hipSuccess) st = YM_HIP_ERROR;
    if (!st) {
        const int64_t nt = ym::tally_slots(l->g);
        ym::launch_fill_zero_u64(l->tally, (size_t)nt, l->stream);
        ym::launch_fill_zero_i32(l->dstatus, 1, l->stream);
        int k = 0;
        for (int s = 0; s < n_sweeps; ++s)
            for (int mu = 0; mu < 4; ++mu)
                for (int p = 0; p < 2; ++p, ++k) {
                    if (per_launch || k == 0) (void)hipEventRecord(ev[per_launch ? 2 * k : 0], l->stream);
                    ym::launch_su3_metropolis(l->prec, l->links, l->g, mu, p, beta, angle, hits, key,
                                              first_sweep + (uint64_t)s, reunit, l->tally, l->dstatus,
                                              l->stream);
                    if (per_launch || k == nl - 1) (void)hipEventRecord(ev[per_launch ? 2 * k + 1 : 1], l->stream);
                }
        ym::launch_tally_sum(l->tally, nt, l->dctr, l->stream);
        unsigned long long acc = 0;
        int status = 0;
        st = hip_status(hipGetLastError());
        if (!st) st = hip_status(hipMemcpyAsync(&acc, l->dctr, sizeof(acc), hipMemcpyDeviceToHost, l->stream));
        if (!st) st = hip_status(hipMemcpyAsync(&status, l->dstatus, sizeof(int), hipMemcpyDeviceToHost, l->stream));
        if (!st) st = finish(l);
        if (!st && status) st = YM_KERNEL_HANDOFF_TIMEOUT;
        if (!st) {
            float sum = 0.f, ms = 0.f;
            for (int i = 0; per_launch && i < nl && !st; ++i) {
                st = hip_status(hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]));
                sum += ms;
            }
            if (!st) st = hip_status(hipEventElapsedTime(total_ms, ev[0], ev[ev.size() - 1]));
            if (per_launch) {
                *kernel_ms_sum = sum;
                *launches = nl;
            }
            if (proposals) *proposals += (uint64_t)n_sweeps * (uint64_t)hits * 4ull * (uint64_t)l->g.volume;
            if (accepts) *accepts += acc;
        }
    }
    for (auto &e : ev)
        if (e) (void)hipEventDestroy(e);
    return st;
}

int32_t ym_su2_heatbath_sweeps(ym_lattice *l, double beta, uint64_t key, uint64_t first_sweep, int32_t n_sweeps,
                               int32_t max_attempts, uint64_t *proposals, uint64_t *updates) {
    if (!l || l->group != 2 || n_sweeps < 0) return YM_INVALID_ARGUMENT;
    if (!std::isfinite(beta) || beta < 0.0 || beta > 8.0) return YM_BETA_OUT_OF_RANGE;
    if (max_attempts <= 0) return YM_ZERO_ATTEMPT_LIMIT;
    if (!all_even(l)) return YM_UNSUPPORTED;
    const int64_t nt = ym::tally_slots(l->g);
    ym::launch_fill_zero_u64(l->tally, (size_t)nt, l->stream);
    ym::launch_fill_zero_u64(l->dctr, 2, l->stream);
    ym::launch_fill_zero_i32(l->dstatus, 1, l->stream);
    for (int k = 0; k < n_sweeps; ++k)
        for (int mu = 0; mu < 4; ++mu)
            for (int p = 0; p < 2; ++p)
                ym::launch_su2_heatbath(l->prec, l->links, l->g, mu, p, beta, key, first_sweep + (uint64_t)k,
                                        max_attempts, l->tally, l->dctr, l->dstatus, l->stream);
    ym::launch_tally_sum(l->tally, nt, l->dctr, l->stream);
    YM_HIP(hipGetLastError());
    unsigned long long ctr[2] = {0, 0};
    int status = 0;
    YM_HIP(hipMemcpyAsync(ctr, l->dctr, sizeof(ctr), hipMemcpyDeviceToHost, l->stream));
    YM_HIP(hipMemcpyAsync(&status, l->dstatus, sizeof(int), hipMemcpyDeviceToHost, l->stream));
    int32_t st = finish(l);
    if (st) return st;
    // counter 1 holds the links that exhausted the attempt limit
    if (proposals) *proposals += ctr[0];
    if (updates) *updates += (uint64_t)n_sweeps * 4ull * (uint64_t)l->g.volume - ctr[1];
    return status ? YM_HEATBATH_ATTEMPT_LIMIT : YM_OK;
}

int32_t ym_su2_overrelax_sweeps(ym_lattice *l, int32_t n_sweeps) {
    if (!l || l->group != 2 || n_sweeps < 0) return YM_INVALID_ARGUMENT;
    if (!all_even(l)) return YM_UNSUPPORTED;
    for (int k = 0; k < n_sweeps; ++k)
        for (int mu = 0; mu < 4; ++mu)
            for (int p = 0; p < 2; ++p) ym::launch_su2_overrelax(l->prec, l->links, l->g, mu, p, l->stream);
    return finish(l);
}

int32_t ym_average_plaquette(const ym_lattice *lc, double *out) {
    ym_lattice *l = const_cast<ym_lattice *>(lc);
    if (!l || !out) return YM_INVALID_ARGUMENT;
    const int nblk = 2 * ((l->g.vh + 255) / 256);
    int32_t st = ensure_part(l, nblk);
    if (st) return st;
    ym::launch_plaquette(l->group, l->prec, l->links, l->g, l->part, l->stream);
    double sum = 0.0;
    st = sum_partials(l, nblk, 1, &sum);
    if (st) return st;
    const double mean = sum / (6.0 * (double)l->g.volume);
    if (!std::isfinite(mean)) return YM_NON_FINITE_STATISTIC;
    *out = mean;
    return YM_OK;
}

}  // extern "C"

// ------------------------------------------------ allocation-free variant
// Per-block partials of ym_average_plaquette_noalloc live in this array: the
// code object defines it in the device memory of every GPU the library runs
// on, where it stays until the process ends.  Its host-side name is only a
// stand-in; kernels get the device address that hipGetSymbolAddress returns
// for the current device, looked up once per device and cached below.
namespace {
constexpr int kNoallocPartials = 1 << 16;  // doubles: 512 KiB per GPU
constexpr int kMaxDevices = 64;
}  // namespace
static __device__ __align__(256) double scratch[kNoallocPartials];
namespace {
std::mutex g_scratch_mutex;
double *g_scratch_address[kMaxDevices] = {};

// Device address of `scratch` on device `dev`, which must be current.
int32_t scratch_address(int dev, double **out) {
    if (dev < 0 || dev >= kMaxDevices) return YM_UNSUPPORTED;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_address[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return YM_HIP_ERROR;
        g_scratch_address[dev] = (double *)p;
    }
    *out = g_scratch_address[dev];
    return YM_OK;
}
}  // namespace

extern "C" {

// ym_average_plaquette with its partials in `scratch`: no hipMalloc, no
// hipFree, no allocating runtime or library call.  Every partial that the
// sum reads is written by the plaquette kernel of the same call, so nothing
// has to be cleared first.  Calls that overlap on one GPU (two host threads,
// or work on two streams) share the array and race.
int32_t ym_average_plaquette_noalloc(const ym_lattice *lc, double *out) {
    ym_lattice *l = const_cast<ym_lattice *>(lc);
    if (!l || !out) return YM_INVALID_ARGUMENT;
    const int nblk = 2 * ((l->g.vh + 255) / 256);
    if (nblk > kNoallocPartials) return YM_WORKSPACE_EXCEEDED;
    int cur = 0;
    YM_HIP(hipGetDevice(&cur));
    if (cur != l->device) YM_HIP(hipSetDevice(l->device));
    double *part = nullptr;
    int32_t st = scratch_address(l->device, &part);
    if (!st) {
        ym::launch_plaquette(l->group, l->prec, l->links, l->g, part, l->stream);
        double sum = 0.0;
        st = sum_partials_from(l, part, nblk, 1, &sum);
        if (!st) {
            const double mean = sum / (6.0 * (double)l->g.volume);
            if (std::isfinite(mean)) *out = mean;
            else st = YM_NON_FINITE_STATISTIC;
        }
    }
    if (cur != l->device) (void)hipSetDevice(cur);
    return st;
}

int32_t ym_wilson_action(const ym_lattice *l, double beta, double *out) {
    if (!l || !out) return YM_INVALID_ARGUMENT;
    if (!std::isfinite(beta)) return YM_NON_FINITE_BETA;
    if (beta < 0.0) return YM_NEGATIVE_BETA;
    double p = 0.0;
    int32_t st = ym_average_plaquette(l, &p);
    if (st) return st;
    const double a = beta * (6.0 * (double)l->g.volume) * (1.0 - p);
    if (!std::isfinite(a)) return YM_NON_FINITE_ACTION;
    *out = a;
    return YM_OK;
}

int32_t ym_staple(const ym_lattice *lc, const int64_t x[4], int32_t mu, double *out) {
    ym_lattice *l = const_cast<ym_lattice *>(lc);
    if (!l || !x || !out) return YM_INVALID_ARGUMENT;
    if (mu < 0 || mu > 3) return YM_DIRECTION_OUT_OF_RANGE;
    int xi[4];
    for (int d = 0; d < 4; ++d) {
        if (x[d] < 0 || x[d] >= l->n[d]) return YM_COORDINATE_OUT_OF_RANGE;
        xi[d] = (int)x[d];
    }
    if (l->group == 3) ym::launch_su3_staple_probe(l->prec, l->links, l->g, xi, mu, l->dres, l->stream);
    else ym::launch_su2_staple_probe(l->prec, l->links, l->g, xi, mu, l->dres, l->stream);
    YM_HIP(hipGetLastError());
    YM_HIP(hipMemcpyAsync(out, l->dres, sizeof(double) * l->ncomp, hipMemcpyDeviceToHost, l->stream));
    return finish(l);
}

int32_t ym_path_holonomy(const ym_lattice *lc, const int64_t start[4], const int32_t *steps, int64_t nsteps,
                         double *out, int64_t end[4]) {
    ym_lattice *l = const_cast<ym_lattice *>(lc);
    if (!l || !start || !out || (nsteps > 0 && !steps)) return YM_INVALID_ARGUMENT;
    if (nsteps <= 0) return YM_EMPTY_PATH;
    if (nsteps > (1 << 20)) return YM_PATH_TOO_LONG;
    int xi[4];
    for (int d = 0; d < 4; ++d) {
        if (start[d] < 0 || start[d] >= l->n[d]) return YM_COORDINATE_OUT_OF_RANGE;
        xi[d] = (int)start[d];
    }
    for (int64_t k = 0; k < nsteps; ++k)
        if (steps[k] == 0 || steps[k] > 4 || steps[k] < -4) return YM_DIRECTION_OUT_OF_RANGE;
    // aux: 22 doubles of result (matrix + end point), then the steps
    int32_t st = ensure_aux(l, sizeof(double) * 22 + sizeof(int) * (size_t)nsteps);
    if (st) return st;
    double *dout = (double *)l->aux;
    int *dsteps = (int *)(dout + 22);
    st = hip_status(hipMemcpyAsync(dsteps, steps, sizeof(int) * nsteps, hipMemcpyHostToDevice, l->stream));
    if (!st) {
        ym::launch_path_holonomy(l->group, l->prec, l->links, l->g, xi, dsteps, (int)nsteps, dout, l->stream);
        st = hip_status(hipGetLastError());
    }
    double h[22];
    if (!st) st = hip_status(hipMemcpyAsync(h, dout, sizeof(h), hipMemcpyDeviceToHost, l->stream));
    if (!st) st = finish(l);
    if (st) return st;
    for (int e = 0; e < l->ncomp; ++e) out[e] = h[e];
    if (end)
        for (int d = 0; d < 4; ++d) end[d] = (int64_t)h[l->ncomp + d];
    return YM_OK;
}

int32_t ym_wilson_loops(const ym_lattice *lc, int32_t rmax, int32_t tmax, int32_t plane_set, double *re, double *im) {
    ym_lattice *l = const_cast<ym_lattice *>(lc);
    if (!l || !re || !im) return YM_INVALID_ARGUMENT;
    if (rmax < 1 || tmax < 1 || rmax > 64 || tmax > 64) return YM_ZERO_RECTANGLE_EXTENT;
    int planes[12][2], np = 0;
    for (int m = 0; m < 4; ++m)
        for (int n = 0; n < 4; ++n) {
            if (m == n) continue;
            const bool temporal = (m < 3 && n == 3), spatial = (m < 3 && n < 3);
            if (plane_set == 0 || (plane_set == 1 && temporal) || (plane_set == 2 && spatial)) {
                planes[np][0] = m;
                planes[np][1] = n;
                ++np;
            }
        }
    if (plane_set < 0 || plane_set > 2) return YM_INVALID_ARGUMENT;
    const int nblk = 2 * ((l->g.vh + 127) / 128);
    const int ncols = rmax * tmax * 2;
    int32_t st = ensure_part(l, (size_t)nblk * ncols);
    if (st) return st;
    ym::launch_fill_zero(l->part, (size_t)nblk * ncols, l->stream);
    for (int k = 0; k < np; ++k)
        ym::launch_wilson_loops(l->group, l->prec, l->links, l->g, planes[k][0], planes[k][1], rmax, tmax, l->part, l->stream);
    std::vector<double> sums(ncols);
    if (ncols > 64) {
        st = ensure_aux(l, sizeof(double) * ncols);
        if (st) return st;
        double *dout = (double *)l->aux;
        ym::launch_sum_partials(l->part, nblk, ncols, nblk, dout, l->stream);
        st = hip_status(hipMemcpyAsync(sums.data(), dout, sizeof(double) * ncols, hipMemcpyDeviceToHost, l->stream));
        if (!st) st = finish(l);
    } else {
        st = sum_partials(l, nblk, ncols, sums.data());
    }
    if (st) return st;
    const double denom = (double)np * (double)l->g.volume;
    for (int r = 0; r < rmax; ++r)
        for (int t = 0; t < tmax; ++t) {
            re[r * tmax + t] = sums[(r * tmax + t) * 2] / denom;
            im[r * tmax + t] = sums[(r * tmax + t) * 2 + 1] / denom;
        }
    return YM_OK;
}

int32_t ym_polyakov_loop(const ym_lattice *lc, int32_t dir, double *re, double *im, double *abs_mean) {
    ym_lattice *l = const_cast<ym_lattice *>(lc);
    if (!l || !re || !im || !abs_mean) return YM_INVALID_ARGUMENT;
    if (dir < 0 || dir > 3) return YM_DIRECTION_OUT_OF_RANGE;
    const int nbase = l->g.volume / l->g.n[dir];
    const int nblk = (nbase + 255) / 256;
    int32_t st = ensure_part(l, (size_t)nblk * 3);
    if (st) return st;
    ym::launch_polyakov(l->group, l->prec, l->links, l->g, dir, l->part, l->stream);
    double s[3];
    st = sum_partials(l, nblk, 3, s);
    if (st) return st;
    *re = s[0] / nbase;
    *im = s[1] / nbase;
    *abs_mean = s[2] / nbase;
    return YM_OK;
}

int32_t ym_su3_timeslice_a1pp(const ym_lattice *lc, double *plaq_t, double *rect_t) {
    ym_lattice *l = const_cast<ym_lattice *>(lc);
    if (!l || !plaq_t || !rect_t) return YM_INVALID_ARGUMENT;
    if (l->group != 3) return YM_UNSUPPORTED;
    for (int d = 0; d < 3; ++d)
        if (l->n[d] < 3) return YM_SPATIAL_EXTENT_TOO_SMALL;
    if (l->n[0] != l->n[1] || l->n[1] != l->n[2]) return YM_SPATIAL_NOT_CUBIC;
    const int nsp = l->g.n[0] * l->g.n[1] * l->g.n[2];
    const int nblk = (nsp + 255) / 256;
    const int T = l->g.n[3];
    const size_t npart = (size_t)nblk * 2 * T;
    int32_t st = ensure_part(l, npart + 2 * (size_t)l->g.volume);
    if (st) return st;
    st = ensure_aux(l, sizeof(double) * 2 * T);
    if (st) return st;
    double *dout = (double *)l->aux;
    ym::launch_a1pp(l->prec, l->links, l->g, l->part + npart, l->part, l->stream);
    ym::launch_sum_partials(l->part, nblk, 2 * T, nblk, dout, l->stream);
    std::vector<double> s(2 * T);
    st = hip_status(hipMemcpyAsync(s.data(), dout, sizeof(double) * 2 * T, hipMemcpyDeviceToHost, l->stream));
    if (!st) st = finish(l);
    if (st) return st;
    for (int t = 0; t < T; ++t) {
        plaq_t[t] = s[2 * t] / (3.0 * nsp);
        rect_t[t] = s[2 * t + 1] / (6.0 * nsp);
    }
    return YM_OK;
}

int32_t ym_su3_stout_spatial(ym_lattice *dst, const ym_lattice *src, double rho, int32_t levels) {
    if (!dst || !src) return YM_INVALID_ARGUMENT;
    if (dst->group != 3 || src->group != 3 || dst->prec != src->prec) return YM_INVALID_ARGUMENT;
    for (int d = 0; d < 4; ++d)
        if (dst->n[d] != src->n[d]) return YM_INVALID_ARGUMENT;
    if (!std::isfinite(rho)) return YM_NON_FINITE_RHO;
    if (rho <= 0.0 || rho > 0.25) return YM_RHO_OUT_OF_RANGE;
    if (levels < 0 || levels > 64) return YM_TOO_MANY_LEVELS;
    for (int d = 0; d < 3; ++d)
        if (src->n[d] < 3) return YM_SPATIAL_EXTENT_TOO_SMALL;
    if (src->n[0] != src->n[1] || src->n[1] != src->n[2]) return YM_SPATIAL_NOT_CUBIC;
    if (dst == src) return YM_INVALID_ARGUMENT;
    YM_HIP(hipStreamSynchronize(src->stream));
    if (levels == 0) return ym_lattice_copy(dst, src);
    if (levels > 1) {
        const int32_t e = ensure_spare(dst);
        if (e) return e;
    }
    void *tmp = dst->spare;
    // The kernels write the spatial links only: the temporal (mu = 3) groups,
    // unchanged by smearing, are copied once into each output buffer (groups
    // (3, 0) and (3, 1) are the last quarter of the tiled field).
    const size_t quarter = src->link_bytes / 4;
    YM_HIP(hipMemcpyAsync((char *)dst->links + 3 * quarter, (const char *)src->links + 3 * quarter, quarter,
                          hipMemcpyDeviceToDevice, dst->stream));
    if (levels > 1)
        YM_HIP(hipMemcpyAsync((char *)tmp + 3 * quarter, (const char *)src->links + 3 * quarter, quarter,
                              hipMemcpyDeviceToDevice, dst->stream));
    // Level k writes dst when (levels-1-k) is even, tmp otherwise, so the
    // buffers alternate and the last level lands in dst.
    const void *cur = src->links;
    for (int k = 0; k < levels; ++k) {
        void *target = ((levels - 1 - k) % 2 == 0) ? dst->links : tmp;
        ym::launch_stout(src->prec, cur, target, src->g, rho, dst->stream);
        cur = target;
    }
    return finish(dst);
}

int32_t ym_su3_blocked_extents(const ym_lattice *src, int64_t coarse[4]) {
    if (!src || !coarse) return YM_INVALID_ARGUMENT;
    if (src->group != 3) return YM_UNSUPPORTED;
    // glueball_blocking.rs:71-94, checks in the reference's order.
    if (src->n[0] != src->n[1] || src->n[1] != src->n[2]) return YM_SPATIAL_NOT_CUBIC;
    if (src->n[0] < 6) return YM_SPATIAL_EXTENT_TOO_SMALL;
    if (src->n[0] % 2 != 0) return YM_SPATIAL_EXTENT_NOT_EVEN;
    for (int d = 0; d < 3; ++d) coarse[d] = src->n[d] / 2;
    coarse[3] = src->n[3];
    return YM_OK;
}

int32_t ym_su3_spatial_block(ym_lattice *dst, const ym_lattice *src) {
    if (!dst || !src || dst == src) return YM_INVALID_ARGUMENT;
    int64_t c[4];
    int32_t st = ym_su3_blocked_extents(src, c);
    if (st) return st;
    if (dst->group != 3 || dst->prec != src->prec) return YM_INVALID_ARGUMENT;
    for (int d = 0; d < 4; ++d)
        if (dst->n[d] != c[d]) return YM_INVALID_ARGUMENT;
    YM_HIP(hipStreamSynchronize(src->stream));
    ym::launch_block2(src->prec, src->links, dst->links, src->g, dst->g, dst->stream);
    YM_HIP(hipGetLastError());
    return finish(dst);
}

int32_t ym_link_residuals(const ym_lattice *lc, double *unit, double *det) {
    ym_lattice *l = const_cast<ym_lattice *>(lc);
    if (!l || !unit || !det) return YM_INVALID_ARGUMENT;
    const int nblk = 8 * ((l->g.vh + 255) / 256);
    int32_t st = ensure_part(l, (size_t)nblk * 2);
    if (st) return st;
    if (l->group == 3) ym::launch_su3_residuals(l->prec, l->links, l->g, l->part, l->stream);
    else ym::launch_su2_residual(l->prec, l->links, l->g, l->part, l->stream);
    ym::launch_max_partials(l->part, nblk, 2, 2, l->dres, l->stream);
    YM_HIP(hipGetLastError());
    YM_HIP(hipMemcpyAsync(l->hres, l->dres, 2 * sizeof(double), hipMemcpyDeviceToHost, l->stream));
    st = finish(l);
    if (st) return st;
    *unit = l->hres[0];
    *det = l->hres[1];
    return YM_OK;
}

int32_t ym_reunitarize(ym_lattice *l) {
    if (!l) return YM_INVALID_ARGUMENT;
    if (l->group == 3) {
        ym::launch_su3_reunit(l->prec, l->links, l->g, l->stream);
        return finish(l);
    }
    ym::launch_su2_normalize(l->prec, l->links, l->g, l->stream);
    return finish(l);
}

int32_t ym_timer_begin(ym_lattice *l) {
    if (!l) return YM_INVALID_ARGUMENT;
    YM_HIP(hipEventRecord(l->ev0, l->stream));
    return YM_OK;
}
int32_t ym_timer_end(ym_lattice *l, float *ms) {
    if (!l || !ms) return YM_INVALID_ARGUMENT;
    YM_HIP(hipEventRecord(l->ev1, l->stream));
    YM_HIP(hipEventSynchronize(l->ev1));
    YM_HIP(hipEventElapsedTime(ms, l->ev0, l->ev1));
    return YM_OK;
}

// su2_ensemble.rs:375-463: mean, variance (/N), autocorrelation with /(N-lag),
// positive-window tau = 1/2 + sum of the leading positive rho, ESS =
// min(N, N / 2 tau), SE = sqrt(var / ESS); sums in index order.
int32_t ym_diagnostics(const double *v, int64_t n, int64_t max_lag, ym_chain_diagnostics *out, double *autocorr) {
    if (!v || !out) return YM_INVALID_ARGUMENT;
    if (n < 2) return YM_MEASUREMENT_COUNT;
    if (max_lag < 0 || max_lag >= n) return YM_LAG_OUT_OF_RANGE;
    for (int64_t i = 0; i < n; ++i)
        if (!std::isfinite(v[i])) return YM_NON_FINITE_STATISTIC;
    const double nd = (double)n;
    double s = -0.0;
    for (int64_t i = 0; i < n; ++i) s += v[i];
    const double mean = s / nd;
    double q = -0.0;
    for (int64_t i = 0; i < n; ++i) {
        const double d = v[i] - mean;
        q += d * d;
    }
    const double var = q / nd;
    if (!std::isfinite(mean) || !std::isfinite(var)) return YM_NON_FINITE_STATISTIC;
    out->mean = mean;
    out->variance = var;
    if (autocorr) autocorr[0] = 1.0;
    if (var == 0.0) {
        if (autocorr)
            for (int64_t k = 1; k <= max_lag; ++k) autocorr[k] = 0.0;
        out->positive_window_lag = 0;
        out->tau_int = 0.5;
        out->ess = nd;
        out->standard_error = 0.0;
        return YM_OK;
    }
    double tau = 0.5;
    int64_t wlag = 0;
    bool open = true;
    for (int64_t lag = 1; lag <= max_lag; ++lag) {
        const int64_t pairs = n - lag;
        double c = -0.0;
        for (int64_t i = 0; i < pairs; ++i) c += (v[i] - mean) * (v[i + lag] - mean);
        const double rho = (c / (double)pairs) / var;
        if (!std::isfinite(rho)) return YM_NON_FINITE_STATISTIC;
        if (autocorr) autocorr[lag] = rho;
        if (open && rho > 0.0) {
            tau += rho;
            wlag = lag;
        } else {
            open = false;
        }
    }
    if (!std::isfinite(tau) || tau < 0.5) return YM_NON_FINITE_STATISTIC;
    const double ess = std::fmin(nd / (2.0 * tau), nd);
    const double se = std::sqrt(var / ess);
    if (!std::isfinite(ess) || !std::isfinite(se)) return YM_NON_FINITE_STATISTIC;
    out->positive_window_lag = wlag;
    out->tau_int = tau;
    out->ess = ess;
    out->standard_error = se;
    return YM_OK;
}

// Sweeps counted in int64 by the chain drivers run in int32-sized calls.
static int32_t su3_sweeps64(ym_lattice *l, const ym_su3_chain_config *c, uint64_t key, uint64_t first, int64_t n,
                            uint64_t *prop, uint64_t *acc) {
    int32_t st = YM_OK;
    for (int64_t done = 0; done < n && !st;) {
        const int32_t k = (int32_t)std::min<int64_t>(n - done, (int64_t)1 << 30);
        st = ym_su3_metropolis_sweeps(l, c->beta, c->proposal_angle, c->hits, key, first + (uint64_t)done, k,
                                      c->reunitarize, prop, acc);
        done += k;
    }
    return st;
}

int32_t ym_su3_run_chain(const ym_su3_chain_config *c, double *meas, ym_su3_chain_report *rep) {
    if (!c || !meas || !rep) return YM_INVALID_ARGUMENT;
    if (!std::isfinite(c->beta) || c->beta < 0.0 || c->beta > 12.0) return YM_BETA_OUT_OF_RANGE;
    if (!std::isfinite(c->proposal_angle) || c->proposal_angle < 1.0e-8 || c->proposal_angle > 1.0) return YM_ANGLE_OUT_OF_RANGE;
    if (c->measurements < 2) return YM_MEASUREMENT_COUNT;
    if (c->sweeps_between_measurements <= 0) return YM_ZERO_STRIDE;
    if (c->autocorrelation_max_lag < 0 || c->autocorrelation_max_lag >= c->measurements) return YM_LAG_OUT_OF_RANGE;
    if (c->thermalization_sweeps < 0 || c->hits < 1) return YM_SCHEDULE_OUT_OF_RANGE;
    // the sweep index (a Philox counter word) must not wrap: total < 2^62
    const int64_t lim = (int64_t)1 << 62;
    if (c->sweeps_between_measurements > lim / c->measurements ||
        c->thermalization_sweeps > lim - c->measurements * c->sweeps_between_measurements)
        return YM_SCHEDULE_OUT_OF_RANGE;
    ym_lattice *l = nullptr;
    int32_t st = ym_lattice_create(3, c->extents, c->precision, &l);
    if (st) return st;
    const uint64_t key = ym_derive_su3_chain_seed(c->master_seed, c->chain_id);
    if (c->hot_start) st = ym_lattice_hot_start(l, key);
    uint64_t prop = 0, acc = 0, sweep = 0;
    if (!st) st = su3_sweeps64(l, c, key, sweep, c->thermalization_sweeps, &prop, &acc);
    sweep += (uint64_t)c->thermalization_sweeps;
    for (int64_t m = 0; !st && m < c->measurements; ++m) {
        st = su3_sweeps64(l, c, key, sweep, c->sweeps_between_measurements, &prop, &acc);
        sweep += (uint64_t)c->sweeps_between_measurements;
        if (!st) st = ym_average_plaquette(l, &meas[m]);
        if (!st && (meas[m] < -0.5 - 1e-12 || meas[m] > 1.0 + 1e-12)) st = YM_NON_FINITE_STATISTIC;
    }
    if (!st) st = ym_diagnostics(meas, c->measurements, c->autocorrelation_max_lag, &rep->diagnostics, nullptr);
    if (!st) st = ym_link_residuals(l, &rep->final_unitarity_residual, &rep->final_determinant_residual);
    if (!st) {
        rep->derived_seed = key;
        rep->total_sweeps = (int64_t)sweep;
        rep->proposals = prop;
        rep->accepts = acc;
        rep->mean_acceptance = prop ? (double)acc / (double)prop : 0.0;
    }
    ym_lattice_destroy(l);
    return st;
}

int32_t ym_su2_markov_sweeps(ym_lattice *l, double beta, uint64_t key, uint64_t first_sweep, int32_t n_sweeps,
                             int32_t overrelaxation_sweeps_per_heatbath, int32_t max_attempts, uint64_t *proposals,
                             uint64_t *updates) {
    if (!l || l->group != 2 || n_sweeps < 0) return YM_INVALID_ARGUMENT;
    if (!std::isfinite(beta) || beta < 0.0 || beta > 8.0) return YM_BETA_OUT_OF_RANGE;
    if (overrelaxation_sweeps_per_heatbath < 0 || overrelaxation_sweeps_per_heatbath > 16) return YM_TOO_MANY_OR;
    if (max_attempts <= 0) return YM_ZERO_ATTEMPT_LIMIT;
    if (!all_even(l)) return YM_UNSUPPORTED;
    const int64_t nt = ym::tally_slots(l->g);
    ym::launch_fill_zero_u64(l->tally, (size_t)nt, l->stream);
    ym::launch_fill_zero_u64(l->dctr, 2, l->stream);
    ym::launch_fill_zero_i32(l->dstatus, 1, l->stream);
    for (int k = 0; k < n_sweeps; ++k) {
        for (int mu = 0; mu < 4; ++mu)
            for (int p = 0; p < 2; ++p)
                ym::launch_su2_heatbath(l->prec, l->links, l->g, mu, p, beta, key, first_sweep + (uint64_t)k,
                                        max_attempts, l->tally, l->dctr, l->dstatus, l->stream);
        for (int o = 0; o < overrelaxation_sweeps_per_heatbath; ++o)
            for (int mu = 0; mu < 4; ++mu)
                for (int p = 0; p < 2; ++p) ym::launch_su2_overrelax(l->prec, l->links, l->g, mu, p, l->stream);
    }
    ym::launch_tally_sum(l->tally, nt, l->dctr, l->stream);
    YM_HIP(hipGetLastError());
    unsigned long long ctr[2] = {0, 0};
    int status = 0;
    YM_HIP(hipMemcpyAsync(ctr, l->dctr, sizeof(ctr), hipMemcpyDeviceToHost, l->stream));
    YM_HIP(hipMemcpyAsync(&status, l->dstatus, sizeof(int), hipMemcpyDeviceToHost, l->stream));
    const int32_t st = finish(l);
    if (st) return st;
    // counter 1 holds the links that exhausted the attempt limit
    if (proposals) *proposals += ctr[0];
    if (updates) *updates += (uint64_t)n_sweeps * 4ull * (uint64_t)l->g.volume - ctr[1];
    return status ? YM_HEATBATH_ATTEMPT_LIMIT : YM_OK;
}

int32_t ym_su2_run_chain(const ym_su2_chain_config *c, double *meas, ym_su2_chain_report *rep) {
    if (!c || !meas || !rep) return YM_INVALID_ARGUMENT;
    // EnsembleConfig::validate (su2_ensemble.rs:70-103), in its order
    if (!std::isfinite(c->beta) || c->beta < 0.0 || c->beta > 8.0) return YM_BETA_OUT_OF_RANGE;
    const int64_t max_sweeps = (int64_t)1 << 20;
    if (c->measurements < 0 || c->sweeps_between_measurements < 0 || c->thermalization_sweeps < 0 ||
        c->overrelaxation_sweeps_per_heatbath < 0 || c->max_heatbath_attempts_per_link < 0 ||
        c->autocorrelation_max_lag < 0)
        return YM_SCHEDULE_OUT_OF_RANGE;
    if (c->measurements > max_sweeps) return YM_SCHEDULE_OUT_OF_RANGE;
    if (c->sweeps_between_measurements > 0 && c->measurements > max_sweeps / c->sweeps_between_measurements)
        return YM_SCHEDULE_OUT_OF_RANGE;
    if (c->thermalization_sweeps > max_sweeps - c->measurements * c->sweeps_between_measurements)
        return YM_SCHEDULE_OUT_OF_RANGE;
    if (c->measurements < 2) return YM_MEASUREMENT_COUNT;
    if (c->sweeps_between_measurements == 0) return YM_ZERO_STRIDE;
    if (c->overrelaxation_sweeps_per_heatbath > 16) return YM_TOO_MANY_OR;
    if (c->max_heatbath_attempts_per_link == 0) return YM_ZERO_ATTEMPT_LIMIT;
    if (c->autocorrelation_max_lag >= c->measurements) return YM_LAG_OUT_OF_RANGE;
    if (c->precision != 32 && c->precision != 64) return YM_INVALID_ARGUMENT;
    // a link needing more than 2^31 - 1 attempts is not a realistic event; the
    // kernels count attempts in int32
    const int32_t max_att = (int32_t)std::min<int64_t>(c->max_heatbath_attempts_per_link, 0x7fffffff);
    const int32_t nor = (int32_t)c->overrelaxation_sweeps_per_heatbath;
    ym_lattice *l = nullptr;
    int32_t st = ym_lattice_create(2, c->extents, c->precision, &l);
    if (st) return st;
    const uint64_t key = ym_derive_chain_seed(c->master_seed, c->chain_id);
    if (c->hot_start) st = ym_lattice_hot_start(l, key);
    uint64_t prop = 0, upd = 0, sweep = 0;
    if (!st && c->thermalization_sweeps > 0)
        st = ym_su2_markov_sweeps(l, c->beta, key, sweep, (int32_t)c->thermalization_sweeps, nor, max_att, &prop, &upd);
    sweep += (uint64_t)c->thermalization_sweeps;
    for (int64_t m = 0; !st && m < c->measurements; ++m) {
        st = ym_su2_markov_sweeps(l, c->beta, key, sweep, (int32_t)c->sweeps_between_measurements, nor, max_att,
                                  &prop, &upd);
        sweep += (uint64_t)c->sweeps_between_measurements;
        if (!st) st = ym_average_plaquette(l, &meas[m]);
    }
    if (!st) st = ym_diagnostics(meas, c->measurements, c->autocorrelation_max_lag, &rep->diagnostics, nullptr);
    double unit = 0.0, det = 0.0;
    if (!st) st = ym_link_residuals(l, &unit, &det);
    if (!st) {
        const double acc = (double)upd / (double)prop;
        if (!std::isfinite(acc) || !std::isfinite(unit)) st = YM_NON_FINITE_STATISTIC;
        rep->derived_seed = key;
        rep->total_heatbath_sweeps = (int64_t)sweep;
        rep->total_overrelaxation_sweeps = (int64_t)sweep * nor;
        rep->heatbath_proposals = prop;
        rep->heatbath_updates = upd;
        rep->mean_heatbath_acceptance = acc;
        rep->final_link_unit_residual = unit;
    }
    ym_lattice_destroy(l);
    return st;
}

}  // extern "C"
