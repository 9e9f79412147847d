// SU(2) heat-bath and over-relaxation kernels for gfx950.
//
// Each kernel is one checkerboard sub-pass (direction MU, parity p); one
// thread owns one link.  The heat-bath follows the reference sampler
// (su2_ensemble.rs:220-281): alpha = beta |V|, r0 from the truncated
// exponential by inverse CDF, accepted with probability sqrt(1 - r0^2), a
// uniform direction on S^2, U <- R V^dagger / |V|.  Attempt a of a link uses
// Philox(site, mu | HB_ATTEMPT << 2 | a << 5, sweep), so the rejection loop
// needs no sequential stream; a link that exhausts `max_attempts` keeps its
// value and raises the status flag (HeatbathAttemptLimitExceeded).
// Over-relaxation is U <- V^dagger U^dagger V^dagger / |V|^2.  Products of
// stored links are in the storage precision; alpha, the sampler and the new
// link are FP64, normalised before the single rounding to storage.
#include "ym_device.hpp"
#include "ym_internal.hpp"

namespace ym {

template <typename Real, int MU>
__device__ __forceinline__ Q4<Real> su2_staple(const typename Vec4<Real>::type *__restrict__ L, const Geom &g,
                                               const int x[4], int p, int s) {
    const int q = p ^ 1;
    Q4<Real> A;
#pragma unroll
    for (int c = 0; c < 4; ++c) A.c[c] = 0;
    int xpm[4];
    shifted(g, x, MU, +1, xpm);
    const int s_pm = slot_of(g, xpm);
#pragma unroll
    for (int nu = 0; nu < 4; ++nu) {
        if (nu == MU) continue;
        int xpn[4], xmn[4], xpmmn[4];
        shifted(g, x, nu, +1, xpn);
        shifted(g, x, nu, -1, xmn);
        shifted(g, xpm, nu, -1, xpmmn);
        const int s_pn = slot_of(g, xpn), s_mn = slot_of(g, xmn), s_pmmn = slot_of(g, xpmmn);
        {
            Q4<Real> X = load2<Real>(L, g.vh, nu, q, s_pm);
            Q4<Real> Y = load2<Real>(L, g.vh, MU, q, s_pn);
            Q4<Real> Z = load2<Real>(L, g.vh, nu, p, s);
            Q4<Real> R = qmul(qmul(X, qdag(Y)), qdag(Z));
#pragma unroll
            for (int c = 0; c < 4; ++c) A.c[c] = A.c[c] + R.c[c];
        }
        {
            Q4<Real> X = load2<Real>(L, g.vh, nu, p, s_pmmn);
            Q4<Real> Y = load2<Real>(L, g.vh, MU, q, s_mn);
            Q4<Real> Z = load2<Real>(L, g.vh, nu, q, s_mn);
            Q4<Real> R = qmul(qmul(qdag(X), qdag(Y)), Z);
#pragma unroll
            for (int c = 0; c < 4; ++c) A.c[c] = A.c[c] + R.c[c];
        }
    }
    return A;
}

// One heat-bath attempt (su2_ensemble.rs:229-250): r0 = 2 u1 - 1 when
// alpha <= 1e-300 (inv_alpha = 0), else r0 = 1 + ln(1 - u1 tm) / alpha with
// tm = 1 - e^{-2 alpha}, the logarithm as log1p(-u1 tm) for small alpha
// (inv_alpha > 2) and as the log of the fused 1 - u1 tm otherwise (that
// argument is >= e^{-1} there, so its rounding error is harmless); accepted
// iff u2^2 <= 1 - r0^2, the reference's u2 <= sqrt(1 - r0^2) without the
// square root.  No division: 1/alpha is formed once per link.
__device__ __forceinline__ bool hb_attempt(uint64_t key, uint32_t site, int mu, int att, uint64_t sweep,
                                           double inv_alpha, double tm, double &r0) {
    uint32_t o[4];
    draw(key, site, mu, KIND_HB_ATTEMPT, (uint32_t)att, sweep, o);
    const double u1 = unit53(o[0], o[1]), u2 = unit53(o[2], o[3]);
    if (inv_alpha == 0.0) {
        r0 = 2.0 * u1 - 1.0;
    } else {
        const double L = inv_alpha > 2.0 ? log1p_dev(-u1 * tm) : log_pos(__builtin_fma(-u1, tm, 1.0));
        r0 = __builtin_fma(L, inv_alpha, 1.0);
        r0 = r0 < -1.0 ? -1.0 : (r0 > 1.0 ? 1.0 : r0);
    }
    return u2 * u2 <= __builtin_fma(-r0, r0, 1.0);
}

// Attempts per lane and round of the cooperative rejection loop.
constexpr int kHbAttemptsPerLane = 2;

// One heat-bath sub-pass: thread = link, staple from the neighbours, then the
// rejection loop run by the whole wavefront together.  Each round the P links
// still without an accepted attempt get k = 64 / 2^ceil(log2 P) lanes each (a
// power of two, so lane -> (link, offset) is a shift and a mask), and lane o
// of link r evaluates the A = kHbAttemptsPerLane attempts next_r + A o ...
// next_r + A o + A - 1 (independent Philox streams: A dependency chains in
// flight per lane).  A link takes the lowest accepted attempt index of the
// round (or moves next on by A k), so every link ends with exactly the
// attempt, r0 and proposal count of its own sequential loop, while a wave
// needs about (total attempts) / (64 A) rounds instead of the largest
// attempt count among its 64 links.  The pending links' parameters and the
// accepted attempt pass through per-wave LDS tables indexed by rank.  No
// block-wide barrier: a wave leaves as soon as its own links are done.  Its
// attempt count goes to its own tally slot; links that exhaust the attempt
// limit (rare) are counted by an atomic.
template <typename Real, int MU>
__global__ __launch_bounds__(256) void k_su2_heatbath(typename Vec4<Real>::type *__restrict__ L, Geom g, int p,
                                                      double beta, uint64_t key, uint64_t sweep, int max_attempts,
                                                      unsigned long long *__restrict__ tally,
                                                      unsigned long long *__restrict__ counters,
                                                      int *__restrict__ status) {
    constexpr int A = kHbAttemptsPerLane, NW = 4;
    __shared__ double2 tab_par[NW][64];  // (1/alpha, tm) of pending link `rank`
    __shared__ int2 tab_idx[NW][64];     // (site, next attempt) of pending link `rank`
    __shared__ double tab_r0[NW][64];    // r0 of the first accepted attempt of link `rank`
    __shared__ int tab_first[NW][64];    // its offset from next
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int tile = (p & 1) ? xcd_tile_rev(blockIdx.x, gridDim.x) : xcd_tile(blockIdx.x, gridDim.x);
    const int s = tile * 256 + threadIdx.x;
    int x[4] = {0, 0, 0, 0};
    const bool has = s < g.vh && site_of(g, p, s, x);
    Q4<double> a;
    double norm = 0.0, inv_alpha = 0.0, tm = 0.0;
    uint32_t site = 0;
    if (has) {
        const Q4<Real> A4 = su2_staple<Real, MU>(L, g, x, p, s);
#pragma unroll
        for (int c = 0; c < 4; ++c) a.c[c] = (double)A4.c[c];
        norm = __builtin_sqrt(qnorm2(a));
        const double alpha = beta * norm;
        site = (uint32_t)lex_of(g, x);
        if (alpha > 1.0e-300) {
            inv_alpha = 1.0 / alpha;
            tm = -expm1_nonpos(-2.0 * alpha);
        }
    }
    bool pending = has, ok = false;
    int next = 0;  // < max_attempts while pending
    unsigned long long nprop = 0;
    double r0 = 0.0;
    const unsigned long long below = (1ull << lane) - 1ull;
    while (true) {
        const unsigned long long pm = __ballot(pending);
        if (pm == 0ull) break;
        const int P = __popcll(pm);
        const int lg = P > 1 ? 32 - __clz(P - 1) : 0;  // ceil(log2 P)
        const int sh = 6 - lg, k = 1 << sh;            // lanes per link
        const int rank = __popcll(pm & below);
        if (pending) {
            tab_par[wave][rank] = make_double2(inv_alpha, tm);
            tab_idx[wave][rank] = make_int2((int)site, next);
        }
        wave_lds_sync();
        const int r = lane >> sh, o = lane & (k - 1);
        bool hit = false;
        double tr0 = 0.0;
        int first = 0;
        if (r < P) {
            const double2 par = tab_par[wave][r];
            const int2 idx = tab_idx[wave][r];
            const int left = max_attempts - idx.y;  // > 0: attempts idx.y + j exist for j < left
            double ra[A];
            bool ha[A];
#pragma unroll
            for (int j = 0; j < A; ++j) {
                const int off = A * o + j;
                ra[j] = 0.0;
                ha[j] = off < left && hb_attempt(key, (uint32_t)idx.x, MU, idx.y + off, sweep, par.x, par.y, ra[j]);
            }
#pragma unroll
            for (int j = A - 1; j >= 0; --j)
                if (ha[j]) {
                    hit = true;
                    tr0 = ra[j];
                    first = A * o + j;
                }
        }
        const unsigned long long am = __ballot(hit);
        // the first accepted lane of each link's group publishes its attempt
        const unsigned long long grp = (k == 64) ? ~0ull : (((1ull << k) - 1ull) << (r << sh));
        if (hit && (am & grp & below) == 0ull) {
            tab_r0[wave][r] = tr0;
            tab_first[wave][r] = first;
        }
        wave_lds_sync();
        if (pending) {
            const unsigned long long mine = (k == 64) ? am : ((am >> (rank << sh)) & ((1ull << k) - 1ull));
            if (mine) {
                ok = true;
                pending = false;
                r0 = tab_r0[wave][rank];
                nprop = (unsigned long long)next + (unsigned long long)tab_first[wave][rank] + 1ull;
            } else if (next >= max_attempts - A * k) {
                pending = false;
                nprop = (unsigned long long)max_attempts;
            } else {
                next += A * k;
            }
        }
        wave_lds_sync();
    }
    if (has && ok) {
        const double radial = __builtin_sqrt(fmax(__builtin_fma(-r0, r0, 1.0), 0.0));
        uint32_t o[4];
        draw(key, site, MU, KIND_HB_DIRECTION, 0, sweep, o);
        const double z = 2.0 * unit53(o[0], o[1]) - 1.0;
        double sn, cs;
        sincos_turns(unit53(o[2], o[3]), sn, cs);
        const double planar = __builtin_sqrt(fmax(__builtin_fma(-z, z, 1.0), 0.0));
        const double rp = radial * planar;
        Q4<double> R;
        R.c[0] = r0; R.c[1] = rp * cs; R.c[2] = rp * sn; R.c[3] = radial * z;
        Q4<double> un;
        if (norm == 0.0) {
            un = R;
        } else {
            const double inv_norm = 1.0 / norm;
            Q4<double> ad;
            ad.c[0] = a.c[0] * inv_norm; ad.c[1] = -(a.c[1] * inv_norm);
            ad.c[2] = -(a.c[2] * inv_norm); ad.c[3] = -(a.c[3] * inv_norm);
            un = qmul(R, ad);
        }
        qnormalize(un);
        Q4<Real> v;
#pragma unroll
        for (int c = 0; c < 4; ++c) v.c[c] = (Real)un.c[c];
        store2<Real>(L, g.vh, MU, p, s, v);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) nprop += __shfl_down(nprop, off, 64);
    const unsigned long long failed = __ballot(has && !ok);
    if (lane == 0) {
        tally[blockIdx.x * NW + wave] += nprop;
        if (failed) {
            atomicAdd(&counters[1], (unsigned long long)__popcll(failed));
            atomicOr(status, 1);
        }
    }
}

template <typename Real, int MU>
__global__ __launch_bounds__(256) void k_su2_overrelax(typename Vec4<Real>::type *__restrict__ L, Geom g, int p) {
    const int tile = (p & 1) ? xcd_tile_rev(blockIdx.x, gridDim.x) : xcd_tile(blockIdx.x, gridDim.x);
    const int s = tile * 256 + threadIdx.x;
    int x[4];
    if (s >= g.vh || !site_of(g, p, s, x)) return;
    const Q4<Real> Af = su2_staple<Real, MU>(L, g, x, p, s);
    Q4<double> a;
#pragma unroll
    for (int c = 0; c < 4; ++c) a.c[c] = (double)Af.c[c];
    const double norm = __builtin_sqrt(qnorm2(a));
    if (norm == 0.0) return;
    Q4<double> ad;
    ad.c[0] = a.c[0] / norm; ad.c[1] = -(a.c[1] / norm); ad.c[2] = -(a.c[2] / norm); ad.c[3] = -(a.c[3] / norm);
    const Q4<Real> Uf = load2<Real>(L, g.vh, MU, p, s);
    Q4<double> ud;
    ud.c[0] = (double)Uf.c[0]; ud.c[1] = -(double)Uf.c[1]; ud.c[2] = -(double)Uf.c[2]; ud.c[3] = -(double)Uf.c[3];
    Q4<double> r = qmul(qmul(ad, ud), ad);
    qnormalize(r);
    Q4<Real> v;
#pragma unroll
    for (int c = 0; c < 4; ++c) v.c[c] = (Real)r.c[c];
    store2<Real>(L, g.vh, MU, p, s, v);
}

template <typename Real>
static void hb_t(void *links, const Geom &g, int mu, int p, double beta, uint64_t key, uint64_t sweep, int max_att,
                 unsigned long long *tally, unsigned long long *ctr, int *status, hipStream_t st) {
    using V4 = typename Vec4<Real>::type;
    const int nb = (g.vh + 255) / 256;
    V4 *L = (V4 *)links;
#define YM_SU2_HB(M)                                                                                     \
    hipLaunchKernelGGL((k_su2_heatbath<Real, M>), dim3(nb), dim3(256), 0, st, L, g, p, beta, key, sweep, \
                       max_att, tally, ctr, status)
    switch (mu) {
    case 0: YM_SU2_HB(0); break;
    case 1: YM_SU2_HB(1); break;
    case 2: YM_SU2_HB(2); break;
    default: YM_SU2_HB(3); break;
    }
#undef YM_SU2_HB
}
template <typename Real>
static void or_t(void *links, const Geom &g, int mu, int p, hipStream_t st) {
    using V4 = typename Vec4<Real>::type;
    const int nb = (g.vh + 255) / 256;
    V4 *L = (V4 *)links;
    switch (mu) {
    case 0: hipLaunchKernelGGL((k_su2_overrelax<Real, 0>), dim3(nb), dim3(256), 0, st, L, g, p); break;
    case 1: hipLaunchKernelGGL((k_su2_overrelax<Real, 1>), dim3(nb), dim3(256), 0, st, L, g, p); break;
    case 2: hipLaunchKernelGGL((k_su2_overrelax<Real, 2>), dim3(nb), dim3(256), 0, st, L, g, p); break;
    default: hipLaunchKernelGGL((k_su2_overrelax<Real, 3>), dim3(nb), dim3(256), 0, st, L, g, p); break;
    }
}
void launch_su2_heatbath(int prec, void *links, const Geom &g, int mu, int p, double beta, uint64_t key, uint64_t sweep,
                         int max_attempts, unsigned long long *tally, unsigned long long *counters, int *status,
                         hipStream_t st) {
    if (prec == 32) hb_t<float>(links, g, mu, p, beta, key, sweep, max_attempts, tally, counters, status, st);
    else hb_t<double>(links, g, mu, p, beta, key, sweep, max_attempts, tally, counters, status, st);
}
void launch_su2_overrelax(int prec, void *links, const Geom &g, int mu, int p, hipStream_t st) {
    if (prec == 32) or_t<float>(links, g, mu, p, st);
    else or_t<double>(links, g, mu, p, st);
}

// ----------------------------------------------------------- hot start
template <typename Real>
__global__ __launch_bounds__(256) void k_su2_hot(typename Vec4<Real>::type *__restrict__ L, Geom g, uint64_t key) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    const int p = blockIdx.y, mu = blockIdx.z;
    int x[4];
    if (s >= g.vh) return;
    Q4<Real> v;
    if (!site_of(g, p, s, x)) {
        v.c[0] = 1; v.c[1] = 0; v.c[2] = 0; v.c[3] = 0;
        store2<Real>(L, g.vh, mu, p, s, v);
        return;
    }
    const uint32_t site = (uint32_t)lex_of(g, x);
    Q4<double> q;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        uint32_t o[4];
        draw(key, site, mu, KIND_HOT_START, (uint32_t)k, 0, o);
        const double u1 = 1.0 - unit53(o[0], o[1]);
        const double u2 = unit53(o[2], o[3]);
        const double r = __builtin_sqrt(-2.0 * log_pos(u1));
        double sn, cs;
        sincos_turns(u2, sn, cs);
        q.c[2 * k] = r * cs;
        q.c[2 * k + 1] = r * sn;
    }
    qnormalize(q);
#pragma unroll
    for (int c = 0; c < 4; ++c) v.c[c] = (Real)q.c[c];
    store2<Real>(L, g.vh, mu, p, s, v);
}
void launch_su2_hot(int prec, void *links, const Geom &g, uint64_t key, hipStream_t st) {
    dim3 grid((g.vh + 255) / 256, 2, 4);
    if (prec == 32) hipLaunchKernelGGL((k_su2_hot<float>), grid, dim3(256), 0, st, (float4 *)links, g, key);
    else hipLaunchKernelGGL((k_su2_hot<double>), grid, dim3(256), 0, st, (double4 *)links, g, key);
}

// ------------------------------------------------------- layout change
template <typename Real>
__global__ __launch_bounds__(256) void k_su2_from_aos(typename Vec4<Real>::type *__restrict__ L, Geom g,
                                                      const double *__restrict__ aos) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    const int p = blockIdx.y, mu = blockIdx.z;
    int x[4];
    if (s >= g.vh) return;
    Q4<Real> v;
    if (!site_of(g, p, s, x)) {
        v.c[0] = 1; v.c[1] = 0; v.c[2] = 0; v.c[3] = 0;
    } else {
        const size_t base = ((size_t)lex_of(g, x) * 4 + mu) * 4;
#pragma unroll
        for (int c = 0; c < 4; ++c) v.c[c] = (Real)aos[base + c];
    }
    store2<Real>(L, g.vh, mu, p, s, v);
}
template <typename Real>
__global__ __launch_bounds__(256) void k_su2_to_aos(const typename Vec4<Real>::type *__restrict__ L, Geom g,
                                                    double *__restrict__ aos) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    const int p = blockIdx.y, mu = blockIdx.z;
    int x[4];
    if (s >= g.vh || !site_of(g, p, s, x)) return;
    const size_t base = ((size_t)lex_of(g, x) * 4 + mu) * 4;
    const Q4<Real> v = load2<Real>(L, g.vh, mu, p, s);
#pragma unroll
    for (int c = 0; c < 4; ++c) aos[base + c] = (double)v.c[c];
}
void launch_su2_from_aos(int prec, void *links, const Geom &g, const double *aos, hipStream_t st) {
    dim3 grid((g.vh + 255) / 256, 2, 4);
    if (prec == 32) hipLaunchKernelGGL((k_su2_from_aos<float>), grid, dim3(256), 0, st, (float4 *)links, g, aos);
    else hipLaunchKernelGGL((k_su2_from_aos<double>), grid, dim3(256), 0, st, (double4 *)links, g, aos);
}
void launch_su2_to_aos(int prec, const void *links, const Geom &g, double *aos, hipStream_t st) {
    dim3 grid((g.vh + 255) / 256, 2, 4);
    if (prec == 32) hipLaunchKernelGGL((k_su2_to_aos<float>), grid, dim3(256), 0, st, (const float4 *)links, g, aos);
    else hipLaunchKernelGGL((k_su2_to_aos<double>), grid, dim3(256), 0, st, (const double4 *)links, g, aos);
}

// Host-layout transfer by lexicographic site range (see ym_su3.hip): thread
// i owns link (site0 + i / 4, mu = i % 4), 4 doubles at aos[4 i].
__device__ __forceinline__ void su2_site_of_lex(const Geom &g, int lex, int x[4]) {
    x[3] = lex % g.n[3];
    int r = lex / g.n[3];
    x[2] = r % g.n[2];
    r /= g.n[2];
    x[1] = r % g.n[1];
    x[0] = r / g.n[1];
}
template <typename Real>
__global__ __launch_bounds__(256) void k_su2_from_aos_range(typename Vec4<Real>::type *__restrict__ L, Geom g,
                                                            const double *__restrict__ aos, int site0, int nlinks) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nlinks) return;
    int x[4];
    su2_site_of_lex(g, site0 + (i >> 2), x);
    Q4<Real> v;
#pragma unroll
    for (int c = 0; c < 4; ++c) v.c[c] = (Real)aos[(size_t)i * 4 + c];
    store2<Real>(L, g.vh, i & 3, parity_of(x), slot_of(g, x), v);
}
template <typename Real>
__global__ __launch_bounds__(256) void k_su2_to_aos_range(const typename Vec4<Real>::type *__restrict__ L, Geom g,
                                                          double *__restrict__ aos, int site0, int nlinks) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nlinks) return;
    int x[4];
    su2_site_of_lex(g, site0 + (i >> 2), x);
    const Q4<Real> v = load2<Real>(L, g.vh, i & 3, parity_of(x), slot_of(g, x));
#pragma unroll
    for (int c = 0; c < 4; ++c) aos[(size_t)i * 4 + c] = (double)v.c[c];
}
void launch_su2_from_aos_range(int prec, void *links, const Geom &g, const double *aos, int site0, int nsites,
                               hipStream_t st) {
    const int nl = 4 * nsites;
    const dim3 grid((nl + 255) / 256);
    if (prec == 32) hipLaunchKernelGGL((k_su2_from_aos_range<float>), grid, dim3(256), 0, st, (float4 *)links, g, aos, site0, nl);
    else hipLaunchKernelGGL((k_su2_from_aos_range<double>), grid, dim3(256), 0, st, (double4 *)links, g, aos, site0, nl);
}
void launch_su2_to_aos_range(int prec, const void *links, const Geom &g, double *aos, int site0, int nsites,
                             hipStream_t st) {
    const int nl = 4 * nsites;
    const dim3 grid((nl + 255) / 256);
    if (prec == 32) hipLaunchKernelGGL((k_su2_to_aos_range<float>), grid, dim3(256), 0, st, (const float4 *)links, g, aos, site0, nl);
    else hipLaunchKernelGGL((k_su2_to_aos_range<double>), grid, dim3(256), 0, st, (const double4 *)links, g, aos, site0, nl);
}

// Identity in every slot (padding included); FP64 normalisation of every
// link, a0^2 + a1^2 + a2^2 + a3^2 by an FMA chain and one reciprocal square
// root per link.
template <typename Real>
__global__ __launch_bounds__(256) void k_su2_identity(typename Vec4<Real>::type *__restrict__ L, Geom g) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= g.vh) return;
    Q4<Real> v;
    v.c[0] = 1; v.c[1] = 0; v.c[2] = 0; v.c[3] = 0;
    store2<Real>(L, g.vh, blockIdx.z, blockIdx.y, s, v);
}
template <typename Real>
__global__ __launch_bounds__(256) void k_su2_normalize(typename Vec4<Real>::type *__restrict__ L, Geom g) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    int x[4];
    if (s >= g.vh || !site_of(g, blockIdx.y, s, x)) return;
    const Q4<Real> v = load2<Real>(L, g.vh, blockIdx.z, blockIdx.y, s);
    Q4<double> q;
#pragma unroll
    for (int c = 0; c < 4; ++c) q.c[c] = (double)v.c[c];
    qnormalize(q);
    Q4<Real> w;
#pragma unroll
    for (int c = 0; c < 4; ++c) w.c[c] = (Real)q.c[c];
    store2<Real>(L, g.vh, blockIdx.z, blockIdx.y, s, w);
}
void launch_su2_identity(int prec, void *links, const Geom &g, hipStream_t st) {
    const dim3 grid((g.vh + 255) / 256, 2, 4);
    if (prec == 32) hipLaunchKernelGGL((k_su2_identity<float>), grid, dim3(256), 0, st, (float4 *)links, g);
    else hipLaunchKernelGGL((k_su2_identity<double>), grid, dim3(256), 0, st, (double4 *)links, g);
}
void launch_su2_normalize(int prec, void *links, const Geom &g, hipStream_t st) {
    const dim3 grid((g.vh + 255) / 256, 2, 4);
    if (prec == 32) hipLaunchKernelGGL((k_su2_normalize<float>), grid, dim3(256), 0, st, (float4 *)links, g);
    else hipLaunchKernelGGL((k_su2_normalize<double>), grid, dim3(256), 0, st, (double4 *)links, g);
}

// Staple at one site (parity tests).
template <typename Real, int MU>
__global__ void k_su2_staple_probe(const typename Vec4<Real>::type *__restrict__ L, Geom g, int x0, int x1, int x2, int x3,
                                   double *out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int x[4] = {x0, x1, x2, x3};
    const Q4<Real> A = su2_staple<Real, MU>(L, g, x, parity_of(x), slot_of(g, x));
    for (int c = 0; c < 4; ++c) out[c] = (double)A.c[c];
}
template <typename Real>
static void probe2_t(const void *links, const Geom &g, const int x[4], int mu, double *out, hipStream_t st) {
    using V4 = typename Vec4<Real>::type;
    const V4 *L = (const V4 *)links;
    switch (mu) {
    case 0: hipLaunchKernelGGL((k_su2_staple_probe<Real, 0>), dim3(1), dim3(64), 0, st, L, g, x[0], x[1], x[2], x[3], out); break;
    case 1: hipLaunchKernelGGL((k_su2_staple_probe<Real, 1>), dim3(1), dim3(64), 0, st, L, g, x[0], x[1], x[2], x[3], out); break;
    case 2: hipLaunchKernelGGL((k_su2_staple_probe<Real, 2>), dim3(1), dim3(64), 0, st, L, g, x[0], x[1], x[2], x[3], out); break;
    default: hipLaunchKernelGGL((k_su2_staple_probe<Real, 3>), dim3(1), dim3(64), 0, st, L, g, x[0], x[1], x[2], x[3], out); break;
    }
}
void launch_su2_staple_probe(int prec, const void *links, const Geom &g, const int x[4], int mu, double *out, hipStream_t st) {
    if (prec == 32) probe2_t<float>(links, g, x, mu, out, st);
    else probe2_t<double>(links, g, x, mu, out, st);
}

}  // namespace ym
