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

// One heat-bath attempt: r0 from the truncated exponential by inverse CDF,
// accepted iff u2 <= sqrt(1 - r0^2) (su2_ensemble.rs:229-250).
__device__ __forceinline__ bool hb_attempt(uint64_t key, uint32_t site, int mu, int att, uint64_t sweep, double alpha,
                                           double tm, double &r0, double &radial) {
    uint32_t o[4];
    draw(key, site, mu, KIND_HB_ATTEMPT, (uint32_t)att, sweep, o);
    const double u1 = unit53(o[0], o[1]), u2 = unit53(o[2], o[3]);
    if (alpha <= 1.0e-300) {
        r0 = 2.0 * u1 - 1.0;
    } else {
        const double y = -log1p_dev(-u1 * tm) / alpha;
        r0 = 1.0 - y;
        r0 = r0 < -1.0 ? -1.0 : (r0 > 1.0 ? 1.0 : r0);
    }
    radial = __builtin_sqrt(fmax(__builtin_fma(-r0, r0, 1.0), 0.0));
    return u2 <= radial;
}

// The rejection loop is run by the whole wavefront together.  Each round the
// P links still without an accepted attempt get k = 64 / P lanes each; lane o
// of a link evaluates its attempt next + o.  The link takes the lowest
// accepted attempt index of the round (or moves next on by k), so every link
// ends with exactly the attempt, r0 and proposal count of its own sequential
// loop, while a wave needs about (total attempts) / 64 rounds instead of the
// largest attempt count among its 64 links.
template <typename Real, int MU>
__global__ __launch_bounds__(256) void k_su2_heatbath(typename Vec4<Real>::type *__restrict__ L, Geom g, int p,
                                                      double beta, uint64_t key, uint64_t sweep, int max_attempts,
                                                      unsigned long long *__restrict__ counters,
                                                      int *__restrict__ status) {
    __shared__ unsigned long long red[4];
    __shared__ int owner[4][64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    // parity-1 sub-passes walk the tiles backwards (the field of a 48^4
    // lattice is larger than the memory-side cache; see k_su3_metropolis_ws)
    const int tile = (p & 1) ? xcd_tile_rev(blockIdx.x, gridDim.x) : xcd_tile(blockIdx.x, gridDim.x);
    const int s = tile * 256 + threadIdx.x;
    unsigned long long nprop = 0, nupd = 0;
    int x[4] = {0, 0, 0, 0};
    const bool has = s < g.vh && site_of(g, p, s, x);
    Q4<double> a;
    double norm = 0.0, alpha = 0.0, tm = 0.0;
    uint32_t site = 0;
    if (has) {
        const Q4<Real> Af = su2_staple<Real, MU>(L, g, x, p, s);
#pragma unroll
        for (int c = 0; c < 4; ++c) a.c[c] = (double)Af.c[c];
        norm = __builtin_sqrt(qnorm2(a));
        alpha = beta * norm;
        site = (uint32_t)lex_of(g, x);
        tm = alpha <= 1.0e-300 ? 0.0 : -expm1_nonpos(-2.0 * alpha);
    }
    bool pending = has, ok = false;
    int next = 0;
    double r0 = 0.0, radial = 0.0;
    while (true) {
        const unsigned long long pm = __ballot(pending);
        if (pm == 0ull) break;
        const int P = __popcll(pm);
        const int k = 64 / P;
        const int rank = __popcll(pm & ((1ull << lane) - 1ull));
        if (pending) owner[wave][rank] = lane;
        wave_lds_sync();
        const int r = lane / k, o = lane - r * k;
        const int src = r < P ? owner[wave][r] : lane;
        const double alpha_o = __shfl(alpha, src), tm_o = __shfl(tm, src);
        const uint32_t site_o = (uint32_t)__shfl((int)site, src);
        const int att = __shfl(next, src) + o;
        double tr0 = 0.0, trad = 0.0;
        const bool hit = r < P && att < max_attempts && hb_attempt(key, site_o, MU, att, sweep, alpha_o, tm_o, tr0, trad);
        const unsigned long long am = __ballot(hit);
        // lanes of this link: [rank * k, rank * k + k)
        const unsigned long long span = (k == 64) ? ~0ull : (((1ull << k) - 1ull) << (rank * k));
        const unsigned long long mine = pending ? (am & span) : 0ull;
        const int first = mine ? __ffsll((long long)mine) - 1 : lane;
        const double r0_f = __shfl(tr0, first), rad_f = __shfl(trad, first);
        if (pending) {
            if (mine) {
                ok = true;
                pending = false;
                r0 = r0_f;
                radial = rad_f;
                nprop = (unsigned long long)(next + (first - rank * k) + 1);
            } else {
                next += k;
                if (next >= max_attempts) {
                    pending = false;
                    nprop = (unsigned long long)max_attempts;
                }
            }
        }
        wave_lds_sync();
    }
    if (has) {
        if (!ok) {
            atomicOr(status, 1);
        } else {
            nupd = 1;
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
                Q4<double> ad;
                ad.c[0] = a.c[0] / norm; ad.c[1] = -(a.c[1] / norm); ad.c[2] = -(a.c[2] / norm); ad.c[3] = -(a.c[3] / norm);
                un = qmul(R, ad);
            }
            qnormalize(un);
            Q4<Real> v;
#pragma unroll
            for (int c = 0; c < 4; ++c) v.c[c] = (Real)un.c[c];
            store2<Real>(L, g.vh, MU, p, s, v);
        }
    }
    const unsigned long long tp = block_sum_u64<256>(nprop, red);
    const unsigned long long tu = block_sum_u64<256>(nupd, red);
    if (threadIdx.x == 0) {
        if (tp) atomicAdd(&counters[0], tp);
        if (tu) atomicAdd(&counters[1], tu);
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
                 unsigned long long *ctr, int *status, hipStream_t st) {
    using V4 = typename Vec4<Real>::type;
    const int nb = (g.vh + 255) / 256;
    V4 *L = (V4 *)links;
    switch (mu) {
    case 0: hipLaunchKernelGGL((k_su2_heatbath<Real, 0>), dim3(nb), dim3(256), 0, st, L, g, p, beta, key, sweep, max_att, ctr, status); break;
    case 1: hipLaunchKernelGGL((k_su2_heatbath<Real, 1>), dim3(nb), dim3(256), 0, st, L, g, p, beta, key, sweep, max_att, ctr, status); break;
    case 2: hipLaunchKernelGGL((k_su2_heatbath<Real, 2>), dim3(nb), dim3(256), 0, st, L, g, p, beta, key, sweep, max_att, ctr, status); break;
    default: hipLaunchKernelGGL((k_su2_heatbath<Real, 3>), dim3(nb), dim3(256), 0, st, L, g, p, beta, key, sweep, max_att, ctr, status); break;
    }
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
                         int max_attempts, unsigned long long *counters, int *status, hipStream_t st) {
    if (prec == 32) hb_t<float>(links, g, mu, p, beta, key, sweep, max_attempts, counters, status, st);
    else hb_t<double>(links, g, mu, p, beta, key, sweep, max_attempts, counters, status, st);
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
