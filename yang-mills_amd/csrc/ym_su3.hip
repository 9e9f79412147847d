// SU(3) link-update kernels for gfx950.
//
// k_su3_metropolis: one checkerboard sub-pass (direction MU, parity p) of
// single-generator Metropolis.  One thread owns one link of the sub-pass; no
// link it reads is written in the same sub-pass, so all threads run
// independently.  Per link:
//   1. staple A (6 terms) in the storage precision with a fixed FMA chain,
//      then held in FP64 with the link U;
//   2. `hits` proposals, generator h mod 8, angle from the reference's paired
//      grid driven by Philox(site, mu, hit, sweep); Re Tr(U' A) - Re Tr(U A)
//      = Re Tr((R - I) U A) for U' = R U from the 2 x 2 block of U A that R
//      touches, each entry part one FMA chain over the current U; accept iff
//      dS <= 0 or u < exp(-dS) in FP64; U <- R U with R replaced by the
//      identity rotation (c, s) = (1, 0) when rejected (branch-free);
//   3. FP64 Gram-Schmidt reunitarisation, then one rounding to storage.
// The CPU replay in the oracle performs the same operations in the same order.
#include "ym_device.hpp"
#include "ym_internal.hpp"
#include "ym_su3_dev.hpp"

namespace ym {

// One sub-pass (MU, p) in one kernel: staple then update per thread; the
// accepted-proposal count is reduced per block and added to the block's
// tally slot.
template <typename Real, int MU>
__global__ __launch_bounds__(256) void k_su3_metropolis(typename Vec2<Real>::type *__restrict__ L, Geom g, int p,
                                                        double nbt, double angle, int hits, uint64_t key,
                                                        uint64_t sweep, int reunit,
                                                        unsigned long long *__restrict__ tally) {
    __shared__ unsigned long long red[4];
    const int s = tile_slot(g, xcd_tile(blockIdx.x, gridDim.x), threadIdx.x);
    unsigned long long nacc = 0;
    int x[4];
    if (s < g.vh && site_of(g, p, s, x)) {
        M3<Real> A;
        su3_staple_at<Real, MU>(L, g, p, s, x, A);
        nacc = su3_update_at<Real, MU>(L, g, p, s, x, A, nbt, angle, hits, key, sweep, reunit);
    }
    unsigned long long tot = block_sum_u64<256>(nacc, red);
    if (threadIdx.x == 0) tally[blockIdx.x] += tot;
}

// Wave-specialised sub-pass for lattices of at least 256 tiles: one block of
// 8 wavefronts per CU walks a sequence of 256-slot tiles.  Producer wave k
// (0-3) computes the staples (and loads the links) of slots 64k..64k+63 of
// each tile into an LDS double buffer; consumer wave k + 4 runs the FP64
// hits of the same 64 slots.  The two waves of a pair hand buffers over by
// LDS flags (wave_handoff_wait / wave_handoff_post), not by a block barrier:
// a producer fills buffer b for tile i once its consumer has read the
// buffer's previous tile i - 2 into registers, and the consumer releases a
// buffer right after reading it, so a producer runs up to two tiles ahead
// of its consumer and no pair waits for the slowest of the four.  Per link
// the arithmetic is that of k_su3_metropolis, so results are identical.
// Tiles are split into 8 contiguous ranges, one per XCD (blocks are dealt
// round-robin over the XCDs), and the blocks of an XCD advance through its
// range side by side, so the tiles in flight on an XCD are neighbours.  The
// producers also stage the link being updated, so the hit waves touch
// global memory only for the final store.  Every sub-pass reads the whole
// link field; sub-passes of parity 1 walk each XCD's range backwards, so a
// sub-pass starts on the tiles its predecessor touched last, which are still
// in that XCD's L2 and in the memory-side cache (a 302 MB field at 32^4 is
// larger than the 256 MB cache: the same direction every time evicts each
// line just before it is needed again).  Tiles are independent, so the
// direction changes timing only.
template <typename Real, int MU, int HITS>
__global__ __launch_bounds__(512) void k_su3_metropolis_ws(typename Vec2<Real>::type *__restrict__ L, Geom g, int p,
                                                           double nbt, double angle, int hits, uint64_t key,
                                                           uint64_t sweep, int reunit, int ntiles,
                                                           unsigned long long *__restrict__ tally,
                                                           int *__restrict__ status) {
    constexpr int NB = 2;  // three or four FP32 buffers measured 2-3% slower
    __shared__ LdsMat<Real> abuf[NB], ubuf[NB];
    __shared__ unsigned long long red[8];
    __shared__ int ready[NB][4], freed[NB][4];  // tile index + 1 last filled / last read, per pair
    const bool producer = threadIdx.x < 256;
    const int lane = threadIdx.x & 255, pair = lane >> 6;
    const bool leader = (threadIdx.x & 63) == 0;
    if (threadIdx.x < 4 * NB) {
        (&ready[0][0])[threadIdx.x] = 0;
        (&freed[0][0])[threadIdx.x] = 0;
    }
    __syncthreads();
    const int xcd = blockIdx.x & 7, j = blockIdx.x >> 3;
    const int nbx = ((int)gridDim.x - 1 - xcd) / 8 + 1;
    const int t0 = (int)(((int64_t)ntiles * xcd) / 8), t1 = (int)(((int64_t)ntiles * (xcd + 1)) / 8);
    const int nsteps = (t1 - t0 - j + nbx - 1) / nbx;
    unsigned long long nacc = 0;
    const bool rev = (p & 1) != 0;
    auto tile_at = [&](int i) { return rev ? t1 - 1 - (j + i * nbx) : t0 + j + i * nbx; };
    if (producer) {
        for (int i = 0; i < nsteps; ++i) {
            const int b = i % NB;
            if (i >= NB) wave_handoff_wait(&freed[b][pair], i - NB + 1, status);
            const int s = tile_slot(g, tile_at(i), lane);
            int x[4];
            if (s < g.vh && site_of(g, p, s, x)) {
                M3<Real> A;
                su3_staple_at<Real, MU>(L, g, p, s, x, A);
                const M3<Real> U = load3_own(L, g.vh, MU, p, s);
                abuf[b].put(lane, A);
                ubuf[b].put(lane, U);
            }
            wave_handoff_post(&ready[b][pair], i + 1, leader);
        }
    } else {
        for (int i = 0; i < nsteps; ++i) {
            const int b = i % NB;
            wave_handoff_wait(&ready[b][pair], i + 1, status);
            const int s = tile_slot(g, tile_at(i), lane);
            int x[4];
            const bool has = s < g.vh && site_of(g, p, s, x);
            M3<Real> A, U;
            if (has) {
                A = abuf[b].get(lane);
                U = ubuf[b].get(lane);
            }
            wave_handoff_post(&freed[b][pair], i + 1, leader);
            if (has) nacc += su3_update_from<Real, MU, HITS>(L, g, p, s, x, U, A, nbt, angle, hits, key, sweep, reunit);
        }
    }
    const unsigned long long tot = block_sum_u64<512>(nacc, red);
    if (threadIdx.x == 0) tally[blockIdx.x] += tot;
}

template <typename Real>
static void launch_metro_t(void *links, const Geom &g, int mu, int p, double beta, double angle, int hits,
                           uint64_t key, uint64_t sweep, int reunit, unsigned long long *tally, int *status,
                           hipStream_t st) {
    using V2 = typename Vec2<Real>::type;
    const double nbt = -(beta / 3.0);
    const int nb = (g.vh + 255) / 256;
    V2 *L = (V2 *)links;
    if (nb >= 256 && g.kper == 0) {
        // one 8-wave block per CU; 10 hits per link (generators 0-7, 0, 1) has
        // its own fully unrolled instance
        const int grid = 256;
#define YM_WS(M, H) hipLaunchKernelGGL((k_su3_metropolis_ws<Real, M, H>), dim3(grid), dim3(512), 0, st, L, g, p, nbt, angle, hits, key, sweep, reunit, nb, tally, status)
        if (hits == 10) {
            switch (mu) {
            case 0: YM_WS(0, 10); break;
            case 1: YM_WS(1, 10); break;
            case 2: YM_WS(2, 10); break;
            default: YM_WS(3, 10); break;
            }
        } else if (hits == 8) {
            switch (mu) {
            case 0: YM_WS(0, 8); break;
            case 1: YM_WS(1, 8); break;
            case 2: YM_WS(2, 8); break;
            default: YM_WS(3, 8); break;
            }
        } else {
            switch (mu) {
            case 0: YM_WS(0, 0); break;
            case 1: YM_WS(1, 0); break;
            case 2: YM_WS(2, 0); break;
            default: YM_WS(3, 0); break;
            }
        }
#undef YM_WS
        return;
    }
    switch (mu) {
    case 0: hipLaunchKernelGGL((k_su3_metropolis<Real, 0>), dim3(nb), dim3(256), 0, st, L, g, p, nbt, angle, hits, key, sweep, reunit, tally); break;
    case 1: hipLaunchKernelGGL((k_su3_metropolis<Real, 1>), dim3(nb), dim3(256), 0, st, L, g, p, nbt, angle, hits, key, sweep, reunit, tally); break;
    case 2: hipLaunchKernelGGL((k_su3_metropolis<Real, 2>), dim3(nb), dim3(256), 0, st, L, g, p, nbt, angle, hits, key, sweep, reunit, tally); break;
    default: hipLaunchKernelGGL((k_su3_metropolis<Real, 3>), dim3(nb), dim3(256), 0, st, L, g, p, nbt, angle, hits, key, sweep, reunit, tally); break;
    }
}
void launch_su3_metropolis(int prec, void *links, const Geom &g, int mu, int p, double beta, double angle, int hits,
                           uint64_t key, uint64_t sweep, int reunit, unsigned long long *tally, int *status,
                           hipStream_t st) {
    if (prec == 32) launch_metro_t<float>(links, g, mu, p, beta, angle, hits, key, sweep, reunit, tally, status, st);
    else launch_metro_t<double>(links, g, mu, p, beta, angle, hits, key, sweep, reunit, tally, status, st);
}

// ----------------------------------------------------------- hot start
// Haar-random SU(3): two rows of complex Gaussians (Box-Muller from Philox),
// Gram-Schmidt, third row = conj(row0 x row1).
template <typename Real>
__global__ __launch_bounds__(256) void k_su3_hot(typename Vec2<Real>::type *__restrict__ L, Geom g, uint64_t key) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    const int p = blockIdx.y, mu = blockIdx.z;
    int x[4];
    if (s >= g.vh) return;
    if (!site_of(g, p, s, x)) {
        store3<Real>(L, g.vh, mu, p, s, identity3<Real>());
        return;
    }
    const uint32_t site = (uint32_t)lex_of(g, x);
    M3<double> u;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        uint32_t o[4];
        draw(key, site, mu, KIND_HOT_START, (uint32_t)k, 0, o);
        const double u1 = 1.0 - unit53(o[0], o[1]);
        const double u2 = unit53(o[2], o[3]);
        const double r = __builtin_sqrt(-2.0 * log_pos(u1));
        double sn, cs;
        sincos_turns(u2, sn, cs);
        u.v[2 * k] = r * cs;
        u.v[2 * k + 1] = r * sn;
    }
#pragma unroll
    for (int c = 12; c < 18; ++c) u.v[c] = 0.0;
    reunitarize3(u);
    M3<Real> o;
#pragma unroll
    for (int c = 0; c < 18; ++c) o.v[c] = (Real)u.v[c];
    store3<Real>(L, g.vh, mu, p, s, o);
}
void launch_su3_hot(int prec, void *links, const Geom &g, uint64_t key, hipStream_t st) {
    dim3 grid((g.vh + 255) / 256, 2, 4);
    if (prec == 32) hipLaunchKernelGGL((k_su3_hot<float>), grid, dim3(256), 0, st, (float2 *)links, g, key);
    else hipLaunchKernelGGL((k_su3_hot<double>), grid, dim3(256), 0, st, (double2 *)links, g, key);
}

// --------------------------------------------------------- staple probe
// Staple at one site, with the sweep's own routine (su3_staple_at), for
// parity tests.
template <typename Real, int MU>
__global__ void k_su3_staple_probe(const typename Vec2<Real>::type *__restrict__ L, Geom g, int x0, int x1, int x2,
                                   int x3, double *out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int x[4] = {x0, x1, x2, x3};
    M3<Real> A;
    su3_staple_at<Real, MU>(L, g, parity_of(x), slot_of(g, x), x, A);
    for (int c = 0; c < 18; ++c) out[c] = (double)A.v[c];
}
template <typename Real>
static void staple_probe_t(const void *links, const Geom &g, const int x[4], int mu, double *out, hipStream_t st) {
    using V2 = typename Vec2<Real>::type;
    const V2 *L = (const V2 *)links;
    switch (mu) {
    case 0: hipLaunchKernelGGL((k_su3_staple_probe<Real, 0>), dim3(1), dim3(64), 0, st, L, g, x[0], x[1], x[2], x[3], out); break;
    case 1: hipLaunchKernelGGL((k_su3_staple_probe<Real, 1>), dim3(1), dim3(64), 0, st, L, g, x[0], x[1], x[2], x[3], out); break;
    case 2: hipLaunchKernelGGL((k_su3_staple_probe<Real, 2>), dim3(1), dim3(64), 0, st, L, g, x[0], x[1], x[2], x[3], out); break;
    default: hipLaunchKernelGGL((k_su3_staple_probe<Real, 3>), dim3(1), dim3(64), 0, st, L, g, x[0], x[1], x[2], x[3], out); break;
    }
}
void launch_su3_staple_probe(int prec, const void *links, const Geom &g, const int x[4], int mu, double *out, hipStream_t st) {
    if (prec == 32) staple_probe_t<float>(links, g, x, mu, out, st);
    else staple_probe_t<double>(links, g, x, mu, out, st);
}

// ------------------------------------------------------ reunitarisation
template <typename Real>
__global__ __launch_bounds__(256) void k_su3_reunit(typename Vec2<Real>::type *__restrict__ L, Geom g) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    const int p = blockIdx.y, mu = blockIdx.z;
    int x[4];
    if (s >= g.vh || !site_of(g, p, s, x)) return;
    M3<Real> a = load3<Real>(L, g.vh, mu, p, s);
    M3<double> u;
#pragma unroll
    for (int c = 0; c < 18; ++c) u.v[c] = (double)a.v[c];
    reunitarize3(u);
#pragma unroll
    for (int c = 0; c < 18; ++c) a.v[c] = (Real)u.v[c];
    store3<Real>(L, g.vh, mu, p, s, a);
}
void launch_su3_reunit(int prec, void *links, const Geom &g, hipStream_t st) {
    dim3 grid((g.vh + 255) / 256, 2, 4);
    if (prec == 32) hipLaunchKernelGGL((k_su3_reunit<float>), grid, dim3(256), 0, st, (float2 *)links, g);
    else hipLaunchKernelGGL((k_su3_reunit<double>), grid, dim3(256), 0, st, (double2 *)links, g);
}

// ------------------------------------------------------------ residuals
// Per block: max over its links of max|U^+U - 1| (both orders) and of
// max(|det U - 1|, |det U^+ - 1|), in FP64 (su3.rs:219-229 semantics).
template <typename Real>
__global__ __launch_bounds__(256) void k_su3_residuals(const typename Vec2<Real>::type *__restrict__ L, Geom g,
                                                       double *__restrict__ part) {
    __shared__ double red[4];
    const int s = blockIdx.x * 256 + threadIdx.x;
    const int p = blockIdx.y, mu = blockIdx.z;
    const int blin = (mu * 2 + p) * gridDim.x + blockIdx.x;
    double ru = 0.0, rd = 0.0;
    int x[4];
    if (s < g.vh && site_of(g, p, s, x)) {
        M3<Real> a = load3<Real>(L, g.vh, mu, p, s);
        M3<double> u;
        for (int c = 0; c < 18; ++c) u.v[c] = (double)a.v[c];
        M3<double> ud = dagger3(u);
        M3<double> l = mul_nn(ud, u), r = mul_nn(u, ud);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double e = (i == j) ? 1.0 : 0.0;
                ru = fmax(ru, hypot(YM_RE(l, i, j) - e, YM_IM(l, i, j)));
                ru = fmax(ru, hypot(YM_RE(r, i, j) - e, YM_IM(r, i, j)));
            }
        // det via cofactor expansion along row 0
        double dr = 0.0, di = 0.0;
        for (int k = 0; k < 3; ++k) {
            const int a1 = (k + 1) % 3, a2 = (k + 2) % 3;
            double cr = YM_RE(u, 1, a1) * YM_RE(u, 2, a2) - YM_IM(u, 1, a1) * YM_IM(u, 2, a2) -
                        (YM_RE(u, 1, a2) * YM_RE(u, 2, a1) - YM_IM(u, 1, a2) * YM_IM(u, 2, a1));
            double ci = YM_RE(u, 1, a1) * YM_IM(u, 2, a2) + YM_IM(u, 1, a1) * YM_RE(u, 2, a2) -
                        (YM_RE(u, 1, a2) * YM_IM(u, 2, a1) + YM_IM(u, 1, a2) * YM_RE(u, 2, a1));
            dr += YM_RE(u, 0, k) * cr - YM_IM(u, 0, k) * ci;
            di += YM_RE(u, 0, k) * ci + YM_IM(u, 0, k) * cr;
        }
        // det(U^+) = conj(det U): both residuals coincide up to rounding.
        rd = hypot(dr - 1.0, di);
    }
    const double mu_ = block_max<256>(ru, red);
    const double md = block_max<256>(rd, red);
    if (threadIdx.x == 0) { part[2 * blin] = mu_; part[2 * blin + 1] = md; }
}
int launch_su3_residuals(int prec, const void *links, const Geom &g, double *part, hipStream_t st) {
    dim3 grid((g.vh + 255) / 256, 2, 4);
    if (prec == 32) hipLaunchKernelGGL((k_su3_residuals<float>), grid, dim3(256), 0, st, (const float2 *)links, g, part);
    else hipLaunchKernelGGL((k_su3_residuals<double>), grid, dim3(256), 0, st, (const double2 *)links, g, part);
    return (int)(grid.x * 8);
}

// ------------------------------------------------------- layout change
// AoS FP64 (reference layout: link site*4+mu, 18 doubles) <-> device planes.
template <typename Real>
__global__ __launch_bounds__(256) void k_su3_from_aos(typename Vec2<Real>::type *__restrict__ L, Geom g,
                                                      const double *__restrict__ aos) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    const int p = blockIdx.y, mu = blockIdx.z;
    int x[4];
    if (s >= g.vh) return;
    if (!site_of(g, p, s, x)) {
        store3<Real>(L, g.vh, mu, p, s, identity3<Real>());
        return;
    }
    const size_t base = ((size_t)lex_of(g, x) * 4 + mu) * 18;
    M3<Real> m;
#pragma unroll
    for (int c = 0; c < 18; ++c) m.v[c] = (Real)aos[base + c];
    store3<Real>(L, g.vh, mu, p, s, m);
}
template <typename Real>
__global__ __launch_bounds__(256) void k_su3_to_aos(const typename Vec2<Real>::type *__restrict__ L, Geom g,
                                                    double *__restrict__ aos) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    const int p = blockIdx.y, mu = blockIdx.z;
    int x[4];
    if (s >= g.vh || !site_of(g, p, s, x)) return;
    const size_t base = ((size_t)lex_of(g, x) * 4 + mu) * 18;
    M3<Real> m = load3<Real>(L, g.vh, mu, p, s);
#pragma unroll
    for (int c = 0; c < 18; ++c) aos[base + c] = (double)m.v[c];
}
void launch_su3_from_aos(int prec, void *links, const Geom &g, const double *aos, hipStream_t st) {
    dim3 grid((g.vh + 255) / 256, 2, 4);
    if (prec == 32) hipLaunchKernelGGL((k_su3_from_aos<float>), grid, dim3(256), 0, st, (float2 *)links, g, aos);
    else hipLaunchKernelGGL((k_su3_from_aos<double>), grid, dim3(256), 0, st, (double2 *)links, g, aos);
}
void launch_su3_to_aos(int prec, const void *links, const Geom &g, double *aos, hipStream_t st) {
    dim3 grid((g.vh + 255) / 256, 2, 4);
    if (prec == 32) hipLaunchKernelGGL((k_su3_to_aos<float>), grid, dim3(256), 0, st, (const float2 *)links, g, aos);
    else hipLaunchKernelGGL((k_su3_to_aos<double>), grid, dim3(256), 0, st, (const double2 *)links, g, aos);
}

// Host-layout transfer by lexicographic site range: thread i of a chunk owns
// link (site0 + i / 4, mu = i % 4), whose 18 doubles sit at aos[18 i]; the
// chunk moves through a fixed staging buffer, so no transfer needs device
// memory proportional to the lattice.  Padding slots (odd N3) are not
// touched; they keep the identity written when the field was created.
__device__ __forceinline__ void site_of_lex(const Geom &g, int lex, int x[4]) {
    x[3] = lex % g.n[3];
    int r = lex / g.n[3];
    x[2] = r % g.n[2];
    r /= g.n[2];
    x[1] = r % g.n[1];
    x[0] = r / g.n[1];
}
template <typename Real>
__global__ __launch_bounds__(256) void k_su3_from_aos_range(typename Vec2<Real>::type *__restrict__ L, Geom g,
                                                            const double *__restrict__ aos, int site0, int nlinks) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nlinks) return;
    int x[4];
    site_of_lex(g, site0 + (i >> 2), x);
    M3<Real> m;
#pragma unroll
    for (int c = 0; c < 18; ++c) m.v[c] = (Real)aos[(size_t)i * 18 + c];
    store3<Real>(L, g.vh, i & 3, parity_of(x), slot_of(g, x), m);
}
template <typename Real>
__global__ __launch_bounds__(256) void k_su3_to_aos_range(const typename Vec2<Real>::type *__restrict__ L, Geom g,
                                                          double *__restrict__ aos, int site0, int nlinks) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nlinks) return;
    int x[4];
    site_of_lex(g, site0 + (i >> 2), x);
    const M3<Real> m = load3<Real>(L, g.vh, i & 3, parity_of(x), slot_of(g, x));
#pragma unroll
    for (int c = 0; c < 18; ++c) aos[(size_t)i * 18 + c] = (double)m.v[c];
}
void launch_su3_from_aos_range(int prec, void *links, const Geom &g, const double *aos, int site0, int nsites,
                               hipStream_t st) {
    const int nl = 4 * nsites;
    const dim3 grid((nl + 255) / 256);
    if (prec == 32) hipLaunchKernelGGL((k_su3_from_aos_range<float>), grid, dim3(256), 0, st, (float2 *)links, g, aos, site0, nl);
    else hipLaunchKernelGGL((k_su3_from_aos_range<double>), grid, dim3(256), 0, st, (double2 *)links, g, aos, site0, nl);
}
void launch_su3_to_aos_range(int prec, const void *links, const Geom &g, double *aos, int site0, int nsites,
                             hipStream_t st) {
    const int nl = 4 * nsites;
    const dim3 grid((nl + 255) / 256);
    if (prec == 32) hipLaunchKernelGGL((k_su3_to_aos_range<float>), grid, dim3(256), 0, st, (const float2 *)links, g, aos, site0, nl);
    else hipLaunchKernelGGL((k_su3_to_aos_range<double>), grid, dim3(256), 0, st, (const double2 *)links, g, aos, site0, nl);
}

// Identity in every slot of every (mu, parity) plane, padding included.
template <typename Real>
__global__ __launch_bounds__(256) void k_su3_identity(typename Vec2<Real>::type *__restrict__ L, Geom g) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= g.vh) return;
    store3<Real>(L, g.vh, blockIdx.z, blockIdx.y, s, identity3<Real>());
}
void launch_su3_identity(int prec, void *links, const Geom &g, hipStream_t st) {
    const dim3 grid((g.vh + 255) / 256, 2, 4);
    if (prec == 32) hipLaunchKernelGGL((k_su3_identity<float>), grid, dim3(256), 0, st, (float2 *)links, g);
    else hipLaunchKernelGGL((k_su3_identity<double>), grid, dim3(256), 0, st, (double2 *)links, g);
}

}  // namespace ym
