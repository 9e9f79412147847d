// Gauge-observable kernels for gfx950: plaquette, averaged rectangular Wilson
// loops, Polyakov loop, per-timeslice A1++ operator sums, residuals and the
// synchronous spatial stout smearing step.  Products run in the storage
// precision; every trace is accumulated in FP64 and reduced per block
// (64-wide wavefront shuffles + LDS), then summed in a fixed order by
// k_sum_partials, so results are deterministic run to run.
#include "ym_device.hpp"
#include "ym_internal.hpp"
#include "ym_obs_dev.hpp"

namespace ym {

// ------------------------------------------------------------ plaquette
// Per site: sum over mu < nu of (1/N) Re Tr P_{mu nu}(x) with
// P = [U_mu(x) U_nu(x+mu)] [U_nu(x) U_mu(x+nu)]^dagger (plaquette_site);
// part[parity * nblk + blk] = block sum.  Used for SU(3) FP64 links, whose
// per-site registers leave no room for a 512-thread block; the other storage
// types use k_plaquette_pp below.
template <class G>
__global__ __launch_bounds__(256) void k_plaquette(const typename G::Store *__restrict__ L, Geom g,
                                                   double *__restrict__ part) {
    __shared__ double red[4];
    const int blk = xcd_tile(blockIdx.x, gridDim.x);  // contiguous tile ranges per XCD
    const int s = blk * 256 + threadIdx.x;
    const int p = blockIdx.y;
    double acc = 0.0;
    if (s < g.vh) acc = plaquette_site<G>(L, g, p, s);
    const double t = block_sum<256>(acc, red);
    if (threadIdx.x == 0) part[blockIdx.y * gridDim.x + blk] = t;
}

// The same partials with both parities of 256 consecutive slots (a tile) in
// one 512-thread block (waves 0-3 parity 0, waves 4-7 parity 1): a site's
// plaquette neighbours are mostly sites of the other parity in the same
// tile, so a tile's links are fetched once for both parity sums.  Visiting
// order: when an x0-slice is a multiple of 8 tiles (tps = tiles per slice),
// each XCD takes the same eighth of every slice (a band of x1) and walks the
// slices in x0 order, so the x0 and x1 neighbours a tile needs are loaded by
// tiles that run at about the same time on the same XCD and are still in
// its L2; otherwise tiles are dealt to the XCDs in contiguous runs
// (xcd_tile).  The partial of a tile is written at the tile's own index and
// each half reduces its 256 sites with block_sum<256>'s tree, so every
// partial is bit-identical to k_plaquette's whatever the visiting order.
__device__ __forceinline__ int plaquette_tile(int b, int nt, int tps) {
    if (tps == 0) return xcd_tile(b, nt);
    const int xcd = b & 7, idx = b >> 3, per = tps >> 3;
    const int slice = idx / per;
    return slice * tps + xcd * per + (idx - slice * per);
}
template <class G>
__global__ __launch_bounds__(512) void k_plaquette_pp(const typename G::Store *__restrict__ L, Geom g, int tps,
                                                      double *__restrict__ part) {
    __shared__ double red[8];
    const int tile = plaquette_tile(blockIdx.x, gridDim.x, tps);
    const int p = threadIdx.x >> 8;
    const int s = tile * 256 + (threadIdx.x & 255);
    double v = 0.0;
    if (s < g.vh) v = plaquette_site<G>(L, g, p, s);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if ((threadIdx.x & 255) == 0) {
        double t = 0.0;
#pragma unroll
        for (int w = 0; w < 4; ++w) t += red[4 * p + w];
        part[p * gridDim.x + tile] = t;
    }
}

// Sum `n` partials in a fixed order into out[0] (one block).
__global__ __launch_bounds__(256) void k_sum_partials(const double *__restrict__ part, int n, int stride,
                                                      double *__restrict__ out) {
    __shared__ double red[4];
    const int col = blockIdx.x;  // independent sums, one per block
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) acc += part[(size_t)col * stride + i];
    const double t = block_sum<256>(acc, red);
    if (threadIdx.x == 0) out[col] = t;
}
__global__ __launch_bounds__(256) void k_max_partials(const double *__restrict__ part, int n, int stride,
                                                      double *__restrict__ out) {
    __shared__ double red[4];
    const int col = blockIdx.x;
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) acc = fmax(acc, part[(size_t)i * stride + col]);
    const double t = block_max<256>(acc, red);
    if (threadIdx.x == 0) out[col] = t;
}
void launch_sum_partials(const double *part, int n, int ncols, int stride, double *out, hipStream_t st) {
    hipLaunchKernelGGL(k_sum_partials, dim3(ncols), dim3(256), 0, st, part, n, stride, out);
}
void launch_max_partials(const double *part, int n, int ncols, int stride, double *out, hipStream_t st) {
    hipLaunchKernelGGL(k_max_partials, dim3(ncols), dim3(256), 0, st, part, n, stride, out);
}

int launch_plaquette(int group, int prec, const void *links, const Geom &g, double *part, hipStream_t st) {
    const int nt = (g.vh + 255) / 256;
    const int slice = g.n[1] * g.n[2] * g.half;  // slots per x0-slice and parity
    const int tps = (slice % 256 == 0 && (slice / 256) % 8 == 0 && g.vh == g.n[0] * slice) ? slice / 256 : 0;
    if (group == 3) {
        if (prec == 32) hipLaunchKernelGGL((k_plaquette_pp<G3<float>>), dim3(nt), dim3(512), 0, st, (const float2 *)links, g, tps, part);
        else hipLaunchKernelGGL((k_plaquette<G3<double>>), dim3(nt, 2), dim3(256), 0, st, (const double2 *)links, g, part);
    } else {
        if (prec == 32) hipLaunchKernelGGL((k_plaquette_pp<G2<float>>), dim3(nt), dim3(512), 0, st, (const float4 *)links, g, tps, part);
        else hipLaunchKernelGGL((k_plaquette_pp<G2<double>>), dim3(nt), dim3(512), 0, st, (const double4 *)links, g, tps, part);
    }
    return 2 * nt;
}

// ------------------------------------------------------- Wilson loops
// For every site x, every plane (mu, nu) of the plane set and every
// 1 <= R <= rmax, 1 <= T <= tmax: the rectangular loop R steps along mu,
// T along nu, R back along mu, T back along nu (path_holonomy.rs:219-260,
// lib.rs:438-481), as an ordered left-to-right product.  Accumulates
// (1/N) Re Tr and (1/N) Im Tr; part[((r*tmax + t)*2 + k) * nblk + blk].
template <class G>
__global__ __launch_bounds__(128) void k_wilson_loops(const typename G::Store *__restrict__ L, Geom g, int plane_mu,
                                                      int plane_nu, int rmax, int tmax, double *__restrict__ part) {
    __shared__ double red[2];
    const int s = blockIdx.x * 128 + threadIdx.x;
    const int p = blockIdx.y;
    const int nblk = gridDim.x * gridDim.y;
    const int blin = blockIdx.y * gridDim.x + blockIdx.x;
    int x[4];
    const bool valid = s < g.vh && site_of(g, p, s, x);
    for (int r = 1; r <= rmax; ++r) {
        for (int t = 1; t <= tmax; ++t) {
            double re = 0.0, im = 0.0;
            if (valid) {
                typename G::Mat h = G::id();
                int c[4] = {x[0], x[1], x[2], x[3]};
                for (int k = 0; k < r; ++k) { h = G::mul(h, G::load(L, g, c, plane_mu)); int y[4]; shifted(g, c, plane_mu, +1, y); c[0]=y[0]; c[1]=y[1]; c[2]=y[2]; c[3]=y[3]; }
                for (int k = 0; k < t; ++k) { h = G::mul(h, G::load(L, g, c, plane_nu)); int y[4]; shifted(g, c, plane_nu, +1, y); c[0]=y[0]; c[1]=y[1]; c[2]=y[2]; c[3]=y[3]; }
                for (int k = 0; k < r; ++k) { int y[4]; shifted(g, c, plane_mu, -1, y); c[0]=y[0]; c[1]=y[1]; c[2]=y[2]; c[3]=y[3]; h = G::mul(h, G::dag(G::load(L, g, c, plane_mu))); }
                for (int k = 0; k < t; ++k) { int y[4]; shifted(g, c, plane_nu, -1, y); c[0]=y[0]; c[1]=y[1]; c[2]=y[2]; c[3]=y[3]; h = G::mul(h, G::dag(G::load(L, g, c, plane_nu))); }
                re = G::retr(h);
                im = G::imtr(h);
            }
            const double tre = block_sum<128>(re, red);
            const double tim = block_sum<128>(im, red);
            if (threadIdx.x == 0) {
                const int idx = ((r - 1) * tmax + (t - 1)) * 2;
                part[(size_t)idx * nblk + blin] += tre;
                part[(size_t)(idx + 1) * nblk + blin] += tim;
            }
        }
    }
}
int launch_wilson_loops(int group, int prec, const void *links, const Geom &g, int mu, int nu, int rmax, int tmax,
                        double *part, hipStream_t st) {
    dim3 grid((g.vh + 127) / 128, 2);
    if (group == 3) {
        if (prec == 32) hipLaunchKernelGGL((k_wilson_loops<G3<float>>), grid, dim3(128), 0, st, (const float2 *)links, g, mu, nu, rmax, tmax, part);
        else hipLaunchKernelGGL((k_wilson_loops<G3<double>>), grid, dim3(128), 0, st, (const double2 *)links, g, mu, nu, rmax, tmax, part);
    } else {
        if (prec == 32) hipLaunchKernelGGL((k_wilson_loops<G2<float>>), grid, dim3(128), 0, st, (const float4 *)links, g, mu, nu, rmax, tmax, part);
        else hipLaunchKernelGGL((k_wilson_loops<G2<double>>), grid, dim3(128), 0, st, (const double4 *)links, g, mu, nu, rmax, tmax, part);
    }
    return (int)(grid.x * grid.y);
}

// ---------------------------------------------------------- Polyakov loop
// One thread per site of the hyperplane x_dir = 0: ordered product of the
// N_dir links along +dir (su3_lattice.rs:185-212, lib.rs:483-505), (1/N) Tr.
// part[(k) * nblk + blk], k = 0 re, 1 im, 2 |P|.
template <class G>
__global__ __launch_bounds__(256) void k_polyakov(const typename G::Store *__restrict__ L, Geom g, int dir, int nbase,
                                                  double *__restrict__ part) {
    __shared__ double red[4];
    const int b = blockIdx.x * 256 + threadIdx.x;
    double re = 0.0, im = 0.0, ab = 0.0;
    if (b < nbase) {
        int x[4];
        int rem = b;
        for (int d = 3; d >= 0; --d) {
            if (d == dir) { x[d] = 0; continue; }
            x[d] = rem % g.n[d];
            rem /= g.n[d];
        }
        typename G::Mat h = G::id();
        for (int k = 0; k < g.n[dir]; ++k) {
            h = G::mul(h, G::load(L, g, x, dir));
            int y[4];
            shifted(g, x, dir, +1, y);
            x[0] = y[0]; x[1] = y[1]; x[2] = y[2]; x[3] = y[3];
        }
        re = G::retr(h);
        im = G::imtr(h);
        ab = __builtin_sqrt(re * re + im * im);
    }
    const double a0 = block_sum<256>(re, red);
    const double a1 = block_sum<256>(im, red);
    const double a2 = block_sum<256>(ab, red);
    if (threadIdx.x == 0) {
        part[0 * gridDim.x + blockIdx.x] = a0;
        part[1 * gridDim.x + blockIdx.x] = a1;
        part[2 * gridDim.x + blockIdx.x] = a2;
    }
}
// ------------------------------------------------------- path holonomy
// One thread walks the path (path_holonomy.rs:132-170, lib.rs:384-436): step
// +(d+1) multiplies by U_d(x) and moves to x + d; step -(d+1) moves to
// x - d and multiplies by U_d(x - d)^dagger; ordered left to right in FP64.
// out: 18 (SU(3)) or 4 (SU(2)) doubles, then the end point (4 doubles).
template <typename Real, int GROUP>
__global__ void k_path_holonomy(const void *__restrict__ links, Geom g, int x0, int x1, int x2, int x3,
                                const int *__restrict__ steps, int nsteps, double *__restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int c[4] = {x0, x1, x2, x3}, y[4];
    if constexpr (GROUP == 3) {
        const auto *L = (const typename Vec2<Real>::type *)links;
        M3<double> h = identity3<double>();
        for (int k = 0; k < nsteps; ++k) {
            const int code = steps[k];
            const int d = (code > 0 ? code : -code) - 1;
            M3<double> seg;
            if (code > 0) {
                const M3<Real> u = load3<Real>(L, g.vh, d, parity_of(c), slot_of(g, c));
                for (int e = 0; e < 18; ++e) seg.v[e] = (double)u.v[e];
                shifted(g, c, d, +1, y);
            } else {
                shifted(g, c, d, -1, y);
                const M3<Real> u = load3<Real>(L, g.vh, d, parity_of(y), slot_of(g, y));
                M3<double> ud;
                for (int e = 0; e < 18; ++e) ud.v[e] = (double)u.v[e];
                seg = dagger3(ud);
            }
            h = mul_nn(h, seg);
            for (int e = 0; e < 4; ++e) c[e] = y[e];
        }
        for (int e = 0; e < 18; ++e) out[e] = h.v[e];
        for (int e = 0; e < 4; ++e) out[18 + e] = c[e];
    } else {
        const auto *L = (const typename Vec4<Real>::type *)links;
        Q4<double> h;
        h.c[0] = 1.0; h.c[1] = 0.0; h.c[2] = 0.0; h.c[3] = 0.0;
        for (int k = 0; k < nsteps; ++k) {
            const int code = steps[k];
            const int d = (code > 0 ? code : -code) - 1;
            Q4<double> seg;
            if (code > 0) {
                const Q4<Real> u = load2<Real>(L, g.vh, d, parity_of(c), slot_of(g, c));
                for (int e = 0; e < 4; ++e) seg.c[e] = (double)u.c[e];
                shifted(g, c, d, +1, y);
            } else {
                shifted(g, c, d, -1, y);
                const Q4<Real> u = load2<Real>(L, g.vh, d, parity_of(y), slot_of(g, y));
                Q4<double> ud;
                for (int e = 0; e < 4; ++e) ud.c[e] = (double)u.c[e];
                seg = qdag(ud);
            }
            h = qmul(h, seg);
            for (int e = 0; e < 4; ++e) c[e] = y[e];
        }
        for (int e = 0; e < 4; ++e) out[e] = h.c[e];
        for (int e = 0; e < 4; ++e) out[4 + e] = c[e];
    }
}
void launch_path_holonomy(int group, int prec, const void *links, const Geom &g, const int x[4], const int *steps,
                          int nsteps, double *out, hipStream_t st) {
    if (group == 3) {
        if (prec == 32) hipLaunchKernelGGL((k_path_holonomy<float, 3>), dim3(1), dim3(64), 0, st, links, g, x[0], x[1], x[2], x[3], steps, nsteps, out);
        else hipLaunchKernelGGL((k_path_holonomy<double, 3>), dim3(1), dim3(64), 0, st, links, g, x[0], x[1], x[2], x[3], steps, nsteps, out);
    } else {
        if (prec == 32) hipLaunchKernelGGL((k_path_holonomy<float, 2>), dim3(1), dim3(64), 0, st, links, g, x[0], x[1], x[2], x[3], steps, nsteps, out);
        else hipLaunchKernelGGL((k_path_holonomy<double, 2>), dim3(1), dim3(64), 0, st, links, g, x[0], x[1], x[2], x[3], steps, nsteps, out);
    }
}

int launch_polyakov(int group, int prec, const void *links, const Geom &g, int dir, double *part, hipStream_t st) {
    const int nbase = g.volume / g.n[dir];
    dim3 grid((nbase + 255) / 256);
    if (group == 3) {
        if (prec == 32) hipLaunchKernelGGL((k_polyakov<G3<float>>), grid, dim3(256), 0, st, (const float2 *)links, g, dir, nbase, part);
        else hipLaunchKernelGGL((k_polyakov<G3<double>>), grid, dim3(256), 0, st, (const double2 *)links, g, dir, nbase, part);
    } else {
        if (prec == 32) hipLaunchKernelGGL((k_polyakov<G2<float>>), grid, dim3(256), 0, st, (const float4 *)links, g, dir, nbase, part);
        else hipLaunchKernelGGL((k_polyakov<G2<double>>), grid, dim3(256), 0, st, (const double4 *)links, g, dir, nbase, part);
    }
    return (int)grid.x;
}

// --------------------------------------------------- A1++ timeslice sums
// Time direction 3.  k_a1pp_sites computes every site's (ps, rs) in slot
// order (a1pp_slot, both parities of 256 slots per 512-thread block);
// k_a1pp then sums them per timeslice: blocks over (L^3 / 256, T), each
// reducing 256 consecutive spatial sites of one timeslice with
// block_sum<256>; part[(t*2 + k)*nblk + blk].
template <typename Real>
__global__ __launch_bounds__(512) void k_a1pp_sites(const typename Vec2<Real>::type *__restrict__ L, Geom g,
                                                    double *__restrict__ sv) {
    // contiguous tile ranges per XCD: neighbouring tiles share one L2
    a1pp_slot<Real>(L, g, threadIdx.x >> 8, xcd_tile(blockIdx.x, gridDim.x) * 256 + (threadIdx.x & 255), sv);
}
__global__ __launch_bounds__(256) void k_a1pp(const double *__restrict__ sv, Geom g, double *__restrict__ part) {
    __shared__ double red[4];
    const int b = blockIdx.x * 256 + threadIdx.x;
    const int t = blockIdx.y;
    const int nsp = g.n[0] * g.n[1] * g.n[2];
    double ps = 0.0, rs = 0.0;
    if (b < nsp) {
        const size_t i = (size_t)t * nsp + b;
        ps = sv[2 * i];
        rs = sv[2 * i + 1];
    }
    const double a0 = block_sum<256>(ps, red);
    const double a1 = block_sum<256>(rs, red);
    if (threadIdx.x == 0) {
        part[(size_t)(t * 2 + 0) * gridDim.x + blockIdx.x] = a0;
        part[(size_t)(t * 2 + 1) * gridDim.x + blockIdx.x] = a1;
    }
}
int launch_a1pp(int prec, const void *links, const Geom &g, double *sv, double *part, hipStream_t st) {
    const int nt = (g.vh + 255) / 256;
    if (prec == 32) hipLaunchKernelGGL((k_a1pp_sites<float>), dim3(nt), dim3(512), 0, st, (const float2 *)links, g, sv);
    else hipLaunchKernelGGL((k_a1pp_sites<double>), dim3(nt), dim3(512), 0, st, (const double2 *)links, g, sv);
    const int nsp = g.n[0] * g.n[1] * g.n[2];
    dim3 grid((nsp + 255) / 256, g.n[3]);
    hipLaunchKernelGGL(k_a1pp, grid, dim3(256), 0, st, (const double *)sv, g, part);
    return (int)grid.x;
}

// ------------------------------------------------------ spatial stout
// One level, out of place (glueball_smearing.rs:137-251): for spatial mu,
// C = rho * sum_{spatial nu != mu} [U_nu(x) U_mu(x+nu) U_nu(x+mu)^+ +
// U_nu(x-nu)^+ U_mu(x-nu) U_nu(x-nu+mu)], Omega = C U^+,
// Q = (i/2)(Omega^+ - Omega) - (i/6) Tr(Omega^+ - Omega), U' = exp(iQ) U.
// Staple products in the storage precision; Q, exp(iQ) (one-norm scaling to
// <= 1/2, Taylor until the term falls below 1e-18, squaring) and U' in FP64.
// Temporal links are copied.
template <typename Real>
__global__ __launch_bounds__(128) void k_stout(const typename Vec2<Real>::type *__restrict__ in,
                                               typename Vec2<Real>::type *__restrict__ out, Geom g, double rho) {
    // logical block: 128-slot block major, then the six (direction, parity)
    // groups; each XCD takes a contiguous range, so the x2- and x1-neighbours
    // of a block (16 and N2 N3 / 2 slots away) and the other groups of the
    // same slots are fetched into that XCD's L2 once
    const int lb = xcd_tile(blockIdx.x, gridDim.x);
    const int sb = lb / 6, r = lb - 6 * sb;
    const int s = sb * 128 + threadIdx.x;
    if (s < g.vh) stout_link<Real>(in, out, g, rho, r & 1, s, r >> 1);
}
// Spatial directions only: the temporal links are the same at every level,
// so the caller copies them into the output buffers once.
void launch_stout(int prec, const void *in, void *out, const Geom &g, double rho, hipStream_t st) {
    dim3 grid(((g.vh + 127) / 128) * 6);
    if (prec == 32) hipLaunchKernelGGL((k_stout<float>), grid, dim3(128), 0, st, (const float2 *)in, (float2 *)out, g, rho);
    else hipLaunchKernelGGL((k_stout<double>), grid, dim3(128), 0, st, (const double2 *)in, (double2 *)out, g, rho);
}

// ------------------------------------------------ spatial stride-2 blocking
// glueball_blocking.rs:105-136: coarse site y embeds at x = (2y0, 2y1, 2y2,
// y3); V_mu(y) = U_mu(x) U_mu(x + mu) for spatial mu, V_3(y) = U_3(x).  One
// thread per coarse site of parity p (grid.y); the fine lattice gf, the
// coarse gc.  The product runs in the storage precision.  Every coarse slot
// is written once; the padding slot of an odd N3 is left alone.
template <typename Real>
__global__ __launch_bounds__(256) void k_block2(const typename Vec2<Real>::type *__restrict__ in,
                                                typename Vec2<Real>::type *__restrict__ out, Geom gf, Geom gc) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    const int p = blockIdx.y;
    int y[4];
    if (s >= gc.vh || !site_of(gc, p, s, y)) return;
    const int x[4] = {2 * y[0], 2 * y[1], 2 * y[2], y[3]};
    const int px = parity_of(x), sx = slot_of(gf, x);
#pragma unroll 1
    for (int mu = 0; mu < 3; ++mu) {
        int xp[4];
        shifted(gf, x, mu, 1, xp);
        const M3<Real> a = load3<Real>(in, gf.vh, mu, px, sx);
        const M3<Real> b = load3<Real>(in, gf.vh, mu, px ^ 1, slot_of(gf, xp));
        store3<Real>(out, gc.vh, mu, p, s, mul_nn(a, b));
    }
    store3<Real>(out, gc.vh, 3, p, s, load3<Real>(in, gf.vh, 3, px, sx));
}
void launch_block2(int prec, const void *in, void *out, const Geom &gf, const Geom &gc, hipStream_t st) {
    dim3 grid((gc.vh + 255) / 256, 2);
    if (prec == 32) hipLaunchKernelGGL((k_block2<float>), grid, dim3(256), 0, st, (const float2 *)in, (float2 *)out, gf, gc);
    else hipLaunchKernelGGL((k_block2<double>), grid, dim3(256), 0, st, (const double2 *)in, (double2 *)out, gf, gc);
}

// ------------------------------------------------------- SU(2) residual
template <typename Real>
__global__ __launch_bounds__(256) void k_su2_residual(const typename Vec4<Real>::type *__restrict__ L, Geom g,
                                                      double *__restrict__ part) {
    __shared__ double red[4];
    const int s = blockIdx.x * 256 + threadIdx.x;
    const int p = blockIdx.y, mu = blockIdx.z;
    const int blin = (mu * 2 + p) * gridDim.x + blockIdx.x;
    double r = 0.0;
    int x[4];
    if (s < g.vh && site_of(g, p, s, x)) {
        const Q4<Real> q = load2<Real>(L, g.vh, mu, p, s);
        double n = 0.0;
        for (int c = 0; c < 4; ++c) n += (double)q.c[c] * (double)q.c[c];
        r = fabs(n - 1.0);
    }
    const double m = block_max<256>(r, red);
    if (threadIdx.x == 0) { part[2 * blin] = m; part[2 * blin + 1] = 0.0; }
}
int launch_su2_residual(int prec, const void *links, const Geom &g, double *part, hipStream_t st) {
    dim3 grid((g.vh + 255) / 256, 2, 4);
    if (prec == 32) hipLaunchKernelGGL((k_su2_residual<float>), grid, dim3(256), 0, st, (const float4 *)links, g, part);
    else hipLaunchKernelGGL((k_su2_residual<double>), grid, dim3(256), 0, st, (const double4 *)links, g, part);
    return (int)(grid.x * 8);
}

}  // namespace ym
