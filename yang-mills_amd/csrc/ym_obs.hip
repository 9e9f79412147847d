// Gauge-observable kernels for gfx950: plaquette, averaged rectangular Wilson
// loops, Polyakov loop, per-timeslice A1++ operator sums, residuals and the
// synchronous spatial stout smearing step.  Products run in the storage
// precision; every trace is accumulated in FP64 and reduced per block
// (64-wide wavefront shuffles + LDS), then summed in a fixed order by
// k_sum_partials, so results are deterministic run to run.
#include "ym_device.hpp"
#include "ym_internal.hpp"

namespace ym {

template <typename Real> struct G3 {
    using Mat = M3<Real>;
    using Store = typename Vec2<Real>::type;
    static constexpr int N = 3;
    __device__ static Mat load(const Store *L, const Geom &g, const int x[4], int mu) {
        return load3<Real>(L, g.vh, mu, parity_of(x), slot_of(g, x));
    }
    __device__ static Mat mul(const Mat &a, const Mat &b) { return mul_nn(a, b); }
    __device__ static Mat dag(const Mat &a) { return dagger3(a); }
    __device__ static Mat id() { return identity3<Real>(); }
    __device__ static double retr(const Mat &a) {
        return ((double)a.v[0] + (double)a.v[8] + (double)a.v[16]) / 3.0;
    }
    __device__ static double imtr(const Mat &a) {
        return ((double)a.v[1] + (double)a.v[9] + (double)a.v[17]) / 3.0;
    }
    // Re Tr(a b^dagger) / N
    __device__ static double retr_nd(const Mat &a, const Mat &b) {
        double t = 0.0;
#pragma unroll
        for (int c = 0; c < 18; ++c) t = __builtin_fma((double)a.v[c], (double)b.v[c], t);
        return t / 3.0;
    }
};
template <typename Real> struct G2 {
    using Mat = Q4<Real>;
    using Store = typename Vec4<Real>::type;
    static constexpr int N = 2;
    __device__ static Mat load(const Store *L, const Geom &g, const int x[4], int mu) {
        return load2<Real>(L, g.vh, mu, parity_of(x), slot_of(g, x));
    }
    __device__ static Mat mul(const Mat &a, const Mat &b) { return qmul(a, b); }
    __device__ static Mat dag(const Mat &a) { return qdag(a); }
    __device__ static Mat id() {
        Mat m;
        m.c[0] = 1; m.c[1] = 0; m.c[2] = 0; m.c[3] = 0;
        return m;
    }
    __device__ static double retr(const Mat &a) { return (double)a.c[0]; }
    __device__ static double imtr(const Mat &) { return 0.0; }
    // (1/2) Re Tr(a b^dagger) = a . b for unit quaternions
    __device__ static double retr_nd(const Mat &a, const Mat &b) {
        double t = (double)a.c[0] * (double)b.c[0];
        t = __builtin_fma((double)a.c[1], (double)b.c[1], t);
        t = __builtin_fma((double)a.c[2], (double)b.c[2], t);
        t = __builtin_fma((double)a.c[3], (double)b.c[3], t);
        return t;
    }
};

// ------------------------------------------------------------ plaquette
// Per site: sum over mu < nu of (1/N) Re Tr P_{mu nu}(x) with
// P = [U_mu(x) U_nu(x+mu)] [U_nu(x) U_mu(x+nu)]^dagger.  part[blk] = sum.
// Also accumulates per-timeslice spatial / temporal sums when tsl != null:
// tsl[(t*2 + kind) * nblk + blk] (kind 0 spatial planes, 1 temporal planes).
template <class G>
__global__ __launch_bounds__(256) void k_plaquette(const typename G::Store *__restrict__ L, Geom g,
                                                   double *__restrict__ part) {
    __shared__ double red[4];
    const int s = blockIdx.x * 256 + threadIdx.x;
    const int p = blockIdx.y;
    double acc = 0.0;
    int x[4];
    if (s < g.vh && site_of(g, p, s, x)) {
#pragma unroll
        for (int mu = 0; mu < 4; ++mu) {
            int xm[4];
            shifted(g, x, mu, +1, xm);
            const typename G::Mat Umu = G::load(L, g, x, mu);
#pragma unroll
            for (int nu = mu + 1; nu < 4; ++nu) {
                int xn[4];
                shifted(g, x, nu, +1, xn);
                const typename G::Mat a = G::mul(Umu, G::load(L, g, xm, nu));
                const typename G::Mat b = G::mul(G::load(L, g, x, nu), G::load(L, g, xn, mu));
                acc += G::retr_nd(a, b);
            }
        }
    }
    const double t = block_sum<256>(acc, red);
    if (threadIdx.x == 0) part[blockIdx.y * gridDim.x + blockIdx.x] = t;
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
    dim3 grid((g.vh + 255) / 256, 2);
    if (group == 3) {
        if (prec == 32) hipLaunchKernelGGL((k_plaquette<G3<float>>), grid, dim3(256), 0, st, (const float2 *)links, g, part);
        else hipLaunchKernelGGL((k_plaquette<G3<double>>), grid, dim3(256), 0, st, (const double2 *)links, g, part);
    } else {
        if (prec == 32) hipLaunchKernelGGL((k_plaquette<G2<float>>), grid, dim3(256), 0, st, (const float4 *)links, g, part);
        else hipLaunchKernelGGL((k_plaquette<G2<double>>), grid, dim3(256), 0, st, (const double4 *)links, g, part);
    }
    return (int)(grid.x * 2);
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
// Time direction 3.  Per spatial site (x0,x1,x2) at slice t: the three 1x1
// spatial plaquettes (0,1),(0,2),(1,2) and the six 2x1 rectangles
// (0,1),(0,2),(1,0),(1,2),(2,0),(2,1) (glueball_operators.rs:24-26, 93-160),
// (1/3) Re Tr each.  Grid (blocks over L^3, T); part[(t*2 + k)*nblk + blk].
template <typename Real>
__device__ __forceinline__ double rect_loop_re(const typename Vec2<Real>::type *__restrict__ L, const Geom &g,
                                               const int x0[4], int m, int n, int r, int t) {
    using G = G3<Real>;
    typename G::Mat h = G::id();
    int c[4] = {x0[0], x0[1], x0[2], x0[3]}, y[4];
    for (int k = 0; k < r; ++k) { h = G::mul(h, G::load(L, g, c, m)); shifted(g, c, m, +1, y); c[0]=y[0]; c[1]=y[1]; c[2]=y[2]; c[3]=y[3]; }
    for (int k = 0; k < t; ++k) { h = G::mul(h, G::load(L, g, c, n)); shifted(g, c, n, +1, y); c[0]=y[0]; c[1]=y[1]; c[2]=y[2]; c[3]=y[3]; }
    for (int k = 0; k < r; ++k) { shifted(g, c, m, -1, y); c[0]=y[0]; c[1]=y[1]; c[2]=y[2]; c[3]=y[3]; h = G::mul(h, G::dag(G::load(L, g, c, m))); }
    for (int k = 0; k < t; ++k) { shifted(g, c, n, -1, y); c[0]=y[0]; c[1]=y[1]; c[2]=y[2]; c[3]=y[3]; h = G::mul(h, G::dag(G::load(L, g, c, n))); }
    return G::retr(h);
}
template <typename Real>
__global__ __launch_bounds__(256) void k_a1pp(const typename Vec2<Real>::type *__restrict__ L, Geom g,
                                              double *__restrict__ part) {
    __shared__ double red[4];
    const int b = blockIdx.x * 256 + threadIdx.x;
    const int t = blockIdx.y;
    const int nsp = g.n[0] * g.n[1] * g.n[2];
    double ps = 0.0, rs = 0.0;
    if (b < nsp) {
        int x[4];
        x[2] = b % g.n[2];
        x[1] = (b / g.n[2]) % g.n[1];
        x[0] = b / (g.n[2] * g.n[1]);
        x[3] = t;
        ps += rect_loop_re<Real>(L, g, x, 0, 1, 1, 1);
        ps += rect_loop_re<Real>(L, g, x, 0, 2, 1, 1);
        ps += rect_loop_re<Real>(L, g, x, 1, 2, 1, 1);
        rs += rect_loop_re<Real>(L, g, x, 0, 1, 2, 1);
        rs += rect_loop_re<Real>(L, g, x, 0, 2, 2, 1);
        rs += rect_loop_re<Real>(L, g, x, 1, 0, 2, 1);
        rs += rect_loop_re<Real>(L, g, x, 1, 2, 2, 1);
        rs += rect_loop_re<Real>(L, g, x, 2, 0, 2, 1);
        rs += rect_loop_re<Real>(L, g, x, 2, 1, 2, 1);
    }
    const double a0 = block_sum<256>(ps, red);
    const double a1 = block_sum<256>(rs, red);
    if (threadIdx.x == 0) {
        part[(size_t)(t * 2 + 0) * gridDim.x + blockIdx.x] = a0;
        part[(size_t)(t * 2 + 1) * gridDim.x + blockIdx.x] = a1;
    }
}
int launch_a1pp(int prec, const void *links, const Geom &g, double *part, hipStream_t st) {
    const int nsp = g.n[0] * g.n[1] * g.n[2];
    dim3 grid((nsp + 255) / 256, g.n[3]);
    if (prec == 32) hipLaunchKernelGGL((k_a1pp<float>), grid, dim3(256), 0, st, (const float2 *)links, g, part);
    else hipLaunchKernelGGL((k_a1pp<double>), grid, dim3(256), 0, st, (const double2 *)links, g, part);
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
__device__ __forceinline__ M3<double> expm_i_hermitian(const M3<double> &Q) {
    // X = i Q
    M3<double> X;
#pragma unroll
    for (int c = 0; c < 9; ++c) { X.v[2 * c] = -Q.v[2 * c + 1]; X.v[2 * c + 1] = Q.v[2 * c]; }
    double norm = 0.0;
    for (int j = 0; j < 3; ++j) {
        double cs = 0.0;
        for (int i = 0; i < 3; ++i) cs += hypot(YM_RE(X, i, j), YM_IM(X, i, j));
        norm = fmax(norm, cs);
    }
    int sq = 0;
    while (norm > 0.5 && sq < 32) { norm *= 0.5; ++sq; }
    const double scale = __builtin_ldexp(1.0, -sq);
#pragma unroll
    for (int c = 0; c < 18; ++c) X.v[c] *= scale;
    M3<double> sum = identity3<double>(), term = identity3<double>();
    for (int k = 1; k <= 64; ++k) {
        term = mul_nn(term, X);
        const double inv = 1.0 / (double)k;
        double mx = 0.0;
#pragma unroll
        for (int c = 0; c < 18; ++c) { term.v[c] *= inv; sum.v[c] += term.v[c]; mx = fmax(mx, fabs(term.v[c])); }
        if (mx <= 1.0e-18) break;
    }
    for (int k = 0; k < sq; ++k) sum = mul_nn(sum, sum);
    return sum;
}

template <typename Real>
__global__ __launch_bounds__(128) void k_stout(const typename Vec2<Real>::type *__restrict__ in,
                                               typename Vec2<Real>::type *__restrict__ out, Geom g, double rho) {
    using G = G3<Real>;
    const int s = blockIdx.x * 128 + threadIdx.x;
    const int p = blockIdx.y, mu = blockIdx.z;
    int x[4];
    if (s >= g.vh) return;
    if (!site_of(g, p, s, x) || mu == 3) {
        store3<Real>(out, g.vh, mu, p, s, load3<Real>(in, g.vh, mu, p, s));
        return;
    }
    M3<double> C;
#pragma unroll
    for (int c = 0; c < 18; ++c) C.v[c] = 0.0;
    int xpm[4];
    shifted(g, x, mu, +1, xpm);
    for (int nu = 0; nu < 3; ++nu) {
        if (nu == mu) continue;
        int xpn[4], xmn[4], xmnpm[4];
        shifted(g, x, nu, +1, xpn);
        shifted(g, x, nu, -1, xmn);
        shifted(g, xmn, mu, +1, xmnpm);
        const typename G::Mat f = G::mul(G::mul(G::load(in, g, x, nu), G::load(in, g, xpn, mu)), G::dag(G::load(in, g, xpm, nu)));
        const typename G::Mat b = G::mul(G::mul(G::dag(G::load(in, g, xmn, nu)), G::load(in, g, xmn, mu)), G::load(in, g, xmnpm, nu));
#pragma unroll
        for (int c = 0; c < 18; ++c) C.v[c] += (double)f.v[c] + (double)b.v[c];
    }
#pragma unroll
    for (int c = 0; c < 18; ++c) C.v[c] *= rho;
    const typename G::Mat Uf = load3<Real>(in, g.vh, mu, p, s);
    M3<double> U;
#pragma unroll
    for (int c = 0; c < 18; ++c) U.v[c] = (double)Uf.v[c];
    const M3<double> Om = mul_nn(C, dagger3(U));
    // A = Omega^+ - Omega (anti-Hermitian); Q = (i/2) A - (i/6) Tr(A) 1
    M3<double> Q;
    double trr = 0.0, tri = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        trr += YM_RE(Om, i, i) - YM_RE(Om, i, i);
        tri += -YM_IM(Om, i, i) - YM_IM(Om, i, i);
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const double ar = YM_RE(Om, j, i) - YM_RE(Om, i, j);
            const double ai = -YM_IM(Om, j, i) - YM_IM(Om, i, j);
            // (i/2)(ar + i ai) = (-ai/2) + i (ar/2)
            double qr = -0.5 * ai, qi = 0.5 * ar;
            if (i == j) {
                // - (i/6)(trr + i tri) = (tri/6) - i (trr/6)
                qr += tri / 6.0;
                qi -= trr / 6.0;
            }
            YM_RE(Q, i, j) = qr;
            YM_IM(Q, i, j) = qi;
        }
    const M3<double> R = expm_i_hermitian(Q);
    const M3<double> Un = mul_nn(R, U);
    M3<Real> o;
#pragma unroll
    for (int c = 0; c < 18; ++c) o.v[c] = (Real)Un.v[c];
    store3<Real>(out, g.vh, mu, p, s, o);
}
void launch_stout(int prec, const void *in, void *out, const Geom &g, double rho, hipStream_t st) {
    dim3 grid((g.vh + 127) / 128, 2, 4);
    if (prec == 32) hipLaunchKernelGGL((k_stout<float>), grid, dim3(128), 0, st, (const float2 *)in, (float2 *)out, g, rho);
    else hipLaunchKernelGGL((k_stout<double>), grid, dim3(128), 0, st, (const double2 *)in, (double2 *)out, g, rho);
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
