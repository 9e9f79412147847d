// Device functions of the gauge observables (group traits, plaquette site
// sum, rectangular loops, A1++ site sums, stout step), shared by the
// single-lattice and the batched-ensemble kernels.
#pragma once

#include "ym_device.hpp"

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

// Sum over mu < nu of (1/N) Re Tr P_{mu nu}(x) for the site of parity p at
// slot s (0 for the padding slot of odd N3), with
// P = [U_mu(x) U_nu(x+mu)] [U_nu(x) U_mu(x+nu)]^dagger.
template <class G>
__device__ __forceinline__ double plaquette_site(const typename G::Store *__restrict__ L, const Geom &g, int p,
                                                 int s) {
    double acc = 0.0;
    int x[4];
    if (!site_of(g, p, s, x)) return acc;
    // The four links at x are loaded once; the two neighbour links of plane
    // k + 1, U_nu(x + mu) and U_mu(x + nu), are in flight while plane k is
    // multiplied.  Planes in the order (0,1) (0,2) (0,3) (1,2) (1,3) (2,3).
    constexpr int PM[6] = {0, 0, 0, 1, 1, 2}, PN[6] = {1, 2, 3, 2, 3, 3};
    typename G::Mat U[4];
#pragma unroll
    for (int mu = 0; mu < 4; ++mu) U[mu] = G::load(L, g, x, mu);
    typename G::Mat A[2], B[2];
    {
        int xm[4], xn[4];
        shifted(g, x, PM[0], +1, xm);
        shifted(g, x, PN[0], +1, xn);
        A[0] = G::load(L, g, xm, PN[0]);
        B[0] = G::load(L, g, xn, PM[0]);
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const int b = k & 1;
        if (k + 1 < 6) {
            int xm[4], xn[4];
            shifted(g, x, PM[k + 1], +1, xm);
            shifted(g, x, PN[k + 1], +1, xn);
            A[b ^ 1] = G::load(L, g, xm, PN[k + 1]);
            B[b ^ 1] = G::load(L, g, xn, PM[k + 1]);
        }
        const typename G::Mat a = G::mul(U[PM[k]], A[b]);
        const typename G::Mat c = G::mul(U[PN[k]], B[b]);
        acc += G::retr_nd(a, c);
    }
    return acc;
}

// FP32 SU(3): every link read as rows 0 and 1 (48 of 72 bytes).  A plane's
// a = U_mu(x) U_nu(x+mu) needs rows 0 and 1 of U_mu(x) and the whole
// neighbour, whose third row is rebuilt as conj(row 0 x row 1); a's own
// third row is rebuilt the same way, and likewise c = U_nu(x) U_mu(x+nu).
template <>
__device__ __forceinline__ double plaquette_site<G3<float>>(const float2 *__restrict__ L, const Geom &g, int p,
                                                            int s) {
    double acc = 0.0;
    int x[4];
    if (!site_of(g, p, s, x)) return acc;
    constexpr int PM[6] = {0, 0, 0, 1, 1, 2}, PN[6] = {1, 2, 3, 2, 3, 3};
    M3<float> U[4];
#pragma unroll
    for (int mu = 0; mu < 4; ++mu) U[mu] = load3_rows01(L, g.vh, mu, p, s);
    M3<float> A[2], B[2];
    {
        int xm[4], xn[4];
        shifted(g, x, PM[0], +1, xm);
        shifted(g, x, PN[0], +1, xn);
        A[0] = load3_rows01(L, g.vh, PN[0], parity_of(xm), slot_of(g, xm));
        B[0] = load3_rows01(L, g.vh, PM[0], parity_of(xn), slot_of(g, xn));
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const int b = k & 1;
        if (k + 1 < 6) {
            int xm[4], xn[4];
            shifted(g, x, PM[k + 1], +1, xm);
            shifted(g, x, PN[k + 1], +1, xn);
            A[b ^ 1] = load3_rows01(L, g.vh, PN[k + 1], parity_of(xm), slot_of(g, xm));
            B[b ^ 1] = load3_rows01(L, g.vh, PM[k + 1], parity_of(xn), slot_of(g, xn));
        }
        M3<float> a, c;
        recon_row2(A[b]);
        mul_nn_rows01(U[PM[k]], A[b], a);
        recon_row2(a);
        recon_row2(B[b]);
        mul_nn_rows01(U[PN[k]], B[b], c);
        recon_row2(c);
        acc += G3<float>::retr_nd(a, c);
    }
    return acc;
}

// (1/3) Re Tr of the SU(3) rectangle from x: R links along M, T along N,
// R back along M (daggered), T back along N (daggered), multiplied left to
// right with mul_nn's FMA chain in the storage precision
// (path_holonomy.rs:219-260).  The product starts from the first link (an
// identity start would add only exact zeros) and only the diagonal of the
// last product is formed.
template <typename Real, int M, int N, int R, int T>
__device__ __forceinline__ double rect_re3(const typename Vec2<Real>::type *__restrict__ L, const Geom &g,
                                           const int x0[4]) {
    using G = G3<Real>;
    int c[4] = {x0[0], x0[1], x0[2], x0[3]}, y[4];
    auto step = [&](int d, int s) {
        shifted(g, c, d, s, y);
        c[0] = y[0]; c[1] = y[1]; c[2] = y[2]; c[3] = y[3];
    };
    if constexpr (sizeof(Real) == 4) {
        // FP32: links read as rows 0 and 1.  Rows 0 and 1 of the running
        // product need only rows 0 and 1 of the product so far and the whole
        // next link (third row rebuilt as conj(row 0 x row 1)); the loop's
        // last product P is formed on its upper 2 x 2 block and
        // Re P(2,2) = Re(P00 P11 - P01 P10), the cofactor of an SU(3) matrix.
        M3<float> h = load3_rows01(L, g.vh, M, parity_of(c), slot_of(g, c)), t, Lk;
        step(M, +1);
        auto fwd = [&](int d) {
            Lk = load3_rows01(L, g.vh, d, parity_of(c), slot_of(g, c));
            recon_row2(Lk);
            mul_nn_rows01(h, Lk, t);
            h = t;
            step(d, +1);
        };
        auto bwd = [&](int d) {
            step(d, -1);
            Lk = load3_rows01(L, g.vh, d, parity_of(c), slot_of(g, c));
            recon_row2(Lk);
            mul_nd_rows01(h, Lk, t);
            h = t;
        };
#pragma unroll
        for (int k = 1; k < R; ++k) fwd(M);
#pragma unroll
        for (int k = 0; k < T; ++k) fwd(N);
#pragma unroll
        for (int k = 0; k < R; ++k) bwd(M);
#pragma unroll
        for (int k = 0; k < T - 1; ++k) bwd(N);
        step(N, -1);
        Lk = load3_rows01(L, g.vh, N, parity_of(c), slot_of(g, c));
        recon_row2(Lk);
        float pr[2][2], pim[2][2];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                float re = YM_RE(h, i, 0) * YM_RE(Lk, j, 0);
                re = fmaR(YM_IM(h, i, 0), YM_IM(Lk, j, 0), re);
                re = fmaR(YM_RE(h, i, 1), YM_RE(Lk, j, 1), re);
                re = fmaR(YM_IM(h, i, 1), YM_IM(Lk, j, 1), re);
                re = fmaR(YM_RE(h, i, 2), YM_RE(Lk, j, 2), re);
                re = fmaR(YM_IM(h, i, 2), YM_IM(Lk, j, 2), re);
                float im = YM_IM(h, i, 0) * YM_RE(Lk, j, 0);
                im = fmaR(-YM_RE(h, i, 0), YM_IM(Lk, j, 0), im);
                im = fmaR(YM_IM(h, i, 1), YM_RE(Lk, j, 1), im);
                im = fmaR(-YM_RE(h, i, 1), YM_IM(Lk, j, 1), im);
                im = fmaR(YM_IM(h, i, 2), YM_RE(Lk, j, 2), im);
                im = fmaR(-YM_RE(h, i, 2), YM_IM(Lk, j, 2), im);
                pr[i][j] = re;
                pim[i][j] = im;
            }
        float d2 = pr[0][0] * pr[1][1];
        d2 = fmaR(-pim[0][0], pim[1][1], d2);
        d2 = fmaR(-pr[0][1], pr[1][0], d2);
        d2 = fmaR(pim[0][1], pim[1][0], d2);
        return ((double)pr[0][0] + (double)pr[1][1] + (double)d2) / 3.0;
    }
    M3<Real> h = G::load(L, g, c, M);
    step(M, +1);
#pragma unroll
    for (int k = 1; k < R; ++k) { h = mul_nn(h, G::load(L, g, c, M)); step(M, +1); }
#pragma unroll
    for (int k = 0; k < T; ++k) { h = mul_nn(h, G::load(L, g, c, N)); step(N, +1); }
#pragma unroll
    for (int k = 0; k < R; ++k) { step(M, -1); h = mul_nn(h, dagger3(G::load(L, g, c, M))); }
#pragma unroll
    for (int k = 0; k < T - 1; ++k) { step(N, -1); h = mul_nn(h, dagger3(G::load(L, g, c, N))); }
    step(N, -1);
    const M3<Real> Y = dagger3(G::load(L, g, c, N));
    Real d[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        Real re = YM_RE(h, i, 0) * YM_RE(Y, 0, i);
        re = fmaR(-YM_IM(h, i, 0), YM_IM(Y, 0, i), re);
        re = fmaR(YM_RE(h, i, 1), YM_RE(Y, 1, i), re);
        re = fmaR(-YM_IM(h, i, 1), YM_IM(Y, 1, i), re);
        re = fmaR(YM_RE(h, i, 2), YM_RE(Y, 2, i), re);
        re = fmaR(-YM_IM(h, i, 2), YM_IM(Y, 2, i), re);
        d[i] = re;
    }
    return ((double)d[0] + (double)d[1] + (double)d[2]) / 3.0;
}

// A1++ values of the site x (time direction 3): the three 1x1 spatial
// plaquettes (0,1),(0,2),(1,2) into ps and the six 2x1 rectangles
// (0,1),(0,2),(1,0),(1,2),(2,0),(2,1) into rs, (1/3) Re Tr each, added in
// that order (glueball_operators.rs:24-26, 93-160).  A scheduling barrier
// after each loop keeps the compiler from hoisting the loads of all nine
// loops at once.
template <typename Real>
__device__ __forceinline__ void a1pp_site_x(const typename Vec2<Real>::type *__restrict__ L, const Geom &g,
                                            const int x[4], double &ps, double &rs) {
    ps += rect_re3<Real, 0, 1, 1, 1>(L, g, x);
    __builtin_amdgcn_sched_barrier(0);
    ps += rect_re3<Real, 0, 2, 1, 1>(L, g, x);
    __builtin_amdgcn_sched_barrier(0);
    ps += rect_re3<Real, 1, 2, 1, 1>(L, g, x);
    __builtin_amdgcn_sched_barrier(0);
    rs += rect_re3<Real, 0, 1, 2, 1>(L, g, x);
    __builtin_amdgcn_sched_barrier(0);
    rs += rect_re3<Real, 0, 2, 2, 1>(L, g, x);
    __builtin_amdgcn_sched_barrier(0);
    rs += rect_re3<Real, 1, 0, 2, 1>(L, g, x);
    __builtin_amdgcn_sched_barrier(0);
    rs += rect_re3<Real, 1, 2, 2, 1>(L, g, x);
    __builtin_amdgcn_sched_barrier(0);
    rs += rect_re3<Real, 2, 0, 2, 1>(L, g, x);
    __builtin_amdgcn_sched_barrier(0);
    rs += rect_re3<Real, 2, 1, 2, 1>(L, g, x);
}

// Per-site A1++ values (ps, rs) of the site at slot s of parity p into
// sv[2 i], sv[2 i + 1] with i = t * N0 N1 N2 + (x0 N1 + x1) N2 + x2
// (timeslice-major spatial order, x2 fastest).  Run over the slots of both
// parities so that a wavefront's link loads read consecutive slots: a
// timeslice is a strided subset of the slots, and visiting it directly
// gathers several times more cache lines per load.
template <typename Real>
__device__ __forceinline__ void a1pp_slot(const typename Vec2<Real>::type *__restrict__ L, const Geom &g, int p, int s,
                                          double *__restrict__ sv) {
    int x[4];
    if (s >= g.vh || !site_of(g, p, s, x)) return;
    double ps = 0.0, rs = 0.0;
    a1pp_site_x<Real>(L, g, x, ps, rs);
    const size_t i = (size_t)x[3] * (size_t)(g.n[0] * g.n[1] * g.n[2]) + (size_t)((x[0] * g.n[1] + x[1]) * g.n[2] + x[2]);
    sv[2 * i] = ps;
    sv[2 * i + 1] = rs;
}

// ------------------------------------------------------ spatial stout
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

// exp(iQ) of a traceless Hermitian 3 x 3 matrix by the Cayley-Hamilton
// theorem (Morningstar & Peardon, Phys. Rev. D 69, 054501 (2004), section
// III): exp(iQ) = f0 I + f1 Q + f2 Q^2 with c0 = det Q = Tr(Q^3)/3,
// c1 = Tr(Q^2)/2, theta = arccos(|c0| / c0max) / 3, c0max = 2 (c1/3)^(3/2),
// u = sqrt(c1/3) cos theta, w = sqrt(c1) sin theta and
//   h0 = (u^2 - w^2) e^{2iu} + e^{-iu} (8 u^2 cos w + 2iu (3u^2 + w^2) xi0(w)),
//   h1 = 2u e^{2iu} - e^{-iu} (2u cos w - i (3u^2 - w^2) xi0(w)),
//   h2 = e^{2iu} - e^{-iu} (cos w + 3iu xi0(w)),
//   f_j = h_j / (9u^2 - w^2),  xi0(w) = sin(w) / w;
// for c0 < 0 the f_j of |c0| give f_j(c0) = (-1)^j conj(f_j(|c0|)).  With
// |c0| the angle stays in [0, pi/6], so 9u^2 - w^2 >= 8u^2 and only Q = 0
// makes the denominator vanish (then exp(iQ) = 1 + iQ - Q^2/2 to the order
// that matters).  FP64 scalars; one matrix product (Q^2).
__device__ __forceinline__ M3<double> expm_i_hermitian_ch(const M3<double> &Q) {
    const M3<double> Q2 = mul_nn(Q, Q);
    double trq3 = 0.0, trq2 = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        trq2 += YM_RE(Q2, i, i);
#pragma unroll
        for (int k = 0; k < 3; ++k) trq3 += YM_RE(Q, i, k) * YM_RE(Q2, k, i) - YM_IM(Q, i, k) * YM_IM(Q2, k, i);
    }
    const double c0 = trq3 / 3.0, c1 = 0.5 * trq2;
    double f0r, f0i, f1r, f1i, f2r, f2i;
    if (!(c1 > 1.0e-30)) {
        f0r = 1.0; f0i = 0.0; f1r = 0.0; f1i = 1.0; f2r = -0.5; f2i = 0.0;
    } else {
        const double neg = c0 < 0.0;
        const double s13 = sqrt(c1 / 3.0);
        const double c0max = 2.0 * (c1 / 3.0) * s13;
        const double ratio = fmin(fabs(c0) / c0max, 1.0);
        const double theta = acos(ratio) / 3.0;
        double st, ct;
        sincos(theta, &st, &ct);
        const double u = s13 * ct, w = sqrt(c1) * st;
        const double u2 = u * u, w2 = w * w;
        double sw, cw, su, cu, s2u, c2u;
        sincos(w, &sw, &cw);
        sincos(u, &su, &cu);
        sincos(2.0 * u, &s2u, &c2u);
        const double xi0 = fabs(w) < 0.05 ? 1.0 - w2 / 6.0 * (1.0 - w2 / 20.0 * (1.0 - w2 / 42.0)) : sw / w;
        // e^{-iu} = cu - i su, e^{2iu} = c2u + i s2u
        // h0
        const double a0r = 8.0 * u2 * cw, a0i = 2.0 * u * (3.0 * u2 + w2) * xi0;
        const double h0r = (u2 - w2) * c2u + (cu * a0r + su * a0i);
        const double h0i = (u2 - w2) * s2u + (cu * a0i - su * a0r);
        // h1 = 2u e^{2iu} - e^{-iu} (2u cos w - i (3u^2 - w^2) xi0)
        const double a1r = 2.0 * u * cw, a1i = -(3.0 * u2 - w2) * xi0;
        const double h1r = 2.0 * u * c2u - (cu * a1r + su * a1i);
        const double h1i = 2.0 * u * s2u - (cu * a1i - su * a1r);
        // h2 = e^{2iu} - e^{-iu} (cos w + 3iu xi0)
        const double a2r = cw, a2i = 3.0 * u * xi0;
        const double h2r = c2u - (cu * a2r + su * a2i);
        const double h2i = s2u - (cu * a2i - su * a2r);
        const double inv = 1.0 / (9.0 * u2 - w2);
        f0r = h0r * inv; f0i = h0i * inv;
        f1r = h1r * inv; f1i = h1i * inv;
        f2r = h2r * inv; f2i = h2i * inv;
        if (neg) {
            f0i = -f0i;
            f1r = -f1r;
            f2i = -f2i;
        }
    }
    M3<double> E;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const double qr = YM_RE(Q, i, j), qi = YM_IM(Q, i, j), pr = YM_RE(Q2, i, j), pi = YM_IM(Q2, i, j);
            double er = f1r * qr - f1i * qi + f2r * pr - f2i * pi;
            double ei = f1r * qi + f1i * qr + f2r * pi + f2i * pr;
            if (i == j) {
                er += f0r;
                ei += f0i;
            }
            YM_RE(E, i, j) = er;
            YM_IM(E, i, j) = ei;
        }
    return E;
}

// One stout step of the spatial link (MU, parity p, slot s < vh) from `in`
// into `out`.  The direction is a template parameter so that every
// neighbour index is a compile-time choice: with a runtime direction the
// site arrays are indexed dynamically, which puts them in scratch memory
// when the geometry comes from memory (the batched kernel's per-chain
// table) instead of kernel arguments.
template <typename Real, int MU>
__device__ __forceinline__ void stout_link_dir(const typename Vec2<Real>::type *__restrict__ in,
                                               typename Vec2<Real>::type *__restrict__ out, const Geom &g, double rho,
                                               int p, int s) {
    using G = G3<Real>;
    constexpr int mu = MU;
    int x[4];
    if (!site_of(g, p, s, x)) {
        store3<Real>(out, g.vh, mu, p, s, load3<Real>(in, g.vh, mu, p, s));
        return;
    }
    M3<double> C;
#pragma unroll
    for (int c = 0; c < 18; ++c) C.v[c] = 0.0;
    int xpm[4];
    shifted(g, x, mu, +1, xpm);
#pragma unroll
    for (int nu = 0; nu < 3; ++nu) {
        if (nu == mu) continue;
        int xpn[4], xmn[4], xmnpm[4];
        shifted(g, x, nu, +1, xpn);
        shifted(g, x, nu, -1, xmn);
        shifted(g, xmn, mu, +1, xmnpm);
        typename G::Mat f, b;
        if constexpr (sizeof(Real) == 4) {
            // FP32: neighbours read as rows 0 and 1, third rows rebuilt as
            // conj(row 0 x row 1) where a product needs them.
            // f = U_nu(x) U_mu(x+nu) U_nu(x+mu)^+ = (X Y) Z^+
            M3<float> X = load3_rows01(in, g.vh, nu, parity_of(x), slot_of(g, x));
            M3<float> Y = load3_rows01(in, g.vh, mu, parity_of(xpn), slot_of(g, xpn));
            M3<float> Z = load3_rows01(in, g.vh, nu, parity_of(xpm), slot_of(g, xpm));
            M3<float> T;
            recon_row2(Y);
            mul_nn_rows01(X, Y, T);
            recon_row2(T);
            recon_row2(Z);
            mul_nd_rows01(T, Z, f);
            recon_row2(f);
            // b = U_nu(x-nu)^+ U_mu(x-nu) U_nu(x-nu+mu) = X^+ (Y Z)
            X = load3_rows01(in, g.vh, nu, parity_of(xmn), slot_of(g, xmn));
            Y = load3_rows01(in, g.vh, mu, parity_of(xmn), slot_of(g, xmn));
            Z = load3_rows01(in, g.vh, nu, parity_of(xmnpm), slot_of(g, xmnpm));
            recon_row2(Z);
            mul_nn_rows01(Y, Z, T);
            recon_row2(T);
            recon_row2(X);
            mul_dn_rows01(X, T, b);
            recon_row2(b);
        } else {
            f = G::mul(G::mul(G::load(in, g, x, nu), G::load(in, g, xpn, mu)), G::dag(G::load(in, g, xpm, nu)));
            b = G::mul(G::mul(G::dag(G::load(in, g, xmn, nu)), G::load(in, g, xmn, mu)), G::load(in, g, xmnpm, nu));
        }
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
    // FP64 storage keeps the reference's scaled Taylor series (the
    // rotation of smear_one_level, glueball_smearing.rs:236-238, is
    // Su3::from_algebra_coefficients -> matrix_exponential, su3.rs:395-450)
    // for parity to 1e-13; FP32 storage uses the closed form, accurate far
    // below its rounding at a fraction of the series' cost
    const M3<double> R = sizeof(Real) == 8 ? expm_i_hermitian(Q) : expm_i_hermitian_ch(Q);
    const M3<double> Un = mul_nn(R, U);
    M3<Real> o;
#pragma unroll
    for (int c = 0; c < 18; ++c) o.v[c] = (Real)Un.v[c];
    store3<Real>(out, g.vh, mu, p, s, o);
}

// The stout step of link (mu, p, s): spatial directions through
// stout_link_dir, the temporal one copied unchanged.
template <typename Real>
__device__ __forceinline__ void stout_link(const typename Vec2<Real>::type *__restrict__ in,
                                           typename Vec2<Real>::type *__restrict__ out, const Geom &g, double rho,
                                           int p, int s, int mu) {
    switch (mu) {
    case 0: stout_link_dir<Real, 0>(in, out, g, rho, p, s); break;
    case 1: stout_link_dir<Real, 1>(in, out, g, rho, p, s); break;
    case 2: stout_link_dir<Real, 2>(in, out, g, rho, p, s); break;
    default: store3<Real>(out, g.vh, mu, p, s, load3<Real>(in, g.vh, mu, p, s)); break;
    }
}

}  // namespace ym
