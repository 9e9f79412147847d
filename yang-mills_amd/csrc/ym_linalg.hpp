// FP64 3x3 complex linear algebra for the learned SU(3) proposal: unitary
// polar factor (Newton iteration), Hermitian eigenvalues (closed form) and
// eigenvectors (cyclic complex Jacobi), the principal logarithm of a unitary
// matrix, the Gell-Mann coordinates of su(3), and the Haar-measure Jacobian
// of the exponential map.  Written for host and device so that the same
// source can be checked on the CPU.  Multiply-adds may be fused here:
// nothing in this file is replayed bit for bit, its callers are checked
// against tolerances.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

#pragma clang fp contract(fast)

namespace ym {
namespace la {

struct cx {
    double r, i;
};
__host__ __device__ __forceinline__ cx cmul(cx a, cx b) { return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
__host__ __device__ __forceinline__ cx cadd(cx a, cx b) { return {a.r + b.r, a.i + b.i}; }
__host__ __device__ __forceinline__ cx csub(cx a, cx b) { return {a.r - b.r, a.i - b.i}; }
__host__ __device__ __forceinline__ cx cconj(cx a) { return {a.r, -a.i}; }
__host__ __device__ __forceinline__ cx cscale(cx a, double s) { return {a.r * s, a.i * s}; }
__host__ __device__ __forceinline__ double cabs2(cx a) { return a.r * a.r + a.i * a.i; }
// s + a b as four fused multiply-adds
__host__ __device__ __forceinline__ cx cmadd(cx s, cx a, cx b) {
    return {__builtin_fma(-a.i, b.i, __builtin_fma(a.r, b.r, s.r)), __builtin_fma(a.i, b.r, __builtin_fma(a.r, b.i, s.i))};
}

struct C3 {
    cx m[3][3];
};

__host__ __device__ __forceinline__ C3 c3_mul(const C3 &a, const C3 &b) {
    C3 c;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            cx s = cmul(a.m[i][0], b.m[0][j]);
            s = cmadd(s, a.m[i][1], b.m[1][j]);
            s = cmadd(s, a.m[i][2], b.m[2][j]);
            c.m[i][j] = s;
        }
    return c;
}
__host__ __device__ __forceinline__ C3 c3_dag(const C3 &a) {
    C3 c;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) c.m[i][j] = cconj(a.m[j][i]);
    return c;
}
__host__ __device__ __forceinline__ C3 c3_identity() {
    C3 c;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) c.m[i][j] = {i == j ? 1.0 : 0.0, 0.0};
    return c;
}
__host__ __device__ __forceinline__ cx c3_det(const C3 &a) {
    const cx t0 = csub(cmul(a.m[1][1], a.m[2][2]), cmul(a.m[1][2], a.m[2][1]));
    const cx t1 = csub(cmul(a.m[1][0], a.m[2][2]), cmul(a.m[1][2], a.m[2][0]));
    const cx t2 = csub(cmul(a.m[1][0], a.m[2][1]), cmul(a.m[1][1], a.m[2][0]));
    return cadd(csub(cmul(a.m[0][0], t0), cmul(a.m[0][1], t1)), cmul(a.m[0][2], t2));
}
__host__ __device__ __forceinline__ cx c3_trace(const C3 &a) { return cadd(cadd(a.m[0][0], a.m[1][1]), a.m[2][2]); }

// Inverse by the adjugate; returns false when |det| is not positive.
__host__ __device__ __forceinline__ bool c3_inverse(const C3 &a, C3 &inv) {
    const cx d = c3_det(a);
    const double n = cabs2(d);
    if (!(n > 0.0)) return false;
    const cx di = {d.r / n, -d.i / n};
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            // cofactor of a(j, i)
            const int r0 = (j + 1) % 3, r1 = (j + 2) % 3, c0 = (i + 1) % 3, c1 = (i + 2) % 3;
            const cx cof = csub(cmul(a.m[r0][c0], a.m[r1][c1]), cmul(a.m[r0][c1], a.m[r1][c0]));
            inv.m[i][j] = cmul(cof, di);
        }
    return true;
}

// Unitary polar factor of a (a = W P, P Hermitian positive): Newton
// iteration W <- (z W + W^{-dagger} / z) / 2 with Frobenius-norm scaling
// z = (||W^{-1}||_F / ||W||_F)^{1/2} (Higham 1986), until the update is
// below 1e-15 (at most 20 steps).  Returns false for a singular input.
__host__ __device__ __forceinline__ bool c3_polar_unitary(const C3 &a, C3 &w) {
    w = a;
    for (int it = 0; it < 20; ++it) {
        C3 inv;
        if (!c3_inverse(w, inv)) return false;
        double fw = 0.0, fi = 0.0;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                fw += cabs2(w.m[i][j]);
                fi += cabs2(inv.m[i][j]);
            }
        const double z = sqrt(sqrt(fi / fw));
        const double zi = 1.0 / z;
        double delta = 0.0;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const cx t = cconj(inv.m[j][i]);
                const cx nw = cadd(cscale(w.m[i][j], 0.5 * z), cscale(t, 0.5 * zi));
                delta += cabs2(csub(nw, w.m[i][j]));
                w.m[i][j] = nw;
            }
        if (!(delta > 1e-30)) break;
    }
    return true;
}

// Eigenvalues of a Hermitian 3x3 matrix, descending (trigonometric solution
// of the characteristic cubic).
__host__ __device__ __forceinline__ void herm_eigvals(const C3 &h, double ev[3]) {
    const double m = (h.m[0][0].r + h.m[1][1].r + h.m[2][2].r) / 3.0;
    const double a = h.m[0][0].r - m, b = h.m[1][1].r - m, c = h.m[2][2].r - m;
    const double o01 = cabs2(h.m[0][1]), o02 = cabs2(h.m[0][2]), o12 = cabs2(h.m[1][2]);
    const double p2 = (a * a + b * b + c * c + 2.0 * (o01 + o02 + o12)) / 6.0;
    if (!(p2 > 1e-300)) {
        ev[0] = ev[1] = ev[2] = m;
        return;
    }
    const double p = sqrt(p2);
    // det(K) / 2 with K = h - m I
    const cx x = cmul(cmul(h.m[0][1], h.m[1][2]), h.m[2][0]);
    const double detk = a * b * c + 2.0 * x.r - a * o12 - b * o02 - c * o01;
    double r = detk / (2.0 * p * p2);
    r = r < -1.0 ? -1.0 : (r > 1.0 ? 1.0 : r);
    const double phi = acos(r) / 3.0;
    const double two_pi_3 = 2.0943951023931954923;
    ev[0] = m + 2.0 * p * cos(phi);
    ev[2] = m + 2.0 * p * cos(phi + two_pi_3);
    ev[1] = 3.0 * m - ev[0] - ev[2];
}

// Cyclic complex Jacobi on a Hermitian matrix: h <- V^dagger h V diagonal,
// columns of v are the eigenvectors.  Sweeps until the off-diagonal norm is
// below 1e-18 of the diagonal (at most 10).
__host__ __device__ __forceinline__ void herm_jacobi(C3 &h, C3 &v) {
    v = c3_identity();
    for (int sweep = 0; sweep < 10; ++sweep) {
        // converged when the off-diagonal part is negligible against the
        // diagonal (quadratic convergence: 3-5 sweeps in FP64)
        const double off = cabs2(h.m[0][1]) + cabs2(h.m[0][2]) + cabs2(h.m[1][2]);
        const double dia = h.m[0][0].r * h.m[0][0].r + h.m[1][1].r * h.m[1][1].r + h.m[2][2].r * h.m[2][2].r;
        if (!(off > 1e-36 * dia) || !(off > 1e-300)) break;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int p = k == 2 ? 1 : 0, q = k == 0 ? 1 : 2;
            const double apq = sqrt(cabs2(h.m[p][q]));
            if (!(apq > 1e-300)) continue;
            const double dpp = h.m[p][p].r, dqq = h.m[q][q].r;
            if (apq <= 1e-18 * (fabs(dpp) + fabs(dqq))) continue;
            // h_pq = apq e^{i alpha}; rotate in the (p, q) plane
            const cx e = {h.m[p][q].r / apq, h.m[p][q].i / apq};
            const double tau = (dqq - dpp) / (2.0 * apq);
            const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
            const double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
            // G = [[c, s e], [-s conj(e), c]] on columns p, q: h <- G^dagger h G
            const cx se = cscale(e, s), sec = cscale(cconj(e), s);
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                const cx hp = h.m[r][p], hq = h.m[r][q];
                h.m[r][p] = csub(cscale(hp, c), cmul(hq, sec));
                h.m[r][q] = cadd(cmul(hp, se), cscale(hq, c));
                const cx vp = v.m[r][p], vq = v.m[r][q];
                v.m[r][p] = csub(cscale(vp, c), cmul(vq, sec));
                v.m[r][q] = cadd(cmul(vp, se), cscale(vq, c));
            }
#pragma unroll
            for (int col = 0; col < 3; ++col) {
                const cx hp = h.m[p][col], hq = h.m[q][col];
                h.m[p][col] = csub(cscale(hp, c), cmul(se, hq));
                h.m[q][col] = cadd(cmul(sec, hp), cscale(hq, c));
            }
            h.m[p][q] = {0.0, 0.0};
            h.m[q][p] = {0.0, 0.0};
            h.m[p][p].i = 0.0;
            h.m[q][q].i = 0.0;
        }
    }
}

// Gell-Mann matrices lambda_1..lambda_8 (su3.rs:325-352 ordering), entries.
// X = sum_a eps_a lambda_a / 2 (Hermitian, traceless).
__host__ __device__ __forceinline__ C3 su3_from_coords(const double eps[8]) {
    const double k8 = 0.28867513459481288225;  // 1 / (2 sqrt 3)
    C3 x;
    x.m[0][0] = {0.5 * eps[2] + k8 * eps[7], 0.0};
    x.m[1][1] = {-0.5 * eps[2] + k8 * eps[7], 0.0};
    x.m[2][2] = {-2.0 * k8 * eps[7], 0.0};
    x.m[0][1] = {0.5 * eps[0], -0.5 * eps[1]};
    x.m[1][0] = {0.5 * eps[0], 0.5 * eps[1]};
    x.m[0][2] = {0.5 * eps[3], -0.5 * eps[4]};
    x.m[2][0] = {0.5 * eps[3], 0.5 * eps[4]};
    x.m[1][2] = {0.5 * eps[5], -0.5 * eps[6]};
    x.m[2][1] = {0.5 * eps[5], 0.5 * eps[6]};
    return x;
}
// eps_a = Tr(X lambda_a) for Hermitian X.
__host__ __device__ __forceinline__ void su3_coords(const C3 &x, double eps[8]) {
    const double s3 = 1.7320508075688772935;
    eps[0] = x.m[0][1].r + x.m[1][0].r;
    eps[1] = x.m[1][0].i - x.m[0][1].i;
    eps[2] = x.m[0][0].r - x.m[1][1].r;
    eps[3] = x.m[0][2].r + x.m[2][0].r;
    eps[4] = x.m[2][0].i - x.m[0][2].i;
    eps[5] = x.m[1][2].r + x.m[2][1].r;
    eps[6] = x.m[2][1].i - x.m[1][2].i;
    eps[7] = (x.m[0][0].r + x.m[1][1].r - 2.0 * x.m[2][2].r) / s3;
}

// exp(i X) for Hermitian X through its eigen-decomposition.
__host__ __device__ __forceinline__ C3 expi_hermitian(const C3 &x) {
    C3 h = x, v;
    herm_jacobi(h, v);
    C3 out;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            cx s = {0.0, 0.0};
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const double ph = h.m[k][k].r;
                const cx e = {cos(ph), sin(ph)};
                s = cmadd(s, cmul(v.m[i][k], e), cconj(v.m[j][k]));
            }
            out.m[i][j] = s;
        }
    return out;
}

// Principal logarithm of a unitary y with det y = 1: y = exp(i X), X
// Hermitian traceless with eigenvalues in (-pi, pi).  The eigenvectors come
// from Jacobi on (y - y^dagger)/(2i) + c (y + y^dagger)/2, whose eigenvalue
// sin(phi) + c cos(phi) separates the phases of a link near the identity.
// Returns false (the link is outside the chart) when an eigenvector does not
// diagonalise y to 1e-9, the phases do not sum to zero, or a phase is within
// 1e-6 of +-pi.  phases (optional) receives the eigenvalues of X.
__host__ __device__ __forceinline__ bool su3_log(const C3 &y, C3 &x, double phases[3]) {
    const double c = 0.3819660112501051518;  // 2 - golden ratio
    C3 h;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const cx a = y.m[i][j], b = cconj(y.m[j][i]);
            // (a - b) / (2i) + c (a + b) / 2
            h.m[i][j] = {0.5 * (a.i - b.i) + 0.5 * c * (a.r + b.r), -0.5 * (a.r - b.r) + 0.5 * c * (a.i + b.i)};
        }
    C3 v;
    herm_jacobi(h, v);
    double ph[3];
    double sum = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        // z = v_k^dagger y v_k and the residual |y v_k - z v_k|
        cx yv[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            cx s = {0.0, 0.0};
#pragma unroll
            for (int j = 0; j < 3; ++j) s = cmadd(s, y.m[i][j], v.m[j][k]);
            yv[i] = s;
        }
        cx z = {0.0, 0.0};
#pragma unroll
        for (int i = 0; i < 3; ++i) z = cmadd(z, cconj(v.m[i][k]), yv[i]);
        double res = 0.0;
#pragma unroll
        for (int i = 0; i < 3; ++i) res += cabs2(csub(yv[i], cmul(z, v.m[i][k])));
        if (!(res <= 1e-18)) return false;
        ph[k] = atan2(z.i, z.r);
        if (!(fabs(ph[k]) < 3.14159165358979)) return false;  // pi - 1e-6
        sum += ph[k];
    }
    if (!(fabs(sum) < 1e-9)) return false;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            cx s = {0.0, 0.0};
#pragma unroll
            for (int k = 0; k < 3; ++k) s = cadd(s, cscale(cmul(v.m[i][k], cconj(v.m[j][k])), ph[k]));
            x.m[i][j] = s;
        }
    if (phases)
#pragma unroll
        for (int k = 0; k < 3; ++k) phases[k] = ph[k];
    return true;
}

// Functions of a Hermitian 3x3 h by Lagrange interpolation on its spectrum.
// With m the mean eigenvalue, K = h - m I and k = ev - m,
//   f(h) = sum_k f(ev_k) P_k,  P_k = (K - k_j)(K - k_l) w_k,
//   w_k = 1 / ((k_k - k_j)(k_k - k_l))   ({j, l} the other two indices),
// so f(h) = c2 K^2 + c1 K + c0 I with c2 = sum f_k w_k,
// c1 = -sum f_k (k_j + k_l) w_k, c0 = sum f_k k_j k_l w_k: one matrix
// product (K^2) for any f.  The shift by m keeps K^2 free of cancellation.
// herm_spectral returns false when two eigenvalues are closer than 1e-3 of
// the spread (the weights lose accuracy there; callers fall back to Jacobi
// or Newton).
struct Spectral {
    C3 K, K2;
    double w[3], s[3], p[3];  // w_k, k_j + k_l, k_j k_l
};
__host__ __device__ __forceinline__ bool herm_spectral(const C3 &h, const double ev[3], Spectral &S) {
    const double spread = ev[0] - ev[2];
    const double g01 = ev[0] - ev[1], g12 = ev[1] - ev[2];
    if (!(spread > 1e-300) || !(g01 > 1e-3 * spread) || !(g12 > 1e-3 * spread)) return false;
    const double m = (h.m[0][0].r + h.m[1][1].r + h.m[2][2].r) / 3.0;
    S.K = h;
#pragma unroll
    for (int d = 0; d < 3; ++d) S.K.m[d][d].r -= m;
    S.K2 = c3_mul(S.K, S.K);
    const double k0 = ev[0] - m, k1 = ev[1] - m, k2 = ev[2] - m;
    S.w[0] = 1.0 / (g01 * spread);
    S.w[1] = -1.0 / (g01 * g12);
    S.w[2] = 1.0 / (spread * g12);
    S.s[0] = k1 + k2;
    S.p[0] = k1 * k2;
    S.s[1] = k0 + k2;
    S.p[1] = k0 * k2;
    S.s[2] = k0 + k1;
    S.p[2] = k0 * k1;
    return true;
}
__host__ __device__ __forceinline__ C3 spectral_real(const Spectral &S, const double f[3]) {
    double c2 = 0.0, c1 = 0.0, c0 = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double fw = f[k] * S.w[k];
        c2 += fw;
        c1 -= fw * S.s[k];
        c0 += fw * S.p[k];
    }
    C3 o;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            o.m[i][j] = {c2 * S.K2.m[i][j].r + c1 * S.K.m[i][j].r + (i == j ? c0 : 0.0),
                         c2 * S.K2.m[i][j].i + c1 * S.K.m[i][j].i};
    return o;
}
__host__ __device__ __forceinline__ C3 spectral_cx(const Spectral &S, const cx f[3]) {
    cx c2 = {0.0, 0.0}, c1 = {0.0, 0.0}, c0 = {0.0, 0.0};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const cx fw = cscale(f[k], S.w[k]);
        c2 = cadd(c2, fw);
        c1 = csub(c1, cscale(fw, S.s[k]));
        c0 = cadd(c0, cscale(fw, S.p[k]));
    }
    C3 o;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            cx t = cmadd(cmul(c2, S.K2.m[i][j]), c1, S.K.m[i][j]);
            if (i == j) t = cadd(t, c0);
            o.m[i][j] = t;
        }
    return o;
}
// Tr(a b)
__host__ __device__ __forceinline__ cx c3_trace_mul(const C3 &a, const C3 &b) {
    cx t = {0.0, 0.0};
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) t = cmadd(t, a.m[i][j], b.m[j][i]);
    return t;
}

// Unitary polar factor of a through the spectrum of h = a^dagger a =
// Q diag(s^2) Q^dagger: w = a h^{-1/2}; falls back to the Newton iteration
// when two singular values nearly coincide or a is nearly singular.  ev: the
// eigenvalues s_k^2 of h (descending), for either path.
__host__ __device__ __forceinline__ bool c3_polar_unitary_fast(const C3 &a, C3 &w, double ev[3]) {
    const C3 h = c3_mul(c3_dag(a), a);
    herm_eigvals(h, ev);
    Spectral S;
    if (ev[2] > 1e-4 * ev[0] && herm_spectral(h, ev, S)) {
        const double f[3] = {1.0 / sqrt(ev[0]), 1.0 / sqrt(ev[1]), 1.0 / sqrt(ev[2])};
        w = c3_mul(a, spectral_real(S, f));
        return true;
    }
    return c3_polar_unitary(a, w);
}

// exp(i X) for Hermitian X, with its eigenvalues x (descending when the
// interpolation applies) and e_k = exp(i x_k); Jacobi when two eigenvalues
// nearly coincide.
__host__ __device__ __forceinline__ C3 expi_hermitian_fast(const C3 &x, double ev[3], cx e[3]) {
    herm_eigvals(x, ev);
    Spectral S;
    if (herm_spectral(x, ev, S)) {
#pragma unroll
        for (int k = 0; k < 3; ++k) e[k] = {cos(ev[k]), sin(ev[k])};
        return spectral_cx(S, e);
    }
    C3 h = x, v;
    herm_jacobi(h, v);
    C3 out;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        ev[k] = h.m[k][k].r;
        e[k] = {cos(ev[k]), sin(ev[k])};
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            cx t = {0.0, 0.0};
#pragma unroll
            for (int k = 0; k < 3; ++k) t = cmadd(t, cmul(v.m[i][k], e[k]), cconj(v.m[j][k]));
            out.m[i][j] = t;
        }
    return out;
}

// Principal logarithm as su3_log, first by interpolation on the spectrum of
// the Hermitian H = (y - y^dagger)/(2i) + c (y + y^dagger)/2 (which commutes
// with y): z_k = Tr(y P_k) = exp(i phi_k); falls back to su3_log when two
// eigenvalues of H nearly coincide or a projector is not rank one
// (|z_k| != 1).  z (unit modulus) receives exp(i phases).
__host__ __device__ __forceinline__ bool su3_log_fast(const C3 &y, C3 &x, double phases[3], cx z[3]) {
    const double c = 0.3819660112501051518;
    C3 h;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const cx a = y.m[i][j], b = cconj(y.m[j][i]);
            h.m[i][j] = {0.5 * (a.i - b.i) + 0.5 * c * (a.r + b.r), -0.5 * (a.r - b.r) + 0.5 * c * (a.i + b.i)};
        }
    double ev[3];
    herm_eigvals(h, ev);
    Spectral S;
    if (herm_spectral(h, ev, S)) {
        const cx t2 = c3_trace_mul(y, S.K2), t1 = c3_trace_mul(y, S.K), t0 = c3_trace(y);
        double ph[3], sum = 0.0;
        bool good = true;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const cx zk = cscale(cadd(csub(t2, cscale(t1, S.s[k])), cscale(t0, S.p[k])), S.w[k]);
            const double n2 = cabs2(zk);
            good = good && fabs(n2 - 1.0) < 1e-9;
            ph[k] = atan2(zk.i, zk.r);
            z[k] = cscale(zk, 1.0 / sqrt(n2));
            sum += ph[k];
        }
        if (good) {
            if (!(fabs(sum) < 1e-9)) return false;
#pragma unroll
            for (int k = 0; k < 3; ++k)
                if (!(fabs(ph[k]) < 3.14159165358979)) return false;
            x = spectral_real(S, ph);
            if (phases)
#pragma unroll
                for (int k = 0; k < 3; ++k) phases[k] = ph[k];
            return true;
        }
    }
    double ph[3];
    if (!su3_log(y, x, ph)) return false;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        z[k] = {cos(ph[k]), sin(ph[k])};
        if (phases) phases[k] = ph[k];
    }
    return true;
}

// log of the Haar density of exp(i X) in the coordinates eps (X = eps.lambda/2)
// relative to Lebesgue measure, up to a constant:
// sum_{j<k} 2 log sinc((x_j - x_k) / 2), x the eigenvalues of X, from x and
// z_k = exp(i x_k):
// sinc^2((x_j - x_k)/2) = |z_j - z_k|^2 / (x_j - x_k)^2, one logarithm for
// the three pairs (1 - d^2/12 for |d| < 1e-6, exact to 1e-25 there).
__host__ __device__ __forceinline__ double haar_log_jacobian(const double x[3], const cx z[3]) {
    double num = 1.0, den = 1.0;
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int k = j + 1; k < 3; ++k) {
            const double d = x[j] - x[k];
            if (fabs(d) < 1e-6) {
                num *= 1.0 - d * d / 12.0;
            } else {
                num *= cabs2(csub(z[j], z[k]));
                den *= d * d;
            }
        }
    return log(num / den);
}

}  // namespace la
}  // namespace ym

#pragma clang fp contract(off)
