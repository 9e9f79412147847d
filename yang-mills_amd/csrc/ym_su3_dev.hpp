// Device functions of the SU(3) single-generator Metropolis update, shared
// by the single-lattice and the batched-ensemble kernels.
#pragma once

#include "ym_device.hpp"

namespace ym {

// The Metropolis decision dS <= 0 or u < exp_neg(dS), exactly that FP64 rule,
// without evaluating exp_neg for almost every proposal.  a = 2^(-dS log2 e)
// by the hardware's FP32 exp2 (v_exp_f32) differs from exp_neg(dS) by less
// than tol = 2^-12 a + 2^-126: the FP32 rounding of the exponent t (|t| <=
// 149 wherever a > 0) moves a by at most 149 * 2^-24 * ln 2 < 2^-17
// relative, v_exp_f32 adds about 1 ulp of FP32 (2^-23), exp_neg is within
// 2 ulp of FP64, and below 2^-126 only the absolute term matters.  So
// u < a - tol decides accept and u >= a + tol decides reject exactly as
// exp_neg would; only u inside that band (probability < 2^-10 per
// proposal) takes the FP64 path.  Each test is one FMA and a compare:
// a (1 - 2^-12) - u > 2^-126 and u - a (1 + 2^-12) >= 2^-126, the FMA
// rounding its exact difference by at most 2^-53 relative (the sign and
// all but 2^-179 of the 2^-126 margin survive), far inside the gap between
// the 2^-17 error and the 2^-12 tolerance.  Decisions, counts and links
// are therefore bit for bit those of the plain rule.
__device__ __forceinline__ bool metropolis_accept(double dS, double u) {
    if (!(dS > 0.0)) return true;
    const float t = (float)(dS * -0x1.71547652b82fep+0);
    const double a = (double)__builtin_amdgcn_exp2f(t);
    if (__builtin_fma(a, 0x1.ffep-1, -u) > 0x1p-126) return true;
    if (__builtin_fma(-a, 0x1.001p+0, u) >= 0x1p-126) return false;
    return u < exp_neg(dS);
}

// Proposal angles up to this bound use the short sincos series: every
// argument of a rotation is then at most 0.5 / 2 (generators 0-6) or
// 2 * 0.5 / (2 sqrt 3) (generator 7) in magnitude, inside its 0.29 range.
constexpr double kShortSeriesMaxAngle = 0.5;

__device__ __forceinline__ void sincos_hit(double x, bool short_series, double &s, double &cm1) {
    if (short_series) sincos_short(x, s, cm1);
    else sincos_small(x, s, cm1);
}

// Parts of M = u a (a the staple in FP64) that a proposal needs, each a
// single FMA chain over k: re_chain continues acc with +-Re (u a)_{rb},
// im_chain with +-Im (u a)_{rb}; FIRST starts the chain with a product,
// NEG subtracts.  A proposal reads rows of the current u, so no copy of
// M is carried or rotated between hits.
template <bool FIRST, bool NEG>
__device__ __forceinline__ double re_chain(const M3<double> &u, const M3<double> &a, int r, int b, double acc) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double ur = NEG ? -YM_RE(u, r, k) : YM_RE(u, r, k);
        const double ui = NEG ? YM_IM(u, r, k) : -YM_IM(u, r, k);
        acc = (FIRST && k == 0) ? ur * YM_RE(a, k, b) : __builtin_fma(ur, YM_RE(a, k, b), acc);
        acc = __builtin_fma(ui, YM_IM(a, k, b), acc);
    }
    return acc;
}
template <bool FIRST, bool NEG>
__device__ __forceinline__ double im_chain(const M3<double> &u, const M3<double> &a, int r, int b, double acc) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double ur = NEG ? -YM_RE(u, r, k) : YM_RE(u, r, k);
        const double ui = NEG ? -YM_IM(u, r, k) : YM_IM(u, r, k);
        acc = (FIRST && k == 0) ? ur * YM_IM(a, k, b) : __builtin_fma(ur, YM_IM(a, k, b), acc);
        acc = __builtin_fma(ui, YM_RE(a, k, b), acc);
    }
    return acc;
}

// One single-generator proposal on the FP64 link u with staple a (FP64):
// generator GEN = h mod 8, angle from the reference's paired grid, accept
// iff dS <= 0 or u < exp(-dS), then u <- R u, where a rejected proposal
// applies the identity rotation (c, s) = (1, 0) so the update is
// branch-free and leaves every entry unchanged.  dS = -(beta/3) Re tr((R -
// 1) u a) from the 2 x 2 block (i, j) of u a (all three diagonal entries
// for generator 7), with t1 = Re M_ii + Re M_jj and t2 as one chain each.
template <int GEN, int MU>
__device__ __forceinline__ void su3_hit(M3<double> &u, const M3<double> &a, uint64_t key, uint32_t site, int h,
                                        uint64_t sweep, double angle, double nbt, unsigned long long &nacc) {
    constexpr int bi = (GEN <= 2 || GEN == 7) ? 0 : (GEN <= 4 ? 0 : 1);
    constexpr int bj = (GEN <= 2 || GEN == 7) ? 1 : 2;
    uint32_t o[4];
    draw(key, site, MU, KIND_SU3_HIT, (uint32_t)h, sweep, o);
    // theta for generator 7; theta / 2 for generators 0-6, drawn directly:
    // (angle / 2) * unit is exactly (angle * unit) / 2, since scaling by 2^-1
    // commutes with the rounding.
    const double ang = angle_from_raw(o[0], o[1], GEN == 7 ? angle : 0.5 * angle);
    const double uacc = unit53(o[2], o[3]);
    const bool short_series = angle <= kShortSeriesMaxAngle;  // uniform over the grid
    double sn, cm1, s2 = 0.0, cm2 = 0.0, dre;
    const double t1 = re_chain<false, false>(u, a, bj, bj, re_chain<true, false>(u, a, bi, bi, 0.0));
    if constexpr (GEN == 7) {
        const double psi = ang * 0x1.279a74590331cp-2;
        sincos_hit(psi, short_series, sn, cm1);
        sincos_hit(-2.0 * psi, short_series, s2, cm2);
        const double t2 = im_chain<false, false>(u, a, 1, 1, im_chain<true, false>(u, a, 0, 0, 0.0));
        const double r22 = re_chain<true, false>(u, a, 2, 2, 0.0);
        const double i22 = im_chain<true, false>(u, a, 2, 2, 0.0);
        const double d1 = __builtin_fma(cm1, t1, -(sn * t2));
        const double d2 = __builtin_fma(cm2, r22, -(s2 * i22));
        dre = d1 + d2;
    } else if constexpr (GEN == 2) {
        sincos_hit(ang, short_series, sn, cm1);
        const double t2 = im_chain<false, true>(u, a, 1, 1, im_chain<true, false>(u, a, 0, 0, 0.0));
        dre = __builtin_fma(cm1, t1, -(sn * t2));
    } else if constexpr (GEN == 0 || GEN == 3 || GEN == 5) {
        sincos_hit(ang, short_series, sn, cm1);
        const double t2 = im_chain<false, false>(u, a, bj, bi, im_chain<true, false>(u, a, bi, bj, 0.0));
        dre = __builtin_fma(cm1, t1, -(sn * t2));
    } else {
        sincos_hit(ang, short_series, sn, cm1);
        const double t2 = re_chain<false, true>(u, a, bi, bj, re_chain<true, false>(u, a, bj, bi, 0.0));
        dre = __builtin_fma(cm1, t1, sn * t2);
    }
    const double dS = nbt * dre;
    const bool accept = metropolis_accept(dS, uacc);
    nacc += accept ? 1ull : 0ull;
    const double c = accept ? 1.0 + cm1 : 1.0;
    const double sa = accept ? sn : 0.0;
    if constexpr (GEN == 7) {
        const double c2 = accept ? 1.0 + cm2 : 1.0;
        const double s2a = accept ? s2 : 0.0;
        rot_phase(u, 0, c, sa); rot_phase(u, 1, c, sa); rot_phase(u, 2, c2, s2a);
    } else if constexpr (GEN == 2) {
        rot_phase(u, 0, c, sa); rot_phase(u, 1, c, -sa);
    } else if constexpr (GEN == 0 || GEN == 3 || GEN == 5) {
        rot_sym(u, bi, bj, c, sa);
    } else {
        rot_asym(u, bi, bj, c, sa);
    }
}

// Hits H .. HITS-1 (generator h mod 8) as straight-line code, for a hit
// count known at compile time: the compiler then schedules consecutive hits
// together (one hit's Philox draw beside the previous hit's FP64 work),
// which a loop over a runtime count with a generator switch prevents.
template <int H, int HITS, int MU>
__device__ __forceinline__ void su3_hits_unrolled(M3<double> &u, const M3<double> &a, uint64_t key, uint32_t site,
                                                  uint64_t sweep, double angle, double nbt,
                                                  unsigned long long &nacc) {
    if constexpr (H < HITS) {
        su3_hit<H % 8, MU>(u, a, key, site, H, sweep, angle, nbt, nacc);
        su3_hits_unrolled<H + 1, HITS, MU>(u, a, key, site, sweep, angle, nbt, nacc);
    }
}

// Staple A(x) of the link U_MU(x), x of parity p at slot s: the six staples
// in the storage precision with the fixed FMA chain (mul_nd/acc_nd forward,
// mul_dd/acc_nn backward, nu ascending).  The three links of term k + 1 are
// loaded before term k is multiplied (FP32 storage), so a wave keeps two
// terms' loads in flight instead of waiting on each term's links in turn; the
// arithmetic and its order are unchanged.
template <typename Real, int MU>
__device__ __forceinline__ void su3_staple_at(const typename Vec2<Real>::type *__restrict__ L, const Geom &g, int p,
                                              int s, const int x[4], M3<Real> &A) {
    const int q = p ^ 1;
#pragma unroll
    for (int c = 0; c < 18; ++c) A.v[c] = 0;
    int xpm[4];
    shifted(g, x, MU, +1, xpm);
    const int s_pm = slot_of(g, xpm);
    // term 2k: forward staple of direction nu_k, X = U_nu(x+mu), Y = U_mu(x+nu), Z = U_nu(x);
    // term 2k+1: backward, X = U_nu(x+mu-nu), Y = U_mu(x-nu), Z = U_nu(x-nu)
    int dir[6], px[6], sx[6], sy[6], pz[6], sz[6];
#pragma unroll
    for (int k = 0, nu = 0; nu < 4; ++nu) {
        if (nu == MU) continue;
        int xpn[4], xmn[4], xpmmn[4];
        shifted(g, x, nu, +1, xpn);
        shifted(g, x, nu, -1, xmn);
        shifted(g, xpm, nu, -1, xpmmn);
        const int s_pn = slot_of(g, xpn), s_mn = slot_of(g, xmn), s_pmmn = slot_of(g, xpmmn);
        dir[k] = nu; px[k] = q; sx[k] = s_pm; sy[k] = s_pn; pz[k] = p; sz[k] = s;
        ++k;
        dir[k] = nu; px[k] = p; sx[k] = s_pmmn; sy[k] = s_mn; pz[k] = q; sz[k] = s_mn;
        ++k;
    }
    if constexpr (sizeof(Real) == 8) {
        // FP64 storage: double-buffered FP64 matrices would not fit in registers
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            M3<Real> T;
            const M3<Real> X = load3<Real>(L, g.vh, dir[k], px[k], sx[k]);
            const M3<Real> Y = load3<Real>(L, g.vh, MU, q, sy[k]);
            if ((k & 1) == 0) mul_nd(X, Y, T);
            else mul_dd(X, Y, T);
            const M3<Real> Z = load3<Real>(L, g.vh, dir[k], pz[k], sz[k]);
            if ((k & 1) == 0) acc_nd(A, T, Z);
            else acc_nn(A, T, Z);
        }
    } else {
        // FP32 storage: rows 0 and 1 of each neighbour (load3_rows01), the third
        // rows rebuilt where a product needs them.  Forward term
        // X Y^+ Z^+ = X (Z Y)^+: Y completed, V = Z Y (rows 0, 1, then row 2
        // rebuilt), P = X V^+ (rows 0, 1, then row 2 rebuilt), A += P.  Backward
        // term X^+ Y^+ Z = (Y X)^+ Z: X completed, W = Y X (rows 0, 1, then row
        // 2 rebuilt), Z completed, A += W^+ Z.
        M3<Real> X[2], Y[2], Z[2];
        X[0] = load3_rows01(L, g.vh, dir[0], px[0], sx[0]);
        Y[0] = load3_rows01(L, g.vh, MU, q, sy[0]);
        Z[0] = load3_rows01(L, g.vh, dir[0], pz[0], sz[0]);
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const int b = k & 1;
            if (k + 1 < 6) {
                X[b ^ 1] = load3_rows01(L, g.vh, dir[k + 1], px[k + 1], sx[k + 1]);
                Y[b ^ 1] = load3_rows01(L, g.vh, MU, q, sy[k + 1]);
                Z[b ^ 1] = load3_rows01(L, g.vh, dir[k + 1], pz[k + 1], sz[k + 1]);
            }
            if (b == 0) {
                M3<Real> V, P;
                recon_row2(Y[b]);
                mul_nn_rows01(Z[b], Y[b], V);
                recon_row2(V);
                mul_nd_rows01(X[b], V, P);
                recon_row2(P);
#pragma unroll
                for (int c = 0; c < 18; ++c) A.v[c] = A.v[c] + P.v[c];
            } else {
                M3<Real> W;
                recon_row2(X[b]);
                mul_nn_rows01(Y[b], X[b], W);
                recon_row2(W);
                recon_row2(Z[b]);
                acc_dn(A, W, Z[b]);
            }
        }
    }
}

// Update of the link U_MU(x) with staple A (held in FP64), `hits`
// proposals, optional FP64 reunitarisation, one rounding to storage.
// Returns the number of accepted proposals.
template <typename Real, int MU, int HITS = 0>
__device__ __forceinline__ unsigned long long su3_update_from(typename Vec2<Real>::type *__restrict__ L,
                                                              const Geom &g, int p, int s, const int x[4],
                                                              const M3<Real> &Uf, const M3<Real> &A, double nbt,
                                                              double angle, int hits, uint64_t key, uint64_t sweep,
                                                              int reunit);

template <typename Real, int MU>
__device__ __forceinline__ unsigned long long su3_update_at(typename Vec2<Real>::type *__restrict__ L, const Geom &g,
                                                            int p, int s, const int x[4], const M3<Real> &A,
                                                            double nbt, double angle, int hits, uint64_t key,
                                                            uint64_t sweep, int reunit) {
    const M3<Real> Uf = load3_own(L, g.vh, MU, p, s);
    return su3_update_from<Real, MU>(L, g, p, s, x, Uf, A, nbt, angle, hits, key, sweep, reunit);
}

// As su3_update_at with the stored link Uf already loaded.  HITS > 0: the
// hit count is HITS (the caller passes hits == HITS) and the hits are
// unrolled; 0: the runtime count `hits`.  The arithmetic is the same.
template <typename Real, int MU, int HITS>
__device__ __forceinline__ unsigned long long su3_update_from(typename Vec2<Real>::type *__restrict__ L,
                                                              const Geom &g, int p, int s, const int x[4],
                                                              const M3<Real> &Uf, const M3<Real> &A, double nbt,
                                                              double angle, int hits, uint64_t key, uint64_t sweep,
                                                              int reunit) {
    unsigned long long nacc = 0;
    M3<double> u, a;
#pragma unroll
    for (int c = 0; c < 18; ++c) {
        u.v[c] = (double)Uf.v[c];
        a.v[c] = (double)A.v[c];
    }
    const uint32_t site = (uint32_t)lex_of(g, x);
    int h = 0;
    if constexpr (HITS > 0) {
        su3_hits_unrolled<0, HITS, MU>(u, a, key, site, sweep, angle, nbt, nacc);
        h = hits;
    }
    for (; h + 8 <= hits; h += 8) {
        su3_hit<0, MU>(u, a, key, site, h + 0, sweep, angle, nbt, nacc);
        su3_hit<1, MU>(u, a, key, site, h + 1, sweep, angle, nbt, nacc);
        su3_hit<2, MU>(u, a, key, site, h + 2, sweep, angle, nbt, nacc);
        su3_hit<3, MU>(u, a, key, site, h + 3, sweep, angle, nbt, nacc);
        su3_hit<4, MU>(u, a, key, site, h + 4, sweep, angle, nbt, nacc);
        su3_hit<5, MU>(u, a, key, site, h + 5, sweep, angle, nbt, nacc);
        su3_hit<6, MU>(u, a, key, site, h + 6, sweep, angle, nbt, nacc);
        su3_hit<7, MU>(u, a, key, site, h + 7, sweep, angle, nbt, nacc);
    }
    for (; h < hits; ++h) {
        switch (h & 7) {
        case 0: su3_hit<0, MU>(u, a, key, site, h, sweep, angle, nbt, nacc); break;
        case 1: su3_hit<1, MU>(u, a, key, site, h, sweep, angle, nbt, nacc); break;
        case 2: su3_hit<2, MU>(u, a, key, site, h, sweep, angle, nbt, nacc); break;
        case 3: su3_hit<3, MU>(u, a, key, site, h, sweep, angle, nbt, nacc); break;
        case 4: su3_hit<4, MU>(u, a, key, site, h, sweep, angle, nbt, nacc); break;
        case 5: su3_hit<5, MU>(u, a, key, site, h, sweep, angle, nbt, nacc); break;
        case 6: su3_hit<6, MU>(u, a, key, site, h, sweep, angle, nbt, nacc); break;
        default: su3_hit<7, MU>(u, a, key, site, h, sweep, angle, nbt, nacc); break;
        }
    }
    if (reunit) reunitarize3(u);
    M3<Real> Uo;
#pragma unroll
    for (int c = 0; c < 18; ++c) Uo.v[c] = (Real)u.v[c];
    store3_nt(L, g.vh, MU, p, s, Uo);
    return nacc;
}

// One 256-link staging buffer in LDS, 16 bytes per lane per access (the
// pairing of the FP32 tile layout; FP64 entries are double2 already), so a
// link is 5 (FP32) or 9 (FP64) b128/b64 LDS accesses instead of 18.
template <typename Real> struct LdsMat;
template <> struct LdsMat<float> {
    float4 q[4][256];
    float2 r[256];
    __device__ __forceinline__ void put(int l, const M3<float> &m) {
#pragma unroll
        for (int k = 0; k < 4; ++k) q[k][l] = make_float4(m.v[4 * k], m.v[4 * k + 1], m.v[4 * k + 2], m.v[4 * k + 3]);
        r[l] = make_float2(m.v[16], m.v[17]);
    }
    __device__ __forceinline__ M3<float> get(int l) const {
        M3<float> m;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const float4 t = q[k][l];
            m.v[4 * k] = t.x; m.v[4 * k + 1] = t.y; m.v[4 * k + 2] = t.z; m.v[4 * k + 3] = t.w;
        }
        const float2 t = r[l];
        m.v[16] = t.x; m.v[17] = t.y;
        return m;
    }
};
template <> struct LdsMat<double> {
    double2 e[9][256];
    __device__ __forceinline__ void put(int l, const M3<double> &m) {
#pragma unroll
        for (int k = 0; k < 9; ++k) e[k][l] = make_double2(m.v[2 * k], m.v[2 * k + 1]);
    }
    __device__ __forceinline__ M3<double> get(int l) const {
        M3<double> m;
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            const double2 t = e[k][l];
            m.v[2 * k] = t.x; m.v[2 * k + 1] = t.y;
        }
        return m;
    }
};


}  // namespace ym
