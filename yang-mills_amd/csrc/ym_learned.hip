// Learned SU(3) link proposal with an exact Metropolis-Hastings correction.
//
// Proposal (no reference counterpart; the reference sampler is plain
// single-generator Metropolis, su3_ensemble.rs:336-391).  For the link
// U = U_mu(x) with staple A (Re Tr(U A) is the link's part of the Wilson
// action, beta/3 times):
//   * W(A) = e^{-i psi/3} polar(A^dagger), psi = arg det polar(A^dagger): the
//     SU(3) element aligned with the staple (a gauge-covariant function of A);
//   * chart: U = exp(i X) W with X = sum_a eps_a lambda_a / 2 Hermitian
//     traceless, eigenvalues in (-pi, pi) (principal logarithm, su3_log);
//   * an MLP maps 12 gauge-invariant features of A (singular values, det
//     phase) to mu_a and log sigma_a (a = 1..8); eps' ~ N(mu, diag sigma^2),
//     U' = exp(i X(eps')) W;
//   * the proposal density with respect to Haar measure is
//     q(U') = N(eps(U'); mu, sigma) / J(eps(U')), J the Haar Jacobian of the
//     exponential map, prod_{j<k} sinc^2((x_j - x_k)/2);
//   * accept with probability min(1, exp[(beta/3) Re Tr((U' - U) A)]
//     q(U) / q(U')).  Because A does not depend on U, the proposal does not
//     either (an independence proposal per link), so this is an exact MH step
//     for the conditional distribution of U.  Proposals eps' whose X has an
//     eigenvalue outside (-pi, pi) are rejected, which keeps eps -> U'
//     one to one on the support of q.  Links whose chart fails (log
//     residual, phase at +-pi) are outside the support of q and never move
//     through this step; `metro_hits` plain Metropolis hits follow on every
//     link, which keeps the chain ergodic.
// The chart is taken of the FP64 Gram-Schmidt projection of the stored
// link (FP32 storage is special unitary only to about 1e-7).
// Random numbers: Philox(site, mu | KIND_LEARNED << 2 | k << 5, sweep) for
// k = 0..3 (eight Box-Muller normals) and k = 4 (accept uniform).
//
// Training: maximum likelihood of the chart coordinates eps of (staple,
// link) pairs from plain chains under N(mu(A), sigma(A)^2), with FP32
// forward/backward kernels, deterministic fixed-order gradient reduction and
// Adam.
#include <algorithm>
#include <cmath>
#include <new>
#include <vector>

#include "../../include/ym_amd.h"
#include "ym_device.hpp"
#include "ym_internal.hpp"
#include "ym_linalg.hpp"
#include "ym_su3_dev.hpp"

namespace ym {

// ------------------------------------------------------------------ MLP
constexpr int kIn = 12, kHid = 64, kOut = 16;
constexpr int oW1 = 0, oB1 = oW1 + kHid * kIn, oW2 = oB1 + kHid, oB2 = oW2 + kHid * kHid, oW3 = oB2 + kHid,
              oB3 = oW3 + kOut * kHid, kParams = oB3 + kOut;
static_assert(kParams == 6032, "parameter layout");
constexpr float kLogSigmaMin = -6.0f, kLogSigmaMax = 1.0f;

// Round to bfloat16 (nearest even) and back: the BF16 inference variant
// rounds weights and layer inputs, accumulating in FP32.
__host__ __device__ __forceinline__ float bf16r(float x) {
    uint32_t u;
    __builtin_memcpy(&u, &x, 4);
    const uint32_t lsb = (u >> 16) & 1u;
    u = (u + 0x7fffu + lsb) & 0xffff0000u;
    float y;
    __builtin_memcpy(&y, &u, 4);
    return y;
}

template <bool BF16>
__device__ __forceinline__ float qz(float x) {
    if constexpr (BF16) return bf16r(x);
    else return x;
}

// out[16] = W3 tanh(W2 tanh(W1 f + b1) + b2) + b3; P points to the 6032
// parameters (LDS or global).
template <bool BF16>
__device__ __forceinline__ void mlp_forward(const float *__restrict__ P, const float f[kIn], float out[kOut]) {
    float fin[kIn];
#pragma unroll
    for (int i = 0; i < kIn; ++i) fin[i] = qz<BF16>(f[i]);
    float h1[kHid];
#pragma unroll
    for (int j = 0; j < kHid; ++j) {
        float a = P[oB1 + j];
#pragma unroll
        for (int i = 0; i < kIn; ++i) a = __builtin_fmaf(qz<BF16>(P[oW1 + j * kIn + i]), fin[i], a);
        h1[j] = qz<BF16>(tanhf(a));
    }
#pragma unroll
    for (int o = 0; o < kOut; ++o) out[o] = P[oB3 + o];
    for (int j = 0; j < kHid; ++j) {
        float a = P[oB2 + j];
#pragma unroll
        for (int i = 0; i < kHid; ++i) a = __builtin_fmaf(qz<BF16>(P[oW2 + j * kHid + i]), h1[i], a);
        const float t = qz<BF16>(tanhf(a));
#pragma unroll
        for (int o = 0; o < kOut; ++o) out[o] = __builtin_fmaf(qz<BF16>(P[oW3 + o * kHid + j]), t, out[o]);
    }
}

// ---------------------------------------------------------- link chart
__device__ __forceinline__ la::C3 to_c3(const M3<double> &m) {
    la::C3 c;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) c.m[i][j] = {YM_RE(m, i, j), YM_IM(m, i, j)};
    return c;
}
__device__ __forceinline__ M3<double> from_c3(const la::C3 &c) {
    M3<double> m;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            YM_RE(m, i, j) = c.m[i][j].r;
            YM_IM(m, i, j) = c.m[i][j].i;
        }
    return m;
}

// Staple-aligned frame W(A) and the 12 features of A.  Returns false when A
// is singular (no frame; the learned step leaves such a link unchanged).
__device__ __forceinline__ bool staple_frame(const la::C3 &A, la::C3 &W, float feat[kIn]) {
    la::C3 Z;
    // ev: eigenvalues of A A^dagger, the squared singular values of A
    double ev[3];
    if (!la::c3_polar_unitary_fast(la::c3_dag(A), Z, ev)) return false;
    const la::cx dz = la::c3_det(Z);
    const double phz = atan2(dz.i, dz.r) / 3.0;
    const la::cx rot = {cos(phz), -sin(phz)};
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) W.m[i][j] = la::cmul(rot, Z.m[i][j]);
    double s[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) s[k] = sqrt(ev[k] > 0.0 ? ev[k] : 0.0) / 3.0;
    const la::cx da = la::c3_det(A);
    const double psi = atan2(da.i, da.r);
    feat[0] = (float)s[0];
    feat[1] = (float)s[1];
    feat[2] = (float)s[2];
    feat[3] = (float)(s[0] * s[1]);
    feat[4] = (float)(s[0] * s[2]);
    feat[5] = (float)(s[1] * s[2]);
    feat[6] = (float)(s[0] * s[1] * s[2]);
    feat[7] = (float)((s[0] + s[1] + s[2]) / 3.0);
    feat[8] = (float)cos(psi);
    feat[9] = (float)sin(psi);
    feat[10] = (float)cos(psi / 3.0);
    feat[11] = (float)sin(psi / 3.0);
    return true;
}

// Chart coordinates of U in the frame W: U W^dagger = exp(i X(eps)).
__device__ __forceinline__ bool chart_coords(const la::C3 &U, const la::C3 &W, double eps[8], double ph[3],
                                             la::cx z[3]) {
    const la::C3 Y = la::c3_mul(U, la::c3_dag(W));
    la::C3 X;
    if (!la::su3_log_fast(Y, X, ph, z)) return false;
    la::su3_coords(X, eps);
    return true;
}

__device__ __forceinline__ double re_tr_mul(const la::C3 &U, const la::C3 &A) {
    double t = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k) t += U.m[i][k].r * A.m[k][i].r - U.m[i][k].i * A.m[k][i].i;
    return t;
}

// log q up to the constant shared by every state of one link.
__device__ __forceinline__ double log_q(const double eps[8], const double ph[3], const la::cx z[3],
                                       const double mu[8], const double inv_sig[8]) {
    double acc = 0.0;
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        const double d = (eps[a] - mu[a]) * inv_sig[a];
        acc -= 0.5 * d * d;
    }
    return acc - la::haar_log_jacobian(ph, z);
}

__device__ __forceinline__ double unit_open(uint32_t lo, uint32_t hi) {
    // (0, 1]: ((raw >> 11) + 1) 2^-53
    const uint64_t v = ((uint64_t)hi << 32) | lo;
    return (double)((v >> 11) + 1ull) * 0x1p-53;
}

// Intermediate values of one MH step, for the single-link probe.
struct MhRecord {
    double mu[8], sig[8], epsn[8];
    double lq0, lq1, dre, lr, uacc;
    int in_chart, accepted;
};

// The learned independence proposal and its MH test for one link whose chart
// (eps0, ph0, z0) in the frame W of its staple A succeeded: eps' from the
// network outputs `out` and four Box-Muller Philox draws, U' = exp(i X(eps')) W,
// log ratio (beta/3) Re Tr((U' - U) A) + log q(U) - log q(U').  Returns true
// when accepted, with U' in Un.  Used by k_su3_learned and by the probe, so
// both evaluate the same arithmetic.
template <bool REC>
__device__ __forceinline__ bool learned_mh(const float out[kOut], const double eps0[8], const double ph0[3],
                                           const la::cx z0[3], const la::C3 &W, const la::C3 &A,
                                           const M3<double> &u, double beta, uint64_t key, uint32_t site, int mu_dir,
                                           uint64_t sweep, la::C3 &Un, MhRecord *rec) {
    double mu[8], sig[8], isig[8];
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        const float ls = fminf(fmaxf(out[8 + a], kLogSigmaMin), kLogSigmaMax);
        mu[a] = (double)out[a];
        // sigma is whatever FP32 exp gives; the same value draws eps' and
        // enters both densities of the MH ratio
        sig[a] = (double)expf(ls);
        isig[a] = 1.0 / sig[a];
    }
    double epsn[8];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        uint32_t o[4];
        draw(key, site, mu_dir, KIND_LEARNED, (uint32_t)k, sweep, o);
        const double u1 = unit_open(o[0], o[1]), u2 = unit53(o[2], o[3]);
        const double r = sqrt(-2.0 * log(u1));
        double sn, cs;
        sincos_turns(u2, sn, cs);
        epsn[2 * k] = mu[2 * k] + sig[2 * k] * (r * cs);
        epsn[2 * k + 1] = mu[2 * k + 1] + sig[2 * k + 1] * (r * sn);
    }
    uint32_t o[4];
    draw(key, site, mu_dir, KIND_LEARNED, 4u, sweep, o);
    const double uacc = unit53(o[0], o[1]);
    // X' = V diag(x) V^dagger; proposals outside the chart (an eigenvalue
    // beyond +-(pi - 1e-6)) are rejected: on the chart eps -> U' is one to
    // one, so the chart coordinates of U' are eps' and its phases the
    // eigenvalues x
    double ph1[3];
    la::cx z1[3];
    const la::C3 E = la::expi_hermitian_fast(la::su3_from_coords(epsn), ph1, z1);
    bool in_chart = true;
#pragma unroll
    for (int k = 0; k < 3; ++k) in_chart = in_chart && fabs(ph1[k]) < 3.14159165358979;
    bool acc = false;
    double dre = 0.0, lq0 = 0.0, lq1 = 0.0, lr = 0.0;
    if (in_chart) {
        Un = la::c3_mul(E, W);
        dre = (beta / 3.0) * (re_tr_mul(Un, A) - re_tr_mul(to_c3(u), A));
        lq0 = log_q(eps0, ph0, z0, mu, isig);
        lq1 = log_q(epsn, ph1, z1, mu, isig);
        lr = dre + lq0 - lq1;
        acc = lr >= 0.0 || uacc < exp(lr);
    }
    if constexpr (REC) {
        for (int a = 0; a < 8; ++a) {
            rec->mu[a] = mu[a];
            rec->sig[a] = sig[a];
            rec->epsn[a] = epsn[a];
        }
        rec->lq0 = lq0;
        rec->lq1 = lq1;
        rec->dre = dre;
        rec->lr = lr;
        rec->uacc = uacc;
        rec->in_chart = in_chart ? 1 : 0;
        rec->accepted = acc ? 1 : 0;
    }
    return acc;
}

// ------------------------------------------------- BF16 MFMA inference
// One wavefront evaluates the MLP for its 64 links with
// v_mfma_f32_32x32x16_bf16 (D 32x32 += A 32x16 * B 16x32, FP32
// accumulation).  Lane layouts (gfx950): A, lane l holds row l % 32, k =
// 8 (l / 32) .. +7; B, lane l holds column l % 32, the same k range; D, lane
// l register r holds row 8 (r / 4) + 4 (l / 32) + r % 4, column l % 32.
// Weights are packed once per call (k_mlp_pack) into the B-operand
// fragments of the 14 MFMAs of one 32-column block: fragment f, lane l holds
// 8 bf16 at packed[f * 64 + l] (f = Cb for layer 1, 2 + 4 Cb + ks for layer
// 2, 10 + ks for layer 3), so each wave loads its weights straight from
// global memory, 16 contiguous bytes per lane, with no block-wide staging or
// barrier.  Activations pass between layers through the wave's own two
// 64 x 64 bf16 LDS tiles (row-chunk XOR swizzle against bank conflicts),
// ordered by wave-local waits only.  Inputs, weights and hidden activations
// are rounded to bf16 (nearest even), biases stay FP32, as in the
// FP32-accumulating restatement.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
constexpr int kFrags = 14;

// tanh for the BF16 path, whose hidden activations are rounded to bfloat16
// (8 significant bits) right after: odd Taylor polynomial to x^5 for
// |x| < 1/16 (relative error < 4e-9), 1 - 2 / (2^(2 |x| log2 e) + 1) above,
// with the hardware exp2 and reciprocal (absolute error a few FP32 ulps).
__device__ __forceinline__ float tanh_bf16_path(float x) {
    const float ax = fabsf(x);
    const float x2 = x * x;
    const float small = x * (1.0f + x2 * (-0.333333333f + x2 * 0.133333333f));
    const float e = __builtin_amdgcn_exp2f(ax * 2.88539008f);
    const float big = 1.0f - 2.0f * __builtin_amdgcn_rcpf(e + 1.0f);
    return ax < 0.0625f ? small : __builtin_copysignf(big, x);
}

__device__ __forceinline__ int swz(int row, int col) {
    return row * 64 + ((((col >> 3) ^ (row & 7)) << 3) | (col & 7));
}

// One block of kFrags x 64 threads... launched as kFrags blocks of 64.
__global__ __launch_bounds__(64) void k_mlp_pack(const float *__restrict__ P, bf16x8 *__restrict__ packed) {
    const int f = blockIdx.x, l = threadIdx.x, lr = l & 31, lh = l >> 5;
    bf16x8 v;
    if (f < 2) {
        const int row = 32 * f + lr;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int k = 8 * lh + i;
            v[i] = (__bf16)(k < kIn ? P[oW1 + row * kIn + k] : 0.0f);
        }
    } else if (f < 10) {
        const int cb = (f - 2) >> 2, ks = (f - 2) & 3, row = 32 * cb + lr;
#pragma unroll
        for (int i = 0; i < 8; ++i) v[i] = (__bf16)P[oW2 + row * kHid + 16 * ks + 8 * lh + i];
    } else {
        const int ks = f - 10, row = lr;
#pragma unroll
        for (int i = 0; i < 8; ++i) v[i] = (__bf16)(row < kOut ? P[oW3 + row * kHid + 16 * ks + 8 * lh + i] : 0.0f);
    }
    packed[f * 64 + l] = v;
}

__device__ __forceinline__ bf16x8 lds_chunk(const __bf16 *__restrict__ t, int row, int chunk) {
    return *reinterpret_cast<const bf16x8 *>(t + row * 64 + ((chunk ^ (row & 7)) << 3));
}

// Must be reached by every lane of the wave.  bufA/bufB: this wave's two
// 64 x 64 bf16 tiles.  Rows of lanes without a link carry zero features.
__device__ __forceinline__ void mlp_wave_bf16(const bf16x8 *__restrict__ packed, const float *__restrict__ P,
                                              __bf16 *__restrict__ bufA, __bf16 *__restrict__ bufB,
                                              const float feat[kIn], float out[kOut]) {
    const int l = threadIdx.x & 63, lr = l & 31, lh = l >> 5;
    bf16x8 wf[kFrags];
#pragma unroll
    for (int f = 0; f < kFrags; ++f) wf[f] = packed[f * 64 + l];
    const float b1[2] = {P[oB1 + lr], P[oB1 + 32 + lr]};
    const float b2[2] = {P[oB2 + lr], P[oB2 + 32 + lr]};
    const float b3 = lr < kOut ? P[oB3 + lr] : 0.0f;
#pragma unroll
    for (int c = 0; c < 16; ++c) bufB[swz(l, c)] = (__bf16)(c < kIn ? feat[c] : 0.0f);
    wave_lds_sync();
    // layer 1: H1 = tanh(F W1^T + b1) -> bufA
#pragma unroll
    for (int R = 0; R < 2; ++R)
#pragma unroll
        for (int Cb = 0; Cb < 2; ++Cb) {
            f32x16 acc = {};
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(lds_chunk(bufB, 32 * R + lr, lh), wf[Cb], acc, 0, 0, 0);
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = 32 * R + 8 * (r >> 2) + 4 * lh + (r & 3), col = 32 * Cb + lr;
                bufA[swz(row, col)] = (__bf16)tanh_bf16_path(acc[r] + b1[Cb]);
            }
        }
    wave_lds_sync();
    // layer 2: H2 = tanh(H1 W2^T + b2) -> bufB
#pragma unroll
    for (int R = 0; R < 2; ++R)
#pragma unroll
        for (int Cb = 0; Cb < 2; ++Cb) {
            f32x16 acc = {};
#pragma unroll
            for (int ks = 0; ks < 4; ++ks)
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(lds_chunk(bufA, 32 * R + lr, 2 * ks + lh),
                                                              wf[2 + 4 * Cb + ks], acc, 0, 0, 0);
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = 32 * R + 8 * (r >> 2) + 4 * lh + (r & 3), col = 32 * Cb + lr;
                bufB[swz(row, col)] = (__bf16)tanh_bf16_path(acc[r] + b2[Cb]);
            }
        }
    wave_lds_sync();
    // layer 3: OUT = H2 W3^T + b3 (columns 16..31 of the block are padding);
    // FP32 results staged in bufA (64 x 16 floats)
    float *ob = reinterpret_cast<float *>(bufA);
#pragma unroll
    for (int R = 0; R < 2; ++R) {
        f32x16 acc = {};
#pragma unroll
        for (int ks = 0; ks < 4; ++ks)
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(lds_chunk(bufB, 32 * R + lr, 2 * ks + lh), wf[10 + ks], acc,
                                                          0, 0, 0);
        if (lr < kOut) {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = 32 * R + 8 * (r >> 2) + 4 * lh + (r & 3);
                ob[row * kOut + lr] = acc[r] + b3;
            }
        }
    }
    wave_lds_sync();
#pragma unroll
    for (int c = 0; c < kOut; ++c) out[c] = ob[l * kOut + c];
    wave_lds_sync();
}

template <bool BF16> struct LearnedLds;
template <> struct LearnedLds<false> {
    float P[kParams];
};
template <> struct LearnedLds<true> {
    __bf16 tiles[4][2][64 * 64];  // per wave of a 256-thread block
};

template <int MU, bool BF16>
__global__ __launch_bounds__(256) void k_su3_learned(float2 *__restrict__ L, Geom g, int p,
                                                     const float *__restrict__ params,
                                                     const bf16x8 *__restrict__ packed, double beta, double angle,
                                                     int metro_hits, uint64_t key, uint64_t sweep,
                                                     unsigned long long *__restrict__ counts) {
    __shared__ LearnedLds<BF16> sh;
    __shared__ unsigned long long red[4];
    if constexpr (!BF16) {
        for (int i = threadIdx.x; i < kParams; i += 256) sh.P[i] = params[i];
        __syncthreads();
    }
    const int s = tile_slot(g, xcd_tile(blockIdx.x, gridDim.x), threadIdx.x);
    unsigned long long tried = 0, accepted = 0, macc = 0;
    int x[4] = {0, 0, 0, 0};
    const bool live = s < g.vh && site_of(g, p, s, x);
    // phase 1 (per link): staple, projected link, frame, features, chart
    M3<double> u;
    la::C3 A, W;
    float feat[kIn];
#pragma unroll
    for (int i = 0; i < kIn; ++i) feat[i] = 0.0f;
    double eps0[8], ph0[3];
    la::cx z0[3];
    bool ok = false;
    uint32_t site = 0;
    if (live) {
        M3<float> Af;
        su3_staple_at<float, MU>(L, g, p, s, x, Af);
        const M3<float> Uf = load3<float>(L, g.vh, MU, p, s);
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) A.m[i][j] = {(double)YM_RE(Af, i, j), (double)YM_IM(Af, i, j)};
#pragma unroll
        for (int c = 0; c < 18; ++c) u.v[c] = (double)Uf.v[c];
        // the chart needs an exactly special-unitary link: the FP32 link is
        // first projected back with the FP64 Gram-Schmidt of the update
        reunitarize3(u);
        site = (uint32_t)lex_of(g, x);
        if (staple_frame(A, W, feat)) ok = chart_coords(to_c3(u), W, eps0, ph0, z0);
        else {
#pragma unroll
            for (int i = 0; i < kIn; ++i) feat[i] = 0.0f;
        }
    }
    // phase 2 (whole block): proposal parameters
    float out[kOut];
    if constexpr (BF16) {
        const int w = threadIdx.x >> 6;
        mlp_wave_bf16(packed, params, sh.tiles[w][0], sh.tiles[w][1], feat, out);
    } else {
        mlp_forward<false>(sh.P, feat, out);
    }
    // phase 3 (per link): proposal, MH test, Metropolis hits, store
    if (live) {
        if (ok) {
            tried = 1;
            la::C3 Un;
            if (learned_mh<false>(out, eps0, ph0, z0, W, A, u, beta, key, site, MU, sweep, Un, nullptr)) {
                accepted = 1;
                u = from_c3(Un);
            }
        }
        if (metro_hits > 0) {
            const M3<double> a = from_c3(A);
            const double nbt = -(beta / 3.0);
            for (int h = 0; h < metro_hits; ++h) {
                switch (h & 7) {
                case 0: su3_hit<0, MU>(u, a, key, site, h, sweep, angle, nbt, macc); break;
                case 1: su3_hit<1, MU>(u, a, key, site, h, sweep, angle, nbt, macc); break;
                case 2: su3_hit<2, MU>(u, a, key, site, h, sweep, angle, nbt, macc); break;
                case 3: su3_hit<3, MU>(u, a, key, site, h, sweep, angle, nbt, macc); break;
                case 4: su3_hit<4, MU>(u, a, key, site, h, sweep, angle, nbt, macc); break;
                case 5: su3_hit<5, MU>(u, a, key, site, h, sweep, angle, nbt, macc); break;
                case 6: su3_hit<6, MU>(u, a, key, site, h, sweep, angle, nbt, macc); break;
                default: su3_hit<7, MU>(u, a, key, site, h, sweep, angle, nbt, macc); break;
                }
            }
        }
        reunitarize3(u);
        M3<float> Uo;
#pragma unroll
        for (int c = 0; c < 18; ++c) Uo.v[c] = (float)u.v[c];
        store3<float>(L, g.vh, MU, p, s, Uo);
    }
    const unsigned long long t0 = block_sum_u64<256>(tried, red);
    const unsigned long long t1 = block_sum_u64<256>(accepted, red);
    const unsigned long long t2 = block_sum_u64<256>(macc, red);
    // tally slots: column j of block b at counts[j * gridDim.x + b]
    if (threadIdx.x == 0) {
        counts[blockIdx.x] += t0;
        counts[gridDim.x + blockIdx.x] += t1;
        counts[2 * gridDim.x + blockIdx.x] += t2;
    }
}

// Training pairs of every link of one (mu, parity) group: features of the
// staple and chart coordinates of the link; valid = 0 where the frame or the
// chart fails.  Row r = base + s.
template <int MU>
__global__ __launch_bounds__(256) void k_learn_pairs(const float2 *__restrict__ L, Geom g, int p, int64_t base,
                                                     float *__restrict__ feats, float *__restrict__ targets,
                                                     int *__restrict__ valid) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= g.vh) return;
    const int64_t r = base + s;
    int x[4];
    int ok = 0;
    float feat[kIn] = {0};
    double eps[8] = {0};
    if (site_of(g, p, s, x)) {
        M3<float> Af;
        su3_staple_at<float, MU>(L, g, p, s, x, Af);
        M3<double> Ad, u;
        const M3<float> Uf = load3<float>(L, g.vh, MU, p, s);
#pragma unroll
        for (int c = 0; c < 18; ++c) {
            Ad.v[c] = (double)Af.v[c];
            u.v[c] = (double)Uf.v[c];
        }
        reunitarize3(u);
        la::C3 W;
        la::cx z[3];
        double ph[3];
        if (staple_frame(to_c3(Ad), W, feat) && chart_coords(to_c3(u), W, eps, ph, z)) ok = 1;
    }
#pragma unroll
    for (int i = 0; i < kIn; ++i) feats[r * kIn + i] = feat[i];
#pragma unroll
    for (int a = 0; a < 8; ++a) targets[r * 8 + a] = (float)eps[a];
    valid[r] = ok;
}

// Probe of one link (tests): staple, frame, chart, features.
template <int MU>
__global__ void k_learn_probe(const float2 *__restrict__ L, Geom g, int p, int s, double *__restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int x[4];
    if (!site_of(g, p, s, x)) return;
    M3<float> Af;
    su3_staple_at<float, MU>(L, g, p, s, x, Af);
    M3<double> Ad, u;
    const M3<float> Uf = load3<float>(L, g.vh, MU, p, s);
    for (int c = 0; c < 18; ++c) {
        Ad.v[c] = (double)Af.v[c];
        u.v[c] = (double)Uf.v[c];
    }
    reunitarize3(u);
    la::C3 W;
    float feat[kIn] = {0};
    double eps[8] = {0}, ph[3] = {0};
    la::cx z[3];
    const int okf = staple_frame(to_c3(Ad), W, feat) ? 1 : 0;
    const int okc = okf && chart_coords(to_c3(u), W, eps, ph, z) ? 1 : 0;
    const M3<double> Wm = from_c3(W);
    for (int c = 0; c < 18; ++c) { out[c] = Ad.v[c]; out[18 + c] = Wm.v[c]; }
    for (int a = 0; a < 8; ++a) out[36 + a] = eps[a];
    for (int k = 0; k < 3; ++k) out[44 + k] = ph[k];
    for (int i = 0; i < kIn; ++i) out[47 + i] = feat[i];
    out[59] = okf;
    out[60] = okc;
    out[61] = okc ? la::haar_log_jacobian(ph, z) : 0.0;
}

// Probe of one learned MH step (tests): the link's chart, the FP32 network's
// outputs, then learned_mh as k_su3_learned evaluates it.  out: [0..7] eps,
// [8..15] mu, [16..23] sigma, [24..31] eps', [32] log q(U), [33] log q(U'),
// [34] (beta/3) Re Tr((U' - U) A), [35] log ratio, [36] accept uniform,
// [37] in chart, [38] accepted, [39] chart ok, [40..57] U'.
template <int MU>
__global__ void k_learn_mh_probe(const float2 *__restrict__ L, Geom g, int p, int s,
                                 const float *__restrict__ params, double beta, uint64_t key, uint64_t sweep,
                                 double *__restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int x[4];
    if (!site_of(g, p, s, x)) return;
    M3<float> Af;
    su3_staple_at<float, MU>(L, g, p, s, x, Af);
    const M3<float> Uf = load3<float>(L, g.vh, MU, p, s);
    la::C3 A, W;
    M3<double> u;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) A.m[i][j] = {(double)YM_RE(Af, i, j), (double)YM_IM(Af, i, j)};
    for (int c = 0; c < 18; ++c) u.v[c] = (double)Uf.v[c];
    reunitarize3(u);
    float feat[kIn] = {0};
    double eps0[8] = {0}, ph0[3] = {0};
    la::cx z0[3];
    bool ok = false;
    if (staple_frame(A, W, feat)) ok = chart_coords(to_c3(u), W, eps0, ph0, z0);
    else
        for (int i = 0; i < kIn; ++i) feat[i] = 0.0f;
    float o[kOut];
    mlp_forward<false>(params, feat, o);
    for (int a = 0; a < 8; ++a) out[a] = eps0[a];
    out[39] = ok ? 1.0 : 0.0;
    if (!ok) return;
    MhRecord rec;
    la::C3 Un;
    learned_mh<true>(o, eps0, ph0, z0, W, A, u, beta, key, (uint32_t)lex_of(g, x), MU, sweep, Un, &rec);
    for (int a = 0; a < 8; ++a) {
        out[8 + a] = rec.mu[a];
        out[16 + a] = rec.sig[a];
        out[24 + a] = rec.epsn[a];
    }
    out[32] = rec.lq0;
    out[33] = rec.lq1;
    out[34] = rec.dre;
    out[35] = rec.lr;
    out[36] = rec.uacc;
    out[37] = rec.in_chart;
    out[38] = rec.accepted;
    if (rec.in_chart) {
        const M3<double> Um = from_c3(Un);
        for (int c = 0; c < 18; ++c) out[40 + c] = Um.v[c];
    }
}

// ------------------------------------------------------------ training
// Per sample: forward, loss, output gradient and the backpropagated
// pre-activation gradients.  rows[k] selects the sample of batch slot k.
__global__ __launch_bounds__(128) void k_mlp_fwd_bwd(const float *__restrict__ P, const float *__restrict__ feats,
                                                     const float *__restrict__ targets, const int *__restrict__ valid,
                                                     const int64_t *__restrict__ rows, int B,
                                                     float *__restrict__ h1o, float *__restrict__ h2o,
                                                     float *__restrict__ dout, float *__restrict__ dz2o,
                                                     float *__restrict__ dz1o, float *__restrict__ loss) {
    const int k = blockIdx.x * 128 + threadIdx.x;
    if (k >= B) return;
    const int64_t r = rows[k];
    const float w = valid[r] ? 1.0f : 0.0f;
    float f[kIn];
#pragma unroll
    for (int i = 0; i < kIn; ++i) f[i] = feats[r * kIn + i];
    float h1[kHid];
#pragma unroll
    for (int j = 0; j < kHid; ++j) {
        float a = P[oB1 + j];
#pragma unroll
        for (int i = 0; i < kIn; ++i) a = __builtin_fmaf(P[oW1 + j * kIn + i], f[i], a);
        h1[j] = tanhf(a);
        h1o[(size_t)k * kHid + j] = h1[j];
    }
    float out[kOut];
#pragma unroll
    for (int o = 0; o < kOut; ++o) out[o] = P[oB3 + o];
    for (int j = 0; j < kHid; ++j) {
        float a = P[oB2 + j];
#pragma unroll
        for (int i = 0; i < kHid; ++i) a = __builtin_fmaf(P[oW2 + j * kHid + i], h1[i], a);
        const float t = tanhf(a);
        h2o[(size_t)k * kHid + j] = t;
#pragma unroll
        for (int o = 0; o < kOut; ++o) out[o] = __builtin_fmaf(P[oW3 + o * kHid + j], t, out[o]);
    }
    // loss = sum_a 0.5 (e_a - mu_a)^2 exp(-2 ls_a) + ls_a, ls clamped
    float l = 0.0f, d[kOut];
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        const float raw = out[8 + a];
        const float ls = fminf(fmaxf(raw, kLogSigmaMin), kLogSigmaMax);
        const float iv = expf(-2.0f * ls);
        const float e = targets[r * 8 + a] - out[a];
        l += 0.5f * e * e * iv + ls;
        d[a] = -e * iv * w;
        d[8 + a] = (raw > kLogSigmaMin && raw < kLogSigmaMax) ? (1.0f - e * e * iv) * w : 0.0f;
    }
    loss[k] = l * w;
#pragma unroll
    for (int o = 0; o < kOut; ++o) dout[(size_t)k * kOut + o] = d[o];
    // dz2_j = (sum_o W3[o][j] d_o) (1 - h2_j^2)
    float dz2[kHid];
#pragma unroll
    for (int j = 0; j < kHid; ++j) {
        float a = 0.0f;
#pragma unroll
        for (int o = 0; o < kOut; ++o) a = __builtin_fmaf(P[oW3 + o * kHid + j], d[o], a);
        const float t = h2o[(size_t)k * kHid + j];
        dz2[j] = a * (1.0f - t * t);
        dz2o[(size_t)k * kHid + j] = dz2[j];
    }
    // dz1_i = (sum_j W2[j][i] dz2_j) (1 - h1_i^2)
#pragma unroll
    for (int i = 0; i < kHid; ++i) {
        float a = 0.0f;
        for (int j = 0; j < kHid; ++j) a = __builtin_fmaf(P[oW2 + j * kHid + i], dz2[j], a);
        dz1o[(size_t)k * kHid + i] = a * (1.0f - h1[i] * h1[i]);
    }
}

// Weight-gradient partials: parameter q (thread) summed over batch chunk
// blockIdx.y of `chunk` samples, in sample order.  partial[y * kParams + q].
__global__ __launch_bounds__(256) void k_mlp_grad_partial(const float *__restrict__ feats,
                                                          const int64_t *__restrict__ rows, int B, int chunk,
                                                          const float *__restrict__ h1, const float *__restrict__ h2,
                                                          const float *__restrict__ dout,
                                                          const float *__restrict__ dz2,
                                                          const float *__restrict__ dz1,
                                                          float *__restrict__ partial) {
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= kParams) return;
    const int k0 = blockIdx.y * chunk, k1 = min(B, k0 + chunk);
    float acc = 0.0f;
    if (q < oB1) {
        const int j = q / kIn, i = q - j * kIn;
        for (int k = k0; k < k1; ++k) acc = __builtin_fmaf(dz1[(size_t)k * kHid + j], feats[rows[k] * kIn + i], acc);
    } else if (q < oW2) {
        const int j = q - oB1;
        for (int k = k0; k < k1; ++k) acc += dz1[(size_t)k * kHid + j];
    } else if (q < oB2) {
        const int j = (q - oW2) / kHid, i = (q - oW2) - j * kHid;
        for (int k = k0; k < k1; ++k) acc = __builtin_fmaf(dz2[(size_t)k * kHid + j], h1[(size_t)k * kHid + i], acc);
    } else if (q < oW3) {
        const int j = q - oB2;
        for (int k = k0; k < k1; ++k) acc += dz2[(size_t)k * kHid + j];
    } else if (q < oB3) {
        const int o = (q - oW3) / kHid, j = (q - oW3) - o * kHid;
        for (int k = k0; k < k1; ++k) acc = __builtin_fmaf(dout[(size_t)k * kOut + o], h2[(size_t)k * kHid + j], acc);
    } else {
        const int o = q - oB3;
        for (int k = k0; k < k1; ++k) acc += dout[(size_t)k * kOut + o];
    }
    partial[(size_t)blockIdx.y * kParams + q] = acc;
}

// grad = (sum of partials in chunk order) / count; then, when step > 0, one
// Adam step (Kingma & Ba, 2015) on the parameters.
__global__ __launch_bounds__(256) void k_mlp_reduce_adam(const float *__restrict__ partial, int nchunks,
                                                         const float *__restrict__ count, float *__restrict__ grad,
                                                         float *__restrict__ P, float *__restrict__ m1,
                                                         float *__restrict__ m2, int step, float lr) {
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= kParams) return;
    float acc = 0.0f;
    for (int c = 0; c < nchunks; ++c) acc += partial[(size_t)c * kParams + q];
    const float n = count[0];
    const float gq = n > 0.0f ? acc / n : 0.0f;
    grad[q] = gq;
    if (step <= 0) return;
    const float b1 = 0.9f, b2 = 0.999f, epsa = 1e-8f;
    const float a = b1 * m1[q] + (1.0f - b1) * gq;
    const float v = b2 * m2[q] + (1.0f - b2) * gq * gq;
    m1[q] = a;
    m2[q] = v;
    const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    P[q] -= lr * (a / c1) / (sqrtf(v / c2) + epsa);
}

// Sum of the batch losses and of the valid weights (one block, fixed order).
__global__ __launch_bounds__(256) void k_mlp_batch_sums(const float *__restrict__ loss, const int *__restrict__ valid,
                                                        const int64_t *__restrict__ rows, int B,
                                                        float *__restrict__ out /* [loss_sum, count] */) {
    __shared__ double red[4];
    double l = 0.0, c = 0.0;
    for (int k = threadIdx.x; k < B; k += 256) {
        l += (double)loss[k];
        c += valid[rows[k]] ? 1.0 : 0.0;
    }
    const double sl = block_sum<256>(l, red);
    const double sc = block_sum<256>(c, red);
    if (threadIdx.x == 0) {
        out[0] = (float)sl;
        out[1] = (float)sc;
    }
}

// rows[k] = (a (first + k) + b) mod n: a permutation of the dataset when
// gcd(a, n) = 1.  a < 2^31 and n < 2^32 keep the product in 64 bits.
__global__ void k_mlp_rows(int64_t *__restrict__ rows, int B, int64_t first, uint64_t a, uint64_t b, int64_t n) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= B) return;
    const uint64_t v = a * (uint64_t)(first + k) + b;
    rows[k] = (int64_t)(v % (uint64_t)n);
}

template <bool BF16>
__global__ __launch_bounds__(256) void k_mlp_infer(const float *__restrict__ params,
                                                   const bf16x8 *__restrict__ packed,
                                                   const float *__restrict__ feats, int64_t n,
                                                   float *__restrict__ out) {
    __shared__ LearnedLds<BF16> sh;
    if constexpr (!BF16) {
        for (int i = threadIdx.x; i < kParams; i += 256) sh.P[i] = params[i];
        __syncthreads();
    }
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    float f[kIn], o[kOut];
#pragma unroll
    for (int i = 0; i < kIn; ++i) f[i] = r < n ? feats[r * kIn + i] : 0.0f;
    if constexpr (BF16) {
        const int w = threadIdx.x >> 6;
        mlp_wave_bf16(packed, params, sh.tiles[w][0], sh.tiles[w][1], f, o);
    } else {
        mlp_forward<false>(sh.P, f, o);
    }
    if (r >= n) return;
#pragma unroll
    for (int k = 0; k < kOut; ++k) out[r * kOut + k] = o[k];
}

}  // namespace ym

struct ym_mlp {
    float *params = nullptr, *m1 = nullptr, *m2 = nullptr, *grad = nullptr;
    int step = 0;
    int64_t cap = 0, n = 0;
    float *feats = nullptr, *targets = nullptr;
    int *valid = nullptr;
    hipStream_t stream = nullptr;
    // BF16 B-operand fragments of the weights, repacked whenever the
    // parameters change (ym_mlp_set_params, ym_mlp_train); sweeps and
    // inference only read them
    ym::bf16x8 *packed = nullptr;
};

namespace {

int32_t hs(hipError_t e) { return e == hipSuccess ? YM_OK : YM_HIP_ERROR; }

// Xavier-uniform weights from SplitMix64, zero biases.
void init_params(std::vector<float> &p, uint64_t seed) {
    uint64_t st = seed;
    auto next = [&]() {
        st += 0x9E3779B97F4A7C15ull;
        uint64_t z = st;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)(z >> 11) * 0x1p-53;
    };
    p.assign(ym::kParams, 0.0f);
    auto fill = [&](int off, int fan_out, int fan_in, double scale) {
        const double lim = scale * std::sqrt(6.0 / (fan_in + fan_out));
        for (int i = 0; i < fan_out * fan_in; ++i) p[off + i] = (float)((2.0 * next() - 1.0) * lim);
    };
    fill(ym::oW1, ym::kHid, ym::kIn, 1.0);
    fill(ym::oW2, ym::kHid, ym::kHid, 1.0);
    fill(ym::oW3, ym::kOut, ym::kHid, 0.1);
}

struct TrainScratch {
    int B = 0, nchunks = 0, chunk = 0;
    int64_t *rows = nullptr;
    float *h1 = nullptr, *h2 = nullptr, *dout = nullptr, *dz2 = nullptr, *dz1 = nullptr, *loss = nullptr;
    float *partial = nullptr, *sums = nullptr;
    int32_t alloc(int b) {
        B = b;
        chunk = 256;
        nchunks = (B + chunk - 1) / chunk;
        int32_t st = YM_OK;
        auto A = [&](void **p, size_t bytes) {
            if (!st && hipMalloc(p, bytes) != hipSuccess) st = YM_ALLOCATION_FAILED;
        };
        A((void **)&rows, sizeof(int64_t) * B);
        A((void **)&h1, sizeof(float) * B * ym::kHid);
        A((void **)&h2, sizeof(float) * B * ym::kHid);
        A((void **)&dout, sizeof(float) * B * ym::kOut);
        A((void **)&dz2, sizeof(float) * B * ym::kHid);
        A((void **)&dz1, sizeof(float) * B * ym::kHid);
        A((void **)&loss, sizeof(float) * B);
        A((void **)&partial, sizeof(float) * nchunks * ym::kParams);
        A((void **)&sums, sizeof(float) * 2);
        return st;
    }
    ~TrainScratch() {
        void *b[] = {rows, h1, h2, dout, dz2, dz1, loss, partial, sums};
        for (void *x : b)
            if (x) (void)hipFree(x);
    }
};

// One batch: rows -> forward/backward -> gradient (and Adam step when
// step > 0).  sums = [loss sum, valid count].
void enqueue_batch(ym_mlp *m, TrainScratch &t, int B, int64_t first, uint64_t a, uint64_t b, int step, float lr) {
    hipLaunchKernelGGL(ym::k_mlp_rows, dim3((B + 255) / 256), dim3(256), 0, m->stream, t.rows, B, first, a, b, m->n);
    hipLaunchKernelGGL(ym::k_mlp_fwd_bwd, dim3((B + 127) / 128), dim3(128), 0, m->stream, m->params, m->feats,
                       m->targets, m->valid, t.rows, B, t.h1, t.h2, t.dout, t.dz2, t.dz1, t.loss);
    const int nch = (B + t.chunk - 1) / t.chunk;
    hipLaunchKernelGGL(ym::k_mlp_grad_partial, dim3((ym::kParams + 255) / 256, nch), dim3(256), 0, m->stream,
                       m->feats, t.rows, B, t.chunk, t.h1, t.h2, t.dout, t.dz2, t.dz1, t.partial);
    hipLaunchKernelGGL(ym::k_mlp_batch_sums, dim3(1), dim3(256), 0, m->stream, t.loss, m->valid, t.rows, B, t.sums);
    hipLaunchKernelGGL(ym::k_mlp_reduce_adam, dim3((ym::kParams + 255) / 256), dim3(256), 0, m->stream, t.partial,
                       nch, t.sums + 1, m->grad, m->params, m->m1, m->m2, step, lr);
}

uint64_t gcd64(uint64_t x, uint64_t y) {
    while (y) {
        const uint64_t r = x % y;
        x = y;
        y = r;
    }
    return x;
}

}  // namespace

extern "C" {

int32_t ym_mlp_create(uint64_t seed, int64_t capacity, ym_mlp **out) {
    if (!out || capacity < 0 || capacity >= ((int64_t)1 << 32)) return YM_INVALID_ARGUMENT;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return YM_NO_DEVICE;
    ym_mlp *m = new (std::nothrow) ym_mlp();
    if (!m) return YM_ALLOCATION_FAILED;
    m->cap = capacity;
    int32_t st = YM_OK;
    auto A = [&](void **p, size_t bytes) {
        if (!st && bytes && hipMalloc(p, bytes) != hipSuccess) st = YM_ALLOCATION_FAILED;
    };
    if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess) st = YM_HIP_ERROR;
    A((void **)&m->params, sizeof(float) * ym::kParams);
    A((void **)&m->m1, sizeof(float) * ym::kParams);
    A((void **)&m->m2, sizeof(float) * ym::kParams);
    A((void **)&m->grad, sizeof(float) * ym::kParams);
    A((void **)&m->feats, sizeof(float) * ym::kIn * capacity);
    A((void **)&m->targets, sizeof(float) * 8 * capacity);
    A((void **)&m->valid, sizeof(int) * capacity);
    A((void **)&m->packed, sizeof(ym::bf16x8) * ym::kFrags * 64);
    if (!st) {
        std::vector<float> p;
        init_params(p, seed);
        st = ym_mlp_set_params(m, p.data());
    }
    if (st) {
        ym_mlp_destroy(m);
        return st;
    }
    *out = m;
    return YM_OK;
}

int32_t ym_mlp_destroy(ym_mlp *m) {
    if (!m) return YM_INVALID_ARGUMENT;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    void *b[] = {m->params, m->m1, m->m2, m->grad, m->feats, m->targets, m->valid, m->packed};
    for (void *x : b)
        if (x) (void)hipFree(x);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
    return YM_OK;
}

int32_t ym_mlp_get_params(const ym_mlp *m, float *params) {
    if (!m || !params) return YM_INVALID_ARGUMENT;
    int32_t st = hs(hipMemcpyAsync(params, m->params, sizeof(float) * ym::kParams, hipMemcpyDeviceToHost, m->stream));
    if (!st) st = hs(hipStreamSynchronize(m->stream));
    return st;
}

int32_t ym_mlp_set_params(ym_mlp *m, const float *params) {
    if (!m || !params) return YM_INVALID_ARGUMENT;
    for (int i = 0; i < ym::kParams; ++i)
        if (!std::isfinite(params[i])) return YM_NON_FINITE_STATISTIC;
    int32_t st = hs(hipMemcpyAsync(m->params, params, sizeof(float) * ym::kParams, hipMemcpyHostToDevice, m->stream));
    if (!st) {
        ym::launch_fill_zero((double *)m->m1, ym::kParams / 2, m->stream);
        ym::launch_fill_zero((double *)m->m2, ym::kParams / 2, m->stream);
        hipLaunchKernelGGL(ym::k_mlp_pack, dim3(ym::kFrags), dim3(64), 0, m->stream, m->params, m->packed);
        m->step = 0;
        st = hs(hipGetLastError());
    }
    if (!st) st = hs(hipStreamSynchronize(m->stream));
    return st;
}

int32_t ym_mlp_dataset_size(const ym_mlp *m, int64_t *n, int64_t *capacity) {
    if (!m || !n) return YM_INVALID_ARGUMENT;
    *n = m->n;
    if (capacity) *capacity = m->cap;
    return YM_OK;
}

int32_t ym_mlp_clear_pairs(ym_mlp *m) {
    if (!m) return YM_INVALID_ARGUMENT;
    m->n = 0;
    return YM_OK;
}

int32_t ym_mlp_collect_pairs(ym_mlp *m, const ym_lattice *lat, int64_t *total) {
    if (!m || !lat) return YM_INVALID_ARGUMENT;
    int32_t group = 0, prec = 0;
    int64_t ext[4];
    int32_t st = ym::lattice_info(lat, &group, &prec, ext);
    if (st) return st;
    if (group != 3 || prec != 32) return YM_UNSUPPORTED;
    for (int d = 0; d < 4; ++d)
        if (ext[d] % 2) return YM_UNSUPPORTED;
    const ym::Geom g = ym::make_geom(ext);
    const int64_t need = (int64_t)8 * g.vh;
    if (m->n + need > m->cap) return YM_TOO_MANY_CONFIGURATIONS;
    const float2 *L = (const float2 *)ym::lattice_links(lat);
    hipStream_t ls = ym::lattice_stream(lat);
    st = hs(hipStreamSynchronize(ls));
    if (st) return st;
    const int nb = (g.vh + 255) / 256;
    for (int mu = 0; mu < 4; ++mu)
        for (int p = 0; p < 2; ++p) {
            const int64_t base = m->n + (int64_t)(mu * 2 + p) * g.vh;
            switch (mu) {
            case 0: hipLaunchKernelGGL(ym::k_learn_pairs<0>, dim3(nb), dim3(256), 0, m->stream, L, g, p, base, m->feats, m->targets, m->valid); break;
            case 1: hipLaunchKernelGGL(ym::k_learn_pairs<1>, dim3(nb), dim3(256), 0, m->stream, L, g, p, base, m->feats, m->targets, m->valid); break;
            case 2: hipLaunchKernelGGL(ym::k_learn_pairs<2>, dim3(nb), dim3(256), 0, m->stream, L, g, p, base, m->feats, m->targets, m->valid); break;
            default: hipLaunchKernelGGL(ym::k_learn_pairs<3>, dim3(nb), dim3(256), 0, m->stream, L, g, p, base, m->feats, m->targets, m->valid); break;
            }
        }
    st = hs(hipGetLastError());
    if (!st) st = hs(hipStreamSynchronize(m->stream));
    if (!st) m->n += need;
    if (total) *total = m->n;
    return st;
}

int32_t ym_mlp_download_pairs(const ym_mlp *m, int64_t first, int64_t count, float *feats, float *targets,
                              int32_t *valid) {
    if (!m || first < 0 || count < 0 || first + count > m->n) return YM_INVALID_ARGUMENT;
    int32_t st = YM_OK;
    if (feats) st = hs(hipMemcpy(feats, m->feats + first * ym::kIn, sizeof(float) * ym::kIn * count, hipMemcpyDeviceToHost));
    if (!st && targets) st = hs(hipMemcpy(targets, m->targets + first * 8, sizeof(float) * 8 * count, hipMemcpyDeviceToHost));
    if (!st && valid) st = hs(hipMemcpy(valid, m->valid + first, sizeof(int) * count, hipMemcpyDeviceToHost));
    return st;
}

int32_t ym_mlp_upload_pairs(ym_mlp *m, const float *feats, const float *targets, const int32_t *valid,
                            int64_t count) {
    if (!m || !feats || !targets || !valid || count < 1 || count > m->cap) return YM_INVALID_ARGUMENT;
    int32_t st = hs(hipMemcpy(m->feats, feats, sizeof(float) * ym::kIn * count, hipMemcpyHostToDevice));
    if (!st) st = hs(hipMemcpy(m->targets, targets, sizeof(float) * 8 * count, hipMemcpyHostToDevice));
    if (!st) st = hs(hipMemcpy(m->valid, valid, sizeof(int) * count, hipMemcpyHostToDevice));
    if (!st) m->n = count;
    return st;
}

int32_t ym_mlp_forward(const ym_mlp *mc, const float *feats, int64_t n, int32_t bf16, float *out) {
    ym_mlp *m = const_cast<ym_mlp *>(mc);
    if (!m || !feats || !out || n < 1) return YM_INVALID_ARGUMENT;
    float *df = nullptr, *dout = nullptr;
    if (hipMalloc((void **)&df, sizeof(float) * ym::kIn * n) != hipSuccess) return YM_ALLOCATION_FAILED;
    if (hipMalloc((void **)&dout, sizeof(float) * ym::kOut * n) != hipSuccess) {
        (void)hipFree(df);
        return YM_ALLOCATION_FAILED;
    }
    int32_t st = hs(hipMemcpyAsync(df, feats, sizeof(float) * ym::kIn * n, hipMemcpyHostToDevice, m->stream));
    const dim3 grid((unsigned)((n + 255) / 256));
    if (!st) {
        if (bf16)
            hipLaunchKernelGGL(ym::k_mlp_infer<true>, grid, dim3(256), 0, m->stream, m->params, m->packed, df, n,
                               dout);
        else
            hipLaunchKernelGGL(ym::k_mlp_infer<false>, grid, dim3(256), 0, m->stream, m->params, m->packed, df, n,
                               dout);
        st = hs(hipGetLastError());
    }
    if (!st) st = hs(hipMemcpyAsync(out, dout, sizeof(float) * ym::kOut * n, hipMemcpyDeviceToHost, m->stream));
    if (!st) st = hs(hipStreamSynchronize(m->stream));
    (void)hipFree(df);
    (void)hipFree(dout);
    return st;
}

int32_t ym_mlp_loss_grad(ym_mlp *m, int64_t first, int64_t count, float *loss, float *grad) {
    if (!m || !loss || !grad || first < 0 || count < 1 || first + count > m->n) return YM_INVALID_ARGUMENT;
    if (count > (1 << 20)) return YM_INVALID_ARGUMENT;
    TrainScratch t;
    int32_t st = t.alloc((int)count);
    // identity permutation (a = 1, b = 0) starting at `first`
    if (!st) {
        enqueue_batch(m, t, (int)count, first, 1, 0, 0, 0.0f);
        st = hs(hipGetLastError());
    }
    float sums[2] = {0, 0};
    if (!st) st = hs(hipMemcpyAsync(sums, t.sums, sizeof(sums), hipMemcpyDeviceToHost, m->stream));
    if (!st) st = hs(hipMemcpyAsync(grad, m->grad, sizeof(float) * ym::kParams, hipMemcpyDeviceToHost, m->stream));
    if (!st) st = hs(hipStreamSynchronize(m->stream));
    if (!st) *loss = sums[1] > 0 ? sums[0] / sums[1] : 0.0f;
    return st;
}

int32_t ym_mlp_train(ym_mlp *m, int32_t epochs, int32_t batch, float lr, uint64_t seed, float *epoch_loss,
                     float *device_ms) {
    if (!m || epochs < 1 || batch < 1 || batch > (1 << 20) || !(lr > 0.0f) || m->n < 1) return YM_INVALID_ARGUMENT;
    TrainScratch t;
    int32_t st = t.alloc(batch);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (!st && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) st = YM_HIP_ERROR;
    if (!st) st = hs(hipEventRecord(e0, m->stream));
    uint64_t s = seed;
    const int64_t nb = m->n / batch;
    for (int ep = 0; ep < epochs && !st; ++ep) {
        // per-epoch affine permutation of the rows
        s += 0x9E3779B97F4A7C15ull;
        const uint64_t amax = std::min<uint64_t>((uint64_t)m->n, 1ull << 31);
        uint64_t a = (s >> 7) % amax;
        while (a == 0 || gcd64(a, (uint64_t)m->n) != 1) a = (a + 1) % amax;
        const uint64_t b = (s >> 17) % (uint64_t)m->n;
        double lsum = 0.0, csum = 0.0;
        for (int64_t k = 0; k < std::max<int64_t>(nb, 1) && !st; ++k) {
            const int B = (int)std::min<int64_t>(batch, m->n);
            enqueue_batch(m, t, B, k * B, a, b, ++m->step, lr);
            if (epoch_loss) {
                float sums[2];
                st = hs(hipMemcpyAsync(sums, t.sums, sizeof(sums), hipMemcpyDeviceToHost, m->stream));
                if (!st) st = hs(hipStreamSynchronize(m->stream));
                lsum += sums[0];
                csum += sums[1];
            }
        }
        if (epoch_loss) epoch_loss[ep] = csum > 0 ? (float)(lsum / csum) : 0.0f;
    }
    if (!st) st = hs(hipEventRecord(e1, m->stream));
    // the BF16 fragments follow the parameters (sweeps only read them)
    if (!st) hipLaunchKernelGGL(ym::k_mlp_pack, dim3(ym::kFrags), dim3(64), 0, m->stream, m->params, m->packed);
    if (!st) st = hs(hipGetLastError());
    if (!st) st = hs(hipStreamSynchronize(m->stream));
    if (!st && device_ms) st = hs(hipEventElapsedTime(device_ms, e0, e1));
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    return st;
}

int32_t ym_su3_learned_sweeps(ym_lattice *lat, const ym_mlp *m, double beta, uint64_t key, uint64_t first_sweep,
                              int32_t n_sweeps, int32_t metro_hits, double angle, int32_t bf16,
                              uint64_t *learned_proposals, uint64_t *learned_accepts, uint64_t *metro_accepts,
                              float *device_ms) {
    if (!lat || !m || n_sweeps < 0 || metro_hits < 0 || metro_hits > 64) return YM_INVALID_ARGUMENT;
    if (!std::isfinite(beta) || beta < 0.0 || beta > 12.0) return YM_BETA_OUT_OF_RANGE;
    if (metro_hits > 0 && (!std::isfinite(angle) || angle < 1.0e-8 || angle > 1.0)) return YM_ANGLE_OUT_OF_RANGE;
    int32_t group = 0, prec = 0;
    int64_t ext[4];
    int32_t st = ym::lattice_info(lat, &group, &prec, ext);
    if (st) return st;
    if (group != 3 || prec != 32) return YM_UNSUPPORTED;
    for (int d = 0; d < 4; ++d)
        if (ext[d] % 2) return YM_UNSUPPORTED;
    const ym::Geom g = ym::make_geom(ext);
    float2 *L = (float2 *)ym::lattice_links(lat);
    hipStream_t ls = ym::lattice_stream(lat);
    // Counters and events of the lattice handle: the network is only read, so
    // several lattices (own streams, own host threads) may sweep with one
    // network at the same time.
    unsigned long long *cnt = ym::lattice_counters(lat);
    unsigned long long *tally = ym::lattice_tally(lat);  // >= 3 nb slots
    hipEvent_t e0, e1;
    ym::lattice_aux_events(lat, &e0, &e1);
    const int nb = (g.vh + 255) / 256;
    st = hs(hipStreamSynchronize(m->stream));
    if (!st) {
        ym::launch_fill_zero_u64(tally, 3 * (size_t)nb, ls);
        st = hs(hipEventRecord(e0, ls));
    }
    for (int k = 0; k < n_sweeps && !st; ++k)
        for (int mu = 0; mu < 4; ++mu)
            for (int p = 0; p < 2; ++p) {
                const uint64_t sw = first_sweep + (uint64_t)k;
#define YM_LEARNED_LAUNCH(MU, BF)                                                                                   \
    hipLaunchKernelGGL((ym::k_su3_learned<MU, BF>), dim3(nb), dim3(256), 0, ls, L, g, p, m->params, m->packed, \
                       beta, angle, metro_hits, key, sw, tally)
                if (bf16) {
                    switch (mu) {
                    case 0: YM_LEARNED_LAUNCH(0, true); break;
                    case 1: YM_LEARNED_LAUNCH(1, true); break;
                    case 2: YM_LEARNED_LAUNCH(2, true); break;
                    default: YM_LEARNED_LAUNCH(3, true); break;
                    }
                } else {
                    switch (mu) {
                    case 0: YM_LEARNED_LAUNCH(0, false); break;
                    case 1: YM_LEARNED_LAUNCH(1, false); break;
                    case 2: YM_LEARNED_LAUNCH(2, false); break;
                    default: YM_LEARNED_LAUNCH(3, false); break;
                    }
                }
#undef YM_LEARNED_LAUNCH
            }
    if (!st) st = hs(hipGetLastError());
    if (!st) st = hs(hipEventRecord(e1, ls));
    for (int j = 0; j < 3 && !st; ++j) ym::launch_tally_sum(tally + (size_t)j * nb, nb, cnt + j, ls);
    unsigned long long h[3] = {0, 0, 0};
    if (!st) st = hs(hipMemcpyAsync(h, cnt, sizeof(h), hipMemcpyDeviceToHost, ls));
    if (!st) st = hs(hipStreamSynchronize(ls));
    if (!st && device_ms) st = hs(hipEventElapsedTime(device_ms, e0, e1));
    if (!st) {
        if (learned_proposals) *learned_proposals += h[0];
        if (learned_accepts) *learned_accepts += h[1];
        if (metro_accepts) *metro_accepts += h[2];
    }
    return st;
}

int32_t ym_su3_link_chart(const ym_lattice *lat, const int64_t x[4], int32_t mu, double *out62) {
    if (!lat || !x || !out62 || mu < 0 || mu > 3) return YM_INVALID_ARGUMENT;
    int32_t group = 0, prec = 0;
    int64_t ext[4];
    int32_t st = ym::lattice_info(lat, &group, &prec, ext);
    if (st) return st;
    if (group != 3 || prec != 32) return YM_UNSUPPORTED;
    for (int d = 0; d < 4; ++d) {
        if (ext[d] % 2) return YM_UNSUPPORTED;
        if (x[d] < 0 || x[d] >= ext[d]) return YM_COORDINATE_OUT_OF_RANGE;
    }
    const ym::Geom g = ym::make_geom(ext);
    const int p = (int)((x[0] + x[1] + x[2] + x[3]) & 1);
    const int64_t row = (x[0] * ext[1] + x[1]) * ext[2] + x[2];
    const int s = (int)((row * g.n3p + x[3]) >> 1);
    double *d = (double *)ym::lattice_aux(lat, sizeof(double) * 62);
    if (!d) return YM_ALLOCATION_FAILED;
    hipStream_t ls = ym::lattice_stream(lat);
    ym::launch_fill_zero(d, 62, ls);
    const float2 *L = (const float2 *)ym::lattice_links(lat);
    switch (mu) {
    case 0: hipLaunchKernelGGL(ym::k_learn_probe<0>, dim3(1), dim3(64), 0, ls, L, g, p, s, d); break;
    case 1: hipLaunchKernelGGL(ym::k_learn_probe<1>, dim3(1), dim3(64), 0, ls, L, g, p, s, d); break;
    case 2: hipLaunchKernelGGL(ym::k_learn_probe<2>, dim3(1), dim3(64), 0, ls, L, g, p, s, d); break;
    default: hipLaunchKernelGGL(ym::k_learn_probe<3>, dim3(1), dim3(64), 0, ls, L, g, p, s, d); break;
    }
    st = hs(hipGetLastError());
    if (!st) st = hs(hipMemcpyAsync(out62, d, sizeof(double) * 62, hipMemcpyDeviceToHost, ls));
    if (!st) st = hs(hipStreamSynchronize(ls));
    return st;
}

int32_t ym_su3_learned_probe(const ym_lattice *lat, const ym_mlp *m, double beta, uint64_t key, uint64_t sweep,
                             const int64_t x[4], int32_t mu, double *out58) {
    if (!lat || !m || !x || !out58 || mu < 0 || mu > 3) return YM_INVALID_ARGUMENT;
    if (!std::isfinite(beta) || beta < 0.0 || beta > 12.0) return YM_BETA_OUT_OF_RANGE;
    int32_t group = 0, prec = 0;
    int64_t ext[4];
    int32_t st = ym::lattice_info(lat, &group, &prec, ext);
    if (st) return st;
    if (group != 3 || prec != 32) return YM_UNSUPPORTED;
    for (int d = 0; d < 4; ++d) {
        if (ext[d] % 2) return YM_UNSUPPORTED;
        if (x[d] < 0 || x[d] >= ext[d]) return YM_COORDINATE_OUT_OF_RANGE;
    }
    const ym::Geom g = ym::make_geom(ext);
    const int p = (int)((x[0] + x[1] + x[2] + x[3]) & 1);
    const int64_t row = (x[0] * ext[1] + x[1]) * ext[2] + x[2];
    const int s = (int)((row * g.n3p + x[3]) >> 1);
    double *d = (double *)ym::lattice_aux(lat, sizeof(double) * 58);
    if (!d) return YM_ALLOCATION_FAILED;
    st = hs(hipStreamSynchronize(m->stream));
    if (st) return st;
    hipStream_t ls = ym::lattice_stream(lat);
    ym::launch_fill_zero(d, 58, ls);
    const float2 *L = (const float2 *)ym::lattice_links(lat);
    switch (mu) {
    case 0: hipLaunchKernelGGL(ym::k_learn_mh_probe<0>, dim3(1), dim3(64), 0, ls, L, g, p, s, m->params, beta, key, sweep, d); break;
    case 1: hipLaunchKernelGGL(ym::k_learn_mh_probe<1>, dim3(1), dim3(64), 0, ls, L, g, p, s, m->params, beta, key, sweep, d); break;
    case 2: hipLaunchKernelGGL(ym::k_learn_mh_probe<2>, dim3(1), dim3(64), 0, ls, L, g, p, s, m->params, beta, key, sweep, d); break;
    default: hipLaunchKernelGGL(ym::k_learn_mh_probe<3>, dim3(1), dim3(64), 0, ls, L, g, p, s, m->params, beta, key, sweep, d); break;
    }
    st = hs(hipGetLastError());
    if (!st) st = hs(hipMemcpyAsync(out58, d, sizeof(double) * 58, hipMemcpyDeviceToHost, ls));
    if (!st) st = hs(hipStreamSynchronize(ls));
    return st;
}

}  // extern "C"
