// Device-side arithmetic for the gfx950 lattice kernels.
//
// Everything here is written with explicit fused multiply-adds and compiled
// with -ffp-contract=off, so the rounding of every operation is fixed by the
// source: a CPU replay that performs the same operations in the same order
// reproduces the GPU chain bit for bit.
//
// Contents
//  * Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) counter-based RNG.
//  * FP64 exp(-d) (Cody-Waite reduction + degree-13 Taylor), sin/cos on
//    |x| <= pi/4 (Taylor to x^19/x^20), sin/cos of 2*pi*t with an exact
//    quarter-turn reduction, log (atanh series), log1p, expm1.
//  * Checkerboard slot geometry: per (mu, parity) a plane of Vh slots; the
//    slot of a site is (row*N3p + x3) >> 1, row = (x0*N1 + x1)*N2 + x2 and
//    N3p = N3 rounded up to even (odd N3 leaves one padding slot per row).
//  * 3x3 complex (SU(3)) and quaternion (SU(2)) products in Real = float or
//    double with a fixed FMA chain.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ym {

// ------------------------------------------------------------------ RNG
// a ^ b ^ c in one gfx950 three-input bit operation (truth table 0x96).
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}

// Round 0 keeps the plain xors: with the counter words that are uniform over
// a wavefront (direction, kind, index, sweep) half of its work is scalar.
// From round 1 on every word varies per lane and each pair of xors is one
// v_bitop3_b32.
__device__ __forceinline__ void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                              uint32_t k0, uint32_t k1, uint32_t out[4]) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = r == 0 ? (uint32_t)(p1 >> 32) ^ c1 ^ k0 : xor3((uint32_t)(p1 >> 32), c1, k0);
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = r == 0 ? (uint32_t)(p0 >> 32) ^ c3 ^ k1 : xor3((uint32_t)(p0 >> 32), c3, k1);
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// Counter layout: (lexicographic site, mu | kind << 2 | index << 5, sweep lo, sweep hi).
enum : uint32_t { KIND_SU3_HIT = 0, KIND_HB_ATTEMPT = 1, KIND_HB_DIRECTION = 2, KIND_HOT_START = 3, KIND_LEARNED = 4 };

__device__ __forceinline__ void draw(uint64_t key, uint32_t site, int mu, uint32_t kind, uint32_t index,
                                     uint64_t sweep, uint32_t out[4]) {
    philox4x32_10(site, (uint32_t)mu | (kind << 2) | (index << 5), (uint32_t)sweep, (uint32_t)(sweep >> 32),
                  (uint32_t)key, (uint32_t)(key >> 32), out);
}
__device__ __forceinline__ double unit53(uint32_t lo, uint32_t hi) {
    uint64_t v = ((uint64_t)hi << 32) | lo;
    return (double)(v >> 11) * 0x1p-53;
}

// Reference angle grid (su3_ensemble.rs:199-213): the high bit selects the
// sign, the low 52 bits an odd multiple of 2^-53 of the proposal angle.
// The odd multiple (2m + 1) 2^-53, m the low 52 bits, is formed without an
// integer conversion: the double with exponent 0 and mantissa m is
// 1 + m 2^-52, subtracting 1 is exact (Sterbenz), and adding 2^-53 is exact
// because (2m + 1) 2^-53 is an integer below 2^53 times a power of two.  So
// angle * unit is the one rounding of angle * (2m + 1) 2^-53, the same
// value as (angle * (2m + 1)) * 2^-53.
__device__ __forceinline__ double angle_from_raw(uint32_t lo, uint32_t hi, double proposal_angle) {
    const uint64_t bits = ((uint64_t)((hi & 0xFFFFFu) | 0x3FF00000u) << 32) | lo;
    const double unit = (__builtin_bit_cast(double, bits) - 1.0) + 0x1p-53;
    const double magnitude = proposal_angle * unit;
    // (hi >> 31) ? magnitude : -magnitude, the negation as a flip of the
    // sign bit (one three-input bit operation instead of compare and select)
    const uint64_t flip = (uint64_t)(~hi & 0x80000000u) << 32;
    return __builtin_bit_cast(double, __builtin_bit_cast(uint64_t, magnitude) ^ flip);
}

// ------------------------------------------------------------ FP64 math
// a * b + c, one rounding, as the three-address v_fma_f64.  The compiler
// turns __builtin_fma with a loop-invariant addend (a series coefficient
// held in registers) into a register copy plus the two-address
// v_fmac_f64; this form needs no copy.  Same operation, same bits.
__device__ __forceinline__ double fma_v3(double a, double b, double c) {
    double r;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}
__device__ __forceinline__ double exp_neg(double d) {
    if (!(d <= 745.0)) return 0.0;
    double x = -d;
    double k = __builtin_rint(x * 0x1.71547652b82fep+0);
    double r = __builtin_fma(-k, 0x1.62e42fefa39efp-1, x);
    r = __builtin_fma(-k, 0x1.abc9e3b39803fp-56, r);
    double p = 0x1.6124613a86d09p-33;
    p = fma_v3(p, r, 0x1.1eed8eff8d898p-29);
    p = fma_v3(p, r, 0x1.ae64567f544e4p-26);
    p = fma_v3(p, r, 0x1.27e4fb7789f5cp-22);
    p = fma_v3(p, r, 0x1.71de3a556c734p-19);
    p = fma_v3(p, r, 0x1.a01a01a01a01ap-16);
    p = fma_v3(p, r, 0x1.a01a01a01a01ap-13);
    p = fma_v3(p, r, 0x1.6c16c16c16c17p-10);
    p = fma_v3(p, r, 0x1.1111111111111p-7);
    p = fma_v3(p, r, 0x1.5555555555555p-5);
    p = fma_v3(p, r, 0x1.5555555555555p-3);
    p = __builtin_fma(p, r, 0x1.0000000000000p-1);
    p = __builtin_fma(p, r, 1.0);
    p = __builtin_fma(p, r, 1.0);
    return __builtin_ldexp(p, (int)k);
}

// sin x and cos x - 1, |x| <= pi/4: Taylor series to x^19 / x^20 by Horner
// in y = -x^2, so every coefficient is a positive 1/n!, the same constants as
// exp_neg's (one register copy each where the compiler keeps them live).
// Rounding is odd-symmetric, so this is bit for bit the series in x^2 with
// alternating coefficients.
__device__ __forceinline__ void sincos_small(double x, double &s, double &cm1) {
    const double y = -(x * x);
    double ps = 0x1.2f49b46814157p-57;  // 1/19!
    ps = fma_v3(ps, y, 0x1.952c77030ad4ap-49);
    ps = fma_v3(ps, y, 0x1.ae7f3e733b81fp-41);
    ps = fma_v3(ps, y, 0x1.6124613a86d09p-33);
    ps = fma_v3(ps, y, 0x1.ae64567f544e4p-26);
    ps = fma_v3(ps, y, 0x1.71de3a556c734p-19);
    ps = fma_v3(ps, y, 0x1.a01a01a01a01ap-13);
    ps = fma_v3(ps, y, 0x1.1111111111111p-7);
    ps = fma_v3(ps, y, 0x1.5555555555555p-3);
    s = __builtin_fma(x * y, ps, x);
    double pc = 0x1.e542ba4020225p-62;  // 1/20!
    pc = fma_v3(pc, y, 0x1.6827863b97d97p-53);
    pc = fma_v3(pc, y, 0x1.ae7f3e733b81fp-45);
    pc = fma_v3(pc, y, 0x1.93974a8c07c9dp-37);
    pc = fma_v3(pc, y, 0x1.1eed8eff8d898p-29);
    pc = fma_v3(pc, y, 0x1.27e4fb7789f5cp-22);
    pc = fma_v3(pc, y, 0x1.a01a01a01a01ap-16);
    pc = fma_v3(pc, y, 0x1.6c16c16c16c17p-10);
    pc = fma_v3(pc, y, 0x1.5555555555555p-5);
    pc = __builtin_fma(pc, y, 0x1p-1);
    cm1 = y * pc;
}

// sin x and cos x - 1 for |x| <= 0.29 (the proposal rotations of angles up
// to 0.5): the same series in y = -x^2 stopped at x^13 / x^12, whose first
// omitted terms are below 2.2e-20 (sin, relative to x) and 7.7e-18
// (cos - 1, relative to x^2 / 2) on that range.
__device__ __forceinline__ void sincos_short(double x, double &s, double &cm1) {
    const double y = -(x * x);
    double ps = 0x1.6124613a86d09p-33;  // 1/13!
    ps = fma_v3(ps, y, 0x1.ae64567f544e4p-26);
    ps = fma_v3(ps, y, 0x1.71de3a556c734p-19);
    ps = fma_v3(ps, y, 0x1.a01a01a01a01ap-13);
    ps = fma_v3(ps, y, 0x1.1111111111111p-7);
    ps = fma_v3(ps, y, 0x1.5555555555555p-3);
    s = __builtin_fma(x * y, ps, x);
    double pc = 0x1.1eed8eff8d898p-29;  // 1/12!
    pc = fma_v3(pc, y, 0x1.27e4fb7789f5cp-22);
    pc = fma_v3(pc, y, 0x1.a01a01a01a01ap-16);
    pc = fma_v3(pc, y, 0x1.6c16c16c16c17p-10);
    pc = fma_v3(pc, y, 0x1.5555555555555p-5);
    pc = __builtin_fma(pc, y, 0x1p-1);
    cm1 = y * pc;
}

// sin(2 pi t), cos(2 pi t) for t in [0, 1).
__device__ __forceinline__ void sincos_turns(double t, double &s, double &c) {
    double q = __builtin_rint(4.0 * t);
    double tr = t - q * 0.25;
    double x = tr * 0x1.921fb54442d18p+2;
    double ss, cm1;
    sincos_small(x, ss, cm1);
    double cc = 1.0 + cm1;
    int quad = ((int)q) & 3;
    if (quad == 0) { s = ss; c = cc; }
    else if (quad == 1) { s = cc; c = -ss; }
    else if (quad == 2) { s = -ss; c = -cc; }
    else { s = -cc; c = ss; }
}

__device__ __forceinline__ double log_pos(double w) {
    int e;
    double m = __builtin_frexp(w, &e);
    if (m < 0x1.6a09e667f3bcdp-1) { m = m * 2.0; e -= 1; }
    double f = (m - 1.0) / (m + 1.0);
    double f2 = f * f;
    double q = 0x1.642c8590b2164p-5;
    q = fma_v3(q, f2, 0x1.8618618618618p-5);
    q = fma_v3(q, f2, 0x1.af286bca1af28p-5);
    q = fma_v3(q, f2, 0x1.e1e1e1e1e1e1ep-5);
    q = fma_v3(q, f2, 0x1.1111111111111p-4);
    q = fma_v3(q, f2, 0x1.3b13b13b13b14p-4);
    q = fma_v3(q, f2, 0x1.745d1745d1746p-4);
    q = fma_v3(q, f2, 0x1.c71c71c71c71cp-4);
    q = fma_v3(q, f2, 0x1.2492492492492p-3);
    q = fma_v3(q, f2, 0x1.999999999999ap-3);
    q = fma_v3(q, f2, 0x1.5555555555555p-2);
    double tf = 2.0 * f;
    double lm = __builtin_fma(tf * f2, q, tf);
    double de = (double)e;
    return __builtin_fma(de, 0x1.62e42fefa39efp-1, __builtin_fma(de, 0x1.abc9e3b39803fp-56, lm));
}

__device__ __forceinline__ double log1p_dev(double y) {
    double w = 1.0 + y;
    if (w == 1.0) return y;
    return log_pos(w) + (y - (w - 1.0)) / w;
}

__device__ __forceinline__ double expm1_nonpos(double x) {
    if (x > -0x1.62e42fefa39efp-2) {
        double p = 0x1.93974a8c07c9dp-37;
        p = fma_v3(p, x, 0x1.6124613a86d09p-33);
        p = fma_v3(p, x, 0x1.1eed8eff8d898p-29);
        p = fma_v3(p, x, 0x1.ae64567f544e4p-26);
        p = fma_v3(p, x, 0x1.27e4fb7789f5cp-22);
        p = fma_v3(p, x, 0x1.71de3a556c734p-19);
        p = fma_v3(p, x, 0x1.a01a01a01a01ap-16);
        p = fma_v3(p, x, 0x1.a01a01a01a01ap-13);
        p = fma_v3(p, x, 0x1.6c16c16c16c17p-10);
        p = fma_v3(p, x, 0x1.1111111111111p-7);
        p = fma_v3(p, x, 0x1.5555555555555p-5);
        p = fma_v3(p, x, 0x1.5555555555555p-3);
        p = __builtin_fma(p, x, 0x1p-1);
        return __builtin_fma(p, x * x, x);
    }
    return exp_neg(-x) - 1.0;
}

// ------------------------------------------------------------- geometry
// Division by a divisor fixed for a launch, as a multiply and a shift
// (Granlund & Montgomery, PLDI 1994, Thm 4.2 with N = 30): for d >= 1 and
// l = ceil(log2 d), m = floor(2^(30+l) / d) + 1 < 2^32 gives
// floor(n * m / 2^(30+l)) = floor(n / d) for every 0 <= n < 2^30.  Slot,
// row and tile indices stay below 2^30 (at most 2^30 sites per lattice).
struct FastDiv {
    uint32_t m, s, d;
};
__host__ inline FastDiv make_fastdiv(int d) {
    FastDiv f{0u, 0u, (uint32_t)d};
    if (d < 1) return f;
    uint32_t l = 0;
    while ((1ull << l) < (uint64_t)d) ++l;
    f.s = 30u + l;
    f.m = (uint32_t)(((1ull << f.s) / (uint64_t)d) + 1ull);
    return f;
}
__host__ __device__ __forceinline__ int fdiv(int n, const FastDiv &f) {
    return (int)(((uint64_t)(uint32_t)n * (uint64_t)f.m) >> f.s);
}

struct Geom {
    int n[4];
    int n3p;   // N3 rounded up to even
    int half;  // N3p / 2: slots per row and parity
    int vh;    // slots per (mu, parity) plane
    int volume;
    int tk;    // 256-slot tiles per (x0, x1) pair, 0 when N2 * half is not a multiple of 256
    int t1;    // x1 strip width of the tile visiting order, 0 = plain order
    int tq;    // 64-slot quarters per (x0, x1) pair for compact tiles, 0 when not available
    int tmode; // 0: a tile is 256 consecutive slots; 1: a tile is 2 x 2 (x0, x1) x 64 slots
    int kper;  // SU(3) sub-pass kernel: 0 wave-specialised from 256 tiles up, 1 one link per thread
    // divisors of site_of / tile_order / tile_slot (set_divisors keeps them current)
    FastDiv dv_half, dv_n2, dv_n1, dv_n0, dv_h0, dv_tk, dv_tq, dv_t1, dv_strip;
};

// Refresh the division constants after n, tk, tq or t1 changed.
__host__ inline void set_divisors(Geom &g) {
    g.dv_half = make_fastdiv(g.half);
    g.dv_n2 = make_fastdiv(g.n[2]);
    g.dv_n1 = make_fastdiv(g.n[1]);
    g.dv_n0 = make_fastdiv(g.n[0]);
    g.dv_h0 = make_fastdiv(g.n[0] >> 1);
    g.dv_tk = make_fastdiv(g.tk);
    g.dv_tq = make_fastdiv(g.tq);
    g.dv_t1 = make_fastdiv(g.t1);
    g.dv_strip = make_fastdiv(g.t1 > 0 ? g.t1 : 1);
}

// Geometry of an N0 x N1 x N2 x N3 lattice (extents already validated).
__host__ inline Geom make_geom(const int64_t n[4]) {
    Geom g{};
    for (int d = 0; d < 4; ++d) g.n[d] = (int)n[d];
    g.n3p = g.n[3] + (g.n[3] & 1);
    g.half = g.n3p / 2;
    g.vh = g.n[0] * g.n[1] * g.n[2] * g.half;
    g.volume = g.n[0] * g.n[1] * g.n[2] * g.n[3];
    const int row_slots = g.n[2] * g.half;
    g.tk = (row_slots % 256 == 0) ? row_slots / 256 : 0;
    g.t1 = (g.tk > 0 && g.n[1] % 4 == 0) ? 4 : 0;
    g.tq = (row_slots % 64 == 0 && g.n[0] % 2 == 0 && g.n[1] % 2 == 0) ? row_slots / 64 : 0;
    g.tmode = 0;
    g.kper = 0;
    set_divisors(g);
    return g;
}

// Order in which the blocks of a sub-pass visit the 256-slot tiles.  In slot
// order the x0-neighbour of a tile lies N1 * N2 * half / 256 tiles away, far
// beyond what one XCD's L2 still holds when that tile is reached.  When every
// (x0, x1) pair owns whole tiles, tiles are visited in strips of t1 values of
// x1, with x0 running inside a strip: x0- and x1-neighbours are then at most
// t1 * tk tiles apart.  A permutation of independent tiles: results do not
// depend on it.
__host__ __device__ __forceinline__ int tile_order(const Geom &g, int t) {
    if (g.t1 == 0) return t;
    int r = fdiv(t, g.dv_tk);
    const int j = t - r * g.tk;
    const int r1 = fdiv(r, g.dv_t1);
    const int x1i = r - r1 * g.t1;
    const int x1t = fdiv(r1, g.dv_n0), x0 = r1 - x1t * g.n[0];
    return (x0 * g.n[1] + x1t * g.t1 + x1i) * g.tk + j;
}

// Site of slot s in parity p; returns false for the padding slot of odd N3.
__device__ __forceinline__ bool site_of(const Geom &g, int p, int s, int x[4]) {
    int row = fdiv(s, g.dv_half);
    int j = s - row * g.half;
    int r2 = fdiv(row, g.dv_n2);
    x[2] = row - r2 * g.n[2];
    x[0] = fdiv(r2, g.dv_n1);
    x[1] = r2 - x[0] * g.n[1];
    x[3] = 2 * j + ((p + x[0] + x[1] + x[2]) & 1);
    return x[3] < g.n[3];
}
__device__ __forceinline__ int slot_of(const Geom &g, const int x[4]) {
    int row = (x[0] * g.n[1] + x[1]) * g.n[2] + x[2];
    return (row * g.n3p + x[3]) >> 1;
}
__device__ __forceinline__ int lex_of(const Geom &g, const int x[4]) {
    return ((x[0] * g.n[1] + x[1]) * g.n[2] + x[2]) * g.n[3] + x[3];
}
__device__ __forceinline__ int parity_of(const int x[4]) { return (x[0] + x[1] + x[2] + x[3]) & 1; }
__device__ __forceinline__ void shifted(const Geom &g, const int x[4], int d, int s, int y[4]) {
    y[0] = x[0]; y[1] = x[1]; y[2] = x[2]; y[3] = x[3];
    int v = x[d] + s;
    if (v < 0) v += g.n[d];
    if (v >= g.n[d]) v -= g.n[d];
    y[d] = v;
}

// Orders this wave's LDS writes before its later LDS reads (no block barrier).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Point-to-point handoff between two waves of one block through an LDS
// flag holding a count.  post: every earlier LDS access of the wave is
// complete (release fence, workgroup scope, LDS only: the wave's global
// stores stay in flight), then one lane stores the count.
// wait: spin (s_sleep between reads) until the flag reaches `target`, then
// acquire.  The partner wave of the same block is always resident and never
// waits on the waiting wave for that flag, so the spin ends; a bound of
// 2^24 reads (about a second) turns a logic error into status bit 2 instead
// of a hang.
__device__ __forceinline__ void wave_handoff_post(int *flag, int count, bool leader) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    if (leader) __hip_atomic_store(flag, count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void wave_handoff_wait(const int *flag, int target, int *status) {
    for (int it = 0; __hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) < target; ++it) {
        if (it >= (1 << 24)) {
            atomicOr(status, 2);
            break;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Logical block index with consecutive tiles kept on one XCD: the hardware
// deals workgroups round-robin over the 8 XCDs (block b -> XCD b % 8), so XCD
// k receives a contiguous range of logical tiles and its L2 sees the
// neighbouring slots of its own tiles.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
    int xcd = b & 7, idx = b >> 3;
    int q = nb >> 3, r = nb & 7;
    return xcd < r ? xcd * (q + 1) + idx : r * (q + 1) + (xcd - r) * q + idx;
}
// The same ranges, each walked backwards (block b takes the tile that
// xcd_tile gives to the last-dispatched block of its XCD, and so on).
__device__ __forceinline__ int xcd_tile_rev(int b, int nb) {
    int xcd = b & 7, idx = b >> 3;
    int q = nb >> 3, r = nb & 7;
    const int cnt = q + (xcd < r ? 1 : 0);
    const int base = xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q;
    return base + cnt - 1 - idx;
}

// Slot of thread tid (0..255) of tile t.  tmode 0: tile_order(t) * 256 +
// tid.  tmode 1 (compact tiles): tile t covers x0 in {2 X0, 2 X0 + 1},
// x1 in {2 X1, 2 X1 + 1} and one 64-slot quarter j of the (x0, x1) rows, so
// half of its x0- and x1-neighbours are inside the tile; tiles are visited
// in strips of t1 values of X1 (t1 = 0 means 1) with X0 running inside a
// strip.  Both are permutations of the Vh slots.
__host__ __device__ __forceinline__ int tile_slot(const Geom &g, int t, int tid) {
    if (g.tmode == 1) {
        const int r = fdiv(t, g.dv_tq);
        const int j = t - r * g.tq;
        const int strip = g.t1 > 0 ? g.t1 : 1;
        const int r1 = fdiv(r, g.dv_strip);
        const int x1i = r - r1 * strip;
        const int h0 = g.n[0] >> 1;
        const int r2 = fdiv(r1, g.dv_h0);
        const int X0 = r1 - r2 * h0, X1 = r2 * strip + x1i;
        const int a = tid >> 7, b = (tid >> 6) & 1;
        return ((2 * X0 + a) * g.n[1] + 2 * X1 + b) * (g.n[2] * g.half) + 64 * j + (tid & 63);
    }
    return tile_order(g, t) * 256 + tid;
}

// ------------------------------------------------------------ SU(3) types
template <typename Real> struct Vec2;
template <> struct Vec2<float> { using type = float2; };
template <> struct Vec2<double> { using type = double2; };

template <typename Real> struct M3 { Real v[18]; };  // row-major, re/im interleaved

#define YM_RE(m, i, j) (m).v[2 * (3 * (i) + (j))]
#define YM_IM(m, i, j) (m).v[2 * (3 * (i) + (j)) + 1]

__device__ __forceinline__ float fmaR(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmaR(double a, double b, double c) { return __builtin_fma(a, b, c); }

// SU(3) links, tiled (AoSoA) layout: the Vh slots of a (mu, parity) group are
// cut into tiles of 64 slots (one wavefront); a tile stores its 9 complex
// entries as 9 consecutive runs of 64 (re, im) pairs, so the entry e of slot s
// is pair ((group * ntile + (s >> 6)) * 9 + e) * 64 + (s & 63), with
// ntile = ceil(Vh / 64).  A wavefront reading one link touches one
// contiguous 9 * 64-pair block (4.5 KB in FP32) instead of 9 planes Vh pairs
// apart, which keeps address translation local.
__host__ __device__ __forceinline__ int ntile_of(int vh) { return (vh + 63) >> 6; }
__host__ __device__ __forceinline__ size_t su3_pairs(int vh) { return (size_t)8 * ntile_of(vh) * 576; }

// Addresses are a group base (mu, parity: uniform across a wavefront in every
// kernel here, so it is formed once in scalar registers) plus a 32-bit
// element index inside the group (a group holds at most 2^23 tiles of 576
// pairs, below 2^32), so no per-lane 64-bit multiply is needed.
template <typename Real>
__device__ __forceinline__ const typename Vec2<Real>::type *link3_ptr(const typename Vec2<Real>::type *L, int vh, int mu,
                                                                     int par, int s) {
    const typename Vec2<Real>::type *grp = L + (size_t)((mu * 2 + par) * ntile_of(vh)) * 576;
    const uint32_t tile = (uint32_t)(s >> 6);
    __builtin_assume(tile < (1u << 23));  // slots < 2^29 (at most 2^30 sites)
    return grp + (tile * 576u + (uint32_t)(s & 63));
}
template <typename Real>
__device__ __forceinline__ M3<Real> load3(const typename Vec2<Real>::type *__restrict__ L, int vh, int mu, int par, int s) {
    M3<Real> m;
    const typename Vec2<Real>::type *base = link3_ptr<Real>(L, vh, mu, par, s);
#pragma unroll
    for (int e = 0; e < 9; ++e) {
        typename Vec2<Real>::type t = base[e * 64];
        m.v[2 * e] = t.x;
        m.v[2 * e + 1] = t.y;
    }
    return m;
}
template <typename Real>
__device__ __forceinline__ void store3(typename Vec2<Real>::type *__restrict__ L, int vh, int mu, int par, int s, const M3<Real> &m) {
    typename Vec2<Real>::type *base = const_cast<typename Vec2<Real>::type *>(link3_ptr<Real>(L, vh, mu, par, s));
#pragma unroll
    for (int e = 0; e < 9; ++e) {
        typename Vec2<Real>::type t;
        t.x = m.v[2 * e];
        t.y = m.v[2 * e + 1];
        base[e * 64] = t;
    }
}

// FP32 storage pairs the entries inside the same 4.5 KB tile: entries
// (0, 1), (2, 3), (4, 5), (6, 7) as 4 runs of 64 float4 (16 bytes per slot)
// and entry 8 as one run of 64 float2.  A link is then 4 dwordx4 + 1 dwordx2
// loads per lane instead of 9 dwordx2, each wavefront load still one
// contiguous block (1 KB or 512 B), so a wave keeps about twice the bytes in
// flight for the same number of outstanding memory instructions.  FP64
// storage keeps the 9 runs of double2 (already 16 bytes per lane).
template <>
__device__ __forceinline__ M3<float> load3<float>(const float2 *__restrict__ L, int vh, int mu, int par, int s) {
    const float2 *grp = L + (size_t)((mu * 2 + par) * ntile_of(vh)) * 576;
    const uint32_t tile = (uint32_t)(s >> 6), lane = (uint32_t)(s & 63);
    __builtin_assume(tile < (1u << 23));  // slots < 2^29 (at most 2^30 sites)
    const float4 *q = reinterpret_cast<const float4 *>(grp) + (tile * 288u + lane);
    M3<float> m;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const float4 t = q[k * 64];
        m.v[4 * k] = t.x;
        m.v[4 * k + 1] = t.y;
        m.v[4 * k + 2] = t.z;
        m.v[4 * k + 3] = t.w;
    }
    const float2 t = grp[tile * 576u + 512u + lane];
    m.v[16] = t.x;
    m.v[17] = t.y;
    return m;
}
template <>
__device__ __forceinline__ void store3<float>(float2 *__restrict__ L, int vh, int mu, int par, int s, const M3<float> &m) {
    float2 *grp = L + (size_t)((mu * 2 + par) * ntile_of(vh)) * 576;
    const uint32_t tile = (uint32_t)(s >> 6), lane = (uint32_t)(s & 63);
    __builtin_assume(tile < (1u << 23));  // slots < 2^29 (at most 2^30 sites)
    float4 *q = reinterpret_cast<float4 *>(grp) + (tile * 288u + lane);
#pragma unroll
    for (int k = 0; k < 4; ++k) q[k * 64] = make_float4(m.v[4 * k], m.v[4 * k + 1], m.v[4 * k + 2], m.v[4 * k + 3]);
    grp[tile * 576u + 512u + lane] = make_float2(m.v[16], m.v[17]);
}

typedef float nt_f4 __attribute__((ext_vector_type(4)));
typedef float nt_f2 __attribute__((ext_vector_type(2)));
// The same FP32 accesses with the non-temporal hint, for the link a sub-pass
// updates: it is read once and written once and no staple of that sub-pass
// reads it, so it should not displace the neighbour links that are reused.
__device__ __forceinline__ M3<float> load3_nt(const float2 *__restrict__ L, int vh, int mu, int par, int s) {
    const float2 *grp = L + (size_t)((mu * 2 + par) * ntile_of(vh)) * 576;
    const uint32_t tile = (uint32_t)(s >> 6), lane = (uint32_t)(s & 63);
    __builtin_assume(tile < (1u << 23));
    const float4 *q = reinterpret_cast<const float4 *>(grp) + (tile * 288u + lane);
    M3<float> m;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const nt_f4 t = __builtin_nontemporal_load(reinterpret_cast<const nt_f4 *>(q + k * 64));
        m.v[4 * k] = t.x;
        m.v[4 * k + 1] = t.y;
        m.v[4 * k + 2] = t.z;
        m.v[4 * k + 3] = t.w;
    }
    const nt_f2 t = __builtin_nontemporal_load(reinterpret_cast<const nt_f2 *>(grp + tile * 576u + 512u + lane));
    m.v[16] = t.x;
    m.v[17] = t.y;
    return m;
}
__device__ __forceinline__ void store3_nt(float2 *__restrict__ L, int vh, int mu, int par, int s, const M3<float> &m) {
    float2 *grp = L + (size_t)((mu * 2 + par) * ntile_of(vh)) * 576;
    const uint32_t tile = (uint32_t)(s >> 6), lane = (uint32_t)(s & 63);
    __builtin_assume(tile < (1u << 23));
    float4 *q = reinterpret_cast<float4 *>(grp) + (tile * 288u + lane);
#pragma unroll
    for (int k = 0; k < 4; ++k)
        __builtin_nontemporal_store(nt_f4{m.v[4 * k], m.v[4 * k + 1], m.v[4 * k + 2], m.v[4 * k + 3]},
                                    reinterpret_cast<nt_f4 *>(q + k * 64));
    __builtin_nontemporal_store(nt_f2{m.v[16], m.v[17]}, reinterpret_cast<nt_f2 *>(grp + tile * 576u + 512u + lane));
}
__device__ __forceinline__ M3<double> load3_nt(const double2 *__restrict__ L, int vh, int mu, int par, int s) {
    return load3<double>(L, vh, mu, par, s);
}
__device__ __forceinline__ void store3_nt(double2 *__restrict__ L, int vh, int mu, int par, int s, const M3<double> &m) {
    store3<double>(L, vh, mu, par, s, m);
}

// T = X Y^dagger
template <typename Real>
__device__ __forceinline__ void mul_nd(const M3<Real> &X, const M3<Real> &Y, M3<Real> &T) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            Real re = YM_RE(X, i, 0) * YM_RE(Y, j, 0);
            re = fmaR(YM_IM(X, i, 0), YM_IM(Y, j, 0), re);
            re = fmaR(YM_RE(X, i, 1), YM_RE(Y, j, 1), re);
            re = fmaR(YM_IM(X, i, 1), YM_IM(Y, j, 1), re);
            re = fmaR(YM_RE(X, i, 2), YM_RE(Y, j, 2), re);
            re = fmaR(YM_IM(X, i, 2), YM_IM(Y, j, 2), re);
            Real im = YM_IM(X, i, 0) * YM_RE(Y, j, 0);
            im = fmaR(-YM_RE(X, i, 0), YM_IM(Y, j, 0), im);
            im = fmaR(YM_IM(X, i, 1), YM_RE(Y, j, 1), im);
            im = fmaR(-YM_RE(X, i, 1), YM_IM(Y, j, 1), im);
            im = fmaR(YM_IM(X, i, 2), YM_RE(Y, j, 2), im);
            im = fmaR(-YM_RE(X, i, 2), YM_IM(Y, j, 2), im);
            YM_RE(T, i, j) = re;
            YM_IM(T, i, j) = im;
        }
}
// T = X^dagger Y^dagger
template <typename Real>
__device__ __forceinline__ void mul_dd(const M3<Real> &X, const M3<Real> &Y, M3<Real> &T) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            Real re = YM_RE(X, 0, i) * YM_RE(Y, j, 0);
            re = fmaR(-YM_IM(X, 0, i), YM_IM(Y, j, 0), re);
            re = fmaR(YM_RE(X, 1, i), YM_RE(Y, j, 1), re);
            re = fmaR(-YM_IM(X, 1, i), YM_IM(Y, j, 1), re);
            re = fmaR(YM_RE(X, 2, i), YM_RE(Y, j, 2), re);
            re = fmaR(-YM_IM(X, 2, i), YM_IM(Y, j, 2), re);
            Real im = (-YM_RE(X, 0, i)) * YM_IM(Y, j, 0);
            im = fmaR(-YM_IM(X, 0, i), YM_RE(Y, j, 0), im);
            im = fmaR(-YM_RE(X, 1, i), YM_IM(Y, j, 1), im);
            im = fmaR(-YM_IM(X, 1, i), YM_RE(Y, j, 1), im);
            im = fmaR(-YM_RE(X, 2, i), YM_IM(Y, j, 2), im);
            im = fmaR(-YM_IM(X, 2, i), YM_RE(Y, j, 2), im);
            YM_RE(T, i, j) = re;
            YM_IM(T, i, j) = im;
        }
}
// A += T Z^dagger
template <typename Real>
__device__ __forceinline__ void acc_nd(M3<Real> &A, const M3<Real> &T, const M3<Real> &Z) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            Real re = YM_RE(A, i, j), im = YM_IM(A, i, j);
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                re = fmaR(YM_RE(T, i, k), YM_RE(Z, j, k), re);
                re = fmaR(YM_IM(T, i, k), YM_IM(Z, j, k), re);
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                im = fmaR(YM_IM(T, i, k), YM_RE(Z, j, k), im);
                im = fmaR(-YM_RE(T, i, k), YM_IM(Z, j, k), im);
            }
            YM_RE(A, i, j) = re;
            YM_IM(A, i, j) = im;
        }
}
// A += T Z
template <typename Real>
__device__ __forceinline__ void acc_nn(M3<Real> &A, const M3<Real> &T, const M3<Real> &Z) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            Real re = YM_RE(A, i, j), im = YM_IM(A, i, j);
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                re = fmaR(YM_RE(T, i, k), YM_RE(Z, k, j), re);
                re = fmaR(-YM_IM(T, i, k), YM_IM(Z, k, j), re);
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                im = fmaR(YM_RE(T, i, k), YM_IM(Z, k, j), im);
                im = fmaR(YM_IM(T, i, k), YM_RE(Z, k, j), im);
            }
            YM_RE(A, i, j) = re;
            YM_IM(A, i, j) = im;
        }
}

// ---- FP32 staple from the first two rows of each neighbour link ----
// A stored SU(3) link determines its third row: row 2 = conj(row 0 x row 1).
// The FP32 staple reads rows 0 and 1 of every neighbour (3 of the 5 loads
// of a link, 48 of 72 bytes) and rebuilds a third row only where a product
// needs it, so that a term costs about what the full product costs.

// Rows 0 and 1 of an FP32 link (entries 0..5: three float4 runs of the tile).
__device__ __forceinline__ M3<float> load3_rows01(const float2 *__restrict__ L, int vh, int mu, int par, int s) {
    const float2 *grp = L + (size_t)((mu * 2 + par) * ntile_of(vh)) * 576;
    const uint32_t tile = (uint32_t)(s >> 6), lane = (uint32_t)(s & 63);
    __builtin_assume(tile < (1u << 23));  // slots < 2^29 (at most 2^30 sites)
    const float4 *q = reinterpret_cast<const float4 *>(grp) + (tile * 288u + lane);
    M3<float> m;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float4 t = q[k * 64];
        m.v[4 * k] = t.x;
        m.v[4 * k + 1] = t.y;
        m.v[4 * k + 2] = t.z;
        m.v[4 * k + 3] = t.w;
    }
#pragma unroll
    for (int c = 12; c < 18; ++c) m.v[c] = 0.0f;
    return m;
}

// Row 2 of M from rows 0 and 1: M(2, c) = conj(a_j b_k - a_k b_j) for
// (c, j, k) cyclic, a = row 0, b = row 1; 8 operations per entry.
template <typename Real>
__device__ __forceinline__ void recon_row2(M3<Real> &M) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const int j = (c + 1) % 3, k = (c + 2) % 3;
        Real re = YM_RE(M, 0, j) * YM_RE(M, 1, k);
        re = fmaR(-YM_IM(M, 0, j), YM_IM(M, 1, k), re);
        re = fmaR(-YM_RE(M, 0, k), YM_RE(M, 1, j), re);
        re = fmaR(YM_IM(M, 0, k), YM_IM(M, 1, j), re);
        Real im = YM_RE(M, 0, j) * YM_IM(M, 1, k);
        im = fmaR(YM_IM(M, 0, j), YM_RE(M, 1, k), im);
        im = fmaR(-YM_RE(M, 0, k), YM_IM(M, 1, j), im);
        im = fmaR(-YM_IM(M, 0, k), YM_RE(M, 1, j), im);
        YM_RE(M, 2, c) = re;
        YM_IM(M, 2, c) = -im;
    }
}

// Rows 0 and 1 of T = X Y (Y complete).
template <typename Real>
__device__ __forceinline__ void mul_nn_rows01(const M3<Real> &X, const M3<Real> &Y, M3<Real> &T) {
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            Real re = YM_RE(X, i, 0) * YM_RE(Y, 0, j);
            re = fmaR(-YM_IM(X, i, 0), YM_IM(Y, 0, j), re);
            re = fmaR(YM_RE(X, i, 1), YM_RE(Y, 1, j), re);
            re = fmaR(-YM_IM(X, i, 1), YM_IM(Y, 1, j), re);
            re = fmaR(YM_RE(X, i, 2), YM_RE(Y, 2, j), re);
            re = fmaR(-YM_IM(X, i, 2), YM_IM(Y, 2, j), re);
            Real im = YM_RE(X, i, 0) * YM_IM(Y, 0, j);
            im = fmaR(YM_IM(X, i, 0), YM_RE(Y, 0, j), im);
            im = fmaR(YM_RE(X, i, 1), YM_IM(Y, 1, j), im);
            im = fmaR(YM_IM(X, i, 1), YM_RE(Y, 1, j), im);
            im = fmaR(YM_RE(X, i, 2), YM_IM(Y, 2, j), im);
            im = fmaR(YM_IM(X, i, 2), YM_RE(Y, 2, j), im);
            YM_RE(T, i, j) = re;
            YM_IM(T, i, j) = im;
        }
}

// Rows 0 and 1 of T = X Y^dagger (Y complete): mul_nd's chain for i < 2.
template <typename Real>
__device__ __forceinline__ void mul_nd_rows01(const M3<Real> &X, const M3<Real> &Y, M3<Real> &T) {
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            Real re = YM_RE(X, i, 0) * YM_RE(Y, j, 0);
            re = fmaR(YM_IM(X, i, 0), YM_IM(Y, j, 0), re);
            re = fmaR(YM_RE(X, i, 1), YM_RE(Y, j, 1), re);
            re = fmaR(YM_IM(X, i, 1), YM_IM(Y, j, 1), re);
            re = fmaR(YM_RE(X, i, 2), YM_RE(Y, j, 2), re);
            re = fmaR(YM_IM(X, i, 2), YM_IM(Y, j, 2), re);
            Real im = YM_IM(X, i, 0) * YM_RE(Y, j, 0);
            im = fmaR(-YM_RE(X, i, 0), YM_IM(Y, j, 0), im);
            im = fmaR(YM_IM(X, i, 1), YM_RE(Y, j, 1), im);
            im = fmaR(-YM_RE(X, i, 1), YM_IM(Y, j, 1), im);
            im = fmaR(YM_IM(X, i, 2), YM_RE(Y, j, 2), im);
            im = fmaR(-YM_RE(X, i, 2), YM_IM(Y, j, 2), im);
            YM_RE(T, i, j) = re;
            YM_IM(T, i, j) = im;
        }
}

// Rows 0 and 1 of T = X^dagger W (X and W complete).
template <typename Real>
__device__ __forceinline__ void mul_dn_rows01(const M3<Real> &X, const M3<Real> &W, M3<Real> &T) {
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            Real re = YM_RE(X, 0, i) * YM_RE(W, 0, j);
            re = fmaR(YM_IM(X, 0, i), YM_IM(W, 0, j), re);
            re = fmaR(YM_RE(X, 1, i), YM_RE(W, 1, j), re);
            re = fmaR(YM_IM(X, 1, i), YM_IM(W, 1, j), re);
            re = fmaR(YM_RE(X, 2, i), YM_RE(W, 2, j), re);
            re = fmaR(YM_IM(X, 2, i), YM_IM(W, 2, j), re);
            Real im = YM_RE(X, 0, i) * YM_IM(W, 0, j);
            im = fmaR(-YM_IM(X, 0, i), YM_RE(W, 0, j), im);
            im = fmaR(YM_RE(X, 1, i), YM_IM(W, 1, j), im);
            im = fmaR(-YM_IM(X, 1, i), YM_RE(W, 1, j), im);
            im = fmaR(YM_RE(X, 2, i), YM_IM(W, 2, j), im);
            im = fmaR(-YM_IM(X, 2, i), YM_RE(W, 2, j), im);
            YM_RE(T, i, j) = re;
            YM_IM(T, i, j) = im;
        }
}

// A += W^dagger Z (W and Z complete), chained into A's entries.
template <typename Real>
__device__ __forceinline__ void acc_dn(M3<Real> &A, const M3<Real> &W, const M3<Real> &Z) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            Real re = YM_RE(A, i, j), im = YM_IM(A, i, j);
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                re = fmaR(YM_RE(W, k, i), YM_RE(Z, k, j), re);
                re = fmaR(YM_IM(W, k, i), YM_IM(Z, k, j), re);
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                im = fmaR(YM_RE(W, k, i), YM_IM(Z, k, j), im);
                im = fmaR(-YM_IM(W, k, i), YM_RE(Z, k, j), im);
            }
            YM_RE(A, i, j) = re;
            YM_IM(A, i, j) = im;
        }
}

// Plain (not FMA-chained) product used by observables: C = X Y.
template <typename Real>
__device__ __forceinline__ M3<Real> mul_nn(const M3<Real> &X, const M3<Real> &Y) {
    M3<Real> C;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            Real re = YM_RE(X, i, 0) * YM_RE(Y, 0, j);
            re = fmaR(-YM_IM(X, i, 0), YM_IM(Y, 0, j), re);
            re = fmaR(YM_RE(X, i, 1), YM_RE(Y, 1, j), re);
            re = fmaR(-YM_IM(X, i, 1), YM_IM(Y, 1, j), re);
            re = fmaR(YM_RE(X, i, 2), YM_RE(Y, 2, j), re);
            re = fmaR(-YM_IM(X, i, 2), YM_IM(Y, 2, j), re);
            Real im = YM_RE(X, i, 0) * YM_IM(Y, 0, j);
            im = fmaR(YM_IM(X, i, 0), YM_RE(Y, 0, j), im);
            im = fmaR(YM_RE(X, i, 1), YM_IM(Y, 1, j), im);
            im = fmaR(YM_IM(X, i, 1), YM_RE(Y, 1, j), im);
            im = fmaR(YM_RE(X, i, 2), YM_IM(Y, 2, j), im);
            im = fmaR(YM_IM(X, i, 2), YM_RE(Y, 2, j), im);
            YM_RE(C, i, j) = re;
            YM_IM(C, i, j) = im;
        }
    return C;
}
template <typename Real>
__device__ __forceinline__ M3<Real> dagger3(const M3<Real> &X) {
    M3<Real> D;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            YM_RE(D, i, j) = YM_RE(X, j, i);
            YM_IM(D, i, j) = -YM_IM(X, j, i);
        }
    return D;
}
// Re Tr(X Y) without forming the product.
template <typename Real>
__device__ __forceinline__ double re_trace_mul(const M3<Real> &X, const M3<Real> &Y) {
    double t = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            t = __builtin_fma((double)YM_RE(X, i, k), (double)YM_RE(Y, k, i), t);
            t = __builtin_fma(-(double)YM_IM(X, i, k), (double)YM_IM(Y, k, i), t);
        }
    return t;
}
template <typename Real>
__device__ __forceinline__ M3<Real> identity3() {
    M3<Real> m;
#pragma unroll
    for (int c = 0; c < 18; ++c) m.v[c] = 0;
    m.v[0] = 1; m.v[8] = 1; m.v[16] = 1;
    return m;
}

// ------------------------------------------------ FP64 link algebra (SU(3))
// rows i,j <- [[c, i s],[i s, c]] (lambda_1 / lambda_4 / lambda_6 block)
__device__ __forceinline__ void rot_sym(M3<double> &m, int i, int j, double c, double s) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double xr = YM_RE(m, i, k), xi = YM_IM(m, i, k), yr = YM_RE(m, j, k), yi = YM_IM(m, j, k);
        YM_RE(m, i, k) = __builtin_fma(c, xr, -(s * yi));
        YM_IM(m, i, k) = __builtin_fma(c, xi, s * yr);
        YM_RE(m, j, k) = __builtin_fma(c, yr, -(s * xi));
        YM_IM(m, j, k) = __builtin_fma(c, yi, s * xr);
    }
}
// rows i,j <- [[c, s],[-s, c]] (lambda_2 / lambda_5 / lambda_7 block)
__device__ __forceinline__ void rot_asym(M3<double> &m, int i, int j, double c, double s) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double xr = YM_RE(m, i, k), xi = YM_IM(m, i, k), yr = YM_RE(m, j, k), yi = YM_IM(m, j, k);
        YM_RE(m, i, k) = __builtin_fma(c, xr, s * yr);
        YM_IM(m, i, k) = __builtin_fma(c, xi, s * yi);
        YM_RE(m, j, k) = __builtin_fma(c, yr, -(s * xr));
        YM_IM(m, j, k) = __builtin_fma(c, yi, -(s * xi));
    }
}
// row i <- (c + i s) row i
__device__ __forceinline__ void rot_phase(M3<double> &m, int i, double c, double s) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double xr = YM_RE(m, i, k), xi = YM_IM(m, i, k);
        YM_RE(m, i, k) = __builtin_fma(c, xr, -(s * xi));
        YM_IM(m, i, k) = __builtin_fma(c, xi, s * xr);
    }
}

// Gram-Schmidt on rows 0 and 1, row 2 = conj(row0 x row1).
__device__ __forceinline__ void reunitarize3(M3<double> &u) {
    double n = YM_RE(u, 0, 0) * YM_RE(u, 0, 0);
    n = __builtin_fma(YM_IM(u, 0, 0), YM_IM(u, 0, 0), n);
    n = __builtin_fma(YM_RE(u, 0, 1), YM_RE(u, 0, 1), n);
    n = __builtin_fma(YM_IM(u, 0, 1), YM_IM(u, 0, 1), n);
    n = __builtin_fma(YM_RE(u, 0, 2), YM_RE(u, 0, 2), n);
    n = __builtin_fma(YM_IM(u, 0, 2), YM_IM(u, 0, 2), n);
    double inv = 1.0 / __builtin_sqrt(n);
#pragma unroll
    for (int k = 0; k < 3; ++k) { YM_RE(u, 0, k) *= inv; YM_IM(u, 0, k) *= inv; }
    double pr = YM_RE(u, 0, 0) * YM_RE(u, 1, 0);
    pr = __builtin_fma(YM_IM(u, 0, 0), YM_IM(u, 1, 0), pr);
    pr = __builtin_fma(YM_RE(u, 0, 1), YM_RE(u, 1, 1), pr);
    pr = __builtin_fma(YM_IM(u, 0, 1), YM_IM(u, 1, 1), pr);
    pr = __builtin_fma(YM_RE(u, 0, 2), YM_RE(u, 1, 2), pr);
    pr = __builtin_fma(YM_IM(u, 0, 2), YM_IM(u, 1, 2), pr);
    double pi = YM_RE(u, 0, 0) * YM_IM(u, 1, 0);
    pi = __builtin_fma(-YM_IM(u, 0, 0), YM_RE(u, 1, 0), pi);
    pi = __builtin_fma(YM_RE(u, 0, 1), YM_IM(u, 1, 1), pi);
    pi = __builtin_fma(-YM_IM(u, 0, 1), YM_RE(u, 1, 1), pi);
    pi = __builtin_fma(YM_RE(u, 0, 2), YM_IM(u, 1, 2), pi);
    pi = __builtin_fma(-YM_IM(u, 0, 2), YM_RE(u, 1, 2), pi);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double ar = YM_RE(u, 0, k), ai = YM_IM(u, 0, k);
        YM_RE(u, 1, k) = __builtin_fma(-pr, ar, __builtin_fma(pi, ai, YM_RE(u, 1, k)));
        YM_IM(u, 1, k) = __builtin_fma(-pr, ai, __builtin_fma(-pi, ar, YM_IM(u, 1, k)));
    }
    n = YM_RE(u, 1, 0) * YM_RE(u, 1, 0);
    n = __builtin_fma(YM_IM(u, 1, 0), YM_IM(u, 1, 0), n);
    n = __builtin_fma(YM_RE(u, 1, 1), YM_RE(u, 1, 1), n);
    n = __builtin_fma(YM_IM(u, 1, 1), YM_IM(u, 1, 1), n);
    n = __builtin_fma(YM_RE(u, 1, 2), YM_RE(u, 1, 2), n);
    n = __builtin_fma(YM_IM(u, 1, 2), YM_IM(u, 1, 2), n);
    inv = 1.0 / __builtin_sqrt(n);
#pragma unroll
    for (int k = 0; k < 3; ++k) { YM_RE(u, 1, k) *= inv; YM_IM(u, 1, k) *= inv; }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int p = (k + 1) % 3, q = (k + 2) % 3;
        double ar = YM_RE(u, 0, p), ai = YM_IM(u, 0, p), br = YM_RE(u, 1, q), bi = YM_IM(u, 1, q);
        double cr = YM_RE(u, 0, q), ci = YM_IM(u, 0, q), dr = YM_RE(u, 1, p), di = YM_IM(u, 1, p);
        double re = ar * br;
        re = __builtin_fma(-ai, bi, re);
        re = __builtin_fma(-cr, dr, re);
        re = __builtin_fma(ci, di, re);
        double im = ar * bi;
        im = __builtin_fma(ai, br, im);
        im = __builtin_fma(-cr, di, im);
        im = __builtin_fma(-ci, dr, im);
        YM_RE(u, 2, k) = re;
        YM_IM(u, 2, k) = -im;
    }
}

// ------------------------------------------------------------ SU(2) types
template <typename Real> struct Vec4;
template <> struct Vec4<float> { using type = float4; };
template <> struct Vec4<double> { using type = double4; };

template <typename Real> struct Q4 { Real c[4]; };

template <typename Real>
__device__ __forceinline__ Q4<Real> load2(const typename Vec4<Real>::type *__restrict__ L, int vh, int mu, int par, int s) {
    typename Vec4<Real>::type t = L[(size_t)(mu * 2 + par) * vh + s];
    Q4<Real> q;
    q.c[0] = t.x; q.c[1] = t.y; q.c[2] = t.z; q.c[3] = t.w;
    return q;
}
template <typename Real>
__device__ __forceinline__ void store2(typename Vec4<Real>::type *__restrict__ L, int vh, int mu, int par, int s, const Q4<Real> &q) {
    typename Vec4<Real>::type t;
    t.x = q.c[0]; t.y = q.c[1]; t.z = q.c[2]; t.w = q.c[3];
    L[(size_t)(mu * 2 + par) * vh + s] = t;
}
// Quaternion product in the reference convention (lib.rs:212-224) as a fixed FMA chain.
template <typename Real>
__device__ __forceinline__ Q4<Real> qmul(const Q4<Real> &a, const Q4<Real> &b) {
    Q4<Real> r;
    Real r0 = a.c[0] * b.c[0];
    r0 = fmaR(-a.c[1], b.c[1], r0); r0 = fmaR(-a.c[2], b.c[2], r0); r0 = fmaR(-a.c[3], b.c[3], r0);
    Real r1 = a.c[0] * b.c[1];
    r1 = fmaR(a.c[1], b.c[0], r1); r1 = fmaR(a.c[2], b.c[3], r1); r1 = fmaR(-a.c[3], b.c[2], r1);
    Real r2 = a.c[0] * b.c[2];
    r2 = fmaR(-a.c[1], b.c[3], r2); r2 = fmaR(a.c[2], b.c[0], r2); r2 = fmaR(a.c[3], b.c[1], r2);
    Real r3 = a.c[0] * b.c[3];
    r3 = fmaR(a.c[1], b.c[2], r3); r3 = fmaR(-a.c[2], b.c[1], r3); r3 = fmaR(a.c[3], b.c[0], r3);
    r.c[0] = r0; r.c[1] = r1; r.c[2] = r2; r.c[3] = r3;
    return r;
}
template <typename Real>
__device__ __forceinline__ Q4<Real> qdag(const Q4<Real> &a) {
    Q4<Real> r;
    r.c[0] = a.c[0]; r.c[1] = -a.c[1]; r.c[2] = -a.c[2]; r.c[3] = -a.c[3];
    return r;
}
__device__ __forceinline__ double qnorm2(const Q4<double> &a) {
    double n = a.c[0] * a.c[0];
    n = __builtin_fma(a.c[1], a.c[1], n);
    n = __builtin_fma(a.c[2], a.c[2], n);
    n = __builtin_fma(a.c[3], a.c[3], n);
    return n;
}
__device__ __forceinline__ void qnormalize(Q4<double> &a) {
    double inv = 1.0 / __builtin_sqrt(qnorm2(a));
#pragma unroll
    for (int c = 0; c < 4; ++c) a.c[c] *= inv;
}

// ----------------------------------------------------- block reductions
// 64-wide wavefront reduction of a double, then across the block's waves
// through LDS.  Every thread of the block must call it.
template <int BLOCK>
__device__ __forceinline__ double block_sum(double v, double *lds) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) lds[wave] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int w = 0; w < BLOCK / 64; ++w) t += lds[w];
    }
    __syncthreads();
    return t;  // valid in thread 0
}
template <int BLOCK>
__device__ __forceinline__ double block_max(double v, double *lds) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) lds[wave] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int w = 0; w < BLOCK / 64; ++w) t = fmax(t, lds[w]);
    }
    __syncthreads();
    return t;
}
template <int BLOCK>
__device__ __forceinline__ unsigned long long block_sum_u64(unsigned long long v, unsigned long long *lds) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) lds[wave] = v;
    __syncthreads();
    unsigned long long t = 0;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int w = 0; w < BLOCK / 64; ++w) t += lds[w];
    }
    __syncthreads();
    return t;
}

// The link a sub-pass updates.  FP32: rows 0 and 1 (non-temporal, as
// load3_nt) and the third row rebuilt as conj(row 0 x row 1) in FP32, so a
// sweep never reads a stored third row.  FP64: the whole link.
__device__ __forceinline__ M3<float> load3_own(const float2 *__restrict__ L, int vh, int mu, int par, int s) {
    const float2 *grp = L + (size_t)((mu * 2 + par) * ntile_of(vh)) * 576;
    const uint32_t tile = (uint32_t)(s >> 6), lane = (uint32_t)(s & 63);
    __builtin_assume(tile < (1u << 23));
    const float4 *q = reinterpret_cast<const float4 *>(grp) + (tile * 288u + lane);
    M3<float> m;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const nt_f4 t = __builtin_nontemporal_load(reinterpret_cast<const nt_f4 *>(q + k * 64));
        m.v[4 * k] = t.x;
        m.v[4 * k + 1] = t.y;
        m.v[4 * k + 2] = t.z;
        m.v[4 * k + 3] = t.w;
    }
    recon_row2(m);
    return m;
}
__device__ __forceinline__ M3<double> load3_own(const double2 *__restrict__ L, int vh, int mu, int par, int s) {
    return load3_nt(L, vh, mu, par, s);
}

}  // namespace ym
