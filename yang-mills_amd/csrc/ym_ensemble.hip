// Batched ensembles of independent SU(3) chains with uneven geometries.
//
// All chains live in one device buffer, chain c at link_off[c] in its own
// tiled layout (the single-lattice layout), so every device function of the
// single-lattice kernels applies unchanged to (links + link_off[c], geometry
// of c).  One batched launch covers every chain: each kernel has a per-chain
// block count and a prefix table begin[0..nchains]; block b belongs to the
// chain c with begin[c] <= b < begin[c+1] (binary search), so a block never
// mixes chains.  A sweep is 8 batched launches (mu = 0..3, parity 0/1) for
// all chains; the measurement (plaquette, A1++ timeslice sums on the raw
// links and on 1..L stout levels, FP64 reductions in a fixed order) is a
// fixed sequence of batched launches.  The sweep index and the output slot
// are read from a device counter that the cycle's last kernel advances, so a
// cycle of `sweeps_per_cycle` sweeps + one measurement is captured once as a
// hipGraph and replayed.
//
// Reference counterpart: generate_independent_configurations
// (glueball_scalar_production.rs:257-297) runs chains one after another on
// the CPU, each with key derive_su3_chain_seed(master, chain_id), cold
// start, metropolis_sweep (su3_ensemble.rs:342-391), then average_plaquette,
// scalar_a1pp_operator_basis (glueball_operators.rs:143) on raw and
// spatial_stout_smear (glueball_smearing.rs:257) links.
#include <algorithm>
#include <cmath>
#include <new>
#include <vector>

#include "../../include/ym_amd.h"
#include "ym_device.hpp"
#include "ym_internal.hpp"
#include "ym_obs_dev.hpp"
#include "ym_su3_dev.hpp"

namespace ym {

struct ChainDesc {
    Geom g;
    uint64_t key;
    uint64_t link_off;  // in link-pair (Vec2) elements
    int nsp;            // spatial sites N0 N1 N2
    int t_off;          // sum of N3 over the chains before this one
    uint64_t sv_off;    // sum of the volumes of the chains before this one
};

// Chain of block b: largest c with begin[c] <= b.
__device__ __forceinline__ int chain_of(const int *__restrict__ begin, int nch, int b) {
    int lo = 0, hi = nch - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (begin[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <typename Real, int MU>
__global__ __launch_bounds__(256) void k_ens_metropolis(typename Vec2<Real>::type *__restrict__ links,
                                                        const ChainDesc *__restrict__ cd,
                                                        const int *__restrict__ begin, int nch, int p, double nbt,
                                                        double angle, int hits, const uint64_t *__restrict__ ctr,
                                                        int sweep_k, int reunit,
                                                        unsigned long long *__restrict__ acc) {
    __shared__ unsigned long long red[4];
    // parity-1 sub-passes walk the tiles backwards: each sub-pass reads every
    // chain's whole field, and starting where the previous one ended finds
    // its last lines still cached
    const int b = (p & 1) ? xcd_tile_rev(blockIdx.x, gridDim.x) : xcd_tile(blockIdx.x, gridDim.x);
    const int c = chain_of(begin, nch, b);
    const Geom g = cd[c].g;
    typename Vec2<Real>::type *L = links + cd[c].link_off;
    const int s = tile_slot(g, b - begin[c], threadIdx.x);
    unsigned long long nacc = 0;
    int x[4];
    if (s < g.vh && site_of(g, p, s, x)) {
        M3<Real> A;
        su3_staple_at<Real, MU>(L, g, p, s, x, A);
        nacc = su3_update_at<Real, MU>(L, g, p, s, x, A, nbt, angle, hits, cd[c].key, ctr[0] + (uint64_t)sweep_k,
                                       reunit);
    }
    const unsigned long long tot = block_sum_u64<256>(nacc, red);
    if (threadIdx.x == 0 && tot) atomicAdd(acc + c, tot);
}

// Wave-specialised form of k_ens_metropolis for batches of at least 512
// tiles (the single-lattice k_su3_metropolis_ws over the concatenated tile
// list of all chains): one 512-thread block per CU walks a sequence of
// tiles; waves 0-3 stage the staples and links of tile i+1 in LDS while
// waves 4-7 run the hits of tile i.  The XCDs take contiguous ranges of the
// tile list, parity-1 sub-passes walk them backwards.  Per link the
// arithmetic is k_ens_metropolis's; accepted proposals are added per chain
// by each consumer wave.
template <typename Real, int MU, int HITS>
__global__ __launch_bounds__(512) void k_ens_metropolis_ws(typename Vec2<Real>::type *__restrict__ links,
                                                           const ChainDesc *__restrict__ cd,
                                                           const int *__restrict__ begin,
                                                           const int *__restrict__ tile_chain, int ntiles,
                                                           int p, double nbt, double angle, int hits,
                                                           const uint64_t *__restrict__ ctr, int sweep_k, int reunit,
                                                           unsigned long long *__restrict__ acc) {
    constexpr int NB = 2;
    __shared__ LdsMat<Real> abuf[NB], ubuf[NB];
    const bool producer = threadIdx.x < 256;
    const int lane = threadIdx.x & 255;
    const int xcd = blockIdx.x & 7, j = blockIdx.x >> 3;
    const int nbx = ((int)gridDim.x - 1 - xcd) / 8 + 1;
    const int t0 = (int)(((int64_t)ntiles * xcd) / 8), t1 = (int)(((int64_t)ntiles * (xcd + 1)) / 8);
    const int nsteps = (t1 - t0 - j + nbx - 1) / nbx;
    const bool rev = (p & 1) != 0;
    const uint64_t sweep = ctr[0] + (uint64_t)sweep_k;
    auto tile_at = [&](int i) { return rev ? t1 - 1 - (j + i * nbx) : t0 + j + i * nbx; };
    auto produce = [&](int i) {
        const int T = tile_at(i);
        const int c = tile_chain[T];
        const Geom g = cd[c].g;
        const typename Vec2<Real>::type *L = links + cd[c].link_off;
        const int s = tile_slot(g, T - begin[c], lane);
        int x[4];
        if (s < g.vh && site_of(g, p, s, x)) {
            M3<Real> A;
            su3_staple_at<Real, MU>(L, g, p, s, x, A);
            const M3<Real> U = load3_own(L, g.vh, MU, p, s);
            abuf[i % NB].put(lane, A);
            ubuf[i % NB].put(lane, U);
        }
    };
    auto consume = [&](int i) {
        const int T = tile_at(i);
        const int c = tile_chain[T];
        const Geom g = cd[c].g;
        typename Vec2<Real>::type *L = links + cd[c].link_off;
        const int s = tile_slot(g, T - begin[c], lane);
        unsigned long long nacc = 0;
        int x[4];
        if (s < g.vh && site_of(g, p, s, x)) {
            const M3<Real> A = abuf[i % NB].get(lane), U = ubuf[i % NB].get(lane);
            nacc = su3_update_from<Real, MU, HITS>(L, g, p, s, x, U, A, nbt, angle, hits, cd[c].key, sweep, reunit);
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) nacc += __shfl_down(nacc, off, 64);
        if ((lane & 63) == 0 && nacc) atomicAdd(acc + c, nacc);
    };
    // Phase t of the block: the producers stage tile t while the consumers
    // update tile t - 1 (n + 1 phases for n tiles); produce and consume have
    // one call site each, so the kernel carries one copy of their code.
    const int nphase = nsteps == 0 ? 0 : nsteps + 1;
    for (int t = 0; t < nphase; ++t) {
        if (producer) {
            if (t < nsteps) produce(t);
        } else if (t > 0) {
            consume(t - 1);
        }
        __syncthreads();
    }
}

// Plaquette partial per block: begin[c+1]-begin[c] = 2 * tiles(c), parity
// major.
template <typename Real>
__global__ __launch_bounds__(256) void k_ens_plaquette(const typename Vec2<Real>::type *__restrict__ links,
                                                       const ChainDesc *__restrict__ cd,
                                                       const int *__restrict__ begin, int nch,
                                                       double *__restrict__ part) {
    __shared__ double red[4];
    const int b = xcd_tile(blockIdx.x, gridDim.x);  // contiguous tile ranges per XCD
    const int c = chain_of(begin, nch, b);
    const Geom g = cd[c].g;
    const int nt = (g.vh + 255) / 256;
    const int lb = b - begin[c];
    const int p = lb / nt, s = (lb - p * nt) * 256 + threadIdx.x;
    double acc = 0.0;
    if (s < g.vh) acc = plaquette_site<G3<Real>>(links + cd[c].link_off, g, p, s);
    const double t = block_sum<256>(acc, red);
    if (threadIdx.x == 0) part[b] = t;
}

// FP32 links: both parities of one 256-slot tile per 512-thread block
// (tiles[c+1]-tiles[c] = ceil(vh/256), the Metropolis table), as
// k_plaquette_pp: a tile's links are fetched once for both parities.  Each
// half writes the partial of its parity at the index k_ens_plaquette uses,
// reduced with the same tree, so results are bit-identical.
template <typename Real>
__global__ __launch_bounds__(512) void k_ens_plaquette_pp(const typename Vec2<Real>::type *__restrict__ links,
                                                          const ChainDesc *__restrict__ cd,
                                                          const int *__restrict__ tiles,
                                                          const int *__restrict__ tile_chain,
                                                          const int *__restrict__ begin,
                                                          double *__restrict__ part) {
    __shared__ double red[8];
    const int b = xcd_tile(blockIdx.x, gridDim.x);  // contiguous tile ranges per XCD
    const int c = tile_chain[b];
    const Geom g = cd[c].g;
    const int nt = (g.vh + 255) / 256;
    const int t = b - tiles[c];
    const int p = threadIdx.x >> 8;
    const int s = t * 256 + (threadIdx.x & 255);
    double v = 0.0;
    if (s < g.vh) v = plaquette_site<G3<Real>>(links + cd[c].link_off, g, p, s);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if ((threadIdx.x & 255) == 0) {
        double sum = 0.0;
#pragma unroll
        for (int w = 0; w < 4; ++w) sum += red[4 * p + w];
        part[begin[c] + p * nt + t] = sum;
    }
}

// One stout level of the spatial links of every chain.  begin[c] is the
// prefix table of 8 * ceil(vh/128) blocks per chain, so chain c owns the 6 k
// spatial blocks from 6 begin[c] / 8 on: 128-slot block major, then the six
// (mu < 3, parity) groups.  Logical blocks go to the XCDs in contiguous
// ranges, so a chain's blocks, and the neighbours they share, stay in one
// XCD's L2.  The temporal links are the same at every level:
// k_ens_copy_temporal puts them into both output buffers once.
template <typename Real>
__global__ __launch_bounds__(128) void k_ens_stout(const typename Vec2<Real>::type *__restrict__ in,
                                                   typename Vec2<Real>::type *__restrict__ out,
                                                   const ChainDesc *__restrict__ cd, const int *__restrict__ begin,
                                                   int nch, double rho) {
    const int b = xcd_tile(blockIdx.x, gridDim.x);
    const int c = chain_of(begin, nch, 8 * (b / 6));
    const Geom g = cd[c].g;
    const int lb = b - 6 * (begin[c] >> 3);
    const int sb = lb / 6, r = lb - 6 * sb;
    const int s = sb * 128 + threadIdx.x;
    const uint64_t off = cd[c].link_off;
    if (s < g.vh) stout_link<Real>(in + off, out + off, g, rho, r & 1, s, r >> 1);
}

// The temporal (mu = 3) links of every chain from `in` into out0 and out1:
// 2 * ceil(vh/128) blocks per chain (block major, then parity).
template <typename Real>
__global__ __launch_bounds__(128) void k_ens_copy_temporal(const typename Vec2<Real>::type *__restrict__ in,
                                                           typename Vec2<Real>::type *__restrict__ out0,
                                                           typename Vec2<Real>::type *__restrict__ out1,
                                                           const ChainDesc *__restrict__ cd,
                                                           const int *__restrict__ begin, int nch) {
    const int b = blockIdx.x;
    const int c = chain_of(begin, nch, 8 * (b >> 1));
    const Geom g = cd[c].g;
    const int lb = b - 2 * (begin[c] >> 3);
    const int s = (lb >> 1) * 128 + threadIdx.x, p = lb & 1;
    const uint64_t off = cd[c].link_off;
    if (s < g.vh) {
        const M3<Real> u = load3<Real>(in + off, g.vh, 3, p, s);
        store3<Real>(out0 + off, g.vh, 3, p, s, u);
        store3<Real>(out1 + off, g.vh, 3, p, s, u);
    }
}

// Per-site A1++ values of every chain (a1pp_slot): both parities of one
// 256-slot tile per 512-thread block, tiles[c+1]-tiles[c] = ceil(vh/256);
// chain c's values at sv + 2 sv_off.
template <typename Real>
__global__ __launch_bounds__(512) void k_ens_a1pp_sites(const typename Vec2<Real>::type *__restrict__ links,
                                                        const ChainDesc *__restrict__ cd,
                                                        const int *__restrict__ tiles,
                                                        const int *__restrict__ tile_chain,
                                                        double *__restrict__ sv) {
    const int b = xcd_tile(blockIdx.x, gridDim.x);  // contiguous tile ranges per XCD
    const int c = tile_chain[b];
    const Geom g = cd[c].g;
    a1pp_slot<Real>(links + cd[c].link_off, g, threadIdx.x >> 8, (b - tiles[c]) * 256 + (threadIdx.x & 255),
                    sv + 2 * cd[c].sv_off);
}

// A1++ partials: begin[c+1]-begin[c] = N3 * ceil(nsp/256), timeslice major;
// part[2 b + k], k = 0 plaquettes, 1 rectangles: block sums of 256
// consecutive spatial sites of one timeslice.
__global__ __launch_bounds__(256) void k_ens_a1pp(const double *__restrict__ sv, const ChainDesc *__restrict__ cd,
                                                  const int *__restrict__ begin, int nch,
                                                  double *__restrict__ part) {
    __shared__ double red[4];
    const int b = blockIdx.x;
    const int c = chain_of(begin, nch, b);
    const int nsp = cd[c].nsp;
    const int nsb = (nsp + 255) / 256;
    const int lb = b - begin[c];
    const int t = lb / nsb, site = (lb - t * nsb) * 256 + threadIdx.x;
    double ps = 0.0, rs = 0.0;
    if (site < nsp) {
        const size_t i = cd[c].sv_off + (size_t)t * nsp + site;
        ps = sv[2 * i];
        rs = sv[2 * i + 1];
    }
    const double a0 = block_sum<256>(ps, red);
    const double a1 = block_sum<256>(rs, red);
    if (threadIdx.x == 0) {
        part[2 * (size_t)b] = a0;
        part[2 * (size_t)b + 1] = a1;
    }
}

// out_plaq[ctr[1] * nch + c] = (sum of chain c's plaquette partials) / (6 V).
__global__ __launch_bounds__(256) void k_ens_reduce_plaq(const double *__restrict__ part,
                                                         const ChainDesc *__restrict__ cd,
                                                         const int *__restrict__ begin, int nch,
                                                         const uint64_t *__restrict__ ctr,
                                                         double *__restrict__ out_plaq) {
    __shared__ double red[4];
    const int c = blockIdx.x;
    double acc = 0.0;
    for (int i = begin[c] + threadIdx.x; i < begin[c + 1]; i += 256) acc += part[i];
    const double t = block_sum<256>(acc, red);
    if (threadIdx.x == 0) out_plaq[ctr[1] * (uint64_t)nch + c] = t / (6.0 * (double)cd[c].g.volume);
}

// Block (c, t): A1++ means of chain c at timeslice t for stout level `level`
// into out_ops[ctr[1] * ops_len + ((level * tsum + t_off[c] + t) * 2 + k)].
__global__ __launch_bounds__(256) void k_ens_reduce_a1pp(const double *__restrict__ part,
                                                         const ChainDesc *__restrict__ cd,
                                                         const int *__restrict__ begin, int level, int tsum,
                                                         uint64_t ops_len, const uint64_t *__restrict__ ctr,
                                                         double *__restrict__ out_ops) {
    __shared__ double red[4];
    const int c = blockIdx.x, t = blockIdx.y;
    const int T = cd[c].g.n[3];
    if (t >= T) return;  // uniform per block
    const int nsb = (cd[c].nsp + 255) / 256;
    const int base = begin[c] + t * nsb;
    double a0 = 0.0, a1 = 0.0;
    for (int i = threadIdx.x; i < nsb; i += 256) {
        a0 += part[2 * (size_t)(base + i)];
        a1 += part[2 * (size_t)(base + i) + 1];
    }
    const double s0 = block_sum<256>(a0, red);
    const double s1 = block_sum<256>(a1, red);
    if (threadIdx.x == 0) {
        const uint64_t o = ctr[1] * ops_len + ((uint64_t)level * tsum + cd[c].t_off + t) * 2;
        out_ops[o] = s0 / (3.0 * (double)cd[c].nsp);
        out_ops[o + 1] = s1 / (6.0 * (double)cd[c].nsp);
    }
}

__global__ void k_ens_advance(uint64_t *ctr, uint64_t sweeps, uint64_t slots) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        ctr[0] += sweeps;
        ctr[1] += slots;
    }
}

// Identity links over the concatenated buffer: every chain occupies whole
// 576-pair tiles.  Entry e of pair j of a tile (load3/store3 layout): FP64
// stores 9 runs of 64 pairs, e = j / 64; FP32 stores entries (2q, 2q+1) as
// float4 run q (pairs 128 q .. 128 q + 127, alternating), then entry 8.
template <typename Real>
__global__ void k_ens_identity(typename Vec2<Real>::type *__restrict__ links, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int j = (int)(i % 576);
    const int e = sizeof(Real) == 8 ? j >> 6 : (j >= 512 ? 8 : 2 * (j >> 7) + (j & 1));
    typename Vec2<Real>::type v;
    v.x = (e == 0 || e == 4 || e == 8) ? (Real)1 : (Real)0;
    v.y = (Real)0;
    links[i] = v;
}

}  // namespace ym

struct ym_ensemble {
    int nch = 0, prec = 0;
    std::vector<ym::ChainDesc> h_cd;
    std::vector<int> h_metro, h_plaq, h_stout, h_a1pp;  // prefix tables, nch + 1 entries
    std::vector<int> h_tile_chain;                       // chain of every Metropolis tile
    int tsum = 0, tmax = 0;
    uint64_t pairs = 0;      // link pairs of all chains
    void *links = nullptr;   // pairs elements of Vec2<Real>
    void *tmp[2] = {nullptr, nullptr};
    ym::ChainDesc *d_cd = nullptr;
    int *d_metro = nullptr, *d_plaq = nullptr, *d_stout = nullptr, *d_a1pp = nullptr;
    int *d_tile_chain = nullptr;  // chain of every Metropolis tile
    double *part = nullptr;  // max(plaq blocks, 2 * a1pp blocks)
    double *sv = nullptr;    // per-site A1++ values, 2 per site of every chain
    uint64_t vsum = 0;       // sites of all chains
    unsigned long long *d_acc = nullptr;
    uint64_t *d_ctr = nullptr;  // [0] next sweep index, [1] output slot
    uint64_t sweep = 0;         // host mirror of ctr[0]
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double *stage = nullptr;   // host-layout staging buffer, created on first transfer
    int64_t stage_sites = 0;
};

namespace {

int32_t hst(hipError_t e) { return e == hipSuccess ? YM_OK : YM_HIP_ERROR; }

template <typename Real>
void enqueue_sweeps(ym_ensemble *E, double beta, double angle, int hits, int nsweeps, int reunit) {
    using V2 = typename ym::Vec2<Real>::type;
    const double nbt = -(beta / 3.0);
    V2 *L = (V2 *)E->links;
    const int ntiles = E->h_metro[E->nch];
    const dim3 grid(ntiles);
    for (int k = 0; k < nsweeps; ++k)
        for (int mu = 0; mu < 4; ++mu)
            for (int p = 0; p < 2; ++p) {
                if (ntiles >= 512) {
                    // 8 and 10 hits per link have fully unrolled instances
#define YM_EWS(M, H) hipLaunchKernelGGL((ym::k_ens_metropolis_ws<Real, M, H>), dim3(256), dim3(512), 0, E->stream, L, E->d_cd, E->d_metro, E->d_tile_chain, ntiles, p, nbt, angle, hits, E->d_ctr, k, reunit, E->d_acc)
#define YM_EWS_MU(H)                    \
    switch (mu) {                       \
    case 0: YM_EWS(0, H); break;        \
    case 1: YM_EWS(1, H); break;        \
    case 2: YM_EWS(2, H); break;        \
    default: YM_EWS(3, H); break;       \
    }
                    if (hits == 8) {
                        YM_EWS_MU(8)
                    } else if (hits == 10) {
                        YM_EWS_MU(10)
                    } else {
                        YM_EWS_MU(0)
                    }
#undef YM_EWS_MU
#undef YM_EWS
                    continue;
                }
                switch (mu) {
                case 0: hipLaunchKernelGGL((ym::k_ens_metropolis<Real, 0>), grid, dim3(256), 0, E->stream, L, E->d_cd, E->d_metro, E->nch, p, nbt, angle, hits, E->d_ctr, k, reunit, E->d_acc); break;
                case 1: hipLaunchKernelGGL((ym::k_ens_metropolis<Real, 1>), grid, dim3(256), 0, E->stream, L, E->d_cd, E->d_metro, E->nch, p, nbt, angle, hits, E->d_ctr, k, reunit, E->d_acc); break;
                case 2: hipLaunchKernelGGL((ym::k_ens_metropolis<Real, 2>), grid, dim3(256), 0, E->stream, L, E->d_cd, E->d_metro, E->nch, p, nbt, angle, hits, E->d_ctr, k, reunit, E->d_acc); break;
                default: hipLaunchKernelGGL((ym::k_ens_metropolis<Real, 3>), grid, dim3(256), 0, E->stream, L, E->d_cd, E->d_metro, E->nch, p, nbt, angle, hits, E->d_ctr, k, reunit, E->d_acc); break;
                }
            }
}

// Plaquette + A1++ on raw and `levels` stout levels, into output slot ctr[1].
template <typename Real>
void enqueue_measure(ym_ensemble *E, int levels, double rho, double *out_plaq, double *out_ops) {
    using V2 = typename ym::Vec2<Real>::type;
    const V2 *L = (const V2 *)E->links;
    const uint64_t ops_len = (uint64_t)(levels + 1) * E->tsum * 2;
    if constexpr (sizeof(Real) == 4)
        hipLaunchKernelGGL((ym::k_ens_plaquette_pp<Real>), dim3(E->h_metro[E->nch]), dim3(512), 0, E->stream, L,
                           E->d_cd, E->d_metro, E->d_tile_chain, E->d_plaq, E->part);
    else
        hipLaunchKernelGGL((ym::k_ens_plaquette<Real>), dim3(E->h_plaq[E->nch]), dim3(256), 0, E->stream, L,
                           E->d_cd, E->d_plaq, E->nch, E->part);
    hipLaunchKernelGGL(ym::k_ens_reduce_plaq, dim3(E->nch), dim3(256), 0, E->stream, E->part, E->d_cd, E->d_plaq,
                       E->nch, E->d_ctr, out_plaq);
    const V2 *cur = L;
    if (levels > 0)
        hipLaunchKernelGGL((ym::k_ens_copy_temporal<Real>), dim3(E->h_stout[E->nch] / 4), dim3(128), 0, E->stream, L,
                           (V2 *)E->tmp[0], (V2 *)E->tmp[1], E->d_cd, E->d_stout, E->nch);
    for (int lev = 0; lev <= levels; ++lev) {
        if (lev > 0) {
            V2 *dst = (V2 *)E->tmp[(lev - 1) & 1];
            hipLaunchKernelGGL((ym::k_ens_stout<Real>), dim3(E->h_stout[E->nch] / 8 * 6), dim3(128), 0, E->stream,
                               cur, dst, E->d_cd, E->d_stout, E->nch, rho);
            cur = dst;
        }
        hipLaunchKernelGGL((ym::k_ens_a1pp_sites<Real>), dim3(E->h_metro[E->nch]), dim3(512), 0, E->stream, cur,
                           E->d_cd, E->d_metro, E->d_tile_chain, E->sv);
        hipLaunchKernelGGL(ym::k_ens_a1pp, dim3(E->h_a1pp[E->nch]), dim3(256), 0, E->stream, E->sv, E->d_cd,
                           E->d_a1pp, E->nch, E->part);
        hipLaunchKernelGGL(ym::k_ens_reduce_a1pp, dim3(E->nch, E->tmax), dim3(256), 0, E->stream, E->part, E->d_cd,
                           E->d_a1pp, lev, E->tsum, ops_len, E->d_ctr, out_ops);
    }
}

int32_t validate_run(const ym_ensemble *E, double beta, double angle, int hits) {
    if (!E || hits < 1) return YM_INVALID_ARGUMENT;
    if (!std::isfinite(beta) || beta < 0.0 || beta > 12.0) return YM_BETA_OUT_OF_RANGE;
    if (!std::isfinite(angle) || angle < 1.0e-8 || angle > 1.0) return YM_ANGLE_OUT_OF_RANGE;
    return YM_OK;
}

int32_t validate_measure(const ym_ensemble *E, int levels, double rho) {
    if (levels < 0 || levels > 64) return YM_TOO_MANY_LEVELS;
    if (levels > 0) {
        if (!std::isfinite(rho)) return YM_NON_FINITE_RHO;
        if (rho <= 0.0 || rho > 0.25) return YM_RHO_OUT_OF_RANGE;
    }
    for (const auto &c : E->h_cd) {
        for (int d = 0; d < 3; ++d)
            if (c.g.n[d] < 3) return YM_SPATIAL_EXTENT_TOO_SMALL;
        if (c.g.n[0] != c.g.n[1] || c.g.n[1] != c.g.n[2]) return YM_SPATIAL_NOT_CUBIC;
    }
    return YM_OK;
}

}  // namespace

extern "C" {

int32_t ym_ensemble_create(int32_t nchains, const int64_t *extents, const uint64_t *keys, int32_t precision,
                           ym_ensemble **out) {
    if (!extents || !keys || !out || nchains < 1) return YM_INVALID_ARGUMENT;
    *out = nullptr;
    if (precision != 32 && precision != 64) return YM_INVALID_ARGUMENT;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return YM_NO_DEVICE;
    ym_ensemble *E = new (std::nothrow) ym_ensemble();
    if (!E) return YM_ALLOCATION_FAILED;
    E->nch = nchains;
    E->prec = precision;
    E->h_cd.resize(nchains);
    E->h_metro.assign(nchains + 1, 0);
    E->h_plaq.assign(nchains + 1, 0);
    E->h_stout.assign(nchains + 1, 0);
    E->h_a1pp.assign(nchains + 1, 0);
    int64_t blocks[4] = {0, 0, 0, 0};
    for (int c = 0; c < nchains; ++c) {
        const int64_t *n = extents + 4 * c;
        int64_t vol = 1;
        for (int d = 0; d < 4; ++d) {
            if (n[d] < 2) { delete E; return YM_INVALID_EXTENT; }
            if (n[d] % 2) { delete E; return YM_UNSUPPORTED; }
            vol *= n[d];
            if (vol > ((int64_t)1 << 30)) { delete E; return YM_LATTICE_TOO_LARGE; }
        }
        ym::ChainDesc &d = E->h_cd[c];
        d.g = ym::make_geom(n);
        d.key = keys[c];
        d.link_off = E->pairs;
        d.nsp = d.g.n[0] * d.g.n[1] * d.g.n[2];
        d.t_off = E->tsum;
        d.sv_off = E->vsum;
        E->vsum += (uint64_t)d.g.volume;
        E->pairs += ym::su3_pairs(d.g.vh);
        E->tsum += d.g.n[3];
        E->tmax = std::max(E->tmax, d.g.n[3]);
        const int64_t nt = (d.g.vh + 255) / 256;
        blocks[0] += nt;
        blocks[1] += 2 * nt;
        blocks[2] += 8 * (int64_t)((d.g.vh + 127) / 128);
        blocks[3] += (int64_t)d.g.n[3] * ((d.nsp + 255) / 256);
        for (int k = 0; k < 4; ++k)
            if (blocks[k] > 0x7fffffff) { delete E; return YM_LATTICE_TOO_LARGE; }
        E->h_metro[c + 1] = (int)blocks[0];
        E->h_plaq[c + 1] = (int)blocks[1];
        E->h_stout[c + 1] = (int)blocks[2];
        E->h_a1pp[c + 1] = (int)blocks[3];
    }
    const size_t pair_bytes = 2 * (size_t)(precision / 8);
    const size_t link_bytes = E->pairs * pair_bytes;
    const size_t nparts = (size_t)std::max(E->h_plaq[nchains], 2 * E->h_a1pp[nchains]);
    int32_t st = YM_OK;
    auto alloc = [&](void **p, size_t bytes, int32_t code) {
        if (!st && hipMalloc(p, bytes) != hipSuccess) st = code;
    };
    if (hipStreamCreateWithFlags(&E->stream, hipStreamNonBlocking) != hipSuccess) st = YM_HIP_ERROR;
    alloc(&E->links, link_bytes, YM_LATTICE_TOO_LARGE);
    alloc(&E->tmp[0], link_bytes, YM_LATTICE_TOO_LARGE);
    alloc(&E->tmp[1], link_bytes, YM_LATTICE_TOO_LARGE);
    alloc((void **)&E->d_cd, sizeof(ym::ChainDesc) * nchains, YM_ALLOCATION_FAILED);
    alloc((void **)&E->d_metro, sizeof(int) * (nchains + 1), YM_ALLOCATION_FAILED);
    alloc((void **)&E->d_tile_chain, sizeof(int) * std::max(1, E->h_metro[nchains]), YM_ALLOCATION_FAILED);
    alloc((void **)&E->d_plaq, sizeof(int) * (nchains + 1), YM_ALLOCATION_FAILED);
    alloc((void **)&E->d_stout, sizeof(int) * (nchains + 1), YM_ALLOCATION_FAILED);
    alloc((void **)&E->d_a1pp, sizeof(int) * (nchains + 1), YM_ALLOCATION_FAILED);
    alloc((void **)&E->part, sizeof(double) * nparts, YM_ALLOCATION_FAILED);
    alloc((void **)&E->sv, sizeof(double) * 2 * E->vsum, YM_ALLOCATION_FAILED);
    alloc((void **)&E->d_acc, sizeof(unsigned long long) * nchains, YM_ALLOCATION_FAILED);
    alloc((void **)&E->d_ctr, sizeof(uint64_t) * 2, YM_ALLOCATION_FAILED);
    if (!st && (hipEventCreate(&E->ev0) != hipSuccess || hipEventCreate(&E->ev1) != hipSuccess)) st = YM_HIP_ERROR;
    if (!st) st = hst(hipMemcpyAsync(E->d_cd, E->h_cd.data(), sizeof(ym::ChainDesc) * nchains,
                                     hipMemcpyHostToDevice, E->stream));
    const std::vector<int> *tabs[4] = {&E->h_metro, &E->h_plaq, &E->h_stout, &E->h_a1pp};
    int *dtabs[4] = {E->d_metro, E->d_plaq, E->d_stout, E->d_a1pp};
    for (int k = 0; k < 4 && !st; ++k)
        st = hst(hipMemcpyAsync(dtabs[k], tabs[k]->data(), sizeof(int) * (nchains + 1), hipMemcpyHostToDevice,
                                E->stream));
    E->h_tile_chain.assign(std::max(1, E->h_metro[nchains]), 0);
    for (int c = 0; c < nchains; ++c)
        for (int t = E->h_metro[c]; t < E->h_metro[c + 1]; ++t) E->h_tile_chain[t] = c;
    if (!st)
        st = hst(hipMemcpyAsync(E->d_tile_chain, E->h_tile_chain.data(), sizeof(int) * E->h_tile_chain.size(),
                                hipMemcpyHostToDevice, E->stream));
    if (!st) st = ym_ensemble_set_identity(E);
    if (st) {
        ym_ensemble_destroy(E);
        return st;
    }
    *out = E;
    return YM_OK;
}

int32_t ym_ensemble_destroy(ym_ensemble *E) {
    if (!E) return YM_INVALID_ARGUMENT;
    if (E->stream) (void)hipStreamSynchronize(E->stream);
    void *bufs[] = {E->links, E->tmp[0], E->tmp[1], E->d_cd, E->d_metro, E->d_tile_chain, E->d_plaq, E->d_stout,
                    E->d_a1pp,
                    E->part, E->sv, E->d_acc, E->d_ctr, E->stage};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    if (E->ev0) (void)hipEventDestroy(E->ev0);
    if (E->ev1) (void)hipEventDestroy(E->ev1);
    if (E->stream) (void)hipStreamDestroy(E->stream);
    delete E;
    return YM_OK;
}

int32_t ym_ensemble_set_identity(ym_ensemble *E) {
    if (!E) return YM_INVALID_ARGUMENT;
    const dim3 grid((unsigned)((E->pairs + 255) / 256));
    if (E->prec == 32) hipLaunchKernelGGL(ym::k_ens_identity<float>, grid, dim3(256), 0, E->stream, (float2 *)E->links, E->pairs);
    else hipLaunchKernelGGL(ym::k_ens_identity<double>, grid, dim3(256), 0, E->stream, (double2 *)E->links, E->pairs);
    ym::launch_fill_zero_u64((unsigned long long *)E->d_ctr, 2, E->stream);
    E->sweep = 0;
    int32_t st = hst(hipGetLastError());
    if (!st) st = hst(hipStreamSynchronize(E->stream));
    return st;
}

int32_t ym_ensemble_chain_count(const ym_ensemble *E, int32_t *nchains, int64_t *total_sites, int64_t *time_sum) {
    if (!E || !nchains) return YM_INVALID_ARGUMENT;
    *nchains = E->nch;
    int64_t v = 0;
    for (const auto &c : E->h_cd) v += c.g.volume;
    if (total_sites) *total_sites = v;
    if (time_sum) *time_sum = E->tsum;
    return YM_OK;
}

// Staging buffer for one chain's host image: sized for the largest chain
// (at most 2^16 sites), created on the first transfer and kept.
static int32_t ens_ensure_stage(ym_ensemble *E) {
    if (E->stage) return YM_OK;
    int64_t vmax = 0;
    for (const auto &c : E->h_cd) vmax = std::max<int64_t>(vmax, c.g.volume);
    const int64_t sites = ym::stage_sites_for(vmax);
    if (hipMalloc((void **)&E->stage, sizeof(double) * (size_t)sites * 4 * 18) != hipSuccess) {
        E->stage = nullptr;
        return YM_ALLOCATION_FAILED;
    }
    E->stage_sites = sites;
    return YM_OK;
}

// Checked as ym_lattice_upload (Su3::from_checked_matrix, su3.rs:242-253).
int32_t ym_ensemble_upload_chain(ym_ensemble *E, int32_t chain, const double *aos, int64_t count) {
    if (!E || !aos || chain < 0 || chain >= E->nch) return YM_INVALID_ARGUMENT;
    const ym::ChainDesc &d = E->h_cd[chain];
    if (count != (int64_t)d.g.volume * 4 * 18) return YM_INVALID_ARGUMENT;
    int32_t st = ym::validate_aos(3, aos, count / 18, ym::link_check_tolerance(3, E->prec));
    if (!st) st = ens_ensure_stage(E);
    if (st) return st;
    const size_t pb = 2 * (size_t)(E->prec / 8);
    return ym::upload_aos_staged(3, E->prec, (char *)E->links + d.link_off * pb, d.g, aos, E->stage, E->stage_sites,
                                 E->stream);
}

int32_t ym_ensemble_download_chain(const ym_ensemble *Ec, int32_t chain, double *aos, int64_t count) {
    ym_ensemble *E = const_cast<ym_ensemble *>(Ec);
    if (!E || !aos || chain < 0 || chain >= E->nch) return YM_INVALID_ARGUMENT;
    const ym::ChainDesc &d = E->h_cd[chain];
    if (count != (int64_t)d.g.volume * 4 * 18) return YM_INVALID_ARGUMENT;
    const int32_t st = ens_ensure_stage(E);
    if (st) return st;
    const size_t pb = 2 * (size_t)(E->prec / 8);
    return ym::download_aos_staged(3, E->prec, (const char *)E->links + d.link_off * pb, d.g, aos, E->stage,
                                   E->stage_sites, E->stream);
}

int32_t ym_ensemble_hot_start(ym_ensemble *E) {
    if (!E) return YM_INVALID_ARGUMENT;
    const size_t pb = 2 * (size_t)(E->prec / 8);
    for (const auto &d : E->h_cd) ym::launch_su3_hot(E->prec, (char *)E->links + d.link_off * pb, d.g, d.key, E->stream);
    int32_t st = hst(hipGetLastError());
    if (!st) st = hst(hipStreamSynchronize(E->stream));
    return st;
}

int32_t ym_ensemble_sweep_index(const ym_ensemble *E, uint64_t *sweep) {
    if (!E || !sweep) return YM_INVALID_ARGUMENT;
    *sweep = E->sweep;
    return YM_OK;
}

int32_t ym_ensemble_ops_length(const ym_ensemble *E, int32_t stout_levels, int64_t *ops_len, int64_t *chain_offsets) {
    if (!E || !ops_len || stout_levels < 0) return YM_INVALID_ARGUMENT;
    *ops_len = (int64_t)(stout_levels + 1) * E->tsum * 2;
    if (chain_offsets)
        for (int c = 0; c < E->nch; ++c) chain_offsets[c] = E->h_cd[c].t_off;
    return YM_OK;
}

int32_t ym_ensemble_run_cycles(ym_ensemble *E, double beta, double angle, int32_t hits, int32_t reunitarize,
                               int32_t sweeps_per_cycle, int32_t stout_levels, double rho, int32_t n_cycles,
                               int32_t use_graph, double *plaquette, double *ops, uint64_t *accepts,
                               float *device_ms) {
    int32_t st = validate_run(E, beta, angle, hits);
    if (st) return st;
    if (sweeps_per_cycle < 0 || n_cycles < 1 || (sweeps_per_cycle == 0 && !plaquette)) return YM_INVALID_ARGUMENT;
    const bool measure = plaquette != nullptr;
    if (measure && (st = validate_measure(E, stout_levels, rho))) return st;
    if (measure && !ops) return YM_INVALID_ARGUMENT;
    const int reunit = (E->prec == 32) ? 1 : (reunitarize != 0);
    const uint64_t ops_len = (uint64_t)(stout_levels + 1) * E->tsum * 2;
    double *d_plaq = nullptr, *d_ops = nullptr;
    if (measure) {
        if (hipMalloc((void **)&d_plaq, sizeof(double) * n_cycles * E->nch) != hipSuccess) return YM_ALLOCATION_FAILED;
        if (hipMalloc((void **)&d_ops, sizeof(double) * n_cycles * ops_len) != hipSuccess) {
            (void)hipFree(d_plaq);
            return YM_ALLOCATION_FAILED;
        }
    }
    // output slot counter restarts for this call; the sweep counter persists
    const uint64_t h_ctr[2] = {E->sweep, 0};
    st = hst(hipMemcpyAsync(E->d_ctr, h_ctr, sizeof(h_ctr), hipMemcpyHostToDevice, E->stream));
    ym::launch_fill_zero_u64(E->d_acc, E->nch, E->stream);
    auto cycle = [&]() {
        if (E->prec == 32) {
            enqueue_sweeps<float>(E, beta, angle, hits, sweeps_per_cycle, reunit);
            if (measure) enqueue_measure<float>(E, stout_levels, rho, d_plaq, d_ops);
        } else {
            enqueue_sweeps<double>(E, beta, angle, hits, sweeps_per_cycle, reunit);
            if (measure) enqueue_measure<double>(E, stout_levels, rho, d_plaq, d_ops);
        }
        hipLaunchKernelGGL(ym::k_ens_advance, dim3(1), dim3(64), 0, E->stream, E->d_ctr, (uint64_t)sweeps_per_cycle,
                           (uint64_t)(measure ? 1 : 0));
    };
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    if (!st) st = hst(hipStreamSynchronize(E->stream));
    if (!st && use_graph) {
        st = hst(hipStreamBeginCapture(E->stream, hipStreamCaptureModeThreadLocal));
        if (!st) {
            cycle();
            const int32_t cst = hst(hipStreamEndCapture(E->stream, &graph));
            st = cst ? cst : hst(hipGetLastError());
        }
        if (!st) st = hst(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    }
    if (!st) st = hst(hipEventRecord(E->ev0, E->stream));
    for (int k = 0; k < n_cycles && !st; ++k) {
        if (use_graph) st = hst(hipGraphLaunch(exec, E->stream));
        else {
            cycle();
            st = hst(hipGetLastError());
        }
    }
    if (!st) st = hst(hipEventRecord(E->ev1, E->stream));
    if (!st) st = hst(hipStreamSynchronize(E->stream));
    if (!st && device_ms) st = hst(hipEventElapsedTime(device_ms, E->ev0, E->ev1));
    if (!st && measure) {
        st = hst(hipMemcpy(plaquette, d_plaq, sizeof(double) * n_cycles * E->nch, hipMemcpyDeviceToHost));
        if (!st) st = hst(hipMemcpy(ops, d_ops, sizeof(double) * n_cycles * ops_len, hipMemcpyDeviceToHost));
    }
    if (!st && accepts) {
        std::vector<unsigned long long> a(E->nch);
        st = hst(hipMemcpy(a.data(), E->d_acc, sizeof(unsigned long long) * E->nch, hipMemcpyDeviceToHost));
        for (int c = 0; c < E->nch && !st; ++c) accepts[c] += a[c];
    }
    if (!st) E->sweep += (uint64_t)sweeps_per_cycle * (uint64_t)n_cycles;
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (d_plaq) (void)hipFree(d_plaq);
    if (d_ops) (void)hipFree(d_ops);
    return st;
}

}  // extern "C"
