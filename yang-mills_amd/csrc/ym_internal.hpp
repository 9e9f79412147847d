// Host-side declarations of the kernel launchers (one per kernel family).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

struct ym_lattice;

namespace ym {

struct Geom;

// Accessors of a lattice handle for the other kernel families.
int32_t lattice_info(const ym_lattice *lat, int32_t *group, int32_t *prec, int64_t ext[4]);
void *lattice_links(const ym_lattice *lat);
unsigned long long *lattice_counters(const ym_lattice *lat);  // 4 device counters of the handle
unsigned long long *lattice_tally(const ym_lattice *lat);     // tally_slots(g) slots of the handle
void lattice_aux_events(const ym_lattice *lat, hipEvent_t *e0, hipEvent_t *e1);
hipStream_t lattice_stream(const ym_lattice *lat);
// The handle's byte buffer, grown to at least `bytes` (nullptr if that fails).
void *lattice_aux(const ym_lattice *lat, size_t bytes);

void launch_su3_metropolis(int prec, void *links, const Geom &g, int mu, int p, double beta, double angle, int hits,
                           uint64_t key, uint64_t sweep, int reunit, unsigned long long *tally, int *status,
                           hipStream_t st);
void launch_su3_hot(int prec, void *links, const Geom &g, uint64_t key, hipStream_t st);
void launch_su3_staple_probe(int prec, const void *links, const Geom &g, const int x[4], int mu, double *out, hipStream_t st);
void launch_su3_reunit(int prec, void *links, const Geom &g, hipStream_t st);
int launch_su3_residuals(int prec, const void *links, const Geom &g, double *part, hipStream_t st);
void launch_su3_from_aos(int prec, void *links, const Geom &g, const double *aos, hipStream_t st);
void launch_su3_to_aos(int prec, const void *links, const Geom &g, double *aos, hipStream_t st);
void launch_su3_from_aos_range(int prec, void *links, const Geom &g, const double *aos, int site0, int nsites,
                               hipStream_t st);
void launch_su3_to_aos_range(int prec, const void *links, const Geom &g, double *aos, int site0, int nsites,
                             hipStream_t st);
void launch_su3_identity(int prec, void *links, const Geom &g, hipStream_t st);

// attempts go to the tally, links that exhaust max_attempts to counters[1]
void launch_su2_heatbath(int prec, void *links, const Geom &g, int mu, int p, double beta, uint64_t key, uint64_t sweep,
                         int max_attempts, unsigned long long *tally, unsigned long long *counters, int *status,
                         hipStream_t st);
void launch_su2_overrelax(int prec, void *links, const Geom &g, int mu, int p, hipStream_t st);
void launch_su2_hot(int prec, void *links, const Geom &g, uint64_t key, hipStream_t st);
void launch_su2_from_aos(int prec, void *links, const Geom &g, const double *aos, hipStream_t st);
void launch_su2_to_aos(int prec, const void *links, const Geom &g, double *aos, hipStream_t st);
void launch_su2_from_aos_range(int prec, void *links, const Geom &g, const double *aos, int site0, int nsites,
                               hipStream_t st);
void launch_su2_to_aos_range(int prec, const void *links, const Geom &g, double *aos, int site0, int nsites,
                             hipStream_t st);
void launch_su2_identity(int prec, void *links, const Geom &g, hipStream_t st);
void launch_su2_normalize(int prec, void *links, const Geom &g, hipStream_t st);
void launch_su2_staple_probe(int prec, const void *links, const Geom &g, const int x[4], int mu, double *out, hipStream_t st);
int launch_su2_residual(int prec, const void *links, const Geom &g, double *part, hipStream_t st);

int launch_plaquette(int group, int prec, const void *links, const Geom &g, double *part, hipStream_t st);
int launch_wilson_loops(int group, int prec, const void *links, const Geom &g, int mu, int nu, int rmax, int tmax,
                        double *part, hipStream_t st);
int launch_polyakov(int group, int prec, const void *links, const Geom &g, int dir, double *part, hipStream_t st);
// sv: 2 * volume doubles of per-site values; part: 2 * N3 * ceil(L^3 / 256)
// partials; returns the partials per (timeslice, k).
int launch_a1pp(int prec, const void *links, const Geom &g, double *sv, double *part, hipStream_t st);
void launch_stout(int prec, const void *in, void *out, const Geom &g, double rho, hipStream_t st);
void launch_block2(int prec, const void *in, void *out, const Geom &gf, const Geom &gc, hipStream_t st);
void launch_path_holonomy(int group, int prec, const void *links, const Geom &g, const int x[4], const int *steps,
                          int nsteps, double *out, hipStream_t st);
void launch_sum_partials(const double *part, int n, int ncols, int stride, double *out, hipStream_t st);
void launch_max_partials(const double *part, int n, int ncols, int stride, double *out, hipStream_t st);
void launch_fill_zero(double *p, size_t n, hipStream_t st);
void launch_fill_zero_u64(unsigned long long *p, size_t n, hipStream_t st);
void launch_fill_zero_i32(int *p, size_t n, hipStream_t st);
void launch_copy16(const void *src, void *dst, size_t bytes, hipStream_t st);

// Counts of the sweep kernels (accepted proposals, heat-bath attempts) go to
// a tally of per-block slots: block b (wave b for the heat-bath sampler) of a
// launch adds its count to slot b with a plain load and store.  One writer
// per slot per launch, launches ordered on one stream: no atomics, which on
// one address cost the sweep kernels tens of microseconds per launch.
// tally_slots(g) slots cover every grid those launchers use on g;
// launch_tally_sum stores the sum of n slots to out[0].
int64_t tally_slots(const Geom &g);
void launch_tally_sum(const unsigned long long *tally, int64_t n, unsigned long long *out, hipStream_t st);

// Host AoS image (reference layout) <-> tiled field of one lattice through a
// device staging buffer of `stage_sites` sites: the image is moved in
// lexicographic site chunks on stream `st`; returns a YM_* status.
int32_t upload_aos_staged(int group, int prec, void *links, const Geom &g, const double *aos, double *stage,
                          int64_t stage_sites, hipStream_t st);
int32_t download_aos_staged(int group, int prec, const void *links, const Geom &g, double *aos, double *stage,
                            int64_t stage_sites, hipStream_t st);
// Sites per staging buffer for a lattice of `volume` sites.
int64_t stage_sites_for(int64_t volume);
// The reference's checks of a host AoS image before it is stored (fail
// closed): SU(3) Su3::from_checked_matrix (su3.rs:242-253) -- finite,
// unitarity and determinant residuals within `tol`; SU(2)
// Su2::from_normalized_components / from_preserved_unit_components
// (lib.rs:155-180) -- finite, non-zero, unit norm within `tol`.  The first
// failing link in site order decides the status.
int32_t validate_aos(int group, const double *aos, int64_t nlinks, double tol);
// Tolerance of those checks for a storage precision: the reference's 5e-12
// (SU(3), su3.rs:20) / 2e-12 (SU(2), lib.rs:20) for FP64 storage, 2^-16 for
// FP32 storage (links rounded to FP32 carry residuals near 1e-7).
double link_check_tolerance(int group, int prec);

}  // namespace ym
