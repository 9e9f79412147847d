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
void lattice_aux_events(const ym_lattice *lat, hipEvent_t *e0, hipEvent_t *e1);
hipStream_t lattice_stream(const ym_lattice *lat);

void launch_su3_metropolis(int prec, void *links, const Geom &g, int mu, int p, double beta, double angle, int hits,
                           uint64_t key, uint64_t sweep, int reunit, unsigned long long *acc, hipStream_t st);
void launch_su3_hot(int prec, void *links, const Geom &g, uint64_t key, hipStream_t st);
void launch_su3_staple_probe(int prec, const void *links, const Geom &g, const int x[4], int mu, double *out, hipStream_t st);
void launch_su3_reunit(int prec, void *links, const Geom &g, hipStream_t st);
int launch_su3_residuals(int prec, const void *links, const Geom &g, double *part, hipStream_t st);
void launch_su3_from_aos(int prec, void *links, const Geom &g, const double *aos, hipStream_t st);
void launch_su3_to_aos(int prec, const void *links, const Geom &g, double *aos, hipStream_t st);

void launch_su2_heatbath(int prec, void *links, const Geom &g, int mu, int p, double beta, uint64_t key, uint64_t sweep,
                         int max_attempts, unsigned long long *counters, int *status, hipStream_t st);
void launch_su2_overrelax(int prec, void *links, const Geom &g, int mu, int p, hipStream_t st);
void launch_su2_hot(int prec, void *links, const Geom &g, uint64_t key, hipStream_t st);
void launch_su2_from_aos(int prec, void *links, const Geom &g, const double *aos, hipStream_t st);
void launch_su2_to_aos(int prec, const void *links, const Geom &g, double *aos, hipStream_t st);
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

}  // namespace ym
