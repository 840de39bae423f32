/*
 * ax_internal.h - declarations shared by the library's translation units.
 */
#ifndef AXIONS_AX_INTERNAL_H
#define AXIONS_AX_INTERNAL_H

#include <cstdarg>

#include <cstdlib>

#include "ax_device.h"

struct ax_lattice;

namespace ax {

/* Work items of a marching kernel: one 64-lane wave per (64-site z chunk,
 * row, x segment of xseg planes), numbered as make_tile_grid describes. */
struct TileGrid {
  int nw;      /* z chunks per row */
  int xseg;    /* x planes marched by one wave */
  int work;    /* nw * ny * nseg */
  int per;     /* XCD-aware order: work items per XCD per x segment, 0 = natural order */
  int rows;    /* 1: within an XCD's share, all chunks of a row together (chunk fastest) */
};

/* Work items of the marching kernels, one per (z chunk, y row, x segment).
 * Workgroups are handed to the eight XCDs round-robin (workgroup w to XCD
 * w mod 8), each XCD with its own L2.  In the natural order (z chunk
 * fastest) XCD x receives chunk x mod 8 of every row: neighbouring chunks of
 * a row, which share the cache lines at their z edges, sit on different
 * XCDs, and with fewer than 8 chunks per row even neighbouring rows do.  The
 * XCD-aware order gives XCD x the x-th eighth of each x segment's items:
 * with ny divisible by 8, a band of ny/8 consecutive rows with all their
 * chunks, chunk fastest (512^3 FP64 reads 15.4 -> 12.9 GB per step, FP32
 * 9.0 -> 6.1 GB); otherwise a run of consecutive rows of one chunk. */
inline TileGrid make_tile_grid(int nx, int ny, int nz, int xseg) {
  TileGrid g;
  g.xseg = (xseg <= 0 || xseg > nx) ? nx : xseg;
  g.nw = (nz + 63) / 64;
  g.work = g.nw * ny * ((nx + g.xseg - 1) / g.xseg);
  const char *o = getenv("AX_TILE_ORDER");        /* "natural" or "chunks" for measurements */
  const bool natural = o && o[0] == 'n';
  g.per = (!natural && (g.nw * ny) % 8 == 0) ? g.nw * ny / 8 : 0;
  g.rows = !(o && o[0] == 'c') && ny % 8 == 0;
  return g;
}

/* (z chunk, y row, x segment) of work item w in the order make_tile_grid chose */
__device__ __forceinline__ void tile_item(const TileGrid &g, int ny, int w, int &zc, int &y, int &seg) {
  if (g.per && g.rows) {
    const int xcd = w & 7, i = w >> 3, s = i / g.per;
    const int item = (s * 8 + xcd) * g.per + (i - s * g.per);   /* chunk + nw * (row + ny * segment) */
    zc = item % g.nw;
    const int r = item / g.nw;
    y = r % ny;
    seg = r / ny;
  } else if (g.per) {
    const int xcd = w & 7, i = w >> 3, s = i / g.per;
    const int item = (s * 8 + xcd) * g.per + (i - s * g.per);   /* row + ny * (chunk + nw * segment) */
    y = item % ny;
    const int r = item / ny;
    zc = r % g.nw;
    seg = r / g.nw;
  } else {
    zc = w % g.nw;
    const int r = w / g.nw;
    y = r % ny;
    seg = r / ny;
  }
}

struct DiagConsts {
  int nx, ny, nz, ex, ez, model, energy_form, ncomp;
  double dx, dy, dz, dt, g, lambda, eta;
  double lambda1, lambda2, lambda12, eta1, eta2, q1, q2;
  double inner; /* (nx-2)(ny-2)(nz-2) as a double */
};

int fail(int code, const char *fmt, ...);
int hip_fail(hipError_t e, const char *what);
int check_params(const ax_params *p);
StepConsts step_consts(const ax_params *p, int n);
DiagConsts diag_consts(const ax_params *p);

template <typename T>
Levels<T> levels_of(ax_lattice *L, int n);
hipStream_t lattice_stream(ax_lattice *L);
const ax_params &lattice_params(const ax_lattice *L);
void *lattice_state(ax_lattice *L);
size_t lattice_sites(const ax_lattice *L);
size_t lattice_pitch(const ax_lattice *L);
int lattice_device(const ax_lattice *L);
int lattice_timestep(const ax_lattice *L);
void lattice_reset_timestep(ax_lattice *L);

int launch_step(ax_lattice *L, int n);
int ensure_winding_buffer(ax_lattice *L);
int ensure_phase_buffer(ax_lattice *L);
hipError_t launch_march_diag(ax_lattice *L, bool phases);
hipError_t launch_phase_winding(ax_lattice *L, int8_t *out, double *res3);
int prepare_faces(ax_lattice *L);
int launch_winding_field(ax_lattice *L, int field, int8_t *out, double *part, double *res3);

template <typename R>
hipError_t launch_step_site(const StepConsts &k, const Levels<R> &L, int model, hipStream_t st);
bool use_tiled_ah(const ax_params &p);
int default_xseg(const ax_params &p);
template <typename R>
hipError_t launch_step_tiled_ah(const StepConsts &k, const Levels<R> &L, int xseg, hipStream_t st);
/* band: the lattice's three per-level band flags (nullptr: none); t the
 * TimeStep; use_band: band[t % 3] was written by the previous step; *fd_used
 * tells whether the kernel kept the flags (FP64 FixedDivisor path) */
template <typename R>
hipError_t launch_step_tiled_moore(const StepConsts &k, const Levels<R> &L, int xseg, hipStream_t st, int *band, int t,
                                   bool use_band, bool *fd_used);

template <typename R>
hipError_t launch_global_step(const ax_params &p, void *state, size_t pitch, int n, hipStream_t st);
template <typename R>
hipError_t launch_global_energy(const ax_params &p, const void *state, size_t pitch, int n, double *part,
                                double *out3, hipStream_t st);

template <typename T>
hipError_t launch_diag(const DiagConsts &k, const Levels<T> &L, double *part, double *out3, hipStream_t st);
/* periodic Abelian-Higgs: energy/Gauss march, with phase != NULL also the
 * FP32 edge values [3][V] (forward edge phase + link angle, ax_wind.h);
 * launch_wind_phase turns them into windings (the field into out when
 * out != NULL, the counts into out3) */
constexpr int kDiagXseg = 32;
bool use_diag_march(const ax_params &p);
bool use_diag_march_windings(const ax_params &p);
size_t diag_part_blocks(const ax_params &p);
template <typename T>
hipError_t launch_diag_march(const DiagConsts &k, const Levels<T> &L, int xseg, float *phase, double *part,
                             double *out3, hipStream_t st);
hipError_t launch_wind_phase(int nx, int ny, int nz, const float *E, int8_t *out, double *part, double *out3,
                             hipStream_t st);
/* periodic lattices: the FP32 edge values of field (fr, fi) of charge q into
 * E, then the windings from them, as launch_wind_phase */
template <typename T>
hipError_t launch_edge_winding(int nx, int ny, int nz, const T *fr, const T *fi, const T *tx, const T *ty,
                               const T *tz, double q, float *E, int8_t *out, double *part, double *out3,
                               hipStream_t st);
template <typename T>
hipError_t launch_winding(int nx, int ny, int nz, const T *fr, const T *fi, const T *tx, const T *ty, const T *tz,
                          double q, int ex, int ez, int8_t *out, double *part, double *out3, hipStream_t st);
template <typename T>
hipError_t launch_crossings(int nx, int ny, int nz, double dx, double dy, double dz, const T *fr, const T *fi,
                            int ex, int ez, int *bcount, long long *boff, long long *total, double *xyz,
                            long long cap, hipStream_t st, bool write);

/* Fused diagnostics (ax_march.h MODE 1/2, FP64 periodic Abelian-Higgs) */
struct FusedDiag;
struct EnergyConsts;
int fused_work_items(int nx, int ny, int nz, int xseg);
hipError_t launch_step_fused(const StepConsts &k, const Levels<double> &L, int xseg, hipStream_t st, int mode,
                             bool p2d, const FusedDiag &fd);
hipError_t launch_sum3(const double *part, int nb, double *out3, hipStream_t st);
hipError_t launch_level_div(int nx, int ny, int nz, const double *tx, const double *ty, const double *tz,
                            const EnergyConsts &ek, bool pow2, double *div, hipStream_t st);
bool diag_pow2(double dx, double dy, double dz, double gg);

/* Boundary work of one TimeStep on a lattice that is not periodic
 * (ax_bound.hip).  launch_neumann_faces sets the faces of level tN from
 * their inner neighbours; launch_absorbing steps the x/y faces and corners
 * and must run before the bulk update, which overwrites the level-tP values
 * they read. */
template <typename R>
hipError_t launch_neumann_faces(const ax_params &p, void *state, size_t pitch, int tN, hipStream_t st);
template <typename R>
hipError_t launch_absorbing(const StepConsts &k, const ax_params &p, void *state, size_t pitch, int tN, int tP,
                            hipStream_t st);

}  // namespace ax

#endif
