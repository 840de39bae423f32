/*
 * ax_step_tiled.hip - Abelian-Higgs and Moore steps of one lattice, one
 * 64-lane wave per (64-site z chunk, y row, x segment) marching along x (the
 * updates themselves are ah_march in ax_march.h and moore_march in
 * ax_march_moore.h).  FP64 and FP32, every boundary (boundary planes and rows
 * are left to the boundary kernels).
 *
 * Work items are numbered in the XCD-aware order of make_tile_grid
 * (ax_internal.h): each XCD takes bands of consecutive rows with all their
 * chunks, so y-neighbour rows and chunk-edge lines are read from its own L2.
 */
#include <cstdlib>
#include <limits>

#include "ax_internal.h"
#include "ax_march.h"
#include "ax_march_moore.h"

namespace ax {

/* The planes [x0, x1) and row y a work item steps; with non-periodic x/y
 * (ex) the boundary planes and rows are left alone (their neighbours would
 * wrap), evolution.cpp:161-186.  Returns false for an empty item. */
__device__ __forceinline__ bool tile_range(const TileGrid &g, int nx, int ny, int ex, int w, int &zc, int &y,
                                           int &x0, int &x1) {
  int seg;
  tile_item(g, ny, w, zc, y, seg);
  x0 = seg * g.xseg;
  x1 = min(nx, x0 + g.xseg);
  if (ex) {
    x0 = max(x0, 1);
    x1 = min(x1, nx - 1);
    if (y == 0 || y == ny - 1) return false;
  }
  return x0 < x1;
}

template <typename R, bool POW2, bool PLAIN, bool BIG, bool PF = false>
__global__ void __launch_bounds__(64) ah_tiled_kernel(MarchConsts<R> k, Levels<R> L, int nx, int ny, int nz,
                                                      TileGrid g, int ex, int ez) {
  const int w = (int)blockIdx.x;
  if (w >= g.work) return;
  int zc, y, x0, x1;
  if (!tile_range(g, nx, ny, ex, w, zc, y, x0, x1)) return;
  ah_march<R, POW2, PLAIN ? 1 : 0, BIG, PF>(k, L, nx, ny, nz, y, zc, x0, x1, ez != 0);
}

/* FD: FP64 with divisors in the FixedDivisor range on a periodic lattice.
 * band[3] are the per-level band flags (ax_march_moore.h), indexed by
 * TimeStep mod 3: this step (TimeStep t) divides through FixedDivisor when
 * use_band is set and band[t % 3] is clear, ORs its output's check into
 * band[(t + 1) % 3] and clears band[(t + 2) % 3] for the next step. */
template <typename R, bool POW2, bool BUF, bool FD>
__global__ void __launch_bounds__(64, 3) moore_tiled_kernel(MooreConsts<R> k, Levels<R> L, int nx, int ny, int nz,
                                                         TileGrid g, int ex, int ez, int *band, int t,
                                                         int use_band) {
  const int w = (int)blockIdx.x;
  if (FD && w == 0 && threadIdx.x == 0) band[(t + 2) % 3] = 0;
  if (w >= g.work) return;
  int zc, y, x0, x1;
  if (!tile_range(g, nx, ny, ex, w, zc, y, x0, x1)) return;
  if constexpr (FD) {
    int *out = band + (t + 1) % 3;
    if (use_band && band[t % 3] == 0)
      moore_march<R, POW2, BUF, true, true>(k, L, nx, ny, nz, y, zc, x0, x1, ez != 0, out);
    else
      moore_march<R, POW2, BUF, false, true>(k, L, nx, ny, nz, y, zc, x0, x1, ez != 0, out);
  } else {
    moore_march<R, POW2, BUF>(k, L, nx, ny, nz, y, zc, x0, x1, ez != 0);
  }
}

template <typename R>
hipError_t launch_step_tiled_moore(const StepConsts &k, const Levels<R> &L, int xseg, hipStream_t st, int *band,
                                   int t, bool use_band, bool *fd_used) {
  const TileGrid g = make_tile_grid(k.nx, k.ny, k.nz, xseg);
  MooreConsts<R> m;
  moore_consts_from<R>(k, m);
  const bool pow2 = is_pow2(k.dx2) && is_pow2(k.dy2) && is_pow2(k.dz2);
  /* buffer path when both levels of a field lie within 4 GiB of its lower plane */
  const size_t lev = (size_t)std::llabs((const char *)L.cur[0] - (const char *)L.prv[0]);
  const bool buf = lev + (size_t)k.nx * k.ny * k.nz * sizeof(R) < ((size_t)1 << 32) &&
                   getenv("AX_BIG_PLANES") == nullptr;   /* forces the global path, for tests */
  /* FP64 divisions by dx^2, dy^2, dz^2 through FixedDivisor, guarded by the
   * band flags (every site written by this kernel: periodic lattices only);
   * with the buffer path only: on the global path (planes of 4 GiB or
   * more) it spills 67 VGPRs at 3 waves per SIMD and is no faster than the
   * site kernel, the plain divisions spill none */
  const bool fd = sizeof(R) == 8 && buf && !pow2 && band != nullptr && !k.ex && !k.ez && fixed_divisor_range(k.dx2) &&
                  fixed_divisor_range(k.dy2) && fixed_divisor_range(k.dz2) && moore_band_charge((int)k.q1) &&
                  moore_band_charge((int)k.q2) && getenv("AX_MOORE_PLAIN_DIV") == nullptr;
  *fd_used = fd;
  dim3 grid(g.work), block(64);
#define AX_MRL(P2, B, FD)                                                                                     \
  hipLaunchKernelGGL((moore_tiled_kernel<R, P2, B, FD>), grid, block, 0, st, m, L, k.nx, k.ny, k.nz, g, k.ex, k.ez, \
                     band, t, use_band ? 1 : 0)
  if (pow2 && buf) AX_MRL(true, true, false);
  else if (pow2) AX_MRL(true, false, false);
  else if (buf && fd) AX_MRL(false, true, true);
  else if (buf) AX_MRL(false, true, false);
  else if (fd) AX_MRL(false, false, true);
  else AX_MRL(false, false, false);
#undef AX_MRL
  return hipGetLastError();
}

template hipError_t launch_step_tiled_moore<double>(const StepConsts &, const Levels<double> &, int, hipStream_t,
                                                    int *, int, bool, bool *);
template hipError_t launch_step_tiled_moore<float>(const StepConsts &, const Levels<float> &, int, hipStream_t, int *,
                                                   int, bool, bool *);

/* Every Abelian-Higgs and Moore lattice marches, whatever its z tail (a
 * last chunk of 1..63 sites runs with its other lanes idle) and whatever
 * its plane size (planes of 4 GiB or more take the 64-bit plane addressing
 * of ah_march and moore_march). */
bool use_tiled_ah(const ax_params &p) {
  return (p.model == AX_MODEL_AH || p.model == AX_MODEL_MOORE) && p.nx >= 3 && p.ny >= 3 &&
         getenv("AX_FORCE_SITE_KERNEL") == nullptr;
}

/* x planes per wave (measured at 512^3 AH, 256^3 Moore, on one box each):
 * Abelian-Higgs FP64 32 (3.875 ms; 8: 3.95, 16: 3.90, 24: 3.88), FP32 16
 * (1.816 ms; 8: 1.834, 12: 1.823, 32: 1.826); Moore FP64 16 (0.963 ms; 4:
 * 0.993, 8: 0.979, 32: 0.974), FP32 8 (0.422 ms; 32: 0.444) */
int default_xseg(const ax_params &p) {
  const char *xe = getenv("AX_TILED_XSEG");
  if (xe) return atoi(xe);
  if (p.model == AX_MODEL_MOORE) return p.precision == 32 ? 8 : 16;
  return p.precision == 32 ? 16 : 32;
}

/* The LDS prefetch of the FP64 march applies: whole 64-site chunks, three
 * field planes within 32-bit offsets, and the [comp][level] plane order its
 * phi and theta_x/theta_y pairs rely on (AX_NO_PREFETCH turns it off). */
template <typename R>
static bool prefetch_ok(const StepConsts &k, const Levels<R> &L) {
  const bool order = L.cur[0] < L.cur[1] && L.cur[2] < L.cur[3] && L.prv[0] < L.prv[1] && L.prv[2] < L.prv[3];
  return k.nz % 64 == 0 && order &&
         (size_t)k.nx * k.ny * k.nz * sizeof(double) * 3 + ((size_t)1 << 20) < ((size_t)1 << 32) &&
         getenv("AX_NO_PREFETCH") == nullptr;
}

template <typename R>
hipError_t launch_step_tiled_ah(const StepConsts &k, const Levels<R> &L, int xseg, hipStream_t st) {
  const TileGrid g = make_tile_grid(k.nx, k.ny, k.nz, xseg);
  MarchConsts<R> m;
  march_consts_from<R>(k, rate_limit<R>((R)k.dt), m);
  const bool pow2 = is_pow2(k.dx2) && is_pow2(k.dy2) && is_pow2(k.dz2);
  /* a field plane of 4 GiB or more: 64-bit plane addressing (AX_BIG_PLANES
   * forces it on smaller lattices, for tests) */
  const bool big = (size_t)k.nx * k.ny * k.nz * sizeof(R) >= ((size_t)1 << 32) || getenv("AX_BIG_PLANES") != nullptr;
  dim3 grid(g.work), block(64);
#define AX_AHL(P2, PL, BG) \
  hipLaunchKernelGGL((ah_tiled_kernel<R, P2, PL, BG>), grid, block, 0, st, m, L, k.nx, k.ny, k.nz, g, k.ex, k.ez)
  /* FP64, whole chunks and three planes within 32-bit offsets: the
   * previous level and the next plane are prefetched into LDS (ah_march; its
   * phi and theta_x/theta_y pairs rely on the [comp][level] plane order).
   * Every variant keeps 3 waves per SIMD with it (161 VGPRs, with the stores
   * issued one plane late). */
  const bool pf = sizeof(R) == 8 && !big && prefetch_ok(k, L);
#define AX_AHPF(P2, PL)                                                                                     \
  hipLaunchKernelGGL((ah_tiled_kernel<double, P2, PL, false, true>), grid, block, 0, st, m, L, k.nx, k.ny, k.nz, g, \
                     k.ex, k.ez)
  if (pf) {
    if constexpr (sizeof(R) == 8) {
      if (pow2 && m.plain) AX_AHPF(true, true);
      else if (pow2) AX_AHPF(true, false);
      else if (m.plain) AX_AHPF(false, true);
      else AX_AHPF(false, false);
    }
  } else if (big) {
    if (pow2 && m.plain) AX_AHL(true, true, true);
    else if (pow2) AX_AHL(true, false, true);
    else if (m.plain) AX_AHL(false, true, true);
    else AX_AHL(false, false, true);
  } else {
    if (pow2 && m.plain) AX_AHL(true, true, false);
    else if (pow2) AX_AHL(true, false, false);
    else if (m.plain) AX_AHL(false, true, false);
    else AX_AHL(false, false, false);
  }
#undef AX_AHL
#undef AX_AHPF
  return hipGetLastError();
}

/* The FP64 periodic Abelian-Higgs step with its current level's divergence
 * stored (MODE 1) or its current level's diagnostics taken (MODE 2), see
 * ah_march.  One partial-sum triple per work item (MODE 2).  PF: the LDS
 * prefetch of the plain step (MODE 1 keeps 3 waves per SIMD with it, 167
 * VGPRs; MODE 2 runs at 2 either way). */
template <bool POW2, bool P2D, int MODE, bool PF>
__global__ void __launch_bounds__(64) ah_fused_kernel(MarchConsts<double> k, Levels<double> L, int nx, int ny, int nz,
                                                      TileGrid g, FusedDiag fd) {
  const int w = (int)blockIdx.x;
  if (w >= g.work) return;
  int zc, y, x0, x1;
  if (!tile_range(g, nx, ny, 0, w, zc, y, x0, x1)) return;
  ah_march<double, POW2, 2, false, PF, MODE, P2D>(k, L, nx, ny, nz, y, zc, x0, x1, false, fd, w);
}

int fused_work_items(int nx, int ny, int nz, int xseg) { return make_tile_grid(nx, ny, nz, xseg).work; }

hipError_t launch_step_fused(const StepConsts &k, const Levels<double> &L, int xseg, hipStream_t st, int mode,
                             bool p2d, const FusedDiag &fd) {
  const TileGrid g = make_tile_grid(k.nx, k.ny, k.nz, xseg);
  MarchConsts<double> m;
  march_consts_from<double>(k, rate_limit<double>(k.dt), m);
  const bool pow2 = is_pow2(k.dx2) && is_pow2(k.dy2) && is_pow2(k.dz2);
  /* the plain step's prefetch conditions (launch_step_tiled_ah) */
  const bool pf = prefetch_ok(k, L);
  dim3 grid(g.work), block(64);
#define AX_FL(P2, PD, MD)                                                                                      \
  do {                                                                                                         \
    if (pf) hipLaunchKernelGGL((ah_fused_kernel<P2, PD, MD, true>), grid, block, 0, st, m, L, k.nx, k.ny, k.nz, g, fd); \
    else hipLaunchKernelGGL((ah_fused_kernel<P2, PD, MD, false>), grid, block, 0, st, m, L, k.nx, k.ny, k.nz, g, fd); \
  } while (0)
  if (mode == 1) {
    if (pow2) AX_FL(true, false, 1);
    else AX_FL(false, false, 1);
  } else {
    if (pow2 && p2d) AX_FL(true, true, 2);
    else if (pow2) AX_FL(true, false, 2);
    else AX_FL(false, false, 2);
  }
#undef AX_FL
  return hipGetLastError();
}

template hipError_t launch_step_tiled_ah<double>(const StepConsts &, const Levels<double> &, int, hipStream_t);
template hipError_t launch_step_tiled_ah<float>(const StepConsts &, const Levels<float> &, int, hipStream_t);

}  // namespace ax
