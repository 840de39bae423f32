/*
 * ax_step_tiled.hip - Abelian-Higgs FP64 step, z-row workgroups marching
 * along x.
 *
 * Same arithmetic, operand order and results as the site kernel
 * (mpi_evolution.cpp:418-501); what changes is how often each sine and
 * cosine is evaluated.  Per site the update needs cos/sin of the three
 * links at s and at s-a, and the sines of the six plaquette angles at s and
 * of their copies at s-b (the second sine of each F term in
 * mpi_evolution.cpp:471-487 is the first one of the neighbour, written with
 * the same operands in the same order, hence the same bits):
 *   - x-backward values come from the previous iteration of the x march
 *     (kept in registers);
 *   - z-backward values come from the lane to the left (one workgroup holds a
 *     whole periodic z row, wave edges go through LDS);
 *   - y-backward values are recomputed.
 * This cuts the trig work from 18 to 12 evaluations per site, and the
 * plaquette sines take the short form of ax_sin when every lane's argument
 * reduces with k = 0 (ax_sin_is_small), which holds almost always because
 * plaquette angles stay below pi/4 while link angles drift.
 *
 * Each field value is read from HBM once per step: plane x+1 is loaded one
 * iteration ahead and carried; the y/z neighbour rows are cache hits.
 *
 * Layout assumptions (checked by use_tiled_ah): nz a multiple of 64,
 * nz <= 512 (one z row per workgroup), periodic boundaries, FP64.
 */
#include <cstdlib>

#include "ax_internal.h"

namespace ax {

constexpr int kXSeg = 32;   /* x planes marched by one workgroup */

struct TiledArgs {
  int nseg;        /* x segments */
  int work;        /* ny * nseg work items */
  int per_xcd;     /* ceil(work / 8) */
  double inv_dx2, inv_dy2, inv_dz2;  /* exact reciprocals when pow2 */
  double lim;      /* largest |d| with d/dt finite (plain-mode guard) */
};

template <bool POW2>
__device__ __forceinline__ double dvd(double a, double d, double inv) {
  if (POW2) return a * inv;   /* d is a power of two: a/d == a*(1/d) exactly */
  return a / d;
}

/* fric*((a-b)/dt) when fric is +0 (plain run): 0*(a-b) has the sign and
 * NaN-ness of the quotient unless (a-b)/dt overflows, which *ovf reports. */
__device__ __forceinline__ double zero_times_rate(double diff, double lim, bool *ovf) {
  *ovf = *ovf || !(fabs(diff) <= lim);
  return 0.0 * diff;
}

template <bool POW2, bool PLAIN>
__global__ void __launch_bounds__(512) ah_tiled_kernel(StepConsts k, Levels<double> L, TiledArgs t) {
  __shared__ double zsh[2][8][4];
  const int NX = k.nx, NY = k.ny, NZ = k.nz;
  const int nw = NZ >> 6;
  const int z = threadIdx.x;
  const int lane = z & 63, wave = z >> 6;
  /* XCD-aware mapping: dispatch puts block b on XCD b%8; give each XCD a
   * contiguous range of (segment, row) items so neighbouring rows share L2. */
  const int w = (blockIdx.x & 7) * t.per_xcd + (blockIdx.x >> 3);
  if (w >= t.work) return;
  const int y = w % NY, seg = w / NY;
  const int x0 = seg * kXSeg;
  const int x1 = min(NX, x0 + kXSeg);
  const int ym = y == 0 ? NY - 1 : y - 1, yp = y == NY - 1 ? 0 : y + 1;
  const int zm = z == 0 ? NZ - 1 : z - 1, zp = z == NZ - 1 ? 0 : z + 1;
  const int PL = NY * NZ;
  const int r0 = y * NZ, rm = ym * NZ, rp = yp * NZ;
  const double *PR = L.cur[0], *PI = L.cur[1], *TX = L.cur[2], *TY = L.cur[3], *TZ = L.cur[4];
  double *OPR = L.prv[0], *OPI = L.prv[1], *OTX = L.prv[2], *OTY = L.prv[3], *OTZ = L.prv[4];
  const double dx2 = k.dx2, dy2 = k.dy2, dz2 = k.dz2, dt = k.dt, dt2 = k.dt2;
  const double hl = k.hl, eta2 = k.eta2, m2gg = k.m2gg, fric = k.fric, damp = k.damp;

  /* ---- prologue: plane x0-1 quantities needed by plane x0 ---- */
  const int xm0 = x0 == 0 ? NX - 1 : x0 - 1;
  double PRm, PIm, cxm, sxm, Qyx_m, Qzx_m;
  {
    const int b = xm0 * PL, bn = x0 * PL;
    PRm = PR[b + r0 + z];
    PIm = PI[b + r0 + z];
    const double X = TX[b + r0 + z], Y = TY[b + r0 + z], Z = TZ[b + r0 + z];
    ax_sincos(X, &sxm, &cxm);
    const double ayx = Y + TX[b + rp + z] - TY[bn + r0 + z] - X;
    const double azx = Z + TX[b + r0 + zp] - TZ[bn + r0 + z] - X;
    Qyx_m = ax_sin(ayx);
    Qzx_m = ax_sin(azx);
  }
  double PRc = PR[x0 * PL + r0 + z], PIc = PI[x0 * PL + r0 + z];
  double TXc = TX[x0 * PL + r0 + z], TYc = TY[x0 * PL + r0 + z], TZc = TZ[x0 * PL + r0 + z];

  for (int x = x0; x < x1; ++x) {
    const int xp = x == NX - 1 ? 0 : x + 1;
    const int b = x * PL, bn = xp * PL;
    const int s = b + r0 + z;
    /* next plane (HBM) and previous level at s (HBM) */
    const double PRn = PR[bn + r0 + z], PIn = PI[bn + r0 + z];
    const double TXn = TX[bn + r0 + z], TYn = TY[bn + r0 + z], TZn = TZ[bn + r0 + z];
    const double oPR = OPR[s], oPI = OPI[s], oX = OTX[s], oY = OTY[s], oZ = OTZ[s];
    /* neighbour rows and columns of plane x (cache) */
    const double PRyp = PR[b + rp + z], PIyp = PI[b + rp + z], PRym = PR[b + rm + z], PIym = PI[b + rm + z];
    const double PRzp = PR[b + r0 + zp], PIzp = PI[b + r0 + zp], PRzm = PR[b + r0 + zm], PIzm = PI[b + r0 + zm];
    const double TXyp = TX[b + rp + z], TZyp = TZ[b + rp + z];
    const double TXzp = TX[b + r0 + zp], TYzp = TY[b + r0 + zp];
    const double TXym = TX[b + rm + z], TYym = TY[b + rm + z], TZym = TZ[b + rm + z];
    const double TYxpym = TY[bn + rm + z], TYymzp = TY[b + rm + zp];

    /* links at s and s-y */
    double cx, sx, cy, sy, cz, sz, cym, sym_;
    ax_sincos(TXc, &sx, &cx);
    ax_sincos(TYc, &sy, &cy);
    ax_sincos(TZc, &sz, &cz);
    ax_sincos(TYym, &sym_, &cym);
    /* plaquette angles, reference operand order */
    const double a_xy = TXc + TYn - TXyp - TYc;
    const double a_xz = TXc + TZn - TXzp - TZc;
    const double a_yx = TYc + TXyp - TYn - TXc;
    const double a_yz = TYc + TZyp - TYzp - TZc;
    const double a_zx = TZc + TXzp - TZn - TXc;
    const double a_zy = TZc + TYzp - TZyp - TYc;
    const double b_xy = TXym + TYxpym - TXc - TYym;
    const double b_zy = TZym + TYymzp - TZc - TYym;
    double Qxy, Qxz, Qyx, Qyz, Qzx, Qzy, Qxy_m, Qzy_m;
    const bool small = ax_sin_is_small(a_xy) && ax_sin_is_small(a_xz) && ax_sin_is_small(a_yx) &&
                       ax_sin_is_small(a_yz) && ax_sin_is_small(a_zx) && ax_sin_is_small(a_zy) &&
                       ax_sin_is_small(b_xy) && ax_sin_is_small(b_zy);
    if (__all(small)) {
      Qxy = ax_sin_small(a_xy); Qxz = ax_sin_small(a_xz); Qyx = ax_sin_small(a_yx);
      Qyz = ax_sin_small(a_yz); Qzx = ax_sin_small(a_zx); Qzy = ax_sin_small(a_zy);
      Qxy_m = ax_sin_small(b_xy); Qzy_m = ax_sin_small(b_zy);
    } else {
      Qxy = ax_sin(a_xy); Qxz = ax_sin(a_xz); Qyx = ax_sin(a_yx);
      Qyz = ax_sin(a_yz); Qzx = ax_sin(a_zx); Qzy = ax_sin(a_zy);
      Qxy_m = ax_sin(b_xy); Qzy_m = ax_sin(b_zy);
    }
    /* z-backward values: lane to the left; wave edge through LDS */
    double czm = __shfl_up(cz, 1, 64), szm = __shfl_up(sz, 1, 64);
    double Qxz_m = __shfl_up(Qxz, 1, 64), Qyz_m = __shfl_up(Qyz, 1, 64);
    const int buf = x & 1;
    if (lane == 63) {
      zsh[buf][wave][0] = cz; zsh[buf][wave][1] = sz;
      zsh[buf][wave][2] = Qxz; zsh[buf][wave][3] = Qyz;
    }
    __syncthreads();
    if (lane == 0) {
      const int src = wave == 0 ? nw - 1 : wave - 1;
      czm = zsh[buf][src][0]; szm = zsh[buf][src][1];
      Qxz_m = zsh[buf][src][2]; Qyz_m = zsh[buf][src][3];
    }

    /* scalar field */
    const double mag2 = PRc * PRc + PIc * PIc;
    const double pot = hl * (mag2 - eta2);
    bool ovf = false;
    double f0, f1;
    if (PLAIN) {
      f0 = zero_times_rate(PRc - oPR, t.lim, &ovf);
      f1 = zero_times_rate(PIc - oPI, t.lim, &ovf);
    } else {
      f0 = fric * ((PRc - oPR) / dt);
      f1 = fric * ((PIc - oPI) / dt);
    }
    const double lxx0 = dvd<POW2>(cx * PRn + sx * PIn - 2 * PRc + cxm * PRm - sxm * PIm, dx2, t.inv_dx2);
    const double lyy0 = dvd<POW2>(cy * PRyp + sy * PIyp - 2 * PRc + cym * PRym - sym_ * PIym, dy2, t.inv_dy2);
    const double lzz0 = dvd<POW2>(cz * PRzp + sz * PIzp - 2 * PRc + czm * PRzm - szm * PIzm, dz2, t.inv_dz2);
    const double lxx1 = dvd<POW2>(cx * PIn + (-sx) * PRn - 2 * PIc + cxm * PIm - (-sxm) * PRm, dx2, t.inv_dx2);
    const double lyy1 = dvd<POW2>(cy * PIyp + (-sy) * PRyp - 2 * PIc + cym * PIym - (-sym_) * PRym, dy2, t.inv_dy2);
    const double lzz1 = dvd<POW2>(cz * PIzp + (-sz) * PRzp - 2 * PIc + czm * PIzm - (-szm) * PRzm, dz2, t.inv_dz2);
    /* gauge field */
    const double jx = cx * (PIc * PRn - PRc * PIn) + sx * (PRc * PRn + PIc * PIn);
    const double jy = cy * (PIc * PRyp - PRc * PIyp) + sy * (PRc * PRyp + PIc * PIyp);
    const double jz = cz * (PIc * PRzp - PRc * PIzp) + sz * (PRc * PRzp + PIc * PIzp);
    const double fxy = dvd<POW2>(Qxy - Qxy_m, dy2, t.inv_dy2);
    const double fxz = dvd<POW2>(Qxz - Qxz_m, dz2, t.inv_dz2);
    const double fyx = dvd<POW2>(Qyx - Qyx_m, dx2, t.inv_dx2);
    const double fyz = dvd<POW2>(Qyz - Qyz_m, dz2, t.inv_dz2);
    const double fzx = dvd<POW2>(Qzx - Qzx_m, dx2, t.inv_dx2);
    const double fzy = dvd<POW2>(Qzy - Qzy_m, dy2, t.inv_dy2);
    double gX, gY, gZ;
    if (PLAIN) {
      gX = zero_times_rate(TXc - oX, t.lim, &ovf);
      gY = zero_times_rate(TYc - oY, t.lim, &ovf);
      gZ = zero_times_rate(TZc - oZ, t.lim, &ovf);
    } else if (k.theta_damp_form == 0) {
      gX = damp * ((TXc - oX) / dt);
      gY = damp * ((TYc - oY) / dt);
      gZ = damp * ((TZc - oZ) / dt);
    } else {
      gX = damp * (TXc - oX) / dt;
      gY = damp * (TYc - oY) / dt;
      gZ = damp * (TZc - oZ) / dt;
    }
    if (PLAIN && __any(ovf)) {
      /* some |a-b|/dt overflows: take the literal products (never in a sane state) */
      f0 = fric * ((PRc - oPR) / dt);
      f1 = fric * ((PIc - oPI) / dt);
      if (k.theta_damp_form == 0) {
        gX = damp * ((TXc - oX) / dt); gY = damp * ((TYc - oY) / dt); gZ = damp * ((TZc - oZ) / dt);
      } else {
        gX = damp * (TXc - oX) / dt; gY = damp * (TYc - oY) / dt; gZ = damp * (TZc - oZ) / dt;
      }
    }
    const double acc0 = lxx0 + lyy0 + lzz0 - pot * PRc - f0;
    const double acc1 = lxx1 + lyy1 + lzz1 - pot * PIc - f1;
    OPR[s] = 2 * PRc - oPR + dt2 * acc0;
    OPI[s] = 2 * PIc - oPI + dt2 * acc1;
    const double ax_ = m2gg * jx - fxy - fxz - gX;
    const double ay_ = m2gg * jy - fyx - fyz - gY;
    const double az_ = m2gg * jz - fzx - fzy - gZ;
    OTX[s] = 2 * TXc - oX + dt2 * ax_;
    OTY[s] = 2 * TYc - oY + dt2 * ay_;
    OTZ[s] = 2 * TZc - oZ + dt2 * az_;

    /* carry to x+1 */
    PRm = PRc; PIm = PIc;
    cxm = cx; sxm = sx; Qyx_m = Qyx; Qzx_m = Qzx;
    PRc = PRn; PIc = PIn; TXc = TXn; TYc = TYn; TZc = TZn;
  }
}

static bool is_pow2(double v) {
  int e;
  double m = frexp(v, &e);
  return v > 0 && m == 0.5;
}

bool use_tiled_ah(const ax_params &p) {
  return p.model == AX_MODEL_AH && p.precision == 64 && p.boundary == AX_BC_PERIODIC && p.nz % 64 == 0 &&
         p.nz <= 512 && p.nx >= 3 && p.ny >= 3 && getenv("AX_FORCE_SITE_KERNEL") == nullptr;
}

/* largest d with d/dt finite */
static double rate_limit(double dt) {
  double d = 1.7976931348623157e308 * (dt < 1 ? dt : 1.0);
  while (!(d / dt <= 1.7976931348623157e308)) d = nextafter(d, 0.0);
  for (;;) {
    double n = nextafter(d, 2e308);
    if (!(n / dt <= 1.7976931348623157e308) || n > 1.7976931348623157e308) break;
    d = n;
  }
  return d;
}

hipError_t launch_step_tiled_ah(const StepConsts &k, const Levels<double> &L, hipStream_t st) {
  TiledArgs t;
  t.nseg = (k.nx + kXSeg - 1) / kXSeg;
  t.work = k.ny * t.nseg;
  t.per_xcd = (t.work + 7) / 8;
  const bool pow2 = is_pow2(k.dx2) && is_pow2(k.dy2) && is_pow2(k.dz2);
  t.inv_dx2 = 1.0 / k.dx2;
  t.inv_dy2 = 1.0 / k.dy2;
  t.inv_dz2 = 1.0 / k.dz2;
  t.lim = rate_limit(k.dt);
  /* plain: both friction coefficients are +0 for this step */
  const bool plain = k.fric == 0.0 && !signbit(k.fric) && k.damp == 0.0 && !signbit(k.damp);
  dim3 grid(8 * t.per_xcd), block(k.nz);
  if (pow2 && plain)
    hipLaunchKernelGGL((ah_tiled_kernel<true, true>), grid, block, 0, st, k, L, t);
  else if (pow2)
    hipLaunchKernelGGL((ah_tiled_kernel<true, false>), grid, block, 0, st, k, L, t);
  else if (plain)
    hipLaunchKernelGGL((ah_tiled_kernel<false, true>), grid, block, 0, st, k, L, t);
  else
    hipLaunchKernelGGL((ah_tiled_kernel<false, false>), grid, block, 0, st, k, L, t);
  return hipGetLastError();
}

}  // namespace ax
