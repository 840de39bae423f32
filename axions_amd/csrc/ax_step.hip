/*
 * ax_step.hip - site-parallel lattice update kernels (one thread per site).
 *
 * These kernels evaluate the reference's per-site expressions with its exact
 * operand order (mpi_evolution.cpp:418-501 for Abelian-Higgs,
 * mooreEvol.cpp:202-337 for Moore) and the in-place two-level update, so the
 * FP64 instantiation is bit-identical to the reference given the shared
 * sin/cos of ax_trig.h.  Periodic neighbours are formed in the index math
 * (no halo).  Each distinct trig argument is evaluated once per thread
 * (ax_sin(x) is by construction the sine half of ax_sincos(x)).
 *
 * They are also the generic path for any lattice shape; the tiled
 * Abelian-Higgs FP64 kernel in ax_step_tiled.hip is used where it applies.
 */
#include "ax_device.h"

namespace ax {

template <typename R>
struct Trig;
template <>
struct Trig<double> {
  __device__ static void sc(double x, double *s, double *c) { ax_sincos(x, s, c); }
  __device__ static double sn(double x) { return ax_sin(x); }
};
template <>
struct Trig<float> {
  __device__ static void sc(float x, float *s, float *c) { ax_sincosf(x, s, c); }
  __device__ static float sn(float x) {
    float s, c;
    ax_sincosf(x, &s, &c);
    return s;
  }
};

template <typename T>
__device__ __forceinline__ T val(const T *p, int i) { return p[i]; }

struct Site {
  int s, xp, xm, yp, ym, zp, zm, xp_ym, xp_zm, xm_yp, yp_zm, xm_zp, ym_zp;
};

__device__ __forceinline__ bool site_of(const StepConsts &k, int s, Site &o, int &i, int &j, int &kk) {
  const int nz = k.nz, ny = k.ny, nx = k.nx;
  kk = s % nz;
  int r = s / nz;
  j = r % ny;
  i = r / ny;
  if ((k.ex && (i == 0 || i == nx - 1 || j == 0 || j == ny - 1)) || (k.ez && (kk == 0 || kk == nz - 1)))
    return false;
  const int ip = wrapp(i, nx), im = wrapm(i, nx), jp = wrapp(j, ny), jm = wrapm(j, ny);
  const int kp = wrapp(kk, nz), km = wrapm(kk, nz);
  auto id = [&](int a, int b, int c) { return (a * ny + b) * nz + c; };
  o.s = s;
  o.xp = id(ip, j, kk); o.xm = id(im, j, kk);
  o.yp = id(i, jp, kk); o.ym = id(i, jm, kk);
  o.zp = id(i, j, kp); o.zm = id(i, j, km);
  o.xp_ym = id(ip, jm, kk); o.xp_zm = id(ip, j, km);
  o.xm_yp = id(im, jp, kk); o.yp_zm = id(i, jp, km);
  o.xm_zp = id(im, j, kp); o.ym_zp = id(i, jm, kp);
  return true;
}

/* Abelian-Higgs fused step at one site, arithmetic type R = storage type. */
template <typename R>
__device__ __forceinline__ void ah_update(const StepConsts &k, const Levels<R> &L, const Site &n) {
  const R *pr = L.cur[0], *pi = L.cur[1], *tx = L.cur[2], *ty = L.cur[3], *tz = L.cur[4];
  const int s = n.s;
  const R dx2 = (R)k.dx2, dy2 = (R)k.dy2, dz2 = (R)k.dz2, dt = (R)k.dt, dt2 = (R)k.dt2;
  /* links at s and s-a */
  R sx, cx, sy, cy, sz, cz, sxm, cxm, sym_, cym, szm, czm;
  Trig<R>::sc(tx[s], &sx, &cx);
  Trig<R>::sc(ty[s], &sy, &cy);
  Trig<R>::sc(tz[s], &sz, &cz);
  Trig<R>::sc(tx[n.xm], &sxm, &cxm);
  Trig<R>::sc(ty[n.ym], &sym_, &cym);
  Trig<R>::sc(tz[n.zm], &szm, &czm);
  const R ar = pr[s], ai = pi[s];
  const R mag2 = ar * ar + ai * ai;
  const R pot = (R)k.hl * (mag2 - (R)k.eta2);
  const R fric = (R)k.fric;
  /* component 0 (Re): sign +1 on the sine terms; component 1 (Im): -1 */
  {
    R lxx = (cx * pr[n.xp] + sx * pi[n.xp] - 2 * ar + cxm * pr[n.xm] - sxm * pi[n.xm]) / dx2;
    R lyy = (cy * pr[n.yp] + sy * pi[n.yp] - 2 * ar + cym * pr[n.ym] - sym_ * pi[n.ym]) / dy2;
    R lzz = (cz * pr[n.zp] + sz * pi[n.zp] - 2 * ar + czm * pr[n.zm] - szm * pi[n.zm]) / dz2;
    R old = L.prv[0][s];
    R vel = (ar - old) / dt;
    R acc = lxx + lyy + lzz - pot * ar - fric * vel;
    L.prv[0][s] = 2 * ar - old + dt2 * acc;
  }
  {
    R lxx = (cx * pi[n.xp] + (-sx) * pr[n.xp] - 2 * ai + cxm * pi[n.xm] - (-sxm) * pr[n.xm]) / dx2;
    R lyy = (cy * pi[n.yp] + (-sy) * pr[n.yp] - 2 * ai + cym * pi[n.ym] - (-sym_) * pr[n.ym]) / dy2;
    R lzz = (cz * pi[n.zp] + (-sz) * pr[n.zp] - 2 * ai + czm * pi[n.zm] - (-szm) * pr[n.zm]) / dz2;
    R old = L.prv[1][s];
    R vel = (ai - old) / dt;
    R acc = lxx + lyy + lzz - pot * ai - fric * vel;
    L.prv[1][s] = 2 * ai - old + dt2 * acc;
  }
  R jx = cx * (ai * pr[n.xp] - ar * pi[n.xp]) + sx * (ar * pr[n.xp] + ai * pi[n.xp]);
  R jy = cy * (ai * pr[n.yp] - ar * pi[n.yp]) + sy * (ar * pr[n.yp] + ai * pi[n.yp]);
  R jz = cz * (ai * pr[n.zp] - ar * pi[n.zp]) + sz * (ar * pr[n.zp] + ai * pi[n.zp]);
  const R X = tx[s], Y = ty[s], Z = tz[s];
  R fxy = (Trig<R>::sn(X + ty[n.xp] - tx[n.yp] - Y) - Trig<R>::sn(tx[n.ym] + ty[n.xp_ym] - X - ty[n.ym])) / dy2;
  R fxz = (Trig<R>::sn(X + tz[n.xp] - tx[n.zp] - Z) - Trig<R>::sn(tx[n.zm] + tz[n.xp_zm] - X - tz[n.zm])) / dz2;
  R fyx = (Trig<R>::sn(Y + tx[n.yp] - ty[n.xp] - X) - Trig<R>::sn(ty[n.xm] + tx[n.xm_yp] - Y - tx[n.xm])) / dx2;
  R fyz = (Trig<R>::sn(Y + tz[n.yp] - ty[n.zp] - Z) - Trig<R>::sn(ty[n.zm] + tz[n.yp_zm] - Y - tz[n.zm])) / dz2;
  R fzx = (Trig<R>::sn(Z + tx[n.zp] - tz[n.xp] - X) - Trig<R>::sn(tz[n.xm] + tx[n.xm_zp] - Z - tx[n.xm])) / dx2;
  R fzy = (Trig<R>::sn(Z + ty[n.zp] - tz[n.yp] - Y) - Trig<R>::sn(tz[n.ym] + ty[n.ym_zp] - Z - ty[n.ym])) / dy2;
  const R m2gg = (R)k.m2gg, damp = (R)k.damp;
  R oX = L.prv[2][s], oY = L.prv[3][s], oZ = L.prv[4][s];
  R ax_, ay_, az_;
  if (k.theta_damp_form == 0) {
    ax_ = m2gg * jx - fxy - fxz - damp * ((X - oX) / dt);
    ay_ = m2gg * jy - fyx - fyz - damp * ((Y - oY) / dt);
    az_ = m2gg * jz - fzx - fzy - damp * ((Z - oZ) / dt);
  } else {
    ax_ = m2gg * jx - fxy - fxz - damp * (X - oX) / dt;
    ay_ = m2gg * jy - fyx - fyz - damp * (Y - oY) / dt;
    az_ = m2gg * jz - fzx - fzy - damp * (Z - oZ) / dt;
  }
  L.prv[2][s] = 2 * X - oX + dt2 * ax_;
  L.prv[3][s] = 2 * Y - oY + dt2 * ay_;
  L.prv[4][s] = 2 * Z - oZ + dt2 * az_;
}

/* Moore fused step at one site (mooreEvol.cpp:202-337, no friction). */
template <typename R>
__device__ __forceinline__ void moore_update(const StepConsts &k, const Levels<R> &L, const Site &n) {
  const R *tx = L.cur[4], *ty = L.cur[5], *tz = L.cur[6];
  const int s = n.s;
  const R dx2 = (R)k.dx2, dy2 = (R)k.dy2, dz2 = (R)k.dz2, dt2 = (R)k.dt2;
  const R q[2] = {(R)k.q1, (R)k.q2};
  R acc[2][2], cur[2][3];
  const R m1 = L.cur[0][s] * L.cur[0][s] + L.cur[1][s] * L.cur[1][s];
  const R m2 = L.cur[2][s] * L.cur[2][s] + L.cur[3][s] * L.cur[3][s];
  const R pot[2] = {(R)k.hl1 * (m1 - (R)k.eta1sq) + (R)k.hl12 * (m2 - (R)k.eta2sq),
                    (R)k.hl2 * (m2 - (R)k.eta2sq) + (R)k.hl12 * (m1 - (R)k.eta1sq)};
#pragma unroll
  for (int f = 0; f < 2; ++f) {
    const R *re = L.cur[2 * f], *im = L.cur[2 * f + 1];
    R sx, cx, sy, cy, sz, cz, sxm, cxm, sym_, cym, szm, czm;
    Trig<R>::sc(q[f] * tx[s], &sx, &cx);
    Trig<R>::sc(q[f] * ty[s], &sy, &cy);
    Trig<R>::sc(q[f] * tz[s], &sz, &cz);
    Trig<R>::sc(q[f] * tx[n.xm], &sxm, &cxm);
    Trig<R>::sc(q[f] * ty[n.ym], &sym_, &cym);
    Trig<R>::sc(q[f] * tz[n.zm], &szm, &czm);
    const R ar = re[s], ai = im[s];
    {
      R lxx = (cx * re[n.xp] + sx * im[n.xp] - 2 * ar + cxm * re[n.xm] - sxm * im[n.xm]) / dx2;
      R lyy = (cy * re[n.yp] + sy * im[n.yp] - 2 * ar + cym * re[n.ym] - sym_ * im[n.ym]) / dy2;
      R lzz = (cz * re[n.zp] + sz * im[n.zp] - 2 * ar + czm * re[n.zm] - szm * im[n.zm]) / dz2;
      acc[f][0] = lxx + lyy + lzz - pot[f] * ar;
    }
    {
      R lxx = (cx * im[n.xp] + (-sx) * re[n.xp] - 2 * ai + cxm * im[n.xm] - (-sxm) * re[n.xm]) / dx2;
      R lyy = (cy * im[n.yp] + (-sy) * re[n.yp] - 2 * ai + cym * im[n.ym] - (-sym_) * re[n.ym]) / dy2;
      R lzz = (cz * im[n.zp] + (-sz) * re[n.zp] - 2 * ai + czm * im[n.zm] - (-szm) * re[n.zm]) / dz2;
      acc[f][1] = lxx + lyy + lzz - pot[f] * ai;
    }
    cur[f][0] = q[f] * (cx * (ai * re[n.xp] - ar * im[n.xp]) + sx * (ar * re[n.xp] + ai * im[n.xp]));
    cur[f][1] = q[f] * (cy * (ai * re[n.yp] - ar * im[n.yp]) + sy * (ar * re[n.yp] + ai * im[n.yp]));
    cur[f][2] = q[f] * (cz * (ai * re[n.zp] - ar * im[n.zp]) + sz * (ar * re[n.zp] + ai * im[n.zp]));
  }
  const R X = tx[s], Y = ty[s], Z = tz[s];
  R fxy = (Trig<R>::sn(X + ty[n.xp] - tx[n.yp] - Y) - Trig<R>::sn(tx[n.ym] + ty[n.xp_ym] - X - ty[n.ym])) / dy2;
  R fxz = (Trig<R>::sn(X + tz[n.xp] - tx[n.zp] - Z) - Trig<R>::sn(tx[n.zm] + tz[n.xp_zm] - X - tz[n.zm])) / dz2;
  R fyx = (Trig<R>::sn(Y + tx[n.yp] - ty[n.xp] - X) - Trig<R>::sn(ty[n.xm] + tx[n.xm_yp] - Y - tx[n.xm])) / dx2;
  R fyz = (Trig<R>::sn(Y + tz[n.yp] - ty[n.zp] - Z) - Trig<R>::sn(ty[n.zm] + tz[n.yp_zm] - Y - tz[n.zm])) / dz2;
  R fzx = (Trig<R>::sn(Z + tx[n.zp] - tz[n.xp] - X) - Trig<R>::sn(tz[n.xm] + tx[n.xm_zp] - Z - tx[n.xm])) / dx2;
  R fzy = (Trig<R>::sn(Z + ty[n.zp] - tz[n.yp] - Y) - Trig<R>::sn(tz[n.ym] + ty[n.ym_zp] - Z - ty[n.ym])) / dy2;
  const R m2gg = (R)k.m2gg;
  R atx = m2gg * (cur[0][0] + cur[1][0]) - fxy - fxz;
  R aty = m2gg * (cur[0][1] + cur[1][1]) - fyx - fyz;
  R atz = m2gg * (cur[0][2] + cur[1][2]) - fzx - fzy;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    R now = L.cur[c][s];
    L.prv[c][s] = 2 * now - L.prv[c][s] + dt2 * acc[c >> 1][c & 1];
  }
  L.prv[4][s] = 2 * X - L.prv[4][s] + dt2 * atx;
  L.prv[5][s] = 2 * Y - L.prv[5][s] + dt2 * aty;
  L.prv[6][s] = 2 * Z - L.prv[6][s] + dt2 * atz;
}

template <typename R, int MODEL>
__global__ void __launch_bounds__(256) step_site_kernel(StepConsts k, Levels<R> L, int V) {
  const int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= V) return;
  Site n;
  int i, j, kk;
  if (!site_of(k, s, n, i, j, kk)) return;
  if (MODEL == AX_MODEL_AH)
    ah_update<R>(k, L, n);
  else
    moore_update<R>(k, L, n);
}

template <typename R>
hipError_t launch_step_site(const StepConsts &k, const Levels<R> &L, int model, hipStream_t st) {
  const int V = k.nx * k.ny * k.nz;
  dim3 grid((V + 255) / 256), block(256);
  if (model == AX_MODEL_AH)
    hipLaunchKernelGGL((step_site_kernel<R, AX_MODEL_AH>), grid, block, 0, st, k, L, V);
  else
    hipLaunchKernelGGL((step_site_kernel<R, AX_MODEL_MOORE>), grid, block, 0, st, k, L, V);
  return hipGetLastError();
}

template hipError_t launch_step_site<double>(const StepConsts &, const Levels<double> &, int, hipStream_t);
template hipError_t launch_step_site<float>(const StepConsts &, const Levels<float> &, int, hipStream_t);

}  // namespace ax
