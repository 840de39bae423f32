/*
 * ax_diag.hip - read-only diagnostic passes over one lattice state.
 *
 *  diag_kernel      energy density and Gauss-law deviation per site
 *                   (evolution.cpp:298-427, Moore mooreEvol.cpp:259-389) in
 *                   FP64 whatever the storage type, plus a finiteness flag
 *                   over every stored value; per-block FP64 partial sums are
 *                   reduced by one block in a fixed order, so the result is
 *                   deterministic for a given lattice shape.
 *  winding_kernel   per-plaquette winding numbers (no reference
 *                   counterpart) and their absolute sum.
 *  crossing kernels the bilinear face-crossing test of
 *                   mpi_evolution.cpp:536-789, compacted in site order.
 *
 * The diagnostics must read the current level of every neighbour and the
 * previous level (the Gauss law needs the neighbours' previous level,
 * evolution.cpp:390-391), so they run as a separate pass before the in-place
 * update of the same TimeStep.
 */
#include "ax_energy.h"
#include "ax_internal.h"
#include "ax_march.h"
#include "ax_wind.h"

namespace ax {


__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

/* Block-wide sum of 256 values in a fixed pairing order; valid in thread 0. */
__device__ __forceinline__ double block_sum256(double v, double *sh) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) sh[w] = v;
  __syncthreads();
  double r = 0;
  if (threadIdx.x == 0) r = (sh[0] + sh[1]) + (sh[2] + sh[3]);
  return r;
}

template <typename T>
__device__ __forceinline__ void site_energy_ah(const DiagConsts &k, const Levels<T> &L, int i, int j, int kk,
                                               double &rho, double &absG) {
  const int nx = k.nx, ny = k.ny, nz = k.nz;
  auto id = [&](int a, int b, int c) { return (a * ny + b) * nz + c; };
  const int s = id(i, j, kk);
  const int sxp = id(wrapp(i, nx), j, kk), sxm = id(wrapm(i, nx), j, kk);
  const int syp = id(i, wrapp(j, ny), kk), sym = id(i, wrapm(j, ny), kk);
  const int szp = id(i, j, wrapp(kk, nz)), szm = id(i, j, wrapm(kk, nz));
  const double dx = k.dx, dy = k.dy, dz = k.dz, dt = k.dt, g = k.g;
  const double pr = ld(L.cur[0], s), pi = ld(L.cur[1], s);
  const double mag2 = pr * pr + pi * pi;
  double vel[2], gx[2], gy[2], gz[2];
  double sxm_, cxm_, sym_, cym_, szm_, czm_;
  ax_sincos(ld(L.cur[2], sxm), &sxm_, &cxm_);
  ax_sincos(ld(L.cur[3], sym), &sym_, &cym_);
  ax_sincos(ld(L.cur[4], szm), &szm_, &czm_);
  {
    vel[0] = (pr - ld(L.prv[0], s)) / dt;
    gx[0] = (pr - cxm_ * ld(L.cur[0], sxm) + sxm_ * ld(L.cur[1], sxm)) / dx;
    gy[0] = (pr - cym_ * ld(L.cur[0], sym) + sym_ * ld(L.cur[1], sym)) / dy;
    gz[0] = (pr - czm_ * ld(L.cur[0], szm) + szm_ * ld(L.cur[1], szm)) / dz;
    vel[1] = (pi - ld(L.prv[1], s)) / dt;
    gx[1] = (pi - cxm_ * ld(L.cur[1], sxm) + (-sxm_) * ld(L.cur[0], sxm)) / dx;
    gy[1] = (pi - cym_ * ld(L.cur[1], sym) + (-sym_) * ld(L.cur[0], sym)) / dy;
    gz[1] = (pi - czm_ * ld(L.cur[1], szm) + (-szm_) * ld(L.cur[0], szm)) / dz;
  }
  const double X = ld(L.cur[2], s), Y = ld(L.cur[3], s), Z = ld(L.cur[4], s);
  const double pxy = X + ld(L.cur[3], sxp) - ld(L.cur[2], syp) - Y;
  const double pxz = X + ld(L.cur[4], sxp) - ld(L.cur[2], szp) - Z;
  const double pyz = Y + ld(L.cur[4], syp) - ld(L.cur[3], szp) - Z;
  const double oX = ld(L.prv[2], s), oY = ld(L.prv[3], s), oZ = ld(L.prv[4], s);
  double tdc, fcont, G = 0;
  const double vv = mag2 - k.eta * k.eta;
  if (k.energy_form == 0) {
    if (g == 0) {
      tdc = 0;
      fcont = 0;
    } else {
      double divN = (X - ld(L.cur[2], sxm)) / (dx * dx) + (Y - ld(L.cur[3], sym)) / (dy * dy) +
                    (Z - ld(L.cur[4], szm)) / (dz * dz);
      double divP = (oX - ld(L.prv[2], sxm)) / (dx * dx) + (oY - ld(L.prv[3], sym)) / (dy * dy) +
                    (oZ - ld(L.prv[4], szm)) / (dz * dz);
      double divT = (divN - divP) / dt;
      G = divT / g - 2 * g * (pi * vel[0] - pr * vel[1]);
      double ex = (X - oX) / (g * dt * dx), ey = (Y - oY) / (g * dt * dy), ez = (Z - oZ) / (g * dt * dz);
      tdc = 0.5 * (ex * ex + ey * ey + ez * ez);
      double bxy = (1 - ax_cos(pxy)) / ((g * dx * dy) * (g * dx * dy));
      double bxz = (1 - ax_cos(pxz)) / ((g * dx * dz) * (g * dx * dz));
      double byz = (1 - ax_cos(pyz)) / ((g * dy * dz) * (g * dy * dz));
      fcont = bxy + bxz + byz;
    }
    double r = vel[0] * vel[0] + vel[1] * vel[1] + gx[0] * gx[0] + gx[1] * gx[1] + gy[0] * gy[0] +
               gy[1] * gy[1] + gz[0] * gz[0] + gz[1] * gz[1] + tdc + fcont + 0.25 * k.lambda * (vv * vv);
    rho = dx * dy * dz * r;
  } else {
    double ex = (X - oX) / dt, ey = (Y - oY) / dt, ez = (Z - oZ) / dt;
    tdc = 0.5 * (ex * ex + ey * ey + ez * ez);
    double bxy = (1 - ax_cos(pxy)) / ((g * dx * dy) * (g * dx * dy));
    double bxz = (1 - ax_cos(pxz)) / ((g * dx * dz) * (g * dx * dz));
    double byz = (1 - ax_cos(pyz)) / ((g * dy * dz) * (g * dy * dz));
    fcont = bxy + bxz + byz;
    rho = (vel[0] * vel[0] + vel[1] * vel[1] + gx[0] * gx[0] + gx[1] * gx[1] + gy[0] * gy[0] +
           gy[1] * gy[1] + gz[0] * gz[0] + gz[1] * gz[1] + tdc + fcont + 0.25 * k.lambda * (vv * vv)) *
          dx * dy * dz;
  }
  absG = fabs(G);
}

template <typename T>
__device__ __forceinline__ void site_energy_moore(const DiagConsts &k, const Levels<T> &L, int i, int j, int kk,
                                                  double &rho, double &absG) {
  const int nx = k.nx, ny = k.ny, nz = k.nz;
  auto id = [&](int a, int b, int c) { return (a * ny + b) * nz + c; };
  const int s = id(i, j, kk);
  const int sxp = id(wrapp(i, nx), j, kk), sxm = id(wrapm(i, nx), j, kk);
  const int syp = id(i, wrapp(j, ny), kk), sym = id(i, wrapm(j, ny), kk);
  const int szp = id(i, j, wrapp(kk, nz)), szm = id(i, j, wrapm(kk, nz));
  const double dx = k.dx, dy = k.dy, dz = k.dz, dt = k.dt, g = k.g;
  const double q[2] = {k.q1, k.q2};
  double v[2][2], gx[2][2], gy[2][2], gz[2][2], m[2];
  for (int f = 0; f < 2; ++f) {
    const T *re = L.cur[2 * f], *im = L.cur[2 * f + 1];
    const double ar = ld(re, s), ai = ld(im, s);
    m[f] = ar * ar + ai * ai;
    double sxm_, cxm_, sym_, cym_, szm_, czm_;
    ax_sincos(q[f] * ld(L.cur[4], sxm), &sxm_, &cxm_);
    ax_sincos(q[f] * ld(L.cur[5], sym), &sym_, &cym_);
    ax_sincos(q[f] * ld(L.cur[6], szm), &szm_, &czm_);
    gx[f][0] = (ar - cxm_ * ld(re, sxm) + sxm_ * ld(im, sxm)) / dx;
    gy[f][0] = (ar - cym_ * ld(re, sym) + sym_ * ld(im, sym)) / dy;
    gz[f][0] = (ar - czm_ * ld(re, szm) + szm_ * ld(im, szm)) / dz;
    gx[f][1] = (ai - cxm_ * ld(im, sxm) + (-sxm_) * ld(re, sxm)) / dx;
    gy[f][1] = (ai - cym_ * ld(im, sym) + (-sym_) * ld(re, sym)) / dy;
    gz[f][1] = (ai - czm_ * ld(im, szm) + (-szm_) * ld(re, szm)) / dz;
    v[f][0] = (ar - ld(L.prv[2 * f], s)) / dt;
    v[f][1] = (ai - ld(L.prv[2 * f + 1], s)) / dt;
  }
  const double X = ld(L.cur[4], s), Y = ld(L.cur[5], s), Z = ld(L.cur[6], s);
  const double oX = ld(L.prv[4], s), oY = ld(L.prv[5], s), oZ = ld(L.prv[6], s);
  double tdc = 0, fcont = 0, G = 0;
  if (g != 0) {
    double divN = (X - ld(L.cur[4], sxm)) / (dx * dx) + (Y - ld(L.cur[5], sym)) / (dy * dy) +
                  (Z - ld(L.cur[6], szm)) / (dz * dz);
    double divP = (oX - ld(L.prv[4], sxm)) / (dx * dx) + (oY - ld(L.prv[5], sym)) / (dy * dy) +
                  (oZ - ld(L.prv[6], szm)) / (dz * dz);
    double divT = (divN - divP) / dt;
    G = divT / g - 2 * g * (q[0] * (ld(L.cur[1], s) * v[0][0] - ld(L.cur[0], s) * v[0][1]) +
                            q[1] * (ld(L.cur[3], s) * v[1][0] - ld(L.cur[2], s) * v[1][1]));
    double ex = (X - oX) / (g * dt * dx), ey = (Y - oY) / (g * dt * dy), ez = (Z - oZ) / (g * dt * dz);
    tdc = 0.5 * (ex * ex + ey * ey + ez * ez);
    double bxy = (1 - ax_cos(X + ld(L.cur[5], sxp) - ld(L.cur[4], syp) - Y)) / ((g * dx * dy) * (g * dx * dy));
    double bxz = (1 - ax_cos(X + ld(L.cur[6], sxp) - ld(L.cur[4], szp) - Z)) / ((g * dx * dz) * (g * dx * dz));
    double byz = (1 - ax_cos(Y + ld(L.cur[6], syp) - ld(L.cur[5], szp) - Z)) / ((g * dy * dz) * (g * dy * dz));
    fcont = bxy + bxz + byz;
  }
  const double w1 = m[0] - k.eta1 * k.eta1, w2 = m[1] - k.eta2 * k.eta2;
  double r = v[0][0] * v[0][0] + v[0][1] * v[0][1] + gx[0][0] * gx[0][0] + gx[0][1] * gx[0][1] +
             gy[0][0] * gy[0][0] + gy[0][1] * gy[0][1] + gz[0][0] * gz[0][0] + gz[0][1] * gz[0][1] +
             v[1][0] * v[1][0] + v[1][1] * v[1][1] + gx[1][0] * gx[1][0] + gx[1][1] * gx[1][1] +
             gy[1][0] * gy[1][0] + gy[1][1] * gy[1][1] + gz[1][0] * gz[1][0] + gz[1][1] * gz[1][1] + tdc +
             fcont + 0.25 * k.lambda1 * (w1 * w1) + 0.25 * k.lambda2 * (w2 * w2) + 0.5 * k.lambda12 * w1 * w2;
  rho = dx * dy * dz * r;
  absG = k.ex ? fabs(G) / k.inner : fabs(G);
}

/* Sites per thread of the winding pass (each thread takes sites 256 apart,
 * so loads stay coalesced); the energy pass takes one site per thread. */
constexpr int kSitesPerThread = 2;
constexpr int kSitesPerBlock = 256 * kSitesPerThread;
constexpr int kDiagSitesPerThread = 1;
constexpr int kDiagSitesPerBlock = 256 * kDiagSitesPerThread;

/* out: per block [energy, sum|G|, nonfinite count] */
template <typename T, int MODEL>
__global__ void __launch_bounds__(256) diag_kernel(DiagConsts k, Levels<T> L, int V, double *part) {
  __shared__ double sh[4];
  double rho = 0, aG = 0, bad = 0;
  for (int q = 0; q < kDiagSitesPerThread; ++q) {
    const int s = blockIdx.x * kDiagSitesPerBlock + q * 256 + threadIdx.x;
    if (s >= V) break;
    const int kk = s % k.nz, r = s / k.nz, j = r % k.ny, i = r / k.ny;
    const bool inner = !((k.ex && (i == 0 || i == k.nx - 1 || j == 0 || j == k.ny - 1)) ||
                         (k.ez && (kk == 0 || kk == k.nz - 1)));
    if (inner) {
      double r1 = 0, g1 = 0;
      if (MODEL == AX_MODEL_AH)
        site_energy_ah<T>(k, L, i, j, kk, r1, g1);
      else
        site_energy_moore<T>(k, L, i, j, kk, r1, g1);
      rho += r1;
      aG += g1;
    }
    for (int c = 0; c < k.ncomp; ++c) {
      if (!isfinite(ld(L.cur[c], s))) bad += 1;
      if (!isfinite(ld(L.prv[c], s))) bad += 1;
    }
  }
  double e = block_sum256(rho, sh);
  double g = block_sum256(aG, sh);
  double b = block_sum256(bad, sh);
  if (threadIdx.x == 0) {
    part[3 * blockIdx.x] = e;
    part[3 * blockIdx.x + 1] = g;
    part[3 * blockIdx.x + 2] = b;
  }
}

/* One block sums nb triples: thread t adds blocks t, t+1024, ... in order,
 * then a fixed tree; deterministic for a given nb. */
__global__ void __launch_bounds__(1024) sum3_kernel(const double *part, int nb, double *out) {
  __shared__ double sh[3][16];
  double v[3] = {0, 0, 0};
  for (int i = threadIdx.x; i < nb; i += 1024)
    for (int c = 0; c < 3; ++c) v[c] += part[3 * i + c];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  for (int c = 0; c < 3; ++c) {
    double x = wave_sum(v[c]);
    if (l == 0) sh[c][w] = x;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    double acc = 0;
    for (int q = 0; q < 16; ++q) acc += sh[threadIdx.x][q];
    out[threadIdx.x] = acc;
  }
}

/* ------------------------------- winding -------------------------------- */
template <typename T>
__global__ void __launch_bounds__(256) winding_kernel(int nx, int ny, int nz, const T *fr, const T *fi,
                                                      const T *tx, const T *ty, const T *tz, double q,
                                                      int ex, int ez, int8_t *out, double *part) {
  __shared__ double sh[4];
  const int V = nx * ny * nz;
  double tot[3] = {0, 0, 0};
  for (int qq = 0; qq < kSitesPerThread; ++qq) {
    const int s = blockIdx.x * kSitesPerBlock + qq * 256 + threadIdx.x;
    if (s >= V) break;
    const int kk = s % nz, r = s / nz, j = r % ny, i = r / ny;
    int w[3];
    site_winding<T>(nx, ny, nz, fr, fi, tx, ty, tz, q, i, j, kk, w);
    /* plaquettes that would wrap across a non-periodic axis do not exist */
    const bool xe = ex && i == nx - 1, ye = ex && j == ny - 1, ze = ez && kk == nz - 1;
    if (xe || ye) w[0] = 0;
    if (xe || ze) w[1] = 0;
    if (ye || ze) w[2] = 0;
#pragma unroll
    for (int pl = 0; pl < 3; ++pl) {
      out[(size_t)pl * V + s] = (int8_t)w[pl];
      tot[pl] += (double)(w[pl] < 0 ? -w[pl] : w[pl]);
    }
  }
  double a = block_sum256(tot[0], sh);
  double b = block_sum256(tot[1], sh);
  double c = block_sum256(tot[2], sh);
  if (threadIdx.x == 0) {
    part[3 * blockIdx.x] = a;
    part[3 * blockIdx.x + 1] = b;
    part[3 * blockIdx.x + 2] = c;
  }
}


/* ------------------------------ crossings ------------------------------- */
/* Solve the face of mpi_evolution.cpp:552-623; returns the number of roots
 * inside the face (0..2) and writes them as (sol1, sol2) pairs. */
__device__ __forceinline__ int face_roots(const double v00[2], const double v10[2], const double v01[2],
                                          const double v11[2], double u, double du, double w, double dw,
                                          double r[4]) {
  double c1[2], c2[2], c3[2], c4[2];
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    c1[c] = v11[c] - v10[c] - v01[c] + v00[c];
    c2[c] = (w + dw) * (v10[c] - v00[c]) + w * (v01[c] - v11[c]);
    c3[c] = (u + du) * (v01[c] - v00[c]) + u * (v10[c] - v11[c]);
    c4[c] = (u + du) * ((w + dw) * v00[c] - w * v01[c]) - u * ((w + dw) * v10[c] - w * v11[c]);
  }
  double a = c1[1] * c3[0] - c1[0] * c3[1];
  double b = c1[1] * c4[0] + c2[1] * c3[0] - c1[0] * c4[1] - c2[0] * c3[1];
  double c = c2[1] * c4[0] - c2[0] * c4[1];
  double disc = b * b - 4 * a * c;
  int n = 0;
  if (a == 0) {
    if (b != 0) {
      double s2 = -c / b;
      if (s2 >= w && s2 <= w + dw) {
        double s1 = -(c3[0] * s2 + c4[0]) / (c1[0] * s2 + c2[0]);
        if (s1 >= u && s1 <= u + du) { r[0] = s1; r[1] = s2; n = 1; }
      }
    }
  } else if (disc >= 0) {
    double s2 = (-b + sqrt(disc)) / (2 * a);
    if (s2 >= w && s2 <= w + dw) {
      double s1 = -(c3[0] * s2 + c4[0]) / (c1[0] * s2 + c2[0]);
      if (s1 >= u && s1 <= u + du) { r[2 * n] = s1; r[2 * n + 1] = s2; ++n; }
    }
    s2 = (-b - sqrt(disc)) / (2 * a);
    if (s2 >= w && s2 <= w + dw) {
      double s1 = -(c3[0] * s2 + c4[0]) / (c1[0] * s2 + c2[0]);
      if (s1 >= u && s1 <= u + du) { r[2 * n] = s1; r[2 * n + 1] = s2; ++n; }
    }
  }
  return n;
}

/* All crossings of site s, in the reference's order; returns count (<= 6). */
template <typename T>
__device__ int site_crossings(int nx, int ny, int nz, double dx, double dy, double dz, const T *fr, const T *fi,
                              int ex, int ez, int s, double pts[18]) {
  const int kk = s % nz, r = s / nz, j = r % ny, i = r / ny;
  /* only the sites the bulk update covers (evolution.cpp:731-743) */
  if ((ex && (i == 0 || i == nx - 1 || j == 0 || j == ny - 1)) || (ez && (kk == 0 || kk == nz - 1))) return 0;
  const int ip = wrapp(i, nx), jp = wrapp(j, ny), kp = wrapp(kk, nz);
  auto id = [&](int a, int b, int c) { return (a * ny + b) * nz + c; };
  const int sx = id(ip, j, kk), sy = id(i, jp, kk), sz = id(i, j, kp);
  const int sxy = id(ip, jp, kk), sxz = id(ip, j, kp), syz = id(i, jp, kp);
  const int x0 = int(0.5 * (nx - 1)), y0 = int(0.5 * (ny - 1)), z0 = int(0.5 * (nz - 1));
  const double x = (i - x0) * dx, y = (j - y0) * dy, z = (kk - z0) * dz;
  const double v[2] = {ld(fr, s), ld(fi, s)}, vx[2] = {ld(fr, sx), ld(fi, sx)};
  const double vy[2] = {ld(fr, sy), ld(fi, sy)}, vz[2] = {ld(fr, sz), ld(fi, sz)};
  const double vxy[2] = {ld(fr, sxy), ld(fi, sxy)}, vxz[2] = {ld(fr, sxz), ld(fi, sxz)};
  const double vyz[2] = {ld(fr, syz), ld(fi, syz)};
  double rr[4];
  int n = 0;
  int m = face_roots(v, vx, vy, vxy, x, dx, y, dy, rr);
  for (int q = 0; q < m; ++q) { pts[3 * n] = rr[2 * q]; pts[3 * n + 1] = rr[2 * q + 1]; pts[3 * n + 2] = z; ++n; }
  m = face_roots(v, vx, vz, vxz, x, dx, z, dz, rr);
  for (int q = 0; q < m; ++q) { pts[3 * n] = rr[2 * q]; pts[3 * n + 1] = y; pts[3 * n + 2] = rr[2 * q + 1]; ++n; }
  m = face_roots(v, vy, vz, vyz, y, dy, z, dz, rr);
  for (int q = 0; q < m; ++q) { pts[3 * n] = x; pts[3 * n + 1] = rr[2 * q]; pts[3 * n + 2] = rr[2 * q + 1]; ++n; }
  return n;
}

/* pass 1: per-block crossing counts */
template <typename T>
__global__ void __launch_bounds__(256) cross_count_kernel(int nx, int ny, int nz, double dx, double dy, double dz,
                                                          const T *fr, const T *fi, int ex, int ez, int *bcount) {
  __shared__ int sh[4];
  const int V = nx * ny * nz;
  const int s = blockIdx.x * 256 + threadIdx.x;
  int n = 0;
  if (s < V) {
    double pts[18];
    n = site_crossings<T>(nx, ny, nz, dx, dy, dz, fr, fi, ex, ez, s, pts);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) n += __shfl_xor(n, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = n;
  __syncthreads();
  if (threadIdx.x == 0) bcount[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

/* pass 2: exclusive scan of the block counts (one block). */
__global__ void __launch_bounds__(1024) scan_kernel(const int *in, long long *out, int n, long long *total) {
  __shared__ long long sh[1024];
  const int t = threadIdx.x;
  const int per = (n + 1023) / 1024;
  const int b0 = t * per, b1 = min(n, b0 + per);
  long long loc = 0;
  for (int i = b0; i < b1; ++i) loc += in[i];
  sh[t] = loc;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    long long v = (t >= off) ? sh[t - off] : 0;
    __syncthreads();
    sh[t] += v;
    __syncthreads();
  }
  long long run = sh[t] - loc;
  for (int i = b0; i < b1; ++i) {
    out[i] = run;
    run += in[i];
  }
  if (t == 1023) *total = sh[1023];
}

/* pass 3: recompute and write points at their global offsets. */
template <typename T>
__global__ void __launch_bounds__(256) cross_write_kernel(int nx, int ny, int nz, double dx, double dy, double dz,
                                                          const T *fr, const T *fi, int ex, int ez,
                                                          const long long *boff, double *xyz, long long cap) {
  __shared__ int sh[256];
  const int V = nx * ny * nz;
  const int s = blockIdx.x * 256 + threadIdx.x;
  double pts[18];
  int n = 0;
  if (s < V) n = site_crossings<T>(nx, ny, nz, dx, dy, dz, fr, fi, ex, ez, s, pts);
  sh[threadIdx.x] = n;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    int v = ((int)threadIdx.x >= off) ? sh[threadIdx.x - off] : 0;
    __syncthreads();
    sh[threadIdx.x] += v;
    __syncthreads();
  }
  long long o = boff[blockIdx.x] + sh[threadIdx.x] - n;
  for (int q = 0; q < n; ++q, ++o) {
    if (o < cap) {
      xyz[3 * o] = pts[3 * q];
      xyz[3 * o + 1] = pts[3 * q + 1];
      xyz[3 * o + 2] = pts[3 * q + 2];
    }
  }
}

/* ------------------- marching diagnostics (Abelian-Higgs) -----------------
 * One 64-lane wave per (64-site z chunk, row, x segment) marching along x,
 * as the step kernel.  Per site it evaluates the energy density and Gauss
 * residual with the expressions of site_energy_ah (the same per-site bits;
 * with x/y or z not periodic only the bulk region's terms are summed, as
 * diag_kernel does),
 * carrying the x-backward field, link and link trig from the previous
 * plane, and, with PH (periodic lattices), the three forward edge phases
 *   delta_a(s) = arg( conj(phi(s)) exp(-i theta_a(s)) phi(s+a) )
 * from the same link sincos, stored as the FP32 edge values
 * e_a(s) = delta_a(s) + theta_a(s) that wind_march_kernel combines
 * (ax_wind.h).  The winding of a plaquette needs four edges of three
 * sites; computing each edge once here instead of nine per site in
 * winding_kernel is the saving. */

/* link_phase_sc (ax_energy.h): link_phase with the link's sin/cos already at hand */

/* P2: dx, dy, dz and g (when g != 0) are powers of two, so every division by
 * dx, dx*dx, g or (g*dx*dy)^2 is the multiplication by its exact reciprocal
 * (the same correctly rounded value).  Divisions by dt stay divisions.
 * FP64 is held to 3 waves per SIMD (the phase-storing variant needs 173
 * VGPRs otherwise, i.e. 2 waves; capped it spills 4 and runs 3.5 % faster
 * per 512^3 call); FP32 keeps the compiler's choice (already 3 waves; the
 * same cap there changed its schedule and cost 36 %). */
template <typename T, bool PH, bool P2>
__global__ void __launch_bounds__(64, sizeof(T) == 8 ? 3 : 1) diag_march_kernel(DiagConsts k, Levels<T> L, TileGrid g, float *phase,
                                                        double *part) {
  const int w = (int)blockIdx.x;
  int zc, y, seg;
  tile_item(g, k.ny, w, zc, y, seg);
  const int x0 = seg * g.xseg, x1 = min(k.nx, x0 + g.xseg);
  const int NX = k.nx, NY = k.ny, NZ = k.nz;
  const int lane = (int)(threadIdx.x & 63);
  const int zr = zc * 64 + lane;
  const bool act = zr < NZ;
  const int z = act ? zr : NZ - 1;
  const int ym = y == 0 ? NY - 1 : y - 1, yp = y == NY - 1 ? 0 : y + 1;
  const bool in_yz = !(k.ex && (y == 0 || y == NY - 1)) && !(k.ez && (z == 0 || z == NZ - 1));
  const uint32_t E = (uint32_t)sizeof(T);
  const uint32_t oz = (uint32_t)z * E;
  const uint32_t ozm = (uint32_t)(z == 0 ? NZ - 1 : z - 1) * E;
  const uint32_t ozp = (uint32_t)(z == NZ - 1 ? 0 : z + 1) * E;
  const uint32_t R0 = (uint32_t)y * NZ * E, RM = (uint32_t)ym * NZ * E, RP = (uint32_t)yp * NZ * E;
  const T *PR = L.cur[0], *PI = L.cur[1], *TX = L.cur[2], *TY = L.cur[3], *TZ = L.cur[4];
  const T *OPR = L.prv[0], *OPI = L.prv[1], *OTX = L.prv[2], *OTY = L.prv[3], *OTZ = L.prv[4];
  /* value of field b in x plane x at (row, lane) byte offsets: the plane's
   * 64-bit address is formed on the scalar unit, so planes of any size work */
  const size_t NYNZ = (size_t)NY * NZ;
  auto ld_ = [NYNZ](const T *b, int x, uint32_t row, uint32_t off) {
    return (double)ldo(b + (size_t)x * NYNZ, row, off);
  };
  const double dx = k.dx, dy = k.dy, dz = k.dz, dt = k.dt, gg = k.g;
  const size_t V = (size_t)NX * NY * NZ;
  const double dxx = dx * dx, dyy = dy * dy, dzz = dz * dz;
  const double bxyd = (gg * dx * dy) * (gg * dx * dy), bxzd = (gg * dx * dz) * (gg * dx * dz),
               byzd = (gg * dy * dz) * (gg * dy * dz);
  const double idx = 1 / dx, idy = 1 / dy, idz = 1 / dz, idxx = 1 / dxx, idyy = 1 / dyy, idzz = 1 / dzz;
  const double igg = 1 / gg, ibxy = 1 / bxyd, ibxz = 1 / bxzd, ibyz = 1 / byzd;
  auto dv = [](double a, double d, double inv) { return P2 ? a * inv : a / d; };
  /* divisions by dt and g dt dx_a as multiplications by the reciprocal: the
   * per-site terms change by at most an ulp, far inside the 1e-12 the sums
   * are held to (the site kernels divide).  Not in the FP64 pass that also
   * stores the edge phases, where the extra live reciprocals cost more
   * spills than the divisions (512^3: 6.50 -> 6.79 ms with them; FP32
   * 5.30 -> 5.01 ms, FP64 without phases 2.67 -> 2.51 ms) */
  constexpr bool RECIP = !(PH && sizeof(T) == 8);
  const double idt = 1 / dt;
  const double iedx = 1 / (gg * dt * dx), iedy = 1 / (gg * dt * dy), iedz = 1 / (gg * dt * dz);
  auto by_dt = [&](double a) { return RECIP ? a * idt : a / dt; };
  auto by_e = [&](double a, double d, double inv) { return RECIP ? a * inv : a / (gg * dt * d); };

  /* carried from plane x-1: phi, theta_x (both levels), sin/cos theta_x */
  const int xm0 = x0 == 0 ? NX - 1 : x0 - 1;
  double PRm = ld_(PR, xm0, R0, oz), PIm = ld_(PI, xm0, R0, oz);
  double TXm = ld_(TX, xm0, R0, oz), oXm = ld_(OTX, xm0, R0, oz);
  double sxm, cxm;
  ax_sincos(TXm, &sxm, &cxm);
  double PRc = ld_(PR, x0, R0, oz), PIc = ld_(PI, x0, R0, oz);
  double X = ld_(TX, x0, R0, oz), Y = ld_(TY, x0, R0, oz), Z = ld_(TZ, x0, R0, oz);
  double rho_sum = 0, g_sum = 0, bad = 0;

  for (int x = x0; x < x1; ++x) {
    const int xp = x == NX - 1 ? 0 : x + 1;
    const double PRn = ld_(PR, xp, R0, oz), PIn = ld_(PI, xp, R0, oz);
    const double TXn = ld_(TX, xp, R0, oz), TYn = ld_(TY, xp, R0, oz), TZn = ld_(TZ, xp, R0, oz);
    const double oPR = ld_(OPR, x, R0, oz), oPI = ld_(OPI, x, R0, oz);
    const double oX = ld_(OTX, x, R0, oz), oY = ld_(OTY, x, R0, oz), oZ = ld_(OTZ, x, R0, oz);
    const double PRyp = ld_(PR, x, RP, oz), PIyp = ld_(PI, x, RP, oz);
    const double TXyp = ld_(TX, x, RP, oz), TZyp = ld_(TZ, x, RP, oz);
    const double PRzp = ld_(PR, x, R0, ozp), PIzp = ld_(PI, x, R0, ozp);
    const double TXzp = ld_(TX, x, R0, ozp), TYzp = ld_(TY, x, R0, ozp);
    const double PRym = ld_(PR, x, RM, oz), PIym = ld_(PI, x, RM, oz);
    const double TYym = ld_(TY, x, RM, oz), oYym = ld_(OTY, x, RM, oz);
    const double PRzm = ld_(PR, x, R0, ozm), PIzm = ld_(PI, x, R0, ozm);
    const double TZzm = ld_(TZ, x, R0, ozm), oZzm = ld_(OTZ, x, R0, ozm);

    /* site_energy_ah at s, operands by name */
    const double pr = PRc, pi = PIc;
    const double mag2 = pr * pr + pi * pi;
    double sym_, cym_, szm_, czm_;
    ax_sincos(TYym, &sym_, &cym_);
    ax_sincos(TZzm, &szm_, &czm_);
    double vel[2], gx[2], gy[2], gz[2];
    vel[0] = by_dt(pr - oPR);
    gx[0] = dv(pr - cxm * PRm + sxm * PIm, dx, idx);
    gy[0] = dv(pr - cym_ * PRym + sym_ * PIym, dy, idy);
    gz[0] = dv(pr - czm_ * PRzm + szm_ * PIzm, dz, idz);
    vel[1] = by_dt(pi - oPI);
    gx[1] = dv(pi - cxm * PIm + (-sxm) * PRm, dx, idx);
    gy[1] = dv(pi - cym_ * PIym + (-sym_) * PRym, dy, idy);
    gz[1] = dv(pi - czm_ * PIzm + (-szm_) * PRzm, dz, idz);
    const double pxy = X + TYn - TXyp - Y;
    const double pxz = X + TZn - TXzp - Z;
    const double pyz = Y + TZyp - TYzp - Z;
    /* cosines of the three plaquette angles; the short k = 0 form when every
     * lane's angles qualify (bit-identical to ax_cos there) */
    auto plaq_cos = [&](double &cxy, double &cxz, double &cyz) {
      if (__all(ax_sin_is_small(pxy) && ax_sin_is_small(pxz) && ax_sin_is_small(pyz))) {
        cxy = ax_cos_kernel(pxy, 0.0); cxz = ax_cos_kernel(pxz, 0.0); cyz = ax_cos_kernel(pyz, 0.0);
      } else {
        cxy = ax_cos(pxy); cxz = ax_cos(pxz); cyz = ax_cos(pyz);
      }
    };
    double tdc, fcont, G = 0, rho;
    const double vv = mag2 - k.eta * k.eta;
    if (k.energy_form == 0) {
      if (gg == 0) {
        tdc = 0;
        fcont = 0;
      } else {
        double divN = dv(X - TXm, dxx, idxx) + dv(Y - TYym, dyy, idyy) + dv(Z - TZzm, dzz, idzz);
        double divP = dv(oX - oXm, dxx, idxx) + dv(oY - oYym, dyy, idyy) + dv(oZ - oZzm, dzz, idzz);
        double divT = by_dt(divN - divP);
        G = dv(divT, gg, igg) - 2 * gg * (pi * vel[0] - pr * vel[1]);
        double ex = by_e(X - oX, dx, iedx), ey = by_e(Y - oY, dy, iedy), ez = by_e(Z - oZ, dz, iedz);
        tdc = 0.5 * (ex * ex + ey * ey + ez * ez);
        double cxy, cxz, cyz;
        plaq_cos(cxy, cxz, cyz);
        double bxy = dv(1 - cxy, bxyd, ibxy);
        double bxz = dv(1 - cxz, bxzd, ibxz);
        double byz = dv(1 - cyz, byzd, ibyz);
        fcont = bxy + bxz + byz;
      }
      double r_ = vel[0] * vel[0] + vel[1] * vel[1] + gx[0] * gx[0] + gx[1] * gx[1] + gy[0] * gy[0] +
                  gy[1] * gy[1] + gz[0] * gz[0] + gz[1] * gz[1] + tdc + fcont + 0.25 * k.lambda * (vv * vv);
      rho = dx * dy * dz * r_;
    } else {
      double ex = by_dt(X - oX), ey = by_dt(Y - oY), ez = by_dt(Z - oZ);
      tdc = 0.5 * (ex * ex + ey * ey + ez * ez);
      double cxy, cxz, cyz;
      plaq_cos(cxy, cxz, cyz);
      double bxy = dv(1 - cxy, bxyd, ibxy);
      double bxz = dv(1 - cxz, bxzd, ibxz);
      double byz = dv(1 - cyz, byzd, ibyz);
      fcont = bxy + bxz + byz;
      rho = (vel[0] * vel[0] + vel[1] * vel[1] + gx[0] * gx[0] + gx[1] * gx[1] + gy[0] * gy[0] +
             gy[1] * gy[1] + gz[0] * gz[0] + gz[1] * gz[1] + tdc + fcont + 0.25 * k.lambda * (vv * vv)) *
            dx * dy * dz;
    }
    double sx, cx;
    ax_sincos(X, &sx, &cx);
    if (PH) {
      double sy, cy, sz, cz;
      ax_sincos(Y, &sy, &cy);
      ax_sincos(Z, &sz, &cz);
      const double dX = link_phase_sc(pr, pi, PRn, PIn, sx, cx);
      const double dY = link_phase_sc(pr, pi, PRyp, PIyp, sy, cy);
      const double dZ = link_phase_sc(pr, pi, PRzp, PIzp, sz, cz);
      if (act) {
        const size_t s = ((size_t)x * NY + y) * NZ + z;
        phase[s] = (float)(dX + X);   /* edge values e_a(s) (ax_wind.h) */
        phase[V + s] = (float)(dY + Y);
        phase[2 * V + s] = (float)(dZ + Z);
      }
    }
    if (act) {
      /* energy and Gauss over the bulk region only when x/y or z are not
       * periodic (the boundary sites' terms, computed from wrapped
       * neighbours, are dropped), finiteness over every site -- as
       * diag_kernel */
      if (in_yz && !(k.ex && (x == 0 || x == NX - 1))) {
        rho_sum += rho;
        g_sum += fabs(G);
      }
      const double vals[10] = {PRc, PIc, X, Y, Z, oPR, oPI, oX, oY, oZ};
#pragma unroll
      for (int q = 0; q < 10; ++q) bad += isfinite(vals[q]) ? 0.0 : 1.0;
    }
    PRm = PRc; PIm = PIc; TXm = X; oXm = oX; sxm = sx; cxm = cx;
    PRc = PRn; PIc = PIn; X = TXn; Y = TYn; Z = TZn;
  }
  const double a = wave_sum(rho_sum), bsum = wave_sum(g_sum), c = wave_sum(bad);
  if (lane == 0) {
    part[3 * w] = a;
    part[3 * w + 1] = bsum;
    part[3 * w + 2] = c;
  }
}

/* Windings of the current level from the edge values the marching
 * diagnostics stored (wind_march, ax_wind.h): one wave per work item of
 * the tile grid, carrying plane x+1's values.  out (the winding field
 * [3][V]) may be null when only the counts are wanted. */
__global__ void __launch_bounds__(64) wind_march_kernel(int nx, int ny, int nz, const float *E, TileGrid g, int8_t *out,
                                                        double *part) {
  const int item = (int)blockIdx.x;
  int zc, y, seg;
  tile_item(g, ny, item, zc, y, seg);
  const int x0 = seg * g.xseg, x1 = min(nx, x0 + g.xseg);
  const WindRow w(nx, ny, nz, y, zc);
  double tot[3] = {0, 0, 0};
  wind_march(w, E, x0, x1, [&](size_t s, const int *wn) {
#pragma unroll
    for (int pl = 0; pl < 3; ++pl) {
      if (out) out[(size_t)pl * w.V + s] = (int8_t)wn[pl];
      tot[pl] += (double)(wn[pl] < 0 ? -wn[pl] : wn[pl]);
    }
  });
  const double a = wave_sum(tot[0]), b = wave_sum(tot[1]), c = wave_sum(tot[2]);
  if ((threadIdx.x & 63) == 0) {
    part[3 * item] = a;
    part[3 * item + 1] = b;
    part[3 * item + 2] = c;
  }
}

/* ------------------ marching diagnostics (Moore model) -------------------
 * Energy density and Gauss residual of site_energy_moore (the same
 * expressions in the same order, so the same per-site terms) by one wave
 * per (64-site z chunk, row, x segment) marching along x: both scalars and
 * theta_x (both levels) of plane x-1 are carried from the previous plane,
 * and plane x+1's current level is loaded once, as the next plane's centre.
 * The site kernel reads the x-neighbour planes, which another pass over the
 * lattice had fetched, a second time.  Sums over the bulk region where x/y
 * or z are not periodic, finiteness over the fourteen stored values of
 * every site, as diag_kernel.  FP64 is held to 3 waves per SIMD (222 VGPRs
 * otherwise; 2 spilled with the cap): 256^3 0.83 -> 0.72 ms per call. */
template <typename T>
__global__ void __launch_bounds__(64, sizeof(T) == 8 ? 3 : 1) moore_diag_march_kernel(DiagConsts k, Levels<T> L,
                                                                                   TileGrid g, double *part) {
  const int w = (int)blockIdx.x;
  int zc, y, seg;
  tile_item(g, k.ny, w, zc, y, seg);
  const int x0 = seg * g.xseg, x1 = min(k.nx, x0 + g.xseg);
  const int NX = k.nx, NY = k.ny, NZ = k.nz;
  const int lane = (int)(threadIdx.x & 63);
  const int zr = zc * 64 + lane;
  const bool act = zr < NZ;
  const int z = act ? zr : NZ - 1;
  const int ym = y == 0 ? NY - 1 : y - 1, yp = y == NY - 1 ? 0 : y + 1;
  const bool in_yz = !(k.ex && (y == 0 || y == NY - 1)) && !(k.ez && (z == 0 || z == NZ - 1));
  const uint32_t E = (uint32_t)sizeof(T);
  const uint32_t oz = (uint32_t)z * E;
  const uint32_t ozm = (uint32_t)(z == 0 ? NZ - 1 : z - 1) * E;
  const uint32_t ozp = (uint32_t)(z == NZ - 1 ? 0 : z + 1) * E;
  const uint32_t R0 = (uint32_t)y * NZ * E, RM = (uint32_t)ym * NZ * E, RP = (uint32_t)yp * NZ * E;
  const size_t NYNZ = (size_t)NY * NZ;
  auto C = [&](int c, int x, uint32_t row, uint32_t off) { return (double)ldo(L.cur[c] + (size_t)x * NYNZ, row, off); };
  auto P = [&](int c, int x, uint32_t row, uint32_t off) { return (double)ldo(L.prv[c] + (size_t)x * NYNZ, row, off); };
  const double dx = k.dx, dy = k.dy, dz = k.dz, dt = k.dt, gg = k.g;
  const double q[2] = {k.q1, k.q2};
  /* plane x-1 (carried): both scalars, theta_x of both levels; plane x */
  const int xm0 = x0 == 0 ? NX - 1 : x0 - 1;
  double Fm[4], TXm = C(4, xm0, R0, oz), oXm = P(4, xm0, R0, oz), Fc[4], Tc[3];
  for (int c = 0; c < 4; ++c) {
    Fm[c] = C(c, xm0, R0, oz);
    Fc[c] = C(c, x0, R0, oz);
  }
  for (int a = 0; a < 3; ++a) Tc[a] = C(4 + a, x0, R0, oz);
  double rho_sum = 0, g_sum = 0, bad = 0;
  for (int x = x0; x < x1; ++x) {
    const int xp = x == NX - 1 ? 0 : x + 1;
    double Fn[4], Tn[3], Fo[4], Fy[4], Fz[4];
    for (int c = 0; c < 4; ++c) {
      Fn[c] = C(c, xp, R0, oz);
      Fo[c] = P(c, x, R0, oz);
      Fy[c] = C(c, x, RM, oz);
      Fz[c] = C(c, x, R0, ozm);
    }
    for (int a = 0; a < 3; ++a) Tn[a] = C(4 + a, xp, R0, oz);
    const double oX = P(4, x, R0, oz), oY = P(5, x, R0, oz), oZ = P(6, x, R0, oz);
    const double TYym = C(5, x, RM, oz), oYym = P(5, x, RM, oz);
    const double TZzm = C(6, x, R0, ozm), oZzm = P(6, x, R0, ozm);
    const double TXyp = C(4, x, RP, oz), TZyp = C(6, x, RP, oz);
    const double TXzp = C(4, x, R0, ozp), TYzp = C(5, x, R0, ozp);
    const double X = Tc[0], Y = Tc[1], Z = Tc[2];

    /* site_energy_moore at s, operands by name */
    double v[2][2], gx[2][2], gy[2][2], gz[2][2], m[2];
    for (int f = 0; f < 2; ++f) {
      const double ar = Fc[2 * f], ai = Fc[2 * f + 1];
      m[f] = ar * ar + ai * ai;
      double sxm_, cxm_, sym_, cym_, szm_, czm_;
      ax_sincos(q[f] * TXm, &sxm_, &cxm_);
      ax_sincos(q[f] * TYym, &sym_, &cym_);
      ax_sincos(q[f] * TZzm, &szm_, &czm_);
      const double rxm = Fm[2 * f], ixm = Fm[2 * f + 1], rym = Fy[2 * f], iym = Fy[2 * f + 1];
      const double rzm = Fz[2 * f], izm = Fz[2 * f + 1];
      gx[f][0] = (ar - cxm_ * rxm + sxm_ * ixm) / dx;
      gy[f][0] = (ar - cym_ * rym + sym_ * iym) / dy;
      gz[f][0] = (ar - czm_ * rzm + szm_ * izm) / dz;
      gx[f][1] = (ai - cxm_ * ixm + (-sxm_) * rxm) / dx;
      gy[f][1] = (ai - cym_ * iym + (-sym_) * rym) / dy;
      gz[f][1] = (ai - czm_ * izm + (-szm_) * rzm) / dz;
      v[f][0] = (ar - Fo[2 * f]) / dt;
      v[f][1] = (ai - Fo[2 * f + 1]) / dt;
    }
    double tdc = 0, fcont = 0, G = 0;
    if (gg != 0) {
      double divN = (X - TXm) / (dx * dx) + (Y - TYym) / (dy * dy) + (Z - TZzm) / (dz * dz);
      double divP = (oX - oXm) / (dx * dx) + (oY - oYym) / (dy * dy) + (oZ - oZzm) / (dz * dz);
      double divT = (divN - divP) / dt;
      G = divT / gg - 2 * gg * (q[0] * (Fc[1] * v[0][0] - Fc[0] * v[0][1]) +
                                q[1] * (Fc[3] * v[1][0] - Fc[2] * v[1][1]));
      double ex = (X - oX) / (gg * dt * dx), ey = (Y - oY) / (gg * dt * dy), ez = (Z - oZ) / (gg * dt * dz);
      tdc = 0.5 * (ex * ex + ey * ey + ez * ez);
      double bxy = (1 - ax_cos(X + Tn[1] - TXyp - Y)) / ((gg * dx * dy) * (gg * dx * dy));
      double bxz = (1 - ax_cos(X + Tn[2] - TXzp - Z)) / ((gg * dx * dz) * (gg * dx * dz));
      double byz = (1 - ax_cos(Y + TZyp - TYzp - Z)) / ((gg * dy * dz) * (gg * dy * dz));
      fcont = bxy + bxz + byz;
    }
    const double w1 = m[0] - k.eta1 * k.eta1, w2 = m[1] - k.eta2 * k.eta2;
    double r = v[0][0] * v[0][0] + v[0][1] * v[0][1] + gx[0][0] * gx[0][0] + gx[0][1] * gx[0][1] +
               gy[0][0] * gy[0][0] + gy[0][1] * gy[0][1] + gz[0][0] * gz[0][0] + gz[0][1] * gz[0][1] +
               v[1][0] * v[1][0] + v[1][1] * v[1][1] + gx[1][0] * gx[1][0] + gx[1][1] * gx[1][1] +
               gy[1][0] * gy[1][0] + gy[1][1] * gy[1][1] + gz[1][0] * gz[1][0] + gz[1][1] * gz[1][1] + tdc +
               fcont + 0.25 * k.lambda1 * (w1 * w1) + 0.25 * k.lambda2 * (w2 * w2) + 0.5 * k.lambda12 * w1 * w2;
    const double rho = dx * dy * dz * r;
    const double absG = k.ex ? fabs(G) / k.inner : fabs(G);
    if (act) {
      if (in_yz && !(k.ex && (x == 0 || x == NX - 1))) {
        rho_sum += rho;
        g_sum += absG;
      }
      const double vals[14] = {Fc[0], Fc[1], Fc[2], Fc[3], X, Y, Z, Fo[0], Fo[1], Fo[2], Fo[3], oX, oY, oZ};
#pragma unroll
      for (int c = 0; c < 14; ++c) bad += isfinite(vals[c]) ? 0.0 : 1.0;
    }
    for (int c = 0; c < 4; ++c) {
      Fm[c] = Fc[c];
      Fc[c] = Fn[c];
    }
    TXm = X;
    oXm = oX;
    for (int a = 0; a < 3; ++a) Tc[a] = Tn[a];
  }
  const double a = wave_sum(rho_sum), b = wave_sum(g_sum), c = wave_sum(bad);
  if (lane == 0) {
    part[3 * w] = a;
    part[3 * w + 1] = b;
    part[3 * w + 2] = c;
  }
}

/* ------------------------------ launchers ------------------------------- */
template <typename T>
hipError_t launch_diag(const DiagConsts &k, const Levels<T> &L, double *part, double *out3, hipStream_t st) {
  const int V = k.nx * k.ny * k.nz;
  int nb = (V + kDiagSitesPerBlock - 1) / kDiagSitesPerBlock;
  if (k.model == AX_MODEL_AH) {
    hipLaunchKernelGGL((diag_kernel<T, AX_MODEL_AH>), dim3(nb), dim3(256), 0, st, k, L, V, part);
  } else if (getenv("AX_DIAG_SITE_KERNEL") == nullptr) {
    /* part holds 3 doubles per work item (diag_part_blocks) */
    const TileGrid g = make_tile_grid(k.nx, k.ny, k.nz, kDiagXseg);
    hipLaunchKernelGGL((moore_diag_march_kernel<T>), dim3(g.work), dim3(64), 0, st, k, L, g, part);
    nb = g.work;
  } else {
    hipLaunchKernelGGL((diag_kernel<T, AX_MODEL_MOORE>), dim3(nb), dim3(256), 0, st, k, L, V, part);
  }
  hipLaunchKernelGGL(sum3_kernel, dim3(1), dim3(1024), 0, st, part, nb, out3);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_winding(int nx, int ny, int nz, const T *fr, const T *fi, const T *tx, const T *ty,
                          const T *tz, double q, int ex, int ez, int8_t *out, double *part, double *out1,
                          hipStream_t st) {
  const int V = nx * ny * nz;
  const int nb = (V + kSitesPerBlock - 1) / kSitesPerBlock;
  hipLaunchKernelGGL((winding_kernel<T>), dim3(nb), dim3(256), 0, st, nx, ny, nz, fr, fi, tx, ty, tz, q, ex, ez,
                     out, part);
  hipLaunchKernelGGL(sum3_kernel, dim3(1), dim3(1024), 0, st, part, nb, out1);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_crossings(int nx, int ny, int nz, double dx, double dy, double dz, const T *fr, const T *fi,
                            int ex, int ez, int *bcount, long long *boff, long long *total, double *xyz,
                            long long cap, hipStream_t st, bool write) {
  const int V = nx * ny * nz;
  const int nb = (V + 255) / 256;
  hipLaunchKernelGGL((cross_count_kernel<T>), dim3(nb), dim3(256), 0, st, nx, ny, nz, dx, dy, dz, fr, fi, ex, ez,
                     bcount);
  hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(1024), 0, st, bcount, boff, nb, total);
  if (write)
    hipLaunchKernelGGL((cross_write_kernel<T>), dim3(nb), dim3(256), 0, st, nx, ny, nz, dx, dy, dz, fr, fi,
                       ex, ez, boff, xyz, cap);
  return hipGetLastError();
}

/* The FP32 edge values of field (fr, fi) of charge q over the work items of
 * the tile grid (phase_march, ax_wind.h), for periodic lattices whose
 * energy and Gauss law take the site kernel (Moore). */
template <typename T>
__global__ void __launch_bounds__(64) edge_march_kernel(int nx, int ny, int nz, const T *fr, const T *fi, const T *tx,
                                                        const T *ty, const T *tz, double q, TileGrid g, float *E) {
  int zc, y, seg;
  tile_item(g, ny, (int)blockIdx.x, zc, y, seg);
  const int x0 = seg * g.xseg, x1 = min(nx, x0 + g.xseg);
  const WindRow w(nx, ny, nz, y, zc);
  (void)phase_march<T>(w, fr, fi, tx, ty, tz, E, x0, x1, q);
}

template <typename T>
hipError_t launch_edge_winding(int nx, int ny, int nz, const T *fr, const T *fi, const T *tx, const T *ty,
                               const T *tz, double q, float *E, int8_t *out, double *part, double *out3,
                               hipStream_t st) {
  const TileGrid g = make_tile_grid(nx, ny, nz, kDiagXseg);
  hipLaunchKernelGGL((edge_march_kernel<T>), dim3(g.work), dim3(64), 0, st, nx, ny, nz, fr, fi, tx, ty, tz, q, g,
                     E);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? launch_wind_phase(nx, ny, nz, E, out, part, out3, st) : e;
}

template <typename T>
hipError_t launch_diag_march(const DiagConsts &k, const Levels<T> &L, int xseg, float *phase, double *part,
                             double *out3, hipStream_t st) {
  const TileGrid g = make_tile_grid(k.nx, k.ny, k.nz, xseg);
  /* every divisor the P2 kernel multiplies by the reciprocal of must be a
   * power of two whose reciprocal is finite (then a*(1/d) == a/d exactly) */
  const double dx = k.dx, dy = k.dy, dz = k.dz, gg = k.g;
  const double dvs[10] = {dx, dy, dz, dx * dx, dy * dy, dz * dz, gg, (gg * dx * dy) * (gg * dx * dy),
                          (gg * dx * dz) * (gg * dx * dz), (gg * dy * dz) * (gg * dy * dz)};
  bool p2 = gg != 0;
  for (double d : dvs) p2 = p2 && is_pow2(fabs(d)) && std::isfinite(1 / d);
#define AX_DM(PH, P2) \
  hipLaunchKernelGGL((diag_march_kernel<T, PH, P2>), dim3(g.work), dim3(64), 0, st, k, L, g, phase, part)
  if (phase && p2) AX_DM(true, true);
  else if (phase) AX_DM(true, false);
  else if (p2) AX_DM(false, true);
  else AX_DM(false, false);
#undef AX_DM
  hipLaunchKernelGGL(sum3_kernel, dim3(1), dim3(1024), 0, st, part, g.work, out3);
  return hipGetLastError();
}

hipError_t launch_sum3(const double *part, int nb, double *out3, hipStream_t st) {
  hipLaunchKernelGGL(sum3_kernel, dim3(1), dim3(1024), 0, st, part, nb, out3);
  return hipGetLastError();
}

/* div[s] = sum_a (theta_a(s) - theta_a(s-a))/dx_a^2 of one level (periodic),
 * the arithmetic of the fused step's stored divergence (ax_march.h). */
template <bool P2>
__global__ void __launch_bounds__(256) level_div_kernel(int nx, int ny, int nz, const double *tx, const double *ty,
                                                        const double *tz, EnergyConsts ek, double *div) {
  const size_t V = (size_t)nx * ny * nz;
  const size_t s = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (s >= V) return;
  const int kk = (int)(s % nz), r = (int)(s / nz), j = r % ny, i = r / ny;
  auto id = [&](int a, int b, int c) { return ((size_t)a * ny + b) * nz + c; };
  div[s] = link_div<P2>(ek, tx[s], tx[id(wrapm(i, nx), j, kk)], ty[s], ty[id(i, wrapm(j, ny), kk)], tz[s],
                        tz[id(i, j, wrapm(kk, nz))]);
}

hipError_t launch_level_div(int nx, int ny, int nz, const double *tx, const double *ty, const double *tz,
                            const EnergyConsts &ek, bool pow2, double *div, hipStream_t st) {
  const size_t V = (size_t)nx * ny * nz;
  const unsigned nb = (unsigned)((V + 255) / 256);
  if (pow2) hipLaunchKernelGGL((level_div_kernel<true>), dim3(nb), dim3(256), 0, st, nx, ny, nz, tx, ty, tz, ek, div);
  else hipLaunchKernelGGL((level_div_kernel<false>), dim3(nb), dim3(256), 0, st, nx, ny, nz, tx, ty, tz, ek, div);
  return hipGetLastError();
}

/* the P2 condition of the diagnostics march: dx, dy, dz, g and the derived
 * divisors are powers of two with finite reciprocals */
bool diag_pow2(double dx, double dy, double dz, double gg) {
  const double dvs[10] = {dx, dy, dz, dx * dx, dy * dy, dz * dz, gg, (gg * dx * dy) * (gg * dx * dy),
                          (gg * dx * dz) * (gg * dx * dz), (gg * dy * dz) * (gg * dy * dz)};
  bool p2 = gg != 0;
  for (double d : dvs) p2 = p2 && is_pow2(fabs(d)) && std::isfinite(1 / d);
  return p2;
}

hipError_t launch_wind_phase(int nx, int ny, int nz, const float *E, int8_t *out, double *part, double *out3,
                             hipStream_t st) {
  /* part holds 3 doubles per work item (diag_part_blocks) */
  const TileGrid g = make_tile_grid(nx, ny, nz, kDiagXseg);
  hipLaunchKernelGGL(wind_march_kernel, dim3(g.work), dim3(64), 0, st, nx, ny, nz, E, g, out, part);
  hipLaunchKernelGGL(sum3_kernel, dim3(1), dim3(1024), 0, st, part, g.work, out3);
  return hipGetLastError();
}

#define AX_INST(T)                                                                                          \
  template hipError_t launch_diag<T>(const DiagConsts &, const Levels<T> &, double *, double *, hipStream_t); \
  template hipError_t launch_winding<T>(int, int, int, const T *, const T *, const T *, const T *, const T *,  \
                                        double, int, int, int8_t *, double *, double *, hipStream_t);        \
  template hipError_t launch_crossings<T>(int, int, int, double, double, double, const T *, const T *, int,   \
                                          int, int *, long long *, long long *, double *, long long,          \
                                          hipStream_t, bool);                                               \
  template hipError_t launch_diag_march<T>(const DiagConsts &, const Levels<T> &, int, float *, double *,     \
                                           double *, hipStream_t);                                          \
  template hipError_t launch_edge_winding<T>(int, int, int, const T *, const T *, const T *, const T *,      \
                                             const T *, double, float *, int8_t *, double *, double *,      \
                                             hipStream_t);
AX_INST(double)
AX_INST(float)
#undef AX_INST

}  // namespace ax
