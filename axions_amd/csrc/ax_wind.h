/*
 * ax_wind.h - per-plaquette winding number of the scalar field at one site
 * (no reference counterpart), shared by the single-lattice and batched
 * diagnostic kernels.
 *
 * For the plaquette spanned by axes a, b at site s:
 *   delta_e = arg( conj(phi(u)) exp(-i q theta_e(u)) phi(u+e) )  per edge,
 *   W = rint( (delta_a(s) + delta_b(s+a) - delta_a(s+b) - delta_b(s) + P) / 2pi )
 * with P = q theta_a(s) + q theta_b(s+a) - q theta_a(s+b) - q theta_b(s) the
 * plaquette angle in the reference's operand order (mpi_evolution.cpp:471-477)
 * and q the field's charge (1 for Abelian-Higgs).  Each
 * delta lies in (-pi, pi], so W is the integer the gauge-invariant phase
 * winds around the plaquette.  All arithmetic is FP64, whatever the storage
 * type T, so FP32 and FP64 states share one definition.
 */
#ifndef AXIONS_AX_WIND_H
#define AXIONS_AX_WIND_H

#include "ax_device.h"

namespace ax {

/* arg( conj(phi(s)) exp(-i theta) phi(s+a) ): the covariant transport of the
 * reference's Laplacian (evolution.cpp:304-311) made gauge invariant. */
__device__ __forceinline__ double link_phase(double cr, double ci, double nr, double ni, double th) {
  double sn, c;
  ax_sincos(th, &sn, &c);
  double wr = c * nr + sn * ni;
  double wi = c * ni - sn * nr;
  double re = cr * wr + ci * wi;
  double im = cr * wi - ci * wr;
  return ax_atan2(im, re);
}

/* Windings of the xy, xz and yz plaquettes spanned from site (i,j,kk). */
template <typename T>
__device__ __forceinline__ void site_winding(int nx, int ny, int nz, const T *fr, const T *fi, const T *tx,
                                             const T *ty, const T *tz, double q, int i, int j, int kk,
                                             int w[3]) {
  const int ip = wrapp(i, nx), jp = wrapp(j, ny), kp = wrapp(kk, nz);
  auto id = [&](int a, int b, int c) { return (a * ny + b) * nz + c; };
  const int s = id(i, j, kk);
  const int sx = id(ip, j, kk), sy = id(i, jp, kk), sz = id(i, j, kp);
  const T *th[3] = {tx, ty, tz};
  const int sa[3] = {sx, sy, sz};
  const int pa[3] = {0, 0, 1}, pb[3] = {1, 2, 2};
  const double c0r = ld(fr, s), c0i = ld(fi, s);
  /* the 9 distinct edges of the 3 plaquettes at s: d0[a] = edge (s, a),
   * d1[a][b] = edge (s+a, b) */
  double pr_a[3], pi_a[3], d0[3], d1[3][3], th0[3], th1[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    pr_a[a] = ld(fr, sa[a]);
    pi_a[a] = ld(fi, sa[a]);
    th0[a] = q * ld(th[a], s);
    d0[a] = link_phase(c0r, c0i, pr_a[a], pi_a[a], th0[a]);
  }
#pragma unroll
  for (int pl = 0; pl < 3; ++pl) {
    const int a = pa[pl], b = pb[pl];
    const int ii = (a == 0 || b == 0) ? ip : i;
    const int jj = (a == 1 || b == 1) ? jp : j;
    const int kq = (a == 2 || b == 2) ? kp : kk;
    const int s_ab = id(ii, jj, kq);
    const double abr = ld(fr, s_ab), abi = ld(fi, s_ab);
    th1[a][b] = q * ld(th[b], sa[a]);
    th1[b][a] = q * ld(th[a], sa[b]);
    d1[a][b] = link_phase(pr_a[a], pi_a[a], abr, abi, th1[a][b]);
    d1[b][a] = link_phase(pr_a[b], pi_a[b], abr, abi, th1[b][a]);
  }
#pragma unroll
  for (int pl = 0; pl < 3; ++pl) {
    const int a = pa[pl], b = pb[pl];
    double P = th0[a] + th1[a][b] - th1[b][a] - th0[b];
    double sum = d0[a] + d1[a][b] - d1[b][a] - d0[b] + P;
    w[pl] = (int)rint(sum * 0.15915494309189535);
  }
}

/* The three forward edge phases delta_a(s), a = x, y, z, of site (i,j,kk)
 * (charge 1): the edges that windings_from_phases combines.  Same
 * arithmetic as link_phase, and as the marching diagnostics that store them
 * with the link's sin/cos computed once. */
template <typename T>
__device__ __forceinline__ void site_edge_phases(int nx, int ny, int nz, const T *fr, const T *fi, const T *tx,
                                                 const T *ty, const T *tz, int i, int j, int kk, double d[3]) {
  auto id = [&](int a, int b, int c) { return (a * ny + b) * nz + c; };
  const int s = id(i, j, kk);
  const int sa[3] = {id(wrapp(i, nx), j, kk), id(i, wrapp(j, ny), kk), id(i, j, wrapp(kk, nz))};
  const T *th[3] = {tx, ty, tz};
  const double cr = ld(fr, s), ci = ld(fi, s);
#pragma unroll
  for (int a = 0; a < 3; ++a) d[a] = link_phase(cr, ci, ld(fr, sa[a]), ld(fi, sa[a]), 1.0 * ld(th[a], s));
}

/* Windings of the xy, xz, yz plaquettes at site s = (i,j,kk) from stored
 * forward edge phases D[a] (V doubles each): the arithmetic of site_winding
 * with q = 1, phases and angles in its order, so the results are its bits. */
template <typename T>
__device__ __forceinline__ void windings_from_phases(int nx, int ny, int nz, const T *tx, const T *ty, const T *tz,
                                                     const double *phase, int i, int j, int kk, int w[3]) {
  auto id = [&](int a, int b, int c) { return (a * ny + b) * nz + c; };
  const size_t V = (size_t)nx * ny * nz;
  const int s = id(i, j, kk);
  const int sa[3] = {id(wrapp(i, nx), j, kk), id(i, wrapp(j, ny), kk), id(i, j, wrapp(kk, nz))};
  const T *th[3] = {tx, ty, tz};
  const double *D[3] = {phase, phase + V, phase + 2 * V};
  const int pa[3] = {0, 0, 1}, pb[3] = {1, 2, 2};
#pragma unroll
  for (int pl = 0; pl < 3; ++pl) {
    const int a = pa[pl], b = pb[pl];
    const double th0a = 1.0 * ld(th[a], s), th0b = 1.0 * ld(th[b], s);
    const double th1ab = 1.0 * ld(th[b], sa[a]), th1ba = 1.0 * ld(th[a], sa[b]);
    const double P = th0a + th1ab - th1ba - th0b;
    const double sum = D[a][s] + D[b][sa[a]] - D[a][sa[b]] - D[b][s] + P;
    w[pl] = (int)rint(sum * 0.15915494309189535);
  }
}

}  // namespace ax

#endif
