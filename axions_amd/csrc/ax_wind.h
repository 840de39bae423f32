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
 * type T, so FP32 and FP64 states share one definition.  The marching
 * passes at the end of this file store FP32 edge values e = delta + q theta
 * and form the same integers from four of them.
 */
#ifndef AXIONS_AX_WIND_H
#define AXIONS_AX_WIND_H

#include "ax_device.h"
#include "ax_march.h"   /* ldo */

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

/* One 64-lane wave per (64-site z chunk zc, row y), marching along x over
 * the planes [x0, x1): lane l holds z = 64 zc + l (lanes past the row end
 * load the row's last site and emit nothing).  Plane x+1's values are loaded
 * once, as the next plane's centre, and carried; the y+1 and z+1 operands
 * are rows that neighbouring waves load too.  Plane addresses are formed on
 * the scalar unit (64-bit), row and lane offsets are 32-bit bytes. */
struct WindRow {
  int nx, ny, nz, y, z;
  bool act;
  size_t nynz, V;
  uint32_t r0, rp, z0, zp;   /* element offsets: row y, row y+1, lane z, z+1 */
  __device__ WindRow(int nx_, int ny_, int nz_, int y_, int zc) : nx(nx_), ny(ny_), nz(nz_), y(y_) {
    const int zr = zc * 64 + (int)(threadIdx.x & 63);
    act = zr < nz;
    z = act ? zr : nz - 1;
    nynz = (size_t)ny * nz;
    V = (size_t)nx * nynz;
    r0 = (uint32_t)(y * nz);
    rp = (uint32_t)((y == ny - 1 ? 0 : y + 1) * nz);
    z0 = (uint32_t)z;
    zp = (uint32_t)(z == nz - 1 ? 0 : z + 1);
  }
  /* field b (element type E) in x plane x at element offsets row, lane */
  template <typename E>
  __device__ double at(const E *b, int x, uint32_t row, uint32_t lane) const {
    return (double)ldo(b + (size_t)x * nynz, row * (uint32_t)sizeof(E), lane * (uint32_t)sizeof(E));
  }
  __device__ size_t site(int x) const { return ((size_t)x * ny + y) * nz + z; }
};

/* The marching passes store, per edge, e_a(s) = delta_a(s) + theta_a(s)
 * rounded to FP32, and form W from four of them:
 *   W = rint( (e_a(s) + e_b(s+a) - e_a(s+b) - e_b(s)) / 2pi ).
 * The sum is site_winding's sum regrouped, and W is the same integer.  Each
 * delta is link_phase's own value (the same FP64 operations as
 * site_winding), which lies within a few ulp of the edge's exact phase up
 * to a multiple of 2 pi; the exact phases of phi telescope around the
 * plaquette, so site_winding's sum lies within rounding of some 2 pi K.
 * The regrouped sum differs from it by rounding only -- half an ulp of
 * |delta + theta| per FP32 e, at most 2^-24 |e| -- far below pi for any
 * link angle below 2^20, so both round to K.  (Non-finite fields are
 * counted separately and fail the run.)  Computing delta any other way
 * (another atan2, FP32) would not keep this: an edge phase near +-pi may
 * then land on the other branch, 2 pi away.  FP32 e halves the bytes the
 * step writes and the winding pass reads. */

/* The edge values e_a(s) = fl32(link_phase(phi(s), phi(s+a), q theta_a(s)) +
 * q theta_a(s)), a = x, y, z, of the sites of one wave's march into
 * E[a V + s]: the edges that wind_march combines.  The arithmetic of
 * link_phase and site_winding (q times the link, q = 1 for Abelian-Higgs),
 * as the marching diagnostics that store them with the link's sin/cos
 * computed once.  Returns the number of the lane's sites holding a
 * non-finite value in phi or a link. */
template <typename T>
__device__ __forceinline__ int phase_march(const WindRow &w, const T *fr, const T *fi, const T *tx, const T *ty,
                                           const T *tz, float *E, int x0, int x1, double q = 1.0) {
  double cr = w.at(fr, x0, w.r0, w.z0), ci = w.at(fi, x0, w.r0, w.z0);
  int bad = 0;
  for (int x = x0; x < x1; ++x) {
    const int xp = x == w.nx - 1 ? 0 : x + 1;
    const double nr = w.at(fr, xp, w.r0, w.z0), ni = w.at(fi, xp, w.r0, w.z0);
    const double yr = w.at(fr, x, w.rp, w.z0), yi = w.at(fi, x, w.rp, w.z0);
    const double zr = w.at(fr, x, w.r0, w.zp), zi = w.at(fi, x, w.r0, w.zp);
    const double X = q * w.at(tx, x, w.r0, w.z0), Y = q * w.at(ty, x, w.r0, w.z0), Z = q * w.at(tz, x, w.r0, w.z0);
    const double dX = link_phase(cr, ci, nr, ni, X);
    const double dY = link_phase(cr, ci, yr, yi, Y);
    const double dZ = link_phase(cr, ci, zr, zi, Z);
    if (w.act) {
      const size_t s = w.site(x);
      E[s] = (float)(dX + X);
      E[w.V + s] = (float)(dY + Y);
      E[2 * w.V + s] = (float)(dZ + Z);
      bad += (!isfinite(cr) || !isfinite(ci) || !isfinite(X) || !isfinite(Y) || !isfinite(Z)) ? 1 : 0;
    }
    cr = nr;
    ci = ni;
  }
  return bad;
}

/* Windings of the xy, xz, yz plaquettes of one wave's march from the stored
 * edge values E[a V + s] = e_a(s) of the same level (above).  emit(s, w) is
 * called for each of the lane's sites. */
template <typename Emit>
__device__ __forceinline__ void wind_march(const WindRow &w, const float *E, int x0, int x1, Emit emit) {
  const float *EX = E, *EY = E + w.V, *EZ = E + 2 * w.V;
  double cX = w.at(EX, x0, w.r0, w.z0), cY = w.at(EY, x0, w.r0, w.z0), cZ = w.at(EZ, x0, w.r0, w.z0);
  for (int x = x0; x < x1; ++x) {
    const int xp = x == w.nx - 1 ? 0 : x + 1;
    /* plane x+1 (next centre), row y+1, site z+1 */
    const double nX = w.at(EX, xp, w.r0, w.z0), nY = w.at(EY, xp, w.r0, w.z0), nZ = w.at(EZ, xp, w.r0, w.z0);
    const double yX = w.at(EX, x, w.rp, w.z0), yZ = w.at(EZ, x, w.rp, w.z0);
    const double zX = w.at(EX, x, w.r0, w.zp), zY = w.at(EY, x, w.r0, w.zp);
    int wn[3];
    wn[0] = (int)rint((cX + nY - yX - cY) * 0.15915494309189535);   /* (a, b) = (x, y) */
    wn[1] = (int)rint((cX + nZ - zX - cZ) * 0.15915494309189535);   /* (x, z) */
    wn[2] = (int)rint((cY + yZ - zY - cZ) * 0.15915494309189535);   /* (y, z) */
    if (w.act) emit(w.site(x), wn);
    cX = nX; cY = nY; cZ = nZ;
  }
}

}  // namespace ax

#endif
