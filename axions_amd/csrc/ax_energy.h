/*
 * ax_energy.h - the per-site energy density and Gauss-law residual of the
 * Abelian-Higgs model (evolution.cpp:386-427, energy form with E_i =
 * dtheta/(g dt dx_i), g != 0), in FP64, shared by the marching diagnostics
 * pass (ax_diag.hip) and the step that takes the diagnostics of its own
 * current level (ax_march.h), so that both give the same per-site bits.
 *
 * Operand order and association are the reference's: vel_c = (phi_c -
 * phi_c^{n-1})/dt, the backward covariant differences (phi_c(s) -
 * cos theta_a(s-a) phi_c(s-a) + sigma_c sin theta_a(s-a) phi_cbar(s-a))/dx_a,
 * div^m = sum_a (theta_a^m(s) - theta_a^m(s-a))/dx_a^2, G = ((div^n -
 * div^{n-1})/dt)/g - 2 g (phi_i vel_r - phi_r vel_i), the electric and
 * magnetic terms, rho = dx dy dz (sum of squares + ... + lambda/4 (|phi|^2 -
 * eta^2)^2).  P2: dx, dy, dz and g are powers of two, so each division by
 * dx, dx^2, g or (g dx dy)^2 is the multiplication by its exact reciprocal
 * (the same correctly rounded value).  RECIP: the divisions by dt and
 * g dt dx_a are multiplications by the rounded reciprocal (at most an ulp
 * per term; the sums are held to 1e-12).
 */
#ifndef AXIONS_AX_ENERGY_H
#define AXIONS_AX_ENERGY_H

#include "ax_device.h"

namespace ax {

struct EnergyConsts {
  double dx, dy, dz, dt, g, lambda, eta;
  double dxx, dyy, dzz, bxyd, bxzd, byzd;
  double idx, idy, idz, idxx, idyy, idzz, igg, ibxy, ibxz, ibyz;
  double idt, iedx, iedy, iedz;
};

__host__ __device__ inline EnergyConsts energy_consts(double dx, double dy, double dz, double dt, double g,
                                                      double lambda, double eta) {
  EnergyConsts k;
  k.dx = dx; k.dy = dy; k.dz = dz; k.dt = dt; k.g = g; k.lambda = lambda; k.eta = eta;
  k.dxx = dx * dx; k.dyy = dy * dy; k.dzz = dz * dz;
  k.bxyd = (g * dx * dy) * (g * dx * dy);
  k.bxzd = (g * dx * dz) * (g * dx * dz);
  k.byzd = (g * dy * dz) * (g * dy * dz);
  k.idx = 1 / dx; k.idy = 1 / dy; k.idz = 1 / dz;
  k.idxx = 1 / k.dxx; k.idyy = 1 / k.dyy; k.idzz = 1 / k.dzz;
  k.igg = 1 / g; k.ibxy = 1 / k.bxyd; k.ibxz = 1 / k.bxzd; k.ibyz = 1 / k.byzd;
  k.idt = 1 / dt;
  k.iedx = 1 / (g * dt * dx); k.iedy = 1 / (g * dt * dy); k.iedz = 1 / (g * dt * dz);
  return k;
}

template <bool P2>
__device__ __forceinline__ double edv(double a, double d, double inv) {
  return P2 ? a * inv : a / d;
}

/* div^m(s) from the level's links at s and at s - a */
template <bool P2>
__device__ __forceinline__ double link_div(const EnergyConsts &k, double X, double Xm, double Y, double Ym, double Z,
                                           double Zm) {
  return edv<P2>(X - Xm, k.dxx, k.idxx) + edv<P2>(Y - Ym, k.dyy, k.idyy) + edv<P2>(Z - Zm, k.dzz, k.idzz);
}

/* cos of the three plaquette angles at s (xy, xz, yz); the short k = 0 form
 * when every lane's angles qualify (bit-identical to ax_cos there).  Every
 * lane of the wave must call it. */
__device__ __forceinline__ void plaquette_cos(double pxy, double pxz, double pyz, double &cxy, double &cxz,
                                              double &cyz) {
  if (__all(ax_sin_is_small(pxy) && ax_sin_is_small(pxz) && ax_sin_is_small(pyz))) {
    cxy = ax_cos_kernel(pxy, 0.0); cxz = ax_cos_kernel(pxz, 0.0); cyz = ax_cos_kernel(pyz, 0.0);
  } else {
    cxy = ax_cos(pxy); cxz = ax_cos(pxz); cyz = ax_cos(pyz);
  }
}

/* Energy density rho and Gauss residual G at s.  Inputs by name: phi (pr,
 * pi) and links (X, Y, Z) at s, both levels; phi at s - a with the backward
 * link's cos/sin; the divergences of both levels; the plaquette cosines. */
template <bool P2, bool RECIP>
__device__ __forceinline__ void energy_gauss(const EnergyConsts &k, double pr, double pi, double oPR, double oPI,
                                             double X, double Y, double Z, double oX, double oY, double oZ,
                                             double PRm, double PIm, double cxm, double sxm, double PRym,
                                             double PIym, double cym, double sym, double PRzm, double PIzm,
                                             double czm, double szm, double divN, double divP, double cxy,
                                             double cxz, double cyz, double &rho, double &G) {
  auto by_dt = [&](double a) { return RECIP ? a * k.idt : a / k.dt; };
  auto by_e = [&](double a, double d, double inv) { return RECIP ? a * inv : a / (k.g * k.dt * d); };
  double vel[2], gx[2], gy[2], gz[2];
  vel[0] = by_dt(pr - oPR);
  gx[0] = edv<P2>(pr - cxm * PRm + sxm * PIm, k.dx, k.idx);
  gy[0] = edv<P2>(pr - cym * PRym + sym * PIym, k.dy, k.idy);
  gz[0] = edv<P2>(pr - czm * PRzm + szm * PIzm, k.dz, k.idz);
  vel[1] = by_dt(pi - oPI);
  gx[1] = edv<P2>(pi - cxm * PIm + (-sxm) * PRm, k.dx, k.idx);
  gy[1] = edv<P2>(pi - cym * PIym + (-sym) * PRym, k.dy, k.idy);
  gz[1] = edv<P2>(pi - czm * PIzm + (-szm) * PRzm, k.dz, k.idz);
  const double mag2 = pr * pr + pi * pi;
  const double vv = mag2 - k.eta * k.eta;
  const double divT = by_dt(divN - divP);
  G = edv<P2>(divT, k.g, k.igg) - 2 * k.g * (pi * vel[0] - pr * vel[1]);
  const double ex = by_e(X - oX, k.dx, k.iedx), ey = by_e(Y - oY, k.dy, k.iedy), ez = by_e(Z - oZ, k.dz, k.iedz);
  const double tdc = 0.5 * (ex * ex + ey * ey + ez * ez);
  const double bxy = edv<P2>(1 - cxy, k.bxyd, k.ibxy);
  const double bxz = edv<P2>(1 - cxz, k.bxzd, k.ibxz);
  const double byz = edv<P2>(1 - cyz, k.byzd, k.ibyz);
  const double fcont = bxy + bxz + byz;
  const double r_ = vel[0] * vel[0] + vel[1] * vel[1] + gx[0] * gx[0] + gx[1] * gx[1] + gy[0] * gy[0] +
                    gy[1] * gy[1] + gz[0] * gz[0] + gz[1] * gz[1] + tdc + fcont + 0.25 * k.lambda * (vv * vv);
  rho = k.dx * k.dy * k.dz * r_;
}

/* arg( conj(phi(s)) exp(-i theta) phi(s+a) ) with the link's sin/cos at
 * hand: the forward edge phase of the winding detector (ax_wind.h) */
__device__ __forceinline__ double link_phase_sc(double cr, double ci, double nr, double ni, double sn, double c) {
  double wr = c * nr + sn * ni;
  double wi = c * ni - sn * nr;
  double re = cr * wr + ci * wi;
  double im = cr * wi - ci * wr;
  return ax_atan2(im, re);
}

}  // namespace ax

#endif
