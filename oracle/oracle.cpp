/*
 * oracle.cpp - TEST INFRASTRUCTURE ONLY.  CPU restatement of the reference's
 * lattice evolution, used as the checker for the HIP library.  Only tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py may load it;
 * the product (axions_amd/) never links or calls anything under oracle/.
 *
 * Every arithmetic expression keeps the reference's operand order and
 * association (the file must be built with -ffp-contract=off), so results
 * are bit-identical to the reference given the same sin/cos.  pow(x,2) in
 * the reference is written x*x here (GCC folds pow(x,2.0) to x*x).
 *
 * Sources restated (file:line in sjcott/Axions):
 *   mt19937 + uniform_real_distribution IC ... mpi_evolution.cpp:287-312
 *   AH fused in-place step ................... mpi_evolution.cpp:381-501
 *   AH two-pass theta friction form .......... evolution.cpp:371-373
 *   AH energy / Gauss (OpenMP form) .......... evolution.cpp:298-427, 1404-1408
 *   AH energy (MPI form) ..................... mpi_evolution.cpp:509-532
 *   face crossings ........................... mpi_evolution.cpp:536-789
 *   Moore step, energy, Gauss ................ mooreEvol.cpp:184-395, 647-697
 *   periodic neighbour indices ............... mpi_evolution.cpp:385-405
 *   global string, momentum-form leapfrog .... leapfrog.cpp:166-347
 *   string length by chaining crossings ...... ic_w_evolution.cpp:961-1087
 *   straight-string profile (SOR) ............ singleSemiAnalytic.cpp:54-134
 *   straight-string lattice .................. straight_ic.cpp:59-146
 * The per-plaquette winding number has no reference counterpart.
 *
 * Pinning (what checks this file, see tests/): compiled reference runs in
 * tests/golden (bit-exact state hashes and sampled sites for AH g=0 and g=1,
 * for Moore with fixed edges and for the global leapfrog), the std::mt19937
 * known-answer value, the
 * libstdc++ <random> routines themselves, and invariants (vacuum, lattice
 * translation, gauge covariance, straight string winding, cube sums).
 */
#ifndef _GNU_SOURCE
#define _GNU_SOURCE
#endif
#include <math.h>
#include <stdint.h>
#include <stddef.h>
#include <string.h>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "axions/axions.h"
#include "axions/ax_trig.h"

namespace {

struct TrigAx {
  static double sn(double x) { return ax_sin(x); }
  static double cs(double x) { return ax_cos(x); }
};
struct TrigLibm {
  static double sn(double x) { return ::sin(x); }
  static double cs(double x) { return ::cos(x); }
};

/* ---------------- mt19937 (C++ [rand.eng.mers], 32-bit parameters) -------- */
struct MT {
  uint32_t s[624];
  int idx;
  explicit MT(uint32_t seed) {
    s[0] = seed;
    for (int i = 1; i < 624; ++i)
      s[i] = 1812433253u * (s[i - 1] ^ (s[i - 1] >> 30)) + (uint32_t)i;
    idx = 624;
  }
  void twist() {
    for (int i = 0; i < 624; ++i) {
      uint32_t y = (s[i] & 0x80000000u) | (s[(i + 1) % 624] & 0x7fffffffu);
      uint32_t v = s[(i + 397) % 624] ^ (y >> 1);
      if (y & 1u) v ^= 0x9908b0dfu;
      s[i] = v;
    }
    idx = 0;
  }
  uint32_t next() {
    if (idx >= 624) twist();
    uint32_t y = s[idx++];
    y ^= (y >> 11);
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= (y >> 18);
    return y;
  }
  /* generate_canonical<double,53> with a 32-bit engine: two draws,
   * sum = d0 + d1*2^32 (rounded), result sum/2^64, clamped below 1. */
  double canonical() {
    double sum = 0.0, tmp = 1.0;
    for (int k = 0; k < 2; ++k) {
      sum += (double)next() * tmp;
      tmp *= 4294967296.0;
    }
    double r = sum / tmp;
    if (r >= 1.0) r = nextafter(1.0, 0.0);
    return r;
  }
};

struct Lat {
  int nx, ny, nz;
  size_t V;
  explicit Lat(const ax_params *p) : nx(p->nx), ny(p->ny), nz(p->nz) {
    V = (size_t)nx * ny * nz;
  }
  size_t id(int i, int j, int k) const { return ((size_t)i * ny + j) * nz + k; }
};

inline double *fld(double *st, const Lat &L, int comp, int lev) {
  return st + ((size_t)comp * 2 + lev) * L.V;
}
inline const double *fld(const double *st, const Lat &L, int comp, int lev) {
  return st + ((size_t)comp * 2 + lev) * L.V;
}

/* neighbour indices of site (i,j,k); periodic wrap as mpi_evolution.cpp:385-405 */
struct Nb {
  int i, j, k, ip, im, jp, jm, kp, km;
  Nb(const Lat &L, int i_, int j_, int k_) : i(i_), j(j_), k(k_) {
    ip = (i + 1) % L.nx; im = (i - 1 + L.nx) % L.nx;
    jp = (j + 1) % L.ny; jm = (j - 1 + L.ny) % L.ny;
    kp = (k + 1) % L.nz; km = (k - 1 + L.nz) % L.nz;
  }
};

/* ------------------------------ AH step ---------------------------------- */
template <class T>
void ah_site(const ax_params *p, const Lat &L, double *st, int tN, int tP, const Nb &n,
             double damp, double fric) {
  const double *pr = fld(st, L, 0, tN), *pi = fld(st, L, 1, tN);
  const double *tx = fld(st, L, 2, tN), *ty = fld(st, L, 3, tN), *tz = fld(st, L, 4, tN);
  double *prP = fld(st, L, 0, tP), *piP = fld(st, L, 1, tP);
  double *txP = fld(st, L, 2, tP), *tyP = fld(st, L, 3, tP), *tzP = fld(st, L, 4, tP);
  const size_t s = L.id(n.i, n.j, n.k);
  const size_t sxp = L.id(n.ip, n.j, n.k), sxm = L.id(n.im, n.j, n.k);
  const size_t syp = L.id(n.i, n.jp, n.k), sym = L.id(n.i, n.jm, n.k);
  const size_t szp = L.id(n.i, n.j, n.kp), szm = L.id(n.i, n.j, n.km);
  const size_t s_xp_ym = L.id(n.ip, n.jm, n.k), s_xp_zm = L.id(n.ip, n.j, n.km);
  const size_t s_xm_yp = L.id(n.im, n.jp, n.k), s_yp_zm = L.id(n.i, n.jp, n.km);
  const size_t s_xm_zp = L.id(n.im, n.j, n.kp), s_ym_zp = L.id(n.i, n.jm, n.kp);
  const double dx = p->dx, dy = p->dy, dz = p->dz, dt = p->dt;
  const double *ph[2] = {pr, pi};
  double *phP[2] = {prP, piP};
  const double sg[2] = {1.0, -1.0};

  double mag2 = pr[s] * pr[s] + pi[s] * pi[s];
  for (int c = 0; c < 2; ++c) {
    const double *a = ph[c], *b = ph[1 - c];
    double lxx = (T::cs(tx[s]) * a[sxp] + sg[c] * T::sn(tx[s]) * b[sxp] - 2 * a[s] +
                  T::cs(tx[sxm]) * a[sxm] - sg[c] * T::sn(tx[sxm]) * b[sxm]) / (dx * dx);
    double lyy = (T::cs(ty[s]) * a[syp] + sg[c] * T::sn(ty[s]) * b[syp] - 2 * a[s] +
                  T::cs(ty[sym]) * a[sym] - sg[c] * T::sn(ty[sym]) * b[sym]) / (dy * dy);
    double lzz = (T::cs(tz[s]) * a[szp] + sg[c] * T::sn(tz[s]) * b[szp] - 2 * a[s] +
                  T::cs(tz[szm]) * a[szm] - sg[c] * T::sn(tz[szm]) * b[szm]) / (dz * dz);
    double vel = (a[s] - phP[c][s]) / dt;
    double acc = lxx + lyy + lzz - 0.5 * p->lambda * (mag2 - p->eta * p->eta) * a[s] - fric * vel;
    phP[c][s] = 2 * a[s] - phP[c][s] + dt * dt * acc;
  }
  double jx = T::cs(tx[s]) * (pi[s] * pr[sxp] - pr[s] * pi[sxp]) +
              T::sn(tx[s]) * (pr[s] * pr[sxp] + pi[s] * pi[sxp]);
  double jy = T::cs(ty[s]) * (pi[s] * pr[syp] - pr[s] * pi[syp]) +
              T::sn(ty[s]) * (pr[s] * pr[syp] + pi[s] * pi[syp]);
  double jz = T::cs(tz[s]) * (pi[s] * pr[szp] - pr[s] * pi[szp]) +
              T::sn(tz[s]) * (pr[s] * pr[szp] + pi[s] * pi[szp]);
  double fxy = (T::sn(tx[s] + ty[sxp] - tx[syp] - ty[s]) -
                T::sn(tx[sym] + ty[s_xp_ym] - tx[s] - ty[sym])) / (dy * dy);
  double fxz = (T::sn(tx[s] + tz[sxp] - tx[szp] - tz[s]) -
                T::sn(tx[szm] + tz[s_xp_zm] - tx[s] - tz[szm])) / (dz * dz);
  double fyx = (T::sn(ty[s] + tx[syp] - ty[sxp] - tx[s]) -
                T::sn(ty[sxm] + tx[s_xm_yp] - ty[s] - tx[sxm])) / (dx * dx);
  double fyz = (T::sn(ty[s] + tz[syp] - ty[szp] - tz[s]) -
                T::sn(ty[szm] + tz[s_yp_zm] - ty[s] - tz[szm])) / (dz * dz);
  double fzx = (T::sn(tz[s] + tx[szp] - tz[sxp] - tx[s]) -
                T::sn(tz[sxm] + tx[s_xm_zp] - tz[s] - tx[sxm])) / (dx * dx);
  double fzy = (T::sn(tz[s] + ty[szp] - tz[syp] - ty[s]) -
                T::sn(tz[sym] + ty[s_ym_zp] - tz[s] - ty[sym])) / (dy * dy);
  const double g = p->g;
  double ax_, ay_, az_;
  if (p->theta_damp_form == 0) {
    double vx = (tx[s] - txP[s]) / dt, vy = (ty[s] - tyP[s]) / dt, vz = (tz[s] - tzP[s]) / dt;
    ax_ = -2 * g * g * jx - fxy - fxz - damp * vx;
    ay_ = -2 * g * g * jy - fyx - fyz - damp * vy;
    az_ = -2 * g * g * jz - fzx - fzy - damp * vz;
  } else {
    ax_ = -2 * g * g * jx - fxy - fxz - damp * (tx[s] - txP[s]) / dt;
    ay_ = -2 * g * g * jy - fyx - fyz - damp * (ty[s] - tyP[s]) / dt;
    az_ = -2 * g * g * jz - fzx - fzy - damp * (tz[s] - tzP[s]) / dt;
  }
  txP[s] = 2 * tx[s] - txP[s] + dt * dt * ax_;
  tyP[s] = 2 * ty[s] - tyP[s] + dt * dt * ay_;
  tzP[s] = 2 * tz[s] - tzP[s] + dt * dt * az_;
}

/* ------------------------------ Moore step ------------------------------- */
template <class T>
void moore_site(const ax_params *p, const Lat &L, double *st, int tN, int tP, const Nb &n) {
  const double *f1[2] = {fld(st, L, 0, tN), fld(st, L, 1, tN)};
  const double *f2[2] = {fld(st, L, 2, tN), fld(st, L, 3, tN)};
  const double *tx = fld(st, L, 4, tN), *ty = fld(st, L, 5, tN), *tz = fld(st, L, 6, tN);
  double *f1P[2] = {fld(st, L, 0, tP), fld(st, L, 1, tP)};
  double *f2P[2] = {fld(st, L, 2, tP), fld(st, L, 3, tP)};
  double *txP = fld(st, L, 4, tP), *tyP = fld(st, L, 5, tP), *tzP = fld(st, L, 6, tP);
  const size_t s = L.id(n.i, n.j, n.k);
  const size_t sxp = L.id(n.ip, n.j, n.k), sxm = L.id(n.im, n.j, n.k);
  const size_t syp = L.id(n.i, n.jp, n.k), sym = L.id(n.i, n.jm, n.k);
  const size_t szp = L.id(n.i, n.j, n.kp), szm = L.id(n.i, n.j, n.km);
  const size_t s_xp_ym = L.id(n.ip, n.jm, n.k), s_xp_zm = L.id(n.ip, n.j, n.km);
  const size_t s_xm_yp = L.id(n.im, n.jp, n.k), s_yp_zm = L.id(n.i, n.jp, n.km);
  const size_t s_xm_zp = L.id(n.im, n.j, n.kp), s_ym_zp = L.id(n.i, n.jm, n.kp);
  const double dx = p->dx, dy = p->dy, dz = p->dz, dt = p->dt;
  const int q1 = p->q1, q2 = p->q2;
  const double sg[2] = {1.0, -1.0};
  double m1 = f1[0][s] * f1[0][s] + f1[1][s] * f1[1][s];
  double m2 = f2[0][s] * f2[0][s] + f2[1][s] * f2[1][s];
  double acc1[2], acc2[2];
  for (int c = 0; c < 2; ++c) {
    const double *a = f1[c], *b = f1[1 - c], *u = f2[c], *w = f2[1 - c];
    double l1x = (T::cs(q1 * tx[s]) * a[sxp] + sg[c] * T::sn(q1 * tx[s]) * b[sxp] - 2 * a[s] +
                  T::cs(q1 * tx[sxm]) * a[sxm] - sg[c] * T::sn(q1 * tx[sxm]) * b[sxm]) / (dx * dx);
    double l2x = (T::cs(q2 * tx[s]) * u[sxp] + sg[c] * T::sn(q2 * tx[s]) * w[sxp] - 2 * u[s] +
                  T::cs(q2 * tx[sxm]) * u[sxm] - sg[c] * T::sn(q2 * tx[sxm]) * w[sxm]) / (dx * dx);
    double l1y = (T::cs(q1 * ty[s]) * a[syp] + sg[c] * T::sn(q1 * ty[s]) * b[syp] - 2 * a[s] +
                  T::cs(q1 * ty[sym]) * a[sym] - sg[c] * T::sn(q1 * ty[sym]) * b[sym]) / (dy * dy);
    double l2y = (T::cs(q2 * ty[s]) * u[syp] + sg[c] * T::sn(q2 * ty[s]) * w[syp] - 2 * u[s] +
                  T::cs(q2 * ty[sym]) * u[sym] - sg[c] * T::sn(q2 * ty[sym]) * w[sym]) / (dy * dy);
    double l1z = (T::cs(q1 * tz[s]) * a[szp] + sg[c] * T::sn(q1 * tz[s]) * b[szp] - 2 * a[s] +
                  T::cs(q1 * tz[szm]) * a[szm] - sg[c] * T::sn(q1 * tz[szm]) * b[szm]) / (dz * dz);
    double l2z = (T::cs(q2 * tz[s]) * u[szp] + sg[c] * T::sn(q2 * tz[s]) * w[szp] - 2 * u[s] +
                  T::cs(q2 * tz[szm]) * u[szm] - sg[c] * T::sn(q2 * tz[szm]) * w[szm]) / (dz * dz);
    acc1[c] = l1x + l1y + l1z -
              (0.5 * p->lambda1 * (m1 - p->eta1 * p->eta1) + 0.5 * p->lambda12 * (m2 - p->eta2 * p->eta2)) * a[s];
    acc2[c] = l2x + l2y + l2z -
              (0.5 * p->lambda2 * (m2 - p->eta2 * p->eta2) + 0.5 * p->lambda12 * (m1 - p->eta1 * p->eta1)) * u[s];
  }
  const double *r1 = f1[0], *i1 = f1[1], *r2 = f2[0], *i2 = f2[1];
  double j1x = q1 * (T::cs(q1 * tx[s]) * (i1[s] * r1[sxp] - r1[s] * i1[sxp]) +
                     T::sn(q1 * tx[s]) * (r1[s] * r1[sxp] + i1[s] * i1[sxp]));
  double j1y = q1 * (T::cs(q1 * ty[s]) * (i1[s] * r1[syp] - r1[s] * i1[syp]) +
                     T::sn(q1 * ty[s]) * (r1[s] * r1[syp] + i1[s] * i1[syp]));
  double j1z = q1 * (T::cs(q1 * tz[s]) * (i1[s] * r1[szp] - r1[s] * i1[szp]) +
                     T::sn(q1 * tz[s]) * (r1[s] * r1[szp] + i1[s] * i1[szp]));
  double j2x = q2 * (T::cs(q2 * tx[s]) * (i2[s] * r2[sxp] - r2[s] * i2[sxp]) +
                     T::sn(q2 * tx[s]) * (r2[s] * r2[sxp] + i2[s] * i2[sxp]));
  double j2y = q2 * (T::cs(q2 * ty[s]) * (i2[s] * r2[syp] - r2[s] * i2[syp]) +
                     T::sn(q2 * ty[s]) * (r2[s] * r2[syp] + i2[s] * i2[syp]));
  double j2z = q2 * (T::cs(q2 * tz[s]) * (i2[s] * r2[szp] - r2[s] * i2[szp]) +
                     T::sn(q2 * tz[s]) * (r2[s] * r2[szp] + i2[s] * i2[szp]));
  double fxy = (T::sn(tx[s] + ty[sxp] - tx[syp] - ty[s]) -
                T::sn(tx[sym] + ty[s_xp_ym] - tx[s] - ty[sym])) / (dy * dy);
  double fxz = (T::sn(tx[s] + tz[sxp] - tx[szp] - tz[s]) -
                T::sn(tx[szm] + tz[s_xp_zm] - tx[s] - tz[szm])) / (dz * dz);
  double fyx = (T::sn(ty[s] + tx[syp] - ty[sxp] - tx[s]) -
                T::sn(ty[sxm] + tx[s_xm_yp] - ty[s] - tx[sxm])) / (dx * dx);
  double fyz = (T::sn(ty[s] + tz[syp] - ty[szp] - tz[s]) -
                T::sn(ty[szm] + tz[s_yp_zm] - ty[s] - tz[szm])) / (dz * dz);
  double fzx = (T::sn(tz[s] + tx[szp] - tz[sxp] - tx[s]) -
                T::sn(tz[sxm] + tx[s_xm_zp] - tz[s] - tx[sxm])) / (dx * dx);
  double fzy = (T::sn(tz[s] + ty[szp] - tz[syp] - ty[s]) -
                T::sn(tz[sym] + ty[s_ym_zp] - tz[s] - ty[sym])) / (dy * dy);
  const double g = p->g;
  double atx = -2 * g * g * (j1x + j2x) - fxy - fxz;
  double aty = -2 * g * g * (j1y + j2y) - fyx - fyz;
  double atz = -2 * g * g * (j1z + j2z) - fzx - fzy;
  for (int c = 0; c < 2; ++c) {
    f1P[c][s] = 2 * f1[c][s] - f1P[c][s] + dt * dt * acc1[c];
    f2P[c][s] = 2 * f2[c][s] - f2P[c][s] + dt * dt * acc2[c];
  }
  txP[s] = 2 * tx[s] - txP[s] + dt * dt * atx;
  tyP[s] = 2 * ty[s] - tyP[s] + dt * dt * aty;
  tzP[s] = 2 * tz[s] - tzP[s] + dt * dt * atz;
}

/* ----------------------- non-periodic boundaries ------------------------
 * R: storage/arithmetic type; T: sin/cos of R.  The reference's loops and
 * expressions: Neumann faces evolution.cpp:192-265 (Moore: q-scaled links,
 * mooreEvol.cpp:114-168); absorbing x/y faces and corners
 * evolution.cpp:449-689. */
template <class R, class T>
void bc_neumann(const ax_params *p, const Lat &L, R *st, int tN) {
  const bool moore = p->model == AX_MODEL_MOORE;
  const int t0 = moore ? 4 : 2, nsc = moore ? 2 : 1;
  const int c[2] = {1, -1};
  auto TH = [&](int a, int i, int j, int k) -> R & { return st[((size_t)(t0 + a) * 2 + tN) * L.V + L.id(i, j, k)]; };
  auto PH = [&](int sc, int comp, int i, int j, int k) -> R & {
    return st[((size_t)(2 * sc + comp) * 2 + tN) * L.V + L.id(i, j, k)];
  };
  auto ang = [&](int sc, R th) -> R { return moore ? (R)(sc == 0 ? p->q1 : p->q2) * th : th; };
  const bool zper = p->boundary_z == AX_BC_PERIODIC;
  const int zEdge1 = zper ? 0 : 1, zEdge2 = zper ? L.nz : L.nz - 1;
  if (p->boundary == AX_BC_NEUMANN) {
    for (int i = 1; i < L.nx - 1; ++i)
      for (int k = zEdge1; k < zEdge2; ++k) {
        for (int comp = 0; comp < 3; ++comp) {
          TH(comp, i, 0, k) = TH(comp, i, 1, k);
          TH(comp, i, L.ny - 1, k) = TH(comp, i, L.ny - 2, k);
        }
        for (int comp = 0; comp < 2; ++comp)
          for (int sc = 0; sc < nsc; ++sc) {
            PH(sc, comp, i, 0, k) = T::cs(ang(sc, TH(1, i, 0, k))) * PH(sc, comp, i, 1, k) +
                                    (R)c[comp] * T::sn(ang(sc, TH(1, i, 0, k))) * PH(sc, comp + c[comp], i, 1, k);
            PH(sc, comp, i, L.ny - 1, k) =
                T::cs(ang(sc, TH(1, i, L.ny - 2, k))) * PH(sc, comp, i, L.ny - 2, k) -
                (R)c[comp] * T::sn(ang(sc, TH(1, i, L.ny - 2, k))) * PH(sc, comp + c[comp], i, L.ny - 2, k);
          }
      }
    for (int j = 0; j < L.ny; ++j)
      for (int k = zEdge1; k < zEdge2; ++k) {
        for (int comp = 0; comp < 3; ++comp) {
          TH(comp, 0, j, k) = TH(comp, 1, j, k);
          TH(comp, L.nx - 1, j, k) = TH(comp, L.nx - 2, j, k);
        }
        for (int comp = 0; comp < 2; ++comp)
          for (int sc = 0; sc < nsc; ++sc) {
            PH(sc, comp, 0, j, k) = T::cs(ang(sc, TH(0, 0, j, k))) * PH(sc, comp, 1, j, k) +
                                    (R)c[comp] * T::sn(ang(sc, TH(0, 0, j, k))) * PH(sc, comp + c[comp], 1, j, k);
            PH(sc, comp, L.nx - 1, j, k) =
                T::cs(ang(sc, TH(0, L.nx - 2, j, k))) * PH(sc, comp, L.nx - 2, j, k) -
                (R)c[comp] * T::sn(ang(sc, TH(0, L.nx - 2, j, k))) * PH(sc, comp + c[comp], L.nx - 2, j, k);
          }
      }
  }
  if (!moore && p->boundary_z == AX_BC_NEUMANN) {
    for (int i = 0; i < L.nz; ++i)          /* sic: x index bounded by nz (evolution.cpp:237) */
      for (int j = 0; j < L.ny; ++j) {
        for (int comp = 0; comp < 3; ++comp) {
          TH(comp, i, j, 0) = TH(comp, i, j, 1);
          TH(comp, i, j, L.nz - 1) = TH(comp, i, j, L.nz - 2);
        }
        for (int comp = 0; comp < 2; ++comp) {
          PH(0, comp, i, j, 0) = T::cs(TH(2, i, j, 0)) * PH(0, comp, i, j, 1) +
                                 (R)c[comp] * T::sn(TH(2, i, j, 0)) * PH(0, comp + c[comp], i, j, 1);
          PH(0, comp, i, j, L.nz - 1) = T::cs(TH(2, i, j, L.nz - 2)) * PH(0, comp, i, j, L.nz - 2) -
                                        (R)c[comp] * T::sn(TH(2, i, j, L.nz - 2)) * PH(0, comp + c[comp], i, j, L.nz - 2);
        }
      }
  }
}

/* Absorbing x/y boundary (AH): second time derivatives of every face and
 * corner site from the pre-update state, then their update
 * prv = 2 cur - prv + dt^2 tt.  Must run before the bulk update. */
template <class R, class T>
void bc_absorbing(const ax_params *p, const Lat &L, R *st, int tN, int tP) {
  const int c[2] = {1, -1};
  const int nx = L.nx, ny = L.ny, nz = L.nz;
  auto PHI = [&](int comp, int lev, int i, int j, int k) -> R & {
    return st[((size_t)comp * 2 + lev) * L.V + L.id(i, j, k)];
  };
  auto TH = [&](int comp, int lev, int i, int j, int k) -> R & {
    return st[((size_t)(2 + comp) * 2 + lev) * L.V + L.id(i, j, k)];
  };
  auto cs = [&](int a, int lev, int i, int j, int k) { return T::cs(TH(a, lev, i, j, k)); };
  auto sn = [&](int a, int lev, int i, int j, int k) { return T::sn(TH(a, lev, i, j, k)); };
  const R dx2 = (R)(p->dx * p->dx), dy2 = (R)(p->dy * p->dy), dz2 = (R)(p->dz * p->dz);
  const R dtdx = (R)(p->dt * p->dx), dtdy = (R)(p->dt * p->dy), dt2 = (R)(p->dt * p->dt);
  const R half = (R)0.5;
  const bool zper = p->boundary_z == AX_BC_PERIODIC;
  const int zEdge1 = zper ? 0 : 1, zEdge2 = zper ? nz : nz - 1;
  const int tNow = tN, tPast = tP;
  std::vector<R> phitt(2 * L.V, 0), thetatt(3 * L.V, 0);
  auto PT = [&](int comp, int i, int j, int k) -> R & { return phitt[(size_t)comp * L.V + L.id(i, j, k)]; };
  auto TT = [&](int comp, int i, int j, int k) -> R & { return thetatt[(size_t)comp * L.V + L.id(i, j, k)]; };
  /* covariant second difference along axis a (0 x, 1 y, 2 z) of comp at (i,j,k) */
  auto lap = [&](int comp, int a, int i, int j, int k, int ip, int jp, int kp, int im, int jm, int km, R d2) {
    return (cs(a, tNow, i, j, k) * PHI(comp, tNow, ip, jp, kp) +
            (R)c[comp] * sn(a, tNow, i, j, k) * PHI(comp + c[comp], tNow, ip, jp, kp) - 2 * PHI(comp, tNow, i, j, k) +
            cs(a, tNow, im, jm, km) * PHI(comp, tNow, im, jm, km) -
            (R)c[comp] * sn(a, tNow, im, jm, km) * PHI(comp + c[comp], tNow, im, jm, km)) / d2;
  };
  for (int k = zEdge1; k < zEdge2; ++k) {
    const int kp = (k + 1) % nz, km = (k - 1 + nz) % nz;
    /* x faces, evolution.cpp:453-517 */
    for (int j = 1; j < ny - 1; ++j) {
      for (int comp = 0; comp < 2; ++comp) {
        const int o = comp + c[comp];
        R phiyy = lap(comp, 1, 0, j, k, 0, j + 1, k, 0, j - 1, k, dy2);
        R phizz = lap(comp, 2, 0, j, k, 0, j, kp, 0, j, km, dz2);
        R phitx = (cs(0, tNow, 0, j, k) * PHI(comp, tNow, 1, j, k) + (R)c[comp] * sn(0, tNow, 0, j, k) * PHI(o, tNow, 1, j, k) -
                   PHI(comp, tNow, 0, j, k) - cs(0, tPast, 0, j, k) * PHI(comp, tPast, 1, j, k) -
                   (R)c[comp] * sn(0, tPast, 0, j, k) * PHI(o, tPast, 1, j, k) + PHI(comp, tPast, 0, j, k)) / dtdx;
        PT(comp, 0, j, k) = phitx + half * (phiyy + phizz);
        phiyy = lap(comp, 1, nx - 1, j, k, nx - 1, j + 1, k, nx - 1, j - 1, k, dy2);
        phizz = lap(comp, 2, nx - 1, j, k, nx - 1, j, kp, nx - 1, j, km, dz2);
        phitx = (PHI(comp, tNow, nx - 1, j, k) - cs(0, tNow, nx - 2, j, k) * PHI(comp, tNow, nx - 2, j, k) +
                 (R)c[comp] * sn(0, tNow, nx - 2, j, k) * PHI(o, tNow, nx - 2, j, k) - PHI(comp, tPast, nx - 1, j, k) +
                 cs(0, tPast, nx - 2, j, k) * PHI(comp, tPast, nx - 2, j, k) -
                 (R)c[comp] * sn(0, tPast, nx - 2, j, k) * PHI(o, tPast, nx - 2, j, k)) / dtdx;
        PT(comp, nx - 1, j, k) = -phitx + half * (phiyy + phizz);
      }
      for (int comp = 0; comp < 3; ++comp) {
        R thetayy = (TH(comp, tNow, 0, j + 1, k) - 2 * TH(comp, tNow, 0, j, k) + TH(comp, tNow, 0, j - 1, k)) / dy2;
        R thetazz = (TH(comp, tNow, 0, j, kp) - 2 * TH(comp, tNow, 0, j, k) + TH(comp, tNow, 0, j, km)) / dz2;
        R thetatx = (TH(comp, tNow, 1, j, k) - TH(comp, tNow, 0, j, k) - TH(comp, tPast, 1, j, k) + TH(comp, tPast, 0, j, k)) / dtdx;
        TT(comp, 0, j, k) = thetatx + half * (thetayy + thetazz);
        thetayy = (TH(comp, tNow, nx - 1, j + 1, k) - 2 * TH(comp, tNow, nx - 1, j, k) + TH(comp, tNow, nx - 1, j - 1, k)) / dy2;
        thetazz = (TH(comp, tNow, nx - 1, j, kp) - 2 * TH(comp, tNow, nx - 1, j, k) + TH(comp, tNow, nx - 1, j, km)) / dz2;
        thetatx = (TH(comp, tNow, nx - 1, j, k) - TH(comp, tNow, nx - 2, j, k) - TH(comp, tPast, nx - 1, j, k) +
                   TH(comp, tPast, nx - 2, j, k)) / dtdx;
        TT(comp, nx - 1, j, k) = -thetatx + half * (thetayy + thetazz);
      }
    }
    /* y faces, evolution.cpp:521-604 */
    for (int i = 1; i < nx - 1; ++i) {
      for (int comp = 0; comp < 2; ++comp) {
        const int o = comp + c[comp];
        R phixx = lap(comp, 0, i, 0, k, i + 1, 0, k, i - 1, 0, k, dx2);
        R phizz = lap(comp, 2, i, 0, k, i, 0, kp, i, 0, km, dz2);
        R phity = (cs(1, tNow, i, 0, k) * PHI(comp, tNow, i, 1, k) + (R)c[comp] * sn(1, tNow, i, 0, k) * PHI(o, tNow, i, 1, k) -
                   PHI(comp, tNow, i, 0, k) - cs(1, tPast, i, 0, k) * PHI(comp, tPast, i, 1, k) -
                   (R)c[comp] * sn(1, tPast, i, 0, k) * PHI(o, tPast, i, 1, k) + PHI(comp, tPast, i, 0, k)) / dtdy;
        PT(comp, i, 0, k) = phity + half * (phixx + phizz);
        phixx = lap(comp, 0, i, ny - 1, k, i + 1, ny - 1, k, i - 1, ny - 1, k, dx2);
        phizz = lap(comp, 2, i, ny - 1, k, i, ny - 1, kp, i, ny - 1, km, dz2);
        phity = (PHI(comp, tNow, i, ny - 1, k) - cs(1, tNow, i, ny - 2, k) * PHI(comp, tNow, i, ny - 2, k) +
                 (R)c[comp] * sn(1, tNow, i, ny - 2, k) * PHI(o, tNow, i, ny - 2, k) - PHI(comp, tPast, i, ny - 1, k) +
                 cs(1, tPast, i, ny - 2, k) * PHI(comp, tPast, i, ny - 2, k) -
                 (R)c[comp] * sn(1, tPast, i, ny - 2, k) * PHI(o, tPast, i, ny - 2, k)) / dtdy;
        PT(comp, i, ny - 1, k) = -phity + half * (phixx + phizz);
      }
      for (int comp = 0; comp < 3; ++comp) {
        R thetaxx = (TH(comp, tNow, i + 1, 0, k) - 2 * TH(comp, tNow, i, 0, k) + TH(comp, tNow, i - 1, 0, k)) / dx2;
        R thetazz = (TH(comp, tNow, i, 0, kp) - 2 * TH(comp, tNow, i, 0, k) + TH(comp, tNow, i, 0, km)) / dz2;
        R thetaty = (TH(comp, tNow, i, 1, k) - TH(comp, tNow, i, 0, k) - TH(comp, tPast, i, 1, k) + TH(comp, tPast, i, 0, k)) / dtdy;
        TT(comp, i, 0, k) = thetaty + half * (thetaxx + thetazz);
        thetaxx = (TH(comp, tNow, i + 1, ny - 1, k) - 2 * TH(comp, tNow, i, ny - 1, k) + TH(comp, tNow, i - 1, ny - 1, k)) / dx2;
        thetazz = (TH(comp, tNow, i, ny - 1, kp) - 2 * TH(comp, tNow, i, ny - 1, k) + TH(comp, tNow, i, ny - 1, km)) / dz2;
        thetaty = (TH(comp, tNow, i, ny - 1, k) - TH(comp, tNow, i, ny - 2, k) - TH(comp, tPast, i, ny - 1, k) +
                   TH(comp, tPast, i, ny - 2, k)) / dtdy;
        TT(comp, i, ny - 1, k) = -thetaty + half * (thetaxx + thetazz);
      }
    }
    /* corners, evolution.cpp:608-683 */
    for (int comp = 0; comp < 2; ++comp) {
      const int o = comp + c[comp];
      /* x,y < 0 */
      R phizz = lap(comp, 2, 0, 0, k, 0, 0, kp, 0, 0, km, dz2);
      R phitx = (cs(0, tNow, 0, 0, k) * PHI(comp, tNow, 1, 0, k) + (R)c[comp] * sn(0, tNow, 0, 0, k) * PHI(o, tNow, 1, 0, k) -
                 PHI(comp, tNow, 0, 0, k) - cs(0, tPast, 0, 0, k) * PHI(comp, tPast, 1, 0, k) -
                 (R)c[comp] * sn(0, tPast, 0, 0, k) * PHI(o, tPast, 1, 0, k) + PHI(comp, tPast, 0, 0, k)) / dtdx;
      R phity = (cs(1, tNow, 0, 0, k) * PHI(comp, tNow, 0, 1, k) + (R)c[comp] * sn(1, tNow, 0, 0, k) * PHI(o, tNow, 0, 1, k) -
                 PHI(comp, tNow, 0, 0, k) - cs(1, tPast, 0, 0, k) * PHI(comp, tPast, 0, 1, k) -
                 (R)c[comp] * sn(1, tPast, 0, 0, k) * PHI(o, tPast, 0, 1, k) + PHI(comp, tPast, 0, 0, k)) / dtdy;
      PT(comp, 0, 0, k) = (phizz + 2 * (phitx + phity)) / 3;
      /* x < 0, y > 0 */
      phizz = lap(comp, 2, 0, ny - 1, k, 0, ny - 1, kp, 0, ny - 1, km, dz2);
      phitx = (cs(0, tNow, 0, ny - 1, k) * PHI(comp, tNow, 1, ny - 1, k) +
               (R)c[comp] * sn(0, tNow, 0, ny - 1, k) * PHI(o, tNow, 1, ny - 1, k) - PHI(comp, tNow, 0, ny - 1, k) -
               cs(0, tPast, 0, ny - 1, k) * PHI(comp, tPast, 1, ny - 1, k) -
               (R)c[comp] * sn(0, tPast, 0, ny - 1, k) * PHI(o, tPast, 1, ny - 1, k) + PHI(comp, tPast, 0, ny - 1, k)) / dtdx;
      phity = (PHI(comp, tNow, 0, ny - 1, k) - cs(1, tNow, 0, ny - 2, k) * PHI(comp, tNow, 0, ny - 2, k) +
               (R)c[comp] * sn(1, tNow, 0, ny - 2, k) * PHI(o, tNow, 0, ny - 2, k) - PHI(comp, tPast, 0, ny - 1, k) +
               cs(1, tPast, 0, ny - 2, k) * PHI(comp, tPast, 0, ny - 2, k) -
               (R)c[comp] * sn(1, tPast, 0, ny - 2, k) * PHI(o, tPast, 0, ny - 2, k)) / dtdy;
      PT(comp, 0, ny - 1, k) = (phizz + 2 * (phitx - phity)) / 3;
      /* x > 0, y < 0 */
      phizz = lap(comp, 2, nx - 1, 0, k, nx - 1, 0, kp, nx - 1, 0, km, dz2);
      phitx = (PHI(comp, tNow, nx - 1, 0, k) - cs(0, tNow, nx - 2, 0, k) * PHI(comp, tNow, nx - 2, 0, k) +
               (R)c[comp] * sn(0, tNow, nx - 2, 0, k) * PHI(o, tNow, nx - 2, 0, k) - PHI(comp, tPast, nx - 1, 0, k) +
               cs(0, tPast, nx - 2, 0, k) * PHI(comp, tPast, nx - 2, 0, k) -
               (R)c[comp] * sn(0, tPast, nx - 2, 0, k) * PHI(o, tPast, nx - 2, 0, k)) / dtdx;
      phity = (cs(1, tNow, nx - 1, 0, k) * PHI(comp, tNow, nx - 1, 1, k) +
               (R)c[comp] * sn(1, tNow, nx - 1, 0, k) * PHI(o, tNow, nx - 1, 1, k) - PHI(comp, tNow, nx - 1, 0, k) -
               cs(1, tPast, nx - 1, 0, k) * PHI(comp, tPast, nx - 1, 1, k) -
               (R)c[comp] * sn(1, tPast, nx - 1, 0, k) * PHI(o, tPast, nx - 1, 1, k) + PHI(comp, tPast, nx - 1, 0, k)) / dtdy;
      PT(comp, nx - 1, 0, k) = (phizz - 2 * (phitx - phity)) / 3;
      /* x, y > 0; the level-tPast cosine is taken at the corner (evolution.cpp:642) */
      phizz = lap(comp, 2, nx - 1, ny - 1, k, nx - 1, ny - 1, kp, nx - 1, ny - 1, km, dz2);
      phitx = (PHI(comp, tNow, nx - 1, ny - 1, k) - cs(0, tNow, nx - 2, ny - 1, k) * PHI(comp, tNow, nx - 2, ny - 1, k) +
               (R)c[comp] * sn(0, tNow, nx - 2, ny - 1, k) * PHI(o, tNow, nx - 2, ny - 1, k) -
               PHI(comp, tPast, nx - 1, ny - 1, k) + cs(0, tPast, nx - 1, ny - 1, k) * PHI(comp, tPast, nx - 2, ny - 1, k) -
               (R)c[comp] * sn(0, tPast, nx - 2, ny - 1, k) * PHI(o, tPast, nx - 2, ny - 1, k)) / dtdx;
      phity = (PHI(comp, tNow, nx - 1, ny - 1, k) - cs(1, tNow, nx - 1, ny - 2, k) * PHI(comp, tNow, nx - 1, ny - 2, k) +
               (R)c[comp] * sn(1, tNow, nx - 1, ny - 2, k) * PHI(o, tNow, nx - 1, ny - 2, k) -
               PHI(comp, tPast, nx - 1, ny - 1, k) + cs(1, tPast, nx - 1, ny - 2, k) * PHI(comp, tPast, nx - 1, ny - 2, k) -
               (R)c[comp] * sn(1, tPast, nx - 1, ny - 2, k) * PHI(o, tPast, nx - 1, ny - 2, k)) / dtdy;
      PT(comp, nx - 1, ny - 1, k) = (phizz - 2 * (phitx + phity)) / 3;
    }
    for (int comp = 0; comp < 3; ++comp) {
      R thetazz = (TH(comp, tNow, 0, 0, kp) - 2 * TH(comp, tNow, 0, 0, k) + TH(comp, tNow, 0, 0, km)) / dz2;
      R thetatx = (TH(comp, tNow, 1, 0, k) - TH(comp, tNow, 0, 0, k) - TH(comp, tPast, 1, 0, k) + TH(comp, tPast, 0, 0, k)) / dtdx;
      R thetaty = (TH(comp, tNow, 0, 1, k) - TH(comp, tNow, 0, 0, k) - TH(comp, tPast, 0, 1, k) + TH(comp, tPast, 0, 0, k)) / dtdy;
      TT(comp, 0, 0, k) = (thetazz + 2 * (thetatx + thetaty)) / 3;
      thetazz = (TH(comp, tNow, 0, ny - 1, kp) - 2 * TH(comp, tNow, 0, ny - 1, k) + TH(comp, tNow, 0, ny - 1, km)) / dz2;
      thetatx = (TH(comp, tNow, 1, ny - 1, k) - TH(comp, tNow, 0, ny - 1, k) - TH(comp, tPast, 1, ny - 1, k) +
                 TH(comp, tPast, 0, ny - 1, k)) / dtdx;
      thetaty = (TH(comp, tNow, 0, ny - 1, k) - TH(comp, tNow, 0, ny - 2, k) - TH(comp, tPast, 0, ny - 1, k) +
                 TH(comp, tPast, 0, ny - 2, k)) / dtdy;
      TT(comp, 0, ny - 1, k) = (thetazz + 2 * (thetatx - thetaty)) / 3;
      thetazz = (TH(comp, tNow, nx - 1, 0, kp) - 2 * TH(comp, tNow, nx - 1, 0, k) + TH(comp, tNow, nx - 1, 0, km)) / dz2;
      thetatx = (TH(comp, tNow, nx - 1, 0, k) - TH(comp, tNow, nx - 2, 0, k) - TH(comp, tPast, nx - 1, 0, k) +
                 TH(comp, tPast, nx - 2, 0, k)) / dtdx;
      thetaty = (TH(comp, tNow, nx - 1, 1, k) - TH(comp, tNow, nx - 1, 0, k) - TH(comp, tPast, nx - 1, 1, k) +
                 TH(comp, tPast, nx - 1, 0, k)) / dtdy;
      TT(comp, nx - 1, 0, k) = (thetazz - 2 * (thetatx - thetaty)) / 3;
      thetazz = (TH(comp, tNow, nx - 1, ny - 1, kp) - 2 * TH(comp, tNow, nx - 1, ny - 1, k) +
                 TH(comp, tNow, nx - 1, ny - 1, km)) / dz2;
      thetatx = (TH(comp, tNow, nx - 1, ny - 1, k) - TH(comp, tNow, nx - 2, ny - 1, k) - TH(comp, tPast, nx - 1, ny - 1, k) +
                 TH(comp, tPast, nx - 2, ny - 1, k)) / dtdx;
      thetaty = (TH(comp, tNow, nx - 1, ny - 1, k) - TH(comp, tNow, nx - 1, ny - 2, k) - TH(comp, tPast, nx - 1, ny - 1, k) +
                 TH(comp, tPast, nx - 1, ny - 2, k)) / dtdy;
      TT(comp, nx - 1, ny - 1, k) = (thetazz - 2 * (thetatx + thetaty)) / 3;
    }
  }
  /* update of the boundary sites (the bulk update covers the interior) */
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j) {
      if (i > 0 && i < nx - 1 && j > 0 && j < ny - 1) continue;
      for (int k = zEdge1; k < zEdge2; ++k) {
        for (int comp = 0; comp < 2; ++comp)
          PHI(comp, tPast, i, j, k) = 2 * PHI(comp, tNow, i, j, k) - PHI(comp, tPast, i, j, k) + dt2 * PT(comp, i, j, k);
        for (int comp = 0; comp < 3; ++comp)
          TH(comp, tPast, i, j, k) = 2 * TH(comp, tNow, i, j, k) - TH(comp, tPast, i, j, k) + dt2 * TT(comp, i, j, k);
      }
    }
}

template <class T>
void run_steps(const ax_params *p, double *st, int ts0, int nsteps, int nthreads) {
  Lat L(p);
  const int e = (p->boundary != AX_BC_PERIODIC) ? 1 : 0, ez = (p->boundary_z != AX_BC_PERIODIC) ? 1 : 0;
  for (int n = ts0; n < ts0 + nsteps; ++n) {
    const int tN = (n + 1) % 2, tP = n % 2;
    const double time = 1 + n * p->dt;
    const double damp = (n < p->damped_nt) ? p->dampFac : 0;
    const double fric = damp + p->alpha * p->scaling / time;
    if (p->boundary == AX_BC_NEUMANN || p->boundary_z == AX_BC_NEUMANN) bc_neumann<double, T>(p, L, st, tN);
    if (p->boundary == AX_BC_ABSORBING) bc_absorbing<double, T>(p, L, st, tN, tP);
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
#endif
    for (int i = e; i < L.nx - e; ++i)
      for (int j = e; j < L.ny - e; ++j)
        for (int k = ez; k < L.nz - ez; ++k) {
          Nb nb(L, i, j, k);
          if (p->model == AX_MODEL_AH)
            ah_site<T>(p, L, st, tN, tP, nb, damp, fric);
          else
            moore_site<T>(p, L, st, tN, tP, nb);
        }
  }
}

/* ---------------------------- diagnostics -------------------------------- */
/* out[0] energy, out[1] mean |G| (periodic: sum/(nx*ny)/nz; fixed edges:
 * mooreEvol.cpp's running sum of |G|/((nx-2)(ny-2)(nz-2)) before its second
 * division), out[2] 1.0 if every stored value is finite. */
template <class T>
void diag_ah(const ax_params *p, const double *st, int n, double *out) {
  Lat L(p);
  const int tN = (n + 1) % 2, tP = n % 2;
  const double *f[2] = {fld(st, L, 0, tN), fld(st, L, 1, tN)};
  const double *fP[2] = {fld(st, L, 0, tP), fld(st, L, 1, tP)};
  const double *th[3] = {fld(st, L, 2, tN), fld(st, L, 3, tN), fld(st, L, 4, tN)};
  const double *thP[3] = {fld(st, L, 2, tP), fld(st, L, 3, tP), fld(st, L, 4, tP)};
  const double dx = p->dx, dy = p->dy, dz = p->dz, dt = p->dt, g = p->g, lam = p->lambda, eta = p->eta;
  const double sg[2] = {1.0, -1.0};
  double energy = 0, dev = 0;
  /* the bulk-update region (evolution.cpp:161-186) */
  const int e = p->boundary != AX_BC_PERIODIC, ez = p->boundary_z != AX_BC_PERIODIC;
  for (int i = e; i < L.nx - e; ++i)
    for (int j = e; j < L.ny - e; ++j)
      for (int k = ez; k < L.nz - ez; ++k) {
        Nb n_(L, i, j, k);
        const size_t s = L.id(i, j, k);
        const size_t sxp = L.id(n_.ip, j, k), sxm = L.id(n_.im, j, k);
        const size_t syp = L.id(i, n_.jp, k), sym = L.id(i, n_.jm, k);
        const size_t szp = L.id(i, j, n_.kp), szm = L.id(i, j, n_.km);
        double mag2 = f[0][s] * f[0][s] + f[1][s] * f[1][s];
        double vel[2], gx[2], gy[2], gz[2];
        for (int c = 0; c < 2; ++c) {
          const double *a = f[c], *b = f[1 - c];
          vel[c] = (a[s] - fP[c][s]) / dt;
          gx[c] = (a[s] - T::cs(th[0][sxm]) * a[sxm] + sg[c] * T::sn(th[0][sxm]) * b[sxm]) / dx;
          gy[c] = (a[s] - T::cs(th[1][sym]) * a[sym] + sg[c] * T::sn(th[1][sym]) * b[sym]) / dy;
          gz[c] = (a[s] - T::cs(th[2][szm]) * a[szm] + sg[c] * T::sn(th[2][szm]) * b[szm]) / dz;
        }
        double pxy = th[0][s] + th[1][sxp] - th[0][syp] - th[1][s];
        double pxz = th[0][s] + th[2][sxp] - th[0][szp] - th[2][s];
        double pyz = th[1][s] + th[2][syp] - th[1][szp] - th[2][s];
        double tdc, fcont, G;
        if (p->energy_form == 0) {
          if (g == 0) {
            G = 0; tdc = 0; fcont = 0;
          } else {
            double divN = (th[0][s] - th[0][sxm]) / (dx * dx) + (th[1][s] - th[1][sym]) / (dy * dy) +
                          (th[2][s] - th[2][szm]) / (dz * dz);
            double divP = (thP[0][s] - thP[0][sxm]) / (dx * dx) + (thP[1][s] - thP[1][sym]) / (dy * dy) +
                          (thP[2][s] - thP[2][szm]) / (dz * dz);
            double divT = (divN - divP) / dt;
            G = divT / g - 2 * g * (f[1][s] * vel[0] - f[0][s] * vel[1]);
            double ex = (th[0][s] - thP[0][s]) / (g * dt * dx);
            double ey = (th[1][s] - thP[1][s]) / (g * dt * dy);
            double ez = (th[2][s] - thP[2][s]) / (g * dt * dz);
            tdc = 0.5 * (ex * ex + ey * ey + ez * ez);
            double bxy = (1 - T::cs(pxy)) / ((g * dx * dy) * (g * dx * dy));
            double bxz = (1 - T::cs(pxz)) / ((g * dx * dz) * (g * dx * dz));
            double byz = (1 - T::cs(pyz)) / ((g * dy * dz) * (g * dy * dz));
            fcont = bxy + bxz + byz;
          }
          dev += fabs(G);
          double vv = mag2 - eta * eta;
          double rho = vel[0] * vel[0] + vel[1] * vel[1] + gx[0] * gx[0] + gx[1] * gx[1] +
                       gy[0] * gy[0] + gy[1] * gy[1] + gz[0] * gz[0] + gz[1] * gz[1] + tdc + fcont +
                       0.25 * lam * (vv * vv);
          energy += dx * dy * dz * rho;
        } else {
          double ex = (th[0][s] - thP[0][s]) / dt;
          double ey = (th[1][s] - thP[1][s]) / dt;
          double ez = (th[2][s] - thP[2][s]) / dt;
          tdc = 0.5 * (ex * ex + ey * ey + ez * ez);
          double bxy = (1 - T::cs(pxy)) / ((g * dx * dy) * (g * dx * dy));
          double bxz = (1 - T::cs(pxz)) / ((g * dx * dz) * (g * dx * dz));
          double byz = (1 - T::cs(pyz)) / ((g * dy * dz) * (g * dy * dz));
          fcont = bxy + bxz + byz;
          double vv = mag2 - eta * eta;
          energy += (vel[0] * vel[0] + vel[1] * vel[1] + gx[0] * gx[0] + gx[1] * gx[1] +
                     gy[0] * gy[0] + gy[1] * gy[1] + gz[0] * gz[0] + gz[1] * gz[1] + tdc + fcont +
                     0.25 * lam * (vv * vv)) * dx * dy * dz;
        }
      }
  /* evolution.cpp:1400-1408 */
  if (p->boundary == AX_BC_PERIODIC) dev = dev / (L.nx * L.ny);
  else dev = dev / ((L.nx - 2) * (L.ny - 2));
  if (p->boundary_z == AX_BC_PERIODIC) dev = dev / L.nz;
  else dev = dev / (L.nz - 2);
  out[0] = energy;
  out[1] = (p->energy_form == 0) ? dev : 0.0;
}

template <class T>
void diag_moore(const ax_params *p, const double *st, int n, double *out) {
  Lat L(p);
  const int tN = (n + 1) % 2, tP = n % 2;
  const double *f1[2] = {fld(st, L, 0, tN), fld(st, L, 1, tN)};
  const double *f2[2] = {fld(st, L, 2, tN), fld(st, L, 3, tN)};
  const double *f1P[2] = {fld(st, L, 0, tP), fld(st, L, 1, tP)};
  const double *f2P[2] = {fld(st, L, 2, tP), fld(st, L, 3, tP)};
  const double *th[3] = {fld(st, L, 4, tN), fld(st, L, 5, tN), fld(st, L, 6, tN)};
  const double *thP[3] = {fld(st, L, 4, tP), fld(st, L, 5, tP), fld(st, L, 6, tP)};
  const double dx = p->dx, dy = p->dy, dz = p->dz, dt = p->dt, g = p->g;
  const int q1 = p->q1, q2 = p->q2;
  const double sg[2] = {1.0, -1.0};
  const bool fixed = p->boundary != AX_BC_PERIODIC;
  const int e = fixed ? 1 : 0;
  const int inner = (L.nx - 2) * (L.ny - 2) * (L.nz - 2);
  double energy = 0, dev = 0;
  for (int i = e; i < L.nx - e; ++i)
    for (int j = e; j < L.ny - e; ++j)
      for (int k = e; k < L.nz - e; ++k) {
        Nb n_(L, i, j, k);
        const size_t s = L.id(i, j, k);
        const size_t sxp = L.id(n_.ip, j, k), sxm = L.id(n_.im, j, k);
        const size_t syp = L.id(i, n_.jp, k), sym = L.id(i, n_.jm, k);
        const size_t szp = L.id(i, j, n_.kp), szm = L.id(i, j, n_.km);
        double m1 = f1[0][s] * f1[0][s] + f1[1][s] * f1[1][s];
        double m2 = f2[0][s] * f2[0][s] + f2[1][s] * f2[1][s];
        double v1[2], v2[2], g1x[2], g1y[2], g1z[2], g2x[2], g2y[2], g2z[2];
        for (int c = 0; c < 2; ++c) {
          const double *a = f1[c], *b = f1[1 - c], *u = f2[c], *w = f2[1 - c];
          g1x[c] = (a[s] - T::cs(q1 * th[0][sxm]) * a[sxm] + sg[c] * T::sn(q1 * th[0][sxm]) * b[sxm]) / dx;
          g1y[c] = (a[s] - T::cs(q1 * th[1][sym]) * a[sym] + sg[c] * T::sn(q1 * th[1][sym]) * b[sym]) / dy;
          g2x[c] = (u[s] - T::cs(q2 * th[0][sxm]) * u[sxm] + sg[c] * T::sn(q2 * th[0][sxm]) * w[sxm]) / dx;
          g2y[c] = (u[s] - T::cs(q2 * th[1][sym]) * u[sym] + sg[c] * T::sn(q2 * th[1][sym]) * w[sym]) / dy;
          g1z[c] = (a[s] - T::cs(q1 * th[2][szm]) * a[szm] + sg[c] * T::sn(q1 * th[2][szm]) * b[szm]) / dz;
          g2z[c] = (u[s] - T::cs(q2 * th[2][szm]) * u[szm] + sg[c] * T::sn(q2 * th[2][szm]) * w[szm]) / dz;
          v1[c] = (a[s] - f1P[c][s]) / dt;
          v2[c] = (u[s] - f2P[c][s]) / dt;
        }
        double tdc, fcont, G;
        if (g == 0) {
          G = 0; tdc = 0; fcont = 0;
        } else {
          double divN = (th[0][s] - th[0][sxm]) / (dx * dx) + (th[1][s] - th[1][sym]) / (dy * dy) +
                        (th[2][s] - th[2][szm]) / (dz * dz);
          double divP = (thP[0][s] - thP[0][sxm]) / (dx * dx) + (thP[1][s] - thP[1][sym]) / (dy * dy) +
                        (thP[2][s] - thP[2][szm]) / (dz * dz);
          double divT = (divN - divP) / dt;
          G = divT / g - 2 * g * (q1 * (f1[1][s] * v1[0] - f1[0][s] * v1[1]) +
                                  q2 * (f2[1][s] * v2[0] - f2[0][s] * v2[1]));
          double ex = (th[0][s] - thP[0][s]) / (g * dt * dx);
          double ey = (th[1][s] - thP[1][s]) / (g * dt * dy);
          double ez = (th[2][s] - thP[2][s]) / (g * dt * dz);
          tdc = 0.5 * (ex * ex + ey * ey + ez * ez);
          double bxy = (1 - T::cs(th[0][s] + th[1][sxp] - th[0][syp] - th[1][s])) / ((g * dx * dy) * (g * dx * dy));
          double bxz = (1 - T::cs(th[0][s] + th[2][sxp] - th[0][szp] - th[2][s])) / ((g * dx * dz) * (g * dx * dz));
          double byz = (1 - T::cs(th[1][s] + th[2][syp] - th[1][szp] - th[2][s])) / ((g * dy * dz) * (g * dy * dz));
          fcont = bxy + bxz + byz;
        }
        if (fixed)
          dev += fabs(G) / inner;
        else
          dev += fabs(G);
        double w1 = m1 - p->eta1 * p->eta1, w2 = m2 - p->eta2 * p->eta2;
        double rho = v1[0] * v1[0] + v1[1] * v1[1] + g1x[0] * g1x[0] + g1x[1] * g1x[1] + g1y[0] * g1y[0] +
                     g1y[1] * g1y[1] + g1z[0] * g1z[0] + g1z[1] * g1z[1] + v2[0] * v2[0] + v2[1] * v2[1] +
                     g2x[0] * g2x[0] + g2x[1] * g2x[1] + g2y[0] * g2y[0] + g2y[1] * g2y[1] + g2z[0] * g2z[0] +
                     g2z[1] * g2z[1] + tdc + fcont + 0.25 * p->lambda1 * (w1 * w1) +
                     0.25 * p->lambda2 * (w2 * w2) + 0.5 * p->lambda12 * w1 * w2;
        energy += dx * dy * dz * rho;
      }
  if (!fixed) {
    dev = dev / (L.nx * L.ny);
    dev = dev / L.nz;
  }
  out[0] = energy;
  out[1] = dev;
}

/* ------------------------------ winding ---------------------------------- */
/* Gauge-invariant phase across the link s -> s+a:
 * delta = arg( conj(phi(s)) * exp(-i theta_a(s)) * phi(s+a) ), the transport
 * used in the covariant Laplacian (evolution.cpp:304-311). */
inline double link_phase(double cr, double ci, double nr, double ni, double th) {
  double c = ax_cos(th), sn = ax_sin(th);
  double wr = c * nr + sn * ni;
  double wi = c * ni - sn * nr;
  double re = cr * wr + ci * wi;
  double im = cr * wi - ci * wr;
  return ax_atan2(im, re);
}

/* winding of plaquette (s; a, b): round((d_a(s) + d_b(s+a) - d_a(s+b) - d_b(s)
 *   + P_ab(s)) / 2pi), P_ab(s) = theta_a(s) + theta_b(s+a) - theta_a(s+b)
 *   - theta_b(s) in the reference's operand order (mpi_evolution.cpp:471-477). */
inline int plaq_wind(double da0, double db1, double da2, double db3, double P) {
  double sum = da0 + db1 - da2 - db3 + P;
  return (int)rint(sum * 0.15915494309189535);
}

void winding_impl(const ax_params *p, const double *st, int n, int8_t *out, int field) {
  Lat L(p);
  const int tN = (n + 1) % 2;
  const int c0 = (p->model == AX_MODEL_AH) ? 0 : 2 * field;
  const int t0 = (p->model == AX_MODEL_AH) ? 2 : 4;
  const double q = (p->model == AX_MODEL_AH) ? 1.0 : (double)(field == 0 ? p->q1 : p->q2);
  const double *fr = fld(st, L, c0, tN), *fi = fld(st, L, c0 + 1, tN);
  const double *th[3] = {fld(st, L, t0, tN), fld(st, L, t0 + 1, tN), fld(st, L, t0 + 2, tN)};
  const bool ex = p->boundary != AX_BC_PERIODIC, ez = p->boundary_z != AX_BC_PERIODIC;
  for (int i = 0; i < L.nx; ++i)
    for (int j = 0; j < L.ny; ++j)
      for (int k = 0; k < L.nz; ++k) {
        Nb n_(L, i, j, k);
        size_t s = L.id(i, j, k);
        size_t sa[3] = {L.id(n_.ip, j, k), L.id(i, n_.jp, k), L.id(i, j, n_.kp)};
        const int pa[3] = {0, 0, 1}, pb[3] = {1, 2, 2};
        for (int pl = 0; pl < 3; ++pl) {
          int a = pa[pl], b = pb[pl];
          size_t s_a = sa[a], s_b = sa[b];
          /* s + a + b */
          int ii = i, jj = j, kk = k;
          if (a == 0 || b == 0) ii = n_.ip;
          if (a == 1 || b == 1) jj = n_.jp;
          if (a == 2 || b == 2) kk = n_.kp;
          size_t s_ab = L.id(ii, jj, kk);
          double da0 = link_phase(fr[s], fi[s], fr[s_a], fi[s_a], q * th[a][s]);
          double db1 = link_phase(fr[s_a], fi[s_a], fr[s_ab], fi[s_ab], q * th[b][s_a]);
          double da2 = link_phase(fr[s_b], fi[s_b], fr[s_ab], fi[s_ab], q * th[a][s_b]);
          double db3 = link_phase(fr[s], fi[s], fr[s_b], fi[s_b], q * th[b][s]);
          double P = q * th[a][s] + q * th[b][s_a] - q * th[a][s_b] - q * th[b][s];
          /* a plaquette wrapping across a non-periodic axis does not exist */
          const bool wx = (a == 0 || b == 0) && ex && i == L.nx - 1, wy = (a == 1 || b == 1) && ex && j == L.ny - 1,
                     wz = (a == 2 || b == 2) && ez && k == L.nz - 1;
          out[(size_t)pl * L.V + s] = (wx || wy || wz) ? 0 : (int8_t)plaq_wind(da0, db1, da2, db3, P);
        }
      }
}

/* ------------------------------ crossings -------------------------------- */
struct Pts {
  double *xyz;
  size_t cap, n;
  void push(double a, double b, double c) {
    if (n < cap) {
      xyz[3 * n] = a; xyz[3 * n + 1] = b; xyz[3 * n + 2] = c;
    }
    ++n;
  }
};

/* One face: bilinear zero of (Re, Im) phi on corners v00=(u,w), v10=(u+du,w),
 * v01=(u,w+dw), v11; returns roots as (sol1 along u, sol2 along w)
 * (mpi_evolution.cpp:552-623). kind selects how the point is stored. */
inline void face(const double v00[2], const double v10[2], const double v01[2], const double v11[2],
                 double u, double du, double w, double dw, int kind, double fixedc, Pts &P) {
  double c1[2], c2[2], c3[2], c4[2];
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
  auto emit = [&](double s1, double s2) {
    if (kind == 0) P.push(s1, s2, fixedc);        /* xy face: (x, y, z) */
    else if (kind == 1) P.push(s1, fixedc, s2);   /* xz face: (x, y, z) */
    else P.push(fixedc, s1, s2);                  /* yz face */
  };
  if (a == 0) {
    if (b != 0) {
      double s2 = -c / b;
      if (s2 >= w && s2 <= w + dw) {
        double s1 = -(c3[0] * s2 + c4[0]) / (c1[0] * s2 + c2[0]);
        if (s1 >= u && s1 <= u + du) emit(s1, s2);
      }
    }
  } else if (disc >= 0) {
    double s2 = (-b + sqrt(disc)) / (2 * a);
    if (s2 >= w && s2 <= w + dw) {
      double s1 = -(c3[0] * s2 + c4[0]) / (c1[0] * s2 + c2[0]);
      if (s1 >= u && s1 <= u + du) emit(s1, s2);
    }
    s2 = (-b - sqrt(disc)) / (2 * a);
    if (s2 >= w && s2 <= w + dw) {
      double s1 = -(c3[0] * s2 + c4[0]) / (c1[0] * s2 + c2[0]);
      if (s1 >= u && s1 <= u + du) emit(s1, s2);
    }
  }
}

void crossings_impl(const ax_params *p, const double *st, int n, Pts &P) {
  Lat L(p);
  const int tN = (n + 1) % 2;
  const double *f[2] = {fld(st, L, 0, tN), fld(st, L, 1, tN)};
  const int x0 = int(0.5 * (L.nx - 1)), y0 = int(0.5 * (L.ny - 1)), z0 = int(0.5 * (L.nz - 1));
  const double dx = p->dx, dy = p->dy, dz = p->dz;
  /* the bulk-update region (evolution.cpp:731-743) */
  const int e = p->boundary != AX_BC_PERIODIC, ez = p->boundary_z != AX_BC_PERIODIC;
  for (int i = e; i < L.nx - e; ++i)
    for (int j = e; j < L.ny - e; ++j)
      for (int k = ez; k < L.nz - ez; ++k) {
        Nb n_(L, i, j, k);
        size_t s = L.id(i, j, k), sx = L.id(n_.ip, j, k), sy = L.id(i, n_.jp, k), sz = L.id(i, j, n_.kp);
        size_t sxy = L.id(n_.ip, n_.jp, k), sxz = L.id(n_.ip, j, n_.kp), syz = L.id(i, n_.jp, n_.kp);
        double x = (i - x0) * dx, y = (j - y0) * dy, z = (k - z0) * dz;
        double v[2], vx[2], vy[2], vz[2], vxy[2], vxz[2], vyz[2];
        for (int c = 0; c < 2; ++c) {
          v[c] = f[c][s]; vx[c] = f[c][sx]; vy[c] = f[c][sy]; vz[c] = f[c][sz];
          vxy[c] = f[c][sxy]; vxz[c] = f[c][sxz]; vyz[c] = f[c][syz];
        }
        face(v, vx, vy, vxy, x, dx, y, dy, 0, z, P);
        face(v, vx, vz, vxz, x, dx, z, dz, 1, y, P);
        face(v, vy, vz, vyz, y, dy, z, dz, 2, x, P);
      }
}

/* ------------------- storage-type-R step (FP32 variant) -------------------
 * The FP32 variant has no reference counterpart; it is defined as the FP64
 * AH step evaluated in the storage type R: the per-step scalars are derived
 * in double exactly as above (dx*dx, dt*dt, 0.5*lambda, eta*eta, (-2*g)*g,
 * damp + alpha*scaling/time), rounded once to R, and every site expression
 * of ah_site is evaluated in R with the same operand order, using the
 * binary32 ax_sincosf for R = float.  ah_site_r<double> is the FP64 step
 * itself (tests check it bit for bit against run_steps<TrigAx>), which pins
 * the operand order of the float instantiation. */
template <class R>
struct RConsts {
  R dx2, dy2, dz2, dt, dt2, hl, eta2, m2gg, damp, fric;
  int form;
};
struct TrigF {
  static float sn(float x) { return ax_sinf(x); }
  static float cs(float x) {
    float s, c;
    ax_sincosf(x, &s, &c);
    return c;
  }
};
template <class R>
struct TrigOf;
template <>
struct TrigOf<double> : TrigAx {};
template <>
struct TrigOf<float> : TrigF {};

template <class R>
void ah_site_r(const Lat &L, R *st, int tN, int tP, const Nb &n, const RConsts<R> &k) {
  typedef TrigOf<R> T;
  auto F = [&](int c, int lev) { return st + ((size_t)c * 2 + lev) * L.V; };
  const R *pr = F(0, tN), *pi = F(1, tN), *tx = F(2, tN), *ty = F(3, tN), *tz = F(4, tN);
  R *P0 = F(0, tP), *P1 = F(1, tP), *PX = F(2, tP), *PY = F(3, tP), *PZ = F(4, tP);
  const size_t s = L.id(n.i, n.j, n.k);
  const size_t xp = L.id(n.ip, n.j, n.k), xm = L.id(n.im, n.j, n.k);
  const size_t yp = L.id(n.i, n.jp, n.k), ym = L.id(n.i, n.jm, n.k);
  const size_t zp = L.id(n.i, n.j, n.kp), zm = L.id(n.i, n.j, n.km);
  const size_t xp_ym = L.id(n.ip, n.jm, n.k), xp_zm = L.id(n.ip, n.j, n.km);
  const size_t xm_yp = L.id(n.im, n.jp, n.k), yp_zm = L.id(n.i, n.jp, n.km);
  const size_t xm_zp = L.id(n.im, n.j, n.kp), ym_zp = L.id(n.i, n.jm, n.kp);
  const R sx = T::sn(tx[s]), cx = T::cs(tx[s]), sy = T::sn(ty[s]), cy = T::cs(ty[s]);
  const R sz = T::sn(tz[s]), cz = T::cs(tz[s]);
  const R sxm = T::sn(tx[xm]), cxm = T::cs(tx[xm]), sym = T::sn(ty[ym]), cym = T::cs(ty[ym]);
  const R szm = T::sn(tz[zm]), czm = T::cs(tz[zm]);
  const R ar = pr[s], ai = pi[s];
  const R mag2 = ar * ar + ai * ai;
  const R pot = k.hl * (mag2 - k.eta2);
  {
    R lxx = (cx * pr[xp] + sx * pi[xp] - 2 * ar + cxm * pr[xm] - sxm * pi[xm]) / k.dx2;
    R lyy = (cy * pr[yp] + sy * pi[yp] - 2 * ar + cym * pr[ym] - sym * pi[ym]) / k.dy2;
    R lzz = (cz * pr[zp] + sz * pi[zp] - 2 * ar + czm * pr[zm] - szm * pi[zm]) / k.dz2;
    R old = P0[s];
    R vel = (ar - old) / k.dt;
    R acc = lxx + lyy + lzz - pot * ar - k.fric * vel;
    P0[s] = 2 * ar - old + k.dt2 * acc;
  }
  {
    R lxx = (cx * pi[xp] + (-sx) * pr[xp] - 2 * ai + cxm * pi[xm] - (-sxm) * pr[xm]) / k.dx2;
    R lyy = (cy * pi[yp] + (-sy) * pr[yp] - 2 * ai + cym * pi[ym] - (-sym) * pr[ym]) / k.dy2;
    R lzz = (cz * pi[zp] + (-sz) * pr[zp] - 2 * ai + czm * pi[zm] - (-szm) * pr[zm]) / k.dz2;
    R old = P1[s];
    R vel = (ai - old) / k.dt;
    R acc = lxx + lyy + lzz - pot * ai - k.fric * vel;
    P1[s] = 2 * ai - old + k.dt2 * acc;
  }
  const R jx = cx * (ai * pr[xp] - ar * pi[xp]) + sx * (ar * pr[xp] + ai * pi[xp]);
  const R jy = cy * (ai * pr[yp] - ar * pi[yp]) + sy * (ar * pr[yp] + ai * pi[yp]);
  const R jz = cz * (ai * pr[zp] - ar * pi[zp]) + sz * (ar * pr[zp] + ai * pi[zp]);
  const R X = tx[s], Y = ty[s], Z = tz[s];
  const R fxy = (T::sn(X + ty[xp] - tx[yp] - Y) - T::sn(tx[ym] + ty[xp_ym] - X - ty[ym])) / k.dy2;
  const R fxz = (T::sn(X + tz[xp] - tx[zp] - Z) - T::sn(tx[zm] + tz[xp_zm] - X - tz[zm])) / k.dz2;
  const R fyx = (T::sn(Y + tx[yp] - ty[xp] - X) - T::sn(ty[xm] + tx[xm_yp] - Y - tx[xm])) / k.dx2;
  const R fyz = (T::sn(Y + tz[yp] - ty[zp] - Z) - T::sn(ty[zm] + tz[yp_zm] - Y - tz[zm])) / k.dz2;
  const R fzx = (T::sn(Z + tx[zp] - tz[xp] - X) - T::sn(tz[xm] + tx[xm_zp] - Z - tx[xm])) / k.dx2;
  const R fzy = (T::sn(Z + ty[zp] - tz[yp] - Y) - T::sn(tz[ym] + ty[ym_zp] - Z - ty[ym])) / k.dy2;
  const R oX = PX[s], oY = PY[s], oZ = PZ[s];
  R ax_, ay_, az_;
  if (k.form == 0) {
    ax_ = k.m2gg * jx - fxy - fxz - k.damp * ((X - oX) / k.dt);
    ay_ = k.m2gg * jy - fyx - fyz - k.damp * ((Y - oY) / k.dt);
    az_ = k.m2gg * jz - fzx - fzy - k.damp * ((Z - oZ) / k.dt);
  } else {
    ax_ = k.m2gg * jx - fxy - fxz - k.damp * (X - oX) / k.dt;
    ay_ = k.m2gg * jy - fyx - fyz - k.damp * (Y - oY) / k.dt;
    az_ = k.m2gg * jz - fzx - fzy - k.damp * (Z - oZ) / k.dt;
  }
  PX[s] = 2 * X - oX + k.dt2 * ax_;
  PY[s] = 2 * Y - oY + k.dt2 * ay_;
  PZ[s] = 2 * Z - oZ + k.dt2 * az_;
}

template <class R>
void run_steps_r(const ax_params *p, R *st, int ts0, int nsteps, int nthreads) {
  Lat L(p);
  const int e = (p->boundary != AX_BC_PERIODIC) ? 1 : 0, ez = (p->boundary_z != AX_BC_PERIODIC) ? 1 : 0;
  for (int n = ts0; n < ts0 + nsteps; ++n) {
    const int tN = (n + 1) % 2, tP = n % 2;
    const double time = 1 + n * p->dt;
    const double damp = (n < p->damped_nt) ? p->dampFac : 0;
    const double fric = damp + p->alpha * p->scaling / time;
    if (p->boundary == AX_BC_NEUMANN || p->boundary_z == AX_BC_NEUMANN) bc_neumann<R, TrigOf<R>>(p, L, st, tN);
    if (p->boundary == AX_BC_ABSORBING) bc_absorbing<R, TrigOf<R>>(p, L, st, tN, tP);
    RConsts<R> k;
    k.dx2 = (R)(p->dx * p->dx); k.dy2 = (R)(p->dy * p->dy); k.dz2 = (R)(p->dz * p->dz);
    k.dt = (R)p->dt; k.dt2 = (R)(p->dt * p->dt);
    k.hl = (R)(0.5 * p->lambda); k.eta2 = (R)(p->eta * p->eta);
    k.m2gg = (R)(-2 * p->g * p->g);
    k.damp = (R)damp; k.fric = (R)fric;
    k.form = p->theta_damp_form;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
#endif
    for (int i = e; i < L.nx - e; ++i)
      for (int j = e; j < L.ny - e; ++j)
        for (int kk = ez; kk < L.nz - ez; ++kk) ah_site_r<R>(L, st, tN, tP, Nb(L, i, j, kk), k);
  }
}

/* ---------------- global string, momentum-form leapfrog ------------------
 * leapfrog.cpp:166-347, one TimeStep:
 *   Neumann copies of the x and y faces (y faces first, then x faces over
 *   all j, :168-193) unless periodic;
 *   phitt = phixx + phiyy + phizz - (|phi|^2 - 1) phi over the updated sites
 *   from the pre-step field (:198-260);
 *   then, site by site in loop order (:266-347): phit += phitt*dt, the energy
 *   density from central differences of the field as the array holds it at
 *   that moment (neighbours earlier in the loop order already updated) with
 *   the momentum rolled back by dt/2, then phi += phit*dt.
 * Single-threaded here, so the energy has the loop order of a one-thread
 * run (with several OpenMP threads the reference's mix of updated and
 * not-yet-updated neighbours depends on the thread schedule).  State layout
 * [Re phi, Im phi, Re pi, Im pi] x sites; R is the storage/arithmetic type,
 * the energy is accumulated in double. */
template <class R>
void global_steps(const ax_params *p, R *st, int nsteps, double *energy) {
  Lat L(p);
  const int nx = L.nx, ny = L.ny, nz = L.nz;
  const bool per = p->boundary == AX_BC_PERIODIC;
  R *ph[2] = {st, st + L.V}, *pt[2] = {st + 2 * L.V, st + 3 * L.V};
  std::vector<R> tt(2 * L.V);
  const R dx = (R)p->dx, dy = (R)p->dy, dz = (R)p->dz, dt = (R)p->dt;
  const R dx2 = (R)(p->dx * p->dx), dy2 = (R)(p->dy * p->dy), dz2 = (R)(p->dz * p->dz);
  const int i0 = per ? 0 : 1, i1 = per ? nx : nx - 1, j0 = per ? 0 : 1, j1 = per ? ny : ny - 1;
  for (int n = 0; n < nsteps; ++n) {
    if (!per) {
      for (int c = 0; c < 2; ++c) {
        R *f = ph[c];
        for (int i = 1; i < nx - 1; ++i)
          for (int k = 0; k < nz; ++k) {
            f[L.id(i, 0, k)] = f[L.id(i, 1, k)];
            f[L.id(i, ny - 1, k)] = f[L.id(i, ny - 2, k)];
          }
        for (int j = 0; j < ny; ++j)
          for (int k = 0; k < nz; ++k) {
            f[L.id(0, j, k)] = f[L.id(1, j, k)];
            f[L.id(nx - 1, j, k)] = f[L.id(nx - 2, j, k)];
          }
      }
    }
    for (int i = i0; i < i1; ++i)
      for (int j = j0; j < j1; ++j)
        for (int k = 0; k < nz; ++k) {
          Nb b(L, i, j, k);
          const size_t s = L.id(i, j, k);
          const R mag = ph[0][s] * ph[0][s] + ph[1][s] * ph[1][s];
          for (int c = 0; c < 2; ++c) {
            const R *f = ph[c];
            R xx = (f[L.id(b.ip, j, k)] - 2 * f[s] + f[L.id(b.im, j, k)]) / dx2;
            R yy = (f[L.id(i, b.jp, k)] - 2 * f[s] + f[L.id(i, b.jm, k)]) / dy2;
            R zz = (f[L.id(i, j, b.kp)] - 2 * f[s] + f[L.id(i, j, b.km)]) / dz2;
            tt[c * L.V + s] = xx + yy + zz - (mag - 1) * f[s];
          }
        }
    double E = 0;
    for (int i = i0; i < i1; ++i)
      for (int j = j0; j < j1; ++j)
        for (int k = 0; k < nz; ++k) {
          Nb b(L, i, j, k);
          const size_t s = L.id(i, j, k);
          const R mag = ph[0][s] * ph[0][s] + ph[1][s] * ph[1][s];
          double gx[2], gy[2], gz[2];
          for (int c = 0; c < 2; ++c) {
            pt[c][s] += tt[c * L.V + s] * dt;
            const R *f = ph[c];
            gx[c] = (double)((f[L.id(b.ip, j, k)] - f[L.id(b.im, j, k)]) / (2 * dx));
            gy[c] = (double)((f[L.id(i, b.jp, k)] - f[L.id(i, b.jm, k)]) / (2 * dy));
            gz[c] = (double)((f[L.id(i, j, b.kp)] - f[L.id(i, j, b.km)]) / (2 * dz));
          }
          double r0 = (double)pt[0][s] - 0.5 * (double)tt[s] * p->dt;
          double r1 = (double)pt[1][s] - 0.5 * (double)tt[L.V + s] * p->dt;
          double dm = (double)mag - 1;
          double rho = r0 * r0 + r1 * r1 + gx[0] * gx[0] + gx[1] * gx[1] + gy[0] * gy[0] + gy[1] * gy[1] +
                       gz[0] * gz[0] + gz[1] * gz[1] + 0.5 * (dm * dm);
          E += p->dx * p->dy * p->dz * rho;
          for (int c = 0; c < 2; ++c) ph[c][s] += pt[c][s] * dt;
        }
    if (energy) energy[n] = E;
  }
}

}  // namespace

/* ================================ C ABI =================================== */
/* With Neumann faces the reference's diagnostics see the faces its step
 * has just set (evolution.cpp:192-265 precede the site loop): a copy of the
 * state with that copy applied. */
template <class T>
const double *with_faces(const ax_params *p, const double *st, int n, std::vector<double> &buf) {
  if (p->boundary != AX_BC_NEUMANN && p->boundary_z != AX_BC_NEUMANN) return st;
  Lat L(p);
  const int nc = (p->model == AX_MODEL_AH) ? 5 : 7;
  buf.assign(st, st + (size_t)nc * 2 * L.V);
  bc_neumann<double, T>(p, L, buf.data(), (n + 1) % 2);
  return buf.data();
}

extern "C" {

int orc_version(void) { return 1; }

void orc_mt19937(uint32_t seed, size_t n, uint32_t *out) {
  MT m(seed);
  for (size_t i = 0; i < n; ++i) out[i] = m.next();
}

void orc_phases(uint32_t seed, size_t n, double *out) {
  MT m(seed);
  const double pi = 4.0 * atan(1.0);
  for (size_t i = 0; i < n; ++i) out[i] = m.canonical() * (2 * pi - 0.0) + 0.0;
}

/* phi = eta*(cos, sin)(phase) as mpi_evolution.cpp:304-305.  GCC at -O1 and
 * above turns that cos/sin pair of one variable into a single glibc sincos()
 * call (its results differ from separate cos/sin for ~0.1% of arguments), so
 * sincos() is called explicitly to reproduce the compiled reference. */
int orc_init_random_phase(const ax_params *p, uint32_t seed, int mode, double *st) {
  Lat L(p);
  MT m(seed);
  const double pi = 4.0 * atan(1.0);
  if (p->model == AX_MODEL_GLOBAL) {
    /* [Re phi, Im phi, Re pi, Im pi]: the same phases, momentum zero */
    memset(st, 0, sizeof(double) * 4 * L.V);
    for (size_t s = 0; s < L.V; ++s) {
      double ph = m.canonical() * (2 * pi - 0.0) + 0.0;
      double sn, cs;
      sincos(ph, &sn, &cs);
      st[s] = p->eta * cs;
      st[L.V + s] = p->eta * sn;
    }
    return 0;
  }
  const int nc = (p->model == AX_MODEL_AH) ? 5 : 7;
  memset(st, 0, sizeof(double) * (size_t)nc * 2 * L.V);
  for (size_t s = 0; s < L.V; ++s) {
    if (p->model == AX_MODEL_AH) {
      double ph = m.canonical() * (2 * pi - 0.0) + 0.0;
      double sn, cs;
      sincos(ph, &sn, &cs);
      double re = p->eta * cs, im = p->eta * sn;
      for (int t = 0; t < 2; ++t) {
        fld(st, L, 0, t)[s] = re;
        fld(st, L, 1, t)[s] = im;
      }
    } else {
      double p1 = m.canonical() * (2 * pi - 0.0) + 0.0;
      double p2 = (mode == 1) ? p1 : m.canonical() * (2 * pi - 0.0) + 0.0;
      double s1, c1, s2, c2;
      sincos(p1, &s1, &c1);
      sincos(p2, &s2, &c2);
      double a = p->eta1 * c1, b = p->eta1 * s1;
      double c = p->eta2 * c2, d = p->eta2 * s2;
      for (int t = 0; t < 2; ++t) {
        fld(st, L, 0, t)[s] = a; fld(st, L, 1, t)[s] = b;
        fld(st, L, 2, t)[s] = c; fld(st, L, 3, t)[s] = d;
      }
    }
  }
  return 0;
}

/* trig_mode 0: the shared ax_trig.h functions; 1: libm sin/cos. */
int orc_step(const ax_params *p, double *st, int ts0, int nsteps, int trig_mode, int nthreads) {
  if (p->model == AX_MODEL_GLOBAL) return -1;
  if (trig_mode == 1)
    run_steps<TrigLibm>(p, st, ts0, nsteps, nthreads);
  else
    run_steps<TrigAx>(p, st, ts0, nsteps, nthreads);
  return 0;
}

/* Abelian-Higgs step in the storage type of p->precision (st is float* for
 * 32, double* for 64), shared trig only. */
int orc_step_r(const ax_params *p, void *st, int ts0, int nsteps, int nthreads) {
  if (p->model != AX_MODEL_AH) return -1;
  if (p->precision == 32)
    run_steps_r<float>(p, (float *)st, ts0, nsteps, nthreads);
  else
    run_steps_r<double>(p, (double *)st, ts0, nsteps, nthreads);
  return 0;
}

/* String length by chaining crossing points, ic_w_evolution.cpp:961-1087,
 * with its all-pairs search over a list it erases from (one thread: ties go
 * to the later point).  out: total, inside the central half-box, outside. */
int orc_string_length(const double *xyz, size_t n, int nx, int ny, int nz, double dx, double dy, double dz,
                      int pxy, int pz, double *out) {
  std::vector<double> X(n), Y(n), Z(n);
  for (size_t i = 0; i < n; ++i) { X[i] = xyz[3 * i]; Y[i] = xyz[3 * i + 1]; Z[i] = xyz[3 * i + 2]; }
  auto d2 = [&](double ax, double ay, double az, double bx, double by, double bz) {
    double d = 0;
    if (fabs(ax - bx) <= 0.5 * nx * dx) d = (ax - bx) * (ax - bx);
    else if (pxy) d = (nx * dx - fabs(ax - bx)) * (nx * dx - fabs(ax - bx));
    if (fabs(ay - by) <= 0.5 * ny * dy) d += (ay - by) * (ay - by);
    else if (pxy) d += (ny * dy - fabs(ay - by)) * (ny * dy - fabs(ay - by));
    if (fabs(az - bz) <= 0.5 * nz * dz) d += (az - bz) * (az - bz);
    else if (pz) d += (nz * dz - fabs(az - bz)) * (nz * dz - fabs(az - bz));
    return d;
  };
  auto inner = [&](size_t f) {
    return X[f] >= -0.25 * nx * dx && X[f] <= 0.25 * nx * dx && Y[f] >= -0.25 * ny * dy && Y[f] <= 0.25 * ny * dy &&
           Z[f] >= -0.25 * nz * dz && Z[f] <= 0.25 * nz * dz;
  };
  double L0 = 0, Lin = 0, Lout = 0;
  size_t f = 0;
  const size_t n0 = n;
  double b[3] = {0, 0, 0};
  if (n0) { b[0] = X[0]; b[1] = Y[0]; b[2] = Z[0]; }
  for (size_t it = 0; it < n0; ++it) {
    bool found = false;
    const double c = dx * dx + dy * dy + dz * dz;
    double best = c * c;
    size_t bi = 0;
    for (size_t j = 0; j < X.size(); ++j) {
      if (j == f) continue;
      const double d = d2(X[f], Y[f], Z[f], X[j], Y[j], Z[j]);
      if (d <= best) { found = true; best = d; bi = j; }
    }
    if (found) {
      const double seg = sqrt(best);
      L0 += seg;
      (inner(f) ? Lin : Lout) += seg;
      X.erase(X.begin() + f); Y.erase(Y.begin() + f); Z.erase(Z.begin() + f);
      f = bi < f ? bi : bi - 1;
    } else {
      if (d2(X[f], Y[f], Z[f], b[0], b[1], b[2]) <= best) {
        const double seg = sqrt(best);
        L0 += seg;
        (inner(f) ? Lin : Lout) += seg;
      }
      X.erase(X.begin() + f); Y.erase(Y.begin() + f); Z.erase(Z.begin() + f);
      f = 0;
      if (X.empty()) break;
      b[0] = X[0]; b[1] = Y[0]; b[2] = Z[0];
    }
  }
  out[0] = L0; out[1] = Lin; out[2] = Lout;
  return 0;
}

/* Straight-string profile by SOR, singleSemiAnalytic.cpp:54-134, one
 * thread: each sweep relaxes j = 1 .. nx-2 in place in increasing j.  par =
 * {h, w, tol, n, lambda, eta, g}; init = nx (phi, A) pairs or NULL for
 * ic_type 0; out = nx (phi, A) pairs.  Returns the sweep count. */
int orc_sor_profile(int nx, int max_iter, const double *par, const double *init, double *out, double *residual) {
  const double h = par[0], w = par[1], tol = par[2], n = par[3], lambda = par[4], eta = par[5], g = par[6];
  std::vector<double> phi(nx), A(nx);
  for (int i = 0; i < nx; ++i) {
    if (init) { phi[i] = init[2 * i]; A[i] = init[2 * i + 1]; }
    else { phi[i] = eta * tanh(0.5 * h * i); A[i] = 0; }
  }
  phi[0] = 0;
  A[0] = 0;
  double tolTest = 1;
  int iterNum = 0;
  while (tolTest > tol && iterNum < max_iter) {
    tolTest = 0;
    iterNum += 1;
    phi[nx - 1] = phi[nx - 2];
    A[nx - 1] = A[nx - 2];
    for (int j = 1; j < nx - 1; ++j) {
      const double q = (n - g * A[j]) / (j * h);
      const double F0 = (q * q + 0.5 * lambda * (phi[j] * phi[j] - eta * eta)) * phi[j] -
                        (phi[j + 1] - phi[j - 1]) / (2 * j * h * h) -
                        (phi[j + 1] - 2 * phi[j] + phi[j - 1]) / (h * h);
      const double F1 = q * q + 0.5 * lambda * (3 * (phi[j] * phi[j]) - eta * eta) + 2 / (h * h);
      const double G0 = -2 * g * (n - g * A[j]) * (phi[j] * phi[j]) + (A[j + 1] - A[j - 1]) / (2 * j * h * h) -
                        (A[j + 1] - 2 * A[j] + A[j - 1]) / (h * h);
      const double G1 = 2 * ((g * phi[j]) * (g * phi[j])) + 2 / (h * h);
      phi[j] += -w * F0 / F1;
      A[j] += -w * G0 / G1;
      if (fabs(F0) > tolTest) tolTest = fabs(F0);
      if (fabs(G0) > tolTest) tolTest = fabs(G0);
    }
  }
  for (int i = 0; i < nx; ++i) { out[2 * i] = phi[i]; out[2 * i + 1] = A[i]; }
  if (residual) *residual = tolTest;
  return iterNum;
}

/* Straight string from a profile, straight_ic.cpp:59-146: the five
 * ic.txt columns (Re phi, Im phi, dx g Ax, dy g Ay, dz g Az) per site in
 * x-major order into st[5][nx*ny*nz].  The nearest profile point being the
 * last one counts as beyond the profile (the reference reads past its table
 * there).  Returns the number of sites beyond the profile. */
long long orc_straight_string(int nx, int ny, int nz, double dx, double dy, double dz, const double *prof,
                              int nprof, double dr, double n, double g, double *st) {
  const size_t V = (size_t)nx * ny * nz;
  const int x0 = (nx - 1) / 2, y0 = (ny - 1) / 2;
  long long off = 0;
  for (int i = 0; i < nx; ++i) {
    const double x = (i - x0) * dx;
    for (int j = 0; j < ny; ++j) {
      const double y = (j - y0) * dy;
      for (int k = 0; k < nz; ++k) {
        const double distance = sqrt(x * x + y * y);
        const int p = (int)round(distance / dr);
        double phiMag, AMag;
        if (p == 0) {
          phiMag = (prof[2 * (p + 1)] * distance - prof[2 * p] * (distance - dr)) / dr;
          AMag = (prof[2 * (p + 1) + 1] * distance - prof[2 * p + 1] * (distance - dr)) / dr;
        } else if (p < nprof - 1) {
          phiMag = (prof[2 * (p - 1)] * (distance - p * dr) * (distance - (p + 1) * dr) -
                    2 * prof[2 * p] * (distance - (p - 1) * dr) * (distance - (p + 1) * dr) +
                    prof[2 * (p + 1)] * (distance - (p - 1) * dr) * (distance - p * dr)) / (2 * dr * dr);
          AMag = (prof[2 * (p - 1) + 1] * (distance - p * dr) * (distance - (p + 1) * dr) -
                  2 * prof[2 * p + 1] * (distance - (p - 1) * dr) * (distance - (p + 1) * dr) +
                  prof[2 * (p + 1) + 1] * (distance - (p - 1) * dr) * (distance - p * dr)) / (2 * dr * dr);
        } else {
          phiMag = 1;
          AMag = g == 0 ? 0 : n / g;
          ++off;
        }
        double f0, f1, a0, a1;
        if (x == 0) { f0 = 0; a1 = 0; }
        else { f0 = phiMag * x / distance; a1 = AMag * x / (distance * distance); }
        if (y == 0) { f1 = 0; a0 = 0; }
        else { f1 = phiMag * y / distance; a0 = -AMag * y / (distance * distance); }
        const size_t s = ((size_t)i * ny + j) * nz + k;
        st[s] = f0;
        st[V + s] = f1;
        st[2 * V + s] = dx * g * a0;
        st[3 * V + s] = dy * g * a1;
        st[4 * V + s] = dz * g * 0.0;
      }
    }
  }
  return off;
}

/* global model (AX_MODEL_GLOBAL): nsteps TimeSteps of leapfrog.cpp on st
 * ([Re phi, Im phi, Re pi, Im pi], float* for precision 32), the energy of
 * each step into energy[0..nsteps) (may be NULL). */
int orc_global_step(const ax_params *p, void *st, int nsteps, double *energy) {
  if (p->model != AX_MODEL_GLOBAL) return -1;
  if (p->precision == 32)
    global_steps<float>(p, (float *)st, nsteps, energy);
  else
    global_steps<double>(p, (double *)st, nsteps, energy);
  return 0;
}

int orc_diag(const ax_params *p, const double *st0, int n, int trig_mode, double *out) {
  if (p->model == AX_MODEL_GLOBAL) return -1;
  std::vector<double> buf;
  const double *st = trig_mode == 1 ? with_faces<TrigLibm>(p, st0, n, buf) : with_faces<TrigAx>(p, st0, n, buf);
  if (p->model == AX_MODEL_AH) {
    if (trig_mode == 1) diag_ah<TrigLibm>(p, st, n, out);
    else diag_ah<TrigAx>(p, st, n, out);
  } else {
    if (trig_mode == 1) diag_moore<TrigLibm>(p, st, n, out);
    else diag_moore<TrigAx>(p, st, n, out);
  }
  Lat L(p);
  const int nc = (p->model == AX_MODEL_AH) ? 5 : 7;
  int fin = 1;
  for (size_t i = 0; i < (size_t)nc * 2 * L.V; ++i)
    if (!isfinite(st[i])) { fin = 0; break; }
  out[2] = fin;
  return 0;
}

/* field: 0 = phi (AH) / phi1 (Moore), 1 = phi2 (Moore) */
int orc_winding(const ax_params *p, const double *st, int n, int field, int8_t *out) {
  if (p->model == AX_MODEL_GLOBAL) return -1;
  std::vector<double> buf;
  winding_impl(p, with_faces<TrigAx>(p, st, n, buf), n, out, field);
  return 0;
}

int orc_crossings(const ax_params *p, const double *st, int n, double *xyz, size_t cap, size_t *count) {
  if (p->model == AX_MODEL_GLOBAL) return -1;
  Pts P{xyz, cap, 0};
  std::vector<double> buf;
  crossings_impl(p, with_faces<TrigAx>(p, st, n, buf), n, P);
  *count = P.n;
  return 0;
}

/* fn: 0 sin, 1 cos, 2 atan2(x, y), 3 sinf, 4 cosf, 5 ax_sin_small (NaN where
 * ax_sin_is_small(x) is false), 6 ax_sinf_small (NaN where ax_sinf_is_small
 * is false), 7 / 8 the cos member of ax_sincos_small / ax_sincosf_small
 * (NaN outside their domains) */
void orc_trig(int fn, const double *x, const double *y, double *out, size_t n) {
  for (size_t i = 0; i < n; ++i) {
    switch (fn) {
      case 0: out[i] = ax_sin(x[i]); break;
      case 1: out[i] = ax_cos(x[i]); break;
      case 2: out[i] = ax_atan2(x[i], y[i]); break;
      case 5: out[i] = ax_sin_is_small(x[i]) ? ax_sin_small(x[i]) : NAN; break;
      case 6: out[i] = ax_sinf_is_small((float)x[i]) ? (double)ax_sinf_small((float)x[i]) : NAN; break;
      case 7: {
        double sv, cv;
        ax_sincos_small(x[i], &sv, &cv);
        out[i] = ax_sin_is_small(x[i]) ? cv : NAN;
        break;
      }
      case 8: {
        float sv, cv;
        ax_sincosf_small((float)x[i], &sv, &cv);
        out[i] = ax_sinf_is_small((float)x[i]) ? (double)cv : NAN;
        break;
      }
      case 9: out[i] = ::sin(x[i]); break;    /* glibc, as the reference calls it */
      case 10: out[i] = ::cos(x[i]); break;
      case 3: { float s, c; ax_sincosf((float)x[i], &s, &c); out[i] = s; break; }
      default: { float s, c; ax_sincosf((float)x[i], &s, &c); out[i] = c; break; }
    }
  }
}

}  // extern "C"
