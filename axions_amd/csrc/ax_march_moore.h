/*
 * ax_march_moore.h - the Moore-model update (mooreEvol.cpp:202-337: two
 * complex scalars with charges q1, q2 on one U(1) link field, no friction)
 * as a 64-lane wave marching along x over a z chunk of one row, like
 * ah_march.  Same arithmetic and operand order as moore_update in
 * ax_step.hip; the cos/sin of q_f theta_x(s-x) and the plaquette sines at
 * s-x come from the previous plane of the march, y/z-backward values are
 * recomputed, the six plaquette sines at s are computed once and take the
 * short k = 0 form when the whole wave qualifies.
 */
#ifndef AXIONS_AX_MARCH_MOORE_H
#define AXIONS_AX_MARCH_MOORE_H

#include "ax_march.h"

namespace ax {

template <typename R>
struct MooreConsts {
  R dx2, dy2, dz2, dt2, m2gg;
  R hl1, hl2, hl12, eta1sq, eta2sq, q1, q2;
  R inv_dx2, inv_dy2, inv_dz2;
};

/* Division by a divisor that is the same for the whole launch, with the
 * bits of the compiler's IEEE division a / d.  That division is: y = rcp(d)
 * refined by two Newton steps, q = a*y, r = fma(-d, q, a), fma(r, y, q),
 * with range scaling (v_div_scale) and special-case repair (v_div_fixup)
 * around it that change nothing while d and a/d are far from the exponent
 * limits.  With d in [2^-100, 2^100] (checked by the launcher) and a = 0 or
 * |a| in [2^-800, 2^600] the steps after y are done here as they stand, and
 * y once per wave (+-0 take a's sign: the bare steps return +0 for -0).
 *
 * The numerators are sums of products of the current level's values and
 * their sines and cosines.  When every value of that level is 0 or has
 * magnitude in [2^-250, 2^40] (in_moore_band) and the charges satisfy
 * |q| <= 2^8 (checked by the launcher), every trig argument -- q*theta or a
 * plaquette angle, a sum of four link values -- is below 2^50 in
 * magnitude, where ax_reduce_pio2 still finds the quadrant exactly and
 * leaves |r| <~ pi/4, so every sine and cosine lies in [-1, 1] (tested on
 * the host up to 2^50).  Then every term of a numerator (a product of a
 * value and a sine or cosine, or 2 times a value) is 0 or a multiple of
 * 2^-605 and at most 2^41 in magnitude, and every numerator (at most five
 * such terms) is 0 or in [2^-605, 2^44]: the step that
 * writes a level checks its outputs against that band and ORs a per-level
 * flag, and the next step divides this way only when the flag of its
 * current level is clear. */
struct FixedDivisor {
  double d, y;
};
__device__ __forceinline__ FixedDivisor fixed_divisor(double d) {
  const double r0 = __builtin_amdgcn_rcp(d);
  const double r1 = fma(r0, fma(-d, r0, 1.0), r0);
  const double y = fma(r1, fma(-d, r1, 1.0), r1);
  /* the same on every lane: kept in scalar registers */
  const uint64_t b = __double_as_longlong(y);
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)b), hi = __builtin_amdgcn_readfirstlane((uint32_t)(b >> 32));
  return {d, __longlong_as_double((long long)(((uint64_t)hi << 32) | lo))};
}
__device__ __forceinline__ double fixed_div(double a, const FixedDivisor &v) {
  const double q = a * v.y;
  return copysign(fma(fma(-v.d, q, a), v.y, q), a);
}
inline bool fixed_divisor_range(double d) { return d >= 0x1p-100 && d <= 0x1p100; }
/* charges for which q*theta stays below 2^50 for every in-band theta */
inline bool moore_band_charge(int q) { return q >= -256 && q <= 256; }
__device__ __forceinline__ bool in_moore_band(double v) {
  const double m = fabs(v);
  return m <= 0x1p40 && (m >= 0x1p-250 || m == 0);
}

template <typename R>
inline void moore_consts_from(const StepConsts &k, MooreConsts<R> &m) {
  m.dx2 = (R)k.dx2; m.dy2 = (R)k.dy2; m.dz2 = (R)k.dz2; m.dt2 = (R)k.dt2; m.m2gg = (R)k.m2gg;
  m.hl1 = (R)k.hl1; m.hl2 = (R)k.hl2; m.hl12 = (R)k.hl12; m.eta1sq = (R)k.eta1sq; m.eta2sq = (R)k.eta2sq;
  m.q1 = (R)k.q1; m.q2 = (R)k.q2;
  m.inv_dx2 = (R)1 / m.dx2; m.inv_dy2 = (R)1 / m.dy2; m.inv_dz2 = (R)1 / m.dz2;
}

/* BUF: the loop's loads and stores are raw buffer accesses, one resource per
 * field spanning both of its levels (planes 2c, 2c+1; the caller checks that
 * two planes fit 32-bit offsets), with the streaming hint on the previous
 * level as in ah_march; otherwise plain global accesses from each x plane's
 * 64-bit address (field planes of any size, 1001^3 FP64 included).
 * FASTDIV (FP64, divisors not powers of two): divisions by dx^2, dy^2,
 * dz^2 through FixedDivisor; the caller has checked the current level's
 * band flag.  CHECK: OR 1 into *bad when a written value lies outside the
 * band (every lane of the wave must reach the end). */
template <typename R, bool POW2, bool BUF = false, bool FASTDIV = false, bool CHECK = false>
__device__ __forceinline__ void moore_march(const MooreConsts<R> &k, const Levels<R> &L, int NX, int NY, int NZ, int y,
                                            int zc, int x0, int x1, bool zfix = false, int *bad = nullptr) {
  typedef TrigR<R> T;
  const int lane = (int)(threadIdx.x & 63);
  const int zr = zc * 64 + lane;
  const bool inz = zr < NZ;
  const int z = inz ? zr : NZ - 1;
  const bool act = inz && !(zfix && (zr == 0 || zr == NZ - 1));
  const int ym = y == 0 ? NY - 1 : y - 1, yp = y == NY - 1 ? 0 : y + 1;
  const uint32_t E = (uint32_t)sizeof(R);
  const uint32_t oz = (uint32_t)z * E;
  const uint32_t ozm = (uint32_t)(z == 0 ? NZ - 1 : z - 1) * E;
  const uint32_t ozp = (uint32_t)(z == NZ - 1 ? 0 : z + 1) * E;
  const uint32_t PLB = (uint32_t)NY * (uint32_t)NZ * E;
  const uint32_t R0 = (uint32_t)y * NZ * E, RM = (uint32_t)ym * NZ * E, RP = (uint32_t)yp * NZ * E;
  const R *F[7] = {L.cur[0], L.cur[1], L.cur[2], L.cur[3], L.cur[4], L.cur[5], L.cur[6]};
  R *O[7] = {L.prv[0], L.prv[1], L.prv[2], L.prv[3], L.prv[4], L.prv[5], L.prv[6]};
  const R *TX = F[4], *TY = F[5], *TZ = F[6];
  const R q[2] = {k.q1, k.q2};
  FixedDivisor fdx = {}, fdy = {}, fdz = {};
  if constexpr (FASTDIV) {
    fdx = fixed_divisor((double)k.dx2);
    fdy = fixed_divisor((double)k.dy2);
    fdz = fixed_divisor((double)k.dz2);
  }
  /* a / d_a for axis a (x, y, z) */
  auto dvx = [&](R a) -> R { if constexpr (FASTDIV) return fixed_div(a, fdx); else return mdiv<R, POW2>(a, k.dx2, k.inv_dx2); };
  auto dvy = [&](R a) -> R { if constexpr (FASTDIV) return fixed_div(a, fdy); else return mdiv<R, POW2>(a, k.dy2, k.inv_dy2); };
  auto dvz = [&](R a) -> R { if constexpr (FASTDIV) return fixed_div(a, fdz); else return mdiv<R, POW2>(a, k.dz2, k.inv_dz2); };
  bool inband = true;
  /* field c's lower plane, and the byte offsets of the two levels from it */
  __amdgpu_buffer_rsrc_t FR[7];
  uint32_t offC = 0, offP = 0;
  if (BUF) {
    for (int c = 0; c < 7; ++c) FR[c] = plane_rsrc(L.cur[c] < L.prv[c] ? (const void *)L.cur[c] : (const void *)L.prv[c]);
    const bool curLow = L.cur[0] < L.prv[0];
    const uint32_t lev = (uint32_t)(curLow ? (const char *)L.prv[0] - (const char *)L.cur[0]
                                           : (const char *)L.cur[0] - (const char *)L.prv[0]);
    offC = curLow ? 0u : lev;
    offP = curLow ? lev : 0u;
  }
  /* current-level value of field c in x plane x at (row byte offset ro,
   * lane offset lo); without BUF the plane's 64-bit address is formed on
   * the scalar unit, so planes of any size work */
  const size_t NYNZ = (size_t)NY * NZ;
  auto LC = [&](int c, int x, uint32_t ro, uint32_t lo) -> R {
    if (BUF) return bld<R>(FR[c], lo, offC + (uint32_t)x * PLB + ro);
    return ldo(F[c] + (size_t)x * NYNZ, ro, lo);
  };
  auto LG = [&](const R *f, int x, uint32_t ro, uint32_t lo) -> R { return ldo(f + (size_t)x * NYNZ, ro, lo); };

  /* carry from plane x-1: phi_f(s-x), cos/sin(q_f theta_x(s-x)), sines at s-x */
  const int xm0 = x0 == 0 ? NX - 1 : x0 - 1;
  R Pm[4], cxm[2], sxm[2], Qyx_m, Qzx_m;
  {
    for (int c = 0; c < 4; ++c) Pm[c] = LG(F[c], xm0, R0, oz);
    const R X = LG(TX, xm0, R0, oz), Y = LG(TY, xm0, R0, oz), Z = LG(TZ, xm0, R0, oz);
    for (int f = 0; f < 2; ++f) T::sc(q[f] * X, &sxm[f], &cxm[f]);
    Qyx_m = T::sn(Y + LG(TX, xm0, RP, oz) - LG(TY, x0, R0, oz) - X);
    Qzx_m = T::sn(Z + LG(TX, xm0, R0, ozp) - LG(TZ, x0, R0, oz) - X);
  }
  R C[7];
  for (int c = 0; c < 7; ++c) C[c] = LG(F[c], x0, R0, oz);

  for (int x = x0; x < x1; ++x) {
    const int xp = x == NX - 1 ? 0 : x + 1;
    const uint32_t so = (uint32_t)x * PLB + R0;   /* BUF only */
    R Nx[7];
    for (int c = 0; c < 7; ++c) Nx[c] = LC(c, xp, R0, oz);
    const R X = C[4], Y = C[5], Z = C[6];
    const R TXyp = LC(4, x, RP, oz), TZyp = LC(6, x, RP, oz);
    const R TXzp = LC(4, x, R0, ozp), TYzp = LC(5, x, R0, ozp);
    const R TXym = LC(4, x, RM, oz), TYym = LC(5, x, RM, oz), TZym = LC(6, x, RM, oz);
    const R TYxpym = LC(5, xp, RM, oz), TYymzp = LC(5, x, RM, ozp);
    const R TXzm = LC(4, x, R0, ozm), TYzm = LC(5, x, R0, ozm), TZzm = LC(6, x, R0, ozm);
    const R TZxpzm = LC(6, xp, R0, ozm), TZypzm = LC(6, x, RP, ozm);

    /* plaquette sines (mooreEvol.cpp:310-333 operand order) */
    const R a_xy = X + Nx[5] - TXyp - Y;
    const R a_xz = X + Nx[6] - TXzp - Z;
    const R a_yx = Y + TXyp - Nx[5] - X;
    const R a_yz = Y + TZyp - TYzp - Z;
    const R a_zx = Z + TXzp - Nx[6] - X;
    const R a_zy = Z + TYzp - TZyp - Y;
    const R b_xy = TXym + TYxpym - X - TYym;
    const R b_zy = TZym + TYymzp - Z - TYym;
    const R b_xz = TXzm + TZxpzm - X - TZzm;
    const R b_yz = TYzm + TZypzm - Y - TZzm;
    R Qxy, Qxz, Qyx, Qyz, Qzx, Qzy, Qxy_m, Qzy_m, Qxz_m, Qyz_m;
    const uint64_t small = small_lanes(a_xy) & small_lanes(a_xz) & small_lanes(a_yx) & small_lanes(a_yz) &
                           small_lanes(a_zx) & small_lanes(a_zy) & small_lanes(b_xy) & small_lanes(b_zy) &
                           small_lanes(b_xz) & small_lanes(b_yz);
    if (small == __builtin_amdgcn_read_exec()) {
      Qxy = T::sn_small(a_xy); Qxz = T::sn_small(a_xz); Qyx = T::sn_small(a_yx);
      Qyz = T::sn_small(a_yz); Qzx = T::sn_small(a_zx); Qzy = T::sn_small(a_zy);
      Qxy_m = T::sn_small(b_xy); Qzy_m = T::sn_small(b_zy); Qxz_m = T::sn_small(b_xz); Qyz_m = T::sn_small(b_yz);
    } else {
      Qxy = T::sn(a_xy); Qxz = T::sn(a_xz); Qyx = T::sn(a_yx);
      Qyz = T::sn(a_yz); Qzx = T::sn(a_zx); Qzy = T::sn(a_zy);
      Qxy_m = T::sn(b_xy); Qzy_m = T::sn(b_zy); Qxz_m = T::sn(b_xz); Qyz_m = T::sn(b_yz);
    }
    const R fxy = dvy(Qxy - Qxy_m);
    const R fxz = dvz(Qxz - Qxz_m);
    const R fyx = dvx(Qyx - Qyx_m);
    const R fyz = dvz(Qyz - Qyz_m);
    const R fzx = dvx(Qzx - Qzx_m);
    const R fzy = dvy(Qzy - Qzy_m);

    const R m1 = C[0] * C[0] + C[1] * C[1];
    const R m2 = C[2] * C[2] + C[3] * C[3];
    const R pot[2] = {k.hl1 * (m1 - k.eta1sq) + k.hl12 * (m2 - k.eta2sq),
                      k.hl2 * (m2 - k.eta2sq) + k.hl12 * (m1 - k.eta1sq)};
    R cur[2][3], acc[2][2], cxs[2], sxs[2];
#pragma unroll
    for (int f = 0; f < 2; ++f) {
      const int cr = 2 * f, ci = 2 * f + 1;
      R sx, cx, sy, cy, sz, cz, sym_, cym, szm, czm;
      T::sc(q[f] * X, &sx, &cx);
      T::sc(q[f] * Y, &sy, &cy);
      T::sc(q[f] * Z, &sz, &cz);
      T::sc(q[f] * TYym, &sym_, &cym);
      T::sc(q[f] * TZzm, &szm, &czm);
      const R ar = C[cr], ai = C[ci];
      const R rxp = Nx[cr], ixp = Nx[ci];
      const R ryp = LC(cr, x, RP, oz), iyp = LC(ci, x, RP, oz);
      const R rym = LC(cr, x, RM, oz), iym = LC(ci, x, RM, oz);
      const R rzp = LC(cr, x, R0, ozp), izp = LC(ci, x, R0, ozp);
      const R rzm = LC(cr, x, R0, ozm), izm = LC(ci, x, R0, ozm);
      {
        R lxx = dvx(cx * rxp + sx * ixp - 2 * ar + cxm[f] * Pm[cr] - sxm[f] * Pm[ci]);
        R lyy = dvy(cy * ryp + sy * iyp - 2 * ar + cym * rym - sym_ * iym);
        R lzz = dvz(cz * rzp + sz * izp - 2 * ar + czm * rzm - szm * izm);
        acc[f][0] = lxx + lyy + lzz - pot[f] * ar;
      }
      {
        R lxx = dvx(cx * ixp + (-sx) * rxp - 2 * ai + cxm[f] * Pm[ci] - (-sxm[f]) * Pm[cr]);
        R lyy = dvy(cy * iyp + (-sy) * ryp - 2 * ai + cym * iym - (-sym_) * rym);
        R lzz = dvz(cz * izp + (-sz) * rzp - 2 * ai + czm * izm - (-szm) * rzm);
        acc[f][1] = lxx + lyy + lzz - pot[f] * ai;
      }
      cur[f][0] = q[f] * (cx * (ai * rxp - ar * ixp) + sx * (ar * rxp + ai * ixp));
      cur[f][1] = q[f] * (cy * (ai * ryp - ar * iyp) + sy * (ar * ryp + ai * iyp));
      cur[f][2] = q[f] * (cz * (ai * rzp - ar * izp) + sz * (ar * rzp + ai * izp));
      cxs[f] = cx;
      sxs[f] = sx;
    }
    const R atx = k.m2gg * (cur[0][0] + cur[1][0]) - fxy - fxz;
    const R aty = k.m2gg * (cur[0][1] + cur[1][1]) - fyx - fyz;
    const R atz = k.m2gg * (cur[0][2] + cur[1][2]) - fzx - fzy;
    /* previous level at s: read once, overwritten with level n+1 */
    auto LP = [&](int c) -> R {
      if (BUF) return bld<R, kStream>(FR[c], oz, offP + so);
      return ldo((const R *)O[c] + (size_t)x * NYNZ, R0, oz);
    };
    R Onew[7];
#pragma unroll
    for (int c = 0; c < 4; ++c) Onew[c] = 2 * C[c] - LP(c) + k.dt2 * acc[c >> 1][c & 1];
    Onew[4] = 2 * X - LP(4) + k.dt2 * atx;
    Onew[5] = 2 * Y - LP(5) + k.dt2 * aty;
    Onew[6] = 2 * Z - LP(6) + k.dt2 * atz;
    if (act) {
      for (int c = 0; c < 7; ++c) {
        if (BUF) bst<R, kStream>(FR[c], oz, offP + so, Onew[c]);
        else sto(O[c] + (size_t)x * NYNZ, R0, oz, Onew[c]);
      }
      if constexpr (CHECK)
        for (int c = 0; c < 7; ++c) inband = inband && in_moore_band((double)Onew[c]);
    }
    /* carry */
    for (int c = 0; c < 4; ++c) Pm[c] = C[c];
    for (int f = 0; f < 2; ++f) { cxm[f] = cxs[f]; sxm[f] = sxs[f]; }
    Qyx_m = Qyx;
    Qzx_m = Qzx;
    for (int c = 0; c < 7; ++c) C[c] = Nx[c];
  }
  if constexpr (CHECK) {
    if (__builtin_amdgcn_ballot_w64(!inband) != 0 && (threadIdx.x & 63) == 0) atomicOr(bad, 1);
  }
}

}  // namespace ax

#endif
