/*
 * ax_trig.h - binary64 / binary32 sin, cos, sincos and atan2 with one IEEE
 * operation sequence shared by host code (g++/clang) and gfx950 device code.
 *
 * Why this exists: the lattice update is chaotic, so a one-ulp difference in a
 * single sin() grows to O(1) within a few hundred steps.  Bit-identical
 * trajectories between the GPU kernels and the CPU oracle therefore need the
 * same trigonometric functions on both sides.  Every routine below is written
 * only with +, -, *, /, fma, rint, compares/selects and sign-bit flips, all
 * of which are exactly specified by IEEE 754, so a host build and a device
 * build that both use -ffp-contract=off produce identical bits.
 *
 * Method (published, not vendor code):
 *   - Cody & Waite argument reduction by pi/2 with a three-part constant;
 *     the first product is exact inside the fma, the second is kept as a
 *     double-double (product error recovered with fma), so r_hi + r_lo
 *     carries about 100 bits of x - k*pi/2 for |x| < 2^20.
 *   - Kernels: sin(r) = r + r^3 S(r^2), cos(r) = 1 - r^2/2 + r^4 C(r^2),
 *     with S and C Chebyshev interpolants (fitted here in 50-digit mpmath on
 *     z in [0, (pi/4 * 1.0001)^2], coefficients rounded to double).  The
 *     cos kernel subtracts r^2/2 in two pieces (the compensation technique
 *     described by K.C. Ng for 4.3BSD/fdlibm) so that 1 - r^2/2 is not
 *     rounded twice.
 *   - atan2: octant reduction to t = min/max in [0, 1]; t > tan(pi/8) is
 *     mapped with atan(t) = pi/4 + atan((t-1)/(t+1)), the argument formed
 *     as (min - max)/(min + max) in one division; kernel u + u^3 A(u^2), A
 *     a Chebyshev interpolant on [0, tan(pi/8)^2].
 * Measured accuracy (tests/test_trig.py, against mpmath at 50 digits):
 * sin/cos max 0.85 ulp for |x| < 2^20 (bound tested: 0.9), atan2 max
 * 1.73 ulp (bound tested: 2.5), sincosf max 1.5 ulp.
 * Arguments with |x| >= 2^20 are reduced with growing absolute error (still
 * deterministic and identical on host and device).
 *
 * Signed zeros follow IEEE libm: sin(+-0) = +-0, cos(+-0) = 1.
 */
#ifndef AXIONS_AX_TRIG_H
#define AXIONS_AX_TRIG_H

#include <math.h>
#include <stdint.h>
#include <string.h>
#ifndef __cplusplus
#include <stdbool.h>
#endif

#if defined(__HIPCC__) || defined(__HIP__)
#define AX_HD __host__ __device__ __forceinline__
#else
#define AX_HD static inline
#endif

AX_HD uint64_t ax_dbits(double v) {
  uint64_t b;
  memcpy(&b, &v, sizeof b);
  return b;
}
AX_HD double ax_bitsd(uint64_t b) {
  double v;
  memcpy(&v, &b, sizeof v);
  return v;
}
AX_HD uint32_t ax_fbits(float v) {
  uint32_t b;
  memcpy(&b, &v, sizeof b);
  return b;
}
AX_HD float ax_bitsf(uint32_t b) {
  float v;
  memcpy(&v, &b, sizeof v);
  return v;
}

/* k mod 4 (0..3) of an integer-valued k with |k| < 2^51: adding 1.5 * 2^52
 * is exact and leaves 2^51 + k as the integer held in the low significand
 * bits, whose two lowest bits are k mod 4.  (Larger |k|, infinities and NaN
 * give some value in 0..3, the same on every IEEE target.) */
AX_HD int ax_mod4(double k) { return (int)(ax_dbits(k + 6755399441055744.0) & 3); }
/* the binary32 form, exact for |k| < 2^22 (1.5 * 2^23) */
AX_HD int ax_mod4f(float k) { return (int)(ax_fbits(k + 12582912.0f) & 3); }

/* pi/2 split: PIO2_1 = RN(pi/2), PIO2_2 = RN(pi/2 - PIO2_1),
 * PIO2_3 = RN(pi/2 - PIO2_1 - PIO2_2). */
#define AX_PIO2_1 1.5707963267948966e+00
#define AX_PIO2_2 6.123233995736766e-17
#define AX_PIO2_3 -1.497384905065873e-33
#define AX_TWO_OVER_PI 6.366197723675814e-01

/* sin kernel tail S(z), z = r^2 */
#define AX_S0 -0.16666666666666666
#define AX_S1 0.008333333333333331
#define AX_S2 -0.0001984126984126506
#define AX_S3 2.7557319219335453e-06
#define AX_S4 -2.5052106231157904e-08
#define AX_S5 1.605853141458901e-10
#define AX_S6 -7.586685070091267e-13
/* cos kernel tail C(z) */
#define AX_C0 0.041666666666666664
#define AX_C1 -0.0013888888888888887
#define AX_C2 2.4801587298763433e-05
#define AX_C3 -2.7557317270560326e-07
#define AX_C4 2.0876146024689747e-09
#define AX_C5 -1.1382614868556325e-11

/* Reduce x to r_hi + r_lo = x - k*pi/2 (|r| <~ pi/4) and quadrant q = k mod 4. */
AX_HD void ax_reduce_pio2(double x, double *rh, double *rl, int *q) {
  /* "+ 0.0" turns k = -0 into +0 so that the first fma returns x unchanged
   * (keeps the sign of a zero argument). */
  double k = rint(x * AX_TWO_OVER_PI) + 0.0;
  double r1 = fma(-k, AX_PIO2_1, x);   /* exact: see header comment */
  double w = k * AX_PIO2_2;
  double we = fma(k, AX_PIO2_2, -w);   /* exact error of k*PIO2_2 */
  double h = r1 - w;
  double l = ((r1 - h) - w) - fma(k, AX_PIO2_3, we);
  *rh = h;
  *rl = l;
  *q = ax_mod4(k);
}

AX_HD double ax_sin_kernel(double rh, double rl) {
  double z = rh * rh;
  double v = z * rh;
  double p = fma(z, AX_S6, AX_S5);
  p = fma(z, p, AX_S4);
  p = fma(z, p, AX_S3);
  p = fma(z, p, AX_S2);
  p = fma(z, p, AX_S1);
  p = fma(z, p, AX_S0);
  /* sin(rh + rl) = sin(rh) + rl*(1 - rh^2/2) + O(rl*rh^3) */
  double t = fma(v, p, fma(-0.5 * z, rl, rl));
  return rh + t;
}

AX_HD double ax_cos_kernel(double rh, double rl) {
  double z = rh * rh;
  double ze = fma(rh, rh, -z);        /* exact error of z */
  double p = fma(z, AX_C5, AX_C4);
  p = fma(z, p, AX_C3);
  p = fma(z, p, AX_C2);
  p = fma(z, p, AX_C1);
  p = fma(z, p, AX_C0);
  double hz = 0.5 * z;
  double w = 1.0 - hz;
  /* cos(rh + rl) = 1 - (z + ze)/2 + z^2 C(z) - rh*rl */
  double tail = fma(z * z, p, -fma(rh, rl, 0.5 * ze));
  return w + (((1.0 - w) - hz) + tail);
}

/* sin and cos of x from one reduction.  ax_sin(x) and ax_cos(x) return
 * exactly the same bits as the corresponding member of ax_sincos(x). */
AX_HD void ax_sincos(double x, double *s, double *c) {
  double rh, rl;
  int q;
  ax_reduce_pio2(x, &rh, &rl, &q);
  double sk = ax_sin_kernel(rh, rl);
  double ck = ax_cos_kernel(rh, rl);
  sk = (x == 0.0) ? x : sk;            /* sin(-0) = -0 */
  double s0 = (q & 1) ? ck : sk;
  double c0 = (q & 1) ? sk : ck;
  /* negation is the sign-bit flip: sin takes it in quadrants 2, 3, cos in 1, 2 */
  *s = ax_bitsd(ax_dbits(s0) ^ ((uint64_t)(q & 2) << 62));
  *c = ax_bitsd(ax_dbits(c0) ^ ((uint64_t)((q + 1) & 2) << 62));
}

AX_HD double ax_sin(double x) {
  double s, c;
  ax_sincos(x, &s, &c);
  return s;
}

AX_HD double ax_cos(double x) {
  double s, c;
  ax_sincos(x, &s, &c);
  return c;
}

/* True when the reduction of x picks k = 0, i.e. |RN(x*2/pi)| <= 0.5.  For
 * such x, ax_sin(x) == ax_sin_small(x) bit for bit (k = 0 makes the
 * reduction return r_hi = x, r_lo = +0 and quadrant 0), so a caller may use
 * the short form when every argument it holds qualifies. */
AX_HD bool ax_sin_is_small(double x) { return fabs(x * AX_TWO_OVER_PI) <= 0.5; }

/* ax_sin_kernel(x, +0) without the r_lo term: for finite z that term is
 * fma(-z/2, +0, +0) = +0, and fma(v, p, +0) rounds like v*p except when
 * v*p = -0, which only happens for x = +0 where both sums give +0 (and the
 * sign copy below returns x's zero anyway). */
AX_HD double ax_sin_small(double x) {
  double z = x * x;
  double v = z * x;
  double p = fma(z, AX_S6, AX_S5);
  p = fma(z, p, AX_S4);
  p = fma(z, p, AX_S3);
  p = fma(z, p, AX_S2);
  p = fma(z, p, AX_S1);
  p = fma(z, p, AX_S0);
  /* for 0 < |x| <= pi/4 the sum has the sign of x (|v p| < |x|), so copying
   * x's sign changes nothing there and gives sin(-0) = -0 */
  return copysign(x + v * p, x);
}

/* Both members of ax_sincos(x) for ax_sin_is_small(x): with k = 0 the
 * reduction returns r_hi = x, r_lo = +0 (x - x, -0 - -0 and fma(+0, c, +0)
 * are +0) and quadrant 0, so sin is ax_sin_small(x) and cos is the cos
 * kernel at (x, +0), bit for bit. */
AX_HD void ax_sincos_small(double x, double *s, double *c) {
  *s = ax_sin_small(x);
  *c = ax_cos_kernel(x, 0.0);
}

/* atan kernel tail A(z), z = u^2, |u| <= tan(pi/8) */
#define AX_A0 -0.3333333333333333
#define AX_A1 0.19999999999995513
#define AX_A2 -0.1428571428466482
#define AX_A3 0.11111111015117626
#define AX_A4 -0.09090904572487347
#define AX_A5 0.07692183059894114
#define AX_A6 -0.06664509603255386
#define AX_A7 0.05858132894427922
#define AX_A8 -0.05085365922344248
#define AX_A9 0.03922923713835942
#define AX_A10 -0.019173922285547925
#define AX_TAN_PI8 0.41421356237309503
#define AX_PIO4 7.853981633974483e-01
#define AX_PIO4_LO 3.061616997868383e-17
#define AX_PIO2_LO 6.123233995736766e-17
#define AX_PI 3.141592653589793e+00
#define AX_PI_LO 1.2246467991473532e-16

/* atan2(y, x) in (-pi, pi]; atan2(+-0, x>=+0) = +-0, atan2(+-0, x<0) = +-pi,
 * atan2(0, 0) = 0 with the sign of y.  NaN in, NaN out. */
AX_HD double ax_atan2(double y, double x) {
  double ax = fabs(x), ay = fabs(y);
  bool swap = ay > ax;
  double mx = swap ? ay : ax;
  double mn = swap ? ax : ay;
  /* t = mn / mx in [0, 1]; above tan(pi/8) the argument (t - 1)/(t + 1) is
   * formed as (mn - mx)/(mn + mx): one division instead of two, and t is not
   * rounded on the way */
  bool big = mn > AX_TAN_PI8 * mx;
  double num = big ? mn - mx : mn;
  double den = big ? mn + mx : mx;
  double u = (mx == 0.0) ? 0.0 : num / den;
  double z = u * u;
  double p = fma(z, AX_A10, AX_A9);
  p = fma(z, p, AX_A8);
  p = fma(z, p, AX_A7);
  p = fma(z, p, AX_A6);
  p = fma(z, p, AX_A5);
  p = fma(z, p, AX_A4);
  p = fma(z, p, AX_A3);
  p = fma(z, p, AX_A2);
  p = fma(z, p, AX_A1);
  p = fma(z, p, AX_A0);
  double a = fma(z * u, p, u);            /* atan(u) */
  a = big ? (AX_PIO4 + (a + AX_PIO4_LO)) : a;            /* atan(t) in [0, pi/4] */
  a = swap ? ((AX_PIO2_1 - a) + AX_PIO2_LO) : a;          /* [0, pi/2] */
  bool xneg = signbit(x);
  a = xneg ? ((AX_PI - a) + AX_PI_LO) : a;                /* [0, pi] */
  a = (x != x || y != y) ? (x + y) : a;
  return signbit(y) ? -a : a;
}

/* ---------------------------------------------------------------------
 * binary32 variants (used by the FP32 storage/arithmetic lattice path).
 * Cody-Waite with a two-part constant whose first part has 11 trailing zero
 * bits, degree-9 sin / degree-8 cos Taylor kernels (|r| <= pi/4), accurate
 * to < 1 ulp for |x| < 2^14.  Same code on host and device.
 * --------------------------------------------------------------------- */
#define AX_F_PIO2_1 1.5703125f
#define AX_F_PIO2_2 4.8382679e-04f
#define AX_F_PIO2_3 2.5633441e-12f
#define AX_F_TWO_OVER_PI 0.63661977f

AX_HD void ax_sincosf(float x, float *s, float *c) {
  float k = rintf(x * AX_F_TWO_OVER_PI) + 0.0f;
  float r = fmaf(-k, AX_F_PIO2_1, x);
  r = fmaf(-k, AX_F_PIO2_2, r);
  r = fmaf(-k, AX_F_PIO2_3, r);
  int q = ax_mod4f(k);
  float z = r * r;
  float ps = fmaf(z, 2.7557319e-06f, -1.9841270e-04f);
  ps = fmaf(z, ps, 8.3333333e-03f);
  ps = fmaf(z, ps, -1.6666667e-01f);
  float sk = fmaf(z * r, ps, r);
  float pc = fmaf(z, -2.7557319e-07f, 2.4801587e-05f);
  pc = fmaf(z, pc, -1.3888889e-03f);
  pc = fmaf(z, pc, 4.1666668e-02f);
  float ck = fmaf(z * z, pc, fmaf(-0.5f, z, 1.0f));
  sk = (x == 0.0f) ? x : sk;
  float s0 = (q & 1) ? ck : sk;
  float c0 = (q & 1) ? sk : ck;
  *s = ax_bitsf(ax_fbits(s0) ^ ((uint32_t)(q & 2) << 30));
  *c = ax_bitsf(ax_fbits(c0) ^ ((uint32_t)((q + 1) & 2) << 30));
}

AX_HD float ax_sinf(float x) {
  float s, c;
  ax_sincosf(x, &s, &c);
  return s;
}

/* True when the reduction in ax_sincosf picks k = 0; then the three fmaf
 * steps return r = x exactly and the quadrant is 0, so ax_sinf(x) ==
 * ax_sinf_small(x) bit for bit. */
AX_HD bool ax_sinf_is_small(float x) { return fabsf(x * AX_F_TWO_OVER_PI) <= 0.5f; }

AX_HD float ax_sinf_small(float x) {
  float z = x * x;
  float ps = fmaf(z, 2.7557319e-06f, -1.9841270e-04f);
  ps = fmaf(z, ps, 8.3333333e-03f);
  ps = fmaf(z, ps, -1.6666667e-01f);
  return copysignf(fmaf(z * x, ps, x), x);   /* sign of x, as in ax_sin_small */
}

/* Both members of ax_sincosf(x) for ax_sinf_is_small(x) (r = x, quadrant 0). */
AX_HD void ax_sincosf_small(float x, float *s, float *c) {
  const float z = x * x;
  *s = ax_sinf_small(x);
  float pc = fmaf(z, -2.7557319e-07f, 2.4801587e-05f);
  pc = fmaf(z, pc, -1.3888889e-03f);
  pc = fmaf(z, pc, 4.1666668e-02f);
  *c = fmaf(z * z, pc, fmaf(-0.5f, z, 1.0f));
}

#endif /* AXIONS_AX_TRIG_H */
