/*
 * ax_march.h - the Abelian-Higgs site update as one 64-lane wave marching
 * along x over a 64-site z chunk of one (y) row.  Shared by the
 * single-lattice tiled kernel (ax_step_tiled.hip) and the batched ensemble
 * kernel (ax_pool.hip).
 *
 * Same arithmetic, operand order and results as the site kernel
 * (mpi_evolution.cpp:418-501) for FP64, and as its FP32 instantiation for
 * FP32; what changes is how often each sine and cosine is evaluated:
 *   - x-backward values (phi(s-x), cos/sin of theta_x(s-x), and the
 *     plaquette sines whose copy at s-x is needed) come from the previous
 *     iteration of the march, kept in registers;
 *   - z-backward trig (cos/sin theta_z(s-z) and the F_xz, F_yz sines at
 *     s-z) is lane - 1's own value, shifted across by DPP; the chunk's first
 *     lane takes values evaluated for the site before the chunk in one
 *     vector pass per few planes (one plane per lane, kept in LDS);
 *   - y-backward values are recomputed from their operands with the same
 *     expressions the neighbour uses for its own forward values;
 *   - the six plaquette sines at s are computed once (the second sine of
 *     each F term in mpi_evolution.cpp:471-487 is the first one of the
 *     neighbour, written with the same operands in the same order), and take
 *     the short k = 0 form of the sine when every lane's argument qualifies,
 *     which holds almost always because plaquette angles stay below pi/4
 *     while link angles drift.
 * Each field value is loaded once per step by the wave that owns it: plane
 * x+1 (current level) is loaded as the next iteration's centre and carried;
 * the y/z neighbour rows are meant to be cache hits (at 512^3 FP64 the L2 ->
 * fabric reads are 1.25x the algorithmic bytes without the LDS prefetch
 * below and 1.6x with it, part of them served by the Infinity Cache).  The loop's loads and stores are raw
 * buffer accesses (bld/bst below) whose addresses cost no vector
 * instructions; the previous level's reads and the new level's writes carry
 * the streaming hint so that they do not push those rows out of L2.
 *
 * A chunk may be partial (nz not a multiple of 64): lanes past the row end
 * load the last site of the row and store nothing; values shifted from
 * lane - 1 only reach active lanes from active lanes, so partial chunks
 * need no special handling beyond that.
 */
#ifndef AXIONS_AX_MARCH_H
#define AXIONS_AX_MARCH_H

#include "ax_device.h"
#include "ax_energy.h"

namespace ax {

template <typename R>
struct TrigR;
template <>
struct TrigR<double> {
  __device__ static void sc(double x, double *s, double *c) { ax_sincos(x, s, c); }
  __device__ static double sn(double x) { return ax_sin(x); }
  __device__ static bool small(double x) { return ax_sin_is_small(x); }
  __device__ static double sn_small(double x) { return ax_sin_small(x); }
};
template <>
struct TrigR<float> {
  __device__ static void sc(float x, float *s, float *c) { ax_sincosf(x, s, c); }
  __device__ static float sn(float x) { return ax_sinf(x); }
  __device__ static bool small(float x) { return ax_sinf_is_small(x); }
  __device__ static float sn_small(float x) { return ax_sinf_small(x); }
};

/* Per-step scalars in the arithmetic type R.  FP64: the values of
 * StepConsts.  FP32: those doubles rounded once to float, exactly as the
 * site kernel's (R)k.x casts do. */
template <typename R>
struct MarchConsts {
  R dx2, dy2, dz2, dt, dt2, hl, eta2, m2gg, damp, fric;
  R inv_dx2, inv_dy2, inv_dz2;  /* exact reciprocals, used when all d^2 are powers of two */
  R lim;                        /* largest |d| with d/dt finite (plain-mode guard) */
  int theta_damp_form;
  int plain;                    /* fric and damp both +0 in R */
};

template <typename R>
__host__ __device__ inline void march_consts_from(const StepConsts &k, R lim, MarchConsts<R> &m) {
  m.dx2 = (R)k.dx2; m.dy2 = (R)k.dy2; m.dz2 = (R)k.dz2;
  m.dt = (R)k.dt; m.dt2 = (R)k.dt2;
  m.hl = (R)k.hl; m.eta2 = (R)k.eta2; m.m2gg = (R)k.m2gg;
  m.damp = (R)k.damp; m.fric = (R)k.fric;
  m.inv_dx2 = (R)1 / m.dx2; m.inv_dy2 = (R)1 / m.dy2; m.inv_dz2 = (R)1 / m.dz2;
  m.lim = lim;
  m.theta_damp_form = k.theta_damp_form;
  m.plain = m.fric == 0 && !signbit(m.fric) && m.damp == 0 && !signbit(m.damp);
}

template <typename R, bool POW2>
__device__ __forceinline__ R mdiv(R a, R d, R inv) {
  if (POW2) return a * inv;   /* d is a power of two: a/d == a*(1/d) exactly */
  return a / d;
}

/* base + wave-uniform row byte offset + per-lane byte offset; both offsets
 * are 32-bit (every field plane is below 4 GiB), which lets the compiler use
 * the scalar-base + 32-bit vector-offset load form. */
typedef __attribute__((address_space(1))) char gchar;
__device__ __forceinline__ gchar *uniform_row(const void *base, uint32_t row) {
  /* readfirstlane marks the row address wave-uniform and keeps the compiler
   * from folding the lane offset into a 64-bit per-lane address; the global
   * address space keeps the loads global (not flat) */
  const uint64_t v = (uint64_t)((const char *)base + row);
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return (gchar *)(((uint64_t)hi << 32) | lo);
}
template <typename R>
__device__ __forceinline__ R ldo(const R *base, uint32_t row, uint32_t lane_off) {
  return *(const __attribute__((address_space(1))) R *)(uniform_row(base, row) + lane_off);
}
template <typename R>
__device__ __forceinline__ void sto(R *base, uint32_t row, uint32_t lane_off, R v) {
  *(__attribute__((address_space(1))) R *)(uniform_row(base, row) + lane_off) = v;
}

/* Raw buffer access for the march's inner loop: one buffer resource per
 * field plane (its base address in SGPRs), a wave-uniform byte offset
 * (plane + row) as the scalar offset and the lane's z offset as the vector
 * offset, so no load or store spends vector instructions on its address.
 * Every plane is below 4 GiB (the tiled kernels' condition), so all offsets
 * fit the 32-bit fields; the range check is disabled (num_records = ~0).
 * kStream (the SLC bit) marks data used once per step -- the previous
 * level's reads and the new level's writes -- as streaming, which leaves L2
 * to the current level's rows that neighbouring waves read again. */
typedef unsigned int ax_u2 __attribute__((ext_vector_type(2)));
constexpr int kStream = 2;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const void *plane) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(plane), 0, -1, 0x00020000);
}
template <typename R, int AUX = 0>
__device__ __forceinline__ R bld(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  if constexpr (sizeof(R) == 8)
    return __builtin_bit_cast(R, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, AUX));
  else
    return __builtin_bit_cast(R, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, AUX));
}
template <typename R, int AUX = 0>
__device__ __forceinline__ void bst(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff, R v) {
  if constexpr (sizeof(R) == 8)
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(ax_u2, v), r, (int)voff, (int)soff, AUX);
  else
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), r, (int)voff, (int)soff, AUX);
}

/* Everything one site update reads for one x plane besides the carried
 * values: plane x+1 at (y,z) and the previous level at s (HBM), and the
 * neighbour rows/columns of plane x and x+1 (cache hits). */
template <typename R>
struct PlaneOps {
  R PRn, PIn, TXn, TYn, TZn;
  R oPR, oPI, oX, oY, oZ;
  R PRyp, PIyp, PRym, PIym, PRzp, PIzp, PRzm, PIzm;
  R TXyp, TZyp, TXzp, TYzp, TXym, TYym, TZym, TYxpym, TYymzp;
};

/* Values carried from plane x-1 into plane x of the march. */
template <typename R>
struct MarchCarry {
  R PRm, PIm;          /* phi(s-x) */
  R cxm, sxm;          /* cos/sin theta_x(s-x) */
  R Qyx_m, Qzx_m;      /* plaquette sines at s-x needed by F_yx, F_zx */
  R TXm;               /* theta_x(s-x) (read only by the divergence of the fused diagnostics) */
};

/* What the site before a chunk (z0 - 1) contributes to the chunk's first
 * lane: sin/cos theta_z there and its F_xz, F_yz plaquette sines.  Every
 * other lane takes them from lane - 1, which evaluates them for its own
 * site with the same operands in the same order. */
template <typename R>
struct ZEdge {
  R s, c, qxz, qyz;
  R tz;                /* theta_z there (the fused diagnostics' divergence) */
};

/* The link trigonometry and plaquette angles of one site update, handed to
 * the diagnostics a fused step takes of its current level. */
template <typename R>
struct SiteTrig {
  R cx, sx, cy, sy, cz, sz;    /* cos/sin theta_a(s) */
  R cym, sym, czm, szm;        /* cos/sin theta_y(s-y), theta_z(s-z) */
  R a_xy, a_xz, a_yz;          /* plaquette angles at s */
};

/* The new previous-level values (level n+1) of one site. */
template <typename R>
struct SiteOut {
  R X, Y, Z, PR, PI;
};

/* Operands of one site update, by name.  The site is s = (x, y, z); "n"
 * is plane x+1, "o" the previous level at s, yp/ym/zp/zm the neighbour
 * rows/columns (xp_ym = (x+1, y-1), ym_zp = (x, y-1, z+1), ...). */
enum Op {
  kPRc, kPIc, kTXc, kTYc, kTZc,
  kPRn, kPIn, kTXn, kTYn, kTZn,
  koPR, koPI, koX, koY, koZ,
  kPRyp, kPIyp, kPRym, kPIym, kPRzp, kPIzp, kPRzm, kPIzm,
  kTXyp, kTZyp, kTXzp, kTYzp, kTXym, kTYym, kTZym, kTYxpym, kTYymzp
};

/* Operands held in registers (the register march). */
template <typename R>
struct RegOps {
  const PlaneOps<R> &D;
  R PRc, PIc, TXc, TYc, TZc;
  __device__ __forceinline__ R operator()(int id) const {
    switch (id) {
      case kPRc: return PRc; case kPIc: return PIc; case kTXc: return TXc; case kTYc: return TYc; case kTZc: return TZc;
      case kPRn: return D.PRn; case kPIn: return D.PIn; case kTXn: return D.TXn; case kTYn: return D.TYn;
      case kTZn: return D.TZn; case koPR: return D.oPR; case koPI: return D.oPI; case koX: return D.oX;
      case koY: return D.oY; case koZ: return D.oZ; case kPRyp: return D.PRyp; case kPIyp: return D.PIyp;
      case kPRym: return D.PRym; case kPIym: return D.PIym; case kPRzp: return D.PRzp; case kPIzp: return D.PIzp;
      case kPRzm: return D.PRzm; case kPIzm: return D.PIzm; case kTXyp: return D.TXyp; case kTZyp: return D.TZyp;
      case kTXzp: return D.TXzp; case kTYzp: return D.TYzp; case kTXym: return D.TXym; case kTYym: return D.TYym;
      case kTZym: return D.TZym; case kTYxpym: return D.TYxpym; default: return D.TYymzp;
    }
  }
};

/* The value of lane-1 (lane 0: `edge`) and of lane+1 (lane 63: `edge`),
 * by DPP wave shifts; every lane of the wave must be active. */
__device__ __forceinline__ double lane_prev(double v, double edge) {
  const long long b = __double_as_longlong(v), e = __double_as_longlong(edge);
  const int lo = __builtin_amdgcn_update_dpp((int)e, (int)b, 0x138, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(e >> 32), (int)(b >> 32), 0x138, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double lane_next(double v, double edge) {
  const long long b = __double_as_longlong(v), e = __double_as_longlong(edge);
  const int lo = __builtin_amdgcn_update_dpp((int)e, (int)b, 0x130, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(e >> 32), (int)(b >> 32), 0x130, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ float lane_prev(float v, float edge) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(edge), __float_as_int(v), 0x138, 0xF, 0xF, false));
}
__device__ __forceinline__ float lane_next(float v, float edge) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(edge), __float_as_int(v), 0x130, 0xF, 0xF, false));
}

/* The lanes whose sine argument a takes the short form. */
template <typename R>
__device__ __forceinline__ uint64_t small_lanes(R a) {
  return __builtin_amdgcn_ballot_w64(TrigR<R>::small(a));
}

/* The lanes where |d| exceeds lim or is NaN. */
template <typename R>
__device__ __forceinline__ uint64_t big_lanes(R d, R lim) {
  return __builtin_amdgcn_ballot_w64(!(fabs(d) <= lim));
}

/* One site of the fused in-place update (mpi_evolution.cpp:418-501) from its
 * operands g(Op), the carry from s-x, which it advances to s, and the
 * chunk-edge values ze (used by lane 0 only).  Every lane of the wave must
 * call it (it votes across the wave with ballots and shifts values between
 * lanes). */
template <typename R, bool POW2, typename G, bool TR = false>
__device__ __forceinline__ void ah_site_update(const MarchConsts<R> &k, bool plain, const G &g, MarchCarry<R> &m,
                                               const ZEdge<R> &ze, SiteOut<R> &o, SiteTrig<R> *tr = nullptr) {
  typedef TrigR<R> T;
  const R dx2 = k.dx2, dy2 = k.dy2, dz2 = k.dz2, dt = k.dt, dt2 = k.dt2;
  const R hl = k.hl, eta2 = k.eta2, m2gg = k.m2gg, fric = k.fric, damp = k.damp;
  /* links at s and s-y; theta_z(s-z) is lane - 1's theta_z(s) */
  const R TXc = g(kTXc), TYc = g(kTYc), TZc = g(kTZc), TYym = g(kTYym);
  R cx, sx, cy, sy, cz, sz, cym, sym_;
  T::sc(TXc, &sx, &cx);
  T::sc(TYc, &sy, &cy);
  T::sc(TZc, &sz, &cz);
  T::sc(TYym, &sym_, &cym);
  const R szm = lane_prev(sz, ze.s), czm = lane_prev(cz, ze.c);
  /* plaquette angles, reference operand order */
  const R TYn = g(kTYn), TZn = g(kTZn), TXyp = g(kTXyp), TXzp = g(kTXzp), TZyp = g(kTZyp), TYzp = g(kTYzp);
  const R a_xy = TXc + TYn - TXyp - TYc;
  const R a_xz = TXc + TZn - TXzp - TZc;
  const R a_yx = TYc + TXyp - TYn - TXc;
  const R a_yz = TYc + TZyp - TYzp - TZc;
  const R a_zx = TZc + TXzp - TZn - TXc;
  const R a_zy = TZc + TYzp - TZyp - TYc;
  const R b_xy = g(kTXym) + g(kTYxpym) - TXc - TYym;
  const R b_zy = g(kTZym) + g(kTYymzp) - TZc - TYym;
  R Qxy, Qxz, Qyx, Qyz, Qzx, Qzy, Qxy_m, Qzy_m;
  /* the wave takes the short form when all its lanes' arguments qualify:
   * one ballot per argument, combined on the scalar unit */
  const uint64_t small = small_lanes(a_xy) & small_lanes(a_xz) & small_lanes(a_yx) & small_lanes(a_yz) &
                         small_lanes(a_zx) & small_lanes(a_zy) & small_lanes(b_xy) & small_lanes(b_zy);
  if (small == __builtin_amdgcn_read_exec()) {
    Qxy = T::sn_small(a_xy); Qxz = T::sn_small(a_xz); Qyx = T::sn_small(a_yx);
    Qyz = T::sn_small(a_yz); Qzx = T::sn_small(a_zx); Qzy = T::sn_small(a_zy);
    Qxy_m = T::sn_small(b_xy); Qzy_m = T::sn_small(b_zy);
  } else {
    Qxy = T::sn(a_xy); Qxz = T::sn(a_xz); Qyx = T::sn(a_yx);
    Qyz = T::sn(a_yz); Qzx = T::sn(a_zx); Qzy = T::sn(a_zy);
    Qxy_m = T::sn(b_xy); Qzy_m = T::sn(b_zy);
  }
  if constexpr (TR) {
    tr->cx = cx; tr->sx = sx; tr->cy = cy; tr->sy = sy; tr->cz = cz; tr->sz = sz;
    tr->cym = cym; tr->sym = sym_; tr->czm = czm; tr->szm = szm;
    tr->a_xy = a_xy; tr->a_xz = a_xz; tr->a_yz = a_yz;
  }
  /* the second sines of F_xz, F_yz (arguments TXzm + TZxpzm - TXc - TZzm and
   * TYzm + TZypzm - TYc - TZzm) are lane - 1's first ones */
  const R Qxz_m = lane_prev(Qxz, ze.qxz), Qyz_m = lane_prev(Qyz, ze.qyz);

  /* gauge field first (frees the plaquette sines early) */
  const R fxy = mdiv<R, POW2>(Qxy - Qxy_m, dy2, k.inv_dy2);
  const R fxz = mdiv<R, POW2>(Qxz - Qxz_m, dz2, k.inv_dz2);
  const R fyx = mdiv<R, POW2>(Qyx - m.Qyx_m, dx2, k.inv_dx2);
  const R fyz = mdiv<R, POW2>(Qyz - Qyz_m, dz2, k.inv_dz2);
  const R fzx = mdiv<R, POW2>(Qzx - m.Qzx_m, dx2, k.inv_dx2);
  const R fzy = mdiv<R, POW2>(Qzy - Qzy_m, dy2, k.inv_dy2);
  const R PRc = g(kPRc), PIc = g(kPIc), PRn = g(kPRn), PIn = g(kPIn);
  const R PRyp = g(kPRyp), PIyp = g(kPIyp), PRzp = g(kPRzp), PIzp = g(kPIzp);
  const R oX = g(koX), oY = g(koY), oZ = g(koZ), oPR = g(koPR), oPI = g(koPI);
  const R jx = cx * (PIc * PRn - PRc * PIn) + sx * (PRc * PRn + PIc * PIn);
  const R jy = cy * (PIc * PRyp - PRc * PIyp) + sy * (PRc * PRyp + PIc * PIyp);
  const R jz = cz * (PIc * PRzp - PRc * PIzp) + sz * (PRc * PRzp + PIc * PIzp);
  R gX, gY, gZ, f0, f1;
  bool literal = !plain;
  if (plain) {
    /* fric*((a-b)/dt) with fric = +0: 0*(a-b) has the sign and NaN-ness of
     * the quotient unless (a-b)/dt overflows */
    const R dX = TXc - oX, dY = TYc - oY, dZ = TZc - oZ, d0 = PRc - oPR, d1 = PIc - oPI;
    gX = (R)0 * dX; gY = (R)0 * dY; gZ = (R)0 * dZ;
    f0 = (R)0 * d0; f1 = (R)0 * d1;
    /* some |a-b| above k.lim (or NaN) on some lane: take the literal
     * products (never in a sane state) */
    literal = (big_lanes(dX, k.lim) | big_lanes(dY, k.lim) | big_lanes(dZ, k.lim) | big_lanes(d0, k.lim) |
               big_lanes(d1, k.lim)) != 0;
  }
  if (literal) {
    f0 = fric * ((PRc - oPR) / dt);
    f1 = fric * ((PIc - oPI) / dt);
    if (k.theta_damp_form == 0) {
      gX = damp * ((TXc - oX) / dt); gY = damp * ((TYc - oY) / dt); gZ = damp * ((TZc - oZ) / dt);
    } else {
      gX = damp * (TXc - oX) / dt; gY = damp * (TYc - oY) / dt; gZ = damp * (TZc - oZ) / dt;
    }
  }
  const R ax_ = m2gg * jx - fxy - fxz - gX;
  const R ay_ = m2gg * jy - fyx - fyz - gY;
  const R az_ = m2gg * jz - fzx - fzy - gZ;
  o.X = 2 * TXc - oX + dt2 * ax_;
  o.Y = 2 * TYc - oY + dt2 * ay_;
  o.Z = 2 * TZc - oZ + dt2 * az_;

  /* scalar field */
  const R PRym = g(kPRym), PIym = g(kPIym), PRzm = g(kPRzm), PIzm = g(kPIzm);
  const R mag2 = PRc * PRc + PIc * PIc;
  const R pot = hl * (mag2 - eta2);
  const R lxx0 = mdiv<R, POW2>(cx * PRn + sx * PIn - 2 * PRc + m.cxm * m.PRm - m.sxm * m.PIm, dx2, k.inv_dx2);
  const R lyy0 = mdiv<R, POW2>(cy * PRyp + sy * PIyp - 2 * PRc + cym * PRym - sym_ * PIym, dy2, k.inv_dy2);
  const R lzz0 = mdiv<R, POW2>(cz * PRzp + sz * PIzp - 2 * PRc + czm * PRzm - szm * PIzm, dz2, k.inv_dz2);
  const R lxx1 = mdiv<R, POW2>(cx * PIn + (-sx) * PRn - 2 * PIc + m.cxm * m.PIm - (-m.sxm) * m.PRm, dx2, k.inv_dx2);
  const R lyy1 = mdiv<R, POW2>(cy * PIyp + (-sy) * PRyp - 2 * PIc + cym * PIym - (-sym_) * PRym, dy2, k.inv_dy2);
  const R lzz1 = mdiv<R, POW2>(cz * PIzp + (-sz) * PRzp - 2 * PIc + czm * PIzm - (-szm) * PRzm, dz2, k.inv_dz2);
  const R acc0 = lxx0 + lyy0 + lzz0 - pot * PRc - f0;
  const R acc1 = lxx1 + lyy1 + lzz1 - pot * PIc - f1;
  o.PR = 2 * PRc - oPR + dt2 * acc0;
  o.PI = 2 * PIc - oPI + dt2 * acc1;

  /* carry to x+1 */
  m.PRm = PRc; m.PIm = PIc;
  m.cxm = cx; m.sxm = sx; m.Qyx_m = Qyx; m.Qzx_m = Qzx;
  m.TXm = TXc;
}

/* The prologue of a march starting at plane x0: phi, cos/sin theta_x and
 * the plaquette sines F_yx, F_zx need at s-x, from plane x0-1 and x0.  The
 * accessor G(array, plane, row, zoff) returns the current-level value at
 * (plane, y + row, z + zoff) for this lane. */
template <typename R, typename Get>
__device__ __forceinline__ void march_prologue(Get G, int xm0, int x0, MarchCarry<R> &m) {
  typedef TrigR<R> T;
  m.PRm = G(0, xm0, 0, 0);
  m.PIm = G(1, xm0, 0, 0);
  const R X = G(2, xm0, 0, 0), Y = G(3, xm0, 0, 0), Z = G(4, xm0, 0, 0);
  T::sc(X, &m.sxm, &m.cxm);
  m.TXm = X;
  const R ayx = Y + G(2, xm0, 1, 0) - G(3, x0, 0, 0) - X;
  const R azx = Z + G(2, xm0, 0, 1) - G(4, x0, 0, 0) - X;
  m.Qyx_m = T::sn(ayx);
  m.Qzx_m = T::sn(azx);
}

/* PLAIN: 0 friction products always evaluated, 1 fric = damp = +0 known at
 * compile time, 2 decided at run time from k.plain (wave-uniform).
 * BIG: a field plane is 4 GiB or more (1001^3 FP64), beyond the 32-bit
 * offsets of the buffer accesses: every access is a global one from the
 * 64-bit address of its x plane (a scalar) plus 32-bit row and lane
 * offsets (one x plane of NY*NZ values is far below 4 GiB). */
typedef __attribute__((address_space(3))) void lds_void;

/* PF (FP64, whole 64-site chunks, planes small enough that three of them
 * fit 32-bit offsets): the ten values of a plane that come from HBM -- the
 * previous level at (x, y) and the current level at (x+1, y) -- are loaded
 * one plane ahead straight into LDS (16 bytes per lane: lanes 0-31 carry
 * one field's 64 sites, lanes 32-63 another's), so their latency overlaps
 * the previous plane's arithmetic without holding VGPRs.  The pairs
 * (previous phi_r, phi_i), (previous theta_x, theta_y), (previous theta_z,
 * current theta_z at x+1), (current phi_r, phi_i at x+1), (current theta_x,
 * theta_y at x+1) each lie within 4 GiB of their lower plane, whose buffer
 * resource the pair's load uses.  Measured at 512^3: 3.747 -> 3.713 ms. */
/* MODE (FP64, periodic): 0 the step alone; 1 the step also stores the
 * divergence of its current level's links (FusedDiag::div), which the next
 * step's diagnostics need for level n-1 once that level is overwritten; 2
 * the step also takes the diagnostics of its current level n -- energy
 * density, Gauss residual (from the stored div of level n-1), finiteness,
 * forward edge phases -- with the arithmetic of the separate diagnostics
 * pass (ax_energy.h), from the link trigonometry the update computes
 * anyway.  P2D: dx, dy, dz and g are powers of two (divisions by them as
 * multiplications by the exact reciprocals). */
struct FusedDiag {
  double *div;     /* [V] divergence of the links of a level */
  float *phase;    /* [3][V] edge values: forward edge phase + link angle, FP32 (ax_wind.h) */
  double *part;    /* [3] per work item: energy, sum |G|, non-finite count */
  EnergyConsts ek;
};

__device__ __forceinline__ double march_wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <typename R, bool POW2, int PLAIN, bool BIG = false, bool PF = false, int MODE = 0, bool P2D = false>
__device__ __forceinline__ void ah_march(const MarchConsts<R> &k, const Levels<R> &L, int NX, int NY, int NZ,
                                         int y, int zc, int x0, int x1, bool zfix = false,
                                         FusedDiag fd = FusedDiag{}, int item = 0) {
  static_assert(!PF || (sizeof(R) == 8 && !BIG), "the prefetch path is FP64 with 32-bit plane offsets");
  static_assert(MODE == 0 || (sizeof(R) == 8 && !BIG), "the fused diagnostics are FP64");
  const int lane = (int)(threadIdx.x & 63);
  const int zr = zc * 64 + lane;
  const bool inz = zr < NZ;
  const int z = inz ? zr : NZ - 1;
  /* zfix: z not periodic, the z = 0 and NZ-1 sites are not stepped */
  const bool act = inz && !(zfix && (zr == 0 || zr == NZ - 1));
  const int ym = y == 0 ? NY - 1 : y - 1, yp = y == NY - 1 ? 0 : y + 1;
  const uint32_t E = (uint32_t)sizeof(R);
  const uint32_t oz = (uint32_t)z * E;
  const uint32_t ozm = (uint32_t)(z == 0 ? NZ - 1 : z - 1) * E;
  const uint32_t ozp = (uint32_t)(z == NZ - 1 ? 0 : z + 1) * E;
  const uint32_t PLB = (uint32_t)NY * (uint32_t)NZ * E;          /* plane bytes */
  const uint32_t R0 = (uint32_t)y * NZ * E, RM = (uint32_t)ym * NZ * E, RP = (uint32_t)yp * NZ * E;
  const R *PR = L.cur[0], *PI = L.cur[1], *TX = L.cur[2], *TY = L.cur[3], *TZ = L.cur[4];
  const bool plain = PLAIN == 1 || (PLAIN == 2 && k.plain);
  /* current level c[0..4] = phi_r, phi_i, theta_x, theta_y, theta_z; previous
   * level o[0..4] (read once, then overwritten with level n+1) */
  __amdgpu_buffer_rsrc_t c[5], o[5];
  for (int f = 0; f < 5; ++f) {
    c[f] = plane_rsrc(L.cur[f]);
    o[f] = plane_rsrc(L.prv[f]);
  }
  const size_t NYNZ = (size_t)NY * NZ;
  /* current level of field f in x plane x at (row byte offset, lane byte
   * offset); previous level (streaming); store of the new level */
  auto LC = [&](int f, int x, uint32_t row, uint32_t lo) -> R {
    if constexpr (BIG) return ldo(L.cur[f] + (size_t)x * NYNZ, row, lo);
    else return bld<R>(c[f], lo, (uint32_t)x * PLB + row);
  };
  auto LP = [&](int f, int x, uint32_t row, uint32_t lo) -> R {
    if constexpr (BIG) return ldo((const R *)L.prv[f] + (size_t)x * NYNZ, row, lo);
    else return bld<R, kStream>(o[f], lo, (uint32_t)x * PLB + row);
  };
  auto SP = [&](int f, int x, uint32_t row, uint32_t lo, R v) {
    if constexpr (BIG) sto(L.prv[f] + (size_t)x * NYNZ, row, lo, v);
    else bst<R, kStream>(o[f], lo, (uint32_t)x * PLB + row, v);
  };

  /* FP32: z+-1 operands come from the neighbouring lanes (DPP) instead of
   * reloading shifted rows; only the two chunk-edge lanes load: lane 0 phi
   * at the site before the chunk, lane `last` (the last active lane) phi,
   * theta_x and theta_y at the site after it (the phi loads of both in the
   * same instructions, separate ones when the chunk has a single site). */
  const int z0 = zc * 64;
  const int last = min(63, NZ - 1 - z0);
  const bool e0 = lane == 0, eL = lane == last;
  const uint32_t oe0 = (uint32_t)(z0 == 0 ? NZ - 1 : z0 - 1) * E;              /* site before the chunk */
  const uint32_t oeL = (uint32_t)(z0 + last + 1 == NZ ? 0 : z0 + last + 1) * E;  /* site after it */
  const uint32_t oe = e0 ? oe0 : oeL;
  /* a one-site chunk (NZ = 64 m + 1): lane 0 is both edges and needs both
   * sites (wave-uniform) */
  const bool single = last == 0;
  auto load_plane = [&](int x, PlaneOps<R> &d, R PRc_, R PIc_, R TXc_, R TYc_, R TZc_) {
    const int xp = x == NX - 1 ? 0 : x + 1;
    if constexpr (!PF) {
      d.PRn = LC(0, xp, R0, oz); d.PIn = LC(1, xp, R0, oz);
      d.TXn = LC(2, xp, R0, oz); d.TYn = LC(3, xp, R0, oz); d.TZn = LC(4, xp, R0, oz);
      d.oPR = LP(0, x, R0, oz); d.oPI = LP(1, x, R0, oz);
      d.oX = LP(2, x, R0, oz); d.oY = LP(3, x, R0, oz); d.oZ = LP(4, x, R0, oz);
    }
    d.PRyp = LC(0, x, RP, oz); d.PIyp = LC(1, x, RP, oz);
    d.PRym = LC(0, x, RM, oz); d.PIym = LC(1, x, RM, oz);
    d.TXyp = LC(2, x, RP, oz); d.TZyp = LC(4, x, RP, oz);
    d.TXym = LC(2, x, RM, oz); d.TYym = LC(3, x, RM, oz); d.TZym = LC(4, x, RM, oz);
    d.TYxpym = LC(3, xp, RM, oz);
    if (sizeof(R) == 8) {
      /* FP64 is issue-bound: the shifts' extra VALU costs more than the
       * reloads (L1 hits) they save; FP32 gains from the fewer requests */
      d.PRzp = LC(0, x, R0, ozp); d.PIzp = LC(1, x, R0, ozp);
      d.PRzm = LC(0, x, R0, ozm); d.PIzm = LC(1, x, R0, ozm);
      d.TXzp = LC(2, x, R0, ozp); d.TYzp = LC(3, x, R0, ozp);
      d.TYymzp = LC(3, x, RM, ozp);
      return;
    }
    R ePR = PRc_, ePI = PIc_, eTX = TXc_, eTY = TYc_, eTYym = TYc_;
    if (e0 || eL) {
      ePR = LC(0, x, R0, oe); ePI = LC(1, x, R0, oe);
    }
    if (eL) {
      eTX = LC(2, x, R0, oeL); eTY = LC(3, x, R0, oeL);
      eTYym = LC(3, x, RM, oeL);
    }
    R pPR = ePR, pPI = ePI;   /* phi after the chunk, for the edge lane */
    if (single) {
      pPR = LC(0, x, R0, oeL); pPI = LC(1, x, R0, oeL);
    }
    d.PRzm = lane_prev(PRc_, ePR); d.PIzm = lane_prev(PIc_, ePI);
    /* the shifts run with every lane active (a DPP read of an inactive lane
     * returns the old operand), then the edge lane takes its own load */
    const R nPR = lane_next(PRc_, ePR), nPI = lane_next(PIc_, ePI), nTX = lane_next(TXc_, eTX);
    const R nTY = lane_next(TYc_, eTY), nTYym = lane_next(d.TYym, eTYym);
    d.PRzp = eL ? pPR : nPR; d.PIzp = eL ? pPI : nPI;
    d.TXzp = eL ? eTX : nTX; d.TYzp = eL ? eTY : nTY;
    d.TYymzp = eL ? eTYym : nTYym;
  };

  /* ---- prologue: plane x0-1 quantities needed by plane x0 ---- */
  const int xm0 = x0 == 0 ? NX - 1 : x0 - 1;
  const R *F[5] = {PR, PI, TX, TY, TZ};
  MarchCarry<R> m;
  march_prologue<R>([&](int a, int x, int dy, int dz) {
    if constexpr (BIG) return ldo(F[a] + (size_t)x * NYNZ, dy ? RP : R0, dz ? ozp : oz);
    else return ldo(F[a], (uint32_t)x * PLB + (dy ? RP : R0), dz ? ozp : oz);
  }, xm0, x0, m);
  const R *PRx0 = PR, *PIx0 = PI, *TXx0 = TX, *TYx0 = TY, *TZx0 = TZ;
  uint32_t b0 = (uint32_t)x0 * PLB;
  if constexpr (BIG) {
    const size_t o0 = (size_t)x0 * NYNZ;
    PRx0 += o0; PIx0 += o0; TXx0 += o0; TYx0 += o0; TZx0 += o0;
    b0 = 0;
  }
  R PRc = ldo(PRx0, b0 + R0, oz), PIc = ldo(PIx0, b0 + R0, oz);
  R TXc = ldo(TXx0, b0 + R0, oz), TYc = ldo(TYx0, b0 + R0, oz), TZc = ldo(TZx0, b0 + R0, oz);

  /* Lane 0's z-backward values (ZEdge) for EB planes at a time: lane j
   * evaluates them at (xb + j, y, z0 - 1) with the operands and expressions
   * that site's own update uses, into LDS (the workgroup is this one wave).
   * One vector evaluation per EB planes instead of a sincos and two sines
   * per plane.  The pass reads its planes' lines ahead of the march; over
   * longer blocks (16 planes: +0.4 GB FP64, +1.7 GB FP32 per 512^3 step)
   * they are evicted from L2 before the march reaches them and are read
   * twice, so the blocks are short.  Measured best: 4 planes in FP32, and
   * in FP64 with the LDS prefetch (512^3: 2 planes 3.697 ms, 3: 3.671, 4:
   * 3.663, 6: 3.668, 8: 3.674, 16: 3.706); 8 in FP64 before it. */
  constexpr int EB = 4;
  __shared__ ZEdge<R> zedge[EB];
  const uint32_t em = R0 + oe0, emP = RP + oe0, ec = R0 + (uint32_t)z0 * E;
  auto edge_block = [&](int xb, int xe) {
    const int xj = min(xb + lane, xe - 1);
    const int xjp = xj == NX - 1 ? 0 : xj + 1;
    R TZe, TXe, TYe, TZxe, TZye, TX0, TY0;
    if constexpr (BIG) {
      /* per-lane planes: 64-bit element offsets */
      const size_t pj = (size_t)xj * NYNZ, pn = (size_t)xjp * NYNZ;
      const size_t iem = em / E, iemP = emP / E, iec = ec / E;
      TZe = TZ[pj + iem]; TXe = TX[pj + iem]; TYe = TY[pj + iem];
      TZxe = TZ[pn + iem]; TZye = TZ[pj + iemP];
      TX0 = TX[pj + iec]; TY0 = TY[pj + iec];
    } else {
      const uint32_t pj = (uint32_t)xj * PLB, pn = (uint32_t)xjp * PLB;
      TZe = bld<R>(c[4], pj + em, 0); TXe = bld<R>(c[2], pj + em, 0); TYe = bld<R>(c[3], pj + em, 0);
      TZxe = bld<R>(c[4], pn + em, 0); TZye = bld<R>(c[4], pj + emP, 0);
      TX0 = bld<R>(c[2], pj + ec, 0); TY0 = bld<R>(c[3], pj + ec, 0);
    }
    ZEdge<R> e;
    TrigR<R>::sc(TZe, &e.s, &e.c);
    e.qxz = TrigR<R>::sn(TXe + TZxe - TX0 - TZe);   /* a_xz at (xj, y, z0 - 1) */
    e.qyz = TrigR<R>::sn(TYe + TZye - TY0 - TZe);   /* a_yz there */
    e.tz = TZe;
    __syncthreads();   /* the previous block's values have been read */
    if (lane < xe - xb) zedge[lane] = e;
    __syncthreads();
  };

  __shared__ double pf[PF ? 2 : 1][5][128];
  /* per lane: the byte distance of its field of each pair from the pair's
   * lower plane (lanes 32-63 take the second field).  Within a level phi_r
   * lies below phi_i and theta_x below theta_y (the [comp][level] plane
   * order; the launcher checks it), so those pairs use the resources of
   * phi_r and theta_x directly -- resources chosen at run time are held in
   * scalar registers of their own, which the loop then spills */
  const bool hi = lane >= 32;
  const uint32_t lz = (uint32_t)(zc * 64) * 8u + 16u * (uint32_t)(lane & 31);
  auto pdelta = [&](const void *F, const void *G) -> uint32_t {
    const char *pF = (const char *)F, *pG = (const char *)G;
    return pF < pG ? (hi ? (uint32_t)(pG - pF) : 0u) : (hi ? 0u : (uint32_t)(pF - pG));
  };
  const bool lo4 = L.prv[4] < L.cur[4];
  const uint32_t d0 = pdelta(L.prv[0], L.prv[1]) + lz, d1 = pdelta(L.prv[2], L.prv[3]) + lz;
  const uint32_t d2 = pdelta(L.prv[4], L.cur[4]) + lz, d3 = pdelta(L.cur[0], L.cur[1]) + lz;
  const uint32_t d4 = pdelta(L.cur[2], L.cur[3]) + lz;
  /* plane x's previous level and plane x+1's current level into pf[buf] */
  auto prefetch = [&](int x, int buf) {
    const int xp = x == NX - 1 ? 0 : x + 1;
    const uint32_t so_ = (uint32_t)x * PLB + R0, sn_ = (uint32_t)xp * PLB + R0;
    const uint32_t s2 = hi ? sn_ : so_;   /* pair 2: plane x (previous level), x+1 (current) */
    __builtin_amdgcn_raw_ptr_buffer_load_lds(o[0], (lds_void *)&pf[buf][0][0], 16, d0 + so_, 0, 0, kStream);
    __builtin_amdgcn_raw_ptr_buffer_load_lds(o[2], (lds_void *)&pf[buf][1][0], 16, d1 + so_, 0, 0, kStream);
    __builtin_amdgcn_raw_ptr_buffer_load_lds(lo4 ? o[4] : c[4], (lds_void *)&pf[buf][2][0], 16, d2 + s2, 0, 0, 0);
    __builtin_amdgcn_raw_ptr_buffer_load_lds(c[0], (lds_void *)&pf[buf][3][0], 16, d3 + sn_, 0, 0, 0);
    __builtin_amdgcn_raw_ptr_buffer_load_lds(c[2], (lds_void *)&pf[buf][4][0], 16, d4 + sn_, 0, 0, 0);
  };
  if constexpr (PF) prefetch(x0, x0 & 1);
  double rho_sum = 0, g_sum = 0, bad = 0;   /* MODE 2 */
  /* DEFER (FP64): plane x's five stores are issued at the top
   * of plane x+1 (after the prefetched values, where there are any, are
   * read).  Issued at the end of plane x they sit in the memory counter
   * behind the prefetch, and the wait for the prefetched values (which the
   * compiler makes a full drain) also waits for them.  It also shortens the
   * live ranges the compiler keeps across the loop: every FP64 variant fits
   * 161 VGPRs with the prefetch, 64-bit plane addressing 170 -> 160, so all
   * keep 3 waves per SIMD.  FP32 is slower with it (+0.5 %). */
  constexpr bool DEFER = sizeof(R) == 8;
  SiteOut<R> pend{};
  auto store_out = [&](int x, const SiteOut<R> &v) {
    SP(2, x, R0, oz, v.X);
    SP(3, x, R0, oz, v.Y);
    SP(4, x, R0, oz, v.Z);
    SP(0, x, R0, oz, v.PR);
    SP(1, x, R0, oz, v.PI);
  };

  for (int x = x0; x < x1; ++x) {
    const int i = (x - x0) % EB;
    if (i == 0) edge_block(x, min(x1, x + EB));
    const ZEdge<R> ze = zedge[i];
    PlaneOps<R> D;
    if constexpr (PF) {
      const int buf = x & 1;
      D.oPR = pf[buf][0][lane]; D.oPI = pf[buf][0][64 + lane];
      D.oX = pf[buf][1][lane]; D.oY = pf[buf][1][64 + lane];
      D.oZ = pf[buf][2][lane]; D.TZn = pf[buf][2][64 + lane];
      D.PRn = pf[buf][3][lane]; D.PIn = pf[buf][3][64 + lane];
      D.TXn = pf[buf][4][lane]; D.TYn = pf[buf][4][64 + lane];
    }
    if constexpr (DEFER) {
      if (x > x0 && act) store_out(x - 1, pend);
    }
    load_plane(x, D, PRc, PIc, TXc, TYc, TZc);
    if constexpr (PF) {
      /* issued after the plane's own loads (the waits for those then leave
       * these in flight) and unconditionally (a branch around them makes
       * every later wait a full drain): the last plane fetches itself again,
       * into the buffer not being read */
      prefetch(x + 1 < x1 ? x + 1 : x, (x + 1) & 1);
    }
    SiteOut<R> out;
    if constexpr (MODE == 0) {
      ah_site_update<R, POW2>(k, plain, RegOps<R>{D, PRc, PIc, TXc, TYc, TZc}, m, ze, out);
    } else {
      const MarchCarry<R> mb = m;   /* the carry from s-x, before the update advances it */
      SiteTrig<R> tr;
      ah_site_update<R, POW2, RegOps<R>, true>(k, plain, RegOps<R>{D, PRc, PIc, TXc, TYc, TZc}, m, ze, out, &tr);
      const size_t si = ((size_t)x * NY + y) * NZ + z;
      const double TZzm = lane_prev((double)TZc, (double)ze.tz);   /* theta_z(s-z): lane - 1's own */
      const double divN = link_div<POW2>(fd.ek, TXc, mb.TXm, TYc, D.TYym, TZc, TZzm);
      if constexpr (MODE == 1) {
        if (act) fd.div[si] = divN;
      } else {
        const double divP = fd.div[si];
        double cxy, cxz, cyz;
        plaquette_cos(tr.a_xy, tr.a_xz, tr.a_yz, cxy, cxz, cyz);
        double rho, G;
        energy_gauss<P2D, false>(fd.ek, PRc, PIc, D.oPR, D.oPI, TXc, TYc, TZc, D.oX, D.oY, D.oZ, mb.PRm, mb.PIm,
                                 mb.cxm, mb.sxm, D.PRym, D.PIym, tr.cym, tr.sym, D.PRzm, D.PIzm, tr.czm, tr.szm,
                                 divN, divP, cxy, cxz, cyz, rho, G);
        const double dX = link_phase_sc(PRc, PIc, D.PRn, D.PIn, tr.sx, tr.cx);
        const double dY = link_phase_sc(PRc, PIc, D.PRyp, D.PIyp, tr.sy, tr.cy);
        const double dZ = link_phase_sc(PRc, PIc, D.PRzp, D.PIzp, tr.sz, tr.cz);
        if (act) {
          const size_t V = (size_t)NX * NY * NZ;
          fd.phase[si] = (float)(dX + TXc);
          fd.phase[V + si] = (float)(dY + TYc);
          fd.phase[2 * V + si] = (float)(dZ + TZc);
          rho_sum += rho;
          g_sum += fabs(G);
          const double vals[10] = {PRc, PIc, TXc, TYc, TZc, D.oPR, D.oPI, D.oX, D.oY, D.oZ};
#pragma unroll
          for (int q = 0; q < 10; ++q) bad += isfinite(vals[q]) ? 0.0 : 1.0;
        }
      }
    }
    if constexpr (DEFER) pend = out;
    else if (act) store_out(x, out);
    PRc = D.PRn; PIc = D.PIn; TXc = D.TXn; TYc = D.TYn; TZc = D.TZn;
  }
  if constexpr (DEFER) {
    if (x1 > x0 && act) store_out(x1 - 1, pend);
  }
  if constexpr (MODE == 2) {
    const double a = march_wave_sum(rho_sum), b = march_wave_sum(g_sum), c = march_wave_sum(bad);
    if (lane == 0) {
      fd.part[3 * (size_t)item] = a;
      fd.part[3 * (size_t)item + 1] = b;
      fd.part[3 * (size_t)item + 2] = c;
    }
  }
}

/* largest d (of type R) with d/dt finite in R */
template <typename R>
inline R rate_limit(R dt) {
  const R big = std::numeric_limits<R>::max();
  R d = big * (dt < 1 ? dt : (R)1);
  while (!(d / dt <= big)) d = std::nextafter(d, (R)0);
  for (;;) {
    R n = std::nextafter(d, std::numeric_limits<R>::infinity());
    if (!(n / dt <= big) || n > big) break;
    d = n;
  }
  return d;
}

inline bool is_pow2(double v) {
  int e;
  double m = frexp(v, &e);
  return v > 0 && m == 0.5;
}

}  // namespace ax

#endif
