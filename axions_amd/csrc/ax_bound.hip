/*
 * ax_bound.hip - the non-periodic boundary conditions of evolution.cpp and
 * mooreEvol.cpp, as small kernels launched around the bulk update on the
 * lattice's stream.
 *
 * Neumann (evolution.cpp:192-265, mooreEvol.cpp:114-168): at the start of a
 * TimeStep the faces of the current level tN are set from their inner
 * neighbours, the links copied and phi parallel-transported along the link:
 *   phi(face_lo) = cos(q th(face_lo)) phi(in) + c sin(q th(face_lo)) phi~(in)
 *   phi(face_hi) = cos(q th(in))      phi(in) - c sin(q th(in))      phi~(in)
 * (c = +1 for Re, -1 for Im; phi~ the other component).  Three launches in
 * the reference's loop order: y faces (interior x), x faces (every y, so the
 * x-y edges take the y faces' new values), z faces.
 *
 * Absorbing (evolution.cpp:449-689, x and y only): the faces and the four
 * z-lines of corners step with one-way wave equations instead of the field
 * equations.  Their second time derivatives read level tP at the face and
 * at its inner neighbour, so they run before the bulk update (which
 * overwrites tP in the interior), and the corners run before the faces
 * (the corner formulas read the faces' tP).  Each kernel evaluates the
 * reference's expressions in its operand order, then applies the update
 * prv = 2 cur - prv + dt^2 tt of evolution.cpp:1385-1398.
 */
#include "ax_internal.h"
#include "ax_march.h"

namespace ax {

namespace {

template <typename R>
struct FaceArgs {
  int nx, ny, nz, z1, z2;   /* z range [z1, z2) of the x/y face loops */
  int nsc;                  /* complex scalars: 1 (AH) or 2 (Moore) */
  R q[2];                   /* link charge per scalar */
  R *f[4];                  /* Re, Im of each scalar, level tN */
  R *th[3];                 /* links, level tN */
};

/* the Neumann copy of one face pair: lo <- in_lo transported by the link
 * at lo, hi <- in_hi transported by the link at in_hi; `a` = link axis */
template <typename R>
__device__ __forceinline__ void neumann_pair(const FaceArgs<R> &A, int a, size_t lo, size_t in_lo, size_t hi,
                                             size_t in_hi) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    A.th[c][lo] = A.th[c][in_lo];
    A.th[c][hi] = A.th[c][in_hi];
  }
  for (int p = 0; p < A.nsc; ++p) {
    R *re = A.f[2 * p], *im = A.f[2 * p + 1];
    R sl, cl, sh, ch;
    TrigR<R>::sc(A.q[p] * A.th[a][lo], &sl, &cl);
    TrigR<R>::sc(A.q[p] * A.th[a][in_hi], &sh, &ch);
    const R r_lo = re[in_lo], i_lo = im[in_lo], r_hi = re[in_hi], i_hi = im[in_hi];
    re[lo] = cl * r_lo + sl * i_lo;
    re[hi] = ch * r_hi - sh * i_hi;
    im[lo] = cl * i_lo + (-sl) * r_lo;
    im[hi] = ch * i_hi - (-sh) * r_hi;
  }
}

template <typename R>
__global__ void __launch_bounds__(256) neumann_y_kernel(FaceArgs<R> A) {
  const int nzr = A.z2 - A.z1;
  const int t = (int)(blockIdx.x * 256 + threadIdx.x);
  if (t >= (A.nx - 2) * nzr) return;
  const int i = 1 + t / nzr, k = A.z1 + t % nzr;
  auto id = [&](int x, int y, int z) { return ((size_t)x * A.ny + y) * A.nz + z; };
  neumann_pair(A, 1, id(i, 0, k), id(i, 1, k), id(i, A.ny - 1, k), id(i, A.ny - 2, k));
}

template <typename R>
__global__ void __launch_bounds__(256) neumann_x_kernel(FaceArgs<R> A) {
  const int nzr = A.z2 - A.z1;
  const int t = (int)(blockIdx.x * 256 + threadIdx.x);
  if (t >= A.ny * nzr) return;
  const int j = t / nzr, k = A.z1 + t % nzr;
  auto id = [&](int x, int y, int z) { return ((size_t)x * A.ny + y) * A.nz + z; };
  neumann_pair(A, 0, id(0, j, k), id(1, j, k), id(A.nx - 1, j, k), id(A.nx - 2, j, k));
}

/* z faces: the reference's loop runs its x index over [0, nz)
 * (evolution.cpp:237); nz <= nx is checked at creation */
template <typename R>
__global__ void __launch_bounds__(256) neumann_z_kernel(FaceArgs<R> A) {
  const int t = (int)(blockIdx.x * 256 + threadIdx.x);
  if (t >= A.nz * A.ny) return;
  const int i = t / A.ny, j = t % A.ny;
  auto id = [&](int x, int y, int z) { return ((size_t)x * A.ny + y) * A.nz + z; };
  neumann_pair(A, 2, id(i, j, 0), id(i, j, 1), id(i, j, A.nz - 1), id(i, j, A.nz - 2));
}

/* ------------------------------ absorbing ------------------------------- */
template <typename R>
struct AbsArgs {
  int nx, ny, nz, z1, z2, zper;
  R dx2, dy2, dz2, dt2, dtdx, dtdy;
  const R *fN[2], *tN[3];    /* level tN */
  R *fP[2], *tP[3];          /* level tP (read, then overwritten) */
};

template <typename R>
struct Abs {
  const AbsArgs<R> &A;
  __device__ Abs(const AbsArgs<R> &a) : A(a) {}
  __device__ size_t id(int x, int y, int z) const { return ((size_t)x * A.ny + y) * A.nz + z; }
  __device__ int kp(int k) const { return A.zper ? (k + 1) % A.nz : k + 1; }
  __device__ int km(int k) const { return A.zper ? (k - 1 + A.nz) % A.nz : k - 1; }

  /* covariant second difference of comp c along link axis a at s:
   * (cos th(s) f(sp) + c sin th(s) f~(sp) - 2 f(s) + cos th(sm) f(sm) - c sin th(sm) f~(sm)) / d2 */
  __device__ R lap(int c, int a, size_t s, size_t sp, size_t sm, R d2) const {
    const R *fa = A.fN[c], *fb = A.fN[1 - c];
    R s0, c0, s1, c1;
    TrigR<R>::sc(A.tN[a][s], &s0, &c0);
    TrigR<R>::sc(A.tN[a][sm], &s1, &c1);
    const R g0 = c == 0 ? s0 : -s0, g1 = c == 0 ? s1 : -s1;
    return (c0 * fa[sp] + g0 * fb[sp] - 2 * fa[s] + c1 * fa[sm] - g1 * fb[sm]) / d2;
  }
  /* one-sided time-space derivative at a low face f with inner neighbour n
   * (link at f): (cos thN(f) fN(n) + c sin thN(f) fN~(n) - fN(f)
   *               - cos thP(f) fP(n) - c sin thP(f) fP~(n) + fP(f)) / d */
  __device__ R dlo(int c, int a, size_t f, size_t n, R d) const {
    const R *aN = A.fN[c], *bN = A.fN[1 - c], *aP = A.fP[c], *bP = A.fP[1 - c];
    R sN, cN, sP, cP;
    TrigR<R>::sc(A.tN[a][f], &sN, &cN);
    TrigR<R>::sc(A.tP[a][f], &sP, &cP);
    const R gN = c == 0 ? sN : -sN, gP = c == 0 ? sP : -sP;
    return (cN * aN[n] + gN * bN[n] - aN[f] - cP * aP[n] - gP * bP[n] + aP[f]) / d;
  }
  /* at a high face f with inner neighbour n (link at n; the level-tP cosine
   * taken at cp, which is n except at one corner):
   * (fN(f) - cos thN(n) fN(n) + c sin thN(n) fN~(n)
   *  - fP(f) + cos thP(cp) fP(n) - c sin thP(n) fP~(n)) / d */
  __device__ R dhi(int c, int a, size_t f, size_t n, size_t cp, R d) const {
    const R *aN = A.fN[c], *bN = A.fN[1 - c], *aP = A.fP[c], *bP = A.fP[1 - c];
    R sN, cN, sP, cP, s2, c2;
    TrigR<R>::sc(A.tN[a][n], &sN, &cN);
    TrigR<R>::sc(A.tP[a][n], &sP, &cP);
    if (cp != n) TrigR<R>::sc(A.tP[a][cp], &s2, &c2);
    else c2 = cP;
    const R gN = c == 0 ? sN : -sN, gP = c == 0 ? sP : -sP;
    return (aN[f] - cN * aN[n] + gN * bN[n] - aP[f] + c2 * aP[n] - gP * bP[n]) / d;
  }
  __device__ R tlap(int c, size_t s, size_t sp, size_t sm, R d2) const {
    const R *t = A.tN[c];
    return (t[sp] - 2 * t[s] + t[sm]) / d2;
  }
  /* (thN(n) - thN(f) - thP(n) + thP(f)) / d */
  __device__ R tlo(int c, size_t f, size_t n, R d) const {
    return (A.tN[c][n] - A.tN[c][f] - A.tP[c][n] + A.tP[c][f]) / d;
  }
  /* (thN(f) - thN(n) - thP(f) + thP(n)) / d */
  __device__ R thi(int c, size_t f, size_t n, R d) const {
    return (A.tN[c][f] - A.tN[c][n] - A.tP[c][f] + A.tP[c][n]) / d;
  }
  __device__ void update(size_t s, const R ptt[2], const R ttt[3]) const {
#pragma unroll
    for (int c = 0; c < 2; ++c) A.fP[c][s] = 2 * A.fN[c][s] - A.fP[c][s] + A.dt2 * ptt[c];
#pragma unroll
    for (int c = 0; c < 3; ++c) A.tP[c][s] = 2 * A.tN[c][s] - A.tP[c][s] + A.dt2 * ttt[c];
  }
};

/* x faces (j in [1, ny-2]) and y faces (i in [1, nx-2]), both sides;
 * evolution.cpp:453-606 */
template <typename R>
__global__ void __launch_bounds__(256) absorbing_faces_kernel(AbsArgs<R> a) {
  const Abs<R> B(a);
  const int nzr = a.z2 - a.z1, nxf = a.ny - 2, nyf = a.nx - 2;
  const int t = (int)(blockIdx.x * 256 + threadIdx.x);
  if (t >= (nxf + nyf) * nzr) return;
  const int k = a.z1 + t % nzr, r = t / nzr;
  const int kp = B.kp(k), km = B.km(k);
  const R h = (R)0.5;
  if (r < nxf) {
    const int j = 1 + r;
    const int X[2] = {0, a.nx - 1}, In[2] = {1, a.nx - 2};
    R ptt[2][2], ttt[2][3];
    for (int side = 0; side < 2; ++side) {
      const int i = X[side];
      const size_t s = B.id(i, j, k), n = B.id(In[side], j, k);
      const size_t yp = B.id(i, j + 1, k), ym = B.id(i, j - 1, k), zp = B.id(i, j, kp), zm = B.id(i, j, km);
      for (int c = 0; c < 2; ++c) {
        const R yy = B.lap(c, 1, s, yp, ym, a.dy2), zz = B.lap(c, 2, s, zp, zm, a.dz2);
        ptt[side][c] = side == 0 ? B.dlo(c, 0, s, n, a.dtdx) + h * (yy + zz)
                                 : -B.dhi(c, 0, s, n, n, a.dtdx) + h * (yy + zz);
      }
      for (int c = 0; c < 3; ++c) {
        const R yy = B.tlap(c, s, yp, ym, a.dy2), zz = B.tlap(c, s, zp, zm, a.dz2);
        ttt[side][c] = side == 0 ? B.tlo(c, s, n, a.dtdx) + h * (yy + zz) : -B.thi(c, s, n, a.dtdx) + h * (yy + zz);
      }
    }
    B.update(B.id(0, j, k), ptt[0], ttt[0]);
    B.update(B.id(a.nx - 1, j, k), ptt[1], ttt[1]);
  } else {
    const int i = 1 + (r - nxf);
    const int Y[2] = {0, a.ny - 1}, In[2] = {1, a.ny - 2};
    R ptt[2][2], ttt[2][3];
    for (int side = 0; side < 2; ++side) {
      const int j = Y[side];
      const size_t s = B.id(i, j, k), n = B.id(i, In[side], k);
      const size_t xp = B.id(i + 1, j, k), xm = B.id(i - 1, j, k), zp = B.id(i, j, kp), zm = B.id(i, j, km);
      for (int c = 0; c < 2; ++c) {
        const R xx = B.lap(c, 0, s, xp, xm, a.dx2), zz = B.lap(c, 2, s, zp, zm, a.dz2);
        ptt[side][c] = side == 0 ? B.dlo(c, 1, s, n, a.dtdy) + h * (xx + zz)
                                 : -B.dhi(c, 1, s, n, n, a.dtdy) + h * (xx + zz);
      }
      for (int c = 0; c < 3; ++c) {
        const R xx = B.tlap(c, s, xp, xm, a.dx2), zz = B.tlap(c, s, zp, zm, a.dz2);
        ttt[side][c] = side == 0 ? B.tlo(c, s, n, a.dtdy) + h * (xx + zz) : -B.thi(c, s, n, a.dtdy) + h * (xx + zz);
      }
    }
    B.update(B.id(i, 0, k), ptt[0], ttt[0]);
    B.update(B.id(i, a.ny - 1, k), ptt[1], ttt[1]);
  }
}

/* the four corner lines, evolution.cpp:608-683:
 * tt = (zz +- 2 (tx +- ty)) / 3 with the faces' one-sided forms */
template <typename R>
__global__ void __launch_bounds__(256) absorbing_corners_kernel(AbsArgs<R> a) {
  const Abs<R> B(a);
  const int nzr = a.z2 - a.z1;
  const int t = (int)(blockIdx.x * 256 + threadIdx.x);
  if (t >= 4 * nzr) return;
  const int k = a.z1 + t % nzr, corner = t / nzr;
  const int kp = B.kp(k), km = B.km(k);
  const int hx = corner >> 1, hy = corner & 1;          /* 0: low side, 1: high side */
  const int i = hx ? a.nx - 1 : 0, j = hy ? a.ny - 1 : 0;
  const int ni = hx ? a.nx - 2 : 1, nj = hy ? a.ny - 2 : 1;
  const size_t s = B.id(i, j, k), nxs = B.id(ni, j, k), nys = B.id(i, nj, k);
  const size_t zp = B.id(i, j, kp), zm = B.id(i, j, km);
  /* the high-high corner's level-tP cosine in its x derivative is taken at
   * the corner itself (evolution.cpp:642) */
  const size_t cpx = (hx && hy) ? s : nxs;
  R ptt[2], ttt[3];
  for (int c = 0; c < 2; ++c) {
    const R zz = B.lap(c, 2, s, zp, zm, a.dz2);
    const R tx = hx ? B.dhi(c, 0, s, nxs, cpx, a.dtdx) : B.dlo(c, 0, s, nxs, a.dtdx);
    const R ty = hy ? B.dhi(c, 1, s, nys, nys, a.dtdy) : B.dlo(c, 1, s, nys, a.dtdy);
    ptt[c] = hx ? (hy ? (zz - 2 * (tx + ty)) / 3 : (zz - 2 * (tx - ty)) / 3)
                : (hy ? (zz + 2 * (tx - ty)) / 3 : (zz + 2 * (tx + ty)) / 3);
  }
  for (int c = 0; c < 3; ++c) {
    const R zz = B.tlap(c, s, zp, zm, a.dz2);
    const R tx = hx ? B.thi(c, s, nxs, a.dtdx) : B.tlo(c, s, nxs, a.dtdx);
    const R ty = hy ? B.thi(c, s, nys, a.dtdy) : B.tlo(c, s, nys, a.dtdy);
    ttt[c] = hx ? (hy ? (zz - 2 * (tx + ty)) / 3 : (zz - 2 * (tx - ty)) / 3)
                : (hy ? (zz + 2 * (tx - ty)) / 3 : (zz + 2 * (tx + ty)) / 3);
  }
  B.update(s, ptt, ttt);
}

inline unsigned blocks_for(long long n) { return (unsigned)((n + 255) / 256); }

}  // namespace

template <typename R>
hipError_t launch_neumann_faces(const ax_params &p, void *state, size_t pitch, int tN, hipStream_t st) {
  FaceArgs<R> A;
  A.nx = p.nx;
  A.ny = p.ny;
  A.nz = p.nz;
  const bool zper = p.boundary_z == AX_BC_PERIODIC;
  A.z1 = zper ? 0 : 1;
  A.z2 = zper ? p.nz : p.nz - 1;
  const bool moore = p.model == AX_MODEL_MOORE;
  A.nsc = moore ? 2 : 1;
  A.q[0] = moore ? (R)p.q1 : (R)1;
  A.q[1] = moore ? (R)p.q2 : (R)1;
  R *b = (R *)state;
  const int nphi = 2 * A.nsc, t0 = nphi;
  for (int c = 0; c < 4; ++c) A.f[c] = b + ((size_t)(c < nphi ? c : 0) * 2 + tN) * pitch;
  for (int c = 0; c < 3; ++c) A.th[c] = b + ((size_t)(t0 + c) * 2 + tN) * pitch;
  const int nzr = A.z2 - A.z1;
  if (p.boundary == AX_BC_NEUMANN) {
    hipLaunchKernelGGL(neumann_y_kernel<R>, dim3(blocks_for((long long)(p.nx - 2) * nzr)), dim3(256), 0, st, A);
    hipLaunchKernelGGL(neumann_x_kernel<R>, dim3(blocks_for((long long)p.ny * nzr)), dim3(256), 0, st, A);
  }
  if (p.boundary_z == AX_BC_NEUMANN)
    hipLaunchKernelGGL(neumann_z_kernel<R>, dim3(blocks_for((long long)p.nz * p.ny)), dim3(256), 0, st, A);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_absorbing(const StepConsts &k, const ax_params &p, void *state, size_t pitch, int tN, int tP,
                            hipStream_t st) {
  AbsArgs<R> a;
  a.nx = p.nx;
  a.ny = p.ny;
  a.nz = p.nz;
  a.zper = p.boundary_z == AX_BC_PERIODIC;
  a.z1 = a.zper ? 0 : 1;
  a.z2 = a.zper ? p.nz : p.nz - 1;
  a.dx2 = (R)k.dx2;
  a.dy2 = (R)k.dy2;
  a.dz2 = (R)k.dz2;
  a.dt2 = (R)k.dt2;
  a.dtdx = (R)(p.dt * p.dx);
  a.dtdy = (R)(p.dt * p.dy);
  R *b = (R *)state;
  for (int c = 0; c < 2; ++c) {
    a.fN[c] = b + ((size_t)c * 2 + tN) * pitch;
    a.fP[c] = b + ((size_t)c * 2 + tP) * pitch;
  }
  for (int c = 0; c < 3; ++c) {
    a.tN[c] = b + ((size_t)(2 + c) * 2 + tN) * pitch;
    a.tP[c] = b + ((size_t)(2 + c) * 2 + tP) * pitch;
  }
  const int nzr = a.z2 - a.z1;
  hipLaunchKernelGGL(absorbing_corners_kernel<R>, dim3(blocks_for(4LL * nzr)), dim3(256), 0, st, a);
  hipLaunchKernelGGL(absorbing_faces_kernel<R>, dim3(blocks_for((long long)(p.nx - 2 + p.ny - 2) * nzr)), dim3(256),
                     0, st, a);
  return hipGetLastError();
}

template hipError_t launch_neumann_faces<double>(const ax_params &, void *, size_t, int, hipStream_t);
template hipError_t launch_neumann_faces<float>(const ax_params &, void *, size_t, int, hipStream_t);
template hipError_t launch_absorbing<double>(const StepConsts &, const ax_params &, void *, size_t, int, int,
                                             hipStream_t);
template hipError_t launch_absorbing<float>(const StepConsts &, const ax_params &, void *, size_t, int, int,
                                            hipStream_t);

}  // namespace ax
