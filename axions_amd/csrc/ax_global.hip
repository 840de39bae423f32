/*
 * ax_global.hip - global (ungauged) string: the momentum-form leapfrog of
 * leapfrog.cpp:166-347 on one lattice, FP64 or FP32.
 *
 * Device layout (planes of `pitch` elements): phi_re at level 0/1, phi_im at
 * level 0/1, pi_re, pi_im.  The field is double-buffered: TimeStep n reads
 * phi from level n%2 and writes level (n+1)%2, so the reference's two
 * passes (all phitt from the pre-step field, then the site-by-site update)
 * become one kernel.  The momentum is updated in place.
 *
 *  global_bc_kernel      Neumann copies of the x and y faces of the current
 *                        field (leapfrog.cpp:168-193): y faces for interior
 *                        x, then x faces over all y, folded into one pass by
 *                        reading the corner sources directly; written to both
 *                        field buffers.
 *  global_step_kernel    phitt = phixx + phiyy + phizz - (|phi|^2 - 1) phi,
 *                        pi += phitt dt, phi_new = phi + pi dt.
 *  global_energy_kernel  the energy the reference prints for that step
 *                        (leapfrog.cpp:301-321): recomputes phitt from the
 *                        pre-step field, takes the momentum rolled back by
 *                        dt/2 and central differences of the field as a
 *                        one-thread run holds it mid-loop -- neighbours
 *                        earlier in loop order already updated, later ones
 *                        not -- accumulated in FP64 per block.
 */
#include "ax_internal.h"

namespace ax {

struct GlobalGeom {
  int nx, ny, nz;
  int periodic;          /* 1: all sites updated, periodic wrap; 0: Neumann x/y faces */
  double dx, dy, dz, dt;
};

template <typename R>
__device__ __forceinline__ R glap(const R *f, size_t s, size_t xp, size_t xm, size_t yp, size_t ym, size_t zp,
                                  size_t zm, R dx2, R dy2, R dz2, R mag) {
  R xx = (f[xp] - 2 * f[s] + f[xm]) / dx2;
  R yy = (f[yp] - 2 * f[s] + f[ym]) / dy2;
  R zz = (f[zp] - 2 * f[s] + f[zm]) / dz2;
  return xx + yy + zz - (mag - 1) * f[s];
}

/* site s = (i,j,k) is updated by the step: interior x/y for Neumann faces */
__device__ __forceinline__ bool g_site(const GlobalGeom &g, int s, int &i, int &j, int &k) {
  k = s % g.nz;
  const int r = s / g.nz;
  j = r % g.ny;
  i = r / g.ny;
  if (g.periodic) return true;
  return i > 0 && i < g.nx - 1 && j > 0 && j < g.ny - 1;
}

template <typename R>
__global__ void __launch_bounds__(256) global_bc_kernel(GlobalGeom g, R *pr, R *pi, R *qr, R *qi) {
  /* boundary sites: the two x faces (all j), then the two y faces at interior x */
  const int nz = g.nz, ny = g.ny, nx = g.nx;
  const int nxf = 2 * ny * nz, nyf = 2 * (nx - 2) * nz;
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= nxf + nyf) return;
  int i, j, k, si, sj;
  if (t < nxf) {
    const int f = t / (ny * nz), r = t % (ny * nz);
    j = r / nz;
    k = r % nz;
    i = f ? nx - 1 : 0;
    si = f ? nx - 2 : 1;
    sj = j == 0 ? 1 : (j == ny - 1 ? ny - 2 : j);   /* the y-face copy already applied at x = si */
  } else {
    const int u = t - nxf;
    const int f = u / ((nx - 2) * nz), r = u % ((nx - 2) * nz);
    i = 1 + r / nz;
    k = r % nz;
    j = f ? ny - 1 : 0;
    si = i;
    sj = f ? ny - 2 : 1;
  }
  const size_t d = ((size_t)i * ny + j) * nz + k, s = ((size_t)si * ny + sj) * nz + k;
  /* into the current field and into the buffer the step writes, whose
   * faces the step leaves alone: after the step the state holds the faces
   * the reference's array holds (copies of the pre-step interior) */
  const R vr = pr[s], vi = pi[s];
  pr[d] = vr;
  pi[d] = vi;
  qr[d] = vr;
  qi[d] = vi;
}

template <typename R>
__global__ void __launch_bounds__(256) global_step_kernel(GlobalGeom g, const R *ar, const R *ai, R *br, R *bi,
                                                          R *mr, R *mi) {
  const int V = g.nx * g.ny * g.nz;
  const int s = blockIdx.x * 256 + threadIdx.x;
  int i, j, k;
  if (s >= V || !g_site(g, s, i, j, k)) return;
  const int ny = g.ny, nz = g.nz;
  auto id = [&](int a, int b, int c) { return ((size_t)a * ny + b) * nz + c; };
  const int ip = wrapp(i, g.nx), im = wrapm(i, g.nx), jp = wrapp(j, ny), jm = wrapm(j, ny);
  const int kp = wrapp(k, nz), km = wrapm(k, nz);
  const size_t xp = id(ip, j, k), xm = id(im, j, k), yp = id(i, jp, k), ym = id(i, jm, k);
  const size_t zp = id(i, j, kp), zm = id(i, j, km);
  const R dx2 = (R)(g.dx * g.dx), dy2 = (R)(g.dy * g.dy), dz2 = (R)(g.dz * g.dz), dt = (R)g.dt;
  const R mag = ar[s] * ar[s] + ai[s] * ai[s];
  const R t0 = glap(ar, s, xp, xm, yp, ym, zp, zm, dx2, dy2, dz2, mag);
  const R t1 = glap(ai, s, xp, xm, yp, ym, zp, zm, dx2, dy2, dz2, mag);
  const R p0 = mr[s] + t0 * dt, p1 = mi[s] + t1 * dt;
  mr[s] = p0;
  mi[s] = p1;
  br[s] = ar[s] + p0 * dt;
  bi[s] = ai[s] + p1 * dt;
}

template <typename R>
__global__ void __launch_bounds__(256) global_energy_kernel(GlobalGeom g, const R *ar, const R *ai, const R *br,
                                                            const R *bi, const R *mr, const R *mi, double *part) {
  __shared__ double sh[2][4];
  const int V = g.nx * g.ny * g.nz;
  const int s = blockIdx.x * 256 + threadIdx.x;
  double rho = 0, bad = 0;
  int i, j, k;
  if (s < V && g_site(g, s, i, j, k)) {
    const int ny = g.ny, nz = g.nz;
    auto id = [&](int a, int b, int c) { return ((size_t)a * ny + b) * nz + c; };
    const int ip = wrapp(i, g.nx), im = wrapm(i, g.nx), jp = wrapp(j, ny), jm = wrapm(j, ny);
    const int kp = wrapp(k, nz), km = wrapm(k, nz);
    const size_t xp = id(ip, j, k), xm = id(im, j, k), yp = id(i, jp, k), ym = id(i, jm, k);
    const size_t zp = id(i, j, kp), zm = id(i, j, km);
    const R dx2 = (R)(g.dx * g.dx), dy2 = (R)(g.dy * g.dy), dz2 = (R)(g.dz * g.dz);
    const R dx = (R)g.dx, dy = (R)g.dy, dz = (R)g.dz;
    const R mag = ar[s] * ar[s] + ai[s] * ai[s];
    const R t[2] = {glap(ar, s, xp, xm, yp, ym, zp, zm, dx2, dy2, dz2, mag),
                    glap(ai, s, xp, xm, yp, ym, zp, zm, dx2, dy2, dz2, mag)};
    /* the value a one-thread loop sees at neighbour n: updated already
     * (new field) iff n is an updated site earlier in loop order */
    auto seen = [&](const R *o, const R *nw, size_t n) -> R {
      int a, b, c;
      const bool upd = g_site(g, (int)n, a, b, c);
      return (upd && n < (size_t)s) ? nw[n] : o[n];
    };
    double gx[2], gy[2], gz[2];
    const R *O[2] = {ar, ai}, *N[2] = {br, bi};
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      gx[c] = (double)((seen(O[c], N[c], xp) - seen(O[c], N[c], xm)) / (2 * dx));
      gy[c] = (double)((seen(O[c], N[c], yp) - seen(O[c], N[c], ym)) / (2 * dy));
      gz[c] = (double)((seen(O[c], N[c], zp) - seen(O[c], N[c], zm)) / (2 * dz));
    }
    const double r0 = (double)mr[s] - 0.5 * (double)t[0] * g.dt;
    const double r1 = (double)mi[s] - 0.5 * (double)t[1] * g.dt;
    const double dm = (double)mag - 1;
    rho = r0 * r0 + r1 * r1 + gx[0] * gx[0] + gx[1] * gx[1] + gy[0] * gy[0] + gy[1] * gy[1] + gz[0] * gz[0] +
          gz[1] * gz[1] + 0.5 * (dm * dm);
    rho = g.dx * g.dy * g.dz * rho;
    bad = (isfinite((double)br[s]) && isfinite((double)bi[s]) && isfinite((double)mr[s]) && isfinite((double)mi[s]))
              ? 0.0 : 1.0;
  }
  /* fixed-order block sums */
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    rho += __shfl_xor(rho, off, 64);
    bad += __shfl_xor(bad, off, 64);
  }
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) {
    sh[0][w] = rho;
    sh[1][w] = bad;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    part[3 * blockIdx.x] = (sh[0][0] + sh[0][1]) + (sh[0][2] + sh[0][3]);
    part[3 * blockIdx.x + 1] = 0;
    part[3 * blockIdx.x + 2] = (sh[1][0] + sh[1][1]) + (sh[1][2] + sh[1][3]);
  }
}

__global__ void __launch_bounds__(1024) global_sum3_kernel(const double *part, int nb, double *out) {
  __shared__ double sh[3][16];
  double a[3] = {0, 0, 0};
  for (int b = threadIdx.x; b < nb; b += 1024)
    for (int c = 0; c < 3; ++c) a[c] += part[3 * b + c];
  for (int c = 0; c < 3; ++c) {
    double v = a[c];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) sh[c][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    double v = 0;
    for (int w = 0; w < 16; ++w) v += sh[threadIdx.x][w];
    out[threadIdx.x] = v;
  }
}

static GlobalGeom geom_of(const ax_params &p) {
  GlobalGeom g;
  g.nx = p.nx; g.ny = p.ny; g.nz = p.nz;
  g.periodic = p.boundary == AX_BC_PERIODIC;
  g.dx = p.dx; g.dy = p.dy; g.dz = p.dz; g.dt = p.dt;
  return g;
}

template <typename R>
hipError_t launch_global_step(const ax_params &p, void *state, size_t pitch, int n, hipStream_t st) {
  R *b = (R *)state;
  const int cur = n % 2, nw = 1 - cur;
  R *ar = b + (size_t)(0 + cur) * pitch, *ai = b + (size_t)(2 + cur) * pitch;
  R *br = b + (size_t)(0 + nw) * pitch, *bi = b + (size_t)(2 + nw) * pitch;
  R *mr = b + 4 * pitch, *mi = b + 5 * pitch;
  const GlobalGeom g = geom_of(p);
  if (!g.periodic) {
    const int nbs = 2 * p.ny * p.nz + 2 * (p.nx - 2) * p.nz;
    hipLaunchKernelGGL((global_bc_kernel<R>), dim3((nbs + 255) / 256), dim3(256), 0, st, g, ar, ai, br, bi);
  }
  const int V = p.nx * p.ny * p.nz;
  hipLaunchKernelGGL((global_step_kernel<R>), dim3((V + 255) / 256), dim3(256), 0, st, g, ar, ai, br, bi, mr, mi);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_global_energy(const ax_params &p, const void *state, size_t pitch, int n, double *part,
                                double *out3, hipStream_t st) {
  const R *b = (const R *)state;
  const int cur = n % 2, nw = 1 - cur;   /* the pre-step field of TimeStep n and its result */
  const GlobalGeom g = geom_of(p);
  const int V = p.nx * p.ny * p.nz;
  const int nb = (V + 255) / 256;
  hipLaunchKernelGGL((global_energy_kernel<R>), dim3(nb), dim3(256), 0, st, g, b + (size_t)cur * pitch,
                     b + (size_t)(2 + cur) * pitch, b + (size_t)nw * pitch, b + (size_t)(2 + nw) * pitch,
                     b + 4 * pitch, b + 5 * pitch, part);
  hipLaunchKernelGGL(global_sum3_kernel, dim3(1), dim3(1024), 0, st, part, nb, out3);
  return hipGetLastError();
}

template hipError_t launch_global_step<double>(const ax_params &, void *, size_t, int, hipStream_t);
template hipError_t launch_global_step<float>(const ax_params &, void *, size_t, int, hipStream_t);
template hipError_t launch_global_energy<double>(const ax_params &, const void *, size_t, int, double *, double *,
                                                 hipStream_t);
template hipError_t launch_global_energy<float>(const ax_params &, const void *, size_t, int, double *, double *,
                                                hipStream_t);

}  // namespace ax
