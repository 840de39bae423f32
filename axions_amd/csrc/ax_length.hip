/*
 * ax_length.hip - string length by chaining string-crossing points, the
 * algorithm of ic_w_evolution.cpp:961-1087, with a cell grid replacing its
 * all-pairs nearest-neighbour search.
 *
 * The reference follows the crossing list greedily: starting from the first
 * point, it repeatedly moves to the nearest remaining point within a cutoff,
 * adds that distance to the length and deletes the point it left; when no
 * point is within the cutoff it adds the cutoff distance if the chain has
 * come back to its starting point, deletes the point and restarts from the
 * first remaining one.  Distances use the periodic minimum image per axis;
 * the cutoff on the squared distance is (dx^2 + dy^2 + dz^2)^2, as written
 * in the reference.  Among equally near points the one latest in list order
 * wins (the reference's `<=` scan with one thread).
 *
 * Here each search visits only the 27 cells around the followed point
 * (cell edge >= the cutoff distance), so a frame of P points costs O(P)
 * instead of O(P^2); the chosen points, the distances and the order of the
 * additions are the reference's, hence the same bits.
 */
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "ax_internal.h"

namespace {

struct Chain {
  int nx, ny, nz;
  double dx, dy, dz;
  bool pxy, pz;
  const double *P;
  size_t n;
  double cut2;

  /* squared distance with the reference's per-axis minimum image
   * (ic_w_evolution.cpp:988-996) */
  double dist2(size_t a, size_t b) const {
    double d = 0;
    const double ax = P[3 * a], ay = P[3 * a + 1], az = P[3 * a + 2];
    const double bx = P[3 * b], by = P[3 * b + 1], bz = P[3 * b + 2];
    if (fabs(ax - bx) <= 0.5 * nx * dx) d = (ax - bx) * (ax - bx);
    else if (pxy) d = (nx * dx - fabs(ax - bx)) * (nx * dx - fabs(ax - bx));
    if (fabs(ay - by) <= 0.5 * ny * dy) d += (ay - by) * (ay - by);
    else if (pxy) d += (ny * dy - fabs(ay - by)) * (ny * dy - fabs(ay - by));
    if (fabs(az - bz) <= 0.5 * nz * dz) d += (az - bz) * (az - bz);
    else if (pz) d += (nz * dz - fabs(az - bz)) * (nz * dz - fabs(az - bz));
    return d;
  }
  double dist2_to(size_t a, const double c[3]) const {
    double d = 0;
    const double ax = P[3 * a], ay = P[3 * a + 1], az = P[3 * a + 2];
    if (fabs(ax - c[0]) <= 0.5 * nx * dx) d = (ax - c[0]) * (ax - c[0]);
    else if (pxy) d = (nx * dx - fabs(ax - c[0])) * (nx * dx - fabs(ax - c[0]));
    if (fabs(ay - c[1]) <= 0.5 * ny * dy) d += (ay - c[1]) * (ay - c[1]);
    else if (pxy) d += (ny * dy - fabs(ay - c[1])) * (ny * dy - fabs(ay - c[1]));
    if (fabs(az - c[2]) <= 0.5 * nz * dz) d += (az - c[2]) * (az - c[2]);
    else if (pz) d += (nz * dz - fabs(az - c[2])) * (nz * dz - fabs(az - c[2]));
    return d;
  }
  bool inner(size_t a) const {
    const double x = P[3 * a], y = P[3 * a + 1], z = P[3 * a + 2];
    return x >= -0.25 * nx * dx && x <= 0.25 * nx * dx && y >= -0.25 * ny * dy && y <= 0.25 * ny * dy &&
           z >= -0.25 * nz * dz && z <= 0.25 * nz * dz;
  }
};

/* Cell grid over one axis of length L with origin lo: cells of edge >= h. */
struct Axis {
  double lo, L;
  int nc;
  bool grid;   /* false: one cell (non-periodic axis, or box too small) */
  int cell(double x) const {
    if (!grid) return 0;
    double u = (x - lo) / L;
    u -= floor(u);
    int c = (int)(u * nc);
    return c < 0 ? 0 : (c >= nc ? nc - 1 : c);
  }
};

}  // namespace

extern "C" int ax_string_length(const double *xyz, size_t n, int32_t nx, int32_t ny, int32_t nz, double dx, double dy,
                                double dz, int32_t periodic_xy, int32_t periodic_z, double *out) {
  if (!out || (n && !xyz)) return ax::fail(AX_EINVAL, "NULL argument");
  if (nx < 1 || ny < 1 || nz < 1 || !(dx > 0 && dy > 0 && dz > 0)) return ax::fail(AX_EINVAL, "bad lattice");
  out[0] = out[1] = out[2] = 0;
  if (n == 0) return AX_OK;
  Chain C{nx, ny, nz, dx, dy, dz, periodic_xy != 0, periodic_z != 0, xyz, n, 0};
  const double c2 = dx * dx + dy * dy + dz * dz;
  C.cut2 = c2 * c2;                          /* pow(dx*dx + dy*dy + dz*dz, 2) */
  const double reach = sqrt(C.cut2) * 1.001 + 1e-12;

  /* cells: a periodic axis gets floor(L / reach) cells (>= 3 to be useful);
   * a non-periodic axis one cell (its far points count by the rule above) */
  const int x0 = (int)(0.5 * (nx - 1)), y0 = (int)(0.5 * (ny - 1)), z0 = (int)(0.5 * (nz - 1));
  Axis A[3];
  const int nn[3] = {nx, ny, nz};
  const double dd[3] = {dx, dy, dz}, org[3] = {(double)x0, (double)y0, (double)z0};
  const bool per[3] = {C.pxy, C.pxy, C.pz};
  for (int a = 0; a < 3; ++a) {
    A[a].L = nn[a] * dd[a];
    A[a].lo = -org[a] * dd[a];
    const int cnt = (int)floor(A[a].L / reach);
    A[a].grid = per[a] && cnt >= 3;
    A[a].nc = A[a].grid ? std::min(cnt, 1 << 10) : 1;
  }
  const size_t ncell = (size_t)A[0].nc * A[1].nc * A[2].nc;
  std::vector<size_t> start(ncell + 1, 0), cellOf(n), list(n);
  for (size_t i = 0; i < n; ++i) {
    const size_t c = ((size_t)A[0].cell(xyz[3 * i]) * A[1].nc + A[1].cell(xyz[3 * i + 1])) * A[2].nc +
                     A[2].cell(xyz[3 * i + 2]);
    cellOf[i] = c;
    start[c + 1]++;
  }
  for (size_t c = 0; c < ncell; ++c) start[c + 1] += start[c];
  {
    std::vector<size_t> fill(start.begin(), start.end() - 1);
    for (size_t i = 0; i < n; ++i) list[fill[cellOf[i]]++] = i;   /* ascending index within a cell */
  }
  std::vector<char> gone(n, 0);
  size_t first = 0;                          /* first remaining point in list order */
  size_t f = 0;                              /* the point being followed */
  double begin[3] = {xyz[0], xyz[1], xyz[2]};
  double total = 0, in = 0, outr = 0;

  for (size_t it = 0; it < n; ++it) {
    /* nearest remaining point within the cutoff; ties -> the latest index */
    double best = C.cut2;
    size_t bi = 0;
    bool found = false;
    int ci[3];
    const double *pf = xyz + 3 * f;
    for (int a = 0; a < 3; ++a) ci[a] = A[a].cell(pf[a]);
    int span[3], lo[3];
    for (int a = 0; a < 3; ++a) {
      span[a] = A[a].grid ? 3 : 1;
      lo[a] = A[a].grid ? ci[a] - 1 : 0;
    }
    for (int u = 0; u < span[0]; ++u)
      for (int v = 0; v < span[1]; ++v)
        for (int w = 0; w < span[2]; ++w) {
          const int cx = (lo[0] + u + A[0].nc) % A[0].nc, cy = (lo[1] + v + A[1].nc) % A[1].nc,
                    cz = (lo[2] + w + A[2].nc) % A[2].nc;
          const size_t c = ((size_t)cx * A[1].nc + cy) * A[2].nc + cz;
          for (size_t q = start[c]; q < start[c + 1]; ++q) {
            const size_t j = list[q];
            if (j == f || gone[j]) continue;
            const double d = C.dist2(f, j);
            if (d < best || (d == best && (!found || j > bi))) {
              if (d <= C.cut2) {
                best = d;
                bi = j;
                found = true;
              }
            }
          }
        }
    if (found) {
      const double seg = sqrt(best);
      total += seg;
      (C.inner(f) ? in : outr) += seg;
      gone[f] = 1;
      f = bi;
    } else {
      /* closed loop: back within the cutoff of its start, counts the cutoff */
      if (C.dist2_to(f, begin) <= C.cut2) {
        const double seg = sqrt(C.cut2);
        total += seg;
        (C.inner(f) ? in : outr) += seg;
      }
      gone[f] = 1;
      while (first < n && gone[first]) ++first;
      if (first >= n) break;
      f = first;
      begin[0] = xyz[3 * f];
      begin[1] = xyz[3 * f + 1];
      begin[2] = xyz[3 * f + 2];
    }
    while (first < n && gone[first]) ++first;
  }
  out[0] = total;
  out[1] = in;
  out[2] = outr;
  return AX_OK;
}
