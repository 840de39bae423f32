/*
 * ax_profile.hip - the straight-string initial-condition pipeline:
 *
 *  1. the radial profile f(r), A(r) of one straight string by successive
 *     over-relaxation (singleSemiAnalytic.cpp:54-134), and
 *  2. the lattice built from it by quadratic interpolation
 *     (straight_ic.cpp:59-146), written straight into a lattice's state.
 *
 * SOR.  The reference relaxes the 1-D grid r_j = j h in place, j = 1 ..
 * nx-2 in increasing order (one thread; its OpenMP loop over the same
 * in-place array races at chunk edges, so only the one-thread order is a
 * defined result): each point uses the new values at j-1 and the previous
 * sweep's values at j and j+1.  Sweep k at point j therefore depends only
 * on sweep k at j-1 and sweep k-1 at j, j+1.  Here lane L of a workgroup
 * runs sweep k0 + L, two points behind lane L-1: at time s lane L updates
 * j = s - 2L, and the two values of sweep k-1 it needs are what lane L-1
 * produced at times s-1 (point j+1) and s-2 (point j).  They move between
 * lanes of a wave by a DPP row-crossing shift and between waves through an
 * LDS ring; the whole workgroup advances in phases of PT time steps, wave w
 * working PT*w steps behind wave 0, with one barrier per phase.  Every
 * sweep sees exactly the operands it sees in the reference loop, so the
 * profile, the sweep count and the stopping residual are the reference's
 * bits.
 *
 * The loop stops after the first sweep whose largest |F| is not above tol
 * (or at max_iter).  A launch runs up to SOR_LANES sweeps; the host reads
 * their residuals, and when sweep k0 + L (L < last) is the stopping one it
 * re-runs that launch from its input with L + 1 lanes.
 */
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "ax_internal.h"

namespace ax {

constexpr int SOR_WAVES = 16;                /* four waves per SIMD hide the divide latency
                                                (30001 points: 1.37 s with 4 waves, 1.04 s with 16) */
constexpr int SOR_LANES = 64 * SOR_WAVES;    /* sweeps in flight per launch */
constexpr int SOR_PT = 64;                   /* time steps per phase */
constexpr int SOR_RING = 3 * SOR_PT;         /* slots: the phase written, and the two read */

struct SorConsts {
  int nx;
  double h, w, n, lambda, eta, g;
  double hh;          /* h*h */
  double two_hh;      /* 2/(h*h) */
  double half_lam;    /* 0.5*lambda */
  double eta2;        /* pow(eta, 2) */
};

/* One point of one sweep, singleSemiAnalytic.cpp:110-123, operands as the
 * reference reads them (pm, am: new j-1; p, a: old j; pp, ap: old j+1).
 * pow(x, 2) is x*x (GCC folds it so at -O1 and above). */
struct SorPoint {
  double phi, A, rphi, rA;
};

__host__ __device__ inline SorPoint sor_point(const SorConsts &k, int j, double pm, double p, double pp, double am,
                                              double a, double ap) {
  const double jh = j * k.h;
  const double den1 = (double)(2 * j) * k.h * k.h;
  const double q = (k.n - k.g * a) / jh;
  const double q2 = q * q;
  const double f0 = (q2 + k.half_lam * (p * p - k.eta2)) * p - (pp - pm) / den1 - (pp - 2 * p + pm) / k.hh;
  const double f1 = q2 + k.half_lam * (3 * (p * p) - k.eta2) + k.two_hh;
  const double gp = k.g * p;
  const double a0 = -2 * k.g * (k.n - k.g * a) * (p * p) + (ap - am) / den1 - (ap - 2 * a + am) / k.hh;
  const double a1 = 2 * (gp * gp) + k.two_hh;
  SorPoint r;
  r.phi = p + -k.w * f0 / f1;
  r.A = a + -k.w * a0 / a1;
  r.rphi = fabs(f0);
  r.rA = fabs(a0);
  return r;
}

/* value of lane-1 (DPP wave_shr:1); lane 0 gets `first` */
__device__ inline double from_prev_lane(double v, double first) {
  const long long bits = __double_as_longlong(v), fb = __double_as_longlong(first);
  const int lo = __builtin_amdgcn_update_dpp((int)fb, (int)bits, 0x138, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(fb >> 32), (int)(bits >> 32), 0x138, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

/* nb sweeps (nb <= SOR_LANES) from (ip, ia) -> (op, oa); res[L] = the
 * largest |F| of sweep L.  One workgroup of SOR_LANES threads. */
__global__ void __launch_bounds__(SOR_LANES) sor_kernel(SorConsts k, const double *__restrict__ ip,
                                                        const double *__restrict__ ia, double *__restrict__ op,
                                                        double *__restrict__ oa, double *__restrict__ res, int nb) {
  /* ring[0]: the input sweep, slot s holds point s+2 (a virtual lane -1);
   * ring[w+1]: lane 63 of wave w, slot s holds its output at time s */
  __shared__ double2 ring[SOR_WAVES + 1][SOR_RING];
  const int lane = (int)threadIdx.x & 63, w = (int)threadIdx.x >> 6;
  const int L = (int)threadIdx.x;
  const int nx = k.nx;
  const int last = nb - 1;
  const int s_end = nx - 2 + 2 * last;                  /* last time any sweep works */
  const int nchunk = (s_end + SOR_PT - 1) / SOR_PT;     /* chunk c: times c*PT+1 .. c*PT+PT */
  auto mod = [](int s) { return ((s % SOR_RING) + SOR_RING) % SOR_RING; };

  /* input slots for times -1 .. PT (points 1 .. PT+2) */
  if (w == 0) {
    for (int s = lane - 1; s <= SOR_PT; s += 64) {
      const int j = s + 2;
      if (j <= nx - 1) ring[0][mod(s)] = make_double2(ip[j], ia[j]);
    }
  }
  if (L == 0) {
    op[0] = 0;
    oa[0] = 0;
  }
  __syncthreads();

  double o1p = 0, o1a = 0, o2p = 0, o2a = 0;            /* own outputs at s-1, s-2 */
  double r = 0;
  const bool live = L < nb;
  for (int P = 0; P < nchunk + SOR_WAVES - 1; ++P) {
    const int c = P - w;
    if (c >= 0 && c < nchunk) {
      for (int t = 1; t <= SOR_PT; ++t) {
        const int s = c * SOR_PT + t;
        const double2 f1 = ring[w][mod(s - 1)], f2 = ring[w][mod(s - 2)];
        const double u1p = from_prev_lane(o1p, f1.x), u1a = from_prev_lane(o1a, f1.y);
        const double u2p = from_prev_lane(o2p, f2.x), u2a = from_prev_lane(o2a, f2.y);
        const int j = s - 2 * L;
        double np = o1p, na = o1a;
        if (live && j >= 1 && j <= nx - 2) {
          const bool edge = j == nx - 2;              /* phi(nx-1) = phi(nx-2) of the previous sweep */
          const SorPoint q = sor_point(k, j, j == 1 ? 0.0 : o1p, u2p, edge ? u2p : u1p, j == 1 ? 0.0 : o1a, u2a,
                                       edge ? u2a : u1a);
          np = q.phi;
          na = q.A;
          if (q.rphi > r) r = q.rphi;
          if (q.rA > r) r = q.rA;
          if (L == last) {
            op[j] = np;
            oa[j] = na;
            if (edge) {
              op[nx - 1] = u2p;
              oa[nx - 1] = u2a;
            }
          }
        }
        o2p = o1p;
        o2a = o1a;
        o1p = np;
        o1a = na;
        if (lane == 63) ring[w + 1][mod(s)] = make_double2(np, na);
      }
    }
    /* wave 0 stages the input for its next chunk (slots of chunk c-2) */
    if (w == 0 && P + 1 < nchunk) {
      const int s = (P + 1) * SOR_PT + 1 + lane;
      const int j = s + 2;
      if (j <= nx - 1) ring[0][mod(s)] = make_double2(ip[j], ia[j]);
    }
    __syncthreads();
  }
  if (live) res[L] = r;
}

/* straight_ic.cpp:63-134 at one (x, y): phi and the gauge field A */
struct StraightSite {
  double f0, f1, a0, a1;
  int offgrid;
};

__host__ __device__ inline StraightSite straight_site(const double *prof, int nprof, double dr, double nw, double g,
                                                      double x, double y) {
  StraightSite o;
  const double d = sqrt(x * x + y * y);
  const long long pc = llround(d / dr);
  double fm, am;
  o.offgrid = 0;
  if (pc == 0) {
    fm = (prof[2] * d - prof[0] * (d - dr)) / dr;
    am = (prof[3] * d - prof[1] * (d - dr)) / dr;
  } else if (pc + 1 < nprof) {
    const int p = (int)pc;
    const double e0 = d - p * dr, em = d - (p - 1) * dr, ep = d - (p + 1) * dr;
    const double den = 2 * dr * dr;
    fm = (prof[2 * (p - 1)] * e0 * ep - 2 * prof[2 * p] * em * ep + prof[2 * (p + 1)] * em * e0) / den;
    am = (prof[2 * (p - 1) + 1] * e0 * ep - 2 * prof[2 * p + 1] * em * ep + prof[2 * (p + 1) + 1] * em * e0) / den;
  } else {
    fm = 1;
    am = g == 0 ? 0 : nw / g;
    o.offgrid = 1;
  }
  const double d2 = d * d;
  if (x == 0) {
    o.f0 = 0;
    o.a1 = 0;
  } else {
    o.f0 = fm * x / d;
    o.a1 = am * x / d2;
  }
  if (y == 0) {
    o.f1 = 0;
    o.a0 = 0;
  } else {
    o.f1 = fm * y / d;
    o.a0 = -am * y / d2;
  }
  return o;
}

struct StraightArgs {
  int nx, ny, nz, nprof, global;
  double dx, dy, dz, dr, nw, g;
  size_t pitch;
};

/* one thread per site; the z column shares one (x, y) value */
template <typename T>
__global__ void __launch_bounds__(256) straight_kernel(StraightArgs a, const double *__restrict__ prof, T *state,
                                                       unsigned long long *offgrid) {
  const size_t V = (size_t)a.nx * a.ny * a.nz;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= V) return;
  const int z = (int)(i % a.nz);
  const size_t xy = i / a.nz;
  const int y = (int)(xy % a.ny), x = (int)(xy / a.ny);
  const int x0 = (a.nx - 1) / 2, y0 = (a.ny - 1) / 2;
  const StraightSite s = straight_site(prof, a.nprof, a.dr, a.nw, a.g, (x - x0) * a.dx, (y - y0) * a.dy);
  if (s.offgrid && z == 0) atomicAdd(offgrid, 1ull);
  if (a.global) {
    /* (Re phi, Im phi) in both parity buffers, momentum 0 */
    for (int b = 0; b < 2; ++b) {
      state[(size_t)(2 * 0 + b) * a.pitch + i] = (T)s.f0;
      state[(size_t)(2 * 1 + b) * a.pitch + i] = (T)s.f1;
    }
    state[(size_t)4 * a.pitch + i] = 0;
    state[(size_t)5 * a.pitch + i] = 0;
    return;
  }
  /* link variables dx*g*A (straight_ic.cpp:146); both levels */
  const T v[5] = {(T)s.f0, (T)s.f1, (T)(a.dx * a.g * s.a0), (T)(a.dy * a.g * s.a1), (T)(a.dz * a.g * 0.0)};
  for (int c = 0; c < 5; ++c)
    for (int t = 0; t < 2; ++t) state[(size_t)(2 * c + t) * a.pitch + i] = v[c];
}

}  // namespace ax

using namespace ax;

namespace {

int sor_check(const ax_sor_params *p) {
  if (!p) return fail(AX_EINVAL, "NULL parameters");
  if (p->nx < 3) return fail(AX_EINVAL, "profile needs nx >= 3 (got %d)", p->nx);
  if (p->nx > (1 << 28)) return fail(AX_EINVAL, "profile nx %d too large", p->nx);
  if (!(p->h > 0)) return fail(AX_EINVAL, "profile spacing h must be > 0");
  return AX_OK;
}

struct DevBuf {
  double *p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

}  // namespace

extern "C" {

void ax_sor_default_params(ax_sor_params *p) {
  if (!p) return;
  memset(p, 0, sizeof *p);
  /* singleSemiAnalytic.cpp:19-32 */
  p->nx = 30001;
  p->h = 0.01;
  p->w = 1.5;
  p->max_iter = 2000000;
  p->tol = 1e-3;
  p->n = 1;
  p->lambda = 1;
  p->eta = 1;
  p->g = 0;
}

int ax_sor_profile(const ax_sor_params *p, const double *init, double *out, int32_t *iterations, double *residual) {
  int rc = sor_check(p);
  if (rc) return rc;
  if (!out) return fail(AX_EINVAL, "NULL output");
  const int nx = p->nx;
  std::vector<double> hp(nx), ha(nx);
  if (init) {
    for (int i = 0; i < nx; ++i) {
      hp[i] = init[2 * i];
      ha[i] = init[2 * i + 1];
    }
  } else {
    for (int i = 0; i < nx; ++i) {          /* ic_type 0 (singleSemiAnalytic.cpp:56-61) */
      hp[i] = p->eta * tanh(0.5 * p->h * i);
      ha[i] = 0;
    }
  }
  hp[0] = 0;                                /* singleSemiAnalytic.cpp:76-78 */
  ha[0] = 0;

  int iters = 0;
  double tol_test = 1;                      /* singleSemiAnalytic.cpp:80 */
  if (tol_test > p->tol && p->max_iter > 0) {
    SorConsts k;
    k.nx = nx;
    k.h = p->h;
    k.w = p->w;
    k.n = p->n;
    k.lambda = p->lambda;
    k.eta = p->eta;
    k.g = p->g;
    k.hh = p->h * p->h;
    k.two_hh = 2 / (p->h * p->h);
    k.half_lam = 0.5 * p->lambda;
    k.eta2 = p->eta * p->eta;

    DevBuf buf, dres;
    hipError_t e = hipMalloc(&buf.p, (size_t)4 * nx * sizeof(double));
    if (e != hipSuccess) return hip_fail(e, "hipMalloc profile");
    e = hipMalloc(&dres.p, SOR_LANES * sizeof(double));
    if (e != hipSuccess) return hip_fail(e, "hipMalloc residuals");
    double *inp = buf.p, *ina = buf.p + nx, *outp = buf.p + 2 * (size_t)nx, *outa = buf.p + 3 * (size_t)nx;
    hipStream_t st;
    e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e != hipSuccess) return hip_fail(e, "hipStreamCreate");
    struct StreamGuard {
      hipStream_t s;
      ~StreamGuard() { (void)hipStreamDestroy(s); }
    } sg{st};
    e = hipMemcpyAsync(inp, hp.data(), nx * sizeof(double), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(ina, ha.data(), nx * sizeof(double), hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return hip_fail(e, "upload profile");
    std::vector<double> hres(SOR_LANES);
    while (iters < p->max_iter) {
      const int nb = std::min(SOR_LANES, p->max_iter - iters);
      hipLaunchKernelGGL(sor_kernel, dim3(1), dim3(SOR_LANES), 0, st, k, inp, ina, outp, outa, dres.p, nb);
      e = hipGetLastError();
      if (e == hipSuccess) e = hipMemcpyAsync(hres.data(), dres.p, nb * sizeof(double), hipMemcpyDeviceToHost, st);
      if (e == hipSuccess) e = hipStreamSynchronize(st);
      if (e != hipSuccess) return hip_fail(e, "SOR sweeps");
      int stop = -1;
      for (int l = 0; l < nb; ++l)
        if (!(hres[l] > p->tol)) {
          stop = l;
          break;
        }
      if (stop >= 0) {
        if (stop < nb - 1) {
          hipLaunchKernelGGL(sor_kernel, dim3(1), dim3(SOR_LANES), 0, st, k, inp, ina, outp, outa, dres.p, stop + 1);
          e = hipGetLastError();
          if (e == hipSuccess) e = hipStreamSynchronize(st);
          if (e != hipSuccess) return hip_fail(e, "SOR sweeps");
        }
        iters += stop + 1;
        tol_test = hres[stop];
        std::swap(inp, outp);
        std::swap(ina, outa);
        break;
      }
      iters += nb;
      tol_test = hres[nb - 1];
      std::swap(inp, outp);
      std::swap(ina, outa);
    }
    e = hipMemcpyAsync(hp.data(), inp, nx * sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(ha.data(), ina, nx * sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "download profile");
  }
  for (int i = 0; i < nx; ++i) {
    out[2 * i] = hp[i];
    out[2 * i + 1] = ha[i];
  }
  if (iterations) *iterations = iters;
  if (residual) *residual = tol_test;
  return AX_OK;
}

int ax_init_straight_string(ax_lattice *L, const double *profile, int32_t nprof, double dr, double n, double g,
                            int64_t *off_grid) {
  if (!L || !profile) return fail(AX_EINVAL, "NULL argument");
  if (nprof < 2) return fail(AX_EINVAL, "profile needs at least 2 points (got %d)", nprof);
  if (!(dr > 0)) return fail(AX_EINVAL, "profile spacing must be > 0");
  const ax_params &p = lattice_params(L);
  if (p.model == AX_MODEL_MOORE) return fail(AX_EINVAL, "straight-string initial condition: AH or global model only");
  StraightArgs a;
  a.nx = p.nx;
  a.ny = p.ny;
  a.nz = p.nz;
  a.nprof = nprof;
  a.global = p.model == AX_MODEL_GLOBAL;
  a.dx = p.dx;
  a.dy = p.dy;
  a.dz = p.dz;
  a.dr = dr;
  a.nw = n;
  a.g = g;
  a.pitch = lattice_pitch(L);
  const size_t V = lattice_sites(L);
  hipStream_t st = lattice_stream(L);
  DevBuf prof, cnt;
  hipError_t e = hipMalloc(&prof.p, (size_t)2 * nprof * sizeof(double) + 256);
  if (e != hipSuccess) return hip_fail(e, "hipMalloc profile");
  unsigned long long *dcount = (unsigned long long *)(prof.p + 2 * (size_t)nprof);
  unsigned long long hcount = 0;
  e = hipMemcpyAsync(prof.p, profile, (size_t)2 * nprof * sizeof(double), hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemsetAsync(dcount, 0, sizeof *dcount, st);
  if (e != hipSuccess) return hip_fail(e, "upload profile");
  const unsigned blocks = (unsigned)((V + 255) / 256);
  if (p.precision == 64)
    hipLaunchKernelGGL(straight_kernel<double>, dim3(blocks), dim3(256), 0, st, a, prof.p,
                       (double *)lattice_state(L), dcount);
  else
    hipLaunchKernelGGL(straight_kernel<float>, dim3(blocks), dim3(256), 0, st, a, prof.p, (float *)lattice_state(L),
                       dcount);
  e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(&hcount, dcount, sizeof hcount, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return hip_fail(e, "straight-string initial condition");
  lattice_reset_timestep(L);
  if (off_grid) *off_grid = (int64_t)(hcount * (unsigned long long)p.nz);
  return AX_OK;
}

}  // extern "C"
