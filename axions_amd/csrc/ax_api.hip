/*
 * ax_api.hip - C ABI of libaxions.so (declared in include/axions/axions.h).
 *
 * Host side of one lattice: parameter checks, device allocation, the
 * seeded initial condition (generated with libstdc++'s own std::mt19937 and
 * uniform_real_distribution, exactly the calls of mpi_evolution.cpp:290-301),
 * the TimeStep loop bookkeeping of mpi_evolution.cpp:354-371 and the
 * diagnostic read-back.  Every kernel is enqueued on the handle's stream.
 */
#ifndef _GNU_SOURCE
#define _GNU_SOURCE
#endif
#include <math.h>
#include <stdarg.h>
#include <string.h>

#include <algorithm>
#include <cstdlib>
#include <cstdio>
#include <new>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "ax_device.h"
#include "ax_energy.h"
#include "ax_internal.h"
#include "ax_march.h"

namespace ax {
thread_local std::string g_last_error;

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

int hip_fail(hipError_t e, const char *what) {
  return fail(AX_EHIP, "%s: %s", what, hipGetErrorString(e));
}

int check_params(const ax_params *p) {
  if (!p) return fail(AX_EINVAL, "params is NULL");
  if (p->model != AX_MODEL_AH && p->model != AX_MODEL_MOORE && p->model != AX_MODEL_GLOBAL)
    return fail(AX_EINVAL, "model %d", p->model);
  if (p->nx < 3 || p->ny < 3 || p->nz < 3) return fail(AX_EINVAL, "lattice %dx%dx%d: each axis needs >= 3 sites",
                                                        p->nx, p->ny, p->nz);
  if ((long long)p->nx * p->ny * p->nz > 2147483647LL / 2)
    return fail(AX_EINVAL, "lattice %dx%dx%d exceeds 2^30 sites", p->nx, p->ny, p->nz);
  if (!(p->dx > 0 && p->dy > 0 && p->dz > 0 && p->dt > 0)) return fail(AX_EINVAL, "spacings must be > 0");
  if (p->precision != 64 && p->precision != 32) return fail(AX_EINVAL, "precision %d (64 or 32)", p->precision);
  {
    const int b = p->boundary, bz = p->boundary_z;
    bool ok;
    if (p->model == AX_MODEL_GLOBAL)
      ok = (b == AX_BC_PERIODIC || b == AX_BC_NEUMANN) && bz == AX_BC_PERIODIC;
    else if (p->model == AX_MODEL_MOORE)
      ok = (b == AX_BC_PERIODIC && bz == AX_BC_PERIODIC) || ((b == AX_BC_FIXED || b == AX_BC_NEUMANN) && bz == AX_BC_FIXED);
    else
      ok = b >= AX_BC_PERIODIC && b <= AX_BC_ABSORBING && bz >= AX_BC_PERIODIC && bz <= AX_BC_NEUMANN;
    if (!ok) return fail(AX_EINVAL, "boundary (%d, z %d) for model %d", b, bz, p->model);
    if (p->model == AX_MODEL_AH && bz == AX_BC_NEUMANN && p->nz > p->nx)
      return fail(AX_EINVAL, "Neumann z faces need nz <= nx (the reference's z-face loop, evolution.cpp:237)");
  }
  if (p->theta_damp_form != 0 && p->theta_damp_form != 1) return fail(AX_EINVAL, "theta_damp_form");
  if (p->energy_form != 0 && p->energy_form != 1) return fail(AX_EINVAL, "energy_form");
  return AX_OK;
}

StepConsts step_consts(const ax_params *p, int n) { return make_step_consts(*p, n); }

DiagConsts diag_consts(const ax_params *p) {
  DiagConsts k;
  k.nx = p->nx; k.ny = p->ny; k.nz = p->nz;
  k.ex = p->boundary != AX_BC_PERIODIC;
  k.ez = p->boundary_z != AX_BC_PERIODIC;
  k.model = p->model;
  k.energy_form = p->energy_form;
  k.ncomp = p->model == AX_MODEL_AH ? 5 : 7;
  k.dx = p->dx; k.dy = p->dy; k.dz = p->dz; k.dt = p->dt; k.g = p->g;
  k.lambda = p->lambda; k.eta = p->eta;
  k.lambda1 = p->lambda1; k.lambda2 = p->lambda2; k.lambda12 = p->lambda12;
  k.eta1 = p->eta1; k.eta2 = p->eta2; k.q1 = p->q1; k.q2 = p->q2;
  k.inner = (double)((p->nx - 2) * (p->ny - 2) * (p->nz - 2));
  return k;
}

}  // namespace ax

using namespace ax;

struct ax_lattice {
  ax_params p;
  int ncomp = 0;
  size_t V = 0;
  size_t pitch = 0;         /* elements from one (component, level) plane to the next */
  size_t esize = 0;
  int timestep = 0;
  int xseg = 32;            /* x planes per wave in the tiled step (ax_set_tile) */
  int device = 0;
  hipStream_t stream = nullptr;
  void *state = nullptr;
  double *part = nullptr;   /* 3 doubles per reduction block (diag_part_blocks) */
  double *dres = nullptr;   /* small device results */
  double *hres = nullptr;   /* pinned host copy */
  int8_t *wind = nullptr;
  float *phase = nullptr;   /* FP32 edge values [3][V] of the marching diagnostics (ax_wind.h) */
  int *bcount = nullptr;
  long long *boff = nullptr;
  long long *ctotal = nullptr;
  double *cxyz = nullptr;
  size_t cxyz_cap = 0;
  int *band = nullptr;      /* Moore: per-level band flags of the FP64 march (ax_march_moore.h), TimeStep mod 3 */
  double *divb = nullptr;   /* fused diagnostics: divergence of the links of level div_for - 1 [V] */
  int div_for = -1;         /* the TimeStep whose diagnostics divb serves (-1: none) */
  double *fpart = nullptr;  /* fused diagnostics: partial sums per step work item */
  size_t fpart_n = 0;
  double *fres = nullptr;   /* fused diagnostics: 8 doubles per diagnostic point */
  size_t fres_n = 0;
  int band_steps = 0;       /* marching steps since the state or TimeStep was last set from the host */
  bool timing = false;                         /* ax_kernel_timer */
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
  int64_t timed_steps = 0;
};

namespace ax {

template <typename T>
Levels<T> levels_of(ax_lattice *L, int n) {
  Levels<T> lv;
  T *base = (T *)L->state;
  const int tN = (n + 1) % 2, tP = n % 2;
  for (int c = 0; c < 7; ++c) {
    int cc = c < L->ncomp ? c : 0;
    lv.cur[c] = base + ((size_t)cc * 2 + tN) * L->pitch;
    lv.prv[c] = base + ((size_t)cc * 2 + tP) * L->pitch;
  }
  return lv;
}
template Levels<double> levels_of<double>(ax_lattice *, int);
template Levels<float> levels_of<float>(ax_lattice *, int);

hipStream_t lattice_stream(ax_lattice *L) { return L->stream; }
const ax_params &lattice_params(const ax_lattice *L) { return L->p; }
void *lattice_state(ax_lattice *L) { return L->state; }
size_t lattice_sites(const ax_lattice *L) { return L->V; }
size_t lattice_pitch(const ax_lattice *L) { return L->pitch; }
int lattice_device(const ax_lattice *L) { return L->device; }
int lattice_timestep(const ax_lattice *L) { return L->timestep; }
void lattice_reset_timestep(ax_lattice *L) {
  L->timestep = 0;
  L->band_steps = 0;
  L->div_for = -1;
}

}  // namespace ax

extern "C" {

void ax_default_params(ax_params *p) {
  if (!p) return;
  memset(p, 0, sizeof *p);
  p->model = AX_MODEL_AH;
  p->nx = p->ny = p->nz = 64;
  p->dx = p->dy = p->dz = 0.5;
  p->dt = 0.1;
  p->lambda = 2; p->eta = 1; p->g = 1;
  p->lambda1 = 2; p->lambda2 = 2; p->lambda12 = 0; p->eta1 = 1; p->eta2 = 1;
  p->q1 = 2; p->q2 = 1;
  p->alpha = 0; p->scaling = 1; p->damped_nt = 0; p->dampFac = 0.5;
  p->theta_damp_form = 0; p->energy_form = 0;
  p->boundary = AX_BC_PERIODIC; p->precision = 64; p->boundary_z = AX_BC_PERIODIC;
}

const char *ax_strerror(int code) {
  switch (code) {
    case AX_OK: return "ok";
    case AX_EINVAL: return "invalid argument";
    case AX_ENOMEM: return "out of memory";
    case AX_EBUDGET: return "HBM budget exceeded";
    case AX_EHIP: return "HIP runtime error";
    case AX_ENONFINITE: return "non-finite value in lattice state";
    case AX_ENOTFOUND: return "unknown request";
    case AX_ENODEV: return "no GPU";
    default: return "unknown error";
  }
}

const char *ax_last_error(void) { return g_last_error.c_str(); }

int ax_state_size(const ax_params *p, int32_t *ncomp, size_t *n) {
  int rc = check_params(p);
  if (rc) return rc;
  const bool glob = p->model == AX_MODEL_GLOBAL;
  int nc = glob ? 4 : (p->model == AX_MODEL_AH ? 5 : 7);
  if (ncomp) *ncomp = nc;
  if (n) *n = (size_t)nc * (glob ? 1 : 2) * p->nx * p->ny * p->nz;
  return AX_OK;
}

int ax_create(const ax_params *p, ax_lattice **out) {
  if (!out) return fail(AX_EINVAL, "out is NULL");
  *out = nullptr;
  int rc = check_params(p);
  if (rc) return rc;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(AX_ENODEV, "no HIP device");
  ax_lattice *L = new (std::nothrow) ax_lattice;
  if (!L) return fail(AX_ENOMEM, "host allocation");
  L->p = *p;
  /* device components: the global model's phi (two buffers) and pi take
   * 3 x 2 planes */
  L->ncomp = p->model == AX_MODEL_AH ? 5 : (p->model == AX_MODEL_MOORE ? 7 : 3);
  L->V = (size_t)p->nx * p->ny * p->nz;
  L->esize = p->precision == 64 ? 8 : 4;
  L->xseg = default_xseg(*p);
  /* Field planes sit (V*esize + 4352) bytes apart instead of exactly V*esize:
   * with power-of-two lattices the plain layout puts the same site of all
   * ten planes on the same L2 set and channel. */
  {
    const char *pe = getenv("AX_PLANE_PAD_BYTES");
    const size_t pad = pe ? (size_t)atol(pe) : 4352;
    L->pitch = L->V + (pad + L->esize - 1) / L->esize;
  }
  (void)hipGetDevice(&L->device);   /* hipGetDeviceCount succeeded above */
  hipError_t e;
  if ((e = hipStreamCreateWithFlags(&L->stream, hipStreamNonBlocking)) != hipSuccess) {
    delete L;
    return hip_fail(e, "hipStreamCreate");
  }
  const size_t bytes = (size_t)L->ncomp * 2 * L->pitch * L->esize;
  const size_t nb = diag_part_blocks(*p);
  if (hipMalloc(&L->state, bytes) != hipSuccess || hipMalloc((void **)&L->part, 3 * nb * sizeof(double)) != hipSuccess ||
      hipMalloc((void **)&L->dres, 16 * sizeof(double)) != hipSuccess ||
      (p->model == AX_MODEL_MOORE && hipMalloc((void **)&L->band, 3 * sizeof(int)) != hipSuccess) ||
      hipHostMalloc((void **)&L->hres, 16 * sizeof(double), hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();   /* clear the sticky error of the failed call */
    ax_destroy(L);
    return fail(AX_ENOMEM, "device allocation of %zu bytes failed", bytes);
  }
  if ((e = hipMemsetAsync(L->state, 0, bytes, L->stream)) != hipSuccess ||
      (e = hipStreamSynchronize(L->stream)) != hipSuccess) {
    ax_destroy(L);
    return hip_fail(e, "clearing the state");
  }
  *out = L;
  return AX_OK;
}

void ax_destroy(ax_lattice *L) {
  if (!L) return;
  if (L->stream) (void)hipStreamSynchronize(L->stream);   /* drain before freeing; errors were reported earlier */
  for (auto &pr : L->ev) {
    (void)hipEventDestroy(pr.first);
    (void)hipEventDestroy(pr.second);
  }
  if (L->state) (void)hipFree(L->state);
  if (L->part) (void)hipFree(L->part);
  if (L->dres) (void)hipFree(L->dres);
  if (L->band) (void)hipFree(L->band);
  if (L->divb) (void)hipFree(L->divb);
  if (L->fpart) (void)hipFree(L->fpart);
  if (L->fres) (void)hipFree(L->fres);
  if (L->hres) (void)hipHostFree(L->hres);
  if (L->wind) (void)hipFree(L->wind);
  if (L->phase) (void)hipFree(L->phase);
  if (L->bcount) (void)hipFree(L->bcount);
  if (L->boff) (void)hipFree(L->boff);
  if (L->ctotal) (void)hipFree(L->ctotal);
  if (L->cxyz) (void)hipFree(L->cxyz);
  if (L->stream) (void)hipStreamDestroy(L->stream);
  delete L;
}

int ax_get_params(const ax_lattice *L, ax_params *out) {
  if (!L || !out) return fail(AX_EINVAL, "NULL argument");
  *out = L->p;
  return AX_OK;
}

int ax_set_timestep(ax_lattice *L, int32_t ts) {
  if (!L || ts < 0) return fail(AX_EINVAL, "bad timestep");
  L->timestep = ts;
  L->band_steps = 0;
  L->div_for = -1;
  return AX_OK;
}

int ax_get_timestep(const ax_lattice *L, int32_t *ts) {
  if (!L || !ts) return fail(AX_EINVAL, "NULL argument");
  *ts = L->timestep;
  return AX_OK;
}

int ax_sync(ax_lattice *L) {
  if (!L) return fail(AX_EINVAL, "NULL lattice");
  hipError_t e = hipStreamSynchronize(L->stream);
  return e == hipSuccess ? AX_OK : hip_fail(e, "hipStreamSynchronize");
}

/* Upload one field plane (V values) from doubles into level lev of comp. */
/* Host state plane m -> device plane.  AH/Moore: identity ([comp][level]).
 * Global model: host (Re phi, Im phi, Re pi, Im pi); phi lives in the buffer
 * of the current TimeStep's parity. */
static int host_planes(const ax_lattice *L) { return L->p.model == AX_MODEL_GLOBAL ? 4 : 2 * L->ncomp; }
static int device_plane(const ax_lattice *L, int m) {
  if (L->p.model != AX_MODEL_GLOBAL) return m;
  return m < 2 ? 2 * m + L->timestep % 2 : 4 + (m - 2);
}

static int upload_plane(ax_lattice *L, int plane, const double *src, std::vector<float> &tmp) {
  char *dst = (char *)L->state + (size_t)plane * L->pitch * L->esize;
  hipError_t e;
  if (L->esize == 8) {
    e = hipMemcpyAsync(dst, src, L->V * 8, hipMemcpyHostToDevice, L->stream);
  } else {
    tmp.resize(L->V);
    for (size_t i = 0; i < L->V; ++i) tmp[i] = (float)src[i];
    e = hipMemcpyAsync(dst, tmp.data(), L->V * 4, hipMemcpyHostToDevice, L->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(L->stream);
  }
  return e == hipSuccess ? AX_OK : hip_fail(e, "upload");
}

int ax_set_state(ax_lattice *L, const double *host, size_t n) {
  if (!L || !host) return fail(AX_EINVAL, "NULL argument");
  const size_t want = (size_t)host_planes(L) * L->V;
  if (n != want) return fail(AX_EINVAL, "state length %zu != %zu", n, want);
  /* no band flag (nor stored divergence) describes the state from here on,
   * even if an upload fails */
  L->band_steps = 0;
  L->div_for = -1;
  std::vector<float> tmp;
  for (int m = 0; m < host_planes(L); ++m) {
    int rc = upload_plane(L, device_plane(L, m), host + (size_t)m * L->V, tmp);
    if (rc) return rc;
  }
  hipError_t e = hipStreamSynchronize(L->stream);
  L->band_steps = 0;
  return e == hipSuccess ? AX_OK : hip_fail(e, "set_state");
}

int ax_get_state(ax_lattice *L, double *host, size_t n) {
  if (!L || !host) return fail(AX_EINVAL, "NULL argument");
  const size_t want = (size_t)host_planes(L) * L->V;
  if (n != want) return fail(AX_EINVAL, "state length %zu != %zu", n, want);
  hipError_t e = hipSuccess;
  std::vector<float> tmp(L->esize == 4 ? L->V : 0);
  for (int m = 0; m < host_planes(L) && e == hipSuccess; ++m) {
    const char *src = (const char *)L->state + (size_t)device_plane(L, m) * L->pitch * L->esize;
    double *dst = host + (size_t)m * L->V;
    if (L->esize == 8) {
      e = hipMemcpyAsync(dst, src, L->V * 8, hipMemcpyDeviceToHost, L->stream);
    } else {
      e = hipMemcpyAsync(tmp.data(), src, L->V * 4, hipMemcpyDeviceToHost, L->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(L->stream);
      for (size_t i = 0; i < L->V; ++i) dst[i] = tmp[i];
    }
  }
  if (e == hipSuccess) e = hipStreamSynchronize(L->stream);
  return e == hipSuccess ? AX_OK : hip_fail(e, "get_state");
}

/* Fill dst[i] = a*cos(ph[i]), dst2[i] = a*sin(ph[i]) with glibc sincos on
 * several host threads (the reference's cos/sin pair of one variable,
 * mpi_evolution.cpp:304-305, is one sincos() call after GCC's -O1+ fusion). */
static void fill_sincos(const std::vector<double> &ph, size_t stride, size_t off, double amp, std::vector<double> &re,
                        std::vector<double> &im) {
  const size_t V = re.size();
  unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; ++t)
    th.emplace_back([&, t] {
      size_t a = V * t / nt, b = V * (t + 1) / nt;
      for (size_t i = a; i < b; ++i) {
        double s, c;
        sincos(ph[i * stride + off], &s, &c);
        re[i] = amp * c;
        im[i] = amp * s;
      }
    });
  for (auto &x : th) x.join();
}

int ax_random_phases(uint32_t seed, size_t n, double *out) {
  if (!out && n) return fail(AX_EINVAL, "NULL output");
  const double pi = 4.0 * atan(1.0);
  std::mt19937 gen(seed);
  std::uniform_real_distribution<double> dist(0.0, 2 * pi);
  for (size_t i = 0; i < n; ++i) out[i] = dist(gen);
  return AX_OK;
}

int ax_init_random_phase(ax_lattice *L, uint32_t seed, int32_t mode) {
  if (!L) return fail(AX_EINVAL, "NULL lattice");
  if (mode != 0 && mode != 1) return fail(AX_EINVAL, "moore_phase_mode %d", mode);
  const double pi = 4.0 * atan(1.0);                       /* mpi_evolution.cpp:12 */
  std::mt19937 gen(seed);
  std::uniform_real_distribution<double> dist(0.0, 2 * pi);
  const bool moore = L->p.model == AX_MODEL_MOORE;
  const size_t per = (moore && mode == 0) ? 2 : 1;
  std::vector<double> ph(L->V * per);
  for (size_t i = 0; i < ph.size(); ++i) ph[i] = dist(gen);
  std::vector<double> re(L->V), im(L->V);
  std::vector<float> tmp;
  hipError_t e = hipMemsetAsync(L->state, 0, (size_t)L->ncomp * 2 * L->pitch * L->esize, L->stream);
  if (e != hipSuccess) return hip_fail(e, "hipMemset");
  const int nfield = moore ? 2 : 1;
  for (int f = 0; f < nfield; ++f) {
    double amp = !moore ? L->p.eta : (f == 0 ? L->p.eta1 : L->p.eta2);
    fill_sincos(ph, per, (per == 2) ? (size_t)f : 0, amp, re, im);
    for (int t = 0; t < 2; ++t) {
      int rc = upload_plane(L, (2 * f) * 2 + t, re.data(), tmp);
      if (!rc) rc = upload_plane(L, (2 * f + 1) * 2 + t, im.data(), tmp);
      if (rc) return rc;
    }
    e = hipStreamSynchronize(L->stream);  /* re/im are reused for the next field */
    if (e != hipSuccess) return hip_fail(e, "init");
  }
  L->timestep = 0;
  L->band_steps = 0;
  L->div_for = -1;
  return AX_OK;
}

int ax_step(ax_lattice *L, int32_t nsteps) {
  if (!L || nsteps < 0) return fail(AX_EINVAL, "bad step request");
  hipEvent_t e0 = nullptr, e1 = nullptr;
  auto drop_events = [&] {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  };
  if (L->timing && nsteps > 0) {
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
      drop_events();
      return fail(AX_EHIP, "event");
    }
    (void)hipEventRecord(e0, L->stream);   /* a failure shows in ax_kernel_time */
  }
  for (int s = 0; s < nsteps; ++s) {
    int rc = launch_step(L, L->timestep);
    if (rc) {
      /* steps before this one were enqueued and counted; no timing for the call */
      drop_events();
      return rc;
    }
    L->timestep++;
  }
  if (e0) {
    (void)hipEventRecord(e1, L->stream);
    L->ev.emplace_back(e0, e1);
    L->timed_steps += nsteps;
  }
  return AX_OK;
}

int ax_set_tile(ax_lattice *L, int32_t xseg) {
  if (!L || xseg < 0) return fail(AX_EINVAL, "bad tile request");
  L->xseg = xseg;
  return AX_OK;
}

int ax_get_tile(const ax_lattice *L, int32_t *xseg, int32_t *tiled) {
  if (!L || !xseg) return fail(AX_EINVAL, "NULL argument");
  *xseg = L->xseg;
  if (tiled) *tiled = use_tiled_ah(L->p) ? 1 : 0;
  return AX_OK;
}

int ax_kernel_timer(ax_lattice *L, int32_t enable) {
  if (!L) return fail(AX_EINVAL, "NULL lattice");
  L->timing = enable != 0;
  return AX_OK;
}

int ax_kernel_time(ax_lattice *L, double *ms, int64_t *steps) {
  if (!L || !ms || !steps) return fail(AX_EINVAL, "NULL argument");
  hipError_t e = hipStreamSynchronize(L->stream);
  if (e != hipSuccess) return hip_fail(e, "sync");
  double tot = 0;
  hipError_t te = hipSuccess;
  for (auto &pr : L->ev) {
    float t = 0;
    const hipError_t x = hipEventElapsedTime(&t, pr.first, pr.second);
    if (x != hipSuccess && te == hipSuccess) te = x;
    tot += t;
    (void)hipEventDestroy(pr.first);
    (void)hipEventDestroy(pr.second);
  }
  L->ev.clear();
  if (te != hipSuccess) {
    L->timed_steps = 0;
    return hip_fail(te, "kernel timer");
  }
  *ms = tot;
  *steps = L->timed_steps;
  L->timed_steps = 0;
  return AX_OK;
}

}  // extern "C"

/* ax_diag from the device results h (0: energy, 1: sum |G|, 2: non-finite
 * count, 4..6: wound plaquettes xy, xz, yz) of TimeStep ts */
static void fill_diag(const ax_lattice *L, const double *h, int ts, bool want_winding, ax_diag *out) {
  const ax_params &p = L->p;
  const bool glob = p.model == AX_MODEL_GLOBAL;
  memset(out, 0, sizeof *out);
  out->timestep = ts;
  out->energy = (glob && ts == 0) ? NAN : h[0];
  double dev = h[1];
  if (p.model != AX_MODEL_MOORE || p.boundary == AX_BC_PERIODIC) {
    /* evolution.cpp:1404-1408 (Moore with fixed edges divides per site) */
    dev = p.boundary == AX_BC_PERIODIC ? dev / (p.nx * p.ny) : dev / ((p.nx - 2) * (p.ny - 2));
    dev = p.boundary_z == AX_BC_PERIODIC ? dev / p.nz : dev / (p.nz - 2);
  }
  out->gauss_mean = ((p.model == AX_MODEL_AH && p.energy_form == 1) || glob) ? 0.0 : dev;
  out->finite = h[2] == 0 ? 1 : 0;
  if (want_winding) {
    const double wxy = h[4], wxz = h[5], wyz = h[6];
    out->wound_plaquettes = (int64_t)(wxy + wxz + wyz);
    out->string_length = (2.0 / 3.0) * (wxy * p.dz + wxz * p.dy + wyz * p.dx);
    out->xi = out->string_length / ((double)L->V * p.dx * p.dy * p.dz);
  }
}

extern "C" {

int ax_diagnostics(ax_lattice *L, int32_t want_winding, ax_diag *out) {
  if (!L || !out) return fail(AX_EINVAL, "NULL argument");
  const ax_params &p = L->p;
  const bool glob = p.model == AX_MODEL_GLOBAL;
  hipError_t e;
  bool march = false, mwind = false;
  if (glob) {
    /* the energy leapfrog.cpp prints for the last TimeStep (its pre-step
     * field is still in the other buffer); none before the first step */
    if (L->timestep > 0)
      e = L->esize == 8 ? launch_global_energy<double>(p, L->state, L->pitch, L->timestep - 1, L->part, L->dres, L->stream)
                        : launch_global_energy<float>(p, L->state, L->pitch, L->timestep - 1, L->part, L->dres, L->stream);
    else
      e = hipMemsetAsync(L->dres, 0, 3 * sizeof(double), L->stream);
  } else {
    DiagConsts k = diag_consts(&p);
    int rc = prepare_faces(L);
    if (rc) return rc;
    march = use_diag_march(p);
    /* the windings come from the march's edge values (periodic), else from
     * launch_winding_field */
    mwind = want_winding && march && use_diag_march_windings(p) && ensure_phase_buffer(L) == AX_OK;
    if (march)
      e = launch_march_diag(L, mwind);
    else if (L->esize == 8)
      e = launch_diag<double>(k, levels_of<double>(L, L->timestep), L->part, L->dres, L->stream);
    else
      e = launch_diag<float>(k, levels_of<float>(L, L->timestep), L->part, L->dres, L->stream);
  }
  if (e != hipSuccess) return hip_fail(e, "diag launch");
  if (want_winding) {
    /* only the count is returned: the march stores no winding field */
    if (mwind) {
      e = launch_phase_winding(L, nullptr, L->dres + 4);
      if (e != hipSuccess) return hip_fail(e, "winding launch");
    } else {
      int rc = ensure_winding_buffer(L);
      if (!rc) rc = launch_winding_field(L, 0, L->wind, L->part, L->dres + 4);
      if (rc) return rc;
    }
  }
  e = hipMemcpyAsync(L->hres, L->dres, 8 * sizeof(double), hipMemcpyDeviceToHost, L->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(L->stream);
  if (e != hipSuccess) return hip_fail(e, "diag readback");
  fill_diag(L, L->hres, L->timestep, want_winding != 0, out);
  return AX_OK;
}

}  // extern "C"

/* The fused diagnostics apply: FP64 periodic Abelian-Higgs with the
 * evolution.cpp energy form and g != 0, on the marching kernels (planes
 * below 4 GiB). */
static bool fused_ok(const ax_lattice *L) {
  const ax_params &p = L->p;
  return p.model == AX_MODEL_AH && p.precision == 64 && p.boundary == AX_BC_PERIODIC &&
         p.boundary_z == AX_BC_PERIODIC && p.energy_form == 0 && p.g != 0 && use_tiled_ah(p) && use_diag_march(p) &&
         (double)L->V * 8.0 < 4294967296.0 && getenv("AX_NO_FUSED_DIAG") == nullptr;
}

/* device buffers of the fused path for up to npoints diagnostic points: the
 * divergence plane, the partial sums, the results and the edge values (the
 * output step stores them whether or not windings are wanted) */
static int ensure_fused(ax_lattice *L, size_t npoints) {
  const size_t items = (size_t)fused_work_items(L->p.nx, L->p.ny, L->p.nz, L->xseg);
  auto grow = [&](double **buf, size_t *have, size_t want, const char *what) -> int {
    if (*have >= want) return AX_OK;
    if (*buf) (void)hipFree(*buf);
    *buf = nullptr;
    *have = 0;
    if (hipMalloc((void **)buf, want * sizeof(double)) != hipSuccess) {
      (void)hipGetLastError();
      return fail(AX_ENOMEM, "%s of %zu doubles", what, want);
    }
    *have = want;
    return AX_OK;
  };
  size_t have_div = L->divb ? L->V : 0;
  int rc = grow(&L->divb, &have_div, L->V, "divergence buffer");
  if (!rc) rc = grow(&L->fpart, &L->fpart_n, 3 * items, "partial sums");
  if (!rc) rc = grow(&L->fres, &L->fres_n, 8 * (npoints ? npoints : 1), "diagnostic results");
  if (!rc) rc = ensure_phase_buffer(L);
  return rc;
}

/* one fused step of mode 1 (store the divergence of the current level) or 2
 * (the current level's diagnostics into fres[8 k ...]) */
static int launch_fused(ax_lattice *L, int mode, size_t k, bool winding) {
  const int n = L->timestep;
  const ax_params &p = L->p;
  FusedDiag fd;
  fd.div = L->divb;
  fd.phase = L->phase;
  fd.part = L->fpart;
  fd.ek = energy_consts(p.dx, p.dy, p.dz, p.dt, p.g, p.lambda, p.eta);
  const StepConsts sk = step_consts(&p, n);
  const Levels<double> lv = levels_of<double>(L, n);
  hipError_t e = hipSuccess;
  if (mode == 2 && L->div_for != n) {
    /* no fused step stored level n-1's divergence: take it from the
     * previous level (intact until this step overwrites it) */
    const bool pow2 = is_pow2(sk.dx2) && is_pow2(sk.dy2) && is_pow2(sk.dz2);
    e = launch_level_div(p.nx, p.ny, p.nz, lv.prv[2], lv.prv[3], lv.prv[4], fd.ek, pow2, L->divb, L->stream);
  }
  if (e == hipSuccess)
    e = launch_step_fused(sk, lv, L->xseg, L->stream, mode, diag_pow2(p.dx, p.dy, p.dz, p.g), fd);
  if (e == hipSuccess && mode == 2) {
    double *res = L->fres + 8 * k;
    e = launch_sum3(L->fpart, fused_work_items(p.nx, p.ny, p.nz, L->xseg), res, L->stream);
    /* windings of level n: its links are intact until the next step */
    if (e == hipSuccess && winding) e = launch_phase_winding(L, nullptr, res + 4);
  }
  if (e != hipSuccess) return hip_fail(e, "fused step launch");
  L->div_for = mode == 1 ? n + 1 : -1;
  L->band_steps = 0;
  return AX_OK;
}

/* plain steps, bracketed by the kernel timer's events when it is on */
static int plain_steps(ax_lattice *L, int nsteps) {
  if (nsteps <= 0) return AX_OK;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (L->timing) {
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
      if (e0) (void)hipEventDestroy(e0);
      if (e1) (void)hipEventDestroy(e1);
      return fail(AX_EHIP, "event");
    }
    (void)hipEventRecord(e0, L->stream);
  }
  for (int s = 0; s < nsteps; ++s) {
    int rc = launch_step(L, L->timestep);
    if (rc) {
      if (e0) (void)hipEventDestroy(e0);
      if (e1) (void)hipEventDestroy(e1);
      return rc;
    }
    L->timestep++;
  }
  if (e0) {
    (void)hipEventRecord(e1, L->stream);
    L->ev.emplace_back(e0, e1);
    L->timed_steps += nsteps;
  }
  return AX_OK;
}

extern "C" {

int ax_step_diag(ax_lattice *L, int32_t nsteps, int32_t every, int32_t want_winding, ax_diag *out, size_t cap,
                 size_t *count) {
  if (!L || nsteps < 0 || every <= 0 || !count || (cap && !out)) return fail(AX_EINVAL, "bad step_diag request");
  *count = 0;
  std::vector<int> at;   /* the TimeSteps whose diagnostics are taken */
  for (int s = 0; s < nsteps && at.size() < cap; ++s)
    if ((L->timestep + s) % every == 0) at.push_back(L->timestep + s);
  if (!fused_ok(L)) {
    /* the separate diagnostics pass before each such step */
    size_t k = 0;
    for (int s = 0; s < nsteps; ++s) {
      if (k < at.size() && L->timestep == at[k]) {
        int rc = ax_diagnostics(L, want_winding, &out[k]);
        if (rc) return rc;
        *count = ++k;
      }
      int rc = plain_steps(L, 1);
      if (rc) return rc;
    }
    return ax_sync(L);   /* the steps after the last output too */
  }
  int rc = ensure_fused(L, at.size());
  if (rc) return rc;
  size_t k = 0;
  int run = 0;   /* plain steps not yet launched */
  for (int s = 0; s < nsteps; ++s) {
    const int t = L->timestep + run;
    const bool diag_here = k < at.size() && t == at[k];
    const bool pre = !diag_here && (t + 1) % every == 0;
    if (!diag_here && !pre) {
      ++run;
      continue;
    }
    rc = plain_steps(L, run);
    run = 0;
    if (rc) return rc;
    rc = launch_fused(L, diag_here ? 2 : 1, k, want_winding != 0);
    if (rc) return rc;
    if (diag_here) ++k;
    L->timestep++;
  }
  rc = plain_steps(L, run);
  if (rc) return rc;
  if (k) {
    std::vector<double> h(8 * k);
    hipError_t e = hipMemcpyAsync(h.data(), L->fres, 8 * k * sizeof(double), hipMemcpyDeviceToHost, L->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(L->stream);
    if (e != hipSuccess) return hip_fail(e, "diagnostic readback");
    for (size_t j = 0; j < k; ++j) fill_diag(L, &h[8 * j], at[j], want_winding != 0, &out[j]);
  } else {
    rc = ax_sync(L);
    if (rc) return rc;
  }
  *count = k;
  return AX_OK;
}

int ax_winding(ax_lattice *L, int8_t *out, size_t n) { return ax_winding_field(L, 0, out, n); }

int ax_winding_field(ax_lattice *L, int32_t field, int8_t *out, size_t n) {
  if (!L || !out) return fail(AX_EINVAL, "NULL argument");
  if (n != 3 * L->V) return fail(AX_EINVAL, "winding length %zu != %zu", n, 3 * L->V);
  if (field < 0 || field > (L->p.model == AX_MODEL_MOORE ? 1 : 0)) return fail(AX_EINVAL, "field %d", field);
  int rc = ensure_winding_buffer(L);
  if (rc) return rc;
  if (use_diag_march_windings(L->p) && ensure_phase_buffer(L) == AX_OK) {
    hipError_t e = launch_march_diag(L, true);
    if (e == hipSuccess) e = launch_phase_winding(L, L->wind, L->dres + 4);
    if (e != hipSuccess) return hip_fail(e, "winding launch");
  } else {
    rc = launch_winding_field(L, field, L->wind, L->part, L->dres + 4);
    if (rc) return rc;
  }
  hipError_t e = hipMemcpyAsync(out, L->wind, n, hipMemcpyDeviceToHost, L->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(L->stream);
  return e == hipSuccess ? AX_OK : hip_fail(e, "winding readback");
}

int ax_crossings(ax_lattice *L, double *xyz, size_t cap, size_t *count) {
  if (!L || !count || (cap && !xyz)) return fail(AX_EINVAL, "NULL argument");
  const ax_params &p = L->p;
  const size_t nb = (L->V + 255) / 256;
  hipError_t e = hipSuccess;
  if (!L->bcount) {
    if (hipMalloc((void **)&L->bcount, nb * sizeof(int)) != hipSuccess ||
        hipMalloc((void **)&L->boff, nb * sizeof(long long)) != hipSuccess ||
        hipMalloc((void **)&L->ctotal, sizeof(long long)) != hipSuccess) {
      (void)hipGetLastError();   /* clear the sticky error of the failed call */
      return fail(AX_ENOMEM, "crossing scratch");
    }
  }
  if (cap > L->cxyz_cap) {
    if (L->cxyz) (void)hipFree(L->cxyz);
    L->cxyz = nullptr;
    L->cxyz_cap = 0;
    if (hipMalloc((void **)&L->cxyz, cap * 3 * sizeof(double)) != hipSuccess) {
      (void)hipGetLastError();   /* clear the sticky error of the failed call */
      return fail(AX_ENOMEM, "crossing output of %zu points", cap);
    }
    L->cxyz_cap = cap;
  }
  const bool glob = p.model == AX_MODEL_GLOBAL;
  const int tN = glob ? L->timestep % 2 : (L->timestep + 1) % 2;
  const int ex = !glob && p.boundary != AX_BC_PERIODIC, ez = !glob && p.boundary_z != AX_BC_PERIODIC;
  int rc = prepare_faces(L);
  if (rc) return rc;
  if (L->esize == 8) {
    const double *b = (const double *)L->state;
    e = launch_crossings<double>(p.nx, p.ny, p.nz, p.dx, p.dy, p.dz, b + (size_t)tN * L->pitch, b + (size_t)(2 + tN) * L->pitch,
                                 ex, ez, L->bcount, L->boff, L->ctotal, L->cxyz, (long long)cap, L->stream, cap > 0);
  } else {
    const float *b = (const float *)L->state;
    e = launch_crossings<float>(p.nx, p.ny, p.nz, p.dx, p.dy, p.dz, b + (size_t)tN * L->pitch, b + (size_t)(2 + tN) * L->pitch,
                                ex, ez, L->bcount, L->boff, L->ctotal, L->cxyz, (long long)cap, L->stream, cap > 0);
  }
  if (e != hipSuccess) return hip_fail(e, "crossing launch");
  long long tot = 0;
  e = hipMemcpyAsync(&tot, L->ctotal, sizeof tot, hipMemcpyDeviceToHost, L->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(L->stream);
  if (e != hipSuccess) return hip_fail(e, "crossing count");
  *count = (size_t)tot;
  size_t nw = std::min(cap, (size_t)tot);
  if (nw) {
    e = hipMemcpyAsync(xyz, L->cxyz, nw * 3 * sizeof(double), hipMemcpyDeviceToHost, L->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(L->stream);
    if (e != hipSuccess) return hip_fail(e, "crossing readback");
  }
  return AX_OK;
}

}  // extern "C"

namespace ax {

/* Neumann faces of the current level, as the coming TimeStep sets them
 * before its site loop (evolution.cpp:192-265), so that energy, crossings and
 * windings see what the reference's loop sees.  Idempotent: the step's own
 * copy rewrites the same values. */
int prepare_faces(ax_lattice *L) {
  const ax_params &p = L->p;
  if (p.model == AX_MODEL_GLOBAL || (p.boundary != AX_BC_NEUMANN && p.boundary_z != AX_BC_NEUMANN)) return AX_OK;
  const int tN = (L->timestep + 1) % 2;
  hipError_t e = L->esize == 8 ? launch_neumann_faces<double>(p, L->state, L->pitch, tN, L->stream)
                               : launch_neumann_faces<float>(p, L->state, L->pitch, tN, L->stream);
  return e == hipSuccess ? AX_OK : hip_fail(e, "Neumann faces launch");
}

/* Abelian-Higgs energy and Gauss law by the diagnostics march (any boundary:
 * it sums the bulk region where x/y or z are not periodic) */
bool use_diag_march(const ax_params &p) {
  return p.model == AX_MODEL_AH && getenv("AX_DIAG_SITE_KERNEL") == nullptr;
}

/* ... and the windings from the edge values it stores (periodic lattices) */
bool use_diag_march_windings(const ax_params &p) {
  return use_diag_march(p) && p.boundary == AX_BC_PERIODIC && p.boundary_z == AX_BC_PERIODIC;
}

/* partial-sum blocks any diagnostic pass of this lattice writes */
size_t diag_part_blocks(const ax_params &p) {
  const size_t V = (size_t)p.nx * p.ny * p.nz;
  const size_t site = (V + 255) / 256;
  const size_t march = (size_t)((p.nz + 63) / 64) * p.ny * ((p.nx + kDiagXseg - 1) / kDiagXseg);
  return site > march ? site : march;
}

int ensure_phase_buffer(ax_lattice *L) {
  if (L->phase) return AX_OK;
  if (hipMalloc((void **)&L->phase, 3 * L->V * sizeof(float)) != hipSuccess) {
    (void)hipGetLastError();   /* clear the sticky error of the failed call */
    L->phase = nullptr;
    return fail(AX_ENOMEM, "edge-phase buffer of %zu bytes", 3 * L->V * sizeof(float));
  }
  return AX_OK;
}

/* energy/Gauss (and, with phases, the edge phases) of the current level */
hipError_t launch_march_diag(ax_lattice *L, bool phases) {
  const DiagConsts k = diag_consts(&L->p);
  float *ph = phases ? L->phase : nullptr;
  return L->esize == 8
             ? launch_diag_march<double>(k, levels_of<double>(L, L->timestep), kDiagXseg, ph, L->part, L->dres, L->stream)
             : launch_diag_march<float>(k, levels_of<float>(L, L->timestep), kDiagXseg, ph, L->part, L->dres, L->stream);
}

/* windings of the current level from the edge values launch_march_diag
 * stored */
hipError_t launch_phase_winding(ax_lattice *L, int8_t *out, double *res3) {
  return launch_wind_phase(L->p.nx, L->p.ny, L->p.nz, L->phase, out, L->part, res3, L->stream);
}

int ensure_winding_buffer(ax_lattice *L) {
  if (L->wind) return AX_OK;
  if (hipMalloc((void **)&L->wind, 3 * L->V) != hipSuccess) {
    (void)hipGetLastError();   /* clear the sticky error of the failed call */
    return fail(AX_ENOMEM, "winding buffer");
  }
  return AX_OK;
}

int launch_winding_field(ax_lattice *L, int field, int8_t *out, double *part, double *res3) {
  const ax_params &p = L->p;
  const bool glob = p.model == AX_MODEL_GLOBAL;
  const int tN = glob ? L->timestep % 2 : (L->timestep + 1) % 2;
  const int c0 = p.model == AX_MODEL_MOORE ? 2 * field : 0;
  const int t0 = p.model == AX_MODEL_MOORE ? 4 : 2;
  /* global model: no links; charge 0 makes every link phase exp(0) (the
   * momentum planes only stand in as finite link operands) */
  const double q = p.model == AX_MODEL_AH ? 1.0 : glob ? 0.0 : (double)(field == 0 ? p.q1 : p.q2);
  hipError_t e;
  auto P = [&](int c) { return glob && c >= 2 ? (size_t)4 * L->pitch : (size_t)(c * 2 + tN) * L->pitch; };
  const int ex = !glob && p.boundary != AX_BC_PERIODIC, ez = !glob && p.boundary_z != AX_BC_PERIODIC;
  int rc = prepare_faces(L);
  if (rc) return rc;
  /* periodic Abelian-Higgs and Moore: the marching edge-value passes
   * (ax_wind.h; the same integers as the site kernel) */
  if (!glob && !ex && !ez && getenv("AX_DIAG_SITE_KERNEL") == nullptr && ensure_phase_buffer(L) == AX_OK) {
    if (L->esize == 8) {
      const double *b = (const double *)L->state;
      e = launch_edge_winding<double>(p.nx, p.ny, p.nz, b + P(c0), b + P(c0 + 1), b + P(t0), b + P(t0 + 1),
                                      b + P(t0 + 2), q, L->phase, out, part, res3, L->stream);
    } else {
      const float *b = (const float *)L->state;
      e = launch_edge_winding<float>(p.nx, p.ny, p.nz, b + P(c0), b + P(c0 + 1), b + P(t0), b + P(t0 + 1),
                                     b + P(t0 + 2), q, L->phase, out, part, res3, L->stream);
    }
    return e == hipSuccess ? AX_OK : hip_fail(e, "winding launch");
  }
  if (L->esize == 8) {
    const double *b = (const double *)L->state;
    e = launch_winding<double>(p.nx, p.ny, p.nz, b + P(c0), b + P(c0 + 1), b + P(t0), b + P(t0 + 1), b + P(t0 + 2), q,
                               ex, ez, out, part, res3, L->stream);
  } else {
    const float *b = (const float *)L->state;
    e = launch_winding<float>(p.nx, p.ny, p.nz, b + P(c0), b + P(c0 + 1), b + P(t0), b + P(t0 + 1), b + P(t0 + 2), q,
                              ex, ez, out, part, res3, L->stream);
  }
  return e == hipSuccess ? AX_OK : hip_fail(e, "winding launch");
}

int launch_step(ax_lattice *L, int n) {
  hipError_t e;
  L->div_for = -1;
  if (L->p.model == AX_MODEL_GLOBAL) {
    e = L->esize == 8 ? launch_global_step<double>(L->p, L->state, L->pitch, n, L->stream)
                      : launch_global_step<float>(L->p, L->state, L->pitch, n, L->stream);
    return e == hipSuccess ? AX_OK : hip_fail(e, "global step launch");
  }
  StepConsts k = step_consts(&L->p, n);
  const bool tiled = use_tiled_ah(L->p), moore = L->p.model == AX_MODEL_MOORE;
  const int tN = (n + 1) % 2, tP = n % 2;
  if (L->p.boundary == AX_BC_NEUMANN || L->p.boundary_z == AX_BC_NEUMANN) {
    e = L->esize == 8 ? launch_neumann_faces<double>(L->p, L->state, L->pitch, tN, L->stream)
                      : launch_neumann_faces<float>(L->p, L->state, L->pitch, tN, L->stream);
    if (e != hipSuccess) return hip_fail(e, "Neumann faces launch");
  }
  if (L->p.boundary == AX_BC_ABSORBING) {
    e = L->esize == 8 ? launch_absorbing<double>(k, L->p, L->state, L->pitch, tN, tP, L->stream)
                      : launch_absorbing<float>(k, L->p, L->state, L->pitch, tN, tP, L->stream);
    if (e != hipSuccess) return hip_fail(e, "absorbing boundary launch");
  }
  /* the band flag of this step's current level is usable when the step
   * before it was a marching step that kept the flags */
  bool banded = false;
  const bool use_band = L->band_steps > 0;
  if (L->esize == 8) {
    Levels<double> lv = levels_of<double>(L, n);
    e = !tiled ? launch_step_site<double>(k, lv, L->p.model, L->stream)
               : moore ? launch_step_tiled_moore<double>(k, lv, L->xseg, L->stream, L->band, n, use_band, &banded)
                       : launch_step_tiled_ah<double>(k, lv, L->xseg, L->stream);
  } else {
    Levels<float> lv = levels_of<float>(L, n);
    e = !tiled ? launch_step_site<float>(k, lv, L->p.model, L->stream)
               : moore ? launch_step_tiled_moore<float>(k, lv, L->xseg, L->stream, L->band, n, use_band, &banded)
                       : launch_step_tiled_ah<float>(k, lv, L->xseg, L->stream);
  }
  /* a failed launch wrote no flag */
  L->band_steps = (banded && e == hipSuccess) ? L->band_steps + 1 : 0;
  return e == hipSuccess ? AX_OK : hip_fail(e, "step launch");
}

}  // namespace ax
