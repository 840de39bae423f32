/*
 * ax_device.h - device-side layout and helpers shared by the lattice kernels.
 *
 * HBM layout of one lattice: a single allocation of ncomp * 2 * V elements
 * of the storage type T (double or float), ordered [component][level][site]
 * with site = (i*ny + j)*nz + k, z fastest -- the reference's Array order
 * (array.cpp:66-70).  Each (component, level) plane is contiguous, so a
 * wavefront reading 64 consecutive z sites of one field issues one
 * coalesced 512 B (FP64) or 256 B (FP32) request.
 */
#ifndef AXIONS_AX_DEVICE_H
#define AXIONS_AX_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "axions/ax_trig.h"
#include "axions/axions.h"

namespace ax {

/* Scalars derived once per step on the host in double precision, with the
 * reference's association (e.g. (0.5*lambda), ((-2*g)*g), (dt*dt)), so a
 * device expression using them rounds exactly like the reference's. */
struct StepConsts {
  int nx, ny, nz;
  int ex, ez;            /* 1: x,y (z) not periodic -- the bulk update skips those faces */
  double dx2, dy2, dz2;  /* dx*dx ... */
  double dt, dt2;        /* dt, dt*dt */
  double hl, eta2;       /* 0.5*lambda, eta*eta */
  double m2gg;           /* (-2*g)*g */
  double damp, fric;     /* damp, damp + alpha*scaling/time */
  int theta_damp_form;
  /* Moore */
  double hl1, hl2, hl12, eta1sq, eta2sq;
  double q1, q2;
};

/* The per-step scalars of TimeStep n.  Evaluated on the host for
 * single-lattice launches and on the device by the batched kernel; both are
 * the same IEEE operations (no contraction), hence the same bits. */
__host__ __device__ inline StepConsts make_step_consts(const ax_params &p, int n) {
  StepConsts k;
  k.nx = p.nx; k.ny = p.ny; k.nz = p.nz;
  k.ex = p.boundary != AX_BC_PERIODIC;
  k.ez = p.boundary_z != AX_BC_PERIODIC;
  k.dx2 = p.dx * p.dx; k.dy2 = p.dy * p.dy; k.dz2 = p.dz * p.dz;
  k.dt = p.dt; k.dt2 = p.dt * p.dt;
  k.hl = 0.5 * p.lambda; k.eta2 = p.eta * p.eta;
  k.m2gg = -2 * p.g * p.g;
  const double time = 1 + n * p.dt;                      /* mpi_evolution.cpp:356 */
  k.damp = (n < p.damped_nt) ? p.dampFac : 0;           /* :367-368 */
  k.fric = k.damp + p.alpha * p.scaling / time;
  k.theta_damp_form = p.theta_damp_form;
  k.hl1 = 0.5 * p.lambda1; k.hl2 = 0.5 * p.lambda2; k.hl12 = 0.5 * p.lambda12;
  k.eta1sq = p.eta1 * p.eta1; k.eta2sq = p.eta2 * p.eta2;
  k.q1 = p.q1; k.q2 = p.q2;
  return k;
}

/* Pointers to the two time levels of every component. */
template <typename T>
struct Levels {
  const T *cur[7];
  T *prv[7];
};

__device__ __forceinline__ double ld(const double *p, size_t i) { return p[i]; }
__device__ __forceinline__ double ld(const float *p, size_t i) { return (double)p[i]; }

__device__ __forceinline__ int wrapm(int i, int n) { return i == 0 ? n - 1 : i - 1; }
__device__ __forceinline__ int wrapp(int i, int n) { return i == n - 1 ? 0 : i + 1; }

}  // namespace ax

#endif
