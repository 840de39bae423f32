/*
 * axions.h - C ABI of the MI355X lattice-evolution library (libaxions.so).
 *
 * The reference (sjcott/Axions) has no library interface: every program is
 * one main() whose parameters are file-scope consts and whose hot path is
 * the body of its TimeStep loop.  Each entry point below names the region of
 * the reference it replaces, so a host program written like the reference's
 * main() can call this library instead of running its own site loops.
 *
 * Conventions
 *  - Plain C types only; no exceptions cross the boundary.  Every function
 *    returns AX_OK (0) or a negative AX_E* code; ax_strerror() names it and
 *    ax_last_error() gives a one-line detail for the calling thread.
 *  - Host arrays use the reference's Array layout (array.cpp:66-70):
 *    [component][level][x][y][z], z fastest.  AH components are
 *    (Re phi, Im phi, theta_x, theta_y, theta_z) as in Data/ic.txt
 *    (evolution.cpp:108); Moore components are (Re phi1, Im phi1, Re phi2,
 *    Im phi2, theta_x, theta_y, theta_z) as in Data/moore_ic.txt
 *    (mooreEvol.cpp:77).  theta_a = g*A_a*dx_a is the link variable.
 *  - Time levels follow the reference: during TimeStep n the current level
 *    is tNow = (n+1)%2 and the update overwrites tPast = n%2
 *    (mpi_evolution.cpp:370-371).
 *  - A handle is used by one host thread at a time; distinct handles are
 *    independent.  ax_step() only enqueues work on the handle's HIP stream.
 */
#ifndef AXIONS_AXIONS_H
#define AXIONS_AXIONS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AX_OK 0
#define AX_EINVAL (-1)     /* bad argument or parameter combination */
#define AX_ENOMEM (-2)     /* device or host allocation failed */
#define AX_EBUDGET (-3)    /* admission refused: HBM budget exceeded */
#define AX_EHIP (-4)       /* a HIP runtime call failed */
#define AX_ENONFINITE (-5) /* a NaN/Inf was found in the lattice state */
#define AX_ENOTFOUND (-6)  /* unknown request id */
#define AX_ENODEV (-7)     /* no usable GPU */

#define AX_MODEL_AH 0      /* Abelian-Higgs: one complex scalar + U(1) links */
#define AX_MODEL_MOORE 1   /* two complex scalars, charges q1,q2, one U(1) */
#define AX_MODEL_GLOBAL 2  /* one complex scalar, no gauge field, momentum (pi)
                              form of the leapfrog (leapfrog.cpp:200-347) */

/* Boundary conditions, per axis group as the reference's xyBC and zBC
 * (evolution.cpp:52-53): ax_params.boundary for x and y, boundary_z for z.
 * Only periodic axes wrap; on the others the faces are not stepped by the
 * bulk update, and the energy, Gauss sum and crossings cover the sites the
 * bulk update covers (evolution.cpp:161-186).
 *   AX_BC_PERIODIC   periodic (mpi_evolution.cpp:385-405)
 *   AX_BC_FIXED      face sites never updated (evolution.cpp "fixed",
 *                    mooreEvol.cpp:187-189)
 *   AX_BC_NEUMANN    before each step the faces of the current level are
 *                    set from their inner neighbours, parallel-transported
 *                    along the link: y faces, then x faces (all y), then z
 *                    faces (evolution.cpp:192-265; Moore with q-scaled links,
 *                    mooreEvol.cpp:114-168).  Global model: plain copies of
 *                    phi and pi, z periodic (leapfrog.cpp:168-193).
 *   AX_BC_ABSORBING  x and y only (AH): the faces and corners step with the
 *                    one-way wave equations of evolution.cpp:449-689.
 * Supported pairs (boundary, boundary_z): AH any x/y mode with z periodic,
 * fixed or Neumann (z Neumann needs nz <= nx: the reference's z-face loop
 * runs its x index up to nz, evolution.cpp:237); Moore (periodic, periodic),
 * (fixed, fixed), (Neumann, fixed); global (periodic, periodic),
 * (Neumann, periodic). */
#define AX_BC_PERIODIC 0
#define AX_BC_FIXED 1
#define AX_BC_NEUMANN 2
#define AX_BC_NEUMANN_XY AX_BC_NEUMANN
#define AX_BC_ABSORBING 3

/* Run parameters.  Names and meanings follow the reference's file-scope
 * consts (evolution.cpp:23-59, mpi_evolution.cpp:17-60, mooreEvol.cpp:19-42).
 * ax_default_params() fills the plain Abelian-Higgs set (dx = 0.5, dt = 0.1,
 * lambda = 2, g = eta = 1, no friction, periodic). */
typedef struct ax_params {
  int32_t model;            /* AX_MODEL_AH or AX_MODEL_MOORE */
  int32_t nx, ny, nz;       /* lattice sites per axis, each >= 3 */
  double dx, dy, dz, dt;
  /* Abelian-Higgs: V = lambda/4 (|phi|^2 - eta^2)^2, gauge coupling g */
  double lambda, eta, g;
  /* Moore (mooreEvol.cpp:27-35); uses g above as the gauge coupling */
  double lambda1, lambda2, lambda12, eta1, eta2;
  int32_t q1, q2;
  /* friction (AH only): PRS term alpha*scaling/time with time = 1 + n*dt,
   * plus damping dampFac for TimeStep < damped_nt (mpi_evolution.cpp:31-38,
   * 356, 367-368).  mooreEvol.cpp has no friction; it is ignored there. */
  double alpha, scaling;
  int32_t damped_nt;
  double dampFac;
  /* 0: theta friction written damp*((theta-thetaPast)/dt), the fused form of
   *    mpi_evolution.cpp:490-497;  1: (damp*(theta-thetaPast))/dt, the
   *    two-pass form of evolution.cpp:371-373.  Identical when damp == 0. */
  int32_t theta_damp_form;
  /* 0: energy as evolution.cpp:403-427 (E_i = dtheta/(g dt dx_i), zero gauge
   *    terms when g == 0);  1: as mpi_evolution.cpp:513-530. */
  int32_t energy_form;
  int32_t boundary;         /* x and y: AX_BC_* (xyBC) */
  int32_t precision;        /* 64: FP64 drop-in; 32: FP32 storage+arithmetic */
  int32_t boundary_z;       /* z: AX_BC_PERIODIC, AX_BC_FIXED or AX_BC_NEUMANN (zBC) */
} ax_params;

/* Diagnostics of the current state (level tNow of the next TimeStep),
 * evaluated as the reference does inside its site loop before the update. */
typedef struct ax_diag {
  int32_t timestep;         /* n: number of steps taken so far */
  int32_t finite;           /* 1 if every stored value is finite */
  double energy;            /* sum dx dy dz rho (evolution.cpp:425) */
  double gauss_mean;        /* sum |G| / (nx ny nz) (evolution.cpp:1404-1408) */
  int64_t wound_plaquettes; /* sum over plaquettes of |winding| */
  double string_length;     /* (2/3) * dx * wound_plaquettes */
  double xi;                /* string length density L / V */
} ax_diag;

typedef struct ax_lattice ax_lattice;

void ax_default_params(ax_params *p);
const char *ax_strerror(int code);
const char *ax_last_error(void);
/* Number of stored components (5 AH, 7 Moore) and the host array length
 * ncomp * 2 * nx * ny * nz that ax_set_state/ax_get_state exchange.  The
 * global model stores its field and momentum instead of two levels: 4
 * components (Re phi, Im phi, Re pi, Im pi), one level each, n = 4 nx ny nz
 * (phi and phit of leapfrog.cpp:83). */
int ax_state_size(const ax_params *p, int32_t *ncomp, size_t *n);

/* Allocate device state (zeroed) on the current HIP device.
 * Replaces the std::vector/Array allocations of main()
 * (mpi_evolution.cpp:94, evolution.cpp:63). */
int ax_create(const ax_params *p, ax_lattice **out);
void ax_destroy(ax_lattice *lat);
int ax_get_params(const ax_lattice *lat, ax_params *out);

/* Seeded random-phase initial condition with TimeStep reset to 0:
 * std::mt19937(seed), uniform_real_distribution<double>(0, 2pi), one draw
 * per site in x-major order, phi = eta*(cos, sin) at both levels, links 0
 * (mpi_evolution.cpp:287-312).  Moore: moore_phase_mode 0 draws an
 * independent phase for phi1 then phi2 at each site; 1 uses one phase for
 * both (moore_mpi_evolution.cpp:304-313). */
int ax_init_random_phase(ax_lattice *lat, uint32_t seed, int32_t moore_phase_mode);
/* The first n phases of that stream (host only, no GPU needed). */
int ax_random_phases(uint32_t seed, size_t n, double *out);

/* Copy the full two-level state in (out of) the reference layout; values
 * are converted to (from) the storage precision.  Replaces the ic.txt
 * reader (evolution.cpp:102-136) and finalField writer
 * (evolution.cpp:1447-1459).  n must equal ax_state_size(). */
int ax_set_state(ax_lattice *lat, const double *host, size_t n);
int ax_get_state(ax_lattice *lat, double *host, size_t n);
int ax_set_timestep(ax_lattice *lat, int32_t timestep);
int ax_get_timestep(const ax_lattice *lat, int32_t *timestep);

/* Advance nsteps TimeSteps: the fused in-place update of
 * mpi_evolution.cpp:381-501 (AH) or mooreEvol.cpp:184-337 + 647-694
 * (Moore) over the whole lattice, periodic wrap in the index math instead
 * of the MPI halo exchange (mpi_evolution.cpp:925-943).  Enqueued only. */
int ax_step(ax_lattice *lat, int32_t nsteps);
/* Wait for all work enqueued on the handle. */
int ax_sync(ax_lattice *lat);
/* Tile shape of the marching step kernel (no reference counterpart: the
 * reference's site loop has no tiles, mpi_evolution.cpp:381-384): xseg =
 * x planes one wave advances through (0 = the whole x extent).  Results do
 * not depend on it.  ax_get_tile also reports whether the lattice uses the
 * marching kernel at all (periodic Abelian-Higgs; otherwise the site
 * kernel). */
int ax_set_tile(ax_lattice *lat, int32_t xseg);
int ax_get_tile(const ax_lattice *lat, int32_t *xseg, int32_t *tiled);
/* Measurement aid (no reference counterpart; the reference times only the
 * whole run, evolution.cpp:1462-1464): while enabled, each ax_step call
 * brackets its back-to-back step kernels with HIP events on the handle's
 * stream.  ax_kernel_time waits, returns the summed GPU milliseconds and
 * the number of steps they cover, and resets both. */
int ax_kernel_timer(ax_lattice *lat, int32_t enable);
int ax_kernel_time(ax_lattice *lat, double *ms, int64_t *steps);

/* Energy, Gauss-law deviation, finiteness and string counts of the current
 * state (evolution.cpp:386-427 + 1404-1408; Moore mooreEvol.cpp:352-389).
 * Synchronous.  want_winding = 0 skips the plaquette pass. */
int ax_diagnostics(ax_lattice *lat, int32_t want_winding, ax_diag *out);

/* nsteps TimeSteps as ax_step, taking at every TimeStep t of the call with
 * t % every == 0 the diagnostics ax_diagnostics would return just before
 * that step (the reference's per-step energy/Gauss sums, evolution.cpp:
 * 386-427, 1404-1408, at its output cadence): out[k] for the k-th such t,
 * at most cap of them, *count = how many.  On FP64 periodic Abelian-Higgs
 * lattices (evolution.cpp energy form, g != 0) the step at t computes them
 * from the values and link trigonometry it uses for its own update, with
 * the arithmetic of the separate pass, the step before it storing the
 * divergence of its links; other lattices take the separate pass.  Same
 * fields, the same energy and Gauss sums to 1e-13 and the same windings as
 * ax_diagnostics + ax_step.  Waits for the device at the end. */
int ax_step_diag(ax_lattice *lat, int32_t nsteps, int32_t every, int32_t want_winding, ax_diag *out, size_t cap,
                 size_t *count);

/* Per-plaquette winding numbers of the current state, 3*nx*ny*nz int8 in
 * [plane][x][y][z] order, plane 0 = xy, 1 = xz, 2 = yz (plaquette spanned
 * from site s along the two axes).  No counterpart in the reference; see
 * the library documentation for the definition. */
int ax_winding(ax_lattice *lat, int8_t *out, size_t n);
/* Same for a chosen scalar: field 0 = phi (AH) or phi1 (Moore, links
 * scaled by q1), field 1 = phi2 (Moore, links scaled by q2). */
int ax_winding_field(ax_lattice *lat, int32_t field, int8_t *out, size_t n);

/* Face crossings of phi = 0 from the bilinear-interpolation test of
 * mpi_evolution.cpp:536-789 / evolution.cpp:723-1016, in the reference's
 * order (site x-major, then faces xy, xz, yz, then + root before - root).
 * Writes min(count, cap) points as x,y,z triples (3*cap doubles) and the
 * total count to *count. */
int ax_crossings(ax_lattice *lat, double *xyz, size_t cap, size_t *count);

/* ax_crossings without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops).  No
 * counterpart in the reference.  The same points in the same order as
 * ax_crossings, written into xyz, a DEVICE buffer of cap points (3*cap
 * doubles) owned by the caller; the total count goes to *count, a DEVICE
 * int64.  Enqueued on the handle's stream only (ax_sync before reading).
 * The lattice's device must be current.  Its per-block scratch is one fixed
 * array per GPU inside the library (6 MiB, held until the process ends), so
 * lattices above 2^27 sites (512^3) are refused (AX_EINVAL), and calls that
 * overlap on one GPU race on that array. */
int ax_crossings_noalloc(ax_lattice *lat, double *xyz, size_t cap, int64_t *count);

/* String length from a list of crossing points (x,y,z triples, e.g. from
 * ax_crossings) by the greedy nearest-neighbour chaining of
 * ic_w_evolution.cpp:961-1087: out[0] = total length, out[1] / out[2] = the
 * parts whose starting point lies inside / outside the central half-box
 * (its splitLength).  Host only; O(n) with a cell grid instead of the
 * reference's O(n^2) search, same result bits.  periodic_xy / periodic_z
 * select the minimum-image distance per axis as the reference's xyBC / zBC. */
int ax_string_length(const double *xyz, size_t n, int32_t nx, int32_t ny, int32_t nz, double dx, double dy,
                     double dz, int32_t periodic_xy, int32_t periodic_z, double *out);

/* ------------------------------------------------------------------------
 * Straight-string initial condition (singleSemiAnalytic.cpp ->
 * straight_ic.cpp).
 *
 * ax_sor_profile: radial profile f(r_i), A(r_i), r_i = i*h, of a straight
 * string of winding n by successive over-relaxation of the static field
 * equations, singleSemiAnalytic.cpp:54-134 in its one-thread order (each
 * sweep relaxes i = 1 .. nx-2 in place; f(0) = A(0) = 0; f(nx-1), A(nx-1)
 * copy i = nx-2 before each sweep).  Sweeps stop after the first one whose
 * largest |F| is not above tol, or after max_iter.  init: nx (f, A) pairs,
 * or NULL for the reference's f = eta*tanh(r/2), A = 0 (its ic_type 0; 1
 * reads the previous output, which is passing it here).  out: nx (f, A)
 * pairs, the rows of Data/SOR_Fields.txt.  Runs on the current HIP device
 * and waits for it; allocates and frees its own buffers. */
typedef struct ax_sor_params {
  int32_t nx;          /* grid points (singleSemiAnalytic.cpp:19) */
  int32_t max_iter;    /* :22 */
  double h;            /* spacing :20 */
  double w;            /* relaxation factor :21 */
  double tol;          /* :23 */
  double n;            /* winding :29 */
  double lambda, eta;  /* :30-31 */
  double g;            /* gauge coupling; 0 = global string :32 */
} ax_sor_params;
void ax_sor_default_params(ax_sor_params *p);
int ax_sor_profile(const ax_sor_params *p, const double *init, double *out, int32_t *iterations, double *residual);

/* Straight string along z through (x0, y0) = ((nx-1)/2, (ny-1)/2) (integer
 * division), interpolated from a profile of nprof (f, A) pairs spaced dr,
 * as straight_ic.cpp:59-146: quadratic through the three profile points
 * around the nearest one (linear next to r = 0); beyond the profile |phi| =
 * 1 and A = n/g (0 if g = 0).  phi = |phi| (x, y)/r, A = A(r) (-y, x)/r^2,
 * stored as the link variables (dx g A_x, dy g A_y, 0) at both levels
 * (a stationary start, as the ic.txt reader of evolution.cpp:102-117); the
 * global model gets phi and zero momentum.  TimeStep is reset to 0.
 * *off_grid (may be NULL) counts sites beyond the profile.  The reference
 * reads one entry past its table when the nearest point is the last one;
 * here that point is also treated as beyond the profile. */
int ax_init_straight_string(ax_lattice *lat, const double *profile, int32_t nprof, double dr, double n, double g,
                            int64_t *off_grid);

/* ------------------------------------------------------------------------
 * Ensemble pool: many independent realizations (seed, N, steps) of one
 * periodic Abelian-Higgs parameter set, served with continuous batching.
 * No counterpart in the reference, which runs one realization per job
 * (mpi_evolution.cpp:48 seed, :354 TimeStep loop) and leaves ensembles to
 * job scripts.
 *
 *  - Requests are admitted in arrival order whenever their state fits in the
 *    pool's HBM budget and a slot is free; admitted lattices stay resident
 *    until they finish.
 *  - All resident lattices advance together: each chunk of chunk_steps
 *    TimeSteps is one replay of a captured hipGraph holding chunk_steps
 *    launches of one persistent kernel that walks every resident lattice's
 *    work items, so lattices of different N share each launch.
 *  - After every chunk the wound-plaquette counts of every resident lattice
 *    are copied back; each request accumulates its xi(t) series
 *    (t = 0, the end of every chunk, and its last step).
 *  - Each lattice evolves exactly as it would alone through ax_step() with
 *    the same parameters (same kernels' arithmetic per site), so results do
 *    not depend on what else is resident.
 * The pool runs its own scheduler thread; the calls below are thread-safe.
 * ---------------------------------------------------------------------- */
typedef struct ax_pool ax_pool;

typedef struct ax_pool_config {
  size_t hbm_budget;       /* device bytes for lattice slots (state + winding edge phases); 0 = 80% of free memory */
  int32_t max_lattices;    /* resident at once, 1..4096 (0 = 1024) */
  int32_t chunk_steps;     /* TimeSteps per graph replay and xi cadence (0 = 10) */
  int32_t ic_threads;      /* host threads building initial conditions (0 = 8) */
  int32_t use_graph;       /* 1 replay a captured hipGraph, 0 plain launches */
  int32_t autotune;        /* 1 time tile shapes per N on first use, 0 use xseg_default */
  int32_t xseg_default;    /* tile x extent when not tuned (0 = 32) */
  int32_t reserved;
} ax_pool_config;

#define AX_POOL_KEEP_STATE 1   /* keep the lattice after it finishes (ax_pool_get_state) */

#define AX_REQ_QUEUED 0
#define AX_REQ_RUNNING 1
#define AX_REQ_DONE 2
#define AX_REQ_FAILED 3        /* non-finite state found; its series stops there */

typedef struct ax_request_info {
  int64_t id;
  uint32_t seed;
  int32_t n, steps, status;
  int32_t nsamples;        /* xi samples so far */
  int32_t xseg;            /* tile x extent used */
  double t_submit, t_admit, t_done;   /* seconds since pool creation (0 if not yet) */
} ax_request_info;

typedef struct ax_pool_stats {
  int64_t submitted, admitted, completed, failed;
  int64_t chunks;                /* graph replays */
  int32_t resident, queued, max_resident, reserved;
  double site_updates;           /* sum over completed steps of N^3 */
  double elapsed;                /* seconds since pool creation */
  size_t budget, in_use, peak_in_use;
  double step_ms;                /* GPU milliseconds of the step launches of those steps
                                    (HIP events around each chunk's steps) */
} ax_pool_stats;

void ax_pool_default_config(ax_pool_config *cfg);
/* tmpl: model AH, periodic, precision 32 or 64; its nx/ny/nz are ignored
 * (each request brings its own N).  Allocates the budget up front. */
int ax_pool_create(const ax_params *tmpl, const ax_pool_config *cfg, ax_pool **out);
/* Stops the scheduler (pending requests are dropped) and frees everything. */
void ax_pool_destroy(ax_pool *pool);
/* Queue a realization: random-phase IC from seed (as ax_init_random_phase),
 * then steps >= 1 TimeSteps on an n^3 lattice.  AX_EBUDGET if its slot
 * alone exceeds the budget. */
int ax_pool_submit(ax_pool *pool, uint32_t seed, int32_t n, int32_t steps, int32_t flags, int64_t *id);
/* Wait until request id (id < 0: every request not yet released) is done
 * or failed: AX_OK, or 1 if timeout_s (< 0: no limit) passed first. */
int ax_pool_wait(ax_pool *pool, int64_t id, double timeout_s);
int ax_pool_info(ax_pool *pool, int64_t id, ax_request_info *out);
/* The xi(t) series so far: timestep, wound plaquettes (sum over the three
 * planes) and xi per sample; writes min(count, cap) entries. */
int ax_pool_xi(ax_pool *pool, int64_t id, int32_t *timestep, int64_t *wound, double *xi, size_t cap,
               size_t *count);
/* Final state of a finished AX_POOL_KEEP_STATE request, reference layout
 * (as ax_get_state). */
int ax_pool_get_state(ax_pool *pool, int64_t id, double *host, size_t n);
/* Forget a finished request (frees a kept state). */
int ax_pool_release(ax_pool *pool, int64_t id);
int ax_pool_stats_get(ax_pool *pool, ax_pool_stats *out);
/* Autotuner: time each candidate tile x extent (4, 8, 16 or 32 planes) with
 * the batched kernel on a table of scratch n^3 lattices (enough sites for
 * four 32-plane work items per workgroup of the persistent grid, up to 256
 * lattices, fewer when the arena lacks room: the pool serves many lattices
 * at once) and
 * keep the fastest for requests of that N.  Runs on its own stream (the
 * scheduler tunes a new N with its chunks drained; a direct call times
 * alongside whatever the pool is running); returns the chosen xseg and the
 * table's time per step. */
int ax_pool_tune(ax_pool *pool, int32_t n, int32_t *xseg, double *ms_per_step);

#ifdef __cplusplus
}
#endif

#endif /* AXIONS_AXIONS_H */
