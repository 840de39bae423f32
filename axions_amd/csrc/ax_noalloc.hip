/*
 * ax_noalloc.hip - ax_crossings_noalloc: the face-crossing detector of
 * ax_crossings (mpi_evolution.cpp:536-789) for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops).
 *
 * ax_crossings allocates its per-block counts and offsets and its point
 * buffer on first use (and the point buffer again whenever the capacity
 * grows).  Here the points go straight into a device buffer the caller owns,
 * and the counts and offsets live in `scratch`, a fixed array in the
 * library's device code: no hipMalloc, hipFree or other allocating call is
 * made.  Every GPU the library runs on holds its own copy of the array
 * until the process ends; its device address is looked up once per device
 * (hipGetSymbolAddress, on the first call with that device current) and
 * only that address is passed to kernels.
 *
 * The array is shared by every call on a GPU: calls that overlap on one GPU
 * (different streams or host threads) race on it.  Nothing in it needs
 * clearing: each call's count pass writes every block count it later reads,
 * and the scan writes every offset.
 */
#include <climits>
#include <mutex>

#include "ax_internal.h"

namespace ax {

/* block counts (int) + block offsets (long long) of up to kNoallocBlocks
 * 256-site blocks: 2^19 blocks = 2^27 sites (512^3), 6 MiB */
constexpr size_t kNoallocBlocks = (size_t)1 << 19;
constexpr size_t kNoallocCountBytes = kNoallocBlocks * sizeof(int);
constexpr size_t kNoallocBytes = kNoallocCountBytes + kNoallocBlocks * sizeof(long long);

static __device__ __align__(256) unsigned char scratch[kNoallocBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static unsigned char *g_scratch_ptr[kMaxDevices];   /* device address of `scratch` per device, or null */

/* The device address of `scratch` on the current device (looked up once). */
static int scratch_address(unsigned char **out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) return fail(AX_EINVAL, "device %d beyond the %d supported", dev, kMaxDevices);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    g_scratch_ptr[dev] = (unsigned char *)ptr;
  }
  *out = g_scratch_ptr[dev];
  return AX_OK;
}

}  // namespace ax

using namespace ax;

extern "C" int ax_crossings_noalloc(ax_lattice *L, double *xyz, size_t cap, int64_t *count) {
  if (!L || !count || (cap && !xyz)) return fail(AX_EINVAL, "NULL argument");
  const ax_params &p = lattice_params(L);
  const size_t V = lattice_sites(L);
  const size_t nb = (V + 255) / 256;
  if (nb > kNoallocBlocks)
    return fail(AX_EINVAL, "lattice of %zu sites exceeds the %zu sites ax_crossings_noalloc supports", V,
                kNoallocBlocks * 256);
  if (cap > (size_t)LLONG_MAX / 3) return fail(AX_EINVAL, "capacity %zu", cap);
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev != lattice_device(L))
    return fail(AX_EINVAL, "the lattice's device must be current");
  unsigned char *base = nullptr;
  int rc = scratch_address(&base);
  if (rc) return rc;
  int *bcount = (int *)base;
  long long *boff = (long long *)(base + kNoallocCountBytes);
  rc = prepare_faces(L);
  if (rc) return rc;
  const bool glob = p.model == AX_MODEL_GLOBAL;
  const int ts = lattice_timestep(L);
  const int tN = glob ? ts % 2 : (ts + 1) % 2;
  const int ex = !glob && p.boundary != AX_BC_PERIODIC, ez = !glob && p.boundary_z != AX_BC_PERIODIC;
  const size_t pitch = lattice_pitch(L);
  hipError_t e;
  if (p.precision == 64) {
    const double *b = (const double *)lattice_state(L);
    e = launch_crossings<double>(p.nx, p.ny, p.nz, p.dx, p.dy, p.dz, b + (size_t)tN * pitch, b + (size_t)(2 + tN) * pitch,
                                 ex, ez, bcount, boff, (long long *)count, xyz, (long long)cap, lattice_stream(L),
                                 cap > 0);
  } else {
    const float *b = (const float *)lattice_state(L);
    e = launch_crossings<float>(p.nx, p.ny, p.nz, p.dx, p.dy, p.dz, b + (size_t)tN * pitch, b + (size_t)(2 + tN) * pitch,
                                ex, ez, bcount, boff, (long long *)count, xyz, (long long)cap, lattice_stream(L),
                                cap > 0);
  }
  return e == hipSuccess ? AX_OK : hip_fail(e, "crossing launch");
}
