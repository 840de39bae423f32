/*
 * ax_pool.hip - ensemble pool: continuous batching of many independent
 * periodic Abelian-Higgs realizations of uneven N on one GPU.
 *
 * Device side
 *   pool_step_kernel   persistent kernel; one launch advances every resident
 *                      lattice by one TimeStep.  Its work items are the
 *                      (z chunk, y row, x segment) marches of ah_march for all
 *                      lattices back to back; a fixed grid of one-wave
 *                      workgroups strides through them.  Which lattices exist,
 *                      where their state lives, their timestep and tile shape
 *                      come from a table in device memory rewritten before
 *                      each chunk, so the launch itself never changes and a
 *                      chunk of K steps is one captured hipGraph replayed
 *                      chunk after chunk.
 *   pool_wind_pass_kernel
 *                      the winding passes over the same work items, each
 *                      item a march along x (ax_wind.h): first the FP32
 *                      edge values (forward edge phase + link angle) of
 *                      every resident lattice into its slot, plus a
 *                      non-finite count, then the wound-plaquette counts
 *                      from them (as the single-lattice diagnostics form
 *                      them), summed per lattice
 *                      with integer atomics (order-independent, so
 *                      deterministic).
 *   pool_zero_kernel   clears those counters (a kernel, not a memset, so the
 *                      graph holds kernel nodes only).
 *
 * Host side
 *   A scheduler thread owns the compute stream: it harvests finished chunks
 *   (xi samples, completions, freed memory), admits queued requests in
 *   arrival order while the HBM arena and slot count allow, and keeps up to
 *   two chunks in flight.  IC worker threads build each request's
 *   random-phase scalar field on the host (the libstdc++ mt19937 stream of
 *   ax_init_random_phase) into pinned staging buffers ahead of admission; the
 *   upload runs on a second stream that the next chunk waits for.
 */
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <random>
#include <thread>
#include <unordered_map>
#include <vector>

#include "ax_internal.h"
#include "ax_march.h"
#include "ax_sched.h"
#include "ax_wind.h"

namespace ax {

struct BatchEntry {
  uint64_t base;        /* device address of plane (component 0, level 0) */
  uint64_t pitch;       /* elements from one plane to the next */
  int32_t n;            /* lattice edge (n^3 sites) */
  int32_t nw, xseg, nseg;
  int32_t work_begin;   /* first work item of this lattice */
  int32_t t0;           /* TimeStep at the start of the chunk */
  int32_t steps;        /* TimeSteps to take in this chunk (<= chunk length) */
  int32_t pre;          /* 1: count windings before the first step (t0 == 0) */
  int32_t pad[2];
  uint64_t phase;       /* device address of the lattice's FP32 edge values [3][n^3] (ax_wind.h) */
};
static_assert(sizeof(BatchEntry) == 64, "entry layout");

struct BatchHeader {
  int32_t count;        /* lattices in the table */
  int32_t work;         /* total work items */
  int32_t pad[2];
};

/* counters per table entry: [phase 0: xy, xz, yz, bad][phase 1: xy, xz, yz, bad] */
constexpr int kCnt = 8;

/* the table entry whose work items include item w */
__device__ __forceinline__ int find_entry(const BatchEntry *ent, int count, int w) {
  int lo = 0, hi = count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    const int b = ent[mid].work_begin;
    if (b <= w) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <typename R>
__device__ __forceinline__ Levels<R> entry_levels(const BatchEntry &E, int t) {
  Levels<R> L;
  R *base = (R *)E.base;
  const int tN = (t + 1) % 2, tP = t % 2;
  for (int c = 0; c < 7; ++c) {
    const int cc = c < 5 ? c : 0;
    L.cur[c] = base + ((size_t)cc * 2 + tN) * E.pitch;
    L.prv[c] = base + ((size_t)cc * 2 + tP) * E.pitch;
  }
  return L;
}

template <typename R, bool POW2>
__global__ void __launch_bounds__(64) pool_step_kernel(ax_params p, R lim, const BatchHeader *__restrict__ hdr,
                                                       const BatchEntry *__restrict__ ent, int j) {
  const int count = hdr->count, work = hdr->work;
  /* XCD-aware order: the grid is resident at once and its workgroups are
   * handed to the eight XCDs round-robin, so workgroup b strides through
   * the w of one residue mod 8 on XCD b mod 8 (the grid is a multiple of
   * 8).  Item it = (w mod 8) * q + w / 8 gives XCD k the k-th eighth of the
   * items -- consecutive chunks and rows of a lattice's segment, whose shared
   * neighbour rows and chunk-edge lines then sit in one L2 -- instead of
   * every eighth item.  The map is a bijection of [0, 8q), so every item is
   * stepped once whatever the grid. */
  const int q = (work + 7) >> 3;
  for (int w = (int)blockIdx.x; w < 8 * q; w += (int)gridDim.x) {
    const int it = xcd_item(w, q);
    if (it >= work) continue;
    const int e = find_entry(ent, count, it);
    const BatchEntry E = ent[e];
    if (j >= E.steps) continue;
    const int local = it - E.work_begin;
    const int zc = local % E.nw, r = local / E.nw;
    const int y = r % E.n, seg = r / E.n;
    const int x0 = seg * E.xseg, x1 = min(E.n, x0 + E.xseg);
    const int t = E.t0 + j;
    MarchConsts<R> m;
    march_consts_from<R>(make_step_consts(p, t), lim, m);
    ah_march<R, POW2, 2>(m, entry_levels<R>(E, t), E.n, E.n, E.n, y, zc, x0, x1);
  }
}

__device__ __forceinline__ int wave_isum(int v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

/* Adds the wave's per-lane counts v[0..nc) to counters dst[0..nc) (lane 0,
 * integer atomics: the totals do not depend on the order). */
template <int NC>
__device__ __forceinline__ void flush_counts(const int *v, unsigned long long *dst) {
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const int tot = wave_isum(v[c]);
    if ((threadIdx.x & 63) == 0 && tot) atomicAdd(&dst[c], (unsigned long long)tot);
  }
}

/* The winding passes march over the step kernel's work items (same table,
 * same XCD-aware order): item -> lattice, z chunk, row, x segment.  A
 * workgroup keeps its counts in registers while consecutive items belong to
 * one lattice and adds them to that lattice's counters when the lattice
 * changes.  which = 0: the state before the chunk's first step (lattices
 * with pre set), 1: the state after the chunk's steps. */
template <typename T, bool WIND>
__global__ void __launch_bounds__(64) pool_wind_pass_kernel(const BatchHeader *__restrict__ hdr,
                                                            const BatchEntry *__restrict__ ent,
                                                            unsigned long long *counts, int which) {
  const int count = hdr->count, work = hdr->work;
  const int q = (work + 7) >> 3;
  int ce = -1;
  int v[3] = {0, 0, 0};
  /* edge values: v[0] counts sites with a non-finite field (counter 3);
   * windings: v[0..2] sum |W| of the xy, xz, yz plaquettes (counters 0..2) */
  auto flush = [&]() {
    if (ce < 0) return;
    unsigned long long *dst = counts + (size_t)ce * kCnt + which * 4;
    if (WIND) flush_counts<3>(v, dst);
    else flush_counts<1>(v, dst + 3);
    v[0] = v[1] = v[2] = 0;
  };
  for (int w = (int)blockIdx.x; w < 8 * q; w += (int)gridDim.x) {
    const int it = xcd_item(w, q);
    if (it >= work) continue;
    const int e = find_entry(ent, count, it);
    const BatchEntry E = ent[e];
    if (which == 0 && !E.pre) continue;
    if (e != ce) {
      flush();
      ce = e;
    }
    const int local = it - E.work_begin;
    const int zc = local % E.nw, r = local / E.nw;
    const int y = r % E.n, seg = r / E.n;
    const int x0 = seg * E.xseg, x1 = min(E.n, x0 + E.xseg);
    const Levels<T> L = entry_levels<T>(E, E.t0 + (which == 0 ? 0 : E.steps));   /* state after t TimeSteps */
    const WindRow row(E.n, E.n, E.n, y, zc);
    if (WIND) {
      wind_march(row, (const float *)E.phase, x0, x1, [&](size_t, const int *wn) {
#pragma unroll
                      for (int c = 0; c < 3; ++c) v[c] += abs(wn[c]);
                    });
    } else {
      v[0] += phase_march<T>(row, L.cur[0], L.cur[1], L.cur[2], L.cur[3], L.cur[4], (float *)E.phase, x0, x1);
    }
  }
  flush();
}

__global__ void __launch_bounds__(256) pool_zero_kernel(const BatchHeader *__restrict__ hdr,
                                                        unsigned long long *counts) {
  const int n = hdr->count * kCnt;
  for (int i = (int)threadIdx.x; i < n; i += 256) counts[i] = 0;
}

}  // namespace ax

using namespace ax;

namespace {

double now_s(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

struct Staging {
  void *host = nullptr;      /* pinned: re plane then im plane */
  size_t cap = 0;
  int busy = 0;              /* 0 free, 1 being filled or waiting, 2 upload in flight */
  hipEvent_t done = nullptr; /* upload finished (busy == 2) */
};

struct Req {
  int64_t id = 0;
  uint32_t seed = 0;
  int n = 0, steps = 0, flags = 0;
  int status = AX_REQ_QUEUED;
  double t_submit = 0, t_admit = 0, t_done = 0;
  int ic = 0;                /* 0 waiting, 1 building, 2 ready */
  int staging = -1;
  int64_t off = -1;          /* arena offset while resident */
  size_t bytes = 0;
  int t = 0;                 /* TimeSteps enqueued */
  bool pre_done = false;
  int xseg = 32;
  hipEvent_t upload = nullptr;
  bool wait_upload = false;
  std::vector<int32_t> ts;
  std::vector<int64_t> wound;
  std::vector<double> xi;
};

struct ChunkRec {
  int64_t seq = 0;
  int buf = 0;
  hipEvent_t done = nullptr;
  hipEvent_t s0 = nullptr, s1 = nullptr;   /* around the chunk's step launches */
  struct Item {
    int64_t id;
    int t0, steps, pre;
  };
  std::vector<Item> items;
  double site_updates = 0;
};

}  // namespace

struct ax_pool {
  ax_params p{};
  ax_pool_config cfg{};
  size_t esize = 8;
  size_t pad_bytes = 4352;
  bool pow2 = false;
  double lim64 = 0;
  float lim32 = 0;
  std::chrono::steady_clock::time_point t0;

  int device = 0;
  hipStream_t cs = nullptr, us = nullptr;   /* compute, upload */
  char *arena_base = nullptr;
  Arena arena;
  BatchHeader *d_hdr = nullptr;
  BatchEntry *d_ent = nullptr;
  unsigned long long *d_cnt = nullptr;
  static constexpr int kBufs = 3;
  char *h_tab[kBufs] = {};          /* pinned: header + entries */
  unsigned long long *h_cnt[kBufs] = {};
  int grid_step = 0, grid_wind = 0;
  /* a chunk is three parts: zero counters + windings of new lattices, the
   * K step launches, windings after the steps; one captured graph each, so
   * that events between them time the step launches alone */
  static constexpr int kParts = 3;
  hipGraphExec_t gexec[kParts] = {};
  hipGraph_t graph[kParts] = {};

  std::mutex mu;
  std::condition_variable cv_sched, cv_ic, cv_done;
  bool stop = false;
  std::string error;
  std::deque<int64_t> queue;        /* not yet admitted, arrival order */
  std::vector<int64_t> resident;    /* admission order */
  std::unordered_map<int64_t, std::unique_ptr<Req>> reqs;
  std::vector<Staging> staging;
  std::deque<ChunkRec> inflight;
  /* arena slots of failed lattices: (offset, last chunk that may still touch
   * it), released once that chunk is harvested, whatever happens to the
   * request meanwhile (ax_pool_release may forget it earlier) */
  std::vector<std::pair<size_t, int64_t>> deferred;
  int64_t next_seq = 0;
  int next_buf = 0;
  int64_t next_id = 1;
  std::map<int, int> tuned;         /* n -> xseg */
  ax_pool_stats st{};
  std::thread sched;
  std::vector<std::thread> icw;

  int launch_chunk_kernels(hipStream_t s, int part);
  void launch_winding(hipStream_t s, int which);
  int capture_graph();
  void scheduler();
  void ic_worker();
  void harvest(ChunkRec &c);
  bool admit_head();
  int enqueue_chunk();
  int tune(int n, int *xseg, double *ms);
  void fail_all(const std::string &why);
  void free_slot(Req &r);
};

namespace {

template <typename R>
void build_ic(uint32_t seed, size_t V, double eta, R *re, R *im) {
  /* the stream and arithmetic of ax_init_random_phase (mpi_evolution.cpp:287-305) */
  const double pi = 4.0 * atan(1.0);
  std::mt19937 gen(seed);
  std::uniform_real_distribution<double> dist(0.0, 2 * pi);
  for (size_t i = 0; i < V; ++i) {
    double s, c;
    sincos(dist(gen), &s, &c);
    re[i] = (R)(eta * c);
    im[i] = (R)(eta * s);
  }
}

}  // namespace

/* edge phases, then windings from them (which: 0 before the steps, 1 after) */
void ax_pool::launch_winding(hipStream_t s, int which) {
  if (esize == 8) {
    hipLaunchKernelGGL((pool_wind_pass_kernel<double, false>), dim3(grid_wind), dim3(64), 0, s, d_hdr, d_ent, d_cnt, which);
    hipLaunchKernelGGL((pool_wind_pass_kernel<double, true>), dim3(grid_wind), dim3(64), 0, s, d_hdr, d_ent, d_cnt, which);
  } else {
    hipLaunchKernelGGL((pool_wind_pass_kernel<float, false>), dim3(grid_wind), dim3(64), 0, s, d_hdr, d_ent, d_cnt, which);
    hipLaunchKernelGGL((pool_wind_pass_kernel<float, true>), dim3(grid_wind), dim3(64), 0, s, d_hdr, d_ent, d_cnt, which);
  }
}

int ax_pool::launch_chunk_kernels(hipStream_t s, int part) {
  const int K = cfg.chunk_steps;
  if (part == 0) {
    hipLaunchKernelGGL(pool_zero_kernel, dim3(1), dim3(256), 0, s, d_hdr, d_cnt);
    launch_winding(s, 0);
  }
  for (int j = 0; part == 1 && j < K; ++j) {
    if (esize == 8) {
      if (pow2) hipLaunchKernelGGL((pool_step_kernel<double, true>), dim3(grid_step), dim3(64), 0, s, p, lim64, d_hdr, d_ent, j);
      else hipLaunchKernelGGL((pool_step_kernel<double, false>), dim3(grid_step), dim3(64), 0, s, p, lim64, d_hdr, d_ent, j);
    } else {
      if (pow2) hipLaunchKernelGGL((pool_step_kernel<float, true>), dim3(grid_step), dim3(64), 0, s, p, lim32, d_hdr, d_ent, j);
      else hipLaunchKernelGGL((pool_step_kernel<float, false>), dim3(grid_step), dim3(64), 0, s, p, lim32, d_hdr, d_ent, j);
    }
  }
  if (part == 2) launch_winding(s, 1);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? AX_OK : hip_fail(e, "pool chunk launch");
}

int ax_pool::capture_graph() {
  for (int part = 0; part < kParts; ++part) {
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return hip_fail(e, "begin capture");
    int rc = launch_chunk_kernels(cs, part);
    e = hipStreamEndCapture(cs, &graph[part]);
    if (rc) return rc;
    if (e != hipSuccess) return hip_fail(e, "end capture");
    e = hipGraphInstantiate(&gexec[part], graph[part], nullptr, nullptr, 0);
    if (e != hipSuccess) return hip_fail(e, "graph instantiate");
  }
  return AX_OK;
}

void ax_pool::free_slot(Req &r) {
  if (r.off >= 0) arena.release((size_t)r.off);
  r.off = -1;
  if (r.upload) {
    (void)hipEventDestroy(r.upload);
    r.upload = nullptr;
  }
  resident.erase(std::remove(resident.begin(), resident.end(), r.id), resident.end());
}

void ax_pool::fail_all(const std::string &why) {
  error = why;
  for (auto &kv : reqs) {
    Req &r = *kv.second;
    if (r.status == AX_REQ_QUEUED || r.status == AX_REQ_RUNNING) {
      r.status = AX_REQ_FAILED;
      r.t_done = now_s(t0);
      st.failed++;
    }
  }
  queue.clear();
  stop = true;
  cv_done.notify_all();
  cv_ic.notify_all();
}

void ax_pool::harvest(ChunkRec &c) {
  const unsigned long long *cnt = h_cnt[c.buf];
  const double dx = p.dx, dy = p.dy, dz = p.dz;
  for (size_t e = 0; e < c.items.size(); ++e) {
    const ChunkRec::Item &it = c.items[e];
    auto f = reqs.find(it.id);
    if (f == reqs.end()) continue;
    Req &r = *f->second;
    if (r.status != AX_REQ_RUNNING) continue;
    const double V = (double)r.n * r.n * r.n;
    for (int ph = 0; ph < 2; ++ph) {
      if (ph == 0 && !it.pre) continue;
      const unsigned long long *v = cnt + e * kCnt + ph * 4;
      if (v[3]) {                          /* non-finite value in the state */
        r.status = AX_REQ_FAILED;
        r.t_done = now_s(t0);
        st.failed++;
        break;
      }
      r.ts.push_back(ph == 0 ? it.t0 : it.t0 + it.steps);
      r.wound.push_back((int64_t)(v[0] + v[1] + v[2]));
      r.xi.push_back(xi_of((double)v[0], (double)v[1], (double)v[2], V, dx, dy, dz));
    }
    if (r.status == AX_REQ_RUNNING && it.t0 + it.steps >= r.steps) {
      r.status = AX_REQ_DONE;
      r.t_done = now_s(t0);
      st.completed++;
    }
    if (r.status == AX_REQ_FAILED) {
      /* chunks enqueued after this one may still include the lattice:
       * free its memory only once the last of them has been harvested */
      r.t = r.steps;
      resident.erase(std::remove(resident.begin(), resident.end(), r.id), resident.end());
      if (r.off >= 0) deferred.emplace_back((size_t)r.off, next_seq - 1);
      r.off = -1;
    } else if (r.status == AX_REQ_DONE) {
      if (r.flags & AX_POOL_KEEP_STATE)
        resident.erase(std::remove(resident.begin(), resident.end(), r.id), resident.end());
      else
        free_slot(r);
    }
  }
  for (size_t i = 0; i < deferred.size();) {
    if (deferred[i].second <= c.seq) {
      arena.release(deferred[i].first);
      deferred[i] = deferred.back();
      deferred.pop_back();
    } else {
      ++i;
    }
  }
  st.site_updates += c.site_updates;
  float ms = 0;
  if (hipEventElapsedTime(&ms, c.s0, c.s1) == hipSuccess) st.step_ms += ms;
  (void)hipEventDestroy(c.s0);
  (void)hipEventDestroy(c.s1);
  (void)hipEventDestroy(c.done);
  cv_done.notify_all();
}

/* Admit the head of the queue if its IC is ready and its state fits. */
bool ax_pool::admit_head() {
  if (queue.empty() || (int)resident.size() >= cfg.max_lattices) return false;
  Req &r = *reqs[queue.front()];
  if (r.ic != 2) return false;
  int64_t off = arena.alloc(r.bytes);
  if (off < 0) return false;
  queue.pop_front();
  r.off = off;
  Staging &sg = staging[r.staging];
  const size_t V = (size_t)r.n * r.n * r.n;
  const size_t pitch = lattice_pitch(V, esize, pad_bytes);
  char *base = arena_base + off;
  const size_t plane = V * esize, pb = pitch * esize;
  /* phi at both levels from the staging buffer, links and their previous
   * level zero: planes (c*2 + level) */
  hipError_t e = hipMemcpyAsync(base + 0 * pb, sg.host, plane, hipMemcpyHostToDevice, us);
  if (e == hipSuccess) e = hipMemcpyAsync(base + 2 * pb, (char *)sg.host + plane, plane, hipMemcpyHostToDevice, us);
  if (e == hipSuccess) e = hipMemcpyAsync(base + 1 * pb, base + 0 * pb, plane, hipMemcpyDeviceToDevice, us);
  if (e == hipSuccess) e = hipMemcpyAsync(base + 3 * pb, base + 2 * pb, plane, hipMemcpyDeviceToDevice, us);
  if (e == hipSuccess) e = hipMemsetAsync(base + 4 * pb, 0, 6 * pb, us);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&r.upload, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventRecord(r.upload, us);
  if (e == hipSuccess && !sg.done) e = hipEventCreateWithFlags(&sg.done, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventRecord(sg.done, us);
  if (e != hipSuccess) {
    fail_all(std::string("pool upload: ") + hipGetErrorString(e));
    return false;
  }
  sg.busy = 2;
  r.staging = -1;
  r.wait_upload = true;
  r.status = AX_REQ_RUNNING;
  r.t_admit = now_s(t0);
  auto tf = tuned.find(r.n);
  r.xseg = tf != tuned.end() ? tf->second : cfg.xseg_default;
  resident.push_back(r.id);
  st.admitted++;
  st.max_resident = std::max(st.max_resident, (int32_t)resident.size());
  return true;
}

int ax_pool::enqueue_chunk() {
  const int K = cfg.chunk_steps;
  const int b = next_buf;
  next_buf = (next_buf + 1) % kBufs;
  BatchHeader *h = (BatchHeader *)h_tab[b];
  BatchEntry *ent = (BatchEntry *)(h_tab[b] + 64);
  ChunkRec c;
  c.buf = b;
  c.seq = next_seq;
  int count = 0, work = 0;
  for (int64_t id : resident) {
    Req &r = *reqs[id];
    if (r.status != AX_REQ_RUNNING || r.t >= r.steps) continue;
    BatchEntry &E = ent[count];
    const size_t V = (size_t)r.n * r.n * r.n;
    E.base = (uint64_t)(uintptr_t)(arena_base + r.off);
    E.pitch = lattice_pitch(V, esize, pad_bytes);
    E.n = r.n;
    E.nw = (r.n + 63) / 64;
    E.xseg = (r.xseg <= 0 || r.xseg > r.n) ? r.n : r.xseg;
    E.nseg = (r.n + E.xseg - 1) / E.xseg;
    E.work_begin = work;
    E.t0 = r.t;
    E.steps = std::min(K, r.steps - r.t);
    E.pre = r.pre_done ? 0 : 1;
    E.pad[0] = E.pad[1] = 0;
    E.phase = (uint64_t)(uintptr_t)(arena_base + r.off + lattice_phase_offset(r.n, esize, pad_bytes));
    work += E.nw * r.n * E.nseg;
    c.items.push_back({id, E.t0, E.steps, E.pre});
    c.site_updates += (double)V * E.steps;
    if (r.wait_upload) {
      hipError_t e = hipStreamWaitEvent(cs, r.upload, 0);
      if (e != hipSuccess) return hip_fail(e, "wait upload");
      r.wait_upload = false;
    }
    r.t += E.steps;
    r.pre_done = true;
    ++count;
  }
  if (count == 0) return AX_OK;
  /* two items per workgroup of the persistent grid (fill_segments) */
  if (work < 2 * grid_step) {
    std::vector<int> nn(count), nw(count), xs(count);
    for (int k = 0; k < count; ++k) {
      nn[k] = ent[k].n;
      nw[k] = ent[k].nw;
      xs[k] = ent[k].xseg;
    }
    fill_segments(count, nn.data(), nw.data(), xs.data(), 2LL * grid_step);
    work = 0;
    for (int k = 0; k < count; ++k) {
      BatchEntry &E = ent[k];
      E.xseg = xs[k];
      E.nseg = (E.n + E.xseg - 1) / E.xseg;
      E.work_begin = work;
      work += E.nw * E.n * E.nseg;
    }
  }
  h->count = count;
  h->work = work;
  h->pad[0] = h->pad[1] = 0;
  hipError_t e = hipMemcpyAsync(d_hdr, h_tab[b], 64 + (size_t)count * sizeof(BatchEntry), hipMemcpyHostToDevice, cs);
  if (e != hipSuccess) return hip_fail(e, "table upload");
  e = hipEventCreate(&c.s0);
  if (e == hipSuccess) e = hipEventCreate(&c.s1);
  if (e != hipSuccess) {
    if (c.s0) (void)hipEventDestroy(c.s0);
    return hip_fail(e, "chunk timing events");
  }
  for (int part = 0; part < kParts; ++part) {
    if (part == 1) e = hipEventRecord(c.s0, cs);
    if (part == 2) e = hipEventRecord(c.s1, cs);
    int rc = e != hipSuccess ? hip_fail(e, "chunk timing event")
             : cfg.use_graph ? (hipGraphLaunch(gexec[part], cs) == hipSuccess ? AX_OK
                                                                              : hip_fail(hipGetLastError(), "graph launch"))
                             : launch_chunk_kernels(cs, part);
    if (rc) {
      (void)hipEventDestroy(c.s0);
      (void)hipEventDestroy(c.s1);
      return rc;
    }
  }
  e = hipMemcpyAsync(h_cnt[b], d_cnt, (size_t)count * kCnt * sizeof(unsigned long long), hipMemcpyDeviceToHost, cs);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c.done, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventRecord(c.done, cs);
  if (e != hipSuccess) {
    (void)hipEventDestroy(c.s0);
    (void)hipEventDestroy(c.s1);
    if (c.done) (void)hipEventDestroy(c.done);
    return hip_fail(e, "chunk record");
  }
  st.chunks++;
  next_seq++;
  inflight.push_back(std::move(c));
  return AX_OK;
}

void ax_pool::scheduler() {
  const hipError_t de = hipSetDevice(device);
  std::unique_lock<std::mutex> lk(mu);
  if (de != hipSuccess) fail_all(std::string("scheduler device: ") + hipGetErrorString(de));
  while (!stop) {
    /* staging buffers whose upload finished go back to the IC workers */
    for (auto &sg : staging)
      if (sg.busy == 2 && hipEventQuery(sg.done) == hipSuccess) {
        sg.busy = 0;
        cv_ic.notify_all();
      }
    /* finished chunks */
    while (!inflight.empty()) {
      hipError_t q = hipEventQuery(inflight.front().done);
      if (q == hipErrorNotReady) break;
      if (q != hipSuccess) {
        fail_all(std::string("chunk: ") + hipGetErrorString(q));
        break;
      }
      harvest(inflight.front());
      inflight.pop_front();
    }
    if (stop) break;
    /* a new N is tuned before its first admission, with the GPU drained */
    const bool head_tune = cfg.autotune && !queue.empty() && !tuned.count(reqs[queue.front()]->n);
    if (head_tune && inflight.empty()) {
      const int n = reqs[queue.front()]->n;
      int xs;
      double ms;
      lk.unlock();
      int rc = tune(n, &xs, &ms);
      lk.lock();
      if (rc) tuned[n] = cfg.xseg_default;
      continue;
    }
    bool admitted = false;
    while (!head_tune && admit_head()) admitted = true;
    if (stop) break;
    if ((int)inflight.size() < 2 && !head_tune) {
      bool any = false;
      for (int64_t id : resident) {
        Req &r = *reqs[id];
        if (r.status == AX_REQ_RUNNING && r.t < r.steps) { any = true; break; }
      }
      if (any) {
        int rc = enqueue_chunk();
        if (rc) {
          fail_all(std::string("enqueue: ") + ax_last_error());
          break;
        }
        continue;
      }
    }
    if (admitted) continue;
    if (!inflight.empty()) {
      hipEvent_t ev = inflight.front().done;
      lk.unlock();
      (void)hipEventSynchronize(ev);   /* its status is read by hipEventQuery above */
      lk.lock();
    } else {
      cv_sched.wait_for(lk, std::chrono::milliseconds(2));
    }
  }
  /* drain what is in flight before the streams go away */
  lk.unlock();
  (void)hipStreamSynchronize(cs);
  (void)hipStreamSynchronize(us);
}

void ax_pool::ic_worker() {
  std::unique_lock<std::mutex> lk(mu);
  for (;;) {
    Req *job = nullptr;
    int sgi = -1;
    cv_ic.wait(lk, [&] {
      if (stop) return true;
      job = nullptr;
      for (int64_t id : queue) {          /* earliest queued request without an IC */
        Req &r = *reqs[id];
        if (r.ic == 0) { job = &r; break; }
      }
      if (!job) return false;
      sgi = -1;
      for (size_t i = 0; i < staging.size(); ++i)
        if (staging[i].busy == 0) { sgi = (int)i; break; }
      return sgi >= 0;
    });
    if (stop) return;
    job->ic = 1;
    job->staging = sgi;
    Staging &sg = staging[sgi];
    sg.busy = 1;
    const size_t V = (size_t)job->n * job->n * job->n;
    const size_t need = 2 * V * esize;
    const uint32_t seed = job->seed;
    const double eta = p.eta;
    void *buf = sg.host;
    const bool grow = sg.cap < need;
    lk.unlock();
    bool ok = true;
    if (grow) {
      if (buf) (void)hipHostFree(buf);
      buf = nullptr;
      ok = hipHostMalloc(&buf, need, hipHostMallocDefault) == hipSuccess;
    }
    if (ok) {
      if (esize == 8) build_ic<double>(seed, V, eta, (double *)buf, (double *)buf + V);
      else build_ic<float>(seed, V, eta, (float *)buf, (float *)buf + V);
    }
    lk.lock();
    if (grow) {
      sg.host = buf;
      sg.cap = ok ? need : 0;
    }
    if (!ok) {
      sg.busy = 0;
      fail_all("pinned staging allocation failed");
      return;
    }
    job->ic = 2;
    cv_sched.notify_all();
  }
}

int ax_pool::tune(int n, int *xseg, double *ms) {
  /* m scratch lattices of edge n, random-phase IC, each candidate timed with
   * the batched kernel over an m-lattice table on its own stream.  The pool
   * serves many lattices at once: timed alone, a small lattice needs short x
   * segments to fill the GPU (64^3 chose 4 planes per wave), while among many
   * resident lattices longer segments pay the march's prologue less often.
   * So the table holds enough sites for four 32-plane work items per
   * workgroup of the persistent grid (enqueue_chunk shortens the segments of
   * a pool that holds fewer), as many lattices as the arena has room for. */
  const size_t V = (size_t)n * n * n;
  const size_t LB = lattice_bytes(n, esize, pad_bytes);
  const size_t target = std::max<size_t>((size_t)1 << 24, (size_t)4 * grid_step * 64 * 32);
  int m = (int)std::max<size_t>(1, std::min<size_t>(256, (target + V - 1) / V));
  int64_t off = -1;
  {
    std::lock_guard<std::mutex> g(mu);
    while (m > 1 && (off = arena.alloc(LB * (size_t)m)) < 0) m /= 2;
    if (off < 0) off = arena.alloc(LB);   /* m == 1 */
  }
  if (off < 0) return fail(AX_EBUDGET, "no room to tune n=%d", n);
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  BatchHeader *dh = nullptr;
  int rc = AX_OK;
  const size_t pitch = lattice_pitch(V, esize, pad_bytes);
  char *base = arena_base + off;
  std::vector<char> ic(2 * V * esize);
  if (esize == 8) build_ic<double>(12345u, V, p.eta, (double *)ic.data(), (double *)ic.data() + V);
  else build_ic<float>(12345u, V, p.eta, (float *)ic.data(), (float *)ic.data() + V);
  /* at most 32 planes per work item: a launch lasts as long as its longest
   * item, so longer segments (which the table's timing can favour) hold the
   * other lattices' workgroups idle at the end of each launch */
  std::vector<int> cand;
  for (int c : {4, 8, 16, 32, n})
    if (c <= n && c <= 32 && std::find(cand.begin(), cand.end(), c) == cand.end()) cand.push_back(c);
  double best = 1e300;
  int bestx = cfg.xseg_default;
  const size_t tab_bytes = 64 + (size_t)m * sizeof(BatchEntry);
  hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = hipMalloc((void **)&dh, tab_bytes);
  const size_t plane = V * esize, pb = pitch * esize;
  if (e == hipSuccess) e = hipMemcpyAsync(base, ic.data(), plane, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(base + 2 * pb, ic.data() + plane, plane, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(base + pb, base, plane, hipMemcpyDeviceToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(base + 3 * pb, base + 2 * pb, plane, hipMemcpyDeviceToDevice, s);
  if (e == hipSuccess) e = hipMemsetAsync(base + 4 * pb, 0, 6 * pb, s);
  /* the other lattices start from the same state (ten planes) */
  for (int k = 1; k < m && e == hipSuccess; ++k)
    e = hipMemcpyAsync(base + (size_t)k * LB, base, 10 * pb, hipMemcpyDeviceToDevice, s);
  std::vector<char> tab(tab_bytes);
  const int reps = 2, warm = 2, timed = 6;
  for (size_t ci = 0; ci < cand.size() && e == hipSuccess; ++ci) {
    std::fill(tab.begin(), tab.end(), 0);
    BatchHeader *h = (BatchHeader *)tab.data();
    BatchEntry *ent = (BatchEntry *)(tab.data() + 64);
    const int nw = (n + 63) / 64, nseg = (n + cand[ci] - 1) / cand[ci];
    for (int k = 0; k < m; ++k) {
      BatchEntry *E = ent + k;
      E->base = (uint64_t)(uintptr_t)(base + (size_t)k * LB);
      E->pitch = pitch;
      E->n = n;
      E->nw = nw;
      E->xseg = cand[ci];
      E->nseg = nseg;
      E->work_begin = k * nw * n * nseg;
      E->steps = 1 << 30;
    }
    h->count = m;
    h->work = m * nw * n * nseg;
    e = hipMemcpyAsync(dh, tab.data(), tab_bytes, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    for (int rep = 0; rep < reps && e == hipSuccess; ++rep) {
      for (int j = 0; j < warm + timed; ++j) {
        if (j == warm) (void)hipEventRecord(e0, s);   /* failures show in the event synchronize below */
        if (esize == 8) {
          if (pow2) hipLaunchKernelGGL((pool_step_kernel<double, true>), dim3(grid_step), dim3(64), 0, s, p, lim64, dh, (const BatchEntry *)((char *)dh + 64), j);
          else hipLaunchKernelGGL((pool_step_kernel<double, false>), dim3(grid_step), dim3(64), 0, s, p, lim64, dh, (const BatchEntry *)((char *)dh + 64), j);
        } else {
          if (pow2) hipLaunchKernelGGL((pool_step_kernel<float, true>), dim3(grid_step), dim3(64), 0, s, p, lim32, dh, (const BatchEntry *)((char *)dh + 64), j);
          else hipLaunchKernelGGL((pool_step_kernel<float, false>), dim3(grid_step), dim3(64), 0, s, p, lim32, dh, (const BatchEntry *)((char *)dh + 64), j);
        }
      }
      (void)hipEventRecord(e1, s);
      e = hipEventSynchronize(e1);
      float t = 0;
      if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
      if (e == hipSuccess && t / timed < best) {
        best = t / timed;
        bestx = cand[ci];
      }
    }
  }
  if (e != hipSuccess) rc = hip_fail(e, "tune");
  if (s) (void)hipStreamSynchronize(s);
  if (dh) (void)hipFree(dh);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (s) (void)hipStreamDestroy(s);
  std::lock_guard<std::mutex> g(mu);
  arena.release((size_t)off);
  if (!rc) {
    tuned[n] = bestx;
    *xseg = bestx;
    *ms = best;
  }
  return rc;
}

extern "C" {

void ax_pool_default_config(ax_pool_config *c) {
  if (!c) return;
  memset(c, 0, sizeof *c);
  c->hbm_budget = 0;
  c->max_lattices = 1024;
  c->chunk_steps = 10;
  c->ic_threads = 8;
  c->use_graph = 1;
  c->autotune = 1;
  c->xseg_default = 32;
}

int ax_pool_create(const ax_params *tmpl, const ax_pool_config *cfg_in, ax_pool **out) {
  if (!tmpl || !out) return fail(AX_EINVAL, "NULL argument");
  *out = nullptr;
  if (tmpl->model != AX_MODEL_AH || tmpl->boundary != AX_BC_PERIODIC || tmpl->boundary_z != AX_BC_PERIODIC)
    return fail(AX_EINVAL, "the pool runs periodic Abelian-Higgs lattices");
  if (tmpl->precision != 64 && tmpl->precision != 32) return fail(AX_EINVAL, "precision %d", tmpl->precision);
  if (!(tmpl->dx > 0 && tmpl->dy > 0 && tmpl->dz > 0 && tmpl->dt > 0)) return fail(AX_EINVAL, "spacings must be > 0");
  ax_pool_config cfg;
  ax_pool_default_config(&cfg);
  if (cfg_in) cfg = *cfg_in;
  if (cfg.max_lattices <= 0) cfg.max_lattices = 1024;
  if (cfg.max_lattices > 4096) return fail(AX_EINVAL, "max_lattices %d > 4096", cfg.max_lattices);
  if (cfg.chunk_steps <= 0) cfg.chunk_steps = 10;
  if (cfg.chunk_steps > 1000) return fail(AX_EINVAL, "chunk_steps %d > 1000", cfg.chunk_steps);
  if (cfg.ic_threads <= 0) cfg.ic_threads = 8;
  if (cfg.ic_threads > 64) cfg.ic_threads = 64;
  if (cfg.xseg_default <= 0) cfg.xseg_default = 32;

  std::unique_ptr<ax_pool> P(new (std::nothrow) ax_pool);
  if (!P) return fail(AX_ENOMEM, "host allocation");
  P->p = *tmpl;
  P->cfg = cfg;
  P->esize = tmpl->precision == 64 ? 8 : 4;
  const char *pe = getenv("AX_PLANE_PAD_BYTES");
  P->pad_bytes = pe ? (size_t)atol(pe) : 4352;
  P->pow2 = is_pow2(tmpl->dx * tmpl->dx) && is_pow2(tmpl->dy * tmpl->dy) && is_pow2(tmpl->dz * tmpl->dz);
  P->lim64 = rate_limit<double>(tmpl->dt);
  P->lim32 = rate_limit<float>((float)tmpl->dt);
  P->t0 = std::chrono::steady_clock::now();
  (void)hipGetDevice(&P->device);

  auto cleanup = [&](int rc) {
    ax_pool_destroy(P.release());
    return rc;
  };
  hipError_t e;
  size_t fr = 0, tot = 0;
  if ((e = hipMemGetInfo(&fr, &tot)) != hipSuccess) return cleanup(hip_fail(e, "hipMemGetInfo"));
  size_t budget = cfg.hbm_budget ? cfg.hbm_budget : (size_t)(0.8 * (double)fr);
  budget = budget / Arena::kAlign * Arena::kAlign;
  if (budget == 0) return cleanup(fail(AX_EINVAL, "zero budget"));
  if ((e = hipStreamCreateWithFlags(&P->cs, hipStreamNonBlocking)) != hipSuccess) return cleanup(hip_fail(e, "stream"));
  if ((e = hipStreamCreateWithFlags(&P->us, hipStreamNonBlocking)) != hipSuccess) return cleanup(hip_fail(e, "stream"));
  if (hipMalloc((void **)&P->arena_base, budget) != hipSuccess) {
    (void)hipGetLastError();   /* clear the sticky error of the failed call */
    return cleanup(fail(AX_ENOMEM, "pool arena of %zu bytes", budget));
  }
  P->arena.reset(budget);
  P->cfg.hbm_budget = budget;
  const size_t tab = 64 + (size_t)cfg.max_lattices * sizeof(BatchEntry);
  if (hipMalloc((void **)&P->d_hdr, tab) != hipSuccess ||
      hipMalloc((void **)&P->d_cnt, (size_t)cfg.max_lattices * kCnt * sizeof(unsigned long long)) != hipSuccess) {
    (void)hipGetLastError();   /* clear the sticky error of the failed call */
    return cleanup(fail(AX_ENOMEM, "pool tables"));
  }
  P->d_ent = (BatchEntry *)((char *)P->d_hdr + 64);
  for (int b = 0; b < ax_pool::kBufs; ++b)
    if (hipHostMalloc((void **)&P->h_tab[b], tab, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&P->h_cnt[b], (size_t)cfg.max_lattices * kCnt * sizeof(unsigned long long),
                      hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();   /* clear the sticky error of the failed call */
      return cleanup(fail(AX_ENOMEM, "pinned pool tables"));
    }
  /* persistent grids: every workgroup slot the device can hold at once */
  int ncu = 0, bs = 0, bw = 0;
  (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, P->device);   /* 0 on failure: grid of 1 per slot */
  if (P->esize == 8)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&bs, P->pow2 ? (const void *)pool_step_kernel<double, true>
                                                                  : (const void *)pool_step_kernel<double, false>, 64, 0);
  else
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&bs, P->pow2 ? (const void *)pool_step_kernel<float, true>
                                                                  : (const void *)pool_step_kernel<float, false>, 64, 0);
  /* the two winding passes share a grid: the smaller of their occupancies */
  int bp = 0;
  if (e == hipSuccess)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&bw, P->esize == 8 ? (const void *)pool_wind_pass_kernel<double, true>
                                                                         : (const void *)pool_wind_pass_kernel<float, true>, 64, 0);
  if (e == hipSuccess)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&bp, P->esize == 8 ? (const void *)pool_wind_pass_kernel<double, false>
                                                                         : (const void *)pool_wind_pass_kernel<float, false>, 64, 0);
  if (e != hipSuccess) return cleanup(hip_fail(e, "occupancy"));
  /* multiples of the 8 XCDs (the XCD-aware item order) */
  P->grid_step = (std::max(1, ncu) * std::max(1, bs) + 7) / 8 * 8;
  P->grid_wind = (std::max(1, ncu) * std::max(1, std::min(bw, bp)) + 7) / 8 * 8;
  if (cfg.use_graph) {
    int rc = P->capture_graph();
    if (rc) return cleanup(rc);
  }
  P->staging.resize((size_t)cfg.ic_threads + 2);
  P->st.budget = budget;
  ax_pool *raw = P.get();
  P->sched = std::thread([raw] { raw->scheduler(); });
  for (int i = 0; i < cfg.ic_threads; ++i) P->icw.emplace_back([raw] { raw->ic_worker(); });
  *out = P.release();
  return AX_OK;
}

void ax_pool_destroy(ax_pool *P) {
  if (!P) return;
  {
    std::lock_guard<std::mutex> g(P->mu);
    P->stop = true;
  }
  P->cv_sched.notify_all();
  P->cv_ic.notify_all();
  P->cv_done.notify_all();
  if (P->sched.joinable()) P->sched.join();
  for (auto &t : P->icw)
    if (t.joinable()) t.join();
  if (P->cs) (void)hipStreamSynchronize(P->cs);
  if (P->us) (void)hipStreamSynchronize(P->us);
  for (auto &c : P->inflight) {
    if (c.done) (void)hipEventDestroy(c.done);
    if (c.s0) (void)hipEventDestroy(c.s0);
    if (c.s1) (void)hipEventDestroy(c.s1);
  }
  for (auto &kv : P->reqs)
    if (kv.second->upload) (void)hipEventDestroy(kv.second->upload);
  for (auto &sg : P->staging) {
    if (sg.done) (void)hipEventDestroy(sg.done);
    if (sg.host) (void)hipHostFree(sg.host);
  }
  for (int part = 0; part < ax_pool::kParts; ++part) {
    if (P->gexec[part]) (void)hipGraphExecDestroy(P->gexec[part]);
    if (P->graph[part]) (void)hipGraphDestroy(P->graph[part]);
  }
  for (int b = 0; b < ax_pool::kBufs; ++b) {
    if (P->h_tab[b]) (void)hipHostFree(P->h_tab[b]);
    if (P->h_cnt[b]) (void)hipHostFree(P->h_cnt[b]);
  }
  if (P->d_hdr) (void)hipFree(P->d_hdr);
  if (P->d_cnt) (void)hipFree(P->d_cnt);
  if (P->arena_base) (void)hipFree(P->arena_base);
  if (P->cs) (void)hipStreamDestroy(P->cs);
  if (P->us) (void)hipStreamDestroy(P->us);
  delete P;
}

int ax_pool_submit(ax_pool *P, uint32_t seed, int32_t n, int32_t steps, int32_t flags, int64_t *id) {
  if (!P || !id) return fail(AX_EINVAL, "NULL argument");
  if (n < 3 || steps < 1 || (flags & ~AX_POOL_KEEP_STATE)) return fail(AX_EINVAL, "bad request n=%d steps=%d", n, steps);
  if ((double)n * n * n * (double)P->esize >= 4294967296.0) return fail(AX_EINVAL, "n=%d too large for the pool", n);
  const size_t bytes = lattice_bytes(n, P->esize, P->pad_bytes);
  std::lock_guard<std::mutex> g(P->mu);
  if (P->stop) return fail(AX_EHIP, "pool stopped: %s", P->error.c_str());
  if (bytes > P->arena.capacity()) return fail(AX_EBUDGET, "n=%d needs %zu bytes, budget %zu", n, bytes, P->arena.capacity());
  std::unique_ptr<Req> r(new Req);
  r->id = P->next_id++;
  r->seed = seed;
  r->n = n;
  r->steps = steps;
  r->flags = flags;
  r->bytes = bytes;
  r->t_submit = now_s(P->t0);
  *id = r->id;
  P->queue.push_back(r->id);
  P->reqs[r->id] = std::move(r);
  P->st.submitted++;
  P->cv_ic.notify_all();
  P->cv_sched.notify_all();
  return AX_OK;
}

int ax_pool_wait(ax_pool *P, int64_t id, double timeout_s) {
  if (!P) return fail(AX_EINVAL, "NULL pool");
  std::unique_lock<std::mutex> lk(P->mu);
  auto finished = [&] {
    if (id >= 0) {
      auto f = P->reqs.find(id);
      return f == P->reqs.end() || f->second->status >= AX_REQ_DONE;
    }
    for (auto &kv : P->reqs)
      if (kv.second->status < AX_REQ_DONE) return false;
    return true;
  };
  if (id >= 0 && !P->reqs.count(id)) return fail(AX_ENOTFOUND, "request %lld", (long long)id);
  const auto until = std::chrono::steady_clock::now() + std::chrono::duration<double>(timeout_s < 0 ? 1e9 : timeout_s);
  while (!finished()) {
    if (P->cv_done.wait_until(lk, until) == std::cv_status::timeout && !finished()) return 1;
  }
  return AX_OK;
}

int ax_pool_info(ax_pool *P, int64_t id, ax_request_info *out) {
  if (!P || !out) return fail(AX_EINVAL, "NULL argument");
  std::lock_guard<std::mutex> g(P->mu);
  auto f = P->reqs.find(id);
  if (f == P->reqs.end()) return fail(AX_ENOTFOUND, "request %lld", (long long)id);
  const Req &r = *f->second;
  out->id = r.id;
  out->seed = r.seed;
  out->n = r.n;
  out->steps = r.steps;
  out->status = r.status;
  out->nsamples = (int32_t)r.ts.size();
  out->xseg = r.xseg;
  out->t_submit = r.t_submit;
  out->t_admit = r.t_admit;
  out->t_done = r.t_done;
  return AX_OK;
}

int ax_pool_xi(ax_pool *P, int64_t id, int32_t *ts, int64_t *wound, double *xi, size_t cap, size_t *count) {
  if (!P || !count || (cap && (!ts || !wound || !xi))) return fail(AX_EINVAL, "NULL argument");
  std::lock_guard<std::mutex> g(P->mu);
  auto f = P->reqs.find(id);
  if (f == P->reqs.end()) return fail(AX_ENOTFOUND, "request %lld", (long long)id);
  const Req &r = *f->second;
  *count = r.ts.size();
  const size_t m = std::min(cap, r.ts.size());
  for (size_t i = 0; i < m; ++i) {
    ts[i] = r.ts[i];
    wound[i] = r.wound[i];
    xi[i] = r.xi[i];
  }
  return AX_OK;
}

int ax_pool_get_state(ax_pool *P, int64_t id, double *host, size_t n) {
  if (!P || !host) return fail(AX_EINVAL, "NULL argument");
  std::unique_lock<std::mutex> lk(P->mu);
  auto f = P->reqs.find(id);
  if (f == P->reqs.end()) return fail(AX_ENOTFOUND, "request %lld", (long long)id);
  Req &r = *f->second;
  if (r.status != AX_REQ_DONE || r.off < 0) return fail(AX_EINVAL, "request %lld has no kept final state", (long long)id);
  const size_t V = (size_t)r.n * r.n * r.n;
  if (n != 10 * V) return fail(AX_EINVAL, "state length %zu != %zu", n, 10 * V);
  const size_t pitch = lattice_pitch(V, P->esize, P->pad_bytes);
  const char *base = P->arena_base + r.off;
  lk.unlock();   /* the slot stays allocated until ax_pool_release */
  hipStream_t s;
  hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  if (e != hipSuccess) return hip_fail(e, "stream");
  std::vector<float> tmp(P->esize == 4 ? V : 0);
  for (int m = 0; m < 10 && e == hipSuccess; ++m) {
    const char *src = base + (size_t)m * pitch * P->esize;
    if (P->esize == 8) {
      e = hipMemcpyAsync(host + (size_t)m * V, src, V * 8, hipMemcpyDeviceToHost, s);
      if (e == hipSuccess) e = hipStreamSynchronize(s);
    } else {
      e = hipMemcpyAsync(tmp.data(), src, V * 4, hipMemcpyDeviceToHost, s);
      if (e == hipSuccess) e = hipStreamSynchronize(s);
      for (size_t i = 0; i < V && e == hipSuccess; ++i) host[(size_t)m * V + i] = tmp[i];
    }
  }
  (void)hipStreamDestroy(s);
  return e == hipSuccess ? AX_OK : hip_fail(e, "pool get_state");
}

int ax_pool_release(ax_pool *P, int64_t id) {
  if (!P) return fail(AX_EINVAL, "NULL pool");
  std::lock_guard<std::mutex> g(P->mu);
  auto f = P->reqs.find(id);
  if (f == P->reqs.end()) return fail(AX_ENOTFOUND, "request %lld", (long long)id);
  Req &r = *f->second;
  if (r.status < AX_REQ_DONE) return fail(AX_EINVAL, "request %lld is not finished", (long long)id);
  P->free_slot(r);
  P->reqs.erase(f);
  P->cv_sched.notify_all();
  return AX_OK;
}

int ax_pool_stats_get(ax_pool *P, ax_pool_stats *out) {
  if (!P || !out) return fail(AX_EINVAL, "NULL argument");
  std::lock_guard<std::mutex> g(P->mu);
  *out = P->st;
  out->resident = (int32_t)P->resident.size();
  out->queued = (int32_t)P->queue.size();
  out->elapsed = now_s(P->t0);
  out->in_use = P->arena.in_use();
  out->peak_in_use = P->arena.peak();
  return AX_OK;
}

int ax_pool_tune(ax_pool *P, int32_t n, int32_t *xseg, double *ms) {
  if (!P || !xseg || !ms || n < 3) return fail(AX_EINVAL, "bad tune request");
  if ((double)n * n * n * (double)P->esize >= 4294967296.0) return fail(AX_EINVAL, "n=%d too large", n);
  int xs = 0;
  double t = 0;
  int rc = P->tune(n, &xs, &t);
  if (rc) return rc;
  *xseg = xs;
  *ms = t;
  return AX_OK;
}

}  // extern "C"
