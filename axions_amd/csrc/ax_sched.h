/*
 * ax_sched.h - host-side bookkeeping of the ensemble pool, free of HIP so it
 * can be unit-tested on a machine without a GPU.
 *
 *  Arena       first-fit sub-allocator over one device allocation of the
 *              pool's HBM budget; blocks are 256-byte aligned and freed
 *              blocks coalesce with their neighbours.
 *  lattice_bytes  device bytes one request of edge N occupies (state and
 *                 the edge phases of its winding pass).
 *  xi_of       string-length density from the three plaquette-plane counts,
 *              the expression ax_diagnostics uses.
 *  fill_segments  the chunk fill: shortens the tuned x segments of a chunk
 *              that holds too few work items for the persistent grid.
 *  xcd_item    the XCD-aware order of the batched step kernel's work items.
 */
#ifndef AXIONS_AX_SCHED_H
#define AXIONS_AX_SCHED_H

#include <stddef.h>
#include <stdint.h>

#include <iterator>
#include <map>

namespace ax {

class Arena {
 public:
  static constexpr size_t kAlign = 256;

  explicit Arena(size_t capacity = 0) { reset(capacity); }

  void reset(size_t capacity) {
    cap_ = capacity / kAlign * kAlign;
    free_.clear();
    used_.clear();
    if (cap_) free_[0] = cap_;
    in_use_ = peak_ = 0;
  }

  /* offset of a block of at least `bytes`, or -1 when no free range fits */
  int64_t alloc(size_t bytes) {
    if (bytes == 0) bytes = 1;
    const size_t need = (bytes + kAlign - 1) / kAlign * kAlign;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < need) continue;
      const size_t off = it->first, len = it->second;
      free_.erase(it);
      if (len > need) free_[off + need] = len - need;
      used_[off] = need;
      in_use_ += need;
      if (in_use_ > peak_) peak_ = in_use_;
      return (int64_t)off;
    }
    return -1;
  }

  /* returns false for an offset that is not an allocated block */
  bool release(size_t off) {
    auto u = used_.find(off);
    if (u == used_.end()) return false;
    size_t len = u->second;
    used_.erase(u);
    in_use_ -= len;
    auto nx = free_.lower_bound(off);
    if (nx != free_.end() && off + len == nx->first) {   /* merge with the next free range */
      len += nx->second;
      nx = free_.erase(nx);
    }
    if (nx != free_.begin()) {                            /* merge with the previous one */
      auto pv = std::prev(nx);
      if (pv->first + pv->second == off) {
        pv->second += len;
        return true;
      }
    }
    free_[off] = len;
    return true;
  }

  size_t capacity() const { return cap_; }
  size_t in_use() const { return in_use_; }
  size_t peak() const { return peak_; }
  size_t largest_free() const {
    size_t m = 0;
    for (auto &f : free_) m = f.second > m ? f.second : m;
    return m;
  }
  size_t free_ranges() const { return free_.size(); }
  size_t blocks() const { return used_.size(); }

 private:
  size_t cap_ = 0, in_use_ = 0, peak_ = 0;
  std::map<size_t, size_t> free_;   /* offset -> length, disjoint, never adjacent */
  std::map<size_t, size_t> used_;   /* offset -> length */
};

/* Arena slot of one n^3 Abelian-Higgs lattice: its state (5 components x 2
 * levels of planes, each plane padded by pad_bytes rounded to the element
 * size), then, 256-byte aligned, the edge values of the winding pass
 * (3 x n^3 floats); the whole block rounded to the arena alignment. */
inline size_t lattice_pitch(size_t sites, size_t esize, size_t pad_bytes) {
  return sites + (pad_bytes + esize - 1) / esize;
}
inline size_t lattice_phase_offset(int n, size_t esize, size_t pad_bytes) {
  const size_t V = (size_t)n * n * n;
  const size_t b = 10 * lattice_pitch(V, esize, pad_bytes) * esize;
  return (b + 255) / 256 * 256;
}
inline size_t lattice_bytes(int n, size_t esize, size_t pad_bytes) {
  const size_t V = (size_t)n * n * n;
  const size_t b = lattice_phase_offset(n, esize, pad_bytes) + 3 * V * sizeof(float);
  return (b + Arena::kAlign - 1) / Arena::kAlign * Arena::kAlign;
}

/* xi = L / V_phys with L = (2/3)(dz W_xy + dy W_xz + dx W_yz), the same
 * expression and operand order as ax_diagnostics. */
inline double xi_of(double wxy, double wxz, double wyz, double sites, double dx, double dy, double dz) {
  const double len = (2.0 / 3.0) * (wxy * dz + wxz * dy + wyz * dx);
  return len / (sites * dx * dy * dz);
}

/* The chunk fill.  The tuned segments are chosen for a full pool; with few
 * lattices resident they leave workgroups of the persistent grid idle.  For
 * count lattices of edge n[k] with nw[k] z chunks of 64 and segments of
 * xseg[k] planes, halve the longest segments (never below floor planes)
 * until the chunk has at least want work items or every segment is at
 * floor; xseg is updated in place and the chunk's work items (nw n nseg
 * summed) returned.  The partition of a lattice does not change its
 * results. */
inline long long fill_segments(int count, const int *n, const int *nw, int *xseg, long long want, int floor = 4) {
  auto items = [&]() {
    long long w = 0;
    for (int k = 0; k < count; ++k) w += (long long)nw[k] * n[k] * ((n[k] + xseg[k] - 1) / xseg[k]);
    return w;
  };
  long long work = items();
  while (work < want) {
    int longest = 0;
    for (int k = 0; k < count; ++k) longest = xseg[k] > longest ? xseg[k] : longest;
    if (longest <= floor) break;
    const int half = longest / 2 > floor ? longest / 2 : floor;
    for (int k = 0; k < count; ++k)
      if (xseg[k] == longest) xseg[k] = half;
    work = items();
  }
  return work;
}

/* The XCD-aware order of a grid of 8q workgroups (a multiple of the eight
 * XCDs, dispatched to them round-robin): the w-th of the grid's items is
 * item (w mod 8) q + w / 8, so XCD k takes the k-th eighth of the items.  A
 * bijection of [0, 8q). */
#if defined(__HIPCC__) || defined(__HIP__)
__host__ __device__
#endif
inline int xcd_item(int w, int q) { return (w & 7) * q + (w >> 3); }

}  // namespace ax

#endif
