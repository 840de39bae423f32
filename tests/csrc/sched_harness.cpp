// Test harness: the pool's host bookkeeping (axions_amd/csrc/ax_sched.h)
// behind a C interface for ctypes, built by tests/test_sched_host.py with g++.
#include "../../axions_amd/csrc/ax_sched.h"

extern "C" {
void *arena_new(size_t cap) { return new ax::Arena(cap); }
void arena_delete(void *a) { delete (ax::Arena *)a; }
long long arena_alloc(void *a, size_t b) { return ((ax::Arena *)a)->alloc(b); }
int arena_release(void *a, size_t off) { return ((ax::Arena *)a)->release(off) ? 1 : 0; }
size_t arena_capacity(void *a) { return ((ax::Arena *)a)->capacity(); }
size_t arena_in_use(void *a) { return ((ax::Arena *)a)->in_use(); }
size_t arena_peak(void *a) { return ((ax::Arena *)a)->peak(); }
size_t arena_largest_free(void *a) { return ((ax::Arena *)a)->largest_free(); }
size_t arena_free_ranges(void *a) { return ((ax::Arena *)a)->free_ranges(); }
size_t lattice_bytes(int n, size_t esize, size_t pad) { return ax::lattice_bytes(n, esize, pad); }
double xi_of(double a, double b, double c, double V, double dx, double dy, double dz) {
  return ax::xi_of(a, b, c, V, dx, dy, dz);
}
long long fill_segments(int count, const int *n, const int *nw, int *xseg, long long want, int floor) {
  return ax::fill_segments(count, n, nw, xseg, want, floor);
}
int xcd_item(int w, int q) { return ax::xcd_item(w, q); }
}
