/*
 * ref_shim.h - force-included (-include) into the reference's translation
 * units by oracle/ref/build_ref.py.  TEST INFRASTRUCTURE ONLY.
 *
 * It changes no arithmetic of the reference.  It does two things:
 *  1. every std::ofstream the reference opens prints with 17 significant
 *     digits instead of 6, so its Data/ and GifData/ files round-trip
 *     doubles exactly;
 *  2. with -DAX_REF_SHARED_TRIG, the reference's sin()/cos() calls are
 *     routed to the ax_sin()/ax_cos() of include/axions/ax_trig.h, the
 *     functions the HIP kernels use, so a gauged (g != 0) run can be
 *     compared bit for bit over its whole length.  Without the define the
 *     reference uses glibc's libm as built.
 * All standard headers the reference includes are included first, so the
 * macros below only reach the reference's own code.
 */
#include <algorithm>
#include <cmath>
#include <complex>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <math.h>
#include <omp.h>
#include <random>
#include <sstream>
#include <string>
#include <sys/time.h>
#include <utility>
#include <vector>

#include "axions/ax_trig.h"

struct ax_ref_ofstream : public std::ofstream {
  template <class... A>
  explicit ax_ref_ofstream(A &&...a) : std::ofstream(std::forward<A>(a)...) {
    this->precision(17);
  }
};
#define ofstream ax_ref_ofstream

#ifdef AX_REF_SHARED_TRIG
#define sin(x) ax_sin(x)
#define cos(x) ax_cos(x)
#endif
