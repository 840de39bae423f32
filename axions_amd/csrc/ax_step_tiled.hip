/*
 * ax_step_tiled.hip - tiled Abelian-Higgs FP64 step (filled in below).
 */
#include "ax_internal.h"

namespace ax {

bool use_tiled_ah(const ax_params &p) { (void)p; return false; }

hipError_t launch_step_tiled_ah(const StepConsts &k, const Levels<double> &L, hipStream_t st) {
  (void)k; (void)L; (void)st;
  return hipErrorNotSupported;
}

}  // namespace ax
