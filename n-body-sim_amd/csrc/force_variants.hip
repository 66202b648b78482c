// The tuned force-kernel variants of one arithmetic mode, selected at build
// time by NBODY_FORCE_MODE (1: MODE_F64, 2: MODE_F32C, 3: MODE_F32R); the
// Makefile builds this file once per mode so the 27 kernels compile in
// parallel.  Variant v = 3 * w + shape: w picks 1, 2 or 4 waves per
// workgroup (one work item per wave), shape the loop shape of force.h.
#include <hip/hip_runtime.h>

#include "force.h"
#include "kernels.h"

#ifndef NBODY_FORCE_MODE
#error "NBODY_FORCE_MODE must be 1, 2 or 3"
#endif

namespace nbody {
namespace {

constexpr int kMode = NBODY_FORCE_MODE;

template <int W, int S>
hipError_t launch_one(const Arena* A, const WorkItem* items, int nitems,
                      const SysDesc* sys, hipStream_t s) {
  const unsigned blocks = (unsigned)((nitems + W - 1) / W);
  hipLaunchKernelGGL((k_force<kMode, false, W, S>), dim3(blocks),
                     dim3(kWave * W), 0, s, A, items, nitems, sys);
  return hipGetLastError();
}

hipError_t launch_variant(int variant, const Arena* A, const WorkItem* items,
                          int nitems, const SysDesc* sys, hipStream_t s) {
  if (nitems <= 0) return hipSuccess;
  switch (variant) {
    case 0: return launch_one<1, 0>(A, items, nitems, sys, s);
    case 1: return launch_one<1, 1>(A, items, nitems, sys, s);
    case 2: return launch_one<1, 2>(A, items, nitems, sys, s);
    case 3: return launch_one<2, 0>(A, items, nitems, sys, s);
    case 4: return launch_one<2, 1>(A, items, nitems, sys, s);
    case 5: return launch_one<2, 2>(A, items, nitems, sys, s);
    case 6: return launch_one<4, 0>(A, items, nitems, sys, s);
    case 7: return launch_one<4, 1>(A, items, nitems, sys, s);
    case 8: return launch_one<4, 2>(A, items, nitems, sys, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

#if NBODY_FORCE_MODE == 1
hipError_t launch_force_f64(int variant, const Arena* A, const WorkItem* items,
                            int nitems, const SysDesc* sys, hipStream_t s) {
  return launch_variant(variant, A, items, nitems, sys, s);
}
#elif NBODY_FORCE_MODE == 2
hipError_t launch_force_f32c(int variant, const Arena* A, const WorkItem* items,
                             int nitems, const SysDesc* sys, hipStream_t s) {
  return launch_variant(variant, A, items, nitems, sys, s);
}
#elif NBODY_FORCE_MODE == 3
hipError_t launch_force_f32r(int variant, const Arena* A, const WorkItem* items,
                             int nitems, const SysDesc* sys, hipStream_t s) {
  return launch_variant(variant, A, items, nitems, sys, s);
}
#else
#error "NBODY_FORCE_MODE must be 1, 2 or 3"
#endif

}  // namespace nbody
