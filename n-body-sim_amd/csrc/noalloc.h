// The allocation-free force evaluation (noalloc.hip).
#pragma once
#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>

namespace nbody {

// Size of the per-device scratch array, and the largest system it serves
// (partial sums of up to 32 source parts per body dominate: about 800 B per
// body, so 65536 bodies take about 53 MB of the 64 MiB).
constexpr size_t kNoallocScratchBytes = (size_t)64 << 20;
constexpr int64_t kNoallocMaxBodies = 65536;

// Looks up the scratch array's address on `device` (current) once; later
// calls for the same device return at once.  Throws Error on failure.
void noalloc_init(int device);

// Queues the force evaluation of n bodies on `stream` (device `device`
// current): m, x, y, z in, ax, ay, az out, all caller-owned device arrays of
// n doubles.  mode: MODE_F64_STRICT, MODE_F64 or MODE_F32C.  The results
// are the bits nbody_calculate_accelerations gives for the same bodies.
// Allocates nothing.  Throws Error on invalid arguments.
void noalloc_accelerations(int device, hipStream_t stream, int mode,
                           int64_t n, const double* m, const double* x,
                           const double* y, const double* z, double g,
                           double eps2, double* ax, double* ay, double* az);

}  // namespace nbody
