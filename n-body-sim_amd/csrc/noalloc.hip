// The allocation-free force evaluation: nbody_calculate_accelerations on
// caller-owned device arrays, with every piece of library memory it needs
// (pointer table, system descriptor, work lists, g*m, the padded position
// snapshot and the split partial sums) carved from one file-scope device
// array.  For callers that cannot allocate device memory: code under HIP
// graph capture, tight serving loops, harnesses that forbid hipMalloc.
//
// The array's device address is looked up once per device with
// hipGetSymbolAddress (at context creation, with the device current) and
// kept per device; host code never passes the array's name to a kernel.  The
// descriptors reach the array through a setup kernel that takes them by
// value, so a call queues only kernel launches: no allocation, no memset, no
// host copy.  Every call on one device shares that device's array, so calls
// that overlap on one device (two streams) race.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <string>

#include "../../include/nbody.h"
#include "engine.h"
#include "kernels.h"
#include "noalloc.h"
#include "nbody_types.h"
#include "tuning.h"

namespace nbody {

namespace {

// The scratch array.  Zero at load; each call overwrites what it reads
// (partials of every part of every target, the snapshot and g*m of every
// body), so nothing has to be cleared between calls.
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};  // scratch's address on each device

struct Layout {  // byte offsets into the scratch array
  size_t arena, sys, spans, items, gm64, gm32, sx, sy, sz, pp, end;
  int64_t len;   // elements per per-body array (n + kPad, rounded)
};

size_t up256(size_t v) { return (v + 255) / 256 * 256; }

Layout layout(int64_t n, int nsplit) {
  Layout L{};
  const int64_t tiles = (n + kWave - 1) / kWave;
  L.len = (int64_t)up256((size_t)(n + kPad) * sizeof(double)) / (int64_t)sizeof(double);
  size_t o = 0;
  auto take = [&](size_t bytes) {
    const size_t at = o;
    o += up256(bytes);
    return at;
  };
  L.arena = take(sizeof(Arena));
  L.sys = take(sizeof(SysDesc));
  L.spans = take(3 * sizeof(ItemSpan));
  L.items = take((size_t)tiles * (nsplit + 2) * sizeof(WorkItem));
  L.gm64 = take((size_t)L.len * sizeof(double));
  L.gm32 = take((size_t)L.len * sizeof(float));
  L.sx = take((size_t)L.len * sizeof(double));
  L.sy = take((size_t)L.len * sizeof(double));
  L.sz = take((size_t)L.len * sizeof(double));
  L.pp = take((size_t)nsplit * L.len * sizeof(Part3));
  L.end = o;
  return L;
}

// Writes the call's descriptors into the scratch array (one thread): the
// pointer table, the system descriptor and the spans of the work lists.
struct Spans {
  ItemSpan s[3];
};

__global__ void k_noalloc_setup(Arena a, SysDesc d, Spans sp, Arena* a_out,
                                SysDesc* d_out, ItemSpan* spans_out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  *a_out = a;
  *d_out = d;
  for (int k = 0; k < 3; ++k) spans_out[k] = sp.s[k];
}

}  // namespace

void noalloc_init(int device) {
  if (device < 0 || device >= kMaxDevices)
    throw Error(NBODY_E_DEVICE, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_ptr[device]) return;
  void* p = nullptr;
  hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)),
            "hipGetSymbolAddress(scratch)");
  if (!p) throw Error(NBODY_E_DEVICE, "scratch array has no device address");
  g_ptr[device] = p;
}

void noalloc_accelerations(int device, hipStream_t stream, int mode,
                           int64_t n, const double* m, const double* x,
                           const double* y, const double* z, double g,
                           double eps2, double* ax, double* ay, double* az) {
  if (n <= 0) throw Error(NBODY_E_INVALID, "n must be >= 1");
  if (n > kNoallocMaxBodies)
    throw Error(NBODY_E_INVALID,
                "n exceeds the allocation-free maximum of " +
                    std::to_string(kNoallocMaxBodies) + " bodies");
  if (mode != MODE_F64_STRICT && mode != MODE_F64 && mode != MODE_F32C)
    throw Error(NBODY_E_UNSUPPORTED,
                "the allocation-free evaluation offers NBODY_PRECISION_F64_STRICT, "
                "F64 and F32C (F32R orders bodies with a sort that needs "
                "temporary memory)");
  if (!m || !x || !y || !z || !ax || !ay || !az)
    throw Error(NBODY_E_INVALID, "null device array");
  if (!std::isfinite(g) || !std::isfinite(eps2) || eps2 < 0.0)
    throw Error(NBODY_E_INVALID, "g and eps2 must be finite, eps2 >= 0");
  void* base;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    base = (device >= 0 && device < kMaxDevices) ? g_ptr[device] : nullptr;
  }
  if (!base) throw Error(NBODY_E_STATE, "scratch not set up on this device");
  // the same split (and below, kernel variant) as the single-system path,
  // so the sums are the same bits
  const int nsplit = mode == MODE_F64_STRICT ? 1 : split_for_size(n);
  const Layout L = layout(n, nsplit);
  if (L.end > kNoallocScratchBytes)
    throw Error(NBODY_E_INVALID, "allocation-free scratch too small for n");
  char* s = (char*)base;
  Arena a{};
  a.m = const_cast<double*>(m);
  a.x = const_cast<double*>(x);
  a.y = const_cast<double*>(y);
  a.z = const_cast<double*>(z);
  a.gm64 = (double*)(s + L.gm64);
  a.gm32 = (float*)(s + L.gm32);
  a.sx[0] = a.sx[1] = (double*)(s + L.sx);
  a.sy[0] = a.sy[1] = (double*)(s + L.sy);
  a.sz[0] = a.sz[1] = (double*)(s + L.sz);
  a.ax = ax;
  a.ay = ay;
  a.az = az;
  a.pp = (Part3*)(s + L.pp);
  a.pstride = L.len;
  a.cap = n;
  SysDesc d{};
  d.off = 0;
  d.n = (int32_t)n;
  d.integ = INTEG_EULER;
  d.dt = 0.0;
  d.eps2 = eps2;
  d.g = g;
  d.eps2f = (float)eps2;
  d.nsplit = (int16_t)nsplit;
  d.flags = 0;
  // work items: [prepare: one per tile | force: tile x part | combine: one
  // per tile, when split], expanded on the device from three spans
  const int32_t tiles = (int32_t)((n + kWave - 1) / kWave);
  const int nforce = tiles * nsplit;
  const int ncombine = nsplit > 1 ? tiles : 0;
  Spans sp{};
  sp.s[0] = ItemSpan{0, 0, tiles, (int16_t)STAGE_ACC, 0, 1};
  sp.s[1] = ItemSpan{tiles, 0, tiles, (int16_t)STAGE_ACC, 0, (int8_t)nsplit};
  sp.s[2] = ItemSpan{tiles + nforce, 0, tiles, (int16_t)STAGE_ACC, 0, 1};
  Arena* d_arena = (Arena*)(s + L.arena);
  SysDesc* d_sys = (SysDesc*)(s + L.sys);
  ItemSpan* d_spans = (ItemSpan*)(s + L.spans);
  WorkItem* d_items = (WorkItem*)(s + L.items);
  hipLaunchKernelGGL(k_noalloc_setup, dim3(1), dim3(64), 0, stream, a, d, sp,
                     d_arena, d_sys, d_spans);
  hip_check(hipGetLastError(), "setup launch");
  hip_check(launch_expand(d_spans, ncombine ? 3 : 2, tiles + nforce + ncombine,
                          d_items, stream),
            "expand launch");
  hip_check(launch_prepare(d_arena, d_items, tiles, d_sys, stream),
            "prepare launch");
  // the variant the single-system path gives a system of n bodies
  hip_check(launch_force(mode, tuning_variant(mode, n), d_arena, d_items + tiles,
                         nforce, d_sys, stream),
            "force launch");
  hip_check(launch_combine(mode, d_arena, d_items + tiles + nforce, ncombine,
                           d_sys, stream),
            "combine launch");
}

}  // namespace nbody
