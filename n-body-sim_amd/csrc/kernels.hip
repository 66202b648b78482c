// gfx950 kernels for the direct-sum gravity step: the launchers, the combine
// and stage kernels, snapshot preparation, diagnostics, gathers and the
// F32R body ordering.  The force kernel itself is in force.h.
//
// The integrator stage is fused into the epilogue (of the force launch, or
// of k_combine when split): it updates the target's master state and writes
// its next source snapshot into the other half of a double buffer, so no
// target of a running launch ever reads a value that launch writes.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <hipcub/hipcub.hpp>
#include <type_traits>

#include "device.h"
#include "force.h"  // the force kernel k_force and its pair loops
#include "kernels.h"
#include "nbody_types.h"

namespace nbody {
namespace {

// Adds the parts' partial sums of each target in part order (a fixed order,
// so results are reproducible) and runs the target's stage epilogue.
template <int MODE>
__global__ __launch_bounds__(kBlock) void k_combine(const Arena* Ap,
                                                    const WorkItem* items,
                                                    int nitems,
                                                    const SysDesc* sysd) {
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int idx = (int)blockIdx.x * kWavesPerBlock + wave;
  if (idx >= nitems) return;
  const WorkItem it = load_uniform(items + idx);
  const CArena* A = (const CArena*)Ap;
  const CSys* sd = (const CSys*)(sysd + it.sys);
  const int64_t end = sd->off + sd->n;
  // the tile's target bodies (in body order in every mode; in MODE_F32R the
  // force items' tiles are slot ranges, but the partials are in body order)
  const int64_t k = sd->off + (int64_t)it.tile * kWave +
                    (int64_t)(threadIdx.x & (kWave - 1));
  if (k >= end) return;
  const int64_t st = A->pstride;
  const Part3* pp = A->pp;
  double ax = pp[k].x;
  if (it.stage == STAGE_POT) {
    for (int p = 1; p < sd->nsplit; ++p) ax = ax + pp[p * st + k].x;
    A->phi[k] = ax;
    return;
  }
  double ay = pp[k].y, az = pp[k].z;
  for (int p = 1; p < sd->nsplit; ++p) {
    const Part3 q = pp[p * st + k];
    ax = ax + q.x;
    ay = ay + q.y;
    az = az + q.z;
  }
  // the body's slot in Hilbert order, for MODE_F32R's slot-ordered snapshot
  const int64_t slot = MODE == MODE_F32R ? sd->off + A->iperm[k] : k;
  epilogue<MODE>(A, sd, it.stage, it.src ^ 1, k, slot, ax, ay, az);
}

// Builds source snapshot it.src of each body from the master state:
// stage 0 -> the positions themselves; STAGE_VV_PRIME -> the speculative
// next positions x + ((v*dt) + ((0.5*a)*dt^2)) from the carried a.  Also
// refreshes gm = g*m in f64 and f32.
__global__ __launch_bounds__(kBlock) void k_prepare(const Arena* Ap,
                                                    const WorkItem* items,
                                                    int nitems,
                                                    const SysDesc* sysd) {
#pragma clang fp contract(off)
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int idx = (int)blockIdx.x * kWavesPerBlock + wave;
  if (idx >= nitems) return;
  const WorkItem it = load_uniform(items + idx);
  const CArena* A = (const CArena*)Ap;
  const CSys* sd = (const CSys*)(sysd + it.sys);
  const int64_t end = sd->off + sd->n;
  const int64_t i = sd->off + (int64_t)it.tile * kWave +
                    (int64_t)(threadIdx.x & (kWave - 1));
  if (i >= end) return;
  const double gm = sd->g * A->m[i];
  A->gm64[i] = gm;
  A->gm32[i] = (float)gm;
  double x = A->x[i], y = A->y[i], z = A->z[i];
  if (it.stage == STAGE_VV_PRIME) {
    const double dt = sd->dt, dt2 = dt * dt;
    x = x + ((A->vx[i] * dt) + ((0.5 * A->ax[i]) * dt2));
    y = y + ((A->vy[i] * dt) + ((0.5 * A->ay[i]) * dt2));
    z = z + ((A->vz[i] * dt) + ((0.5 * A->az[i]) * dt2));
  }
  const int d = it.src;
  A->sx[d][i] = x;
  A->sy[d][i] = y;
  A->sz[d][i] = z;
}

// One integrator stage on the acceleration arrays (StageOp), with the
// reference's expression trees (src/integrators/cpu/euler.rs:69-80,
// velocity_verlet.rs:82-101, runge_kutta.rs:149-281) written exactly as the
// fused epilogue writes them, so a step run stage by stage around separate
// force evaluations gives the fused step's bits.  Staged RK4 moves the
// master positions to each stage's evaluation point and keeps the step's
// start in xs.
__global__ __launch_bounds__(kBlock) void k_stage(const Arena* Ap,
                                                  const WorkItem* items,
                                                  int nitems,
                                                  const SysDesc* sysd,
                                                  int op) {
#pragma clang fp contract(off)
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int idx = (int)blockIdx.x * kWavesPerBlock + wave;
  if (idx >= nitems) return;
  const WorkItem it = load_uniform(items + idx);
  const CArena* A = (const CArena*)Ap;
  const CSys* sd = (const CSys*)(sysd + it.sys);
  const int64_t end = sd->off + sd->n;
  const int64_t i = sd->off + (int64_t)it.tile * kWave +
                    (int64_t)(threadIdx.x & (kWave - 1));
  if (i >= end) return;
  const double dt = sd->dt;
  const double ax = A->ax[i], ay = A->ay[i], az = A->az[i];
  double* X[3] = {A->x, A->y, A->z};
  double* V[3] = {A->vx, A->vy, A->vz};
  const double a[3] = {ax, ay, az};
  switch (op) {
    case OP_EULER:
      for (int c = 0; c < 3; ++c) {
        const double v = V[c][i] + a[c] * dt;
        V[c][i] = v;
        X[c][i] = X[c][i] + v * dt;
      }
      break;
    case OP_VV_DRIFT: {
      const double dt2 = dt * dt;
      for (int c = 0; c < 3; ++c) {
        A->rk[c][i] = a[c];
        X[c][i] = X[c][i] + ((V[c][i] * dt) + ((0.5 * a[c]) * dt2));
      }
      break;
    }
    case OP_VV_KICK:
      for (int c = 0; c < 3; ++c)
        V[c][i] = V[c][i] + ((0.5 * (A->rk[c][i] + a[c])) * dt);
      break;
    case OP_RK4_1: {
      const double h = dt / 2.0;
      for (int c = 0; c < 3; ++c) {
        const double v = V[c][i], x0 = X[c][i];
        const double k2 = v + h * a[c];
        A->xs[c][i] = x0;
        A->rk[c][i] = v + (2.0 * k2);
        A->rk[3 + c][i] = a[c];
        A->rk[6 + c][i] = k2;
        X[c][i] = x0 + h * v;
      }
      break;
    }
    case OP_RK4_2: {
      const double h = dt / 2.0;
      for (int c = 0; c < 3; ++c) {
        const double k3 = V[c][i] + h * a[c];
        const double k2 = A->rk[6 + c][i];
        A->rk[c][i] = A->rk[c][i] + (2.0 * k3);
        A->rk[3 + c][i] = A->rk[3 + c][i] + (2.0 * a[c]);
        A->rk[6 + c][i] = k3;
        X[c][i] = A->xs[c][i] + h * k2;
      }
      break;
    }
    case OP_RK4_3:
      for (int c = 0; c < 3; ++c) {
        const double k3 = A->rk[6 + c][i];
        A->rk[c][i] = A->rk[c][i] + (V[c][i] + dt * a[c]);
        A->rk[3 + c][i] = A->rk[3 + c][i] + (2.0 * a[c]);
        X[c][i] = A->xs[c][i] + dt * k3;
      }
      break;
    case OP_RK4_4: {
      const double s6 = dt / 6.0;
      for (int c = 0; c < 3; ++c) {
        X[c][i] = A->xs[c][i] + s6 * A->rk[c][i];
        V[c][i] = V[c][i] + s6 * (A->rk[3 + c][i] + a[c]);
      }
      break;
    }
    default:
      break;
  }
}

// Per-system f64 diagnostics, one workgroup per listed system, fixed
// summation order (thread t takes bodies t, t+256, ...; then a fixed LDS
// tree), so results are reproducible bit for bit.  Needs phi from a
// potential pass.  out[s*9 + DIAG_*].
// One workgroup of kDiagBlock threads per listed system; thread t sums the
// bodies t, t + kDiagBlock, ..., then each wave reduces by a fixed
// butterfly and thread q adds the waves' values of quantity q in wave order:
// a fixed order, so the result is reproducible.
constexpr int kDiagBlock = 1024;
__global__ __launch_bounds__(kDiagBlock) void k_diag(const Arena* Ap,
                                                     const SysDesc* sysd,
                                                     const int32_t* list,
                                                     double* out) {
  constexpr int kWaves = kDiagBlock / kWave;
  __shared__ double red[DIAG_COUNT][kWaves];
  const int s = list[blockIdx.x];
  const CArena* A = (const CArena*)Ap;
  const CSys* sd = (const CSys*)(sysd + s);
  double acc[DIAG_COUNT];
#pragma unroll
  for (int q = 0; q < DIAG_COUNT; ++q) acc[q] = 0.0;
  for (int64_t k = threadIdx.x; k < sd->n; k += kDiagBlock) {
    const int64_t i = sd->off + k;
    const double m = A->m[i];
    const double vx = A->vx[i], vy = A->vy[i], vz = A->vz[i];
    acc[DIAG_K] += 0.5 * m * ((vx * vx + vy * vy) + vz * vz);
    acc[DIAG_U] += m * A->phi[i];
    acc[DIAG_PX] += m * vx;
    acc[DIAG_PY] += m * vy;
    acc[DIAG_PZ] += m * vz;
    acc[DIAG_MX] += m * A->x[i];
    acc[DIAG_MY] += m * A->y[i];
    acc[DIAG_MZ] += m * A->z[i];
    acc[DIAG_M] += m;
  }
#pragma unroll
  for (int q = 0; q < DIAG_COUNT; ++q)
    for (int d = kWave / 2; d > 0; d >>= 1) acc[q] += __shfl_xor(acc[q], d, kWave);
  const int wave = (int)(threadIdx.x / kWave);
  if ((threadIdx.x & (kWave - 1)) == 0) {
#pragma unroll
    for (int q = 0; q < DIAG_COUNT; ++q) red[q][wave] = acc[q];
  }
  __syncthreads();
  if (threadIdx.x < DIAG_COUNT) {
    double v = red[threadIdx.x][0];
    for (int w = 1; w < kWaves; ++w) v += red[threadIdx.x][w];
    if (threadIdx.x == DIAG_U) v = -0.5 * v;  // each pair counted twice
    out[(int64_t)blockIdx.x * DIAG_COUNT + threadIdx.x] = v;
  }
}

// Packs state rows of listed systems into a staging buffer: item.pad_ is
// the system's base row in the staging layout [field][total]; fields 0..2
// are x, y, z and 3..5 (when nfields == 6) vx, vy, vz.
__global__ __launch_bounds__(kBlock) void k_gather(const Arena* Ap,
                                                   const WorkItem* items,
                                                   int nitems,
                                                   const SysDesc* sysd,
                                                   double* stage,
                                                   int64_t total,
                                                   int nfields) {
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int idx = (int)blockIdx.x * kWavesPerBlock + wave;
  if (idx >= nitems) return;
  const WorkItem it = load_uniform(items + idx);
  const CArena* A = (const CArena*)Ap;
  const CSys* sd = (const CSys*)(sysd + it.sys);
  const int64_t k = (int64_t)it.tile * kWave + (threadIdx.x & (kWave - 1));
  if (k >= sd->n) return;
  const int64_t i = sd->off + k, r = (int64_t)it.pad_ + k;
  stage[r] = A->x[i];
  stage[total + r] = A->y[i];
  stage[2 * total + r] = A->z[i];
  if (nfields == 6) {
    stage[3 * total + r] = A->vx[i];
    stage[4 * total + r] = A->vy[i];
    stage[5 * total + r] = A->vz[i];
  }
}

// Writes zeros; used instead of hipMemsetAsync so it can sit in a graph.
// Expands item spans into work items: thread i finds the span holding item
// i (the spans cover [0, total) in order) by binary search.
__global__ __launch_bounds__(256) void k_expand(const ItemSpan* spans,
                                                int nspans, int total,
                                                WorkItem* out) {
  const int i = (int)blockIdx.x * 256 + (int)threadIdx.x;
  if (i >= total) return;
  int lo = 0, hi = nspans - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (spans[mid].first <= i)
      lo = mid;
    else
      hi = mid - 1;
  }
  const ItemSpan sp = spans[lo];
  const int k = i - sp.first;
  WorkItem w;
  w.sys = sp.sys;
  w.tile = k / sp.parts;
  w.stage = sp.stage;
  w.src = sp.src;
  w.part = (int8_t)(k % sp.parts);
  w.pad_ = 0;
  out[i] = w;
}

// Zeroes bodies [off, off + n) of a group of arena arrays in one launch:
// CLEAR_RETIRED the positions, both snapshots and g*m (a freed range, so
// stale values can never be read as sources), CLEAR_ACC the acceleration
// arrays (a new system).
__global__ __launch_bounds__(256) void k_clear_range(const Arena* Ap,
                                                     int64_t off, int64_t n,
                                                     int what) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const CArena* A = (const CArena*)Ap;
  const int64_t k = off + i;
  if (what == CLEAR_ACC) {
    A->ax[k] = 0.0;
    A->ay[k] = 0.0;
    A->az[k] = 0.0;
    return;
  }
  A->x[k] = 0.0;
  A->y[k] = 0.0;
  A->z[k] = 0.0;
  for (int b = 0; b < 2; ++b) {
    A->sx[b][k] = 0.0;
    A->sy[b][k] = 0.0;
    A->sz[b][k] = 0.0;
  }
  A->gm64[k] = 0.0;
  A->gm32[k] = 0.f;
}

// Device-side probe run once at context creation (the reference runs a
// gpu_init_check kernel, cuda/kernels.cu:4-9): writes a known value.
__global__ void k_probe(int* out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) out[0] = 0x6e626479;
}

inline unsigned blocks_for(int nitems) {
  return (unsigned)((nitems + kWavesPerBlock - 1) / kWavesPerBlock);
}

// ---- MODE_F32C target order -------------------------------------------------
__device__ __forceinline__ uint64_t spread16(uint32_t v) {  // 16 bits -> 48
  uint64_t x = v & 0xffffu;
  x = (x | (x << 16)) & 0x0000ff0000ffull;
  x = (x | (x << 8)) & 0x00f00f00f00full;
  x = (x | (x << 4)) & 0x0c30c30c30c3ull;
  x = (x | (x << 2)) & 0x249249249249ull;
  return x;
}

__device__ __forceinline__ uint32_t quant16(double v, double o, double inv) {
  const double q = fmin(fmax((v - o) * inv + 0.5, 0.0), 0.99999);  // NaN -> 0
  return (uint32_t)(q * 65536.0);
}

// 3-D Hilbert index of 16-bit coordinates: Skilling's transform from axes
// to the transposed index (J. Skilling, "Programming the Hilbert curve",
// AIP Conf. Proc. 707, 381, 2004), then the bits interleaved with axis 0
// most significant.  Consecutive indices are adjacent cells, so runs of 64
// bodies are compact (a Z-order curve jumps across octant boundaries, which
// cut through the dense centre of a system).
__device__ __forceinline__ uint64_t hilbert48(uint32_t a, uint32_t b,
                                              uint32_t c) {
  uint32_t X[3] = {a, b, c};
  for (uint32_t Q = 1u << 15; Q > 1; Q >>= 1) {
    const uint32_t P = Q - 1;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      if (X[i] & Q) {
        X[0] ^= P;  // invert
      } else {      // exchange
        const uint32_t t = (X[0] ^ X[i]) & P;
        X[0] ^= t;
        X[i] ^= t;
      }
    }
  }
  X[1] ^= X[0];  // Gray encode
  X[2] ^= X[1];
  uint32_t t = 0;
  for (uint32_t Q = 1u << 15; Q > 1; Q >>= 1)
    if (X[2] & Q) t ^= Q - 1;
  X[0] ^= t;
  X[1] ^= t;
  X[2] ^= t;
  return (spread16(X[0]) << 2) | (spread16(X[1]) << 1) | spread16(X[2]);
}

// One block of 64 per (tile, listed system): key = rank << 48 | Hilbert
// index of the body's position in the system's box; value = local index.
__global__ __launch_bounds__(kWave) void k_sort_keys(const Arena* Ap,
                                                     const SortSys* tab,
                                                     uint64_t* keys,
                                                     int32_t* vals) {
  const int r = blockIdx.y;
  const SortSys t = tab[r];
  const int l = (int)blockIdx.x * kWave + (int)threadIdx.x;
  if (l >= t.n) return;
  const CArena* A = (const CArena*)Ap;
  const int64_t i = t.off + l;
  const uint64_t code = hilbert48(quant16(A->x[i], t.ox, t.ix),
                                  quant16(A->y[i], t.oy, t.iy),
                                  quant16(A->z[i], t.oz, t.iz));
  keys[t.base + l] = ((uint64_t)r << 48) | code;
  vals[t.base + l] = l;
}

// Sorted keys are grouped by rank, so system r's order is slots
// [base, base + n) of the sorted values.
__global__ __launch_bounds__(kWave) void k_sort_scatter(const Arena* Ap,
                                                        const SortSys* tab,
                                                        const int32_t* vals) {
  const SortSys t = tab[blockIdx.y];
  const int k = (int)blockIdx.x * kWave + (int)threadIdx.x;
  if (k >= t.n) return;
  const CArena* A = (const CArena*)Ap;
  const int64_t q = t.off + k;
  int32_t v;
  if (vals) {
    v = vals[t.base + k];
    A->perm[q] = v;
    A->iperm[t.off + v] = k;
  } else {
    v = A->perm[q];
  }
  // the slot-ordered copy of both snapshot buffers (only the current one
  // matters; the next force round rewrites the other) and of g*m
  const int64_t i = t.off + v;
  const float g = A->gm32[i];
  for (int b = 0; b < 2; ++b)
    A->hs[b][q] = HSrc{A->sx[b][i], A->sy[b][i], A->sz[b][i], g, 0.f};
}

}  // namespace

hipError_t launch_force(int mode, int variant, const Arena* A,
                        const WorkItem* items, int nitems, const SysDesc* sys,
                        hipStream_t s) {
  if (nitems <= 0) return hipSuccess;
  if (variant < 0 || variant >= kForceVariants) return hipErrorInvalidValue;
  switch (mode) {
    case MODE_F64_STRICT:  // one shape: the reference's sequential sum
      hipLaunchKernelGGL((k_force<MODE_F64_STRICT, false, 1, 0>),
                         dim3(blocks_for(nitems)), dim3(kWave), 0, s, A, items,
                         nitems, sys);
      return hipGetLastError();
    case MODE_F64:
      return launch_force_f64(variant, A, items, nitems, sys, s);
    case MODE_F32C:
      return launch_force_f32c(variant, A, items, nitems, sys, s);
    case MODE_F32R:
      return launch_force_f32r(variant, A, items, nitems, sys, s);
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t launch_combine(int mode, const Arena* A, const WorkItem* items,
                          int nitems, const SysDesc* sys, hipStream_t s) {
  if (nitems <= 0) return hipSuccess;
  const dim3 grid(blocks_for(nitems)), block(kBlock);
  switch (mode) {
    case MODE_F64_STRICT:
      hipLaunchKernelGGL(k_combine<MODE_F64_STRICT>, grid, block, 0, s, A,
                         items, nitems, sys);
      break;
    case MODE_F64:
      hipLaunchKernelGGL(k_combine<MODE_F64>, grid, block, 0, s, A, items,
                         nitems, sys);
      break;
    case MODE_F32R:
      hipLaunchKernelGGL(k_combine<MODE_F32R>, grid, block, 0, s, A, items,
                         nitems, sys);
      break;
    case MODE_F32C:
      hipLaunchKernelGGL(k_combine<MODE_F32C>, grid, block, 0, s, A, items,
                         nitems, sys);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_potential(const Arena* A, const WorkItem* items, int nitems,
                            const SysDesc* sys, hipStream_t s) {
  if (nitems <= 0) return hipSuccess;
  hipLaunchKernelGGL((k_force<MODE_F64, true, 1, 0>), dim3(blocks_for(nitems)),
                     dim3(kWave), 0, s, A, items, nitems, sys);
  return hipGetLastError();
}

hipError_t launch_prepare(const Arena* A, const WorkItem* items, int nitems,
                          const SysDesc* sys, hipStream_t s) {
  if (nitems <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_prepare, dim3(blocks_for(nitems)), dim3(kBlock), 0, s,
                     A, items, nitems, sys);
  return hipGetLastError();
}

hipError_t launch_stage(const Arena* A, const WorkItem* items, int nitems,
                        const SysDesc* sys, int op, hipStream_t s) {
  if (nitems <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_stage, dim3(blocks_for(nitems)), dim3(kBlock), 0, s, A,
                     items, nitems, sys, op);
  return hipGetLastError();
}

hipError_t launch_diag(const Arena* A, const SysDesc* sys, const int32_t* list,
                       int nlist, double* out, hipStream_t s) {
  if (nlist <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_diag, dim3(nlist), dim3(kDiagBlock), 0, s, A, sys, list,
                     out);
  return hipGetLastError();
}

hipError_t launch_gather(const Arena* A, const WorkItem* items, int nitems,
                         const SysDesc* sys, double* stage, int64_t total,
                         int nfields, hipStream_t s) {
  if (nitems <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather, dim3(blocks_for(nitems)), dim3(kBlock), 0, s, A,
                     items, nitems, sys, stage, total, nfields);
  return hipGetLastError();
}

hipError_t launch_sort_keys(const Arena* A, const SortSys* tab, int nsys,
                            int max_tiles, uint64_t* keys, int32_t* vals,
                            hipStream_t s) {
  if (nsys <= 0 || max_tiles <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sort_keys, dim3(max_tiles, nsys), dim3(kWave), 0, s, A,
                     tab, keys, vals);
  return hipGetLastError();
}

hipError_t launch_sort_scatter(const Arena* A, const SortSys* tab, int nsys,
                               int max_tiles, const int32_t* vals,
                               hipStream_t s) {
  if (nsys <= 0 || max_tiles <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sort_scatter, dim3(max_tiles, nsys), dim3(kWave), 0, s,
                     A, tab, vals);
  return hipGetLastError();
}

hipError_t sort_pairs(void* tmp, size_t& bytes, const uint64_t* kin,
                      uint64_t* kout, const int32_t* vin, int32_t* vout, int n,
                      int end_bit, hipStream_t s) {
  return hipcub::DeviceRadixSort::SortPairs(tmp, bytes, kin, kout, vin, vout,
                                            n, 0, end_bit, s);
}

hipError_t launch_expand(const ItemSpan* spans, int nspans, int total,
                         WorkItem* out, hipStream_t s) {
  if (total <= 0 || nspans <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_expand, dim3((total + 255) / 256), dim3(256), 0, s,
                     spans, nspans, total, out);
  return hipGetLastError();
}

hipError_t launch_clear_range(const Arena* A, int64_t off, int64_t n, int what,
                              hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_clear_range, dim3((unsigned)((n + 255) / 256)),
                     dim3(256), 0, s, A, off, n, what);
  return hipGetLastError();
}

hipError_t launch_probe(int* out, hipStream_t s) {
  hipLaunchKernelGGL(k_probe, dim3(1), dim3(64), 0, s, out);
  return hipGetLastError();
}

}  // namespace nbody
