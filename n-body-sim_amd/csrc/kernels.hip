// gfx950 kernels for the direct-sum gravity step.
//
// One wavefront owns one work item: 64 consecutive target slots of one
// system (one target per lane) and one part of its source range.
//   * F64_STRICT walks all sources in index order, one part, so each target
//     sums j = 0..n-1 with j == i skipped, the reference's sequential order
//     (src/gravity/cpu/newton.rs:104-124).
//   * F64 and F32C read sources through address-space-4 pointers: the
//     addresses are wave-uniform, so they land in SGPRs via s_load and the
//     scalar cache broadcasts them (no LDS, no barriers).
//   * F32R walks targets and sources in Hilbert order; each 64-source chunk
//     is differenced against the wave's origin once, one source per lane,
//     and staged in the wave's own LDS slice.
// With the source range split into parts, every part writes partial sums
// and k_combine adds them in part order, so results do not depend on
// scheduling.
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
#include "kernels.h"
#include "nbody_types.h"

namespace nbody {
namespace {


// Source-loop shape: bodies per chunk for the f64 and f32 paths and chunks
// per compensated group.
#ifndef NBODY_U64
#define NBODY_U64 8
#endif
#ifndef NBODY_U32
#define NBODY_U32 8
#endif
#ifndef NBODY_KC
#define NBODY_KC 4
#endif
constexpr int kU64 = NBODY_U64, kU32 = NBODY_U32, kKC = NBODY_KC;

struct Bounds {  // chunk-aligned source regions for one work item
  int64_t a0, a1;  // unmasked, before the diagonal block
  int64_t b0, b1;  // masked (j == i possible, or j >= end)
  int64_t c0, c1;  // unmasked, after the diagonal block
  int64_t d0, d1;  // masked tail (j >= end possible)
};

template <int U>
__device__ __forceinline__ Bounds make_bounds(int64_t off, int64_t n,
                                              int64_t t0) {
  Bounds b;
  const int64_t e_up = off + ((n + U - 1) / U) * U;
  const int64_t nfull = off + (n / U) * U;
  b.a0 = off;
  b.a1 = t0;
  b.b0 = t0;
  b.b1 = min(t0 + (int64_t)kWave, e_up);
  b.c0 = b.b1;
  b.c1 = max(b.b1, nfull);
  b.d0 = b.c1;
  b.d1 = e_up;
  return b;
}

// ---- binary64, reference operation order (no contraction, IEEE sqrt/div).
template <bool MASK>
__device__ __forceinline__ void sweep_strict(const cf64* X, const cf64* Y,
                                             const cf64* Z, const cf64* G,
                                             int64_t j0, int64_t j1,
                                             int64_t i, int64_t end,
                                             double xi, double yi, double zi,
                                             double eps2, double& ax,
                                             double& ay, double& az) {
#pragma clang fp contract(off)
  constexpr int U = 4;
  for (int64_t j = j0; j < j1; j += U) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const double dx = X[j + u] - xi;
      const double dy = Y[j + u] - yi;
      const double dz = Z[j + u] - zi;
      const double r2 = ((dx * dx) + (dy * dy)) + (dz * dz);
      const double inv = 1.0 / __builtin_sqrt(r2 + eps2);
      const double inv3 = (inv * inv) * inv;
      const double k = G[j + u] * inv3;  // G[j] = g * m_j, same bits
      const double tx = k * dx, ty = k * dy, tz = k * dz;
      if (MASK) {
        const int64_t jj = j + u;
        const bool ok = (jj != i) && (jj < end);
        ax = ok ? ax + tx : ax;
        ay = ok ? ay + ty : ay;
        az = ok ? az + tz : az;
      } else {
        ax = ax + tx;
        ay = ay + ty;
        az = az + tz;
      }
    }
  }
}

// ---- fast paths: one loop over the part's source range ------------------
//
// The sources [jlo, jhi) of a part are walked in groups of KC chunks of U
// bodies.  A group that overlaps the wave's own targets [t0, t0+64) or runs
// past the system's end takes the masked variant (a uniform scalar branch);
// all other groups run unmasked.  Each chunk's sources arrive through scalar
// loads; with several waves per SIMD (split source ranges) their latency is
// covered by the other waves.

// Chunk-aligned source range of one part: the system's chunks [0, nchunks)
// cut into nsplit contiguous runs, part p taking [p*nchunks/nsplit,
// (p+1)*nchunks/nsplit).  Parts cover every chunk once, in order.
template <int U>
__device__ __forceinline__ void part_range(int64_t off, int64_t n, int nsplit,
                                           int part, int64_t& lo,
                                           int64_t& hi) {
  const uint32_t nchunks = (uint32_t)((n + U - 1) / U);
  lo = off + (int64_t)((uint64_t)nchunks * (uint32_t)part / (uint32_t)nsplit) * U;
  hi = off +
       (int64_t)((uint64_t)nchunks * (uint32_t)(part + 1) / (uint32_t)nsplit) * U;
}

template <bool M>
using Masked = std::integral_constant<bool, M>;

// Source bundles: the arrays a chunk reads and its register buffer.
struct SrcF64 {  // f64 positions and g*m
  const cf64 *X, *Y, *Z, *G;
  template <int U>
  struct Buf {
    double x[U], y[U], z[U], g[U];
  };
  template <int U>
  __device__ __forceinline__ void load(int64_t j, Buf<U>& b) const {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      b.x[u] = X[j + u];
      b.y[u] = Y[j + u];
      b.z[u] = Z[j + u];
      b.g[u] = G[j + u];
    }
  }
};

struct SrcX64 {  // f64 positions, f32 g*m
  const cf64 *X, *Y, *Z;
  const cf32* G;
  template <int U>
  struct Buf {
    double x[U], y[U], z[U];
    float g[U];
  };
  template <int U>
  __device__ __forceinline__ void load(int64_t j, Buf<U>& b) const {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      b.x[u] = X[j + u];
      b.y[u] = Y[j + u];
      b.z[u] = Z[j + u];
      b.g[u] = G[j + u];
    }
  }
};

// chunk(masked, buf, j) consumes one chunk; flush() runs after every group.
// Reads stay below hi <= off + roundup(n, U) <= off + span.
template <int U, int KC, typename Src, typename Chunk, typename Flush>
__device__ __forceinline__ void drive(const Src& src, int64_t jlo, int64_t jhi,
                                      int64_t t0, int64_t end, Chunk&& chunk,
                                      Flush&& flush) {
  auto step = [&](int64_t j, auto masked) {
    typename Src::template Buf<U> b;
    src.template load<U>(j, b);
    chunk(masked, b, j);
  };
  constexpr int GRP = KC * U;
  int64_t j = jlo;
  for (; j + GRP <= jhi; j += GRP) {
    const bool m = (j < t0 + kWave && j + GRP > t0) || (j + GRP > end);
    // not unrolled: bounds the live source registers to one chunk
    if (m) {
#pragma unroll 1
      for (int g = 0; g < KC; ++g) step(j + g * U, Masked<true>{});
    } else {
#pragma unroll 1
      for (int g = 0; g < KC; ++g) step(j + g * U, Masked<false>{});
    }
    flush();
  }
  for (; j < jhi; j += U) {  // remaining chunks, always masked
    step(j, Masked<true>{});
    flush();
  }
}

// binary64 pair terms: rsq + refinement, fused multiply-add (POT: potential).
template <bool POT, int U>
struct ChunkF64 {
  int64_t i, end;
  double xi, yi, zi, eps2;
  double &ax, &ay, &az;
  template <typename MK>
  __device__ __forceinline__ void operator()(
      MK, const SrcF64::Buf<U>& b, int64_t j) const {
    constexpr bool MASK = MK::value;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      double dx = b.x[u] - xi;
      double dy = b.y[u] - yi;
      double dz = b.z[u] - zi;
      const int64_t jj = j + u;
      const bool ok = !MASK || ((jj != i) && (jj < end));
      if (MASK) {  // masked sources contribute exactly zero
        dx = ok ? dx : 0.0;
        dy = ok ? dy : 0.0;
        dz = ok ? dz : 0.0;
      }
      const double r2 = fma(dz, dz, fma(dy, dy, fma(dx, dx, eps2)));
      double k = b.g[u] * (POT ? rsqrt_f64(r2) : inv_cube_f64(r2));
      if (MASK) k = ok ? k : 0.0;
      if (POT) {
        ax += k;
      } else {
        ax = fma(k, dx, ax);
        ay = fma(k, dy, ay);
        az = fma(k, dz, az);
      }
    }
  }
};

// ---- binary32 pair terms, compensated accumulation.  Each group of KC
// chunks is summed in f32, then added to the running sum with Kahan's
// compensated summation (Kahan 1965), which keeps the running sum's rounding
// error at the level of one group sum instead of growing with n.
struct Kahan3 {
  float sx, sy, sz, cx, cy, cz;
};

__device__ __forceinline__ void kahan_add(float& s, float& c, float v) {
  const float y = v - c;
  const float t = s + y;
  c = (t - s) - y;
  s = t;
}

// MODE_F32C: f64 positions, each difference rounded once to f32 -- exact
// to f64 rounding at any distance from the origin, so close pairs keep full
// f32 accuracy.
template <typename Src, typename S, int U>
struct ChunkF32 {
  int64_t i, end;
  S xi, yi, zi;
  float eps2;
  float &px, &py, &pz;
  template <typename MK>
  __device__ __forceinline__ void operator()(
      MK, const typename Src::template Buf<U>& b, int64_t j) const {
    constexpr bool MASK = MK::value;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      float dx = (float)(b.x[u] - xi);
      float dy = (float)(b.y[u] - yi);
      float dz = (float)(b.z[u] - zi);
      const int64_t jj = j + u;
      const bool ok = !MASK || ((jj != i) && (jj < end));
      if (MASK) {  // masked sources contribute exactly zero
        dx = ok ? dx : 0.f;
        dy = ok ? dy : 0.f;
        dz = ok ? dz : 0.f;
      }
      const float r2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, eps2)));
      const float inv = __builtin_amdgcn_rsqf(r2);
      float k = b.g[u] * ((inv * inv) * inv);
      if (MASK) k = ok ? k : 0.f;
      px = fmaf(k, dx, px);
      py = fmaf(k, dy, py);
      pz = fmaf(k, dz, pz);
    }
  }
};

// MODE_F32R sources read from LDS before use, per batch (near chunks: half):
// 4 measured about 1 % faster than 8 and 2 % faster than 16.  The batch
// changes only when loads issue, not the summation order.
#ifndef NBODY_LDS_BATCH
#define NBODY_LDS_BATCH 4
#endif
#ifndef NBODY_NEAR_K
#define NBODY_NEAR_K 8.0f
#endif
constexpr int kLdsBatch = NBODY_LDS_BATCH;

// v of lane L (an inactive lane holds its clamped body: any body will do).
__device__ __forceinline__ double lane_f64(double v, int L) {
  const int64_t b = __builtin_bit_cast(int64_t, v);
  const int32_t lo = __builtin_amdgcn_readlane((int32_t)b, L);
  const int32_t hi = __builtin_amdgcn_readlane((int32_t)(b >> 32), L);
  return __builtin_bit_cast(double, ((int64_t)hi << 32) | (uint32_t)lo);
}

// ---- MODE_F32R: relative coordinates, near/far chunks -------------------
//
// Targets and sources are both walked in Hilbert order (perm), so a wave's
// 64 targets and each 64-source chunk are compact.  Every coordinate is taken
// relative to the wave's origin o (its middle target, lane 32): s = f32(x - o)
// from an f64 difference, plus the residual r = f32((x - o) - s).
//
// Far chunks use dx = s_j - s_i, a pure f32 pair loop.  Its error per
// coordinate is at most 2^-24 (|x_i - o| + |x_j - o| + |dx|).  A chunk is far
// when every source j in it has R + |s_j| <= K * D_j, where R bounds |s_i|
// over the wave's targets and D_j is the distance from s_j to the targets'
// bounding box (so D_j <= |dx| for every target).  Then each far pair's
// difference is within (K + 1) 2^-24 |dx| of exact, whatever the shape of
// the tile or where the pair sits.
//
// Near chunks (a source close to the tile compared with the distances to
// the origin: the wave's own chunk, its neighbours along the curve, a tight
// pair far from the origin) use dx = (s_j - s_i) + (r_j - r_i), which is
// within about 2^-23 |dx| + 2^-48 (|x_i - o| + |x_j - o|) of exact, like
// MODE_F32C's one rounding of an f64 difference.  The chunk's class is a
// wave vote, so each chunk runs one loop.
constexpr float kNearK = NBODY_NEAR_K;

struct TargetR {
  float x, y, z;        // s_i
  float rx, ry, rz;     // r_i
  float lox, loy, loz;  // bounding box of the wave's s_i
  float hix, hiy, hiz;
  float rad;            // max |s_i| over the wave
};

// Orders one wave's LDS accesses around a staging step.  Each wave stages
// into its own LDS slice and a wave's LDS operations are performed in the
// order issued, so no workgroup barrier is needed; this only stops the
// compiler from moving memory accesses across the point.  (A workgroup
// barrier would also make the wave wait for every outstanding global load.)
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) v = fminf(v, __shfl_xor(v, d, kWave));
  return v;
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) v = fmaxf(v, __shfl_xor(v, d, kWave));
  return v;
}

__device__ __forceinline__ float uniform_f32(float v) {  // into an SGPR
  return __builtin_bit_cast(
      float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int32_t, v)));
}

// True unless the bound R + |s| <= K * D holds for source s (see above).
// Overflow or NaN in the test makes the source near (the exact path).  The
// test runs in f32 with v_sqrt_f32 (about 1 ulp); its rounding moves K by a
// few parts in 10^7.
__device__ __forceinline__ bool near_source(const TargetR& t, float sx,
                                            float sy, float sz) {
  const float ex = fmaxf(fmaxf(t.lox - sx, sx - t.hix), 0.f);
  const float ey = fmaxf(fmaxf(t.loy - sy, sy - t.hiy), 0.f);
  const float ez = fmaxf(fmaxf(t.loz - sz, sz - t.hiz), 0.f);
  const float d2 = fmaf(ez, ez, fmaf(ey, ey, ex * ex));
  const float a = t.rad + __builtin_amdgcn_sqrtf(fmaf(sz, sz, fmaf(sy, sy, sx * sx)));
  return !(a * a <= (kNearK * kNearK) * d2);
}

// The value in a VGPR.  Uniform values otherwise stay in SGPRs, and on
// gfx950 an f32 VALU instruction with an SGPR operand issues at about 4.2
// cycles per wave64 against about 2.4 with VGPR operands only (measured).
__device__ __forceinline__ float in_vgpr(float v) {
  // a DPP move with the identity lane pattern (quad_perm [0,1,2,3])
  return __builtin_bit_cast(
      float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0xe4, 0xf,
                                      0xf, false));
}

// One staged chunk of 64 sources against the lane's target.  RES adds the
// residual correction (near chunks).  Group sums of 32 sources are added with
// Kahan compensation.  NOMASK (SYS_F32_SELF_FINITE): eps2 > 0 and the self
// term's k is finite, so a source at the target's own position (itself)
// contributes exactly zero.  Otherwise (zero softening, or g*m*eps^-3 beyond
// the f32 range) k is clamped to +-FLT_MAX (fminf/fmaxf return the number
// when k is NaN from 0 * inf), so a pair with a zero difference adds exactly
// zero: the target's own term, and coincident bodies, which the reference's
// f64 path gives NaN at zero softening and zero when softened.  Sources past
// the system's end stage as zero mass.
template <bool NOMASK, bool RES>
__device__ __forceinline__ void chunk_f32r(const float4* stage,
                                           const float4* res,
                                           const TargetR& t, float eps2_s,
                                           Kahan3& acc) {
  constexpr int B = RES ? kLdsBatch / 2 : kLdsBatch;
  const float eps2 = in_vgpr(eps2_s);
#pragma unroll 1
  for (int q = 0; q < kWave; q += 32) {
    float px = 0.f, py = 0.f, pz = 0.f;
#pragma unroll
    for (int u0 = 0; u0 < 32; u0 += B) {
      float4 s[B], r[B];  // issue the batch's reads before any use
#pragma unroll
      for (int u = 0; u < B; ++u) {
        s[u] = stage[q + u0 + u];
        if (RES) r[u] = res[q + u0 + u];
      }
#pragma unroll
      for (int u = 0; u < B; ++u) {
        float dx = s[u].x - t.x, dy = s[u].y - t.y, dz = s[u].z - t.z;
        if (RES) {
          dx = dx + (r[u].x - t.rx);
          dy = dy + (r[u].y - t.ry);
          dz = dz + (r[u].z - t.rz);
        }
        const float r2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, eps2)));
        const float inv = __builtin_amdgcn_rsqf(r2);
        float k = s[u].w * ((inv * inv) * inv);
        // k clamped to the finite range: a zero difference then adds
        // exactly zero (an overflowing or 0 * inf k times 0)
        if (!NOMASK) k = fmaxf(fminf(k, FLT_MAX), -FLT_MAX);
        px = fmaf(k, dx, px);
        py = fmaf(k, dy, py);
        pz = fmaf(k, dz, pz);
      }
    }
    kahan_add(acc.sx, acc.cx, px);
    kahan_add(acc.sy, acc.cy, py);
    kahan_add(acc.sz, acc.cz, pz);
  }
}

// Sources in slots [jlo, jhi) (64-aligned, Hilbert order), read from the
// slot-ordered snapshot.  Each chunk is loaded once per wave, one source per
// lane, differenced in f64 against the origin, rounded to s, voted near or
// far, and staged in LDS (with the residuals r when near); every lane then
// reads the sources back as broadcasts.  Sources past the system's end stage
// as zero mass at the origin.
template <bool NOMASK>
__device__ __forceinline__ void sweep_f32r(float4* stage, float4* res,
                                           const double* X, const double* Y,
                                           const double* Z, const float* G,
                                           int64_t jlo, int64_t jhi,
                                           int64_t end, double ox, double oy,
                                           double oz, const TargetR& t,
                                           float eps2, Kahan3& acc) {
  const int lane = (int)(threadIdx.x & (kWave - 1));
  for (int64_t j0 = jlo; j0 < jhi; j0 += kWave) {
    const int64_t j = j0 + lane;
    const bool valid = j < end;
    double dx = 0.0, dy = 0.0, dz = 0.0;
    float g = 0.f;
    if (valid) {
      dx = X[j] - ox;
      dy = Y[j] - oy;
      dz = Z[j] - oz;
      g = G[j];
    }
    const float sx = (float)dx, sy = (float)dy, sz = (float)dz;
    stage[lane] = {sx, sy, sz, g};
    const bool near = valid && near_source(t, sx, sy, sz);
    const bool any_near = __ballot(near) != 0;  // wave-uniform
    if (any_near)
      res[lane] = {(float)(dx - (double)sx), (float)(dy - (double)sy),
                   (float)(dz - (double)sz), 0.f};
    wave_sync();
    if (any_near)
      chunk_f32r<NOMASK, true>(stage, res, t, eps2, acc);
    else
      chunk_f32r<NOMASK, false>(stage, res, t, eps2, acc);
    wave_sync();
  }
}

template <int MODE, bool POT>
__global__ __launch_bounds__(kBlock) void k_force(const Arena* Ap,
                                                  const WorkItem* items,
                                                  int nitems,
                                                  const SysDesc* sysd) {
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int idx = (int)blockIdx.x * kWavesPerBlock + wave;
  if (idx >= nitems) return;
  const WorkItem it = load_uniform(items + idx);
  const CArena* A = (const CArena*)Ap;
  const CSys* sd = (const CSys*)(sysd + it.sys);
  const int64_t off = sd->off, n = sd->n, end = off + n;
  const int64_t t0 = off + (int64_t)it.tile * kWave;
  const int64_t slot = t0 + (int64_t)(threadIdx.x & (kWave - 1));
  const bool active = slot < end;
  // target body: the tile slot itself, or in MODE_F32R the slot's body in
  // Hilbert order
  const int64_t il = MODE == MODE_F32R ? off + A->perm[active ? slot : end - 1]
                                       : (active ? slot : end - 1);
  const int64_t i = MODE == MODE_F32R ? il : slot;
  const int src = it.src;
  double ax = 0.0, ay = 0.0, az = 0.0;

  const int nsplit = sd->nsplit, part = it.part;
  if (MODE == MODE_F32R) {
    __shared__ float4 stage[kWavesPerBlock][2][kWave];  // staged s, residuals
    const int64_t ts = active ? slot : end - 1;  // in this tile
    const double px0 = A->hx[src][ts], py0 = A->hy[src][ts],
                 pz0 = A->hz[src][ts];
    const double ox = lane_f64(px0, 32);  // the middle target
    const double oy = lane_f64(py0, 32);
    const double oz = lane_f64(pz0, 32);
    const double tx = px0 - ox, ty = py0 - oy, tz = pz0 - oz;
    TargetR t;
    t.x = (float)tx;
    t.y = (float)ty;
    t.z = (float)tz;
    t.rx = (float)(tx - (double)t.x);
    t.ry = (float)(ty - (double)t.y);
    t.rz = (float)(tz - (double)t.z);
    t.lox = uniform_f32(wave_min(t.x));
    t.loy = uniform_f32(wave_min(t.y));
    t.loz = uniform_f32(wave_min(t.z));
    t.hix = uniform_f32(wave_max(t.x));
    t.hiy = uniform_f32(wave_max(t.y));
    t.hiz = uniform_f32(wave_max(t.z));
    t.rad = uniform_f32(wave_max(
        __builtin_amdgcn_sqrtf(fmaf(t.z, t.z, fmaf(t.y, t.y, t.x * t.x)))));
    Kahan3 acc = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    int64_t jlo, jhi;
    part_range<kWave>(off, n, nsplit, part, jlo, jhi);
    const float e2 = sd->eps2f;
    if (sd->flags & SYS_F32_SELF_FINITE)
      sweep_f32r<true>(stage[wave][0], stage[wave][1], A->hx[src], A->hy[src],
                       A->hz[src], A->hg, jlo, jhi, end, ox, oy, oz, t, e2,
                       acc);
    else
      sweep_f32r<false>(stage[wave][0], stage[wave][1], A->hx[src], A->hy[src],
                        A->hz[src], A->hg, jlo, jhi, end, ox, oy, oz, t, e2,
                        acc);
    ax = (double)acc.sx - (double)acc.cx;
    ay = (double)acc.sy - (double)acc.cy;
    az = (double)acc.sz - (double)acc.cz;
  } else if (MODE == MODE_F32C) {
    constexpr int U = kU32, KC = kKC;
    const SrcX64 src_bundle{(const cf64*)A->sx[src], (const cf64*)A->sy[src],
                            (const cf64*)A->sz[src], (const cf32*)A->gm32};
    const double xi = A->sx[src][il], yi = A->sy[src][il], zi = A->sz[src][il];
    Kahan3 acc = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    float px = 0.f, py = 0.f, pz = 0.f;
    int64_t jlo, jhi;
    part_range<U>(off, n, nsplit, part, jlo, jhi);
    ChunkF32<SrcX64, double, U> ch{i,  end, xi, yi, zi, in_vgpr(sd->eps2f),
                                   px, py, pz};
    drive<U, KC>(src_bundle, jlo, jhi, t0, end, ch, [&] {
      kahan_add(acc.sx, acc.cx, px);
      kahan_add(acc.sy, acc.cy, py);
      kahan_add(acc.sz, acc.cz, pz);
      px = py = pz = 0.f;
    });
    ax = (double)acc.sx - (double)acc.cx;
    ay = (double)acc.sy - (double)acc.cy;
    az = (double)acc.sz - (double)acc.cz;
  } else {
    const cf64* X = (const cf64*)A->sx[src];
    const cf64* Y = (const cf64*)A->sy[src];
    const cf64* Z = (const cf64*)A->sz[src];
    const cf64* G = (const cf64*)A->gm64;
    const double xi = A->sx[src][il], yi = A->sy[src][il], zi = A->sz[src][il];
    const double e2 = sd->eps2;
    if (MODE == MODE_F64_STRICT) {
      const Bounds b = make_bounds<4>(off, n, t0);
      sweep_strict<false>(X, Y, Z, G, b.a0, b.a1, i, end, xi, yi, zi, e2, ax,
                          ay, az);
      sweep_strict<true>(X, Y, Z, G, b.b0, b.b1, i, end, xi, yi, zi, e2, ax,
                         ay, az);
      sweep_strict<false>(X, Y, Z, G, b.c0, b.c1, i, end, xi, yi, zi, e2, ax,
                          ay, az);
      sweep_strict<true>(X, Y, Z, G, b.d0, b.d1, i, end, xi, yi, zi, e2, ax,
                         ay, az);
    } else {
      constexpr int U = kU64;
      int64_t jlo, jhi;
      part_range<U>(off, n, nsplit, part, jlo, jhi);
      ChunkF64<POT, U> ch{i, end, xi, yi, zi, e2, ax, ay, az};
      drive<U, kKC>(SrcF64{X, Y, Z, G}, jlo, jhi, t0, end, ch, [] {});
    }
  }
  if (!active) return;
  if (sd->nsplit > 1) {  // partial sum; k_combine finishes the target
    // partials in body order in every mode (in MODE_F32R a wave's 64 writes
    // are scattered, so k_combine can walk bodies with coalesced accesses)
    const int64_t q = (int64_t)it.part * A->pstride + i;
    A->px[q] = ax;
    if (!POT) {
      A->py[q] = ay;
      A->pz[q] = az;
    }
    return;
  }
  if (POT) {
    A->phi[i] = ax;
    return;
  }
  epilogue<MODE>(A, sd, it.stage, src ^ 1, i, slot, ax, ay, az);
}

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
  double ax = A->px[k];
  if (it.stage == STAGE_POT) {
    for (int p = 1; p < sd->nsplit; ++p) ax = ax + A->px[p * st + k];
    A->phi[k] = ax;
    return;
  }
  double ay = A->py[k], az = A->pz[k];
  for (int p = 1; p < sd->nsplit; ++p) {
    ax = ax + A->px[p * st + k];
    ay = ay + A->py[p * st + k];
    az = az + A->pz[p * st + k];
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

__global__ void k_fill_zero(double* p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
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
  const uint64_t code = hilbert48(quant16(A->x[i], t.ox, t.inv),
                                  quant16(A->y[i], t.oy, t.inv),
                                  quant16(A->z[i], t.oz, t.inv));
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
  for (int b = 0; b < 2; ++b) {
    A->hx[b][q] = A->sx[b][i];
    A->hy[b][q] = A->sy[b][i];
    A->hz[b][q] = A->sz[b][i];
  }
  A->hg[q] = A->gm32[i];
}

}  // namespace

hipError_t launch_force(int mode, const Arena* A, const WorkItem* items,
                        int nitems, const SysDesc* sys, hipStream_t s) {
  if (nitems <= 0) return hipSuccess;
  const dim3 grid(blocks_for(nitems)), block(kBlock);
  switch (mode) {
    case MODE_F64_STRICT:
      hipLaunchKernelGGL((k_force<MODE_F64_STRICT, false>), grid, block, 0, s,
                         A, items, nitems, sys);
      break;
    case MODE_F64:
      hipLaunchKernelGGL((k_force<MODE_F64, false>), grid, block, 0, s, A,
                         items, nitems, sys);
      break;
    case MODE_F32R:
      hipLaunchKernelGGL((k_force<MODE_F32R, false>), grid, block, 0, s, A,
                         items, nitems, sys);
      break;
    case MODE_F32C:
      hipLaunchKernelGGL((k_force<MODE_F32C, false>), grid, block, 0, s, A,
                         items, nitems, sys);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
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
  hipLaunchKernelGGL((k_force<MODE_F64, true>), dim3(blocks_for(nitems)),
                     dim3(kBlock), 0, s, A, items, nitems, sys);
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

hipError_t launch_fill_zero(double* p, int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)((n + 255) / 256)),
                     dim3(256), 0, s, p, n);
  return hipGetLastError();
}

hipError_t launch_probe(int* out, hipStream_t s) {
  hipLaunchKernelGGL(k_probe, dim3(1), dim3(64), 0, s, out);
  return hipGetLastError();
}

}  // namespace nbody
