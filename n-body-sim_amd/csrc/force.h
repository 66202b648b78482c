// The force kernel k_force and its pair loops, shared by kernels.hip (the
// strict and potential evaluations) and force_variants.hip (the tuned
// variants of MODE_F64, MODE_F32C and MODE_F32R, one object per mode).
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
// scheduling.  The integrator stage is fused into the epilogue (of this
// kernel, or of k_combine when split).
#pragma once
#include <hip/hip_runtime.h>

#include <cfloat>
#include <type_traits>

#include "device.h"
#include "nbody_types.h"

namespace nbody {
namespace {

// Loop shapes of the force kernel, chosen per size bucket by the tuning
// table (variant = 3 * waves-per-workgroup index + shape):
//   U   sources per chunk (scalar-loaded) in MODE_F64 and MODE_F32C; the
//       source parts are whole chunks, so U moves their boundaries;
//   KC  chunks per group: MODE_F32C adds each group's f32 sum to the
//       compensated running sum (so U and KC set its summation order, and U
//       sets MODE_F64's through the part boundaries);
//   B   sources read from LDS per batch in MODE_F32R (near chunks: B / 2);
//       its chunks are always 64 sources, so its sums do not depend on the
//       shape.
// Shape 0 is the one measured best on one large system.
template <int S>
struct Shape;
template <>
struct Shape<0> {
  static constexpr int U = 8, KC = 4, B = 4;
};
template <>
struct Shape<1> {
  static constexpr int U = 4, KC = 8, B = 2;
};
template <>
struct Shape<2> {
  static constexpr int U = 8, KC = 2, B = 8;
};
constexpr int kShapes = 3;
constexpr int kWaveCounts[3] = {1, 2, 4};  // waves per workgroup by index

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

// MODE_F32R reads sources from LDS before use, per batch of Shape::B (near
// chunks: half); the batch changes only when loads issue, not the summation
// order.
#ifndef NBODY_NEAR_K
#define NBODY_NEAR_K 8.0f
#endif

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

// Source classes by the bound R + |s| <= K * D (see above): near unless it
// holds at K = kNearK; matrix-core class (far enough for r^2 from the Gram
// form on the matrix cores, below) when it holds at K = kMatK, D^2 >= 2^-60,
// and the source's terms stay in range there: |g| max(|s|, R, 1) <= 2^100
// D^2 min(D^2, 1) <= 2^100 D^3 (so w g, w g |s| and R w g stay below 2^100
// in magnitude with w = r^-3 <= D^-3) and |g| |s| <= 2^120 (the staged g s).  Overflow or NaN in a test makes the
// source near (the exact path) or keeps it out of the matrix-core class.
// The tests run in f32 with v_sqrt_f32 (about 1 ulp); its rounding moves K
// by a few parts in 10^7.
struct SrcClass {
  bool near, mat;
};

#ifndef NBODY_F32R_MATRIX
#define NBODY_F32R_MATRIX 1
#endif
constexpr float kMatK = 4.0f;
constexpr float kMatMinD2 = 0x1p-60f;

__device__ __forceinline__ SrcClass classify_source(const TargetR& t, float sx,
                                                    float sy, float sz,
                                                    float g) {
  const float ex = fmaxf(fmaxf(t.lox - sx, sx - t.hix), 0.f);
  const float ey = fmaxf(fmaxf(t.loy - sy, sy - t.hiy), 0.f);
  const float ez = fmaxf(fmaxf(t.loz - sz, sz - t.hiz), 0.f);
  const float d2 = fmaf(ez, ez, fmaf(ey, ey, ex * ex));
  const float sn = __builtin_amdgcn_sqrtf(fmaf(sz, sz, fmaf(sy, sy, sx * sx)));
  const float a = t.rad + sn;
  const float a2 = a * a;
  const bool near = !(a2 <= (kNearK * kNearK) * d2);
  if (!NBODY_F32R_MATRIX) return {near, false};
  const float d3 = d2 * fminf(d2, 1.f);  // <= D^3
  const float ga = fabsf(g);  // masses may be negative
  const float gs = ga * sn;
  const bool mat = (a2 <= (kMatK * kMatK) * d2) && (d2 >= kMatMinD2) &&
                   (ga * fmaxf(fmaxf(sn, t.rad), 1.f) <= 0x1p100f * d3) &&
                   (gs <= 0x1p120f);
  return {near, mat};
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

// ---- MODE_F32R, matrix-core class ---------------------------------------
//
// For a chunk whose every source is far at K = kMatK, r^2 comes from the Gram
// form r^2 = q + p - 2 s.t (q = |s|^2 + eps^2 per source, p = |t|^2 per
// target) on v_mfma_f32_16x16x32_bf16.  Each f32 value is split exactly into
// three bf16 pieces, v = v1 + v2 + v3 (round to nearest each: |v2| <= 2^-8 |v|,
// |v3| <= 2^-16 |v|), and the K = 32 slots of one source and one target hold
// the products v_a * w_b with a + b <= 4 of each coordinate (the dropped ones
// are below 2^-24 |s||t|), q's pieces against 1 and 1 against p's pieces.
// The matrix core forms each product exactly and adds in f32, rounding at
// every addition (measured), so r^2 carries at most about 31 roundings of
// quantities below (|s| + |t|)^2 + eps^2 <= K^2 r^2: a relative error below
// 31 K^2 2^-24 = 3.0e-5 at K = 4 in the worst case.  Sources placed at the
// class boundary give at most 26 units of 2^-24 in any summation order
// (tests/test_f32r_matrix_model.py), and typical pairs a few units.
//
// Per 16-source block and per group c of 16 targets, one instruction gives
// lane (grp = l/16, col = l%16) r^2 of targets 16c + col against sources
// 16b + 4grp + r in its four result registers; the lane then adds
// w (g s, g) with w = r^-3 into its four-wide sums of target 16c + col: 7 VALU
// instructions per pair (v_rsq_f32, two multiplies, four FMAs) against 13 in
// the f32 loop, and the matrix instructions overlap the VALU.  After every two
// such chunks (and before any other class) the four lane groups' sums of each
// target are added through LDS, and a_i gains (sum w g s) - t_i (sum w g).
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));

constexpr uint32_t kBf16One = 0x3f80u;  // 1.0 in bf16

// bf16 of a (low half) and b (high half), each rounded to nearest even
__device__ __forceinline__ uint32_t pack_bf16(float a, float b) {
  typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
  const bf16x2_t v = {(__bf16)a, (__bf16)b};
  return __builtin_bit_cast(uint32_t, v);
}
__device__ __forceinline__ float bf16_low(uint32_t p) {
  return __builtin_bit_cast(float, p << 16);
}
__device__ __forceinline__ float bf16_high(uint32_t p) {
  return __builtin_bit_cast(float, p & 0xffff0000u);
}

// a = a1 + a2 + a3 and b = b1 + b2 + b3 exactly, packed pairwise: h = (a1, b1),
// m = (a2, b2), l = (a3, b3).  Each residual is exact in f32 (Sterbenz), and
// the last has at most 8 significant bits, so it is exact in bf16 (values
// within f32's normal range).
__device__ __forceinline__ void split3(float a, float b, uint32_t& h,
                                       uint32_t& m, uint32_t& l) {
  // (the empty asm statements keep the compiler from converting a and ra a
  // second time to read back their low halves)
  h = pack_bf16(a, b);
  __asm__ volatile("" : "+v"(h));
  const float ra = a - bf16_low(h), rb = b - bf16_high(h);
  m = pack_bf16(ra, rb);
  __asm__ volatile("" : "+v"(m));
  l = pack_bf16(ra - bf16_low(m), rb - bf16_high(m));
}

// The lane's target as the B operands of the four target groups: lane (grp,
// col) holds K dwords 4grp..4grp+3 of target 16c + col.  Target dwords
// (pieces of -2t, pieces of p = |t|^2):
//   (x1,y1) (x2,y2) (x3,y3) (x1,y1) | (x2,y2) (x1,y1) (z1,1) (z2,0) |
//   (z3,0) (z1,1) (z2,0) (z1,1) | (p1,p2) (p3,0) 0 0
// Source dwords (pieces of s, of q = |s|^2 + eps^2), slot for slot:
//   (x1,y1) (x1,y1) (x1,y1) (x2,y2) | (x2,y2) (x3,y3) (z1,q1) (z1,q1) |
//   (z1,q1) (z2,q2) (z2,q2) (z3,q3) | (1,1) (1,0) 0 0
// The last source quad is the same for every source and stays in LDS once.
// kq: the wave's quad area (at least 257 quads); leaves quad 256 holding it.
__device__ __forceinline__ void mat_targets(u32x4_t* kq, const TargetR& t,
                                            u32x4_t (&Bt)[4]) {
  const int lane = (int)(threadIdx.x & (kWave - 1));
  const int grp = lane >> 4, col = lane & 15;
  uint32_t xh, xm, xl, zh, zm, zl;
  split3(-2.f * t.x, -2.f * t.y, xh, xm, xl);
  split3(-2.f * t.z, fmaf(t.z, t.z, fmaf(t.y, t.y, t.x * t.x)), zh, zm, zl);
  const uint32_t z1 = (zh & 0xffffu) | (kBf16One << 16), z2 = zm & 0xffffu,
                 z3 = zl & 0xffffu;
  kq[4 * lane + 0] = u32x4_t{xh, xm, xl, xh};
  kq[4 * lane + 1] = u32x4_t{xm, xh, z1, z2};
  kq[4 * lane + 2] = u32x4_t{z3, z1, z2, z1};
  kq[4 * lane + 3] = u32x4_t{(zh >> 16) | (zm & 0xffff0000u), zl >> 16, 0u, 0u};
  wave_sync();
#pragma unroll
  for (int c = 0; c < 4; ++c) Bt[c] = kq[4 * (16 * c + col) + grp];
  wave_sync();
  if (lane == 0) kq[256] = u32x4_t{kBf16One | (kBf16One << 16), kBf16One, 0u, 0u};
}

// Stages the lane's source (s, q, g) for the matrix-core class: its three
// K quads at kq[3 lane ..] and (g s, g) at rec[lane].
__device__ __forceinline__ void mat_stage(u32x4_t* kq, float4* rec, float sx,
                                          float sy, float sz, float g,
                                          float eps2) {
  const int lane = (int)(threadIdx.x & (kWave - 1));
  const float q = fmaf(sz, sz, fmaf(sy, sy, fmaf(sx, sx, eps2)));
  uint32_t h, m, l, zh, zm, zl;
  split3(sx, sy, h, m, l);
  split3(sz, q, zh, zm, zl);
  kq[3 * lane + 0] = u32x4_t{h, h, h, m};
  kq[3 * lane + 1] = u32x4_t{m, l, zh, zh};
  kq[3 * lane + 2] = u32x4_t{zh, zm, zm, zl};
  rec[lane] = {g * sx, g * sy, g * sz, g};
}

// One staged chunk of 64 sources (matrix-core class) against the wave's 64
// targets, added to the lane's four-wide sums a[c] of targets 16c + col over
// its sources 16b + 4grp + r.  Reads kq quads abase + 48 b (lane groups 0-2:
// their source's quad; group 3: the constant quad 256) and rec.
__device__ __forceinline__ void chunk_f32r_mat(const u32x4_t* kq,
                                               const float4* rec,
                                               const u32x4_t (&Bt)[4],
                                               f32x4_t (&a)[4]) {
  const int lane = (int)(threadIdx.x & (kWave - 1));
  const int grp = lane >> 4, col = lane & 15;
  const int abase = grp < 3 ? 3 * col + grp : 256;
  const int astride = grp < 3 ? 48 : 0;
#pragma unroll 1
  for (int b = 0; b < 4; ++b) {
    const bf16x8_t As = __builtin_bit_cast(bf16x8_t, kq[abase + astride * b]);
    f32x4_t r2[4];
#pragma unroll
    for (int c = 0; c < 4; ++c)
      r2[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
          As, __builtin_bit_cast(bf16x8_t, Bt[c]), f32x4_t{0.f, 0.f, 0.f, 0.f},
          0, 0, 0);
    // one source record at a time, against the four target groups
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const float4 sd = rec[16 * b + 4 * grp + r];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const float inv = __builtin_amdgcn_rsqf(r2[c][r]);
        const float w = (inv * inv) * inv;
        a[c][0] = fmaf(w, sd.x, a[c][0]);
        a[c][1] = fmaf(w, sd.y, a[c][1]);
        a[c][2] = fmaf(w, sd.z, a[c][2]);
        a[c][3] = fmaf(w, sd.w, a[c][3]);
      }
    }
  }
}

// Adds the four lane groups' sums of each target through LDS (kq[0, 256),
// free once the last chunk's quads are read), gives the lane's own target
// (sum w g s) - t (sum w g) to the compensated sums, and clears a.
__device__ __forceinline__ void flush_f32r_mat(u32x4_t* kq, f32x4_t (&a)[4],
                                               const TargetR& t, Kahan3& acc) {
  const int lane = (int)(threadIdx.x & (kWave - 1));
  const int grp = lane >> 4, col = lane & 15;
  float4* red = (float4*)kq;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    red[4 * (16 * c + col) + grp] = __builtin_bit_cast(float4, a[c]);
    a[c] = f32x4_t{0.f, 0.f, 0.f, 0.f};
  }
  wave_sync();
  float4 f = red[4 * lane];
#pragma unroll
  for (int g2 = 1; g2 < 4; ++g2) {
    const float4 v = red[4 * lane + g2];
    f.x += v.x;
    f.y += v.y;
    f.z += v.z;
    f.w += v.w;
  }
  kahan_add(acc.sx, acc.cx, fmaf(-t.x, f.w, f.x));
  kahan_add(acc.sy, acc.cy, fmaf(-t.y, f.w, f.y));
  kahan_add(acc.sz, acc.cz, fmaf(-t.z, f.w, f.z));
  wave_sync();  // the sums are read before kq is staged again
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
template <bool NOMASK, bool RES, int BATCH>
__device__ __forceinline__ void chunk_f32r(const float4* stage,
                                           const float4* res,
                                           const TargetR& t, float eps2_s,
                                           Kahan3& acc) {
  constexpr int B = RES ? (BATCH > 1 ? BATCH / 2 : 1) : BATCH;
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
template <bool NOMASK, int BATCH>
__device__ __forceinline__ void sweep_f32r(float4* stage, float4* res,
                                           u32x4_t* kq, const u32x4_t (&Bt)[4],
                                           const HSrc* H, int64_t jlo,
                                           int64_t jhi,
                                           int64_t end, double ox, double oy,
                                           double oz, const TargetR& t,
                                           float eps2, Kahan3& acc) {
  const int lane = (int)(threadIdx.x & (kWave - 1));
  f32x4_t a[4];  // matrix-core class sums, cleared by each flush
#pragma unroll
  for (int c = 0; c < 4; ++c) a[c] = f32x4_t{0.f, 0.f, 0.f, 0.f};
  int pend = 0;  // matrix-core chunks in a[] (wave-uniform)
  for (int64_t j0 = jlo; j0 < jhi; j0 += kWave) {
    const int64_t j = j0 + lane;
    const bool valid = j < end;
    double dx = 0.0, dy = 0.0, dz = 0.0;
    float g = 0.f;
    if (valid) {
      const HSrc h = H[j];
      dx = h.x - ox;
      dy = h.y - oy;
      dz = h.z - oz;
      g = h.g;
    }
    const float sx = (float)dx, sy = (float)dy, sz = (float)dz;
    const SrcClass cl = classify_source(t, sx, sy, sz, g);
    // wave-uniform: every source valid and in the matrix-core class
    const bool all_mat = __ballot(!(valid && cl.mat)) == 0;
    if (all_mat) {
      mat_stage(kq, stage, sx, sy, sz, g, eps2);
      wave_sync();
      chunk_f32r_mat(kq, stage, Bt, a);
      wave_sync();
      ++pend;
    }
    // lanes sum 32 sources per target before the cross-lane addition; the
    // other classes reuse kq, so they flush first
    if (pend == 2 || (pend && !all_mat)) {
      flush_f32r_mat(kq, a, t, acc);
      pend = 0;
    }
    if (all_mat) continue;
    stage[lane] = {sx, sy, sz, g};
    const bool any_near = __ballot(valid && cl.near) != 0;  // wave-uniform
    if (any_near)
      res[lane] = {(float)(dx - (double)sx), (float)(dy - (double)sy),
                   (float)(dz - (double)sz), 0.f};
    wave_sync();
    if (any_near)
      chunk_f32r<NOMASK, true, BATCH>(stage, res, t, eps2, acc);
    else
      chunk_f32r<NOMASK, false, BATCH>(stage, res, t, eps2, acc);
    wave_sync();
  }
  if (pend) flush_f32r_mat(kq, a, t, acc);
}

// W waves per workgroup, loop shape S (see Shape).
template <int MODE, bool POT, int W, int S>
__global__ __launch_bounds__(kWave * W) void k_force(const Arena* Ap,
                                                     const WorkItem* items,
                                                     int nitems,
                                                     const SysDesc* sysd) {
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int idx = (int)blockIdx.x * W + wave;
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
    // per wave: 64 staged sources, then an area holding either the near
    // chunks' residuals (64) or the matrix-core class's K quads and cross-lane
    // sums (257 quads)
    __shared__ float4 lds[W][kWave + 257];
    float4* stage = lds[wave];
    u32x4_t* kq = (u32x4_t*)(lds[wave] + kWave);
    const int64_t ts = active ? slot : end - 1;  // in this tile
    const HSrc h0 = A->hs[src][ts];
    const double px0 = h0.x, py0 = h0.y, pz0 = h0.z;
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
    u32x4_t Bt[4];
    mat_targets(kq, t, Bt);
    Kahan3 acc = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    int64_t jlo, jhi;
    part_range<kWave>(off, n, nsplit, part, jlo, jhi);
    const float e2 = sd->eps2f;
    constexpr int B = Shape<S>::B;
    if (sd->flags & SYS_F32_SELF_FINITE)
      sweep_f32r<true, B>(stage, (float4*)kq, kq, Bt, A->hs[src], jlo, jhi,
                          end, ox, oy, oz, t, e2, acc);
    else
      sweep_f32r<false, B>(stage, (float4*)kq, kq, Bt, A->hs[src], jlo, jhi,
                           end, ox, oy, oz, t, e2, acc);
    ax = (double)acc.sx - (double)acc.cx;
    ay = (double)acc.sy - (double)acc.cy;
    az = (double)acc.sz - (double)acc.cz;
  } else if (MODE == MODE_F32C) {
    constexpr int U = Shape<S>::U, KC = Shape<S>::KC;
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
      constexpr int U = Shape<S>::U, KC = Shape<S>::KC;
      int64_t jlo, jhi;
      part_range<U>(off, n, nsplit, part, jlo, jhi);
      ChunkF64<POT, U> ch{i, end, xi, yi, zi, e2, ax, ay, az};
      drive<U, KC>(SrcF64{X, Y, Z, G}, jlo, jhi, t0, end, ch, [] {});
    }
  }
  if (!active) return;
  if (sd->nsplit > 1) {  // partial sum; k_combine finishes the target
    // partials in body order in every mode (in MODE_F32R a wave's 64 writes
    // are scattered, so k_combine can walk bodies with coalesced accesses)
    Part3* q = A->pp + ((int64_t)it.part * A->pstride + i);
    if (POT)
      q->x = ax;
    else
      *q = Part3{ax, ay, az};
    return;
  }
  if (POT) {
    A->phi[i] = ax;
    return;
  }
  epilogue<MODE>(A, sd, it.stage, src ^ 1, i, slot, ax, ay, az);
}

}  // namespace
}  // namespace nbody
