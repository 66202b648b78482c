// gfx950 kernels for the direct-sum gravity step.
//
// One wavefront owns one work item: 64 consecutive target bodies of one
// system, one target per lane.  The wave walks that system's source bodies in
// index order.  Source addresses are wave-uniform, so sources are read through
// address-space-4 pointers and land in SGPRs via s_load (the scalar cache
// broadcasts them; no LDS, no barriers).  Each lane keeps its own running
// sums, so the j order per target is the reference's sequential order
// (src/gravity/cpu/newton.rs:104-124): j = 0..n-1 with j == i skipped.
//
// The integrator stage for the system is fused into the epilogue of the same
// launch: the wave updates its own bodies' master state and writes their next
// source snapshot into the other half of a double buffer, so no target of the
// running launch ever reads a value this launch writes.
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "nbody_types.h"

namespace nbody {
namespace {

typedef __attribute__((address_space(4))) const double cf64;
typedef __attribute__((address_space(4))) const float cf32;
typedef __attribute__((address_space(4))) const int32_t ci32;

constexpr int kWavesPerBlock = kBlock / kWave;

// Wave-uniform load of a small descriptor through the constant address space
// (scalar loads into SGPRs).
template <typename T>
__device__ __forceinline__ T load_uniform(const T* p) {
  static_assert(sizeof(T) % 4 == 0, "dword-sized descriptors only");
  T v;
  const ci32* s = (const ci32*)p;
  int32_t* d = (int32_t*)&v;
#pragma unroll
  for (int k = 0; k < (int)(sizeof(T) / 4); ++k) d[k] = s[k];
  return v;
}

// 1/sqrt(x) in binary64: v_rsq_f64 seed refined by one step of the
// third-order series (1 - e)^(-1/2) = 1 + e/2 + 3e^2/8 + O(e^3) with
// e = 1 - x*y^2 (Householder iteration of order 3).
__device__ __forceinline__ double rsqrt_f64(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  const double e = fma(-x * y, y, 1.0);
  return fma(y * e, fma(e, 0.375, 0.5), y);
}

// Optimisation barrier on a register value (an empty asm statement, no
// instruction is emitted).  Applied to the double-buffered source registers at
// the end of a chunk it stops the compiler from re-deriving them as fresh
// loads at the top of the next chunk, so the next chunk's scalar loads are in
// flight while the current chunk computes.
template <typename T>
__device__ __forceinline__ void keep(T& v) {
  asm volatile("" : "+s"(v));
}

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

// Restricts the regions to the source range of one part: the system's
// chunks [0, nchunks) are cut into nsplit contiguous runs, part p taking
// chunks [p*nchunks/nsplit, (p+1)*nchunks/nsplit).  Parts cover every chunk
// once, in order.
template <int U>
__device__ __forceinline__ Bounds clip_part(Bounds b, int64_t off, int64_t n,
                                            int nsplit, int part) {
  if (nsplit <= 1) return b;
  const int64_t nchunks = (n + U - 1) / U;
  const int64_t lo = off + (nchunks * part / nsplit) * U;
  const int64_t hi = off + (nchunks * (part + 1) / nsplit) * U;
  auto clip = [&](int64_t& a, int64_t& e) {
    a = max(a, lo);
    e = min(e, hi);
    if (e < a) e = a;
  };
  clip(b.a0, b.a1);
  clip(b.b0, b.b1);
  clip(b.c0, b.c1);
  clip(b.d0, b.d1);
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

// ---- binary64, fast: rsq + refinement, fused multiply-add.  Sources for
// the next chunk are requested before the current chunk is computed.
template <bool MASK, bool POT>
__device__ __forceinline__ void sweep_f64(const cf64* X, const cf64* Y,
                                          const cf64* Z, const cf64* G,
                                          int64_t j0, int64_t j1, int64_t i,
                                          int64_t end, double xi, double yi,
                                          double zi, double eps2, double& ax,
                                          double& ay, double& az) {
  constexpr int U = 4;
  if (j0 >= j1) return;
  double cx[U], cy[U], cz[U], cg[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    cx[u] = X[j0 + u];
    cy[u] = Y[j0 + u];
    cz[u] = Z[j0 + u];
    cg[u] = G[j0 + u];
  }
  for (int64_t j = j0; j < j1; j += U) {
    // prefetch: j + 2U - 1 < end_up + U <= cap + kPad, always in bounds
    double nx[U], ny[U], nz[U], ng[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      nx[u] = X[j + U + u];
      ny[u] = Y[j + U + u];
      nz[u] = Z[j + U + u];
      ng[u] = G[j + U + u];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      double dx = cx[u] - xi;
      double dy = cy[u] - yi;
      double dz = cz[u] - zi;
      // masked sources contribute exactly zero, whatever values they hold
      const int64_t jj = j + u;
      const bool ok = !MASK || ((jj != i) && (jj < end));
      if (MASK) {
        dx = ok ? dx : 0.0;
        dy = ok ? dy : 0.0;
        dz = ok ? dz : 0.0;
      }
      const double r2 = fma(dz, dz, fma(dy, dy, fma(dx, dx, eps2)));
      const double inv = rsqrt_f64(r2);
      double k = POT ? cg[u] * inv : cg[u] * ((inv * inv) * inv);
      if (MASK) k = ok ? k : 0.0;
      if (POT) {
        ax += k;
      } else {
        ax = fma(k, dx, ax);
        ay = fma(k, dy, ay);
        az = fma(k, dz, az);
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      cx[u] = nx[u];
      cy[u] = ny[u];
      cz[u] = nz[u];
      cg[u] = ng[u];
      keep(cx[u]);
      keep(cy[u]);
      keep(cz[u]);
      keep(cg[u]);
    }
  }
}

// ---- binary32 pair terms, compensated accumulation.  Each chunk of U terms
// is summed in f32, then added to the running sum with Kahan's compensated
// summation (Kahan 1965), which keeps the running sum's rounding error at the
// level of one chunk sum instead of growing with n.
struct Kahan3 {
  float sx, sy, sz, cx, cy, cz;
};

__device__ __forceinline__ void kahan_add(float& s, float& c, float v) {
  const float y = v - c;
  const float t = s + y;
  c = (t - s) - y;
  s = t;
}

template <bool MASK>
__device__ __forceinline__ void sweep_f32(const cf32* X, const cf32* Y,
                                          const cf32* Z, const cf32* G,
                                          int64_t j0, int64_t j1, int64_t i,
                                          int64_t end, float xi, float yi,
                                          float zi, float eps2, Kahan3& acc) {
  constexpr int U = 8;
  if (j0 >= j1) return;
  float cx[U], cy[U], cz[U], cg[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    cx[u] = X[j0 + u];
    cy[u] = Y[j0 + u];
    cz[u] = Z[j0 + u];
    cg[u] = G[j0 + u];
  }
  for (int64_t j = j0; j < j1; j += U) {
    float nx[U], ny[U], nz[U], ng[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      nx[u] = X[j + U + u];
      ny[u] = Y[j + U + u];
      nz[u] = Z[j + U + u];
      ng[u] = G[j + U + u];
    }
    float px = 0.f, py = 0.f, pz = 0.f;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      float dx = cx[u] - xi;
      float dy = cy[u] - yi;
      float dz = cz[u] - zi;
      const int64_t jj = j + u;
      const bool ok = !MASK || ((jj != i) && (jj < end));
      if (MASK) {
        dx = ok ? dx : 0.f;
        dy = ok ? dy : 0.f;
        dz = ok ? dz : 0.f;
      }
      const float r2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, eps2)));
      const float inv = __builtin_amdgcn_rsqf(r2);
      float k = cg[u] * ((inv * inv) * inv);
      if (MASK) k = ok ? k : 0.f;
      px = fmaf(k, dx, px);
      py = fmaf(k, dy, py);
      pz = fmaf(k, dz, pz);
    }
    kahan_add(acc.sx, acc.cx, px);
    kahan_add(acc.sy, acc.cy, py);
    kahan_add(acc.sz, acc.cz, pz);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      cx[u] = nx[u];
      cy[u] = ny[u];
      cz[u] = nz[u];
      cg[u] = ng[u];
      keep(cx[u]);
      keep(cy[u]);
      keep(cz[u]);
      keep(cg[u]);
    }
  }
}

// Writes the next source snapshot of body i (f64 copy, or f32 re-centred).
template <int MODE>
__device__ __forceinline__ void put_snapshot(const Arena& A, int dst,
                                             int64_t i, const SysDesc& sd,
                                             double x, double y, double z) {
  if (MODE == MODE_F32C) {
    A.fx[dst][i] = (float)(x - sd.ox);
    A.fy[dst][i] = (float)(y - sd.oy);
    A.fz[dst][i] = (float)(z - sd.oz);
  } else {
    A.sx[dst][i] = x;
    A.sy[dst][i] = y;
    A.sz[dst][i] = z;
  }
}

// Integrator stage for body i, given its fresh acceleration (ax, ay, az).
// Expression trees follow the reference exactly; contraction is off so the
// stage arithmetic is bit-identical to the reference for identical inputs.
template <int MODE>
__device__ __forceinline__ void epilogue(const Arena& A, const SysDesc& sd,
                                         int stage, int dst, int64_t i,
                                         double ax, double ay, double az) {
#pragma clang fp contract(off)
  const double dt = sd.dt;
  switch (stage) {
    case STAGE_ACC: {
      A.ax[i] = ax;
      A.ay[i] = ay;
      A.az[i] = az;
      break;
    }
    case STAGE_EULER: {  // src/integrators/cpu/euler.rs:69-80
      const double vx = A.vx[i] + ax * dt;
      const double vy = A.vy[i] + ay * dt;
      const double vz = A.vz[i] + az * dt;
      const double x = A.x[i] + vx * dt;
      const double y = A.y[i] + vy * dt;
      const double z = A.z[i] + vz * dt;
      A.vx[i] = vx;
      A.vy[i] = vy;
      A.vz[i] = vz;
      A.x[i] = x;
      A.y[i] = y;
      A.z[i] = z;
      put_snapshot<MODE>(A, dst, i, sd, x, y, z);
      break;
    }
    case STAGE_VV_PRIME: {  // velocity_verlet.rs:82-86, first half only
      const double dt2 = dt * dt;
      A.ax[i] = ax;
      A.ay[i] = ay;
      A.az[i] = az;
      const double x = A.x[i] + ((A.vx[i] * dt) + ((0.5 * ax) * dt2));
      const double y = A.y[i] + ((A.vy[i] * dt) + ((0.5 * ay) * dt2));
      const double z = A.z[i] + ((A.vz[i] * dt) + ((0.5 * az) * dt2));
      put_snapshot<MODE>(A, dst, i, sd, x, y, z);
      break;
    }
    case STAGE_VV: {  // velocity_verlet.rs:82-86 then :97-101, a_n carried
      const double dt2 = dt * dt;
      const double px = A.ax[i], py = A.ay[i], pz = A.az[i];  // a_n
      double vx = A.vx[i], vy = A.vy[i], vz = A.vz[i];
      const double x = A.x[i] + ((vx * dt) + ((0.5 * px) * dt2));
      const double y = A.y[i] + ((vy * dt) + ((0.5 * py) * dt2));
      const double z = A.z[i] + ((vz * dt) + ((0.5 * pz) * dt2));
      vx = vx + ((0.5 * (px + ax)) * dt);
      vy = vy + ((0.5 * (py + ay)) * dt);
      vz = vz + ((0.5 * (pz + az)) * dt);
      A.x[i] = x;
      A.y[i] = y;
      A.z[i] = z;
      A.vx[i] = vx;
      A.vy[i] = vy;
      A.vz[i] = vz;
      A.ax[i] = ax;
      A.ay[i] = ay;
      A.az[i] = az;
      put_snapshot<MODE>(A, dst, i, sd, x + ((vx * dt) + ((0.5 * ax) * dt2)),
                         y + ((vy * dt) + ((0.5 * ay) * dt2)),
                         z + ((vz * dt) + ((0.5 * az) * dt2)));
      break;
    }
    // RK4 (src/integrators/cpu/runge_kutta.rs:149-281).  rk[0..2] running
    // sum of the position slopes, rk[3..5] of the velocity slopes, rk[6..8]
    // the next stage's position slope k_{s+1,r}.
    case STAGE_RK0: {
      const double h = dt / 2.0;
      const double vx = A.vx[i], vy = A.vy[i], vz = A.vz[i];
      const double k2x = vx + h * ax, k2y = vy + h * ay, k2z = vz + h * az;
      A.rk[0][i] = vx + (2.0 * k2x);
      A.rk[1][i] = vy + (2.0 * k2y);
      A.rk[2][i] = vz + (2.0 * k2z);
      A.rk[3][i] = ax;
      A.rk[4][i] = ay;
      A.rk[5][i] = az;
      A.rk[6][i] = k2x;
      A.rk[7][i] = k2y;
      A.rk[8][i] = k2z;
      put_snapshot<MODE>(A, dst, i, sd, A.x[i] + h * vx, A.y[i] + h * vy,
                         A.z[i] + h * vz);
      break;
    }
    case STAGE_RK1: {
      const double h = dt / 2.0;
      const double vx = A.vx[i], vy = A.vy[i], vz = A.vz[i];
      const double k3x = vx + h * ax, k3y = vy + h * ay, k3z = vz + h * az;
      const double k2x = A.rk[6][i], k2y = A.rk[7][i], k2z = A.rk[8][i];
      A.rk[0][i] = A.rk[0][i] + (2.0 * k3x);
      A.rk[1][i] = A.rk[1][i] + (2.0 * k3y);
      A.rk[2][i] = A.rk[2][i] + (2.0 * k3z);
      A.rk[3][i] = A.rk[3][i] + (2.0 * ax);
      A.rk[4][i] = A.rk[4][i] + (2.0 * ay);
      A.rk[5][i] = A.rk[5][i] + (2.0 * az);
      A.rk[6][i] = k3x;
      A.rk[7][i] = k3y;
      A.rk[8][i] = k3z;
      put_snapshot<MODE>(A, dst, i, sd, A.x[i] + h * k2x, A.y[i] + h * k2y,
                         A.z[i] + h * k2z);
      break;
    }
    case STAGE_RK2: {
      const double vx = A.vx[i], vy = A.vy[i], vz = A.vz[i];
      const double k3x = A.rk[6][i], k3y = A.rk[7][i], k3z = A.rk[8][i];
      A.rk[0][i] = A.rk[0][i] + (vx + dt * ax);
      A.rk[1][i] = A.rk[1][i] + (vy + dt * ay);
      A.rk[2][i] = A.rk[2][i] + (vz + dt * az);
      A.rk[3][i] = A.rk[3][i] + (2.0 * ax);
      A.rk[4][i] = A.rk[4][i] + (2.0 * ay);
      A.rk[5][i] = A.rk[5][i] + (2.0 * az);
      put_snapshot<MODE>(A, dst, i, sd, A.x[i] + dt * k3x, A.y[i] + dt * k3y,
                         A.z[i] + dt * k3z);
      break;
    }
    case STAGE_RK3: {
      const double s6 = dt / 6.0;
      const double x = A.x[i] + s6 * A.rk[0][i];
      const double y = A.y[i] + s6 * A.rk[1][i];
      const double z = A.z[i] + s6 * A.rk[2][i];
      A.vx[i] = A.vx[i] + s6 * (A.rk[3][i] + ax);
      A.vy[i] = A.vy[i] + s6 * (A.rk[4][i] + ay);
      A.vz[i] = A.vz[i] + s6 * (A.rk[5][i] + az);
      A.x[i] = x;
      A.y[i] = y;
      A.z[i] = z;
      put_snapshot<MODE>(A, dst, i, sd, x, y, z);
      break;
    }
    default:
      break;
  }
}

template <int MODE, bool POT>
__global__ __launch_bounds__(kBlock) void k_force(Arena A,
                                                  const WorkItem* items,
                                                  int nitems,
                                                  const SysDesc* sysd) {
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int idx = (int)blockIdx.x * kWavesPerBlock + wave;
  if (idx >= nitems) return;
  const WorkItem it = load_uniform(items + idx);
  const SysDesc sd = load_uniform(sysd + it.sys);
  const int64_t off = sd.off, n = sd.n, end = off + n;
  const int64_t t0 = off + (int64_t)it.tile * kWave;
  const int64_t i = t0 + (int64_t)(threadIdx.x & (kWave - 1));
  const bool active = i < end;
  const int64_t il = active ? i : end - 1;
  const int src = it.src;
  double ax = 0.0, ay = 0.0, az = 0.0;

  if (MODE == MODE_F32C) {
    const cf32* X = (const cf32*)A.fx[src];
    const cf32* Y = (const cf32*)A.fy[src];
    const cf32* Z = (const cf32*)A.fz[src];
    const cf32* G = (const cf32*)A.gm32;
    const float xi = A.fx[src][il], yi = A.fy[src][il], zi = A.fz[src][il];
    const float e2 = sd.eps2f;
    Kahan3 acc = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    const Bounds b =
        clip_part<8>(make_bounds<8>(off, n, t0), off, n, sd.nsplit, it.part);
    sweep_f32<false>(X, Y, Z, G, b.a0, b.a1, i, end, xi, yi, zi, e2, acc);
    sweep_f32<true>(X, Y, Z, G, b.b0, b.b1, i, end, xi, yi, zi, e2, acc);
    sweep_f32<false>(X, Y, Z, G, b.c0, b.c1, i, end, xi, yi, zi, e2, acc);
    sweep_f32<true>(X, Y, Z, G, b.d0, b.d1, i, end, xi, yi, zi, e2, acc);
    ax = (double)acc.sx - (double)acc.cx;
    ay = (double)acc.sy - (double)acc.cy;
    az = (double)acc.sz - (double)acc.cz;
  } else {
    const cf64* X = (const cf64*)A.sx[src];
    const cf64* Y = (const cf64*)A.sy[src];
    const cf64* Z = (const cf64*)A.sz[src];
    const cf64* G = (const cf64*)A.gm64;
    const double xi = A.sx[src][il], yi = A.sy[src][il], zi = A.sz[src][il];
    const double e2 = sd.eps2;
    const Bounds b =
        clip_part<4>(make_bounds<4>(off, n, t0), off, n, sd.nsplit, it.part);
    if (MODE == MODE_F64_STRICT) {
      sweep_strict<false>(X, Y, Z, G, b.a0, b.a1, i, end, xi, yi, zi, e2, ax,
                          ay, az);
      sweep_strict<true>(X, Y, Z, G, b.b0, b.b1, i, end, xi, yi, zi, e2, ax,
                         ay, az);
      sweep_strict<false>(X, Y, Z, G, b.c0, b.c1, i, end, xi, yi, zi, e2, ax,
                          ay, az);
      sweep_strict<true>(X, Y, Z, G, b.d0, b.d1, i, end, xi, yi, zi, e2, ax,
                         ay, az);
    } else {
      sweep_f64<false, POT>(X, Y, Z, G, b.a0, b.a1, i, end, xi, yi, zi, e2,
                            ax, ay, az);
      sweep_f64<true, POT>(X, Y, Z, G, b.b0, b.b1, i, end, xi, yi, zi, e2,
                           ax, ay, az);
      sweep_f64<false, POT>(X, Y, Z, G, b.c0, b.c1, i, end, xi, yi, zi, e2,
                            ax, ay, az);
      sweep_f64<true, POT>(X, Y, Z, G, b.d0, b.d1, i, end, xi, yi, zi, e2,
                           ax, ay, az);
    }
  }
  if (!active) return;
  if (sd.nsplit > 1) {  // partial sum; k_combine finishes the target
    const int64_t q = (int64_t)it.part * A.pstride + i;
    A.px[q] = ax;
    if (!POT) {
      A.py[q] = ay;
      A.pz[q] = az;
    }
    return;
  }
  if (POT) {
    A.phi[i] = ax;
    return;
  }
  epilogue<MODE>(A, sd, it.stage, src ^ 1, i, ax, ay, az);
}

// Adds the parts' partial sums of each target in part order (a fixed order,
// so results are reproducible) and runs the target's stage epilogue.
template <int MODE>
__global__ __launch_bounds__(kBlock) void k_combine(Arena A,
                                                    const WorkItem* items,
                                                    int nitems,
                                                    const SysDesc* sysd) {
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int idx = (int)blockIdx.x * kWavesPerBlock + wave;
  if (idx >= nitems) return;
  const WorkItem it = load_uniform(items + idx);
  const SysDesc sd = load_uniform(sysd + it.sys);
  const int64_t end = sd.off + sd.n;
  const int64_t i = sd.off + (int64_t)it.tile * kWave +
                    (int64_t)(threadIdx.x & (kWave - 1));
  if (i >= end) return;
  const int64_t st = A.pstride;
  double ax = A.px[i];
  if (it.stage == STAGE_POT) {
    for (int p = 1; p < sd.nsplit; ++p) ax = ax + A.px[p * st + i];
    A.phi[i] = ax;
    return;
  }
  double ay = A.py[i], az = A.pz[i];
  for (int p = 1; p < sd.nsplit; ++p) {
    ax = ax + A.px[p * st + i];
    ay = ay + A.py[p * st + i];
    az = az + A.pz[p * st + i];
  }
  epilogue<MODE>(A, sd, it.stage, it.src ^ 1, i, ax, ay, az);
}

// Builds source snapshot it.src of each body from the master state:
// stage 0 -> the positions themselves; STAGE_VV_PRIME -> the speculative
// next positions x + ((v*dt) + ((0.5*a)*dt^2)) from the carried a.  Also
// refreshes gm = g*m in f64 and f32.
__global__ __launch_bounds__(kBlock) void k_prepare(Arena A,
                                                    const WorkItem* items,
                                                    int nitems,
                                                    const SysDesc* sysd) {
#pragma clang fp contract(off)
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int idx = (int)blockIdx.x * kWavesPerBlock + wave;
  if (idx >= nitems) return;
  const WorkItem it = load_uniform(items + idx);
  const SysDesc sd = load_uniform(sysd + it.sys);
  const int64_t end = sd.off + sd.n;
  const int64_t i = sd.off + (int64_t)it.tile * kWave +
                    (int64_t)(threadIdx.x & (kWave - 1));
  if (i >= end) return;
  const double gm = sd.g * A.m[i];
  A.gm64[i] = gm;
  A.gm32[i] = (float)gm;
  double x = A.x[i], y = A.y[i], z = A.z[i];
  if (it.stage == STAGE_VV_PRIME) {
    const double dt = sd.dt, dt2 = dt * dt;
    x = x + ((A.vx[i] * dt) + ((0.5 * A.ax[i]) * dt2));
    y = y + ((A.vy[i] * dt) + ((0.5 * A.ay[i]) * dt2));
    z = z + ((A.vz[i] * dt) + ((0.5 * A.az[i]) * dt2));
  }
  const int d = it.src;
  A.sx[d][i] = x;
  A.sy[d][i] = y;
  A.sz[d][i] = z;
  A.fx[d][i] = (float)(x - sd.ox);
  A.fy[d][i] = (float)(y - sd.oy);
  A.fz[d][i] = (float)(z - sd.oz);
}

// Per-system f64 diagnostics, one workgroup per listed system, fixed
// summation order (thread t takes bodies t, t+256, ...; then a fixed LDS
// tree), so results are reproducible bit for bit.  Needs phi from a
// potential pass.  out[s*9 + DIAG_*].
__global__ __launch_bounds__(kBlock) void k_diag(Arena A, const SysDesc* sysd,
                                                 const int32_t* list,
                                                 double* out) {
  __shared__ double red[DIAG_COUNT][kBlock];
  const int s = list[blockIdx.x];
  const SysDesc sd = sysd[s];
  double acc[DIAG_COUNT];
#pragma unroll
  for (int q = 0; q < DIAG_COUNT; ++q) acc[q] = 0.0;
  for (int64_t k = threadIdx.x; k < sd.n; k += kBlock) {
    const int64_t i = sd.off + k;
    const double m = A.m[i];
    const double vx = A.vx[i], vy = A.vy[i], vz = A.vz[i];
    acc[DIAG_K] += 0.5 * m * ((vx * vx + vy * vy) + vz * vz);
    acc[DIAG_U] += m * A.phi[i];
    acc[DIAG_PX] += m * vx;
    acc[DIAG_PY] += m * vy;
    acc[DIAG_PZ] += m * vz;
    acc[DIAG_MX] += m * A.x[i];
    acc[DIAG_MY] += m * A.y[i];
    acc[DIAG_MZ] += m * A.z[i];
    acc[DIAG_M] += m;
  }
#pragma unroll
  for (int q = 0; q < DIAG_COUNT; ++q) red[q][threadIdx.x] = acc[q];
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
#pragma unroll
      for (int q = 0; q < DIAG_COUNT; ++q)
        red[q][threadIdx.x] += red[q][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x < DIAG_COUNT) {
    double v = red[threadIdx.x][0];
    if (threadIdx.x == DIAG_U) v = -0.5 * v;  // each pair counted twice
    out[(int64_t)blockIdx.x * DIAG_COUNT + threadIdx.x] = v;
  }
}

// Writes zeros; used instead of hipMemsetAsync so it can sit in a graph.
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

}  // namespace

hipError_t launch_force(int mode, const Arena& A, const WorkItem* items,
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
    case MODE_F32C:
      hipLaunchKernelGGL((k_force<MODE_F32C, false>), grid, block, 0, s, A,
                         items, nitems, sys);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_combine(int mode, const Arena& A, const WorkItem* items,
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
    case MODE_F32C:
      hipLaunchKernelGGL(k_combine<MODE_F32C>, grid, block, 0, s, A, items,
                         nitems, sys);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_potential(const Arena& A, const WorkItem* items, int nitems,
                            const SysDesc* sys, hipStream_t s) {
  if (nitems <= 0) return hipSuccess;
  hipLaunchKernelGGL((k_force<MODE_F64, true>), dim3(blocks_for(nitems)),
                     dim3(kBlock), 0, s, A, items, nitems, sys);
  return hipGetLastError();
}

hipError_t launch_prepare(const Arena& A, const WorkItem* items, int nitems,
                          const SysDesc* sys, hipStream_t s) {
  if (nitems <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_prepare, dim3(blocks_for(nitems)), dim3(kBlock), 0, s,
                     A, items, nitems, sys);
  return hipGetLastError();
}

hipError_t launch_diag(const Arena& A, const SysDesc* sys, const int32_t* list,
                       int nlist, double* out, hipStream_t s) {
  if (nlist <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_diag, dim3(nlist), dim3(kBlock), 0, s, A, sys, list,
                     out);
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
