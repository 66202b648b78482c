// Device code shared by the gfx950 kernel files: constant-address-space
// views of the arena and system table, wave-uniform loads, the f64 inverse
// square root, and the integrator stage epilogue that every force kernel
// runs for its targets.
#pragma once
#include <hip/hip_runtime.h>

#include "nbody_types.h"

namespace nbody {
namespace {

typedef __attribute__((address_space(4))) const double cf64;
typedef __attribute__((address_space(4))) const float cf32;
typedef __attribute__((address_space(4))) const int32_t ci32;
// Arena and system descriptors are read through constant-address-space
// pointers at the point of use: each field becomes a scalar load where it is
// needed instead of a kernel-argument load kept live in SGPRs for the whole
// kernel (which would crowd out the source registers of the hot loops).
typedef __attribute__((address_space(4))) const Arena CArena;
typedef __attribute__((address_space(4))) const SysDesc CSys;

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

// x^(-3/2) in binary64 with the refinement folded into the cube: with the
// v_rsq_f64 seed y and e = 1 - x*y^2 (|e| about 2^-22),
//   (y (1 - e)^(-1/2))^3 = y^3 (1 - e)^(-3/2) = y^3 (1 + 1.5 e + O(e^2)),
// so the cube takes the correction as one factor instead of refining y
// first (the dropped e^2 term is below 2^-43 relative).  Six f64 operations;
// e comes from one fused multiply-add against the inline constant 1.0, and
// the factor needs only one non-inline constant (one scalar operand).
__device__ __forceinline__ double inv_cube_f64(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  const double u = y * y;
  const double e = fma(-x, u, 1.0);
  const double c = fma(e, 1.5, 1.0);
  return (y * u) * c;
}

// Writes the next source snapshot of body i (f64 positions, every mode);
// MODE_F32R also writes it at the body's slot in Hilbert order.
template <int MODE>
__device__ __forceinline__ void put_snapshot(const CArena* A, int dst,
                                             int64_t i, int64_t slot,
                                             double x, double y, double z) {
  A->sx[dst][i] = x;
  A->sy[dst][i] = y;
  A->sz[dst][i] = z;
  if (MODE == MODE_F32R) {  // one 24-byte piece of the body's slot record
    HSrc* h = A->hs[dst] + slot;
    h->x = x;
    h->y = y;
    h->z = z;
  }
}

// Integrator stage for body i (at `slot` in MODE_F32R's Hilbert order; any
// value in the other modes), given its fresh acceleration (ax, ay, az).
// Expression trees follow the reference exactly; contraction is off so the
// stage arithmetic is bit-identical to the reference for identical inputs.
template <int MODE>
__device__ __forceinline__ void epilogue(const CArena* A, const CSys* sd,
                                         int stage, int dst, int64_t i,
                                         int64_t slot, double ax, double ay,
                                         double az) {
#pragma clang fp contract(off)
  const double dt = sd->dt;
  switch (stage) {
    case STAGE_ACC: {
      A->ax[i] = ax;
      A->ay[i] = ay;
      A->az[i] = az;
      break;
    }
    case STAGE_EULER: {  // src/integrators/cpu/euler.rs:69-80
      const double vx = A->vx[i] + ax * dt;
      const double vy = A->vy[i] + ay * dt;
      const double vz = A->vz[i] + az * dt;
      const double x = A->x[i] + vx * dt;
      const double y = A->y[i] + vy * dt;
      const double z = A->z[i] + vz * dt;
      A->vx[i] = vx;
      A->vy[i] = vy;
      A->vz[i] = vz;
      A->x[i] = x;
      A->y[i] = y;
      A->z[i] = z;
      put_snapshot<MODE>(A, dst, i, slot, x, y, z);
      break;
    }
    case STAGE_VV_PRIME: {  // velocity_verlet.rs:82-86, first half only
      const double dt2 = dt * dt;
      A->ax[i] = ax;
      A->ay[i] = ay;
      A->az[i] = az;
      const double x = A->x[i] + ((A->vx[i] * dt) + ((0.5 * ax) * dt2));
      const double y = A->y[i] + ((A->vy[i] * dt) + ((0.5 * ay) * dt2));
      const double z = A->z[i] + ((A->vz[i] * dt) + ((0.5 * az) * dt2));
      put_snapshot<MODE>(A, dst, i, slot, x, y, z);
      break;
    }
    case STAGE_VV: {  // velocity_verlet.rs:82-86 then :97-101, a_n carried
      const double dt2 = dt * dt;
      const double px = A->ax[i], py = A->ay[i], pz = A->az[i];  // a_n
      double vx = A->vx[i], vy = A->vy[i], vz = A->vz[i];
      const double x = A->x[i] + ((vx * dt) + ((0.5 * px) * dt2));
      const double y = A->y[i] + ((vy * dt) + ((0.5 * py) * dt2));
      const double z = A->z[i] + ((vz * dt) + ((0.5 * pz) * dt2));
      vx = vx + ((0.5 * (px + ax)) * dt);
      vy = vy + ((0.5 * (py + ay)) * dt);
      vz = vz + ((0.5 * (pz + az)) * dt);
      A->x[i] = x;
      A->y[i] = y;
      A->z[i] = z;
      A->vx[i] = vx;
      A->vy[i] = vy;
      A->vz[i] = vz;
      A->ax[i] = ax;
      A->ay[i] = ay;
      A->az[i] = az;
      put_snapshot<MODE>(A, dst, i, slot, x + ((vx * dt) + ((0.5 * ax) * dt2)),
                         y + ((vy * dt) + ((0.5 * ay) * dt2)),
                         z + ((vz * dt) + ((0.5 * az) * dt2)));
      break;
    }
    // RK4 (src/integrators/cpu/runge_kutta.rs:149-281).  rk[0..2] running
    // sum of the position slopes, rk[3..5] of the velocity slopes, rk[6..8]
    // the next stage's position slope k_{s+1,r}.
    case STAGE_RK0: {
      const double h = dt / 2.0;
      const double vx = A->vx[i], vy = A->vy[i], vz = A->vz[i];
      const double k2x = vx + h * ax, k2y = vy + h * ay, k2z = vz + h * az;
      A->rk[0][i] = vx + (2.0 * k2x);
      A->rk[1][i] = vy + (2.0 * k2y);
      A->rk[2][i] = vz + (2.0 * k2z);
      A->rk[3][i] = ax;
      A->rk[4][i] = ay;
      A->rk[5][i] = az;
      A->rk[6][i] = k2x;
      A->rk[7][i] = k2y;
      A->rk[8][i] = k2z;
      put_snapshot<MODE>(A, dst, i, slot, A->x[i] + h * vx, A->y[i] + h * vy,
                         A->z[i] + h * vz);
      break;
    }
    case STAGE_RK1: {
      const double h = dt / 2.0;
      const double vx = A->vx[i], vy = A->vy[i], vz = A->vz[i];
      const double k3x = vx + h * ax, k3y = vy + h * ay, k3z = vz + h * az;
      const double k2x = A->rk[6][i], k2y = A->rk[7][i], k2z = A->rk[8][i];
      A->rk[0][i] = A->rk[0][i] + (2.0 * k3x);
      A->rk[1][i] = A->rk[1][i] + (2.0 * k3y);
      A->rk[2][i] = A->rk[2][i] + (2.0 * k3z);
      A->rk[3][i] = A->rk[3][i] + (2.0 * ax);
      A->rk[4][i] = A->rk[4][i] + (2.0 * ay);
      A->rk[5][i] = A->rk[5][i] + (2.0 * az);
      A->rk[6][i] = k3x;
      A->rk[7][i] = k3y;
      A->rk[8][i] = k3z;
      put_snapshot<MODE>(A, dst, i, slot, A->x[i] + h * k2x, A->y[i] + h * k2y,
                         A->z[i] + h * k2z);
      break;
    }
    case STAGE_RK2: {
      const double vx = A->vx[i], vy = A->vy[i], vz = A->vz[i];
      const double k3x = A->rk[6][i], k3y = A->rk[7][i], k3z = A->rk[8][i];
      A->rk[0][i] = A->rk[0][i] + (vx + dt * ax);
      A->rk[1][i] = A->rk[1][i] + (vy + dt * ay);
      A->rk[2][i] = A->rk[2][i] + (vz + dt * az);
      A->rk[3][i] = A->rk[3][i] + (2.0 * ax);
      A->rk[4][i] = A->rk[4][i] + (2.0 * ay);
      A->rk[5][i] = A->rk[5][i] + (2.0 * az);
      put_snapshot<MODE>(A, dst, i, slot, A->x[i] + dt * k3x, A->y[i] + dt * k3y,
                         A->z[i] + dt * k3z);
      break;
    }
    case STAGE_RK3: {
      const double s6 = dt / 6.0;
      const double x = A->x[i] + s6 * A->rk[0][i];
      const double y = A->y[i] + s6 * A->rk[1][i];
      const double z = A->z[i] + s6 * A->rk[2][i];
      A->vx[i] = A->vx[i] + s6 * (A->rk[3][i] + ax);
      A->vy[i] = A->vy[i] + s6 * (A->rk[4][i] + ay);
      A->vz[i] = A->vz[i] + s6 * (A->rk[5][i] + az);
      A->x[i] = x;
      A->y[i] = y;
      A->z[i] = z;
      put_snapshot<MODE>(A, dst, i, slot, x, y, z);
      break;
    }
    default:
      break;
  }
}

}  // namespace
}  // namespace nbody
