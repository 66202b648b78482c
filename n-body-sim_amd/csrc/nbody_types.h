// Types shared by the host engine and the gfx950 kernels.
//
// Memory layout in HBM (one arena per engine, structure of arrays, f64 state):
//   m, gm64 (= G*m), gm32         bodies' masses, G folded in once per system
//   x, y, z, vx, vy, vz           master state, binary64
//   sx/sy/sz[2]                   f64 source snapshot, double-buffered
//   perm, iperm                   MODE_F32R body order (Hilbert curve)
//                                 and its inverse
//   hs[2]                         MODE_F32R: the snapshot and g*m in that
//                                 order, one 32-byte record per slot, so
//                                 sources read coalesced and a body's
//                                 scattered update is one record
//   ax, ay, az                    carried acceleration (velocity Verlet) or the
//                                 output of a plain force evaluation
//   rk[9]                         RK4 running sums and next-stage velocity
//   xs[3]                         staged RK4: positions at the step's start
//   phi                           per-body potential (energy diagnostics)
//   pp[kMaxSplit]                 partial sums of split source ranges,
//                                 one 24-byte record per body and part
//                                 (body order in every mode)
// Systems occupy disjoint [off, off+n) ranges of every array.  Every array has
// kPad elements of zeroed slack past the arena end so a source chunk that
// straddles the last body stays in bounds.
#pragma once
#include <stdint.h>

namespace nbody {

constexpr int kWave = 64;        // gfx950 wavefront
#ifndef NBODY_BLOCK
#define NBODY_BLOCK 64
#endif
// threads per workgroup: one wave per workgroup measured best (finer-grained
// dispatch of uneven work items); each wave runs one work item
constexpr int kBlock = NBODY_BLOCK;
constexpr int kPad = 128;        // slack elements after the arena
constexpr int kMaxSplit = 32;    // source-range parts per target tile
// Force-kernel variants of the tuned modes (MODE_F64, MODE_F32C, MODE_F32R):
// variant = 3 * w + shape, w indexing 1, 2 or 4 waves per workgroup (one
// work item per wave) and shape the loop shape of force.h (chunk and group
// sizes, LDS batch).  Variant 0 is one wave per workgroup, shape 0.
constexpr int kForceVariants = 9;

enum Mode : int32_t {            // force arithmetic (public NBODY_PRECISION_*)
  MODE_F64_STRICT = 0,           // reference operation order, no contraction,
                                 // correctly rounded sqrt and divide
  MODE_F64 = 1,                  // f64, rsq + refinement, fused multiply-add
  MODE_F32C = 2,                 // f64 positions, each difference rounded
                                 // once to f32, f32 pair terms, compensated
                                 // accumulation
  MODE_F32R = 3,                 // bodies in Hilbert order, coordinates
                                 // relative to each wave's middle target
                                 // (f32 value + f32 residual), sources staged
                                 // in LDS, f32 pair terms (residuals added
                                 // for near chunks), compensated accumulation
};

enum Integ : int32_t { INTEG_EULER = 0, INTEG_VV = 1, INTEG_RK4 = 2 };

// Force law of an engine's systems: the direct sum, or the Barnes-Hut octree
// (one tree per system, rebuilt for every evaluation).
enum Gravity : int32_t { GRAV_DIRECT = 0, GRAV_BARNES_HUT = 1 };

// What the epilogue of one force evaluation does for a system.
enum Stage : int32_t {
  STAGE_ACC = 0,       // store a into ax/ay/az; state unchanged
  STAGE_EULER = 1,     // v += a*dt; x += v*dt; snapshot <- x
  STAGE_VV_PRIME = 2,  // carry a; snapshot <- x + ((v*dt) + ((0.5*a)*dt2))
  STAGE_VV = 3,        // finish step (x, v), carry a, snapshot <- next x
  STAGE_RK0 = 4,       // RK4 stages 1..4
  STAGE_RK1 = 5,
  STAGE_RK2 = 6,
  STAGE_RK3 = 7,
  STAGE_POT = 8,       // store potential sum_j G m_j / sqrt(r2 + eps2) in phi
};

struct SysDesc {       // one per resident system (48 B)
  int64_t off;         // first body in the arena
  int32_t n;           // body count
  int32_t integ;       // Integ
  double dt;
  double eps2;         // softening squared, f64
  double g;            // gravitational constant
  float eps2f;         // eps2 rounded to f32
  int16_t nsplit;      // source-range parts per tile (1: no split)
  int16_t flags;       // SysFlag bits
};
static_assert(sizeof(SysDesc) == 48, "SysDesc layout");

enum SysFlag : int16_t {
  // eps2f > 0 and the f32 self term g*m_i * rsq(eps2f)^3 is finite for every
  // body of the system, so in MODE_F32R a target's own pair (dx = dy = dz =
  // 0) contributes exactly zero without a mask.  Without it (zero softening,
  // or a softening so small that g*m*eps^-3 overflows f32) MODE_F32R masks
  // the pairs with a zero difference.
  SYS_F32_SELF_FINITE = 1,
};

struct WorkItem {      // one wave: 64 consecutive targets of one system
  int32_t sys;         // index into the SysDesc table
  int32_t tile;        // targets [off + 64*tile, off + 64*tile + 64) clipped to n
  int16_t stage;       // Stage
  int8_t src;          // snapshot buffer read (0/1); the epilogue writes src^1
  int8_t part;         // source-range part (0 .. nsplit-1)
  int32_t pad_ = 0;
};
static_assert(sizeof(WorkItem) == 16, "WorkItem layout");

// A run of work items of one system, expanded on the device (k_expand):
// item first + k is (sys, tile k / parts, stage, src, part k % parts) for
// k < tiles * parts, in the order the host would list them.
struct ItemSpan {
  int32_t first;       // index of the run's first item
  int32_t sys;
  int32_t tiles;
  int16_t stage;
  int8_t src;
  int8_t parts;        // source-range parts (1 for combine items)
};
static_assert(sizeof(ItemSpan) == 16, "ItemSpan layout");

// One body's partial acceleration sum over one source part (k_force ->
// k_combine); in the potential pass only x is used.
struct Part3 {
  double x, y, z;
};
static_assert(sizeof(Part3) == 24, "Part3 layout");

// One MODE_F32R source in slot order: its snapshot position and g*m in f32.
struct HSrc {
  double x, y, z;
  float g;
  float pad_;
};
static_assert(sizeof(HSrc) == 32, "HSrc layout");

struct Arena {         // device pointers, passed by value to every kernel
  double *m, *gm64;
  float *gm32;
  double *x, *y, *z, *vx, *vy, *vz;
  double *sx[2], *sy[2], *sz[2];
  double *ax, *ay, *az;
  double *rk[9];
  double *xs[3];        // staged RK4: positions at the start of the step
  double *phi;
  Part3* pp;            // per-part partial sums: [part * pstride + i]
  int32_t* perm;        // MODE_F32R order: slot off + k holds body
                        // off + perm[off + k] (Hilbert order)
  int32_t* iperm;       // its inverse: body off + v sits at slot
                        // off + iperm[off + v]
  HSrc* hs[2];          // MODE_F32R: sx/sy/sz[b] and gm32 in slot order
  int64_t pstride;     // cap + kPad
  int64_t cap;         // bodies (excluding the kPad slack)
};

// One system of a target-ordering pass (MODE_F32R): its bodies' Hilbert
// codes are keyed (rank << 48 | code) so one sort orders every system.
struct SortSys {
  int64_t off;         // arena offset
  int32_t n, base;     // bodies; first slot in the pass's key arrays
  double ox, oy, oz;   // quantisation origin: position 0.5 of each axis
  double ix, iy, iz;   // per-axis scale: 1 / the box's width along the axis
};

// One integrator stage applied on its own to the acceleration arrays (ax,
// ay, az), for callers that evaluate the forces separately (the public
// NBODY_STAGE_* values).
enum StageOp : int32_t {
  OP_EULER = 0,     // v += a*dt; x += v*dt; one step done
  OP_VV_DRIFT = 1,  // keep a as a_n (rk[0..2]); x += (v*dt) + ((0.5*a)*dt2)
  OP_VV_KICK = 2,   // v += (0.5*(a_n + a))*dt; one step done
  OP_RK4_1 = 3,     // after a1 = F(x0): xs <- x0, sums, x <- x0 + h*v
  OP_RK4_2 = 4,     // after a2: x <- x0 + h*k2r
  OP_RK4_3 = 5,     // after a3: x <- x0 + dt*k3r
  OP_RK4_4 = 6,     // after a4: x, v <- the step's result; one step done
};

// Per-system diagnostics written by the reduction kernel (9 doubles).
enum Diag : int32_t {
  DIAG_K = 0, DIAG_U = 1, DIAG_PX = 2, DIAG_PY = 3, DIAG_PZ = 4,
  DIAG_MX = 5, DIAG_MY = 6, DIAG_MZ = 7, DIAG_M = 8, DIAG_COUNT = 9
};

}  // namespace nbody
