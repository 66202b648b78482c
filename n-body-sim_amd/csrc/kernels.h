// Host-side launchers for the kernels in kernels.hip.
#pragma once
#include <hip/hip_runtime_api.h>

#include "nbody_types.h"

namespace nbody {

// One force evaluation over a work list, with each item's integrator stage
// fused into the epilogue.  mode: Mode; variant: as above (MODE_F64_STRICT
// has one kernel and ignores it).  A, items, sys: device pointers.
hipError_t launch_force(int mode, int variant, const Arena* A,
                        const WorkItem* items, int nitems, const SysDesc* sys,
                        hipStream_t s);
// The per-mode variant launchers (force_variants.hip, one object per mode).
hipError_t launch_force_f64(int variant, const Arena* A, const WorkItem* items,
                            int nitems, const SysDesc* sys, hipStream_t s);
hipError_t launch_force_f32c(int variant, const Arena* A, const WorkItem* items,
                             int nitems, const SysDesc* sys, hipStream_t s);
hipError_t launch_force_f32r(int variant, const Arena* A, const WorkItem* items,
                             int nitems, const SysDesc* sys, hipStream_t s);
// Sums the partial results of split source ranges in part order, then runs
// each target's stage epilogue (or stores its potential for STAGE_POT).
hipError_t launch_combine(int mode, const Arena* A, const WorkItem* items,
                          int nitems, const SysDesc* sys, hipStream_t s);
// Per-body softened potential sum_j g*m_j / sqrt(r^2 + eps2) into A.phi (f64).
hipError_t launch_potential(const Arena* A, const WorkItem* items, int nitems,
                            const SysDesc* sys, hipStream_t s);
// Snapshot (and gm) rebuild from the master state.
hipError_t launch_prepare(const Arena* A, const WorkItem* items, int nitems,
                          const SysDesc* sys, hipStream_t s);
// One integrator stage (StageOp) on the acceleration arrays of the listed
// tiles' bodies.
hipError_t launch_stage(const Arena* A, const WorkItem* items, int nitems,
                        const SysDesc* sys, int op, hipStream_t s);
// Per-system diagnostics (Diag layout), one workgroup per listed system.
hipError_t launch_diag(const Arena* A, const SysDesc* sys, const int32_t* list,
                       int nlist, double* out, hipStream_t s);
// Packs x, y, z (and vx, vy, vz when nfields == 6) of listed systems into
// stage[field * total + row]; each item's pad_ holds its system's base row.
hipError_t launch_gather(const Arena* A, const WorkItem* items, int nitems,
                         const SysDesc* sys, double* stage, int64_t total,
                         int nfields, hipStream_t s);
// MODE_F32R body order: keys (rank << 48 | Hilbert index) and local indices
// of the listed systems' bodies, a radix sort (sort_pairs: hipCUB; query the
// temporary size with tmp == nullptr), then perm from the sorted values.
// launch_sort_scatter also rebuilds the slot-ordered snapshot copy (hs: the
// positions and g*m) from perm; with vals == nullptr it keeps perm and only
// rebuilds the copy (after the body-order snapshot or g*m was rewritten).
hipError_t launch_sort_keys(const Arena* A, const SortSys* tab, int nsys,
                            int max_tiles, uint64_t* keys, int32_t* vals,
                            hipStream_t s);
hipError_t launch_sort_scatter(const Arena* A, const SortSys* tab, int nsys,
                               int max_tiles, const int32_t* vals,
                               hipStream_t s);
hipError_t sort_pairs(void* tmp, size_t& bytes, const uint64_t* kin,
                      uint64_t* kout, const int32_t* vin, int32_t* vout, int n,
                      int end_bit, hipStream_t s);
// Expands nspans item spans (covering items [0, total) in order) into out.
hipError_t launch_expand(const ItemSpan* spans, int nspans, int total,
                         WorkItem* out, hipStream_t s);
// Zeroes bodies [off, off + n) of the arrays named by what (ClearWhat).
enum ClearWhat : int { CLEAR_RETIRED = 0, CLEAR_ACC = 1 };
hipError_t launch_clear_range(const Arena* A, int64_t off, int64_t n, int what,
                              hipStream_t s);
hipError_t launch_probe(int* out, hipStream_t s);

}  // namespace nbody
