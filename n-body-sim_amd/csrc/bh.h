// Barnes-Hut gravity on the GPU: the octree of one system, rebuilt for every
// force evaluation, with the structure of the reference's octree
// (src/gravity/cpu/barnes_hut/octree.rs) so the traversal meets the same
// nodes in the same order and, in the strict arithmetic, gives the same bits.
//
// Build, all on the system's stream:
//   bbox    min/max of the source snapshot, root width = max extent * 1.001
//   codes   21-bit quantised positions interleaved (Morton), index payload
//   sort    radix sort of (code, index): a stable sort on the code, so ties
//           keep index order as the reference's sort of (code, index) does
//   chains  the nodes that start at sorted position k form a chain of
//           levels [leaf(k), top(k)], found from the highest differing bit
//           against k-1 and k+1; a scan of chain lengths gives every node
//           its preorder index, and the total (read back: the one
//           synchronisation of a build) sizes the node arrays
//   nodes   body range, width and empty child slots of each node; then each
//           node writes itself into its parent's slot for its octant
//   masses  centres of mass level by level from the leaves up, summed in the
//           reference's order (bodies of a leaf, then children in octant
//           order)
// Walk: one wavefront per 64 consecutive targets in Morton order, a shared
// stack of (node, lane mask) entries in LDS.  Strict walk: each lane applies
// the reference's per-target rules; the wave opens a node's children for the
// lanes that need them, pushing them in octant order and popping the last
// first, so every lane sees its own nodes in the reference's order.  Fast
// walk: the stack holds nodes to open; the children of a popped node are
// read together and judged at once (far nodes and one-body leaves add their
// term straight away), and only children some lane must open are pushed --
// the same set of terms, summed in another order.  The integrator stage
// epilogue runs at the end of either walk.
#pragma once
#include <hip/hip_runtime_api.h>

#include <cstdint>

#include "nbody_types.h"

namespace nbody {

// One octree node, 128 bytes: the fields every visit reads in the first 64
// (one scalar load), the child slots (read when the node is opened) after.
struct alignas(128) BhNode {
  double cx, cy, cz;      // centre of mass
  double gm;              // g * mass, the product the pair term takes
  double w2;              // width * width, as the opening test squares it
  int32_t start, count;   // bodies [start, start + count) in Morton order
  int32_t leaf, pad0;
  double mass;            // total mass (summed into the parent)
  int32_t child[8];       // preorder ids by octant, -1 where empty
  int32_t pad1[8];
};
static_assert(sizeof(BhNode) == 128, "node record layout");

// One source body in Morton order: position and g * m (32 bytes).
struct BhBody {
  double x, y, z, gm;
};

// Device arrays of one tree.
struct BhTree {
  const BhBody* body;      // sorted sources
  const int32_t* perm;     // sorted position -> body index
  const BhNode* node;      // preorder nodes
};

class BarnesHut {
 public:
  explicit BarnesHut(int64_t n);
  ~BarnesHut();
  BarnesHut(const BarnesHut&) = delete;
  BarnesHut& operator=(const BarnesHut&) = delete;

  // Builds the octree of snapshot `src` of system `sys` (n bodies at arena
  // offset desc.off), then walks it for every body and runs stage `stage`,
  // whose epilogue writes snapshot src ^ 1.  strict: the reference's
  // operation order, IEEE sqrt and division, no contraction; otherwise f64
  // with the rsq-based inverse square root and fused multiply-adds.
  void evaluate(hipStream_t s, bool strict, const Arena* d_arena,
                const SysDesc* d_sys, int sys, const SysDesc& desc, int stage,
                int src, double theta);

  // Nodes of the last build (synchronises).
  int64_t node_count(hipStream_t s);

 private:
  void reserve_nodes(int64_t cap);
  int64_t n_ = 0, cap_ = 0;
  void* mem_ = nullptr;    // per-body arrays
  void* nodes_ = nullptr;  // per-node arrays (grow only)
  int32_t* h_last_ = nullptr;  // pinned: last base and count of the scan
  uint64_t *keys_in_ = nullptr, *keys_ = nullptr;
  int32_t *idx_in_ = nullptr, *perm_ = nullptr;
  BhBody* body_ = nullptr;
  double* bm_ = nullptr;  // sorted masses (leaf centres of mass)
  int32_t *cnt_ = nullptr, *base_ = nullptr;
  int8_t *top_ = nullptr, *bot_ = nullptr;
  BhNode* node_ = nullptr;
  double* bbox_part_ = nullptr;  // 6 per block of the bounding-box pass
  double* bbox_ = nullptr;       // min x, min y, min z, root width
  void* temp_ = nullptr;
  size_t temp_bytes_ = 0;
};

}  // namespace nbody
