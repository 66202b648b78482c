// gfx950 Barnes-Hut: octree build and walk (bh.h describes the scheme).
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include "../../include/nbody.h"
#include "bh.h"
#include "device.h"
#include "engine.h"

namespace nbody {
namespace {

constexpr int kDepth = 21;                // octree.rs MAX_DEPTH
constexpr int kStack = 8 * (kDepth + 2);  // > 1 + 7*21 + 8 entries ever live
constexpr int kThreads = 256;
constexpr int kBboxBlocks = 256;

// ---- bounding box (octree.rs:434-468) ----------------------------------------
__global__ __launch_bounds__(kThreads) void k_bh_bbox_part(const Arena* Ap,
                                                           int64_t off, int n,
                                                           int src,
                                                           double* part) {
  const CArena* A = (const CArena*)Ap;
  const double *X = A->sx[src], *Y = A->sy[src], *Z = A->sz[src];
  __shared__ double s[6][kThreads];
  double v[6] = {HUGE_VAL, HUGE_VAL, HUGE_VAL, -HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n;
       i += gridDim.x * kThreads) {
    const double p[3] = {X[off + i], Y[off + i], Z[off + i]};
    for (int c = 0; c < 3; ++c) {
      v[c] = p[c] < v[c] ? p[c] : v[c];
      v[3 + c] = p[c] > v[3 + c] ? p[c] : v[3 + c];
    }
  }
  for (int c = 0; c < 6; ++c) s[c][threadIdx.x] = v[c];
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int c = 0; c < 3; ++c) {
        const double a = s[c][threadIdx.x + w], b = s[3 + c][threadIdx.x + w];
        s[c][threadIdx.x] = a < s[c][threadIdx.x] ? a : s[c][threadIdx.x];
        s[3 + c][threadIdx.x] = b > s[3 + c][threadIdx.x] ? b : s[3 + c][threadIdx.x];
      }
    __syncthreads();
  }
  if (threadIdx.x < 6) part[blockIdx.x * 6 + threadIdx.x] = s[threadIdx.x][0];
}

// Minimum corner and root width = max extent * 1.001 (find_bounding_box).
__global__ __launch_bounds__(64) void k_bh_bbox_final(const double* part,
                                                      int nparts,
                                                      double* bbox) {
#pragma clang fp contract(off)
  __shared__ double s[6][64];
  double v[6] = {HUGE_VAL, HUGE_VAL, HUGE_VAL, -HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
  for (int p = threadIdx.x; p < nparts; p += 64)
    for (int c = 0; c < 3; ++c) {
      const double a = part[p * 6 + c], b = part[p * 6 + 3 + c];
      v[c] = a < v[c] ? a : v[c];
      v[3 + c] = b > v[3 + c] ? b : v[3 + c];
    }
  for (int c = 0; c < 6; ++c) s[c][threadIdx.x] = v[c];
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int t = 1; t < 64; ++t)
    for (int c = 0; c < 3; ++c) {
      v[c] = s[c][t] < v[c] ? s[c][t] : v[c];
      v[3 + c] = s[3 + c][t] > v[3 + c] ? s[3 + c][t] : v[3 + c];
    }
  const double dx = v[3] - v[0], dy = v[4] - v[1], dz = v[5] - v[2];
  double w = dx < dy ? dy : dx;  // f64::max on finite values
  w = w < dz ? dz : w;
  bbox[0] = v[0];
  bbox[1] = v[1];
  bbox[2] = v[2];
  bbox[3] = w * 1.001;
}

// ---- Morton codes (octree.rs:470-579) -----------------------------------------
__device__ __forceinline__ uint64_t quantize(double v, double mn, double w) {
#pragma clang fp contract(off)
  const double q = ((v - mn) / w) * 2097152.0;
  if (!(q >= 0.0)) return 0;  // `as u64` saturates: NaN and negatives to 0
  if (q >= 18446744073709551616.0) return ~0ull;
  return (uint64_t)q;
}

__device__ __forceinline__ uint64_t spread(uint64_t x) {
  x &= 0x1fffffull;
  x = (x | (x << 32)) & 0x1f00000000ffffull;
  x = (x | (x << 16)) & 0x1f0000ff0000ffull;
  x = (x | (x << 8)) & 0x100f00f00f00f00full;
  x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
  x = (x | (x << 2)) & 0x1249249249249249ull;
  return x;
}

__global__ __launch_bounds__(kThreads) void k_bh_codes(const Arena* Ap,
                                                       int64_t off, int n,
                                                       int src,
                                                       const double* bbox,
                                                       uint64_t* keys,
                                                       int32_t* idx) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const CArena* A = (const CArena*)Ap;
  const double w = bbox[3];
  const uint64_t qx = quantize(A->sx[src][off + i], bbox[0], w);
  const uint64_t qy = quantize(A->sy[src][off + i], bbox[1], w);
  const uint64_t qz = quantize(A->sz[src][off + i], bbox[2], w);
  keys[i] = spread(qx) | (spread(qy) << 1) | (spread(qz) << 2);
  idx[i] = i;
}

// Sorted copies of the sources the walk reads.
__global__ __launch_bounds__(kThreads) void k_bh_gather(const Arena* Ap,
                                                        int64_t off, int n,
                                                        int src,
                                                        const int32_t* perm,
                                                        BhBody* body,
                                                        double* bm) {
  const int k = blockIdx.x * kThreads + threadIdx.x;
  if (k >= n) return;
  const CArena* A = (const CArena*)Ap;
  const int64_t i = off + perm[k];
  body[k] = BhBody{A->sx[src][i], A->sy[src][i], A->sz[src][i], A->gm64[i]};
  bm[k] = A->m[i];
}

// ---- node chains ------------------------------------------------------------
// Node (k, b): the bodies of Morton run [k, end) sharing code >> 3b, at bit
// level b (21 = root).  Position k starts the nodes of levels
// bot(k) <= b <= top(k): top from the highest bit where code k differs from
// code k-1 (the root for k = 0), bot where the run of k becomes a single
// body (or level 0).  The preorder index of (k, b) is base(k) + top(k) - b.
__device__ __forceinline__ int hibit(uint64_t x) { return 63 - __clzll(x); }

__global__ __launch_bounds__(kThreads) void k_bh_chain(const uint64_t* keys,
                                                       int n, int32_t* cnt,
                                                       int8_t* top,
                                                       int8_t* bot) {
  const int k = blockIdx.x * kThreads + threadIdx.x;
  if (k >= n) return;
  const uint64_t c = keys[k];
  int t;
  if (k == 0) {
    t = kDepth;
  } else {
    const uint64_t x = keys[k - 1] ^ c;
    t = x == 0 ? -1 : min(kDepth - 1, hibit(x) / 3);
  }
  int b = 0;
  if (t >= 0) {
    if (k == n - 1) {
      b = t;
    } else {
      const uint64_t y = c ^ keys[k + 1];
      b = y == 0 ? 0 : min(t, hibit(y) / 3);
    }
  }
  top[k] = (int8_t)t;
  bot[k] = (int8_t)b;
  cnt[k] = t < 0 ? 0 : t - b + 1;
}

// End of the run of k at level b: first index whose code >> 3b differs.
__device__ __forceinline__ int run_end(const uint64_t* keys, int n, int k,
                                       int b) {
  const uint64_t p = keys[k] >> (3 * b);
  int lo = k + 1, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if ((keys[mid] >> (3 * b)) == p)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// Start of the run containing k at level b.
__device__ __forceinline__ int run_start(const uint64_t* keys, int k, int b) {
  const uint64_t p = keys[k] >> (3 * b);
  int lo = 0, hi = k;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if ((keys[mid] >> (3 * b)) < p)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

__global__ __launch_bounds__(kThreads) void k_bh_nodes(
    const uint64_t* keys, int n, const int32_t* base, const int8_t* top,
    const int8_t* bot, const double* bbox, BhNode* node) {
#pragma clang fp contract(off)
  const int k = blockIdx.x * kThreads + threadIdx.x;
  if (k >= n) return;
  const int t = top[k], lo = bot[k];
  if (t < 0) return;
  const double root = bbox[3];
  for (int b = t; b >= lo; --b) {
    BhNode& nd = node[base[k] + (t - b)];
    nd.start = k;
    nd.count = (b == lo && (k == n - 1 || keys[k + 1] != keys[k]))
                   ? 1
                   : run_end(keys, n, k, b) - k;
    const double w = ldexp(root, b - kDepth);  // root width halved per level
    nd.w2 = w * w;                             // octree.rs:227-228
    nd.leaf = b == lo;
    for (int c = 0; c < 8; ++c) nd.child[c] = -1;
  }
}

// Each node below the root writes itself into its parent's octant slot.
__global__ __launch_bounds__(kThreads) void k_bh_link(const uint64_t* keys,
                                                      int n,
                                                      const int32_t* base,
                                                      const int8_t* top,
                                                      const int8_t* bot,
                                                      BhNode* node) {
  const int k = blockIdx.x * kThreads + threadIdx.x;
  if (k >= n) return;
  const int t = top[k], lo = bot[k];
  if (t < 0) return;
  const uint64_t c = keys[k];
  for (int b = min(t, kDepth - 1); b >= lo; --b) {
    const int id = base[k] + (t - b);
    int parent;
    if (b < t) {
      parent = id - 1;
    } else {
      const int s = run_start(keys, k, b + 1);
      parent = base[s] + (top[s] - (b + 1));
    }
    node[parent].child[(int)((c >> (3 * b)) & 7)] = id;
  }
}

// Centres of mass of the nodes of level L (octree.rs:381-432): a leaf sums
// its bodies in Morton order, an internal node its children in octant
// order; each total divides only when the mass is positive.  Also g * mass,
// the product the pair term forms (newton.rs:139).
__global__ __launch_bounds__(kThreads) void k_bh_mass(
    int level, int n, double g, const int32_t* base, const int8_t* top,
    const int8_t* bot, const BhBody* body, const double* bm, BhNode* node) {
#pragma clang fp contract(off)
  const int k = blockIdx.x * kThreads + threadIdx.x;
  if (k >= n) return;
  const int t = top[k];
  if (t < 0 || level > t || level < bot[k]) return;
  BhNode& nd = node[base[k] + (t - level)];
  double tm = 0.0, sx = 0.0, sy = 0.0, sz = 0.0;
  if (nd.leaf) {
    const int s = nd.start, e = s + nd.count;
    for (int b = s; b < e; ++b) {
      tm = tm + bm[b];
      sx = sx + bm[b] * body[b].x;
      sy = sy + bm[b] * body[b].y;
      sz = sz + bm[b] * body[b].z;
    }
  } else {
    for (int c = 0; c < 8; ++c) {
      const int ch = nd.child[c];
      if (ch < 0) continue;
      const BhNode& cn = node[ch];
      tm = tm + cn.mass;
      sx = sx + cn.mass * cn.cx;
      sy = sy + cn.mass * cn.cy;
      sz = sz + cn.mass * cn.cz;
    }
  }
  if (tm > 0.0) {
    sx = sx / tm;
    sy = sy / tm;
    sz = sz / tm;
  }
  nd.mass = tm;
  nd.gm = g * tm;
  nd.cx = sx;
  nd.cy = sy;
  nd.cz = sz;
}

// ---- walk (octree.rs:159-254) --------------------------------------------------
template <bool STRICT>
__device__ __forceinline__ void add_pair(double gm, double dx, double dy,
                                         double dz, double eps2, bool ok,
                                         double& ax, double& ay, double& az) {
#pragma clang fp contract(off)
  if (STRICT) {  // newton.rs:128-141
    const double r2 = ((dx * dx) + (dy * dy)) + (dz * dz);
    const double inv = 1.0 / __builtin_sqrt(r2 + eps2);
    const double k = gm * ((inv * inv) * inv);
    ax = ok ? ax + k * dx : ax;
    ay = ok ? ay + k * dy : ay;
    az = ok ? az + k * dz : az;
  } else {
    const double r2 = fma(dz, dz, fma(dy, dy, fma(dx, dx, eps2)));
    const double k = ok ? gm * inv_cube_f64(r2) : 0.0;
    ax = fma(k, dx, ax);
    ay = fma(k, dy, ay);
    az = fma(k, dz, az);
  }
}

__device__ __forceinline__ uint64_t uniform64(uint64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

// The fields of a node record every visit reads (the first 52 bytes).
struct NodeHot {
  double cx, cy, cz, gm, w2;
  int32_t start, count, leaf;
};
struct NodeKids {
  int32_t c[8];
};

// Strict walk: the reference's traversal per lane (octree.rs:159-254),
// its arithmetic and its order.
__global__ __launch_bounds__(kWave) void k_bh_walk_strict(const Arena* Ap,
                                                          const SysDesc* sysd,
                                                          int sys, int stage,
                                                          int src, BhTree T,
                                                          double theta2,
                                                          int n) {
#pragma clang fp contract(off)
  __shared__ int32_t st_node[kStack];
  __shared__ uint64_t st_mask[kStack];
  const int lane = (int)threadIdx.x;
  const int k = (int)blockIdx.x * kWave + lane;
  const bool valid = k < n;
  const int kk = valid ? k : n - 1;
  const CArena* A = (const CArena*)Ap;
  const CSys* sd = (const CSys*)(sysd + sys);
  const double eps2 = sd->eps2;
  const BhBody me = T.body[kk];
  const double tx = me.x, ty = me.y, tz = me.z;
  double ax = 0.0, ay = 0.0, az = 0.0;
  const uint64_t live = __ballot(valid);
  if (lane == 0) {  // the root, for every live lane
    st_node[0] = 0;
    st_mask[0] = live;
  }
  __syncthreads();
  int sp = 1;
  while (sp > 0) {
    --sp;
    const int id = __builtin_amdgcn_readfirstlane(st_node[sp]);
    const uint64_t mask = uniform64(st_mask[sp]);
    __syncthreads();
    const NodeHot nd = load_uniform((const NodeHot*)(T.node + id));
    const bool act = (mask >> lane) & 1;
    const double ndx = nd.cx - tx, ndy = nd.cy - ty, ndz = nd.cz - tz;
    const double d2 = ((ndx * ndx) + (ndy * ndy)) + (ndz * ndz);
    const bool here = act && d2 != 0.0;  // a node at the target is skipped
    if (nd.leaf) {
      if (__ballot(here)) {
        for (int b = nd.start; b < nd.start + nd.count; ++b) {
          const BhBody s = load_uniform(T.body + b);
          add_pair<true>(s.gm, s.x - tx, s.y - ty, s.z - tz, eps2,
                           here && b != k, ax, ay, az);
        }
      }
      continue;
    }
    const bool far = (nd.w2 / d2) < theta2;  // octree.rs:227-229
    bool near = false;
    if (here) {
      if (far)  // far enough: one mass at the centre
        add_pair<true>(nd.gm, ndx, ndy, ndz, eps2, true, ax, ay, az);
      else
        near = true;
    }
    const uint64_t open = __ballot(near);
    if (open) {
      const NodeKids kids = load_uniform((const NodeKids*)T.node[id].child);
      for (int c = 0; c < 8; ++c) {
        if (kids.c[c] >= 0) {
          if (lane == 0) {
            st_node[sp] = kids.c[c];
            st_mask[sp] = open;
          }
          ++sp;
        }
      }
      __syncthreads();
    }
  }
  if (valid) {
    const int64_t i = sd->off + T.perm[k];
    epilogue<MODE_F64_STRICT>(A, sd, stage, src ^ 1, i, i, ax, ay, az);
  }
}

// Fast walk (f64, fused multiply-add, rsq-based inverse square root).
// Judges one node for every lane: d^2 == 0 skips; a one-body leaf adds its
// body (through the node's centre of mass and g*m, which equal the body's to
// rounding) unless it is the lane's own; a larger leaf adds its bodies; an
// internal node adds its mass when w^2 < theta^2 d^2 and otherwise joins the
// returned open mask.
__device__ __forceinline__ uint64_t judge_fast(const NodeHot& r, const BhTree& T,
                                               int k, double tx, double ty,
                                               double tz, double eps2,
                                               double theta2, bool act,
                                               double& ax, double& ay,
                                               double& az) {
  const double ndx = r.cx - tx, ndy = r.cy - ty, ndz = r.cz - tz;
  const double d2 = fma(ndz, ndz, fma(ndy, ndy, ndx * ndx));
  const bool here = act && d2 != 0.0;
  if (r.leaf) {
    if (r.count == 1) {
      add_pair<false>(r.gm, ndx, ndy, ndz, eps2, here && r.start != k, ax, ay,
                      az);
    } else if (__ballot(here)) {
      for (int b = r.start; b < r.start + r.count; ++b) {
        const BhBody s = load_uniform(T.body + b);
        add_pair<false>(s.gm, s.x - tx, s.y - ty, s.z - tz, eps2,
                        here && b != k, ax, ay, az);
      }
    }
    return 0;
  }
  const bool far = r.w2 < theta2 * d2;
  add_pair<false>(r.gm, ndx, ndy, ndz, eps2, here && far, ax, ay, az);
  return __ballot(here && !far);
}

__global__ __launch_bounds__(kWave) void k_bh_walk_fast(const Arena* Ap,
                                                        const SysDesc* sysd,
                                                        int sys, int stage,
                                                        int src, BhTree T,
                                                        double theta2, int n) {
  __shared__ int32_t st_node[kStack];
  __shared__ uint64_t st_mask[kStack];
  const int lane = (int)threadIdx.x;
  const int k = (int)blockIdx.x * kWave + lane;
  const bool valid = k < n;
  const int kk = valid ? k : n - 1;
  const CArena* A = (const CArena*)Ap;
  const CSys* sd = (const CSys*)(sysd + sys);
  const double eps2 = sd->eps2;
  const BhBody me = T.body[kk];
  const double tx = me.x, ty = me.y, tz = me.z;
  double ax = 0.0, ay = 0.0, az = 0.0;
  int sp = 0;
  {
    const NodeHot root = load_uniform((const NodeHot*)T.node);
    const uint64_t open = judge_fast(root, T, k, tx, ty, tz, eps2, theta2,
                                     valid, ax, ay, az);
    if (open) {
      if (lane == 0) {
        st_node[0] = 0;
        st_mask[0] = open;
      }
      sp = 1;
    }
  }
  __syncthreads();
  while (sp > 0) {
    --sp;
    const int id = __builtin_amdgcn_readfirstlane(st_node[sp]);
    const uint64_t mask = uniform64(st_mask[sp]);
    __syncthreads();
    const bool act = (mask >> lane) & 1;
    const NodeKids kids = load_uniform((const NodeKids*)T.node[id].child);
    // four children per batch: their records load together (an empty slot
    // reads the root's record and is then skipped)
#pragma unroll
    for (int c0 = 0; c0 < 8; c0 += 4) {
      NodeHot r[4];
#pragma unroll
      for (int q = 0; q < 4; ++q)
        r[q] = load_uniform((const NodeHot*)(T.node + max(kids.c[c0 + q], 0)));
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int ch = kids.c[c0 + q];
        if (ch < 0) continue;
        const uint64_t open = judge_fast(r[q], T, k, tx, ty, tz, eps2, theta2,
                                         act, ax, ay, az);
        if (open) {
          if (lane == 0) {
            st_node[sp] = ch;
            st_mask[sp] = open;
          }
          ++sp;
        }
      }
    }
    __syncthreads();
  }
  if (valid) {
    const int64_t i = sd->off + T.perm[k];
    epilogue<MODE_F64>(A, sd, stage, src ^ 1, i, i, ax, ay, az);
  }
}

inline int blocks(int64_t n, int t) { return (int)((n + t - 1) / t); }

}  // namespace

BarnesHut::BarnesHut(int64_t n) : n_(n) {
  if (n <= 0 || n > (1 << 26))
    throw Error(NBODY_E_INVALID, "barnes-hut supports 1 .. 2^26 bodies");
  // hipCUB temporary storage: the larger of the sort and the scan
  size_t sort_bytes = 0, scan_bytes = 0;
  hip_check(hipcub::DeviceRadixSort::SortPairs(
                nullptr, sort_bytes, (uint64_t*)nullptr, (uint64_t*)nullptr,
                (int32_t*)nullptr, (int32_t*)nullptr, (int)n, 0, 63),
            "radix sort sizing");
  hip_check(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes,
                                             (int32_t*)nullptr,
                                             (int32_t*)nullptr, (int)n),
            "scan sizing");
  temp_bytes_ = std::max(sort_bytes, scan_bytes);
  // one allocation for the per-body arrays, carved in 256-byte pieces
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t at = off;
    off += (bytes + 255) & ~(size_t)255;
    return at;
  };
  const size_t o_keys_in = take(8 * n), o_keys = take(8 * n);
  const size_t o_idx_in = take(4 * n), o_perm = take(4 * n);
  const size_t o_body = take(sizeof(BhBody) * n), o_bm = take(8 * n);
  const size_t o_cnt = take(4 * n), o_base = take(4 * n);
  const size_t o_top = take(n), o_bot = take(n);
  const size_t o_bbox = take(8 * (6 * kBboxBlocks + 4));
  const size_t o_temp = take(temp_bytes_);
  hip_check(hipMalloc(&mem_, off), "hipMalloc(barnes-hut)");
  char* p = (char*)mem_;
  keys_in_ = (uint64_t*)(p + o_keys_in);
  keys_ = (uint64_t*)(p + o_keys);
  idx_in_ = (int32_t*)(p + o_idx_in);
  perm_ = (int32_t*)(p + o_perm);
  body_ = (BhBody*)(p + o_body);
  bm_ = (double*)(p + o_bm);
  cnt_ = (int32_t*)(p + o_cnt);
  base_ = (int32_t*)(p + o_base);
  top_ = (int8_t*)(p + o_top);
  bot_ = (int8_t*)(p + o_bot);
  bbox_part_ = (double*)(p + o_bbox);
  bbox_ = bbox_part_ + 6 * kBboxBlocks;
  temp_ = p + o_temp;
  hip_check(hipHostMalloc((void**)&h_last_, 2 * sizeof(int32_t),
                          hipHostMallocDefault),
            "hipHostMalloc(node count)");
  reserve_nodes(3 * n + 64);
}

BarnesHut::~BarnesHut() {
  (void)hipFree(nodes_);
  (void)hipFree(mem_);
  (void)hipHostFree(h_last_);
}

// Node records for `cap` nodes (grow only; the previous ones are unused
// once the stream has drained, which the callers' synchronisation ensures).
void BarnesHut::reserve_nodes(int64_t cap) {
  if (cap <= cap_) return;
  (void)hipFree(nodes_);
  nodes_ = nullptr;
  cap_ = cap;
  hip_check(hipMalloc(&nodes_, sizeof(BhNode) * (size_t)cap_),
            "hipMalloc(barnes-hut nodes)");
  node_ = (BhNode*)nodes_;
}

void BarnesHut::evaluate(hipStream_t s, bool strict, const Arena* d_arena,
                         const SysDesc* d_sys, int sys, const SysDesc& desc,
                         int stage, int src, double theta) {
  const int n = (int)desc.n;
  if (desc.n != n_) throw Error(NBODY_E_STATE, "tree sized for another n");
  const int64_t off = desc.off;
  const int nb = std::min(kBboxBlocks, blocks(n, kThreads));
  hipLaunchKernelGGL(k_bh_bbox_part, dim3(nb), dim3(kThreads), 0, s, d_arena,
                     off, n, src, bbox_part_);
  hipLaunchKernelGGL(k_bh_bbox_final, dim3(1), dim3(64), 0, s, bbox_part_, nb,
                     bbox_);
  hipLaunchKernelGGL(k_bh_codes, dim3(blocks(n, kThreads)), dim3(kThreads), 0,
                     s, d_arena, off, n, src, bbox_, keys_in_, idx_in_);
  size_t bytes = temp_bytes_;
  hip_check(hipcub::DeviceRadixSort::SortPairs(temp_, bytes, keys_in_, keys_,
                                               idx_in_, perm_, n, 0, 63, s),
            "radix sort");
  hipLaunchKernelGGL(k_bh_gather, dim3(blocks(n, kThreads)), dim3(kThreads), 0,
                     s, d_arena, off, n, src, perm_, body_, bm_);
  hipLaunchKernelGGL(k_bh_chain, dim3(blocks(n, kThreads)), dim3(kThreads), 0,
                     s, keys_, n, cnt_, top_, bot_);
  bytes = temp_bytes_;
  hip_check(hipcub::DeviceScan::ExclusiveSum(temp_, bytes, cnt_, base_, n, s),
            "scan");
  // the node count decides the node arrays: read it back (one sync)
  const int64_t total = node_count(s);
  if (total < 1 || total > 22 * (int64_t)n + 1)
    throw Error(NBODY_E_DEVICE, "barnes-hut node count out of range");
  reserve_nodes(total);
  hipLaunchKernelGGL(k_bh_nodes, dim3(blocks(n, kThreads)), dim3(kThreads), 0,
                     s, keys_, n, base_, top_, bot_, bbox_, node_);
  hipLaunchKernelGGL(k_bh_link, dim3(blocks(n, kThreads)), dim3(kThreads), 0,
                     s, keys_, n, base_, top_, bot_, node_);
  for (int level = 0; level <= kDepth; ++level)
    hipLaunchKernelGGL(k_bh_mass, dim3(blocks(n, kThreads)), dim3(kThreads), 0,
                       s, level, n, desc.g, base_, top_, bot_, body_, bm_,
                       node_);
  BhTree T{body_, perm_, node_};
  const double theta2 = theta * theta;  // octree.rs:229
  if (strict)
    hipLaunchKernelGGL(k_bh_walk_strict, dim3(blocks(n, kWave)), dim3(kWave), 0,
                       s, d_arena, d_sys, sys, stage, src, T, theta2, n);
  else
    hipLaunchKernelGGL(k_bh_walk_fast, dim3(blocks(n, kWave)), dim3(kWave), 0,
                       s, d_arena, d_sys, sys, stage, src, T, theta2, n);
  hip_check(hipGetLastError(), "barnes-hut launch");
}

int64_t BarnesHut::node_count(hipStream_t s) {
  hip_check(hipMemcpyAsync(&h_last_[0], base_ + n_ - 1, 4,
                           hipMemcpyDeviceToHost, s),
            "hipMemcpy(node count)");
  hip_check(hipMemcpyAsync(&h_last_[1], cnt_ + n_ - 1, 4,
                           hipMemcpyDeviceToHost, s),
            "hipMemcpy(node count)");
  hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
  return (int64_t)h_last_[0] + h_last_[1];
}

}  // namespace nbody
