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

typedef __attribute__((address_space(4))) const int8_t ci8;

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
  if (threadIdx.x != 0) return;
  double v[6];
  for (int c = 0; c < 6; ++c) v[c] = part[c];
  for (int p = 1; p < nparts; ++p)
    for (int c = 0; c < 3; ++c) {
      const double a = part[p * 6 + c], b = part[p * 6 + 3 + c];
      v[c] = a < v[c] ? a : v[c];
      v[3 + c] = b > v[3 + c] ? b : v[3 + c];
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
__global__ __launch_bounds__(kThreads) void k_bh_gather(
    const Arena* Ap, int64_t off, int n, int src, const int32_t* perm,
    double* bx, double* by, double* bz, double* bm, double* bg) {
  const int k = blockIdx.x * kThreads + threadIdx.x;
  if (k >= n) return;
  const CArena* A = (const CArena*)Ap;
  const int64_t i = off + perm[k];
  bx[k] = A->sx[src][i];
  by[k] = A->sy[src][i];
  bz[k] = A->sz[src][i];
  bm[k] = A->m[i];
  bg[k] = A->gm64[i];
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
    const int8_t* bot, const double* bbox, int32_t* start, int32_t* count,
    double* width, int8_t* leaf, int32_t* child) {
  const int k = blockIdx.x * kThreads + threadIdx.x;
  if (k >= n) return;
  const int t = top[k], lo = bot[k];
  if (t < 0) return;
  const double root = bbox[3];
  for (int b = t; b >= lo; --b) {
    const int id = base[k] + (t - b);
    start[id] = k;
    count[id] = (b == lo && (k == n - 1 || keys[k + 1] != keys[k]))
                    ? 1
                    : run_end(keys, n, k, b) - k;
    width[id] = ldexp(root, b - kDepth);  // root width halved per level
    leaf[id] = (int8_t)(b == lo);
    for (int c = 0; c < 8; ++c) child[(int64_t)id * 8 + c] = -1;
  }
}

// Each node below the root writes itself into its parent's octant slot.
__global__ __launch_bounds__(kThreads) void k_bh_link(const uint64_t* keys,
                                                      int n,
                                                      const int32_t* base,
                                                      const int8_t* top,
                                                      const int8_t* bot,
                                                      int32_t* child) {
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
    child[(int64_t)parent * 8 + (int)((c >> (3 * b)) & 7)] = id;
  }
}

// Centres of mass of the nodes of level L (octree.rs:381-432): a leaf sums
// its bodies in Morton order, an internal node its children in octant
// order; each total divides only when the mass is positive.
__global__ __launch_bounds__(kThreads) void k_bh_mass(
    int level, int n, const int32_t* base, const int8_t* top,
    const int8_t* bot, const int32_t* start, const int32_t* count,
    const int8_t* leaf, const int32_t* child, const double* bm,
    const double* bx, const double* by, const double* bz, double* mass,
    double* cx, double* cy, double* cz) {
#pragma clang fp contract(off)
  const int k = blockIdx.x * kThreads + threadIdx.x;
  if (k >= n) return;
  const int t = top[k];
  if (t < 0 || level > t || level < bot[k]) return;
  const int id = base[k] + (t - level);
  double tm = 0.0, sx = 0.0, sy = 0.0, sz = 0.0;
  if (leaf[id]) {
    const int s = start[id], e = s + count[id];
    for (int b = s; b < e; ++b) {
      tm = tm + bm[b];
      sx = sx + bm[b] * bx[b];
      sy = sy + bm[b] * by[b];
      sz = sz + bm[b] * bz[b];
    }
  } else {
    for (int c = 0; c < 8; ++c) {
      const int ch = child[(int64_t)id * 8 + c];
      if (ch < 0) continue;
      tm = tm + mass[ch];
      sx = sx + mass[ch] * cx[ch];
      sy = sy + mass[ch] * cy[ch];
      sz = sz + mass[ch] * cz[ch];
    }
  }
  if (tm > 0.0) {
    sx = sx / tm;
    sy = sy / tm;
    sz = sz / tm;
  }
  mass[id] = tm;
  cx[id] = sx;
  cy[id] = sy;
  cz[id] = sz;
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
    const double inv = rsqrt_f64(r2);
    const double k = ok ? gm * ((inv * inv) * inv) : 0.0;
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

template <bool STRICT>
__global__ __launch_bounds__(kWave) void k_bh_walk(const Arena* Ap,
                                                   const SysDesc* sysd,
                                                   int sys, int stage,
                                                   int src, BhTree T,
                                                   double theta2, int n) {
#pragma clang fp contract(off)
  __shared__ int32_t st_node[kStack];
  __shared__ uint64_t st_mask[kStack];
  const int lane = (int)threadIdx.x;
  const int k = (int)blockIdx.x * kWave + lane;
  const bool valid = k < n;
  const int kk = valid ? k : n - 1;
  const CArena* A = (const CArena*)Ap;
  const CSys* sd = (const CSys*)(sysd + sys);
  const double eps2 = sd->eps2, g = sd->g;
  const double tx = T.bx[kk], ty = T.by[kk], tz = T.bz[kk];
  const cf64 *CX = (const cf64*)T.cx, *CY = (const cf64*)T.cy,
             *CZ = (const cf64*)T.cz, *M = (const cf64*)T.mass,
             *W = (const cf64*)T.width, *BX = (const cf64*)T.bx,
             *BY = (const cf64*)T.by, *BZ = (const cf64*)T.bz,
             *BG = (const cf64*)T.bg;
  const ci32 *START = (const ci32*)T.start, *COUNT = (const ci32*)T.count,
             *CHILD = (const ci32*)T.child;
  const ci8* LEAF = (const ci8*)T.leaf;
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
    const int node = __builtin_amdgcn_readfirstlane(st_node[sp]);
    const uint64_t mask = uniform64(st_mask[sp]);
    __syncthreads();
    const bool act = (mask >> lane) & 1;
    const double ndx = CX[node] - tx, ndy = CY[node] - ty, ndz = CZ[node] - tz;
    const double d2 = ((ndx * ndx) + (ndy * ndy)) + (ndz * ndz);
    const bool here = act && d2 != 0.0;  // a node at the target is skipped
    if (LEAF[node]) {
      if (__ballot(here)) {
        const int s = START[node], e = s + COUNT[node];
        for (int b = s; b < e; ++b)
          add_pair<STRICT>(BG[b], BX[b] - tx, BY[b] - ty, BZ[b] - tz, eps2,
                           here && b != k, ax, ay, az);
      }
      continue;
    }
    const double w = W[node];
    bool near = false;
    if (here) {
      if (((w * w) / d2) < theta2) {  // far enough: one mass at the centre
        add_pair<STRICT>(g * M[node], ndx, ndy, ndz, eps2, true, ax, ay, az);
      } else {
        near = true;
      }
    }
    const uint64_t open = __ballot(near);
    if (open) {
      for (int c = 0; c < 8; ++c) {
        const int ch = CHILD[(int64_t)node * 8 + c];
        if (ch >= 0) {
          if (lane == 0) {
            st_node[sp] = ch;
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
    epilogue<STRICT ? MODE_F64_STRICT : MODE_F64>(A, sd, stage, src ^ 1, i, ax,
                                                  ay, az);
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
  const size_t o_b = take(8 * 5 * n);
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
  bx_ = (double*)(p + o_b);
  by_ = bx_ + n;
  bz_ = by_ + n;
  bm_ = bz_ + n;
  bg_ = bm_ + n;
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

// Node arrays for `cap` nodes (grow only; the previous ones are unused
// once the stream has drained, which the callers' synchronisation ensures).
void BarnesHut::reserve_nodes(int64_t cap) {
  if (cap <= cap_) return;
  (void)hipFree(nodes_);
  nodes_ = nullptr;
  cap_ = cap;
  const size_t per = 8 * 5 + 4 + 4 + 4 * 8 + 1;  // bytes per node
  hip_check(hipMalloc(&nodes_, per * (size_t)cap_ + 5 * 256),
            "hipMalloc(barnes-hut nodes)");
  char* p = (char*)nodes_;
  auto next = [&](size_t bytes) {
    char* at = p;
    p += (bytes + 255) & ~(size_t)255;
    return at;
  };
  cx_ = (double*)next(8 * 5 * (size_t)cap_);
  cy_ = cx_ + cap_;
  cz_ = cy_ + cap_;
  mass_ = cz_ + cap_;
  width_ = mass_ + cap_;
  start_ = (int32_t*)next(4 * (size_t)cap_);
  count_ = (int32_t*)next(4 * (size_t)cap_);
  child_ = (int32_t*)next(4 * 8 * (size_t)cap_);
  leaf_ = (int8_t*)next((size_t)cap_);
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
                     s, d_arena, off, n, src, perm_, bx_, by_, bz_, bm_, bg_);
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
                     s, keys_, n, base_, top_, bot_, bbox_, start_, count_,
                     width_, leaf_, child_);
  hipLaunchKernelGGL(k_bh_link, dim3(blocks(n, kThreads)), dim3(kThreads), 0,
                     s, keys_, n, base_, top_, bot_, child_);
  for (int level = 0; level <= kDepth; ++level)
    hipLaunchKernelGGL(k_bh_mass, dim3(blocks(n, kThreads)), dim3(kThreads), 0,
                       s, level, n, base_, top_, bot_, start_, count_, leaf_,
                       child_, bm_, bx_, by_, bz_, mass_, cx_, cy_, cz_);
  BhTree T{bx_, by_, bz_, bm_, bg_, perm_, cx_, cy_, cz_, mass_, width_,
           start_, count_, child_, leaf_};
  const double theta2 = theta * theta;  // octree.rs:229
  if (strict)
    hipLaunchKernelGGL(k_bh_walk<true>, dim3(blocks(n, kWave)), dim3(kWave), 0,
                       s, d_arena, d_sys, sys, stage, src, T, theta2, n);
  else
    hipLaunchKernelGGL(k_bh_walk<false>, dim3(blocks(n, kWave)), dim3(kWave),
                       0, s, d_arena, d_sys, sys, stage, src, T, theta2, n);
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
