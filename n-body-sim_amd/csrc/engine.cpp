// Host engine implementation (see engine.h).
#include "engine.h"

#include "bh.h"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "../../include/nbody.h"
#include "kernels.h"
#include "tuning.h"

namespace nbody {

void hip_check(hipError_t e, const char* what) {
  if (e == hipSuccess) return;
  const int code = (e == hipErrorOutOfMemory) ? NBODY_E_NOMEM : NBODY_E_DEVICE;
  throw Error(code, std::string(what) + ": " + hipGetErrorString(e));
}

KernelTimer::~KernelTimer() {
  for (hipEvent_t e : ev_) (void)hipEventDestroy(e);
}

void KernelTimer::begin(hipStream_t s) {
  if (used_ + 2 > ev_.size()) {
    if (ev_.size() >= 4096) harvest();  // pool full: wait, then reuse
    while (used_ + 2 > ev_.size()) {
      hipEvent_t e;
      hip_check(hipEventCreate(&e), "hipEventCreate");
      ev_.push_back(e);
    }
  }
  hip_check(hipEventRecord(ev_[used_++], s), "hipEventRecord");
}

void KernelTimer::end(hipStream_t s, Kind kind) {
  hip_check(hipEventRecord(ev_[used_++], s), "hipEventRecord");
  kind_.push_back(kind);
}

void KernelTimer::harvest() {
  for (size_t k = 0; k < kind_.size(); ++k) {
    hip_check(hipEventSynchronize(ev_[2 * k + 1]), "hipEventSynchronize");
    float t = 0.f;
    hip_check(hipEventElapsedTime(&t, ev_[2 * k], ev_[2 * k + 1]),
              "hipEventElapsedTime");
    ms[kind_[k]] += t;
    ++launches[kind_[k]];
  }
  kind_.clear();
  used_ = 0;
}

void KernelTimer::reset() {
  harvest();
  ms[0] = ms[1] = 0.0;
  launches[0] = launches[1] = 0;
}

namespace {
constexpr int64_t kAlign = 64;  // system ranges start on a wave boundary

int64_t round_up(int64_t v, int64_t a) { return (v + a - 1) / a * a; }
}  // namespace


Engine::Engine(int device, hipStream_t stream, int mode, int64_t capacity,
               int max_systems)
    : stream_(stream),
      mode_(mode),
      cap_(round_up(std::max<int64_t>(capacity, 1), kAlign)),
      max_sys_(std::max(max_systems, 1)) {
  (void)device;  // the stream belongs to it; every call runs on the stream
  if (mode < MODE_F64_STRICT || mode > MODE_F32R)
    throw Error(NBODY_E_INVALID, "unknown precision");
  alloc_arena();
  hip_check(hipMalloc((void**)&d_sys_, sizeof(SysDesc) * max_sys_),
            "hipMalloc(system table)");
  hip_check(hipMalloc((void**)&d_list_, sizeof(int32_t) * max_sys_),
            "hipMalloc(system list)");
  hip_check(hipMalloc((void**)&d_diag_, sizeof(double) * DIAG_COUNT * max_sys_),
            "hipMalloc(diagnostics)");
  hip_check(hipEventCreateWithFlags(&ev_gather_, hipEventDisableTiming),
            "hipEventCreate");
  hip_check(hipEventCreateWithFlags(&ev_boundary_, hipEventDisableTiming),
            "hipEventCreate");
  h_sys_.assign(max_sys_, SysDesc{});
  sys_.assign(max_sys_, Sys{});
  free_[0] = cap_;
}

Engine::~Engine() {
  drop_graphs();
  (void)hipStreamSynchronize(stream_);
  (void)hipFree(d_items_);
  (void)hipFree(d_step_items_);
  for (Staging& st : stage_) {
    (void)hipFree(st.d);
    (void)hipHostFree(st.h);
  }
  (void)hipFree(d_traj_);
  (void)hipFree(d_snap_items_);
  (void)hipFree(d_step_spans_);
  (void)hipFree(d_sort_tab_);
  (void)hipFree(d_keys_);
  (void)hipFree(d_vals_);
  (void)hipFree(d_sort_tmp_);
  (void)hipFree(d_diag_);
  (void)hipFree(d_list_);
  (void)hipFree(d_sys_);
  (void)hipFree(d_arena_);
  (void)hipFree(arena_mem_);
  if (ev_gather_) (void)hipEventDestroy(ev_gather_);
  if (ev_boundary_) (void)hipEventDestroy(ev_boundary_);
}

// Carves every arena array out of one block starting at `base`; returns the
// bytes used.  Called once with base = nullptr to size the block, then on the
// allocation itself, so the size can never disagree with the layout.
size_t Engine::carve(char* base) {
  const int64_t len = cap_ + kPad;
  const size_t dbytes = round_up(len * 8, 256);
  const size_t fbytes = round_up(len * 4, 256);
  size_t used = 0;
  auto D = [&](int count = 1) {
    double* r = (double*)(base + used);
    used += dbytes * count;
    return r;
  };
  auto F = [&]() {
    float* r = (float*)(base + used);
    used += fbytes;
    return r;
  };
  A_.m = D();
  A_.gm64 = D();
  A_.x = D();
  A_.y = D();
  A_.z = D();
  A_.vx = D();
  A_.vy = D();
  A_.vz = D();
  for (int b = 0; b < 2; ++b) {
    A_.sx[b] = D();
    A_.sy[b] = D();
    A_.sz[b] = D();
  }
  A_.ax = D();
  A_.ay = D();
  A_.az = D();
  for (int q = 0; q < 9; ++q) A_.rk[q] = D();
  for (int q = 0; q < 3; ++q) A_.xs[q] = D();
  A_.phi = D();
  A_.pp = (Part3*)D(3 * kMaxSplit);  // record k of part p at p * pstride + k
  A_.pstride = (int64_t)(dbytes / sizeof(double));
  for (int b = 0; b < 2; ++b) A_.hs[b] = (HSrc*)D(4);
  A_.perm = (int32_t*)F();
  A_.iperm = (int32_t*)F();
  A_.gm32 = F();
  A_.cap = cap_;
  return used;
}

void Engine::alloc_arena() {
  // Every array gets cap + kPad elements, each array 256-byte aligned.
  arena_bytes_ = carve(nullptr);
  hip_check(hipMalloc(&arena_mem_, arena_bytes_), "hipMalloc(arena)");
  // Zero everything once: slack and never-used ranges must hold finite values.
  hip_check(hipMemsetAsync(arena_mem_, 0, arena_bytes_, stream_),
            "hipMemset(arena)");
  if (carve((char*)arena_mem_) != arena_bytes_)
    throw Error(NBODY_E_DEVICE, "arena layout mismatch");
  // kernels read the arena's pointers from one device-resident copy
  hip_check(hipMalloc((void**)&d_arena_, sizeof(Arena)), "hipMalloc(arena table)");
  hip_check(hipMemcpyAsync(d_arena_, &A_, sizeof(Arena), hipMemcpyHostToDevice,
                           stream_),
            "hipMemcpy(arena table)");
  hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
}

// Host-side guard run on every work list before it is uploaded: each item
// must name a live system, a tile inside it and a snapshot buffer 0 or 1.
void Engine::validate_items(const std::vector<WorkItem>& items) const {
  // items come in runs of one system: its limits are looked up per run
  int cur = -1;
  int64_t tiles = 0;
  int parts = 1;
  for (const WorkItem& it : items) {
    if (it.sys != cur) {
      if (it.sys < 0 || it.sys >= max_sys_ || !sys_[it.sys].live)
        throw Error(NBODY_E_STATE, "work item names no live system");
      const Sys& s = sys_[it.sys];
      if (s.off < 0 || s.off + s.span > cap_ || s.span < s.p.n)
        throw Error(NBODY_E_STATE, "system range outside the arena");
      cur = it.sys;
      tiles = (s.p.n + kWave - 1) / kWave;
      parts = std::max(1, split_of(it.sys));
    }
    if (it.tile < 0 || it.tile >= tiles || (it.src & ~1) != 0 ||
        it.stage < STAGE_ACC || it.stage > STAGE_POT || it.part < 0 ||
        it.part >= parts)
      throw Error(NBODY_E_STATE, "work item out of range");
  }
}

void Engine::drop_graphs() {
  for (auto& g : gexec_)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
  gkey_ = -1;
  gids_.clear();
}

void Engine::check_id(int id) const {
  if (id < 0 || id >= max_sys_ || !sys_[id].live)
    throw Error(NBODY_E_NOTFOUND, "unknown system id");
}

int Engine::nsystems() const {
  int c = 0;
  for (auto& s : sys_) c += s.live;
  return c;
}

const SystemParams& Engine::params(int id) const {
  check_id(id);
  return sys_[id].p;
}

int64_t Engine::steps_done(int id) const {
  check_id(id);
  return sys_[id].steps;
}

static void validate(const SystemParams& p) {
  if (p.n <= 0) throw Error(NBODY_E_INVALID, "system must have n >= 1 bodies");
  if (p.n > INT32_MAX) throw Error(NBODY_E_INVALID, "n too large");
  if (p.integ < INTEG_EULER || p.integ > INTEG_RK4)
    throw Error(NBODY_E_INVALID, "unknown integrator");
  if (!std::isfinite(p.dt) || !std::isfinite(p.g) || !std::isfinite(p.eps2) ||
      p.eps2 < 0.0)
    throw Error(NBODY_E_INVALID, "dt, g and eps2 must be finite, eps2 >= 0");
}

int Engine::add(const SystemParams& p, const double* m, const double* x,
                const double* y, const double* z, const double* vx,
                const double* vy, const double* vz) {
  step_sig_.clear();
  validate(p);
  if (!m || !x || !y || !z || !vx || !vy || !vz)
    throw Error(NBODY_E_INVALID, "null body array");
  const double* arrs[7] = {m, x, y, z, vx, vy, vz};
  for (auto a : arrs)
    for (int64_t i = 0; i < p.n; ++i)
      if (!std::isfinite(a[i]))
        throw Error(NBODY_E_INVALID, "non-finite body value");
  int id = -1;
  for (int k = 0; k < max_sys_; ++k)
    if (!sys_[k].live) {
      id = k;
      break;
    }
  if (id < 0) throw Error(NBODY_E_NOSPACE, "system table full");
  const int64_t need = round_up(p.n, kAlign);
  const int64_t off = take_range(free_, p.n);
  if (off < 0) throw Error(NBODY_E_NOSPACE, "arena full");

  Sys& s = sys_[id];
  s = Sys{};
  s.live = true;
  s.p = p;
  s.off = off;
  s.span = need;
  s.split = std::min(split_for_size(p.n), split_cap_);
  for (int m = 0; m < 4; ++m) s.variant[m] = tuning_variant(m, p.n);
  // centre of mass and its velocity, f64 in index order (the centre of the
  // MODE_F32R ordering box, advanced by vcom * elapsed time), and the
  // largest distance from it along each axis (the box size)
  double M = 0, c[3] = {0, 0, 0}, v[3] = {0, 0, 0};
  for (int64_t i = 0; i < p.n; ++i) {
    M += m[i];
    c[0] += m[i] * x[i];
    c[1] += m[i] * y[i];
    c[2] += m[i] * z[i];
    v[0] += m[i] * vx[i];
    v[1] += m[i] * vy[i];
    v[2] += m[i] * vz[i];
  }
  for (int q = 0; q < 3; ++q) {
    s.com[q] = (M != 0.0) ? c[q] / M : 0.0;
    s.vcom[q] = (M != 0.0) ? v[q] / M : 0.0;
  }
  for (int64_t i = 0; i < p.n; ++i) {
    s.ext3[0] = std::max(s.ext3[0], std::fabs(x[i] - s.com[0]));
    s.ext3[1] = std::max(s.ext3[1], std::fabs(y[i] - s.com[1]));
    s.ext3[2] = std::max(s.ext3[2], std::fabs(z[i] - s.com[2]));
    s.mmax = std::max(s.mmax, std::fabs(m[i]));
  }
  const size_t bytes = sizeof(double) * p.n;
  double* dst[7] = {A_.m, A_.x, A_.y, A_.z, A_.vx, A_.vy, A_.vz};
  for (int q = 0; q < 7; ++q)
    hip_check(hipMemcpyAsync(dst[q] + off, arrs[q], bytes,
                             hipMemcpyHostToDevice, stream_),
              "hipMemcpy(upload)");
  hip_check(launch_clear_range(d_arena_, off, p.n, CLEAR_ACC, stream_),
            "clear launch");
  // the caller's arrays may be released as soon as this returns
  hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  table_dirty_ = true;
  drop_graphs();
  return id;
}

void Engine::remove(int id) {
  step_sig_.clear();
  check_id(id);
  trees_.erase(id);
  Sys& s = sys_[id];
  // Clear the range so stale values can never be read as sources (one
  // launch for every array).
  hip_check(launch_clear_range(d_arena_, s.off, s.span, CLEAR_RETIRED, stream_),
            "clear launch");
  int64_t off = s.off, len = s.span;
  auto nx = free_.lower_bound(off);
  if (nx != free_.end() && nx->first == off + len) {
    len += nx->second;
    free_.erase(nx);
  }
  auto pv = free_.lower_bound(off);
  if (pv != free_.begin()) {
    --pv;
    if (pv->first + pv->second == off) {
      off = pv->first;
      len += pv->second;
      free_.erase(pv);
    }
  }
  free_[off] = len;
  s = Sys{};
  table_dirty_ = true;
  drop_graphs();
}

void Engine::set_params(int id, const SystemParams& p) {
  step_sig_.clear();
  check_id(id);
  SystemParams q = p;
  q.n = sys_[id].p.n;
  validate(q);
  Sys& s = sys_[id];
  const bool g_changed = q.g != s.p.g;
  if (q.integ != s.p.integ || q.dt != s.p.dt || q.eps2 != s.p.eps2 ||
      g_changed) {
    if (s.vv_primed) s.prepared = false;  // snapshot holds next positions
    s.vv_primed = false;
  }
  if (g_changed) s.prepared = false;
  s.p = q;
  table_dirty_ = true;
  drop_graphs();
}

void Engine::set_mode(int mode) {
  step_sig_.clear();
  if (mode < MODE_F64_STRICT || mode > MODE_F32R)
    throw Error(NBODY_E_INVALID, "unknown precision");
  if (mode == mode_) return;
  mode_ = mode;
  for (auto& s : sys_) {
    s.vv_primed = false;
    s.prepared = false;
    s.sorted_at = -1;
  }
  table_dirty_ = true;  // split counts depend on the mode
  drop_graphs();
}

void Engine::set_gravity(int gravity, double theta) {
  if (gravity != GRAV_DIRECT && gravity != GRAV_BARNES_HUT)
    throw Error(NBODY_E_INVALID, "unknown gravity method");
  if (gravity == GRAV_BARNES_HUT && !(std::isfinite(theta) && theta >= 0.0))
    throw Error(NBODY_E_INVALID, "theta must be finite and >= 0");
  if (gravity == gravity_ && (gravity == GRAV_DIRECT || theta == theta_)) return;
  gravity_ = gravity;
  theta_ = theta;
  step_sig_.clear();
  for (auto& s : sys_) {
    s.vv_primed = false;
    s.prepared = false;
    s.sorted_at = -1;
  }
  drop_graphs();
}

void Engine::bh_round(int id, int stage, int src) {
  upload_table();
  std::unique_ptr<BarnesHut>& t = trees_[id];
  if (!t) t.reset(new BarnesHut(sys_[id].p.n));
  t->evaluate(stream_, mode_ == MODE_F64_STRICT, d_arena_, d_sys_, id,
              h_sys_[id], stage, src, theta_);
}

void Engine::bh_step(const std::vector<int>& ids, int64_t t) {
  for (int id : ids) {
    const Sys& s = sys_[id];
    if (s.p.integ == INTEG_RK4) {
      for (int r = 0; r < 4; ++r) bh_round(id, STAGE_RK0 + r, s.cur ^ (r & 1));
    } else {
      bh_round(id, s.p.integ == INTEG_EULER ? STAGE_EULER : STAGE_VV,
               s.cur ^ (int)(t & 1));
    }
  }
}

int64_t Engine::tree_nodes(int id) {
  check_id(id);
  auto it = trees_.find(id);
  return it == trees_.end() ? 0 : it->second->node_count(stream_);
}

SysDesc Engine::desc_of(int id) const {
  const Sys& s = sys_[id];
  SysDesc d{};
  d.off = s.off;
  d.n = (int32_t)s.p.n;
  d.integ = s.p.integ;
  d.dt = s.p.dt;
  d.eps2 = s.p.eps2;
  d.g = s.p.g;
  d.eps2f = (float)s.p.eps2;
  d.nsplit = (int16_t)split_of(id);
  d.flags = 0;
  if (d.eps2f > 0.0f) {
    // the self term's k = g*m * rsq(eps2f)^3 in f32: finite with a margin
    // for the rounding of g*m and the approximate v_rsq_f32 (a few ulp)
    const double inv = 1.0 / std::sqrt((double)d.eps2f);
    const double kmax = std::fabs(s.p.g) * s.mmax * ((inv * inv) * inv);
    if (kmax < 0x1p120) d.flags |= SYS_F32_SELF_FINITE;
  }
  return d;
}

// Source-range parts per tile for one system: looked up by the bucket
// floor(log2 n) in the process-wide split table (tuning.cpp), then clamped so
// every part keeps at least 64 sources.  Depends only on the system's own
// size and the arithmetic mode (never on what shares the launch), so a
// system's summation order -- and its bits -- are the same alone or in a
// batch.  Strict mode keeps the reference's single sequential sum.
int Engine::split_of(int id) const {
  return mode_ == MODE_F64_STRICT ? 1 : sys_[id].split;
}

// The force-kernel variant of a system in the current mode: chosen by its
// own size when it was uploaded, so (like the split) it never depends on
// what shares its launches.
int Engine::variant_of(int id) const { return sys_[id].variant[mode_]; }

// The split a system gets for its lifetime, fixed when it is uploaded so the
// host work lists and the device descriptor can never disagree.
int split_for_size(int64_t n) {
  int k = tuning_split(n);
  while (k > 1 && n / k < 64) --k;
  return std::max(1, std::min(k, kMaxSplit));
}

void Engine::upload_table() {
  if (!table_dirty_) return;
  for (int k = 0; k < max_sys_; ++k) {
    if (!sys_[k].live) continue;
    h_sys_[k] = desc_of(k);
  }
  hip_check(hipMemcpyAsync(d_sys_, h_sys_.data(), sizeof(SysDesc) * max_sys_,
                           hipMemcpyHostToDevice, stream_),
            "hipMemcpy(system table)");
  // the host copy must stay unchanged until the copy has been consumed
  hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  table_dirty_ = false;
}

void Engine::ensure_items(size_t n) {
  if (n <= items_cap_) return;
  hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  if (d_items_) (void)hipFree(d_items_);
  d_items_ = nullptr;
  items_cap_ = std::max<size_t>(n, 2 * items_cap_);
  hip_check(hipMalloc((void**)&d_items_, sizeof(WorkItem) * items_cap_),
            "hipMalloc(work list)");
}

// Force items of one system: one per (tile, part).  With combine != null the
// system's combine items (one per tile) are appended there when it is split.
void Engine::tiles_of(int id, int stage, int src, std::vector<WorkItem>& out,
                      std::vector<WorkItem>* combine) const {
  const int64_t tiles = (sys_[id].p.n + kWave - 1) / kWave;
  const int parts = combine ? split_of(id) : 1;
  for (int64_t t = 0; t < tiles; ++t)
    for (int q = 0; q < parts; ++q)
      out.push_back(WorkItem{id, (int32_t)t, (int16_t)stage, (int8_t)src,
                             (int8_t)q});
  if (combine && parts > 1)
    for (int64_t t = 0; t < tiles; ++t)
      combine->push_back(
          WorkItem{id, (int32_t)t, (int16_t)stage, (int8_t)src, (int8_t)0});
}

// One force round outside the stepping loop (prime, plain accelerations,
// potential): uploads the force items and, for split systems, the combine
// items, launches both, and waits.
void Engine::run_round(std::vector<WorkItem>& items,
                       std::vector<WorkItem>& combine, bool potential) {
  if (items.empty()) return;
  sort_items(items);
  validate_items(items);
  validate_items(combine);
  std::vector<WorkItem> all(items);
  all.insert(all.end(), combine.begin(), combine.end());
  ensure_items(all.size());
  hip_check(hipMemcpyAsync(d_items_, all.data(), sizeof(WorkItem) * all.size(),
                           hipMemcpyHostToDevice, stream_),
            "hipMemcpy(work list)");
  const int nf = (int)items.size(), nc = (int)combine.size();
  // one force launch per kernel variant (items are grouped by variant)
  for (int a = 0; a < nf;) {
    const int v = variant_of(items[a].sys);
    int b = a;
    while (b < nf && variant_of(items[b].sys) == v) ++b;
    force_launch(d_items_ + a, b - a, v, potential);
    a = b;
  }
  combine_launch(d_items_ + nf, nc);
  // the pageable host list must not change before the copy is consumed
  hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
}

// Rebuilds gm and the current snapshot of the listed systems that were
// never prepared (or whose snapshot a parameter or mode change invalidated).
void Engine::prepare(const std::vector<int>& ids) {
  std::vector<WorkItem> items;
  std::vector<int> fresh;
  for (int id : ids) {
    Sys& s = sys_[id];
    if (s.prepared) continue;
    fresh.push_back(id);
    const int stage = (s.p.integ == INTEG_VV && s.vv_primed) ? STAGE_VV_PRIME
                                                             : STAGE_ACC;
    tiles_of(id, stage, s.cur, items);
    s.prepared = true;
  }
  upload_table();
  if (items.empty()) return;
  validate_items(items);
  ensure_items(items.size());
  hip_check(hipMemcpyAsync(d_items_, items.data(),
                           sizeof(WorkItem) * items.size(),
                           hipMemcpyHostToDevice, stream_),
            "hipMemcpy(work list)");
  hip_check(launch_prepare(d_arena_, d_items_, (int)items.size(), d_sys_, stream_),
            "prepare launch");
  // MODE_F32R reads sources from the slot-ordered copy: refresh it where the
  // order is already set (otherwise the next step's sort builds it)
  if (mode_ == MODE_F32R && gravity_ == GRAV_DIRECT) {
    std::vector<SortSys> tab;
    int max_tiles = 0;
    for (int id : fresh) {
      const Sys& s = sys_[id];
      if (s.sorted_at < 0) continue;
      SortSys e{};
      e.off = s.off;
      e.n = (int32_t)s.p.n;
      tab.push_back(e);
      max_tiles = std::max<int>(max_tiles, (int)((e.n + kWave - 1) / kWave));
    }
    if (!tab.empty()) {
      upload_sort_tab(tab);
      hip_check(launch_sort_scatter(d_arena_, d_sort_tab_, (int)tab.size(),
                                    max_tiles, nullptr, stream_),
                "snapshot copy launch");
    }
  }
  hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
}

// Work lists of one step for the listed systems: force[r] and combine[r]
// for rounds 0..3, with each system's snapshot parity as it will be when that
// round runs, given `flip` extra steps already taken by Euler /
// velocity-Verlet systems (RK4 systems return to the same buffer after their
// four stages).
// The item spans of one step's lists: force and combine runs of every
// listed system per round (RK4 systems in all four rounds, Euler and VV
// ones in the first), longest systems first.  flip: the snapshot parity of
// the second of two consecutive steps.
void Engine::step_spans(const std::vector<int>& order, int flip,
                        std::vector<ItemSpan> (&force)[4][kForceVariants],
                        std::vector<ItemSpan> (&combine)[4]) const {
  for (int r = 0; r < 4; ++r) {
    for (auto& f : force[r]) f.clear();
    combine[r].clear();
    for (int id : order) {
      const Sys& s = sys_[id];
      int stage, src;
      if (s.p.integ == INTEG_RK4) {
        stage = STAGE_RK0 + r;
        src = s.cur ^ (r & 1);
      } else if (r == 0) {
        stage = s.p.integ == INTEG_EULER ? STAGE_EULER : STAGE_VV;
        src = s.cur ^ flip;
      } else {
        continue;
      }
      const int32_t tiles = (int32_t)((s.p.n + kWave - 1) / kWave);
      const int parts = split_of(id);
      force[r][variant_of(id)].push_back(
          ItemSpan{0, id, tiles, (int16_t)stage, (int8_t)src, (int8_t)parts});
      if (parts > 1)
        combine[r].push_back(
            ItemSpan{0, id, tiles, (int16_t)stage, (int8_t)src, (int8_t)1});
    }
  }
}

std::vector<int> Engine::by_size(const std::vector<int>& ids) const {
  std::vector<int> order(ids);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return sys_[a].p.n > sys_[b].p.n;
  });
  return order;
}

void Engine::sort_items(std::vector<WorkItem>& items) const {
  // grouped by kernel variant (one launch each), longest systems first
  // within a group: the dispatcher then packs the short tail last
  std::stable_sort(items.begin(), items.end(),
                   [&](const WorkItem& a, const WorkItem& b) {
                     const int va = variant_of(a.sys), vb = variant_of(b.sys);
                     if (va != vb) return va < vb;
                     return sys_[a.sys].p.n > sys_[b.sys].p.n;
                   });
}

// Primes velocity-Verlet systems that carry no acceleration yet.
void Engine::prime(const std::vector<int>& ids) {
  std::vector<WorkItem> items, combine;
  for (int id : ids) {
    const Sys& s = sys_[id];
    if (s.p.integ != INTEG_VV || s.vv_primed) continue;
    if (gravity_ == GRAV_BARNES_HUT)
      bh_round(id, STAGE_VV_PRIME, s.cur);
    else
      tiles_of(id, STAGE_VV_PRIME, s.cur, items, &combine);
  }
  if (!items.empty()) run_round(items, combine);
  for (int id : ids) {
    Sys& s = sys_[id];
    if (s.p.integ == INTEG_VV && !s.vv_primed) {
      s.vv_primed = true;
      s.cur ^= 1;
    }
  }
}

// Uploads the lists of step A (current parity) and step B (after one step)
// into the step-list device buffer.  Slot k = which*8 + round*2 + (0 force,
// 1 combine).  The lists depend only on the listed systems' ids, sizes,
// integrators, snapshot parity and the mode; when none of those changed
// since the last upload the device copy is reused as it is.
void Engine::upload_step_lists(const std::vector<int>& ids, StepLists& L) {
  std::vector<int64_t> sig;
  sig.reserve(4 * ids.size() + 2);
  sig.push_back(mode_);
  sig.push_back((int64_t)ids.size());
  for (int id : ids) {
    const Sys& s = sys_[id];
    sig.push_back(id);
    sig.push_back(s.off);
    sig.push_back(s.p.n * 4 + s.p.integ);
    sig.push_back(s.cur);
  }
  if (sig == step_sig_ && d_step_items_) {
    L = step_lists_;
    return;
  }
  for (int id : ids) {
    check_id(id);
    const Sys& s = sys_[id];
    if (s.off < 0 || s.off + s.span > cap_ || s.span < s.p.n)
      throw Error(NBODY_E_STATE, "system range outside the arena");
  }
  const std::vector<int> order = by_size(ids);
  std::vector<ItemSpan> fa[4][kForceVariants], ca[4], fb[4][kForceVariants], cb[4];
  step_spans(order, 0, fa, ca);
  step_spans(order, 1, fb, cb);
  // one span array for every list; each list's items follow the previous
  std::vector<ItemSpan> spans;
  int64_t total = 0;
  for (int k = 0; k < StepLists::kLists; ++k) {
    const int which = k / (4 * StepLists::kPer), r = (k / StepLists::kPer) % 4,
              v = k % StepLists::kPer;
    const std::vector<ItemSpan>& l =
        v == kForceVariants ? (which == 0 ? ca[r] : cb[r])
                            : (which == 0 ? fa[r][v] : fb[r][v]);
    L.off[k] = (int)total;
    for (ItemSpan sp : l) {
      sp.first = (int32_t)total;
      total += (int64_t)sp.tiles * sp.parts;
      if (total > INT32_MAX) throw Error(NBODY_E_INVALID, "work list too long");
      spans.push_back(sp);
    }
    L.cnt[k] = (int)(total - L.off[k]);
  }
  // the previous lists may still be in use by queued launches
  hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  if ((size_t)total > step_items_cap_) {
    (void)hipFree(d_step_items_);
    d_step_items_ = nullptr;
    step_items_cap_ = std::max((size_t)total, 2 * step_items_cap_);
    hip_check(hipMalloc((void**)&d_step_items_,
                        sizeof(WorkItem) * step_items_cap_),
              "hipMalloc(step lists)");
    drop_graphs();  // graphs hold the old buffer address
  }
  if (spans.size() > step_spans_cap_) {
    (void)hipFree(d_step_spans_);
    d_step_spans_ = nullptr;
    step_spans_cap_ = std::max(spans.size(), 2 * step_spans_cap_);
    hip_check(hipMalloc((void**)&d_step_spans_,
                        sizeof(ItemSpan) * step_spans_cap_),
              "hipMalloc(step spans)");
  }
  if (!spans.empty()) {
    hip_check(hipMemcpyAsync(d_step_spans_, spans.data(),
                             sizeof(ItemSpan) * spans.size(),
                             hipMemcpyHostToDevice, stream_),
              "hipMemcpy(step spans)");
    hip_check(launch_expand(d_step_spans_, (int)spans.size(), (int)total,
                            d_step_items_, stream_),
              "expand launch");
    // the spans are host memory of this call
    hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  }
  step_sig_ = std::move(sig);
  step_lists_ = L;
}

void Engine::force_launch(const WorkItem* items, int nitems, int variant,
                          bool potential) {
  if (nitems <= 0) return;
  const bool timed = timer_ && timer_->on;
  if (timed) timer_->begin(stream_);
  hip_check(potential ? launch_potential(d_arena_, items, nitems, d_sys_, stream_)
                      : launch_force(mode_, variant, d_arena_, items, nitems,
                                     d_sys_, stream_),
            "force launch");
  if (timed) timer_->end(stream_, KernelTimer::FORCE);
}

void Engine::combine_launch(const WorkItem* items, int nitems) {
  if (nitems <= 0) return;
  const bool timed = timer_ && timer_->on;
  if (timed) timer_->begin(stream_);
  hip_check(launch_combine(mode_, d_arena_, items, nitems, d_sys_, stream_),
            "combine launch");
  if (timed) timer_->end(stream_, KernelTimer::COMBINE);
}

void Engine::launch_step(const StepLists& L, int which) {
  for (int r = 0; r < 4; ++r) {
    for (int v = 0; v < kForceVariants; ++v) {
      const int k = StepLists::at(which, r, v);
      force_launch(d_step_items_ + L.off[k], L.cnt[k], v);
    }
    const int c = StepLists::at(which, r, kForceVariants);
    combine_launch(d_step_items_ + L.off[c], L.cnt[c]);
  }
}

void Engine::finish_steps(const std::vector<int>& ids, int64_t nsteps) {
  for (int id : ids) {
    Sys& s = sys_[id];
    if (s.p.integ != INTEG_RK4 && (nsteps & 1)) s.cur ^= 1;
    s.steps += nsteps;
    s.stage_next = -1;  // whole steps end any staged sequence
    s.acc_fresh = false;
  }
}

void Engine::step(const std::vector<int>& ids, int64_t nsteps) {
  run_steps(ids, nsteps, false, [](int64_t) {});
}

template <typename Hook>
void Engine::run_steps(const std::vector<int>& ids, int64_t nsteps,
                       bool graph, Hook&& after_step) {
  if (nsteps < 0) throw Error(NBODY_E_INVALID, "nsteps < 0");
  if (graph && gravity_ != GRAV_DIRECT)
    throw Error(NBODY_E_UNSUPPORTED, "graph stepping runs the direct sum only");
  if (timer_ && timer_->on) graph = false;  // launches bracketed one by one
  for (int id : ids) check_id(id);
  if (nsteps == 0 || ids.empty()) return;
  prepare(ids);
  if (gravity_ == GRAV_BARNES_HUT) {
    prime(ids);
    for (int64_t t = 0; t < nsteps; ++t) {
      bh_step(ids, t);
      after_step(t);
    }
    finish_steps(ids, nsteps);
    return;
  }
  const std::vector<SortPass> plan = plan_sorts(ids, nsteps);
  size_t next = 0;
  run_sorts(plan, next, 0);
  prime(ids);
  StepLists L;
  upload_step_lists(ids, L);
  if (graph) {
    // one step of each snapshot parity (lists A and B) captured once into
    // two HIP graphs and replayed alternately; they reference only the
    // device work-list buffer and counts, so they stay valid while the
    // composition (and therefore every count) is unchanged
    int64_t key = (int64_t)(uintptr_t)d_step_items_ * 31 + mode_;
    for (int k = 0; k < StepLists::kLists; ++k) key = key * 1000003 + L.cnt[k];
    if (key != gkey_ || ids != gids_ || !gexec_[0] || !gexec_[1]) {
      drop_graphs();
      for (int which = 0; which < 2; ++which) {
        hipGraph_t g = nullptr;
        hip_check(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal),
                  "hipStreamBeginCapture");
        try {
          launch_step(L, which);
        } catch (...) {
          (void)hipStreamEndCapture(stream_, &g);
          if (g) (void)hipGraphDestroy(g);
          throw;
        }
        hip_check(hipStreamEndCapture(stream_, &g), "hipStreamEndCapture");
        hipError_t e = hipGraphInstantiate(&gexec_[which], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        hip_check(e, "hipGraphInstantiate");
      }
      gkey_ = key;
      gids_ = ids;
    }
  }
  for (int64_t t = 0; t < nsteps; ++t) {
    run_sorts(plan, next, t);
    if (graph)
      hip_check(hipGraphLaunch(gexec_[t & 1], stream_), "hipGraphLaunch");
    else
      launch_step(L, (int)(t & 1));
    after_step(t);
  }
  finish_steps(ids, nsteps);
}

void Engine::ensure_stage(int which, size_t need) {
  Staging& st = stage_[which];
  if (need <= st.cap) return;
  hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  (void)hipFree(st.d);
  (void)hipHostFree(st.h);
  st.d = nullptr;
  st.h = nullptr;
  st.cap = std::max(need, 2 * st.cap);
  hip_check(hipMalloc((void**)&st.d, sizeof(double) * st.cap),
            "hipMalloc(gather staging)");
  hip_check(hipHostMalloc((void**)&st.h, sizeof(double) * st.cap,
                          hipHostMallocDefault),
            "hipHostMalloc(gather staging)");
}

const double* Engine::step_snapshots(const std::vector<int>& ids,
                                     int64_t nsteps, bool graph,
                                     const std::vector<Snap>& snaps,
                                     std::vector<int64_t>& rows,
                                     int64_t& total,
                                     const std::vector<int64_t>* counts,
                                     const std::function<bool(int64_t)>& boundary,
                                     const std::function<void(int64_t, const double*)>& deliver,
                                     int64_t* ran) {
  rows.clear();
  total = 0;
  if (ran) *ran = nsteps;
  if (counts) {
    if (counts->size() != ids.size())
      throw Error(NBODY_E_INVALID, "one step count per system");
    for (int64_t c : *counts)
      if (c < 0 || c > nsteps)
        throw Error(NBODY_E_INVALID, "step count outside the call");
  }
  auto count_of = [&](int id) {
    if (!counts) return nsteps;
    for (size_t q = 0; q < ids.size(); ++q)
      if (ids[q] == id) return (*counts)[q];
    return (int64_t)-1;
  };
  // gather lists, grouped by step: [first[t], first[t+1]) are step t's items
  std::vector<Snap> order(snaps);
  std::stable_sort(order.begin(), order.end(),
                   [](const Snap& a, const Snap& b) { return a.t < b.t; });
  std::vector<WorkItem> items;
  std::vector<std::pair<int64_t, std::pair<int, int>>> at;  // t -> range
  std::vector<int64_t> row_of(snaps.size());
  for (const Snap& sn : order) {
    check_id(sn.id);
    if (sn.t < 0 || sn.t >= nsteps || (counts && sn.t >= count_of(sn.id)))
      throw Error(NBODY_E_INVALID, "snapshot step outside the call");
  }
  for (size_t k = 0; k < order.size(); ++k) {
    const Snap& sn = order[k];
    if (at.empty() || at.back().first != sn.t)
      at.push_back({sn.t, {(int)items.size(), (int)items.size()}});
    const int64_t n = sys_[sn.id].p.n;
    const int64_t tiles = (n + kWave - 1) / kWave;
    for (int64_t t = 0; t < tiles; ++t) {
      WorkItem w{sn.id, (int32_t)t, (int16_t)STAGE_ACC, (int8_t)0, (int8_t)0};
      w.pad_ = (int32_t)total;
      items.push_back(w);
    }
    at.back().second.second = (int)items.size();
    rows.push_back(total);
    total += n;
  }
  const std::vector<int64_t> srow(rows);  // rows in step order
  // rows in the caller's order
  {
    std::vector<int64_t> r(snaps.size());
    std::vector<size_t> idx(snaps.size());
    for (size_t k = 0; k < idx.size(); ++k) idx[k] = k;
    std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) {
      return snaps[a].t < snaps[b].t;
    });
    for (size_t k = 0; k < idx.size(); ++k) r[idx[k]] = rows[k];
    rows.swap(r);
  }
  if (total > INT32_MAX) throw Error(NBODY_E_INVALID, "snapshots too large");
  if (!items.empty()) {
    validate_items(items);
    ensure_stage(0, (size_t)total * 3);
    if (items.size() > snap_items_cap_) {
      hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
      (void)hipFree(d_snap_items_);
      d_snap_items_ = nullptr;
      snap_items_cap_ = std::max(items.size(), 2 * snap_items_cap_);
      hip_check(hipMalloc((void**)&d_snap_items_,
                          sizeof(WorkItem) * snap_items_cap_),
                "hipMalloc(snapshot lists)");
    }
    hip_check(hipMemcpyAsync(d_snap_items_, items.data(),
                             sizeof(WorkItem) * items.size(),
                             hipMemcpyHostToDevice, stream_),
              "hipMemcpy(snapshot lists)");
    hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  }
  // rows [0, on_host) of every coordinate are copied to the host, or queued
  // to be (step order)
  int64_t on_host = 0;
  auto copy_before = [&](int64_t t) {  // snapshots taken before step t
    size_t k = 0;
    while (k < order.size() && order[k].t < t) ++k;
    const int64_t end = k < order.size() ? srow[k] : total;
    if (end <= on_host) return;
    for (int f = 0; f < 3; ++f)
      hip_check(hipMemcpyAsync(stage_[0].h + f * total + on_host,
                               stage_[0].d + f * total + on_host,
                               sizeof(double) * (end - on_host),
                               hipMemcpyDeviceToHost, stream_),
                "hipMemcpy(snapshots)");
    on_host = end;
  };
  int64_t stop = nsteps;
  size_t next = 0;
  auto gather_after = [&](int64_t t) {
    for (; next < at.size() && at[next].first == t; ++next) {
      const int lo = at[next].second.first, hi = at[next].second.second;
      hip_check(launch_gather(d_arena_, d_snap_items_ + lo, hi - lo, d_sys_,
                              stage_[0].d, total, 3, stream_),
                "gather launch");
    }
  };
  if (!counts) {
    run_steps(ids, nsteps, graph, gather_after);
  } else {
    // segments [t0, e) between the distinct counts, each over the systems
    // whose count reaches e
    std::vector<int64_t> ends(counts->begin(), counts->end());
    std::sort(ends.begin(), ends.end());
    ends.erase(std::unique(ends.begin(), ends.end()), ends.end());
    const bool uniform = ends.size() == 1;
    int64_t t0 = 0, pending = -1;  // the boundary whose deliver is due
    auto flush = [&] {
      if (pending < 0) return;
      const int64_t t = pending;
      pending = -1;
      hip_check(hipEventSynchronize(ev_boundary_), "hipEventSynchronize");
      if (deliver) deliver(t, total ? stage_[0].h : nullptr);
    };
    for (int64_t e : ends) {
      if (e <= t0) continue;
      std::vector<int> active;
      for (size_t q = 0; q < ids.size(); ++q)
        if ((*counts)[q] >= e) active.push_back(ids[q]);
      run_steps(active, e - t0, graph && uniform,
                [&](int64_t t) { gather_after(t0 + t); });
      flush();  // the previous boundary's host work, while this segment runs
      t0 = e;
      if (t0 < nsteps && boundary) {
        copy_before(t0);
        const bool cut = boundary(t0);
        hip_check(hipEventRecord(ev_boundary_, stream_), "hipEventRecord");
        pending = t0;
        if (cut) {
          stop = t0;
          break;
        }
      }
    }
    flush();
  }
  if (ran) *ran = stop;
  if (total == 0) return nullptr;
  copy_before(stop);
  hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  return stage_[0].h;
}

void Engine::accelerations(int id, double* ax, double* ay, double* az) {
  check_id(id);
  Sys& s = sys_[id];
  if (s.vv_primed) {  // the snapshot holds next-step positions: rebuild
    s.vv_primed = false;
    s.prepared = false;
  }
  prepare({id});
  if (gravity_ == GRAV_BARNES_HUT) {
    bh_round(id, STAGE_ACC, s.cur);
  } else {
    const std::vector<SortPass> plan = plan_sorts({id}, 1);
    size_t next = 0;
    run_sorts(plan, next, 0);
    std::vector<WorkItem> items, combine;
    tiles_of(id, STAGE_ACC, s.cur, items, &combine);
    run_round(items, combine);
  }
  s.vv_primed = false;  // the carried acceleration was overwritten
  s.acc_fresh = true;   // a staged integrator step may use it
  if (ax || ay || az) {
    const size_t bytes = sizeof(double) * s.p.n;
    if (ax)
      hip_check(hipMemcpyAsync(ax, A_.ax + s.off, bytes, hipMemcpyDeviceToHost,
                               stream_),
                "hipMemcpy(acc)");
    if (ay)
      hip_check(hipMemcpyAsync(ay, A_.ay + s.off, bytes, hipMemcpyDeviceToHost,
                               stream_),
                "hipMemcpy(acc)");
    if (az)
      hip_check(hipMemcpyAsync(az, A_.az + s.off, bytes, hipMemcpyDeviceToHost,
                               stream_),
                "hipMemcpy(acc)");
    hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  }
}

void Engine::check_stage(int id, int op) const {
  check_id(id);
  if (op < OP_EULER || op > OP_RK4_4)
    throw Error(NBODY_E_INVALID, "unknown integrator stage");
  const Sys& s = sys_[id];
  // the stages of one step come in their order, each after an evaluation
  // of the accelerations at the positions the previous stage left
  if (!s.acc_fresh)
    throw Error(NBODY_E_STATE,
                "integrator stage without accelerations evaluated since the "
                "last stage or step");
  const bool first = op == OP_EULER || op == OP_VV_DRIFT || op == OP_RK4_1;
  if (s.stage_next >= 0 ? op != s.stage_next : !first)
    throw Error(NBODY_E_STATE, "integrator stage out of order");
}

void Engine::integrate_stage(int id, int op) {
  check_stage(id, op);
  Sys& s = sys_[id];
  upload_table();
  std::vector<WorkItem> items;
  tiles_of(id, STAGE_ACC, s.cur, items);
  validate_items(items);
  ensure_items(items.size());
  hip_check(hipMemcpyAsync(d_items_, items.data(), sizeof(WorkItem) * items.size(),
                           hipMemcpyHostToDevice, stream_),
            "hipMemcpy(work list)");
  hip_check(launch_stage(d_arena_, d_items_, (int)items.size(), d_sys_, op, stream_),
            "stage launch");
  // the pageable host list must not change before the copy is consumed
  hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  s.prepared = false;    // positions moved: rebuild the snapshot
  s.vv_primed = false;   // a carried acceleration no longer applies
  if (op == OP_EULER || op == OP_VV_KICK || op == OP_RK4_4) ++s.steps;
  s.stage_next = op == OP_VV_DRIFT ? OP_VV_KICK
                 : (op >= OP_RK4_1 && op < OP_RK4_4) ? op + 1
                                                     : -1;
  s.acc_fresh = false;
}

void Engine::diagnostics(const std::vector<int>& ids, double* out) {
  for (int id : ids) check_id(id);
  if (ids.empty()) return;
  // the potential pass reads the f64 snapshot of the current positions
  const int saved = mode_;
  std::vector<WorkItem> items;
  std::vector<int32_t> list(ids.begin(), ids.end());
  // f64 snapshot of the master positions into the spare buffer
  std::vector<WorkItem> prep;
  for (int id : ids) tiles_of(id, STAGE_ACC, sys_[id].cur ^ 1, prep);
  upload_table();
  validate_items(prep);
  ensure_items(prep.size());
  hip_check(hipMemcpyAsync(d_items_, prep.data(), sizeof(WorkItem) * prep.size(),
                           hipMemcpyHostToDevice, stream_),
            "hipMemcpy(work list)");
  hip_check(launch_prepare(d_arena_, d_items_, (int)prep.size(), d_sys_, stream_),
            "prepare launch");
  hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  std::vector<WorkItem> combine;
  for (int id : ids) tiles_of(id, STAGE_POT, sys_[id].cur ^ 1, items, &combine);
  run_round(items, combine, /*potential=*/true);
  (void)saved;
  hip_check(hipMemcpyAsync(d_list_, list.data(), sizeof(int32_t) * list.size(),
                           hipMemcpyHostToDevice, stream_),
            "hipMemcpy(list)");
  hip_check(launch_diag(d_arena_, d_sys_, d_list_, (int)list.size(), d_diag_,
                        stream_),
            "diag launch");
  hip_check(hipMemcpyAsync(out, d_diag_, sizeof(double) * DIAG_COUNT * ids.size(),
                           hipMemcpyDeviceToHost, stream_),
            "hipMemcpy(diag)");
  hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
}

void Engine::download(int id, double* m, double* x, double* y, double* z,
                      double* vx, double* vy, double* vz) {
  check_id(id);
  const Sys& s = sys_[id];
  const size_t bytes = sizeof(double) * s.p.n;
  double* dst[7] = {m, x, y, z, vx, vy, vz};
  const double* src[7] = {A_.m, A_.x, A_.y, A_.z, A_.vx, A_.vy, A_.vz};
  for (int q = 0; q < 7; ++q)
    if (dst[q])
      hip_check(hipMemcpyAsync(dst[q], src[q] + s.off, bytes,
                               hipMemcpyDeviceToHost, stream_),
                "hipMemcpy(download)");
  hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
}

const double* Engine::gather(const std::vector<int>& ids, int nfields,
                             std::vector<int64_t>& rows, int64_t& total) {
  const double* h = gather_async(ids, nfields, rows, total);
  hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  return h;
}

const double* Engine::gather_async(const std::vector<int>& ids, int nfields,
                                   std::vector<int64_t>& rows,
                                   int64_t& total) {
  if (nfields != 3 && nfields != 6)
    throw Error(NBODY_E_INVALID, "gather: nfields must be 3 or 6");
  // the previous gather's list and host view are reused
  if (gather_queued_) {
    hip_check(hipEventSynchronize(ev_gather_), "hipEventSynchronize");
    gather_queued_ = false;
  }
  rows.clear();
  total = 0;
  std::vector<WorkItem>& items = gather_items_;
  items.clear();
  for (int id : ids) {
    check_id(id);
    rows.push_back(total);
    const int64_t tiles = (sys_[id].p.n + kWave - 1) / kWave;
    for (int64_t t = 0; t < tiles; ++t) {
      WorkItem w{id, (int32_t)t, (int16_t)STAGE_ACC, (int8_t)0, (int8_t)0};
      w.pad_ = (int32_t)total;
      items.push_back(w);
    }
    total += sys_[id].p.n;
  }
  if (total == 0) return nullptr;
  if (total > INT32_MAX) throw Error(NBODY_E_INVALID, "gather too large");
  const size_t need = (size_t)total * nfields;
  ensure_stage(1, need);
  validate_items(items);
  ensure_items(items.size());
  hip_check(hipMemcpyAsync(d_items_, items.data(), sizeof(WorkItem) * items.size(),
                           hipMemcpyHostToDevice, stream_),
            "hipMemcpy(work list)");
  upload_table();
  hip_check(launch_gather(d_arena_, d_items_, (int)items.size(), d_sys_,
                          stage_[1].d, total, nfields, stream_),
            "gather launch");
  hip_check(hipMemcpyAsync(stage_[1].h, stage_[1].d, sizeof(double) * need,
                           hipMemcpyDeviceToHost, stream_),
            "hipMemcpy(gather)");
  hip_check(hipEventRecord(ev_gather_, stream_), "hipEventRecord");
  gather_queued_ = true;
  return stage_[1].h;
}

// The step loop of step() with one gather launch after every step.  The
// gather of step s (s counted within the current buffer fill) writes field
// f at d_traj_[f * C * n + s * n + i], C = steps per fill, so each field of
// a fill is one contiguous run of s-major rows and leaves in one copy.
void Engine::record(int id, int64_t nsteps, double* x, double* y, double* z,
                    size_t max_bytes) {
  if (nsteps < 0) throw Error(NBODY_E_INVALID, "nsteps < 0");
  check_id(id);
  if (nsteps == 0) return;
  if (!x || !y || !z) throw Error(NBODY_E_INVALID, "null trajectory array");
  const int64_t n = sys_[id].p.n;
  const int64_t per_step = 3 * n;  // doubles
  int64_t C = (int64_t)(max_bytes / (sizeof(double) * (size_t)per_step));
  C = std::max<int64_t>(1, std::min(C, nsteps));
  if ((size_t)(C * per_step) > traj_cap_) {
    hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
    (void)hipFree(d_traj_);
    d_traj_ = nullptr;
    traj_cap_ = (size_t)(C * per_step);
    hip_check(hipMalloc((void**)&d_traj_, sizeof(double) * traj_cap_),
              "hipMalloc(trajectory)");
  }
  std::vector<WorkItem> gitems;
  const int64_t tiles = (n + kWave - 1) / kWave;
  for (int64_t t = 0; t < tiles; ++t) {
    WorkItem w{id, (int32_t)t, (int16_t)STAGE_ACC, (int8_t)0, (int8_t)0};
    w.pad_ = 0;
    gitems.push_back(w);
  }
  prepare({id});
  const bool bh = gravity_ == GRAV_BARNES_HUT;
  std::vector<SortPass> plan;
  size_t next = 0;
  if (!bh) {
    plan = plan_sorts({id}, nsteps);
    run_sorts(plan, next, 0);
  }
  prime({id});
  StepLists L;
  if (!bh) upload_step_lists({id}, L);
  validate_items(gitems);
  ensure_items(gitems.size());
  upload_table();
  hip_check(hipMemcpyAsync(d_items_, gitems.data(),
                           sizeof(WorkItem) * gitems.size(),
                           hipMemcpyHostToDevice, stream_),
            "hipMemcpy(work list)");
  double* dst[3] = {x, y, z};
  int64_t done = 0;
  while (done < nsteps) {
    const int64_t c = std::min(C, nsteps - done);
    for (int64_t s = 0; s < c; ++s) {
      if (bh) {
        bh_step({id}, done + s);
      } else {
        run_sorts(plan, next, done + s);
        launch_step(L, (int)((done + s) & 1));
      }
      hip_check(launch_gather(d_arena_, d_items_, (int)gitems.size(), d_sys_,
                              d_traj_ + s * n, C * n, 3, stream_),
                "gather launch");
    }
    for (int f = 0; f < 3; ++f)
      hip_check(hipMemcpyAsync(dst[f] + done * n, d_traj_ + f * C * n,
                               sizeof(double) * (size_t)(c * n),
                               hipMemcpyDeviceToHost, stream_),
                "hipMemcpy(trajectory)");
    hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
    done += c;
  }
  finish_steps({id}, nsteps);
}

// step() with the launches of one step of each snapshot parity captured
// into HIP graphs and replayed (see run_steps).
void Engine::step_graph(const std::vector<int>& ids, int64_t nsteps) {
  run_steps(ids, nsteps, true, [](int64_t) {});
}

// The box in which MODE_F32R's target order quantises a system (k_sort_keys:
// position 0.5 + (v - o) * scale per axis, 16 bits, then the Hilbert index).
// Width 4 * ext about the centre of mass, advanced by vcom * time.  An axis
// along which the system is thin (extent at most ext / 8, the normal of a
// disc) is scaled 16 times more coarsely and its centre moved off the
// box's mid-plane, to 0.5 + 2^-7: the slab then lies inside one cell of the
// curve at every level down to cells of 2^-6 of the box.  Centred, the
// curve's first split would cut the disc along its mid-plane, and every
// 64-body run would hold bodies of one half of its thickness only, over
// twice the area.  Depends only on the system's own state (as the order
// must, so a system orders alike alone and batched).
static void sort_box(const double com[3], const double vcom[3],
                     const double ext3[3], double time, SortSys& e) {
  const double ext = std::max({ext3[0], ext3[1], ext3[2]});
  const double inv = ext > 0.0 ? 1.0 / (4.0 * ext) : 1.0;
  double o[3], sc[3];
  for (int q = 0; q < 3; ++q) {
    const double c = com[q] + vcom[q] * time;
    if (ext > 0.0 && ext3[q] <= 0.125 * ext) {
      sc[q] = inv * 0x1p-4;
      o[q] = c - 0x1p-7 / sc[q];  // c maps to 0.5 + 2^-7
    } else {
      sc[q] = inv;
      o[q] = c;
    }
  }
  e.ox = o[0];
  e.oy = o[1];
  e.oz = o[2];
  e.ix = sc[0];
  e.iy = sc[1];
  e.iz = sc[2];
}

// Target-ordering passes of one call (see SortPass).  Runs only in
// MODE_F32R with the direct sum; otherwise the plan is empty.
std::vector<Engine::SortPass> Engine::plan_sorts(const std::vector<int>& ids,
                                                 int64_t nsteps) {
  std::vector<SortPass> plan;
  if (mode_ != MODE_F32R || gravity_ != GRAV_DIRECT || ids.empty()) return plan;
  std::vector<SortSys> tab;
  std::vector<int64_t> last(ids.size(), -2);
  for (int64_t t = 0; t < nsteps; ++t) {
    SortPass p{t, 0, 0, 0, (int)tab.size()};
    for (size_t q = 0; q < ids.size(); ++q) {
      const Sys& s = sys_[ids[q]];
      const int64_t T = s.steps + t;
      const int64_t prev = last[q] >= -1 ? last[q] : s.sorted_at;
      const bool due = (t == 0 && prev < 0) || (T % kResortEvery == 0 && prev != T);
      if (!due) continue;
      last[q] = T;
      SortSys e;
      e.off = s.off;
      e.n = (int32_t)s.p.n;
      e.base = p.total;
      const double time = (double)T * s.p.dt;  // the centre moves with vcom
      sort_box(s.com, s.vcom, s.ext3, time, e);
      tab.push_back(e);
      ++p.nsys;
      p.total += e.n;
      p.max_tiles = std::max<int>(p.max_tiles, (int)((e.n + kWave - 1) / kWave));
    }
    if (p.nsys > 0) plan.push_back(p);
  }
  if (plan.empty()) return plan;
  for (size_t q = 0; q < ids.size(); ++q)
    if (last[q] >= 0) sys_[ids[q]].sorted_at = last[q];
  upload_sort_tab(tab);
  return plan;
}

// Copies a table of SortSys rows to d_sort_tab_ (grown as needed) and waits
// for the copy (the rows are host memory of the caller).
void Engine::upload_sort_tab(const std::vector<SortSys>& tab) {
  if (tab.size() > sort_tab_cap_) {
    hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
    (void)hipFree(d_sort_tab_);
    d_sort_tab_ = nullptr;
    sort_tab_cap_ = std::max(tab.size(), 2 * sort_tab_cap_);
    hip_check(hipMalloc((void**)&d_sort_tab_, sizeof(SortSys) * sort_tab_cap_),
              "hipMalloc(sort table)");
  }
  if (!d_keys_) {
    const size_t len = (size_t)cap_ + kPad;
    hip_check(hipMalloc((void**)&d_keys_, 2 * sizeof(uint64_t) * len),
              "hipMalloc(sort keys)");
    hip_check(hipMalloc((void**)&d_vals_, 2 * sizeof(int32_t) * len),
              "hipMalloc(sort values)");
    hip_check(sort_pairs(nullptr, sort_tmp_bytes_, nullptr, nullptr, nullptr,
                         nullptr, (int)len, 64, stream_),
              "radix sort sizing");
    hip_check(hipMalloc(&d_sort_tmp_, sort_tmp_bytes_), "hipMalloc(sort scratch)");
  }
  // the table is host memory of this call: copy, then wait for the copy
  hip_check(hipMemcpyAsync(d_sort_tab_, tab.data(), sizeof(SortSys) * tab.size(),
                           hipMemcpyHostToDevice, stream_),
            "hipMemcpy(sort table)");
  hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
}

void Engine::run_sorts(const std::vector<SortPass>& plan, size_t& next,
                       int64_t t) {
  const size_t len = (size_t)cap_ + kPad;
  for (; next < plan.size() && plan[next].t == t; ++next) {
    const SortPass& p = plan[next];
    int bits = 48;
    while ((1 << (bits - 48)) < p.nsys) ++bits;  // rank bits above the code
    hip_check(launch_sort_keys(d_arena_, d_sort_tab_ + p.tab, p.nsys,
                               p.max_tiles, d_keys_, d_vals_, stream_),
              "sort keys launch");
    size_t bytes = sort_tmp_bytes_;
    hip_check(sort_pairs(d_sort_tmp_, bytes, d_keys_, d_keys_ + len, d_vals_,
                         d_vals_ + len, p.total, bits, stream_),
              "radix sort");
    hip_check(launch_sort_scatter(d_arena_, d_sort_tab_ + p.tab, p.nsys,
                                  p.max_tiles, d_vals_ + len, stream_),
              "sort scatter launch");
  }
}

}  // namespace nbody
