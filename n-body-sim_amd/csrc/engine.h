// Host engine: one device arena holding any number of independent systems,
// stepped together in shared launches.
//
// The single-system C API (a DeviceBodies counterpart), the batched API and
// the request server are all built on this one class.  A step of the engine
// is up to four force rounds: round 0 carries every listed system (the only
// evaluation of Euler and carried-acceleration velocity Verlet, RK4's first
// stage); rounds 1-3 carry only RK4 systems.  Each round is one launch over a
// work list of (system, 64-target tile) items, longest systems first.
#pragma once
#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "nbody_types.h"

namespace nbody {

class BarnesHut;

// Error carrying a public NBODY_E_* code.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void hip_check(hipError_t e, const char* what);

// Optional timing of force and combine launches: while `on`, each launch is
// bracketed by two HIP events on the engine's stream (from a pool, reused
// after every harvest) and harvest() adds up the elapsed times per kind.
// One timer serves every engine of a context; use from one thread.
class KernelTimer {
 public:
  enum Kind { FORCE = 0, COMBINE = 1 };
  KernelTimer() = default;
  KernelTimer(const KernelTimer&) = delete;
  KernelTimer& operator=(const KernelTimer&) = delete;
  ~KernelTimer();
  bool on = false;
  void begin(hipStream_t s);          // before a launch
  void end(hipStream_t s, Kind kind);  // after it
  void harvest();                     // waits for the recorded launches
  void reset();                       // harvest, then zero the sums
  double ms[2] = {0.0, 0.0};
  int64_t launches[2] = {0, 0};

 private:
  std::vector<hipEvent_t> ev_;
  std::vector<Kind> kind_;
  size_t used_ = 0;  // events recorded since the last harvest
};

// Source-range parts of a system of n bodies (tuning table by floor(log2 n),
// clamped so every part keeps at least 64 sources, at most kMaxSplit); the
// engine clamps it further by its split cap.
int split_for_size(int64_t n);

struct SystemParams {
  int64_t n = 0;
  int integ = INTEG_EULER;
  double dt = 0.0, g = 1.0, eps2 = 0.0;
};

class Engine {
 public:
  // capacity: bodies the arena can hold; max_systems: table size.
  Engine(int device, hipStream_t stream, int mode, int64_t capacity,
         int max_systems);
  ~Engine();
  Engine(const Engine&) = delete;
  Engine& operator=(const Engine&) = delete;

  // Changing the arithmetic rebuilds every snapshot and drops carried state.
  void set_mode(int mode);
  // Force law (GRAV_*) and Barnes-Hut opening angle.  A change drops
  // carried accelerations and rebuilds the snapshots, as a mode change does.
  // Barnes-Hut evaluates in f64: the strict arithmetic in MODE_F64_STRICT,
  // the fast f64 arithmetic in every other mode.
  void set_gravity(int gravity, double theta);
  // Largest source-range split given to systems uploaded afterwards (the
  // tuning table's per-size choice is clamped to it).  A system keeps its
  // split for life, so its summation order never depends on what shares its
  // launches; the cap lets an API form with many resident systems (already
  // enough waves) use fewer, longer parts than one with a single system.
  void set_split_cap(int cap) { split_cap_ = cap < 1 ? 1 : cap; }
  // Octree nodes of a system's last Barnes-Hut build (synchronises).
  int64_t tree_nodes(int id);
  // Launch timing (may be null).  While it is on, graph stepping launches
  // the same kernels directly so that each one can be bracketed.
  void set_timer(KernelTimer* t) { timer_ = t; }

  // Uploads a system; returns its id.  Throws Error(NBODY_E_NOSPACE) when no
  // contiguous range of the arena is free.
  int add(const SystemParams& p, const double* m, const double* x,
          const double* y, const double* z, const double* vx,
          const double* vy, const double* vz);
  void remove(int id);
  // The arena's free ranges (offset -> length), and the placement rule of
  // add(): the first range, in offset order, that holds n bodies (rounded
  // up to a wave).  take_range() carves that range from `free` and returns
  // its offset, or -1 (leaving `free` as it was) when none holds n.
  // Replaying it on a copy tells which of several uploads will fit.
  std::map<int64_t, int64_t> free_ranges() const { return free_; }
  static int64_t take_range(std::map<int64_t, int64_t>& free, int64_t n) {
    const int64_t need = (n + kWave - 1) / kWave * kWave;
    auto it = free.begin();
    for (; it != free.end(); ++it)
      if (it->second >= need) break;
    if (it == free.end()) return -1;
    const int64_t off = it->first, len = it->second;
    free.erase(it);
    if (len > need) free[off + need] = len - need;
    return off;
  }
  int64_t capacity() const { return cap_; }
  int nsystems() const;
  const SystemParams& params(int id) const;
  int64_t steps_done(int id) const;
  void set_params(int id, const SystemParams& p);

  // Advances every listed system by nsteps of its own integrator.
  void step(const std::vector<int>& ids, int64_t nsteps);
  // One snapshot of a stepping call: positions of system `id` after step t
  // (0-based) of the call.
  struct Snap {
    int64_t t;
    int id;
  };
  // step() (graph = false) or step_graph() that also gathers the listed
  // snapshots on the device as the steps run -- a gather launch after each
  // step that has some -- and copies them to pinned host memory in one
  // transfer at the end.  Returns a host view valid until the next call:
  // coordinate f of snapshot k starts at base[f * total + rows[k]].
  // With `counts` (one per id, each in [0, nsteps]) system ids[q] takes only
  // counts[q] of the call's steps: the call runs in segments between the
  // distinct counts, each over the systems still stepping, so a short system
  // drops out without ending the call (graphs are used only when every
  // count is nsteps).  `boundary`, when given, is called at each boundary
  // t0 < nsteps between segments, once the steps before t0 are queued (the
  // systems with count t0 have taken their last step on the stream; nothing
  // has been waited for).  It may queue device work of its own
  // (gather_async) but must not read results.  When it returns true the
  // call stops there: *ran receives the steps run (systems took
  // min(count, *ran)) and snapshots at or after it hold no data.
  // `deliver(t0, hv)` follows each boundary once the snapshots taken before
  // step t0, and what the boundary queued, have reached the host (hv holds
  // them; null when the call has no snapshots).  It runs after the next
  // segment's launches are queued, so the host's handling overlaps the
  // device's stepping, and always before the next boundary.  Synchronous.
  const double* step_snapshots(
      const std::vector<int>& ids, int64_t nsteps, bool graph,
      const std::vector<Snap>& snaps, std::vector<int64_t>& rows,
      int64_t& total, const std::vector<int64_t>* counts = nullptr,
      const std::function<bool(int64_t)>& boundary = nullptr,
      const std::function<void(int64_t, const double*)>& deliver = nullptr,
      int64_t* ran = nullptr);
  // F(x) of one system into the device acceleration arrays (and optionally
  // to the host).  Drops a velocity-Verlet carry of that system.
  void accelerations(int id, double* ax, double* ay, double* az);
  // One integrator stage (StageOp) of one system on its acceleration arrays
  // (the result of accelerations()), with the system's dt.  Stages that move
  // positions invalidate the snapshot; the stages that end a step count it.
  // check_stage throws NBODY_E_STATE unless accelerations() ran since the
  // last stage or step and op is the stage that comes next (a first stage --
  // Euler, VV drift, RK4 stage 1 -- when no staged step is under way).
  void check_stage(int id, int op) const;
  void integrate_stage(int id, int op);
  // Per-system diagnostics, DIAG_COUNT doubles per listed system.
  void diagnostics(const std::vector<int>& ids, double* out);
  // Downloads the master state (any pointer may be null).  Synchronous.
  void download(int id, double* m, double* x, double* y, double* z,
                double* vx, double* vy, double* vz);
  // Gathers the state rows of several systems in one device pass and one
  // device-to-host copy into library-owned pinned memory.  nfields 3:
  // positions; 6: positions and velocities.  Returns a host view valid until
  // the next gather: field f of system ids[k] starts at
  // base[f * total + rows[k]].  Synchronous.
  const double* gather(const std::vector<int>& ids, int nfields,
                       std::vector<int64_t>& rows, int64_t& total);
  // gather() without the wait: the same view, valid once the stream has
  // passed this point (a deliver of step_snapshots() when called from its
  // boundary, or any synchronous call).  A later gather first waits for
  // this one.
  const double* gather_async(const std::vector<int>& ids, int nfields,
                             std::vector<int64_t>& rows, int64_t& total);

  // Graph-captured multi-step loop for a fixed set of systems (batched API).
  void step_graph(const std::vector<int>& ids, int64_t nsteps);

  // Advances one system nsteps steps and copies its positions after every
  // step to the host: step s of x at x[s*n .. s*n + n), likewise y and z.
  // Positions are kept in a device trajectory buffer of at most
  // max_bytes (at least one step) and copied out once per filled buffer.
  // Synchronous.
  void record(int id, int64_t nsteps, double* x, double* y, double* z,
              size_t max_bytes);


 private:
  struct Sys {
    bool live = false;
    SystemParams p;
    int64_t off = 0, span = 0;  // arena range (span: rounded allocation)
    int cur = 0;                // snapshot buffer holding the next sources
    bool vv_primed = false;
    bool prepared = false;
    int64_t steps = 0;
    int split = 1;              // source-range parts (fixed at upload)
    int variant[4] = {0, 0, 0, 0};  // force-kernel variant per Mode (fixed
                                    // at upload, from the tuning table)
    double com[3] = {0, 0, 0};  // centre of mass at upload (sort box centre,
    double vcom[3] = {0, 0, 0};  // advanced with the centre-of-mass velocity)
    double ext3[3] = {0, 0, 0};  // largest distance from com along each axis
                                 // (the sort box's size)
    double mmax = 0.0;          // largest |m| (bounds the f32 self term)
    int64_t sorted_at = -1;     // step count of the last body ordering
    // staged integration (integrate_stage): the stage that must come next
    // (-1: any first stage of a step) and whether accelerations were
    // evaluated since the last stage
    int stage_next = -1;
    bool acc_fresh = false;
  };
  // MODE_F32R body ordering.  A system's bodies are put in Hilbert order
  // (perm, and the slot-ordered snapshot copy) before its first F32R
  // evaluation and before every step whose count is a multiple of
  // kResortEvery -- a schedule of the system's own step count, so results do
  // not depend on how steps are split across calls or on the other systems
  // stepped with it.  plan_sorts() uploads the tables of every pass one call
  // needs (pass t runs before that call's step t).
  struct SortPass {
    int64_t t;
    int nsys, max_tiles, total, tab;
  };
  static constexpr int64_t kResortEvery = 32;
  std::vector<SortPass> plan_sorts(const std::vector<int>& ids, int64_t nsteps);
  void upload_sort_tab(const std::vector<SortSys>& tab);
  void run_sorts(const std::vector<SortPass>& plan, size_t& next, int64_t t);

  size_t carve(char* base);
  void alloc_arena();
  void validate_items(const std::vector<WorkItem>& items) const;
  void upload_table();
  void ensure_items(size_t n);
  void prepare(const std::vector<int>& ids);
  SysDesc desc_of(int id) const;
  void tiles_of(int id, int stage, int src, std::vector<WorkItem>& out,
                std::vector<WorkItem>* combine = nullptr) const;
  int split_of(int id) const;
  int variant_of(int id) const;
  // Device offsets/counts of one step's lists: for each step parity
  // (which) and force round, one force list per kernel variant, then the
  // combine list (v == kForceVariants).
  struct StepLists {
    static constexpr int kPer = kForceVariants + 1;
    static constexpr int kLists = 2 * 4 * kPer;
    static int at(int which, int round, int v) { return (which * 4 + round) * kPer + v; }
    int off[kLists] = {};
    int cnt[kLists] = {};
  };
  void run_round(std::vector<WorkItem>& items, std::vector<WorkItem>& combine,
                 bool potential = false);
  void sort_items(std::vector<WorkItem>& items) const;
  std::vector<int> by_size(const std::vector<int>& ids) const;
  void step_spans(const std::vector<int>& order, int flip,
                  std::vector<ItemSpan> (&force)[4][kForceVariants],
                  std::vector<ItemSpan> (&combine)[4]) const;
  void prime(const std::vector<int>& ids);
  void upload_step_lists(const std::vector<int>& ids, StepLists& L);
  void launch_step(const StepLists& L, int which);
  void finish_steps(const std::vector<int>& ids, int64_t nsteps);
  // The stepping loop of step() / step_graph(); after_step(t) (may be empty)
  // runs on the host after step t's launches are queued.
  template <typename Hook>
  void run_steps(const std::vector<int>& ids, int64_t nsteps, bool graph,
                 Hook&& after_step);
  void ensure_stage(int which, size_t doubles);
  void check_id(int id) const;
  // One Barnes-Hut force evaluation of a system with its stage epilogue.
  void bh_round(int id, int stage, int src);
  // One step of every listed system with Barnes-Hut forces (parity t & 1).
  void bh_step(const std::vector<int>& ids, int64_t t);

  // launch_force / launch_combine of nitems items, bracketed by the timer
  void force_launch(const WorkItem* items, int nitems, int variant,
                    bool potential = false);
  void combine_launch(const WorkItem* items, int nitems);

  hipStream_t stream_;
  KernelTimer* timer_ = nullptr;
  int mode_;
  int gravity_ = GRAV_DIRECT;
  int split_cap_ = kMaxSplit;
  double theta_ = 0.5;
  std::map<int, std::unique_ptr<BarnesHut>> trees_;
  int64_t cap_;
  int max_sys_;
  Arena A_{};
  void* arena_mem_ = nullptr;
  Arena* d_arena_ = nullptr;  // device copy of A_
  size_t arena_bytes_ = 0;
  SysDesc* d_sys_ = nullptr;
  std::vector<SysDesc> h_sys_;
  WorkItem* d_items_ = nullptr;
  size_t items_cap_ = 0;
  int32_t* d_list_ = nullptr;
  double* d_diag_ = nullptr;
  struct Staging {              // gather staging: device + pinned host
    double* d = nullptr;
    double* h = nullptr;
    size_t cap = 0;             // doubles
  };
  // [0]: step_snapshots(), [1]: gather(), so a gather between the segments
  // of a stepping call leaves that call's snapshots intact
  Staging stage_[2];
  std::vector<WorkItem> gather_items_;  // host side of the last gather list
  hipEvent_t ev_gather_ = nullptr;      // the last gather's copy is done
  bool gather_queued_ = false;
  hipEvent_t ev_boundary_ = nullptr;    // a boundary's copies are done
  WorkItem* d_snap_items_ = nullptr;  // gather lists of step_snapshots()
  size_t snap_items_cap_ = 0;
  double* d_traj_ = nullptr;    // trajectory buffer of record()
  size_t traj_cap_ = 0;         // doubles
  // body ordering (MODE_F32R): keys/values in and out, sort scratch
  SortSys* d_sort_tab_ = nullptr;
  size_t sort_tab_cap_ = 0;
  uint64_t* d_keys_ = nullptr;  // 2 * cap: in, out
  int32_t* d_vals_ = nullptr;   // 2 * cap: in, out
  void* d_sort_tmp_ = nullptr;
  size_t sort_tmp_bytes_ = 0;
  std::vector<Sys> sys_;
  std::map<int64_t, int64_t> free_;  // offset -> length
  bool table_dirty_ = true;

  // step lists: their own device buffer, reused while the signature holds
  WorkItem* d_step_items_ = nullptr;
  size_t step_items_cap_ = 0;
  ItemSpan* d_step_spans_ = nullptr;  // spans k_expand turns into the lists
  size_t step_spans_cap_ = 0;
  std::vector<int64_t> step_sig_;
  StepLists step_lists_;

  // graph cache for step_graph: one step of each snapshot parity
  hipGraphExec_t gexec_[2] = {nullptr, nullptr};
  std::vector<int> gids_;
  int64_t gkey_ = -1;
  void drop_graphs();
};

}  // namespace nbody
