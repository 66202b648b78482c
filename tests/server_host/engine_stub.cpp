// Test-only CPU stand-in for the parts of nbody::Engine the request server
// uses, and for the four HIP runtime calls it makes, so that server.cpp can
// be built with g++ and run under ThreadSanitizer without a GPU.  Nothing
// here is part of the library.
//
// The stub keeps each system's state in host memory and advances it by a
// free drift, x += v * dt per step (no forces): the server's logic does not
// depend on the physics, and the test can predict every result bit for bit.
// step_snapshots() follows the real engine's contract (segments between the
// distinct step counts, boundary and deliver callbacks, snapshot rows in step
// order), and each step sleeps briefly, as a device would be busy.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <map>
#include <thread>
#include <vector>

#include "../../include/nbody.h"
#include "../../n-body-sim_amd/csrc/bh.h"
#include "../../n-body-sim_amd/csrc/engine.h"

extern "C" {
hipError_t hipGetDeviceCount(int* count) {
  *count = 1;
  return hipSuccess;
}
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned int) {
  *s = nullptr;
  return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
const char* hipGetErrorString(hipError_t) { return "stub"; }
}

// how long an upload (add) takes; a test lengthens it to reach its window
int g_upload_us = 100;

namespace nbody {

// the engine's tree table is never filled here; its destructor needs this
BarnesHut::~BarnesHut() {}

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw Error(NBODY_E_DEVICE, what);
}

namespace {
struct HostSys {
  std::vector<double> x, y, z, vx, vy, vz;
};
struct StubState {
  std::map<int, HostSys> sys;
  std::vector<double> snap_stage, gather_stage;
};
std::map<const Engine*, StubState> g_state;  // the worker thread's engines
StubState& state(const Engine* e) { return g_state[e]; }
}  // namespace

Engine::Engine(int, hipStream_t stream, int mode, int64_t capacity,
               int max_systems)
    : stream_(stream),
      mode_(mode),
      cap_((std::max<int64_t>(capacity, 1) + kWave - 1) / kWave * kWave),
      max_sys_(std::max(max_systems, 1)) {
  sys_.assign(max_sys_, Sys{});
  free_[0] = cap_;
  state(this);
}

Engine::~Engine() { g_state.erase(this); }

int Engine::nsystems() const {
  int c = 0;
  for (auto& s : sys_) c += s.live;
  return c;
}

int64_t Engine::steps_done(int id) const {
  if (id < 0 || id >= max_sys_ || !sys_[id].live)
    throw Error(NBODY_E_NOTFOUND, "unknown system id");
  return sys_[id].steps;
}

int Engine::add(const SystemParams& p, const double* m, const double* x,
                const double* y, const double* z, const double* vx,
                const double* vy, const double* vz) {
  (void)m;
  int id = -1;
  for (int k = 0; k < max_sys_; ++k)
    if (!sys_[k].live) {
      id = k;
      break;
    }
  if (id < 0) throw Error(NBODY_E_NOSPACE, "system table full");
  const int64_t off = take_range(free_, p.n);
  if (off < 0) throw Error(NBODY_E_NOSPACE, "arena full");
  Sys& s = sys_[id];
  s = Sys{};
  s.live = true;
  s.p = p;
  s.off = off;
  s.span = (p.n + kWave - 1) / kWave * kWave;
  HostSys& h = state(this).sys[id];
  h.x.assign(x, x + p.n);
  h.y.assign(y, y + p.n);
  h.z.assign(z, z + p.n);
  h.vx.assign(vx, vx + p.n);
  h.vy.assign(vy, vy + p.n);
  h.vz.assign(vz, vz + p.n);
  std::this_thread::sleep_for(std::chrono::microseconds(g_upload_us));  // upload
  return id;
}

void Engine::remove(int id) {
  if (id < 0 || id >= max_sys_ || !sys_[id].live)
    throw Error(NBODY_E_NOTFOUND, "unknown system id");
  Sys& s = sys_[id];
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
  state(this).sys.erase(id);
}

const double* Engine::step_snapshots(
    const std::vector<int>& ids, int64_t nsteps, bool graph,
    const std::vector<Snap>& snaps, std::vector<int64_t>& rows, int64_t& total,
    const std::vector<int64_t>* counts,
    const std::function<bool(int64_t)>& boundary,
    const std::function<void(int64_t, const double*)>& deliver, int64_t* ran) {
  (void)graph;
  StubState& st = state(this);
  rows.assign(snaps.size(), 0);
  total = 0;
  if (ran) *ran = nsteps;
  if (counts && counts->size() != ids.size())
    throw Error(NBODY_E_INVALID, "one step count per system");
  auto count_of = [&](int id) {
    if (!counts) return nsteps;
    for (size_t q = 0; q < ids.size(); ++q)
      if (ids[q] == id) return (*counts)[q];
    return (int64_t)-1;
  };
  // rows in step order, as the engine lays out its staging
  std::vector<size_t> order(snaps.size());
  for (size_t k = 0; k < order.size(); ++k) order[k] = k;
  std::stable_sort(order.begin(), order.end(),
                   [&](size_t a, size_t b) { return snaps[a].t < snaps[b].t; });
  for (size_t k : order) {
    const Snap& sn = snaps[k];
    if (sn.t < 0 || sn.t >= nsteps || sn.t >= count_of(sn.id))
      throw Error(NBODY_E_INVALID, "snapshot step outside the call");
    rows[k] = total;
    total += sys_[sn.id].p.n;
  }
  st.snap_stage.assign((size_t)(3 * total), 0.0);
  std::vector<int64_t> ends;
  if (counts)
    ends.assign(counts->begin(), counts->end());
  else
    ends.push_back(nsteps);
  std::sort(ends.begin(), ends.end());
  ends.erase(std::unique(ends.begin(), ends.end()), ends.end());
  int64_t t0 = 0, stop = nsteps;
  for (int64_t e : ends) {
    if (e <= t0) continue;
    for (int64_t t = t0; t < e; ++t) {
      for (size_t q = 0; q < ids.size(); ++q) {
        if (count_of(ids[q]) < e) continue;
        Sys& s = sys_[ids[q]];
        HostSys& h = st.sys.at(ids[q]);
        const double dt = s.p.dt;
        for (int64_t i = 0; i < s.p.n; ++i) {
          h.x[i] = h.x[i] + h.vx[i] * dt;
          h.y[i] = h.y[i] + h.vy[i] * dt;
          h.z[i] = h.z[i] + h.vz[i] * dt;
        }
        ++s.steps;
      }
      for (size_t k = 0; k < snaps.size(); ++k) {
        if (snaps[k].t != t) continue;
        const HostSys& h = st.sys.at(snaps[k].id);
        for (size_t i = 0; i < h.x.size(); ++i) {
          st.snap_stage[rows[k] + i] = h.x[i];
          st.snap_stage[total + rows[k] + i] = h.y[i];
          st.snap_stage[2 * total + rows[k] + i] = h.z[i];
        }
      }
      std::this_thread::sleep_for(std::chrono::microseconds(30));  // a step
    }
    t0 = e;
    if (t0 < nsteps && boundary) {
      const bool cut = boundary(t0);
      if (deliver) deliver(t0, total ? st.snap_stage.data() : nullptr);
      if (cut) {
        stop = t0;
        break;
      }
    }
  }
  if (ran) *ran = stop;
  return total ? st.snap_stage.data() : nullptr;
}

const double* Engine::gather(const std::vector<int>& ids, int nfields,
                             std::vector<int64_t>& rows, int64_t& total) {
  return gather_async(ids, nfields, rows, total);
}

const double* Engine::gather_async(const std::vector<int>& ids, int nfields,
                                   std::vector<int64_t>& rows,
                                   int64_t& total) {
  StubState& st = state(this);
  rows.clear();
  total = 0;
  for (int id : ids) {
    rows.push_back(total);
    total += sys_.at(id).p.n;
  }
  st.gather_stage.assign((size_t)(nfields * total), 0.0);
  for (size_t k = 0; k < ids.size(); ++k) {
    const HostSys& h = st.sys.at(ids[k]);
    const std::vector<double>* f[6] = {&h.x, &h.y, &h.z, &h.vx, &h.vy, &h.vz};
    for (int c = 0; c < nfields; ++c)
      std::copy(f[c]->begin(), f[c]->end(), st.gather_stage.begin() + c * total + rows[k]);
  }
  return total ? st.gather_stage.data() : nullptr;
}

}  // namespace nbody
