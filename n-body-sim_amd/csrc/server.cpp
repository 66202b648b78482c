// Served form: a request loop with continuous batching.
//
// One worker thread owns the engine.  Requests wait in a FIFO until their
// device footprint (a contiguous arena range plus a table slot) is free;
// admission is first-fit in arrival order, and a request that does not fit
// may be bypassed by later ones (each later request admitted past it counts
// once) at most kMaxBypass times before admission stops behind it (no
// starvation).  All resident systems advance together in
// epochs of snapshot_every steps; a system whose request ends inside an
// epoch stops stepping at its last step while the others go on, reports
// DONE at that point, and is retired at the epoch's end; the epoch is cut at
// such a step when other requests wait for admission.  Snapshots are taken
// at multiples of snapshot_every of each system's own steps, by gather
// launches between the epoch's steps, so every system is snapshotted and
// retired at the same own-step counts whether it runs alone or in a mix (its
// target reordering in MODE_F32R also follows its own step count).  New
// requests join at the next epoch.  Snapshots taken before a request's end
// leave the device with its final state and are queued as events while the
// next segment runs; the rest leave at the epoch's end.  When the queued
// snapshot bytes exceed the configured cap the worker pauses stepping until
// the client drains events (backpressure).
#include "server.h"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "engine.h"

namespace nbody {

namespace {
constexpr int kMaxBypass = 64;

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

struct Request {
  uint64_t id = 0;
  SystemParams p;
  int64_t steps = 0;
  bool want_snapshots = false;
  std::vector<double> m, x, y, z, vx, vy, vz;
  double t_submit = 0;
  int bypassed = 0;
};

struct Active {
  Request req;
  int sys = -1;
  double t_admit = 0;
  bool cancel = false;
  bool done_sent = false;  // DONE emitted inside an epoch; only removal left
};

struct Event {
  nbody_event ev{};
  std::vector<double> x, y, z, vx, vy, vz;
  size_t bytes() const {
    return sizeof(double) * (x.size() + y.size() + z.size() + vx.size() +
                             vy.size() + vz.size());
  }
};

int mode_of(int precision) {
  switch (precision) {
    case NBODY_PRECISION_F64_STRICT:
      return MODE_F64_STRICT;
    case NBODY_PRECISION_F64:
      return MODE_F64;
    case NBODY_PRECISION_F32C:
      return MODE_F32C;
    case NBODY_PRECISION_F32R:
      return MODE_F32R;
  }
  throw Error(NBODY_E_INVALID, "unknown precision");
}

int64_t evals_per_step(int integ) { return integ == INTEG_RK4 ? 4 : 1; }
}  // namespace

std::vector<size_t> admission_pass(const std::vector<int64_t>& sizes,
                                   std::vector<int>& bypassed,
                                   std::map<int64_t, int64_t> free,
                                   int free_slots, int max_bypass) {
  std::vector<size_t> take;
  std::vector<size_t> passed;  // positions that did not fit
  for (size_t i = 0; i < sizes.size() && free_slots > 0; ++i) {
    if (Engine::take_range(free, sizes[i]) < 0) {
      passed.push_back(i);
      continue;
    }
    bool starving = false;
    for (size_t w : passed) starving = starving || bypassed[w] >= max_bypass;
    if (starving) break;
    for (size_t w : passed) ++bypassed[w];
    --free_slots;
    take.push_back(i);
  }
  return take;
}

class Server {
 public:
  explicit Server(const nbody_server_config& cfg) : cfg_(cfg) {
    if (cfg.capacity_bodies <= 0 || cfg.max_systems <= 0 ||
        cfg.snapshot_every <= 0)
      throw Error(NBODY_E_INVALID,
                  "capacity_bodies, max_systems and snapshot_every must be > 0");
    mode_ = mode_of(cfg.precision);
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
      throw Error(NBODY_E_DEVICE, "no GPU available");
    if (cfg.device < 0 || cfg.device >= count)
      throw Error(NBODY_E_INVALID, "device index out of range");
    hip_check(hipSetDevice(cfg.device), "hipSetDevice");
    hip_check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking),
              "hipStreamCreate");
    try {
      eng_ = std::make_unique<Engine>(cfg.device, stream_, mode_,
                                      cfg.capacity_bodies, cfg.max_systems);
    } catch (...) {
      (void)hipStreamDestroy(stream_);
      throw;
    }
    worker_ = std::thread([this] { run(); });
  }

  ~Server() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    cv_work_.notify_all();
    cv_events_.notify_all();
    if (worker_.joinable()) worker_.join();
    (void)hipSetDevice(cfg_.device);
    eng_.reset();
    (void)hipStreamDestroy(stream_);
  }

  uint64_t submit(const nbody_request& r) {
    if (r.n <= 0) throw Error(NBODY_E_INVALID, "request needs n >= 1");
    if (!(r.dt > 0) || !std::isfinite(r.dt))
      throw Error(NBODY_E_INVALID, "dt must be finite and > 0");
    if (r.steps < 0) throw Error(NBODY_E_INVALID, "steps < 0");
    if (r.integrator < NBODY_INTEGRATOR_EULER ||
        r.integrator > NBODY_INTEGRATOR_RUNGE_KUTTA)
      throw Error(NBODY_E_INVALID, "unknown integrator");
    if (!r.masses || !r.pos_x || !r.pos_y || !r.pos_z || !r.vel_x ||
        !r.vel_y || !r.vel_z)
      throw Error(NBODY_E_INVALID, "null body array");
    if ((r.n + 63) / 64 * 64 > eng_->capacity())
      throw Error(NBODY_E_NOSPACE, "request larger than the server arena");
    Request q;
    q.p.n = r.n;
    q.p.integ = r.integrator;  // NBODY_INTEGRATOR_* == Integ values
    q.p.dt = r.dt;
    q.p.g = r.g;
    q.p.eps2 = r.eps2;
    q.steps = r.steps;
    q.want_snapshots = r.want_snapshots != 0;
    auto cp = [&](const double* a, std::vector<double>& v) {
      v.assign(a, a + r.n);
      for (double d : v)
        if (!std::isfinite(d))
          throw Error(NBODY_E_INVALID, "non-finite body value");
    };
    cp(r.masses, q.m);
    cp(r.pos_x, q.x);
    cp(r.pos_y, q.y);
    cp(r.pos_z, q.z);
    cp(r.vel_x, q.vx);
    cp(r.vel_y, q.vy);
    cp(r.vel_z, q.vz);
    if (!std::isfinite(r.g) || !std::isfinite(r.eps2) || r.eps2 < 0)
      throw Error(NBODY_E_INVALID, "g and eps2 must be finite, eps2 >= 0");
    q.t_submit = now_s();
    uint64_t id;
    {
      std::lock_guard<std::mutex> lk(mu_);
      id = q.id = ++next_id_;
      queue_.push_back(std::move(q));
      ++submitted_;
    }
    cv_work_.notify_all();
    return id;
  }

  void poll(int timeout_ms, nbody_event* ev, double* x, double* y, double* z,
            double* vx, double* vy, double* vz, int64_t capacity) {
    std::unique_lock<std::mutex> lk(mu_);
    auto ready = [&] { return !events_.empty() || stop_; };
    if (timeout_ms < 0)
      cv_events_.wait(lk, ready);
    else
      cv_events_.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready);
    if (events_.empty()) throw Error(NBODY_E_TIMEOUT, "no event");
    Event& e = events_.front();
    *ev = e.ev;
    const bool has_data = !e.x.empty();
    if (has_data && (x || y || z || vx || vy || vz) && capacity < e.ev.n)
      throw Error(NBODY_E_NOSPACE, "output arrays smaller than event n");
    auto put = [&](double* dst, const std::vector<double>& v) {
      if (dst && !v.empty()) std::memcpy(dst, v.data(), sizeof(double) * v.size());
    };
    put(x, e.x);
    put(y, e.y);
    put(z, e.z);
    put(vx, e.vx);
    put(vy, e.vy);
    put(vz, e.vz);
    queued_bytes_ -= e.bytes();
    events_.pop_front();
    lk.unlock();
    cv_work_.notify_all();  // may release backpressure
  }

  void cancel(uint64_t id) {
    std::lock_guard<std::mutex> lk(mu_);
    for (auto it = queue_.begin(); it != queue_.end(); ++it)
      if (it->id == id) {
        Event e;
        fill(e.ev, NBODY_EVENT_CANCELLED, *it, 0, 0.0, "cancelled while queued");
        push_event(std::move(e));
        queue_.erase(it);
        return;
      }
    // taken from the queue, its upload still running: cancelled once resident
    auto ad = admitting_.find(id);
    if (ad != admitting_.end()) {
      ad->second = true;
      return;
    }
    auto a = active_.find(id);
    if (a == active_.end()) throw Error(NBODY_E_NOTFOUND, "unknown request id");
    a->second.cancel = true;
  }

  void stats(double* out) {
    std::lock_guard<std::mutex> lk(mu_);
    out[0] = (double)submitted_;
    out[1] = (double)admitted_;
    out[2] = (double)completed_;
    out[3] = (double)active_.size();
    out[4] = (double)queue_.size();
    out[5] = (double)epoch_steps_;
    out[6] = interactions_;
    out[7] = (double)failed_;
  }

 private:
  static void fill(nbody_event& ev, int kind, const Request& q, int64_t step,
                   double t_admit, const char* msg) {
    ev.kind = kind;
    ev.request_id = q.id;
    ev.step = step;
    ev.n = q.p.n;
    ev.t_submit = q.t_submit;
    ev.t_admit = t_admit;
    ev.t_event = now_s();
    std::strncpy(ev.message, msg ? msg : "", sizeof(ev.message) - 1);
    ev.message[sizeof(ev.message) - 1] = '\0';
  }

  void push_event(Event&& e) {  // mu_ held
    queued_bytes_ += e.bytes();
    events_.push_back(std::move(e));
    cv_events_.notify_all();
  }

  // Admits queued requests that fit (worker thread, mu_ held).  Placement
  // is replayed on a copy of the arena's free ranges with add()'s own rule,
  // in admission order, so every admitted request gets a range when it is
  // uploaded (ranges taken by earlier admissions of the same pass are gone).
  // Admitting a request bypasses every earlier one that did not fit; once one
  // of those has been bypassed kMaxBypass times, admission stops behind it.
  // Admitted ids stay in admitting_ until they are resident, so a cancel
  // that arrives during the upload is kept.
  std::vector<Request> take_admissible() {
    std::vector<int64_t> sizes;
    std::vector<int> bypassed;
    for (const Request& q : queue_) {
      sizes.push_back(q.p.n);
      bypassed.push_back(q.bypassed);
    }
    const std::vector<size_t> take =
        admission_pass(sizes, bypassed, eng_->free_ranges(),
                       cfg_.max_systems - eng_->nsystems(), kMaxBypass);
    for (size_t k = 0; k < queue_.size(); ++k) queue_[k].bypassed = bypassed[k];
    std::vector<Request> out;
    for (size_t k : take) {
      admitting_[queue_[k].id] = false;
      out.push_back(std::move(queue_[k]));
    }
    // erase from the back so earlier positions stay valid
    for (size_t q = take.size(); q-- > 0;)
      queue_.erase(queue_.begin() + (std::ptrdiff_t)take[q]);
    return out;
  }

  void run() {
    (void)hipSetDevice(cfg_.device);
    std::vector<int> prev_ids;
    for (;;) {
      std::vector<Request> admit;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_work_.wait(lk, [&] {
          if (stop_) return true;
          const bool backpressured =
              cfg_.max_snapshot_bytes > 0 &&
              queued_bytes_ > (size_t)cfg_.max_snapshot_bytes;
          if (backpressured) return false;
          return !active_.empty() || !queue_.empty();
        });
        if (stop_) return;
        admit = take_admissible();
      }
      // upload admitted requests (the engine is owned by this thread)
      for (auto& q : admit) {
        Active a;
        a.t_admit = now_s();
        try {
          a.sys = eng_->add(q.p, q.m.data(), q.x.data(), q.y.data(),
                            q.z.data(), q.vx.data(), q.vy.data(), q.vz.data());
        } catch (const std::exception& ex) {
          Event e;
          fill(e.ev, NBODY_EVENT_FAILED, q, 0, a.t_admit, ex.what());
          std::lock_guard<std::mutex> lk(mu_);
          admitting_.erase(q.id);
          push_event(std::move(e));
          ++failed_;
          continue;
        }
        for (auto* v : {&q.m, &q.x, &q.y, &q.z, &q.vx, &q.vy, &q.vz})
          std::vector<double>().swap(*v);  // free the host copy
        a.req = std::move(q);
        std::lock_guard<std::mutex> lk(mu_);
        ++admitted_;
        const uint64_t id = a.req.id;
        auto ad = admitting_.find(id);
        if (ad != admitting_.end()) {
          a.cancel = ad->second;  // cancelled during the upload
          admitting_.erase(ad);
        }
        active_.emplace(id, std::move(a));
      }
      // retire cancelled and finished requests; the epoch is snapshot_every
      // steps, and a request that ends inside it stops stepping at its last
      // step (the epoch goes on for the others)
      retire_finished();
      std::vector<int> ids;
      std::map<int, int64_t> left;  // system -> steps in this epoch
      int64_t k = 0;
      {
        std::lock_guard<std::mutex> lk(mu_);
        for (auto& kv : active_) {
          const Active& a = kv.second;
          const int64_t c = std::min<int64_t>(
              cfg_.snapshot_every, a.req.steps - eng_->steps_done(a.sys));
          ids.push_back(a.sys);
          left[a.sys] = c;
          k = std::max(k, c);
        }
      }
      if (ids.empty() || k <= 0) continue;
      std::sort(ids.begin(), ids.end());
      std::vector<int64_t> counts;
      for (int id : ids) counts.push_back(left[id]);
      // snapshots at own-step multiples of snapshot_every, gathered on the
      // device as the epoch runs (the final state goes with DONE instead)
      std::vector<Engine::Snap> snaps;
      std::vector<const Active*> snap_req;
      std::vector<int64_t> snap_step;  // the request's own step count
      {
        std::lock_guard<std::mutex> lk(mu_);
        for (int64_t t = 0; t < k; ++t)
          for (auto& kv : active_) {
            const Active& a = kv.second;
            if (t >= left[a.sys]) continue;
            const int64_t T = eng_->steps_done(a.sys) + t + 1;
            if (a.req.want_snapshots && T < a.req.steps &&
                T % cfg_.snapshot_every == 0) {
              snaps.push_back({t, a.sys});
              snap_req.push_back(&a);
              snap_step.push_back(T);
            }
          }
      }
      std::vector<int64_t> rows;
      int64_t total = 0, ran = k;
      const double* h = nullptr;
      std::vector<char> snap_sent(snaps.size(), 0);
      // queues the snapshot events of steps before t not sent yet, in step
      // order (a request's snapshots always precede its DONE)
      auto emit_snaps = [&](int64_t t, const double* hv) {
        if (!hv) return;
        std::vector<Event> evs;
        for (size_t q = 0; q < snaps.size(); ++q) {
          if (snap_sent[q] || snaps[q].t >= t) continue;
          snap_sent[q] = 1;
          const Active& a = *snap_req[q];
          const int64_t n = a.req.p.n;
          Event e;
          fill(e.ev, NBODY_EVENT_SNAPSHOT, a.req, snap_step[q], a.t_admit, "");
          e.x.assign(hv + rows[q], hv + rows[q] + n);
          e.y.assign(hv + total + rows[q], hv + total + rows[q] + n);
          e.z.assign(hv + 2 * total + rows[q], hv + 2 * total + rows[q] + n);
          evs.push_back(std::move(e));
        }
        std::lock_guard<std::mutex> lk(mu_);
        for (auto& e : evs) push_event(std::move(e));
      };
      // requests that took their last step at a boundary: their final state
      // is gathered there and reported, after the snapshots taken so far,
      // while the next segment runs
      Final fin;
      try {
        // a graph pays off only while the resident set stays the same; the
        // epoch is cut where a request ends if others wait for admission
        h = eng_->step_snapshots(
            ids, k, ids == prev_ids, snaps, rows, total, &counts,
            [&](int64_t) {
              fin = Final{};
              {
                std::lock_guard<std::mutex> lk(mu_);
                for (auto& kv : active_) {
                  const Active& a = kv.second;
                  if (!a.done_sent && !a.cancel &&
                      eng_->steps_done(a.sys) >= a.req.steps)
                    fin.ids.push_back(kv.first);
                }
              }
              gather_final(fin, false);
              std::lock_guard<std::mutex> lk(mu_);
              return !queue_.empty();
            },
            [&](int64_t t0, const double* hv) {
              emit_snaps(t0, hv);
              emit_final(fin);
              fin = Final{};
            },
            &ran);
      } catch (const std::exception& ex) {
        std::lock_guard<std::mutex> lk(mu_);
        for (auto it = active_.begin(); it != active_.end();) {
          Event e;
          fill(e.ev, NBODY_EVENT_FAILED, it->second.req,
               eng_->steps_done(it->second.sys), it->second.t_admit, ex.what());
          try {
            eng_->remove(it->second.sys);
          } catch (...) {
          }
          push_event(std::move(e));
          ++failed_;
          it = active_.erase(it);
        }
        prev_ids.clear();
        continue;
      }
      prev_ids = ids;
      for (auto& kv : left) kv.second = std::min(kv.second, ran);
      emit_snaps(ran, h);  // snapshots after a cut were not taken
      double inter = 0;
      std::lock_guard<std::mutex> lk(mu_);
      epoch_steps_ += ran;
      for (auto& kv : active_) {
        const double n = (double)kv.second.req.p.n;
        const auto c = left.find(kv.second.sys);
        if (c == left.end()) continue;
        inter += n * (n - 1.0) *
                 (double)(c->second * evals_per_step(kv.second.req.p.integ));
      }
      interactions_ += inter;
    }
  }

  // Emits the final events of cancelled and finished requests (final state
  // gathered for all of them in one pass) and frees their arena ranges.
  void retire_finished() {
    std::vector<uint64_t> fin, canc, sent;
    {
      std::lock_guard<std::mutex> lk(mu_);
      for (auto& kv : active_) {
        const Active& a = kv.second;
        if (a.done_sent)
          sent.push_back(kv.first);
        else if (a.cancel)
          canc.push_back(kv.first);
        else if (eng_->steps_done(a.sys) >= a.req.steps)
          fin.push_back(kv.first);
      }
    }
    if (fin.empty() && canc.empty() && sent.empty()) return;
    emit_done(fin);
    std::vector<Event> evs;
    std::lock_guard<std::mutex> lk(mu_);
    for (uint64_t id : canc) {
      Active& a = active_.at(id);
      Event e;
      fill(e.ev, NBODY_EVENT_CANCELLED, a.req, eng_->steps_done(a.sys),
           a.t_admit, "cancelled");
      evs.push_back(std::move(e));
    }
    for (auto* lst : {&fin, &canc, &sent})
      for (uint64_t id : *lst) {
        try {
          eng_->remove(active_.at(id).sys);
        } catch (...) {
        }
        active_.erase(id);
      }
    for (auto& e : evs) push_event(std::move(e));
  }

  // Final states of finished requests, gathered in one pass.
  struct Final {
    std::vector<uint64_t> ids;
    std::vector<int64_t> rows;
    int64_t total = 0;
    const double* h = nullptr;  // valid once the gather has completed
    std::string err;
  };

  // Queues (wait = false) or runs the gather of fin's final states.
  void gather_final(Final& fin, bool wait) {
    if (fin.ids.empty()) return;
    std::vector<int> sys;
    {
      std::lock_guard<std::mutex> lk(mu_);
      for (uint64_t id : fin.ids) sys.push_back(active_.at(id).sys);
    }
    try {
      fin.h = wait ? eng_->gather(sys, 6, fin.rows, fin.total)
                   : eng_->gather_async(sys, 6, fin.rows, fin.total);
    } catch (const std::exception& ex) {
      fin.h = nullptr;
      fin.err = ex.what();
    }
  }

  // DONE (or FAILED) events of the listed finished requests; marks them
  // sent.
  void emit_done(const std::vector<uint64_t>& ids) {
    Final fin;
    fin.ids = ids;
    gather_final(fin, true);
    emit_final(fin);
  }

  void emit_final(const Final& fin) {
    if (fin.ids.empty()) return;
    const double* h = fin.h;
    const std::vector<int64_t>& rows = fin.rows;
    const int64_t total = fin.total;
    std::lock_guard<std::mutex> lk(mu_);
    for (size_t q = 0; q < fin.ids.size(); ++q) {
      Active& a = active_.at(fin.ids[q]);
      Event e;
      const int64_t done = eng_->steps_done(a.sys);
      if (h) {
        fill(e.ev, NBODY_EVENT_DONE, a.req, done, a.t_admit, "");
        const int64_t n = a.req.p.n;
        std::vector<double>* f[6] = {&e.x, &e.y, &e.z, &e.vx, &e.vy, &e.vz};
        for (int c = 0; c < 6; ++c)
          f[c]->assign(h + c * total + rows[q], h + c * total + rows[q] + n);
        ++completed_;
      } else {
        fill(e.ev, NBODY_EVENT_FAILED, a.req, done, a.t_admit, fin.err.c_str());
        ++failed_;
      }
      a.done_sent = true;
      push_event(std::move(e));
    }
  }

  nbody_server_config cfg_;
  int mode_ = MODE_F32C;
  hipStream_t stream_ = nullptr;
  std::unique_ptr<Engine> eng_;
  std::thread worker_;
  std::mutex mu_;
  std::condition_variable cv_work_, cv_events_;
  std::deque<Request> queue_;
  std::map<uint64_t, bool> admitting_;  // taken from the queue, uploading:
                                        // id -> cancel requested
  std::map<uint64_t, Active> active_;
  std::deque<Event> events_;
  size_t queued_bytes_ = 0;
  bool stop_ = false;
  uint64_t next_id_ = 0;
  int64_t submitted_ = 0, admitted_ = 0, completed_ = 0, failed_ = 0;
  int64_t epoch_steps_ = 0;
  double interactions_ = 0;
};

Server* server_create(const nbody_server_config& cfg) {
  return new Server(cfg);
}
uint64_t server_submit(Server* s, const nbody_request& req) {
  return s->submit(req);
}
void server_poll(Server* s, int timeout_ms, nbody_event* ev, double* x,
                 double* y, double* z, double* vx, double* vy, double* vz,
                 int64_t capacity) {
  s->poll(timeout_ms, ev, x, y, z, vx, vy, vz, capacity);
}
void server_cancel(Server* s, uint64_t id) { s->cancel(id); }
void server_stats(Server* s, double* out) { s->stats(out); }
void server_destroy(Server* s) { delete s; }

}  // namespace nbody
