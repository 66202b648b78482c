// Host tests of the request server (n-body-sim_amd/csrc/server.cpp) on the
// CPU stub engine (engine_stub.cpp), built with -fsanitize=thread by
// tests/test_server_host.py.
//   admission  the admission pass on built inputs: placement replayed on a
//              copy of the free ranges, bypasses counted only when a later
//              request is admitted, the starvation limit, the slot limit
//   stress     threads submitting, polling and cancelling at once, with
//              snapshot backpressure; every request's events are checked
//              (order, snapshot steps, final state bit for bit, cancel
//              semantics)
// Exit status 0 when every check passes; each failure prints a line.
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "../../include/nbody.h"
#include "../../n-body-sim_amd/csrc/engine.h"
#include "../../n-body-sim_amd/csrc/server.h"

using nbody::Error;

extern int g_upload_us;  // engine_stub.cpp

namespace {
int g_fail = 0;

void check(bool ok, const char* what, long a = 0, long b = 0) {
  if (!ok) {
    std::printf("FAIL %s (%ld, %ld)\n", what, a, b);
    ++g_fail;
  }
}

double now_s() {  // the server's clock
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// ---- admission ------------------------------------------------------------
void test_admission() {
  using Free = std::map<int64_t, int64_t>;
  const int kMax = 64;
  {  // two holes of 128: three requests of 100 each fit one at a time, but
     // only two fit together; the replay keeps the third queued
    std::vector<int> byp(3, 0);
    auto t = nbody::admission_pass({100, 100, 100}, byp, Free{{0, 128}, {512, 128}},
                                   16, kMax);
    check(t == std::vector<size_t>({0, 1}), "replay takes two of three");
    check(byp[2] == 0, "nothing admitted behind the third", byp[2]);
  }
  {  // bypasses count only when a later request is admitted past
    std::vector<int> byp(4, 0);
    auto t = nbody::admission_pass({1000, 64, 64, 2000}, byp, Free{{0, 128}}, 16, kMax);
    check(t == std::vector<size_t>({1, 2}), "small ones admitted");
    check(byp[0] == 2, "bypassed by each admission", byp[0]);
    check(byp[3] == 0, "nothing admitted behind the last", byp[3]);
  }
  {  // a pass that admits nothing bypasses nothing (snapshot_every = 1 runs
     // one pass per step)
    std::vector<int> byp(1, 0);
    for (int k = 0; k < 200; ++k)
      check(nbody::admission_pass({1000}, byp, Free{{0, 64}}, 16, kMax).empty(),
            "too large stays queued");
    check(byp[0] == 0, "no bypass without an admission", byp[0]);
    // then a small request behind it is admitted every pass until the limit
    std::vector<int> b2(2, 0);
    int admitted = 0;
    for (int k = 0; k < 100; ++k)
      admitted += (int)nbody::admission_pass({1000, 64}, b2, Free{{0, 64}}, 16, kMax).size();
    check(admitted == kMax, "admissions past a waiting request stop at the limit", admitted);
    check(b2[0] == kMax, "waiting request bypassed exactly kMax times", b2[0]);
  }
  {  // at the limit nothing behind it is taken, even what would fit
    std::vector<int> byp = {kMax, 0, 0};
    auto t = nbody::admission_pass({1000, 64, 64}, byp, Free{{0, 4096}}, 16, kMax);
    check(t == std::vector<size_t>({0, 1, 2}), "the waiting request first when it fits");
    std::vector<int> b3 = {kMax, 0};
    check(nbody::admission_pass({5000, 64}, b3, Free{{0, 4096}}, 16, kMax).empty(),
          "blocked behind a starving request");
  }
  {  // the table-slot limit
    std::vector<int> byp(3, 0);
    auto t = nbody::admission_pass({64, 64, 64}, byp, Free{{0, 1 << 20}}, 2, kMax);
    check(t == std::vector<size_t>({0, 1}), "at most free_slots admitted");
  }
}

// ---- stress -----------------------------------------------------------------
struct Req {
  int64_t n = 0, steps = 0;
  double dt = 0;
  bool snaps = false;
  std::vector<double> x, y, z, vx, vy, vz;
};

struct Seen {
  std::vector<int64_t> snap_steps;
  int terminal = 0;  // event kind
  int terminals = 0;
  int64_t step = -1;
  double t_event = 0;
  bool state_ok = false;
  bool after_terminal = false;  // an event arrived after the terminal one
};

void test_stress(uint64_t seed) {
  nbody_server_config cfg{};
  cfg.device = 0;
  cfg.precision = NBODY_PRECISION_F64;
  cfg.capacity_bodies = 4096;
  cfg.max_systems = 12;
  cfg.snapshot_every = 10;
  cfg.max_snapshot_bytes = 96 << 10;  // backpressure
  nbody::Server* srv = nbody::server_create(cfg);
  const int kThreads = 3, kPer = 40, kTotal = kThreads * kPer;
  std::mutex mu;
  std::map<uint64_t, Req> reqs;
  std::map<uint64_t, Seen> seen;
  std::vector<uint64_t> ids;
  struct CancelRec {
    uint64_t id;
    int code;
    double t;
  };
  std::vector<CancelRec> cancels;
  std::atomic<int> terminal_count{0};
  std::atomic<bool> submit_done{false};

  auto submitter = [&](int tno) {
    std::mt19937_64 rng(seed * 131 + tno);
    for (int k = 0; k < kPer; ++k) {
      Req r;
      r.n = 16 + (int64_t)(rng() % 600);
      r.steps = 5 + (int64_t)(rng() % 80);
      r.dt = (rng() % 2) ? 1e-3 : 5e-4;
      r.snaps = rng() % 3 != 0;
      std::vector<double> m(r.n, 1.0);
      auto fill = [&](std::vector<double>& v) {
        v.resize(r.n);
        for (auto& e : v) e = (double)(int64_t)(rng() % 2001 - 1000) / 997.0;
      };
      fill(r.x); fill(r.y); fill(r.z); fill(r.vx); fill(r.vy); fill(r.vz);
      nbody_request q{};
      q.n = r.n;
      q.integrator = (int)(rng() % 3);
      q.gravity_method = NBODY_GRAVITY_NEWTON_PARALLEL;
      q.g = 1.0;
      q.dt = r.dt;
      q.eps2 = 1e-4;
      q.steps = r.steps;
      q.want_snapshots = r.snaps;
      q.masses = m.data();
      q.pos_x = r.x.data(); q.pos_y = r.y.data(); q.pos_z = r.z.data();
      q.vel_x = r.vx.data(); q.vel_y = r.vy.data(); q.vel_z = r.vz.data();
      std::lock_guard<std::mutex> lk(mu);  // the id must be known before its events
      const uint64_t id = nbody::server_submit(srv, q);
      reqs[id] = std::move(r);
      ids.push_back(id);
    }
  };

  auto poller = [&]() {
    std::vector<double> b[6];
    for (auto& v : b) v.resize(1024);
    while (terminal_count.load() < kTotal) {
      nbody_event ev{};
      try {
        nbody::server_poll(srv, 50, &ev, b[0].data(), b[1].data(), b[2].data(),
                           b[3].data(), b[4].data(), b[5].data(), 1024);
      } catch (const Error& e) {
        check(e.code == NBODY_E_TIMEOUT, "poll error", e.code);
        continue;
      }
      std::lock_guard<std::mutex> lk(mu);
      auto it = reqs.find(ev.request_id);
      check(it != reqs.end(), "event for an unknown id", (long)ev.request_id);
      if (it == reqs.end()) continue;
      const Req& r = it->second;
      Seen& s = seen[ev.request_id];
      if (s.terminals) s.after_terminal = true;
      check(ev.n == r.n, "event n", (long)ev.n, (long)r.n);
      if (ev.kind == NBODY_EVENT_SNAPSHOT) {
        s.snap_steps.push_back(ev.step);
        continue;
      }
      s.terminal = ev.kind;
      ++s.terminals;
      s.step = ev.step;
      s.t_event = ev.t_event;
      if (ev.kind == NBODY_EVENT_DONE) {
        // the stub drifts x += v * dt per step: predict the final state
        bool ok = ev.step == r.steps;
        for (int64_t i = 0; i < r.n && ok; ++i) {
          double x = r.x[i], y = r.y[i], z = r.z[i];
          for (int64_t t = 0; t < r.steps; ++t) {
            x = x + r.vx[i] * r.dt;
            y = y + r.vy[i] * r.dt;
            z = z + r.vz[i] * r.dt;
          }
          ok = x == b[0][i] && y == b[1][i] && z == b[2][i] && r.vx[i] == b[3][i] &&
               r.vy[i] == b[4][i] && r.vz[i] == b[5][i];
        }
        s.state_ok = ok;
      }
      ++terminal_count;
    }
  };

  auto canceller = [&]() {
    std::mt19937_64 rng(seed * 977 + 5);
    while (!submit_done.load() || terminal_count.load() < kTotal) {
      uint64_t id = 0;
      {
        std::lock_guard<std::mutex> lk(mu);
        if (!ids.empty() && rng() % 3 == 0)
          id = ids[ids.size() - 1 - rng() % std::min<size_t>(ids.size(), 8)];
      }
      if (id) {
        int code = NBODY_OK;
        try {
          nbody::server_cancel(srv, id);
        } catch (const Error& e) {
          code = e.code;
        }
        std::lock_guard<std::mutex> lk(mu);
        cancels.push_back({id, code, now_s()});
      }
      std::this_thread::sleep_for(std::chrono::microseconds(200 + rng() % 800));
      if (submit_done.load() && rng() % 50 == 0) break;
    }
  };

  // one poller: events are checked in the order the server queued them
  std::vector<std::thread> th;
  th.emplace_back(poller);
  std::thread canc(canceller);
  std::vector<std::thread> sub;
  for (int k = 0; k < kThreads; ++k) sub.emplace_back(submitter, k);
  for (auto& t : sub) t.join();
  submit_done = true;
  canc.join();
  for (auto& t : th) t.join();
  double st[8];
  nbody::server_stats(srv, st);
  nbody::server_destroy(srv);

  check((int)seen.size() == kTotal, "every request reported", (long)seen.size());
  int done = 0, cancelled = 0;
  for (auto& kv : seen) {
    const Req& r = reqs.at(kv.first);
    const Seen& s = kv.second;
    check(s.terminals == 1, "one terminal event", (long)kv.first, s.terminals);
    check(!s.after_terminal, "no event after the terminal one", (long)kv.first);
    check(s.terminal == NBODY_EVENT_DONE || s.terminal == NBODY_EVENT_CANCELLED,
          "DONE or CANCELLED", (long)kv.first, s.terminal);
    int64_t last = 0;
    for (int64_t t : s.snap_steps) {
      check(t > last && t % 10 == 0 && t < r.steps, "snapshot steps", (long)t, (long)last);
      last = t;
    }
    if (s.terminal == NBODY_EVENT_DONE) {
      ++done;
      check(s.state_ok, "final state", (long)kv.first);
      const int64_t want = r.snaps ? (r.steps - 1) / 10 : 0;
      check((int64_t)s.snap_steps.size() == want, "every snapshot before DONE",
            (long)s.snap_steps.size(), (long)want);
    } else {
      ++cancelled;
    }
  }
  for (const CancelRec& c : cancels) {
    const Seen& s = seen[c.id];
    check(c.code == NBODY_OK || c.code == NBODY_E_NOTFOUND, "cancel code", c.code);
    if (c.code == NBODY_E_NOTFOUND)  // only once the request was retired
      check(s.terminals == 1 && s.t_event <= c.t, "NOTFOUND only after retirement",
            (long)c.id, s.terminal);
  }
  check(st[0] == kTotal, "submitted counter", (long)st[0]);
  check(st[2] == done, "completed counter", (long)st[2], done);
  check(st[3] == 0 && st[4] == 0, "nothing left resident or queued", (long)st[3], (long)st[4]);
  std::printf("stress seed %llu: %d done, %d cancelled, %zu cancel calls\n",
              (unsigned long long)seed, done, cancelled, cancels.size());
}
// A cancel that arrives while the request is being uploaded (taken from the
// queue, not yet resident) must be kept: OK, then CANCELLED once resident.
void test_cancel_while_admitting() {
  nbody_server_config cfg{};
  cfg.precision = NBODY_PRECISION_F64;
  cfg.capacity_bodies = 1024;
  cfg.max_systems = 4;
  cfg.snapshot_every = 10;
  nbody::Server* srv = nbody::server_create(cfg);
  g_upload_us = 3000;
  std::vector<double> m(64, 1.0), v(64, 0.5);
  int hit = 0;
  for (int attempt = 0; attempt < 10 && hit < 3; ++attempt) {
    nbody_request q{};
    q.n = 64;
    q.integrator = NBODY_INTEGRATOR_EULER;
    q.gravity_method = NBODY_GRAVITY_NEWTON_PARALLEL;
    q.g = 1.0;
    q.dt = 1e-3;
    q.steps = 1 << 30;  // never finishes by itself
    q.masses = m.data();
    q.pos_x = q.pos_y = q.pos_z = v.data();
    q.vel_x = q.vel_y = q.vel_z = v.data();
    const uint64_t id = nbody::server_submit(srv, q);
    std::this_thread::sleep_for(std::chrono::microseconds(800));  // mid-upload
    int code = NBODY_OK;
    try {
      nbody::server_cancel(srv, id);
    } catch (const Error& e) {
      code = e.code;
    }
    check(code == NBODY_OK, "cancel during the upload is kept", attempt, code);
    nbody_event ev{};
    bool got = false;
    for (int k = 0; k < 200 && !got; ++k) {
      try {
        nbody::server_poll(srv, 50, &ev, nullptr, nullptr, nullptr, nullptr,
                           nullptr, nullptr, 0);
        got = ev.request_id == id;
      } catch (const Error& e) {
        check(e.code == NBODY_E_TIMEOUT, "poll error", e.code);
      }
    }
    check(got && ev.kind == NBODY_EVENT_CANCELLED, "cancelled request retired",
          attempt, ev.kind);
    hit += got && ev.t_admit > 0.0;  // it had been admitted: the window was hit
  }
  check(hit > 0, "some cancel reached a request being uploaded", hit);
  g_upload_us = 100;
  nbody::server_destroy(srv);
  std::printf("cancel while admitting: %d of the attempts hit the upload\n", hit);
}
}  // namespace

int main(int argc, char** argv) {
  const std::string what = argc > 1 ? argv[1] : "all";
  if (what == "admission" || what == "all") test_admission();
  if (what == "cancel" || what == "all") test_cancel_while_admitting();
  if (what == "stress" || what == "all")
    for (uint64_t seed = 1; seed <= 6; ++seed) test_stress(seed);
  std::printf("%s\n", g_fail ? "FAILED" : "ok");
  return g_fail ? 1 : 0;
}
