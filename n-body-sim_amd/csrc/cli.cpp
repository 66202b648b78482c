// n-body-sim: the simulator's command-line front end over the C ABI.
//
// Flags, defaults and method strings are those of the reference binary
// (src/args.rs:19-64, src/constants.rs:5-22), so the reference GUI's runner
// (gui/runners.py:94-114) can start this program unchanged.  The run follows
// NBodySim::run (src/n_body_sim.rs:40-59): a simulation loop that records
// the positions at t = 0 and after every step (src/simulation_runner.rs:
// 28-52), handing batches of snapshots to a writer thread that writes .csv
// or .nbody by the output extension (src/output.rs:73-139).
//
// Every force evaluation runs on the GPU.  Without --gpu the reference runs
// its f64 CPU code; here that selection runs the GPU in the F64_STRICT
// arithmetic, whose results are bit-identical to that CPU code for newton,
// newton-parallel and barnes-hut, so the output files are identical byte for
// byte.  With --gpu the reference's GPU kernel (rsqrt, contracted FMAs) is
// replaced by the fast F64 arithmetic, for all three integrators and for
// barnes-hut too (the reference offers newton-parallel with Euler only
// there).  --precision overrides either choice.
#include <algorithm>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/nbody.h"

namespace {

struct Args {
  std::string initial_conditions_path = "initial_conditions.csv";
  std::string output_data_path = "output.csv";
  double g_constant = 1.0;
  double time_step = 0.01;
  unsigned long long num_steps = 10000;
  double softening_factor = 0.005;
  double theta = 0.5;
  std::string gravity = "newton";
  std::string integrator = "euler";
  bool progress = false;
  bool gpu = false;
  std::string precision;  // empty: f64-strict, or f64 with --gpu
  int device = 0;
};

const char* kUsage = "Usage: n-body-sim [OPTIONS]";

void print_help() {
  std::printf(
      "%s\n\n"
      "Options:\n"
      "  -i, --initial-conditions-path <INITIAL_CONDITIONS_PATH>\n"
      "          Path to CSV file containing initial conditions for each body "
      "[default: initial_conditions.csv]\n"
      "  -o, --output-data-path <OUTPUT_DATA_PATH>\n"
      "          Path to csv or nbody file to save simulation output data "
      "[default: output.csv]\n"
      "  -g, --g-constant <G_CONSTANT>\n"
      "          The gravitational constant to use in force calculations "
      "[default: 1]\n"
      "  -t, --time-step <TIME_STEP>\n"
      "          Simulation time step in seconds [default: 0.01]\n"
      "  -n, --num-steps <NUM_STEPS>\n"
      "          Number of time steps to simulate [default: 10000]\n"
      "      --softening-factor <SOFTENING_FACTOR>\n"
      "          The softening factor to avoid numerical instability as "
      "distances approach zero [default: 0.005]\n"
      "      --theta <THETA>\n"
      "          Theta value for Barnes-Hut gravity calculation method "
      "[default: 0.5]\n"
      "      --gravity <GRAVITY>\n"
      "          Gravity force calculation method [default: newton] "
      "[possible values: newton, newton-parallel, barnes-hut]\n"
      "      --integrator <INTEGRATOR>\n"
      "          Integrator for computing next-step state [default: euler] "
      "[possible values: euler, velocity-verlet, runge-kutta]\n"
      "      --progress\n"
      "          Flag to enable percent progress output to stdout\n"
      "      --gpu\n"
      "          Flag to select the fast GPU arithmetic (all integrators, "
      "newton-parallel and barnes-hut)\n"
      "      --precision <PRECISION>\n"
      "          Force arithmetic [default: f64-strict, f64 with --gpu] "
      "[possible values: f64-strict, f64, f32c, f32r]\n"
      "      --device <DEVICE>\n"
      "          GPU index [default: 0]\n"
      "  -h, --help\n"
      "          Print help\n",
      kUsage);
}

[[noreturn]] void usage_error(const std::string& msg) {
  std::fprintf(stderr, "error: %s\n\n%s\n\nFor more information, try '--help'.\n",
               msg.c_str(), kUsage);
  std::exit(2);
}

[[noreturn]] void fail(const std::string& msg, int code = 1) {
  std::fprintf(stderr, "Error: %s\n", msg.c_str());
  std::exit(code);
}

std::string last_error() { return nbody_last_error(); }

// Numeric flags use Rust's f64 grammar, as clap's parse of an f64 field.
bool parse_double(const std::string& s, double* out) {
  return nbody_parse_f64(s.c_str(), out) == NBODY_OK;
}

bool parse_usize(const std::string& s, unsigned long long* out) {
  size_t i = s.size() > 0 && s[0] == '+' ? 1 : 0;
  if (i == s.size()) return false;
  unsigned long long v = 0;
  for (; i < s.size(); ++i) {
    if (s[i] < '0' || s[i] > '9') return false;
    const unsigned long long d = (unsigned long long)(s[i] - '0');
    if (v > (~0ULL - d) / 10) return false;
    v = v * 10 + d;
  }
  *out = v;
  return true;
}

struct Opt {
  char short_name;
  const char* long_name;
  const char* value_name;  // null: flag
};

const Opt kOpts[] = {
    {'i', "initial-conditions-path", "INITIAL_CONDITIONS_PATH"},
    {'o', "output-data-path", "OUTPUT_DATA_PATH"},
    {'g', "g-constant", "G_CONSTANT"},
    {'t', "time-step", "TIME_STEP"},
    {'n', "num-steps", "NUM_STEPS"},
    {0, "softening-factor", "SOFTENING_FACTOR"},
    {0, "theta", "THETA"},
    {0, "gravity", "GRAVITY"},
    {0, "integrator", "INTEGRATOR"},
    {0, "progress", nullptr},
    {0, "gpu", nullptr},
    {0, "precision", "PRECISION"},
    {0, "device", "DEVICE"},
    {'h', "help", nullptr},
};

std::string shown(const Opt& o) {
  std::string s = std::string("--") + o.long_name;
  if (o.value_name) s += std::string(" <") + o.value_name + ">";
  return s;
}

void apply(Args& a, const Opt& o, const std::string& v) {
  const std::string name = o.long_name;
  auto bad = [&](const char* extra) {
    usage_error("invalid value '" + v + "' for '" + shown(o) + "'" + extra);
  };
  auto one_of = [&](std::initializer_list<const char*> vals) {
    for (const char* x : vals)
      if (v == x) return;
    std::string list;
    for (const char* x : vals) list += std::string(list.empty() ? "" : ", ") + x;
    bad(("\n  [possible values: " + list + "]").c_str());
  };
  if (name == "initial-conditions-path") {
    a.initial_conditions_path = v;
  } else if (name == "output-data-path") {
    a.output_data_path = v;
  } else if (name == "g-constant") {
    if (!parse_double(v, &a.g_constant)) bad(": invalid float literal");
  } else if (name == "time-step") {
    if (!parse_double(v, &a.time_step)) bad(": invalid float literal");
  } else if (name == "num-steps") {
    if (!parse_usize(v, &a.num_steps)) bad(": invalid digit found in string");
  } else if (name == "softening-factor") {
    if (!parse_double(v, &a.softening_factor)) bad(": invalid float literal");
  } else if (name == "theta") {
    if (!parse_double(v, &a.theta)) bad(": invalid float literal");
  } else if (name == "gravity") {
    one_of({"newton", "newton-parallel", "barnes-hut"});
    a.gravity = v;
  } else if (name == "integrator") {
    one_of({"euler", "velocity-verlet", "runge-kutta"});
    a.integrator = v;
  } else if (name == "precision") {
    one_of({"f64-strict", "f64", "f32c", "f32r"});
    a.precision = v;
  } else if (name == "device") {
    unsigned long long d = 0;
    if (!parse_usize(v, &d) || d > 1024) bad(": invalid device index");
    a.device = (int)d;
  }
}

Args parse_args(int argc, char** argv) {
  Args a;
  std::vector<bool> seen(sizeof(kOpts) / sizeof(kOpts[0]), false);
  for (int k = 1; k < argc; ++k) {
    std::string arg = argv[k];
    const Opt* o = nullptr;
    std::string value;
    bool has_value = false;
    if (arg.size() > 2 && arg[0] == '-' && arg[1] == '-') {
      std::string name = arg.substr(2);
      const size_t eq = name.find('=');
      if (eq != std::string::npos) {
        value = name.substr(eq + 1);
        name = name.substr(0, eq);
        has_value = true;
      }
      for (const Opt& c : kOpts)
        if (name == c.long_name) o = &c;
      if (!o) usage_error("unexpected argument '" + arg + "' found");
    } else if (arg.size() >= 2 && arg[0] == '-' && arg[1] != '-') {
      for (const Opt& c : kOpts)
        if (c.short_name && arg[1] == c.short_name) o = &c;
      if (!o) usage_error("unexpected argument '" + arg + "' found");
      if (arg.size() > 2) {
        value = arg.substr(arg[2] == '=' ? 3 : 2);
        has_value = true;
      }
    } else {
      usage_error("unexpected argument '" + arg + "' found");
    }
    const size_t idx = (size_t)(o - kOpts);
    if (std::string(o->long_name) == "help") {
      print_help();
      std::exit(0);
    }
    if (seen[idx])
      usage_error("the argument '" + shown(*o) +
                  "' cannot be used multiple times");
    seen[idx] = true;
    if (!o->value_name) {
      if (has_value)
        usage_error("unexpected value '" + value + "' for '--" +
                    o->long_name + "' found; no more were expected");
      if (std::string(o->long_name) == "progress") a.progress = true;
      if (std::string(o->long_name) == "gpu") a.gpu = true;
      continue;
    }
    if (!has_value) {
      if (k + 1 >= argc)
        usage_error("a value is required for '" + shown(*o) +
                    "' but none was supplied");
      value = argv[++k];
    }
    apply(a, *o, value);
  }
  return a;
}

// One batch of snapshots for the writer: count snapshots of n bodies,
// positions snapshot-major (snapshot s of x at x[s*n .. s*n+n)).
struct Batch {
  std::vector<double> times, x, y, z;
  int64_t n = 0;
};

// Bounded hand-over between the simulation loop and the writer thread
// (the reference's mpsc channel, src/n_body_sim.rs:29-31); a bound of two
// batches keeps memory flat while the GPU steps ahead of the writer.
class Channel {
 public:
  explicit Channel(size_t depth) : depth_(depth) {}
  bool send(std::unique_ptr<Batch> b) {
    std::unique_lock<std::mutex> lk(mu_);
    cv_.wait(lk, [&] { return q_.size() < depth_ || closed_; });
    if (closed_) return false;
    q_.push_back(std::move(b));
    cv_.notify_all();
    return true;
  }
  std::unique_ptr<Batch> recv() {
    std::unique_lock<std::mutex> lk(mu_);
    cv_.wait(lk, [&] { return !q_.empty() || closed_; });
    if (q_.empty()) return nullptr;
    auto b = std::move(q_.front());
    q_.pop_front();
    cv_.notify_all();
    return b;
  }
  void close() {
    std::lock_guard<std::mutex> lk(mu_);
    closed_ = true;
    cv_.notify_all();
  }

 private:
  size_t depth_;
  std::mutex mu_;
  std::condition_variable cv_;
  std::deque<std::unique_ptr<Batch>> q_;
  bool closed_ = false;
};

int integrator_id(const std::string& s) {
  if (s == "euler") return NBODY_INTEGRATOR_EULER;
  if (s == "velocity-verlet") return NBODY_INTEGRATOR_VELOCITY_VERLET;
  return NBODY_INTEGRATOR_RUNGE_KUTTA;
}

int precision_id(const std::string& s) {
  if (s == "f64-strict") return NBODY_PRECISION_F64_STRICT;
  if (s == "f64") return NBODY_PRECISION_F64;
  if (s == "f32c") return NBODY_PRECISION_F32C;
  return NBODY_PRECISION_F32R;
}

std::string method_name(const std::string& s) {  // Debug form of the enum
  if (s == "newton") return "Newton";
  if (s == "newton-parallel") return "NewtonParallel";
  return "BarnesHut";
}

}  // namespace

int main(int argc, char** argv) {
  const Args a = parse_args(argc, argv);

  // Method selection first: src/args.rs:111-123 panics for newton on the
  // GPU (exit status 101 from the panic).  newton and newton-parallel are the
  // same direct sum (bit for bit); barnes-hut is the octree method.
  if (a.gpu && a.gravity == "newton") {
    std::fprintf(stderr,
                 "GPU-accelerated version of %s gravity method is not "
                 "implemented yet\n",
                 method_name(a.gravity).c_str());
    return 101;
  }
  const int gravity = a.gravity == "barnes-hut" ? NBODY_GRAVITY_BARNES_HUT
                                                 : NBODY_GRAVITY_NEWTON_PARALLEL;

  // Initial conditions (src/input.rs:20-33).
  int64_t n = 0;
  const char* icp = a.initial_conditions_path.c_str();
  if (nbody_load_initial_conditions(icp, &n, nullptr, nullptr, nullptr,
                                    nullptr, nullptr, nullptr, nullptr) != 0)
    fail(last_error());
  std::vector<double> ic[7];
  for (auto& v : ic) v.resize((size_t)n);
  if (nbody_load_initial_conditions(icp, &n, ic[0].data(), ic[1].data(),
                                    ic[2].data(), ic[3].data(), ic[4].data(),
                                    ic[5].data(), ic[6].data()) != 0)
    fail(last_error());

  // The writer creates the file and rejects unknown extensions before any
  // step runs.
  nbody_writer* writer = nullptr;
  if (nbody_writer_open(a.output_data_path.c_str(), &writer) != 0)
    fail(last_error());

  Channel chan(2);
  std::string write_error;
  std::thread writer_thread([&] {
    while (auto b = chan.recv()) {
      for (size_t s = 0; s < b->times.size(); ++s) {
        const size_t o = s * (size_t)b->n;
        if (nbody_writer_append(writer, b->times[s], b->n, b->x.data() + o,
                                b->y.data() + o, b->z.data() + o) != 0) {
          write_error = last_error();
          chan.close();
          return;
        }
      }
    }
  });

  const double eps2 = a.softening_factor * a.softening_factor;  // args.rs:115
  const int precision = precision_id(
      !a.precision.empty() ? a.precision : (a.gpu ? "f64" : "f64-strict"));
  const int integ = integrator_id(a.integrator);
  const uint64_t num_steps = a.num_steps;
  const uint64_t one_percent = std::max<uint64_t>(num_steps / 100, 1);

  nbody_ctx* ctx = nullptr;
  nbody_bodies* bodies = nullptr;
  std::string sim_error;
  auto finish = [&](int code) {
    chan.close();
    writer_thread.join();
    if (bodies) nbody_bodies_destroy(bodies);
    if (ctx) nbody_ctx_destroy(ctx);
    const int rc = nbody_writer_close(writer);
    if (code == 0 && !write_error.empty()) fail(write_error);
    if (code == 0 && rc != 0) fail(last_error());
    if (code != 0) fail(sim_error, code);
    return 0;
  };

  if (n > 0) {
    if (nbody_ctx_create(a.device, &ctx) != 0 ||
        nbody_bodies_create(ctx, n, ic[0].data(), ic[1].data(), ic[2].data(),
                            ic[3].data(), ic[4].data(), ic[5].data(),
                            ic[6].data(), &bodies) != 0 ||
        nbody_bodies_set_theta(bodies, a.theta) != 0) {
      sim_error = last_error();
      return finish(1);
    }
  }

  // Snapshot at t = 0: the initial positions as uploaded.
  double time = 0.0;
  {
    auto b = std::make_unique<Batch>();
    b->n = n;
    b->times.push_back(time);
    b->x = ic[1];
    b->y = ic[2];
    b->z = ic[3];
    if (!chan.send(std::move(b))) return finish(0);
  }

  // Steps in chunks of about 16 MB of positions (the reference's batch size,
  // src/simulation_runner.rs:12-16), each chunk one call that records the
  // positions after every step on the device.
  const uint64_t per_step = (uint64_t)std::max<int64_t>(n, 1) * 24;
  const uint64_t chunk = std::max<uint64_t>(1, 16000000 / per_step);
  uint64_t step = 0;
  while (step < num_steps) {
    const uint64_t k = std::min(chunk, num_steps - step);
    auto b = std::make_unique<Batch>();
    b->n = n;
    b->times.resize(k);
    b->x.resize(k * (uint64_t)n);
    b->y.resize(k * (uint64_t)n);
    b->z.resize(k * (uint64_t)n);
    if (n > 0 &&
        nbody_step_record(ctx, bodies, integ, gravity, precision, a.time_step, a.g_constant, eps2,
                          (int64_t)k, b->x.data(), b->y.data(), b->z.data(),
                          0) != 0) {
      sim_error = last_error();
      return finish(1);
    }
    for (uint64_t s = 0; s < k; ++s) {
      time += a.time_step;  // accumulated as src/simulation/gpu_simulation.rs:48
      b->times[s] = time;
    }
    if (!chan.send(std::move(b))) break;  // the writer failed
    if (a.progress) {
      for (uint64_t s = step; s < step + k; ++s)
        if (s % one_percent == 0)
          std::printf("%llu\n",
                      (unsigned long long)((s + 1) * 100 / num_steps));
      std::fflush(stdout);
    }
    step += k;
  }
  return finish(0);
}
