// C ABI (include/nbody.h) over the host engine.  No exception crosses the
// boundary: every entry point converts failures to an NBODY_E_* code and
// records the message for nbody_last_error().
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/nbody.h"
#include "engine.h"
#include "formats.h"
#include "kernels.h"
#include "noalloc.h"
#include "server.h"
#include "tuning.h"

using nbody::Engine;
using nbody::Error;
using nbody::hip_check;

namespace {
thread_local std::string g_last_error;

template <typename F>
int guarded(F&& f) {
  try {
    f();
    g_last_error.clear();
    return NBODY_OK;
  } catch (const Error& e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_last_error = "host allocation failed";
    return NBODY_E_NOMEM;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return NBODY_E_DEVICE;
  }
}

void require(bool ok, int code, const char* msg) {
  if (!ok) throw Error(code, msg);
}
}  // namespace

struct nbody_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string name;
  int cus = 0;
  hipEvent_t ev[8] = {};
  nbody::KernelTimer timer;  // force/combine launch timing of its children
  // Bodies and batches made from this context share its stream.  Destroying
  // the context while some live only marks it closed; the last of them to be
  // destroyed frees it, so destroy order does not matter to callers.
  int children = 0;
  bool closed = false;
  void use() const {
    if (closed) throw Error(NBODY_E_STATE, "context was destroyed");
    set_device();
  }
  void set_device() const { hip_check(hipSetDevice(device), "hipSetDevice"); }
};

struct nbody_bodies {
  nbody_ctx* ctx = nullptr;
  std::unique_ptr<Engine> eng;
  int id = -1;
  int64_t n = 0;
  double theta = 0.5;  // src/constants.rs DEFAULT_THETA
};

struct nbody_writer {
  std::unique_ptr<nbody::SnapshotWriter> w;
};

struct nbody_batch {
  nbody_ctx* ctx = nullptr;
  std::unique_ptr<Engine> eng;
  std::vector<int> ids;
  std::vector<int64_t> n, offs;
  int64_t total = 0;
};

static_assert(NBODY_STAGE_EULER == nbody::OP_EULER &&
                  NBODY_STAGE_VV_DRIFT == nbody::OP_VV_DRIFT &&
                  NBODY_STAGE_VV_KICK == nbody::OP_VV_KICK &&
                  NBODY_STAGE_RK4_1 == nbody::OP_RK4_1 &&
                  NBODY_STAGE_RK4_4 == nbody::OP_RK4_4,
              "public stage ids are the kernel's");

namespace {
int mode_of(int precision) {
  switch (precision) {
    case NBODY_PRECISION_F64_STRICT:
      return nbody::MODE_F64_STRICT;
    case NBODY_PRECISION_F64:
      return nbody::MODE_F64;
    case NBODY_PRECISION_F32C:
      return nbody::MODE_F32C;
    case NBODY_PRECISION_F32R:
      return nbody::MODE_F32R;
  }
  throw Error(NBODY_E_INVALID, "unknown precision");
}

int integ_of(int integrator) {
  switch (integrator) {
    case NBODY_INTEGRATOR_EULER:
      return nbody::INTEG_EULER;
    case NBODY_INTEGRATOR_VELOCITY_VERLET:
      return nbody::INTEG_VV;
    case NBODY_INTEGRATOR_RUNGE_KUTTA:
      return nbody::INTEG_RK4;
  }
  throw Error(NBODY_E_INVALID, "unknown integrator");
}

// The GPU offers the direct sum only, as the reference's
// GravityMethod::gpu_create (src/args.rs:111-123), which panics on the others.
void check_gravity(int gravity_method) {
  if (gravity_method == NBODY_GRAVITY_NEWTON_PARALLEL) return;
  if (gravity_method == NBODY_GRAVITY_NEWTON ||
      gravity_method == NBODY_GRAVITY_BARNES_HUT)
    throw Error(NBODY_E_UNSUPPORTED,
                "GPU-accelerated version of this gravity method is not "
                "implemented; use newton-parallel");
  throw Error(NBODY_E_INVALID, "unknown gravity method");
}

// Any finite step is accepted, as by the reference's integrators (a
// negative step runs the system backwards, zero leaves it in place).
// Device bodies also run Barnes-Hut (the reference has it on the CPU only,
// src/gravity.rs:14); newton stays the CPU method as in gpu_create.
int bodies_gravity(int gravity_method) {
  if (gravity_method == NBODY_GRAVITY_BARNES_HUT) return nbody::GRAV_BARNES_HUT;
  check_gravity(gravity_method);
  return nbody::GRAV_DIRECT;
}

void check_dt(double dt) {
  require(std::isfinite(dt), NBODY_E_INVALID, "time step must be finite");
}
}  // namespace

extern "C" {

const char* nbody_last_error(void) { return g_last_error.c_str(); }

const char* nbody_version(void) { return "nbody_amd 0.1 (gfx950)"; }

int nbody_ctx_create(int device, nbody_ctx** out) {
  return guarded([&] {
    require(out != nullptr, NBODY_E_INVALID, "null output pointer");
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
      throw Error(NBODY_E_DEVICE,
                  "no GPU available: the GPU gravity path needs a visible "
                  "MI355X (gfx950) device");
    require(device >= 0 && device < count, NBODY_E_INVALID,
            "device index out of range");
    auto c = std::make_unique<nbody_ctx>();
    c->device = device;
    c->use();
    hipDeviceProp_t prop;
    hip_check(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
    c->name = prop.gcnArchName;
    c->cus = prop.multiProcessorCount;
    hip_check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking),
              "hipStreamCreate");
    // the allocation-free entry point's scratch address on this device
    nbody::noalloc_init(device);
    // init check: one probe launch must run and write its marker
    int* d = nullptr;
    int h = 0;
    hip_check(hipMalloc((void**)&d, sizeof(int)), "hipMalloc(probe)");
    hipError_t e = nbody::launch_probe(d, c->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(&h, d, sizeof(int), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(d);
    if (e != hipSuccess || h != 0x6e626479) {
      (void)hipStreamDestroy(c->stream);
      throw Error(NBODY_E_DEVICE,
                  std::string("GPU initialisation check failed: ") +
                      hipGetErrorString(e) +
                      " (is the library built for this GPU's architecture?)");
    }
    *out = c.release();
  });
}

}  // extern "C"

namespace {
void free_ctx(nbody_ctx* ctx) {
  ctx->set_device();
  (void)hipStreamSynchronize(ctx->stream);
  for (auto& e : ctx->ev)
    if (e) (void)hipEventDestroy(e);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

// Called after a child object is gone.
void release_child(nbody_ctx* ctx) {
  if (--ctx->children == 0 && ctx->closed) free_ctx(ctx);
}
}  // namespace

extern "C" {

int nbody_ctx_destroy(nbody_ctx* ctx) {
  return guarded([&] {
    if (!ctx) return;
    require(!ctx->closed, NBODY_E_STATE, "context already destroyed");
    if (ctx->children > 0) {
      ctx->set_device();
      hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
      ctx->closed = true;
      return;
    }
    free_ctx(ctx);
  });
}

int nbody_ctx_synchronize(nbody_ctx* ctx) {
  return guarded([&] {
    require(ctx != nullptr, NBODY_E_INVALID, "null context");
    ctx->use();
    hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  });
}

int nbody_ctx_device_info(nbody_ctx* ctx, char* name, int name_len,
                          int* compute_units) {
  return guarded([&] {
    require(ctx != nullptr, NBODY_E_INVALID, "null context");
    if (name && name_len > 0) {
      std::strncpy(name, ctx->name.c_str(), (size_t)name_len - 1);
      name[name_len - 1] = '\0';
    }
    if (compute_units) *compute_units = ctx->cus;
  });
}

int nbody_ctx_record(nbody_ctx* ctx, int slot) {
  return guarded([&] {
    require(ctx && slot >= 0 && slot < 8, NBODY_E_INVALID, "bad event slot");
    ctx->use();
    if (!ctx->ev[slot])
      hip_check(hipEventCreate(&ctx->ev[slot]), "hipEventCreate");
    hip_check(hipEventRecord(ctx->ev[slot], ctx->stream), "hipEventRecord");
  });
}

int nbody_ctx_elapsed(nbody_ctx* ctx, int a, int b, float* ms) {
  return guarded([&] {
    require(ctx && ms && a >= 0 && a < 8 && b >= 0 && b < 8 && ctx->ev[a] &&
                ctx->ev[b],
            NBODY_E_INVALID, "bad event slots");
    ctx->use();
    hip_check(hipEventSynchronize(ctx->ev[b]), "hipEventSynchronize");
    hip_check(hipEventElapsedTime(ms, ctx->ev[a], ctx->ev[b]),
              "hipEventElapsedTime");
  });
}

int nbody_ctx_kernel_timing(nbody_ctx* ctx, int enable) {
  return guarded([&] {
    require(ctx != nullptr, NBODY_E_INVALID, "null context");
    ctx->use();
    ctx->timer.reset();
    ctx->timer.on = enable != 0;
  });
}

int nbody_ctx_kernel_times(nbody_ctx* ctx, double out[4]) {
  return guarded([&] {
    require(ctx && out, NBODY_E_INVALID, "null argument");
    ctx->use();
    ctx->timer.harvest();
    out[0] = ctx->timer.ms[nbody::KernelTimer::FORCE];
    out[1] = (double)ctx->timer.launches[nbody::KernelTimer::FORCE];
    out[2] = ctx->timer.ms[nbody::KernelTimer::COMBINE];
    out[3] = (double)ctx->timer.launches[nbody::KernelTimer::COMBINE];
  });
}

// ---- tuning ---------------------------------------------------------------
int nbody_tuning_set_split(const int* parts, int count) {
  return guarded([&] {
    require(parts && count > 0, NBODY_E_INVALID, "bad tuning table");
    nbody::tuning_set(parts, count);
  });
}

int nbody_tuning_get_split(int* parts, int count) {
  return guarded([&] {
    require(parts && count > 0, NBODY_E_INVALID, "bad tuning table");
    nbody::tuning_get(parts, count);
  });
}

int nbody_tuning_set_variant(int precision, const int* variants, int count) {
  return guarded([&] {
    require(variants && count > 0, NBODY_E_INVALID, "bad tuning table");
    const int mode = mode_of(precision);
    require(mode != nbody::MODE_F64_STRICT, NBODY_E_INVALID,
            "NBODY_PRECISION_F64_STRICT has one kernel");
    for (int b = 0; b < count; ++b)
      require(variants[b] >= 0 && variants[b] < nbody::kForceVariants,
              NBODY_E_INVALID, "variant out of range");
    nbody::tuning_set_variant(mode, variants, count);
  });
}

int nbody_tuning_get_variant(int precision, int* variants, int count) {
  return guarded([&] {
    require(variants && count > 0, NBODY_E_INVALID, "bad tuning table");
    nbody::tuning_get_variant(mode_of(precision), variants, count);
  });
}

// ---- device bodies ------------------------------------------------------
int nbody_bodies_create(nbody_ctx* ctx, int64_t n, const double* masses,
                        const double* pos_x, const double* pos_y,
                        const double* pos_z, const double* vel_x,
                        const double* vel_y, const double* vel_z,
                        nbody_bodies** out) {
  return guarded([&] {
    require(ctx && out, NBODY_E_INVALID, "null argument");
    *out = nullptr;
    require(n > 0, NBODY_E_INVALID, "bodies must hold n >= 1 bodies");
    ctx->use();
    auto b = std::make_unique<nbody_bodies>();
    b->ctx = ctx;
    b->n = n;
    b->eng = std::make_unique<Engine>(ctx->device, ctx->stream,
                                      nbody::MODE_F64, n, 1);
    b->eng->set_timer(&ctx->timer);
    nbody::SystemParams p;
    p.n = n;
    b->id = b->eng->add(p, masses, pos_x, pos_y, pos_z, vel_x, vel_y, vel_z);
    ++ctx->children;
    *out = b.release();
  });
}

int nbody_bodies_download(nbody_ctx* ctx, nbody_bodies* b, double* masses,
                          double* pos_x, double* pos_y, double* pos_z,
                          double* vel_x, double* vel_y, double* vel_z) {
  return guarded([&] {
    require(ctx && b && b->ctx == ctx, NBODY_E_INVALID, "bad context/bodies");
    ctx->use();
    b->eng->download(b->id, masses, pos_x, pos_y, pos_z, vel_x, vel_y, vel_z);
  });
}

int nbody_bodies_destroy(nbody_bodies* b) {
  return guarded([&] {
    if (!b) return;
    nbody_ctx* ctx = b->ctx;
    ctx->set_device();
    delete b;
    release_child(ctx);
  });
}

int64_t nbody_bodies_len(const nbody_bodies* b) { return b ? b->n : 0; }

int nbody_bodies_set_theta(nbody_bodies* b, double theta) {
  return guarded([&] {
    require(b != nullptr, NBODY_E_INVALID, "null bodies");
    require(std::isfinite(theta) && theta >= 0.0, NBODY_E_INVALID,
            "theta must be finite and >= 0");
    b->theta = theta;
  });
}

int nbody_bodies_tree_nodes(nbody_bodies* b, int64_t* nodes) {
  return guarded([&] {
    require(b && nodes, NBODY_E_INVALID, "null argument");
    b->ctx->use();
    *nodes = b->eng->tree_nodes(b->id);
  });
}

// ---- gravity / integrator ----------------------------------------------
int nbody_calculate_accelerations(nbody_ctx* ctx, nbody_bodies* b,
                                  int gravity_method, double g, double eps2,
                                  int precision, double* acc_x, double* acc_y,
                                  double* acc_z) {
  return guarded([&] {
    require(ctx && b && b->ctx == ctx, NBODY_E_INVALID, "bad context/bodies");
    const int grav = bodies_gravity(gravity_method);
    ctx->use();
    b->eng->set_mode(mode_of(precision));
    b->eng->set_gravity(grav, b->theta);
    nbody::SystemParams p = b->eng->params(b->id);
    p.g = g;
    p.eps2 = eps2;
    b->eng->set_params(b->id, p);
    b->eng->accelerations(b->id, acc_x, acc_y, acc_z);
  });
}

int nbody_calculate_accelerations_noalloc(
    nbody_ctx* ctx, int64_t n, const double* masses, const double* pos_x,
    const double* pos_y, const double* pos_z, double g, double eps2,
    int precision, double* acc_x, double* acc_y, double* acc_z) {
  return guarded([&] {
    require(ctx != nullptr, NBODY_E_INVALID, "null context");
    const int mode = mode_of(precision);
    ctx->use();
    nbody::noalloc_accelerations(ctx->device, ctx->stream, mode, n, masses,
                                 pos_x, pos_y, pos_z, g, eps2, acc_x, acc_y,
                                 acc_z);
  });
}

int nbody_integrate_stage(nbody_ctx* ctx, nbody_bodies* b, int stage,
                          double dt) {
  return guarded([&] {
    require(ctx && b && b->ctx == ctx, NBODY_E_INVALID, "bad context/bodies");
    require(stage >= NBODY_STAGE_EULER && stage <= NBODY_STAGE_RK4_4,
            NBODY_E_INVALID, "unknown integrator stage");
    check_dt(dt);
    ctx->use();
    b->eng->check_stage(b->id, stage);  // before dt changes
    nbody::SystemParams p = b->eng->params(b->id);
    p.dt = dt;
    b->eng->set_params(b->id, p);
    b->eng->integrate_stage(b->id, stage);
  });
}

int nbody_step(nbody_ctx* ctx, nbody_bodies* b, int integrator,
               int gravity_method, int precision, double dt, double g,
               double eps2, int64_t nsteps) {
  return guarded([&] {
    require(ctx && b && b->ctx == ctx, NBODY_E_INVALID, "bad context/bodies");
    const int grav = bodies_gravity(gravity_method);
    check_dt(dt);
    require(nsteps >= 0, NBODY_E_INVALID, "nsteps < 0");
    ctx->use();
    b->eng->set_mode(mode_of(precision));
    b->eng->set_gravity(grav, b->theta);
    nbody::SystemParams p = b->eng->params(b->id);
    p.integ = integ_of(integrator);
    p.dt = dt;
    p.g = g;
    p.eps2 = eps2;
    b->eng->set_params(b->id, p);
    b->eng->step({b->id}, nsteps);
  });
}

int nbody_step_record(nbody_ctx* ctx, nbody_bodies* b, int integrator,
                      int gravity_method, int precision, double dt, double g,
                      double eps2, int64_t nsteps, double* pos_x,
                      double* pos_y, double* pos_z, int64_t max_stage_bytes) {
  return guarded([&] {
    require(ctx && b && b->ctx == ctx, NBODY_E_INVALID, "bad context/bodies");
    const int grav = bodies_gravity(gravity_method);
    check_dt(dt);
    require(nsteps >= 0, NBODY_E_INVALID, "nsteps < 0");
    require(nsteps == 0 || (pos_x && pos_y && pos_z), NBODY_E_INVALID,
            "null trajectory array");
    require(max_stage_bytes >= 0, NBODY_E_INVALID, "max_stage_bytes < 0");
    ctx->use();
    b->eng->set_mode(mode_of(precision));
    b->eng->set_gravity(grav, b->theta);
    nbody::SystemParams p = b->eng->params(b->id);
    p.integ = integ_of(integrator);
    p.dt = dt;
    p.g = g;
    p.eps2 = eps2;
    b->eng->set_params(b->id, p);
    const size_t cap = max_stage_bytes > 0 ? (size_t)max_stage_bytes
                                           : (size_t)64 << 20;
    b->eng->record(b->id, nsteps, pos_x, pos_y, pos_z, cap);
  });
}

namespace {
void pack_diag(const double* d, double* out) {
  out[0] = d[nbody::DIAG_K];
  out[1] = d[nbody::DIAG_U];
  out[2] = d[nbody::DIAG_K] + d[nbody::DIAG_U];
  out[3] = d[nbody::DIAG_PX];
  out[4] = d[nbody::DIAG_PY];
  out[5] = d[nbody::DIAG_PZ];
  const double M = d[nbody::DIAG_M];
  out[6] = M != 0.0 ? d[nbody::DIAG_MX] / M : 0.0;
  out[7] = M != 0.0 ? d[nbody::DIAG_MY] / M : 0.0;
  out[8] = M != 0.0 ? d[nbody::DIAG_MZ] / M : 0.0;
  out[9] = M;
}
}  // namespace

int nbody_energy_momentum(nbody_ctx* ctx, nbody_bodies* b, double g,
                          double eps2, double out[10]) {
  return guarded([&] {
    require(ctx && b && b->ctx == ctx && out, NBODY_E_INVALID, "bad argument");
    ctx->use();
    nbody::SystemParams p = b->eng->params(b->id);
    if (p.g != g || p.eps2 != eps2) {
      p.g = g;
      p.eps2 = eps2;
      b->eng->set_params(b->id, p);
    }
    double d[nbody::DIAG_COUNT];
    b->eng->diagnostics({b->id}, d);
    pack_diag(d, out);
  });
}

// ---- batched form -------------------------------------------------------
int nbody_batch_create(nbody_ctx* ctx, int nsys, const int64_t* n,
                       const int* integrator, const double* dt,
                       const double* g, const double* eps2, int precision,
                       const double* masses, const double* pos_x,
                       const double* pos_y, const double* pos_z,
                       const double* vel_x, const double* vel_y,
                       const double* vel_z, nbody_batch** out) {
  return guarded([&] {
    require(ctx && out && n && integrator && dt && g && eps2, NBODY_E_INVALID,
            "null argument");
    *out = nullptr;
    require(nsys > 0, NBODY_E_INVALID, "batch needs at least one system");
    ctx->use();
    auto bt = std::make_unique<nbody_batch>();
    bt->ctx = ctx;
    int64_t cap = 0;
    for (int s = 0; s < nsys; ++s) {
      require(n[s] > 0, NBODY_E_INVALID, "every system needs n >= 1");
      check_dt(dt[s]);
      bt->n.push_back(n[s]);
      bt->offs.push_back(bt->total);
      bt->total += n[s];
      cap += (n[s] + 63) / 64 * 64;
    }
    bt->eng = std::make_unique<Engine>(ctx->device, ctx->stream,
                                       mode_of(precision), cap, nsys);
    // hundreds of resident systems fill the GPU with waves already: the
    // batched form keeps each system to at most 4 source parts (measured on
    // 512 star systems of 64..8192 bodies: 3.3 % faster than 8 in f32r,
    // 1.8 % in f32c; 6-12 % faster than 16 or 32)
    bt->eng->set_split_cap(4);
    bt->eng->set_timer(&ctx->timer);
    for (int s = 0; s < nsys; ++s) {
      nbody::SystemParams p;
      p.n = n[s];
      p.integ = integ_of(integrator[s]);
      p.dt = dt[s];
      p.g = g[s];
      p.eps2 = eps2[s];
      const int64_t o = bt->offs[s];
      bt->ids.push_back(bt->eng->add(p, masses + o, pos_x + o, pos_y + o,
                                     pos_z + o, vel_x + o, vel_y + o,
                                     vel_z + o));
    }
    ++ctx->children;
    *out = bt.release();
  });
}

int nbody_batch_step(nbody_ctx* ctx, nbody_batch* bt, int64_t nsteps) {
  return guarded([&] {
    require(ctx && bt && bt->ctx == ctx, NBODY_E_INVALID, "bad context/batch");
    require(nsteps >= 0, NBODY_E_INVALID, "nsteps < 0");
    ctx->use();
    bt->eng->step_graph(bt->ids, nsteps);
  });
}

int nbody_batch_download(nbody_ctx* ctx, nbody_batch* bt, double* pos_x,
                         double* pos_y, double* pos_z, double* vel_x,
                         double* vel_y, double* vel_z) {
  return guarded([&] {
    require(ctx && bt && bt->ctx == ctx, NBODY_E_INVALID, "bad context/batch");
    ctx->use();
    for (size_t s = 0; s < bt->ids.size(); ++s) {
      const int64_t o = bt->offs[s];
      auto at = [o](double* p) { return p ? p + o : nullptr; };
      bt->eng->download(bt->ids[s], nullptr, at(pos_x), at(pos_y), at(pos_z),
                        at(vel_x), at(vel_y), at(vel_z));
    }
  });
}

int nbody_batch_energy(nbody_ctx* ctx, nbody_batch* bt, double* out) {
  return guarded([&] {
    require(ctx && bt && bt->ctx == ctx && out, NBODY_E_INVALID,
            "bad argument");
    ctx->use();
    std::vector<double> d(nbody::DIAG_COUNT * bt->ids.size());
    bt->eng->diagnostics(bt->ids, d.data());
    for (size_t s = 0; s < bt->ids.size(); ++s)
      pack_diag(&d[s * nbody::DIAG_COUNT], out + 10 * s);
  });
}

int nbody_batch_destroy(nbody_batch* bt) {
  return guarded([&] {
    if (!bt) return;
    nbody_ctx* ctx = bt->ctx;
    ctx->set_device();
    delete bt;
    release_child(ctx);
  });
}

// ---- served form --------------------------------------------------------
int nbody_server_create(const nbody_server_config* cfg, nbody_server** out) {
  return guarded([&] {
    require(cfg && out, NBODY_E_INVALID, "null argument");
    *out = nullptr;
    *out = reinterpret_cast<nbody_server*>(nbody::server_create(*cfg));
  });
}

int nbody_server_submit(nbody_server* s, const nbody_request* req,
                        uint64_t* request_id) {
  return guarded([&] {
    require(s && req && request_id, NBODY_E_INVALID, "null argument");
    check_gravity(req->gravity_method);
    *request_id = nbody::server_submit(reinterpret_cast<nbody::Server*>(s), *req);
  });
}

int nbody_server_poll(nbody_server* s, int timeout_ms, nbody_event* ev,
                      double* pos_x, double* pos_y, double* pos_z,
                      double* vel_x, double* vel_y, double* vel_z,
                      int64_t capacity) {
  return guarded([&] {
    require(s && ev, NBODY_E_INVALID, "null argument");
    nbody::server_poll(reinterpret_cast<nbody::Server*>(s), timeout_ms, ev,
                       pos_x, pos_y, pos_z, vel_x, vel_y, vel_z, capacity);
  });
}

int nbody_server_cancel(nbody_server* s, uint64_t request_id) {
  return guarded([&] {
    require(s != nullptr, NBODY_E_INVALID, "null server");
    nbody::server_cancel(reinterpret_cast<nbody::Server*>(s), request_id);
  });
}

int nbody_server_stats(nbody_server* s, double out[8]) {
  return guarded([&] {
    require(s && out, NBODY_E_INVALID, "null argument");
    nbody::server_stats(reinterpret_cast<nbody::Server*>(s), out);
  });
}

int nbody_server_destroy(nbody_server* s) {
  return guarded([&] {
    if (s) nbody::server_destroy(reinterpret_cast<nbody::Server*>(s));
  });
}

// ---- files ----------------------------------------------------------------
int nbody_writer_open(const char* path, nbody_writer** out) {
  return guarded([&] {
    require(path && out, NBODY_E_INVALID, "null argument");
    *out = nullptr;
    auto w = std::make_unique<nbody_writer>();
    w->w = std::make_unique<nbody::SnapshotWriter>(path);
    *out = w.release();
  });
}

int nbody_writer_append(nbody_writer* w, double time, int64_t n,
                        const double* pos_x, const double* pos_y,
                        const double* pos_z) {
  return guarded([&] {
    require(w != nullptr, NBODY_E_INVALID, "null writer");
    w->w->append(time, n, pos_x, pos_y, pos_z);
  });
}

int nbody_writer_close(nbody_writer* w) {
  std::unique_ptr<nbody_writer> owned(w);
  return guarded([&] {
    if (owned) owned->w->close();
  });
}

int nbody_format_f64(double v, char buf[32]) {
  return guarded([&] {
    require(buf != nullptr, NBODY_E_INVALID, "null buffer");
    const int len = nbody::format_f64(v, buf);
    buf[len] = '\0';
  });
}

int nbody_parse_f64(const char* text, double* out) {
  return guarded([&] {
    require(text && out, NBODY_E_INVALID, "null argument");
    require(nbody::parse_f64(text, std::strlen(text), out), NBODY_E_INVALID,
            "invalid float literal");
  });
}

int nbody_load_initial_conditions(const char* path, int64_t* n,
                                  double* masses, double* pos_x,
                                  double* pos_y, double* pos_z, double* vel_x,
                                  double* vel_y, double* vel_z) {
  return guarded([&] {
    require(path && n, NBODY_E_INVALID, "null argument");
    const nbody::InitialConditions ic = nbody::load_initial_conditions(path);
    const int64_t count = (int64_t)ic.size();
    double* dst[7] = {masses, pos_x, pos_y, pos_z, vel_x, vel_y, vel_z};
    bool all = true;
    for (double* d : dst) all = all && d != nullptr;
    const int64_t cap = *n;
    *n = count;
    if (!all) return;
    require(cap >= count, NBODY_E_NOSPACE,
            "initial conditions hold more bodies than the arrays");
    const std::vector<double>* src[7] = {&ic.m,  &ic.x,  &ic.y, &ic.z,
                                         &ic.vx, &ic.vy, &ic.vz};
    for (int c = 0; c < 7; ++c)
      if (count) std::memcpy(dst[c], src[c]->data(), sizeof(double) * count);
  });
}

}  // extern "C"
