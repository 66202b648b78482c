/*
 * nbody.h -- C ABI of the MI355X (gfx950) direct-sum gravity + integrator
 * library (libnbody_amd.so).
 *
 * The entry points are what the reference's GPU path binds through its Rust
 * runtime; each cites the reference interface it replaces.  Conventions:
 *   - every call returns int: NBODY_OK (0) or a negative NBODY_E_* code;
 *     nbody_last_error() returns the calling thread's last message;
 *   - no exceptions or C++ types cross the ABI; host arrays are caller-owned
 *     and sized by the explicit counts; device memory is library-owned;
 *   - arrays are structure-of-arrays binary64, the reference's layout
 *     (src/body.rs:52-61);
 *   - a context and everything created from it are used by one thread at a
 *     time (the reference moves its simulation onto a single thread,
 *     src/n_body_sim.rs:47-48); the server API is the exception and is
 *     thread-safe;
 *   - work is queued on the context's stream; calls that return host data
 *     (download, energy, accelerations with host outputs) synchronise.
 *
 * A Rust maintainer binds these 1:1 in an `extern "C"` block in the
 * reference's src/gpu.rs.
 */
#ifndef NBODY_AMD_H
#define NBODY_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define NBODY_OK 0
#define NBODY_E_INVALID (-1)      /* bad argument (N=0, non-finite value or dt, ...) */
#define NBODY_E_UNSUPPORTED (-2)  /* method/integrator combination not offered on GPU */
#define NBODY_E_DEVICE (-3)       /* HIP runtime error / no usable GPU */
#define NBODY_E_NOMEM (-4)        /* device allocation failed */
#define NBODY_E_NOSPACE (-5)      /* arena or admission budget full */
#define NBODY_E_NOTFOUND (-6)     /* unknown request or system id */
#define NBODY_E_STATE (-7)        /* call not valid in the object's state */
#define NBODY_E_TIMEOUT (-8)      /* poll timed out with no event */

/* ---- method selection (src/args.rs:84-187) ---------------------------- */
/* GravityMethod, same order as the reference enum (src/args.rs:84-89). */
#define NBODY_GRAVITY_NEWTON 0           /* not offered on GPU (src/gravity.rs:12) */
#define NBODY_GRAVITY_NEWTON_PARALLEL 1  /* the GPU direct sum */
#define NBODY_GRAVITY_BARNES_HUT 2       /* the GPU octree, device bodies only
                                            (the batched and served forms run
                                            the direct sum) */
/* IntegratorMethod (src/args.rs:140-144). */
#define NBODY_INTEGRATOR_EULER 0
#define NBODY_INTEGRATOR_VELOCITY_VERLET 1
#define NBODY_INTEGRATOR_RUNGE_KUTTA 2
/* Arithmetic of the pair terms (no reference counterpart: the reference GPU
 * kernel is f64 only, cuda/kernels.cu:11-52). */
#define NBODY_PRECISION_F64_STRICT 0  /* reference operation order, IEEE sqrt and divide, no FMA: bit-identical to the f64 CPU path */
#define NBODY_PRECISION_F64 1         /* f64, rsq + refinement, FMA; same j order */
#define NBODY_PRECISION_F32C 2        /* f64 positions, each difference rounded once to f32, f32 pair terms, compensated sums; f64 state */
#define NBODY_PRECISION_F32R 3        /* bodies in Hilbert order, coordinates relative to each wave's middle target split into an f32 value and an f32 residual; f32 pair terms, the residuals added for source chunks near the targets, r^2 of chunks far from them from bf16-split Gram products on the matrix cores; compensated sums; <= 1e-5 relative; faster than F32C for systems above about 1500 bodies */
/* Zero softening (eps2 == 0) with two distinct bodies at one point:
 * F64_STRICT, F64 and F32C follow the reference (1/sqrt(0) * 0 gives NaN
 * for both bodies).  F32R excludes a target's own term by its zero
 * difference, so such bodies get the finite sum over their other sources. */
/* Range of the f32 forms (F32C, F32R): r^2 and k = g*m*(r^2 + eps2)^-1.5
 * are formed in f32.
 *  - Pair distances must stay below about 1e18 (r^2 in f32 range), and
 *    beyond about 4e12 the factor r^-3 leaves f32's normal range (it
 *    vanishes beyond about 1e15): systems wider than about 1e12 length
 *    units need NBODY_PRECISION_F64.
 *  - Tiny softening: once g*m*eps^-3 exceeds FLT_MAX (eps below about
 *    1.4e-13 for g*m = 1), k overflows at close range.  F32C skips the
 *    target's own index.  F32R drops its own pair by the zero difference:
 *    it runs unmasked only while g*max(m)*eps^-3 < 2^120 (the self pair
 *    then adds exactly 0 * k); otherwise it clamps k to +-FLT_MAX, so a pair
 *    at zero difference adds zero and every other pair stays finite.
 *  - Distinct bodies closer than the softening can also overflow k in F32C
 *    when g*m*eps^-3 exceeds FLT_MAX; use NBODY_PRECISION_F64 there. */

typedef struct nbody_ctx nbody_ctx;
typedef struct nbody_bodies nbody_bodies;
typedef struct nbody_batch nbody_batch;
typedef struct nbody_server nbody_server;

/* Message of the calling thread's last failed call ("" if none). */
const char *nbody_last_error(void);
/* Library version string. */
const char *nbody_version(void);

/* ---- context: replaces CudaManager::new (src/gpu/cuda_manager.rs:21-54)
 * including its gpu_init_check launch (:56-61, cuda/kernels.cu:4-9). */
int nbody_ctx_create(int device, nbody_ctx **out);
int nbody_ctx_destroy(nbody_ctx *ctx);
/* Waits for all queued work (the reference's stream.synchronize(),
 * src/simulation/gpu_simulation.rs:54). */
int nbody_ctx_synchronize(nbody_ctx *ctx);
/* Device name and compute-unit count, for reports. */
int nbody_ctx_device_info(nbody_ctx *ctx, char *name, int name_len,
                          int *compute_units);

/* Timing on the context's stream (the stream every kernel of this context
 * runs on): record event slot 0..7 now; elapsed milliseconds between two
 * recorded slots (synchronises on the later one). */
int nbody_ctx_record(nbody_ctx *ctx, int slot);
int nbody_ctx_elapsed(nbody_ctx *ctx, int slot_from, int slot_to, float *ms);
/* Launch timing (no reference counterpart; the reference times whole steps
 * under criterion, benches/gpu_integrator_bench.rs:155-162): with enable != 0
 * every force and combine launch of this context's bodies and batches is
 * bracketed by two HIP events on the context's stream (batched stepping then
 * launches without its graph).  Enabling or disabling zeroes the sums.
 * nbody_ctx_kernel_times waits for the recorded launches and returns out[0]
 * force-launch milliseconds, out[1] force launches, out[2] combine-launch
 * milliseconds, out[3] combine launches, summed since timing was enabled. */
int nbody_ctx_kernel_timing(nbody_ctx *ctx, int enable);
int nbody_ctx_kernel_times(nbody_ctx *ctx, double out[4]);

/* ---- tuning (no reference counterpart): source-range parts per target
 * tile for systems of n bodies, by bucket b = floor(log2 n), b = 0..31.
 * Loaded once per process from tuning/gfx950.txt next to the library (or the
 * file named by NBODY_AMD_TUNING), else a measured default rule.  The parts
 * set the summation order of NBODY_PRECISION_F64 / F32C / F32R (never of
 * F64_STRICT), so a fixed table gives fixed bits; a change applies to
 * objects created afterwards. */
int nbody_tuning_set_split(const int *parts, int count);
int nbody_tuning_get_split(int *parts, int count);
/* The force-kernel variant per bucket for one precision (F64, F32C or
 * F32R; F64_STRICT has one kernel): variant = 3 * w + shape, w picking 1, 2
 * or 4 waves per workgroup and shape one of three loop shapes (sources per
 * chunk and chunks per compensated group in F64 and F32C, sources per LDS
 * batch in F32R).  Loaded with the split table ("variant <precision>
 * <bucket> <variant>" lines); default 0 (one wave per workgroup, shape 0).
 * The shape sets the summation order of F64 and F32C (source parts are
 * aligned to the chunk size; F32C also adds its compensated groups of
 * chunks), so, like the split, a system keeps the variant of its size from
 * its creation on. */
int nbody_tuning_set_variant(int precision, const int *variants, int count);
int nbody_tuning_get_variant(int precision, int *variants, int count);

/* ---- device bodies: replaces DeviceBodies (src/gpu/device_bodies.rs:7-55).
 * create = DeviceBodies::new (:23-36, upload of the 7 SoA arrays plus zeroed
 * accelerations); download = DeviceBodies::download (:40-50) plus the
 * synchronisation the reference leaves to the caller. */
int nbody_bodies_create(nbody_ctx *ctx, int64_t n, const double *masses,
                        const double *pos_x, const double *pos_y,
                        const double *pos_z, const double *vel_x,
                        const double *vel_y, const double *vel_z,
                        nbody_bodies **out);
int nbody_bodies_download(nbody_ctx *ctx, nbody_bodies *b, double *masses,
                          double *pos_x, double *pos_y, double *pos_z,
                          double *vel_x, double *vel_y, double *vel_z);
int nbody_bodies_destroy(nbody_bodies *b);
int64_t nbody_bodies_len(const nbody_bodies *b);
/* Barnes-Hut opening angle used when gravity_method is
 * NBODY_GRAVITY_BARNES_HUT: replaces the theta of BarnesHutGravity::new
 * (src/gravity/cpu/barnes_hut.rs:15-21); default 0.5
 * (src/constants.rs:21-22). */
int nbody_bodies_set_theta(nbody_bodies *b, double theta);
/* Octree nodes of the last Barnes-Hut build of these bodies (0 before any);
 * the reference's node_count (octree.rs:353-357).  Synchronous. */
int nbody_bodies_tree_nodes(nbody_bodies *b, int64_t *nodes);

/* ---- gravity: replaces GpuGravity::calculate_accelerations
 * (src/gravity.rs:30-36; src/gravity/gpu/newton_parallel.rs:20-46).
 * gravity_method NBODY_GRAVITY_NEWTON_PARALLEL is the direct sum;
 * NBODY_GRAVITY_BARNES_HUT is the reference's Barnes-Hut octree
 * (src/gravity/cpu/barnes_hut.rs:25-51, rebuilt for every evaluation, with
 * the bodies' theta; in F64_STRICT bit-identical to the reference's CPU
 * Barnes-Hut, in the other precisions f64 with the fast inverse square root).
 * NBODY_GRAVITY_NEWTON returns NBODY_E_UNSUPPORTED, where the reference
 * panics (src/args.rs:117-120).
 * eps2 is the squared softening as the reference GPU path takes it
 * (src/args.rs:115).  Results stay in the device acceleration arrays
 * (DeviceBodies::acc_*); when acc_x/acc_y/acc_z are non-null they are also
 * copied to the host (synchronous). */
int nbody_calculate_accelerations(nbody_ctx *ctx, nbody_bodies *b,
                                  int gravity_method, double g, double eps2,
                                  int precision, double *acc_x,
                                  double *acc_y, double *acc_z);

/* ---- gravity on caller-owned device arrays, allocating nothing.  No
 * counterpart in the reference: this is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  masses, pos_* (inputs) and acc_*
 * (outputs) are DEVICE pointers to n doubles each, owned by the caller.
 * The evaluation is queued on the context's stream and returns without
 * waiting; its results are the bits nbody_calculate_accelerations gives
 * for the same bodies and arguments (direct sum).  The library's working
 * memory (work lists, g*m, a padded position snapshot, split partial sums)
 * is one static device array of 64 MiB per GPU, held until the process
 * ends: n is at most 65536 (NBODY_E_INVALID beyond), and precision is
 * NBODY_PRECISION_F64_STRICT, F64 or F32C (F32R: NBODY_E_UNSUPPORTED, its
 * body ordering needs temporary memory).  Every call on one GPU shares that
 * array, so calls that overlap on one GPU (contexts on different streams)
 * race: order them. */
int nbody_calculate_accelerations_noalloc(nbody_ctx *ctx, int64_t n,
                                          const double *masses,
                                          const double *pos_x,
                                          const double *pos_y,
                                          const double *pos_z, double g,
                                          double eps2, int precision,
                                          double *acc_x, double *acc_y,
                                          double *acc_z);

/* ---- integrator: replaces GpuIntegrator::step
 * (src/integrators/gpu/gpu_integrator.rs:8-15; src/integrators/gpu/euler.rs:17-43)
 * called nsteps times.  All three integrators are offered (the reference GPU
 * path has Euler only, src/args.rs:165-174).  Velocity Verlet carries
 * a(n+1) into the next step, which gives the same bits as the reference's
 * recomputation (src/integrators/cpu/velocity_verlet.rs:76). */
int nbody_step(nbody_ctx *ctx, nbody_bodies *b, int integrator,
               int gravity_method, int precision, double dt, double g,
               double eps2, int64_t nsteps);

/* ---- one integrator stage on the device acceleration arrays: replaces
 * the reference's euler_step launch (src/integrators/gpu/euler.rs:26-41,
 * cuda/kernels.cu:54-77) and adds the stages of velocity Verlet and RK4, so
 * that a GpuIntegrator can step through the unchanged GpuGravity trait
 * (src/gravity.rs:30-36): gravity.calculate_accelerations (here
 * nbody_calculate_accelerations, which leaves a(x) in the device arrays),
 * then the stage, in the reference's order:
 *   euler:           F, EULER
 *   velocity-verlet: F, VV_DRIFT, F, VV_KICK  (two evaluations per step,
 *                    src/integrators/cpu/velocity_verlet.rs:76,:90)
 *   runge-kutta:     F, RK4_1, F, RK4_2, F, RK4_3, F, RK4_4
 * RK4_1..RK4_3 move the bodies to the next stage's evaluation point (the
 * step's start is kept on the device) and RK4_4 writes the step's result.
 * The expression trees are those of nbody_step, so a step run this way gives
 * nbody_step's bits for the same precision.  (Except velocity Verlet in
 * NBODY_PRECISION_F32R at a step where the bodies are re-sorted, every 32
 * steps of their own count: the re-evaluated a(n) is summed in the new
 * order, where nbody_step carries the value summed in the old one; the two
 * agree to rounding.)  dt is the step (the same for
 * every stage of a step).  Synchronous.  NBODY_E_INVALID for an unknown
 * stage or a non-finite dt; NBODY_E_STATE when no
 * nbody_calculate_accelerations ran for these bodies since the last stage or
 * step, or when the stage is not the one that comes next (a staged step
 * starts with EULER, VV_DRIFT or RK4_1 and continues in the order above; a
 * whole step by nbody_step ends any staged step under way), and dt is then
 * left unchanged. */
#define NBODY_STAGE_EULER 0     /* v += a dt; x += v dt */
#define NBODY_STAGE_VV_DRIFT 1  /* keep a as a_n; x += (v dt) + ((0.5 a) dt^2) */
#define NBODY_STAGE_VV_KICK 2   /* v += (0.5 (a_n + a)) dt */
#define NBODY_STAGE_RK4_1 3
#define NBODY_STAGE_RK4_2 4
#define NBODY_STAGE_RK4_3 5
#define NBODY_STAGE_RK4_4 6
int nbody_integrate_stage(nbody_ctx *ctx, nbody_bodies *b, int stage,
                          double dt);

/* ---- stepping with a snapshot after every step: replaces the loop of
 * SimulationRunner::run (src/simulation_runner.rs:28-52) over
 * GpuSimulation::step and ::snapshot (src/simulation/gpu_simulation.rs:45-65),
 * which copies all seven state arrays to the host after every step.  Here
 * positions of step s (1-based step s+1) land at pos_x[s*n .. s*n+n) and
 * likewise pos_y/pos_z, each an array of nsteps*n doubles; the steps run back
 * to back and the positions leave the device in one copy per field for every
 * `max_stage_bytes` of device staging (0: 64 MiB; at least one step).
 * Synchronous.  Same arguments and errors as nbody_step. */
int nbody_step_record(nbody_ctx *ctx, nbody_bodies *b, int integrator,
                      int gravity_method, int precision, double dt, double g,
                      double eps2, int64_t nsteps, double *pos_x,
                      double *pos_y, double *pos_z, int64_t max_stage_bytes);

/* ---- diagnostics (no reference counterpart): f64 on the GPU.
 * out[0] kinetic K, out[1] softened potential U, out[2] E = K + U,
 * out[3..5] momentum, out[6..8] centre of mass, out[9] total mass. */
int nbody_energy_momentum(nbody_ctx *ctx, nbody_bodies *b, double g,
                          double eps2, double out[10]);

/* ---- scenario generators (no reference counterpart; the star system
 * follows gui/generators.py:11-98).  Deterministic for a seed (own
 * xoshiro256** stream).  Output arrays hold n values each.  Every
 * parameter must be finite (NBODY_E_INVALID otherwise), so a generator
 * that returns NBODY_OK wrote finite bodies. */
/* Plummer sphere: scale radius a, total mass, G; equal masses, virial
 * velocities, centre-of-mass frame. */
int nbody_scenario_plummer(int64_t n, uint64_t seed, double a,
                           double total_mass, double g, double *masses,
                           double *pos_x, double *pos_y, double *pos_z,
                           double *vel_x, double *vel_y, double *vel_z);
/* Star system: central mass at rest plus n-1 bodies on near-circular
 * Keplerian orbits; body mass log-uniform [m_lo, m_hi], radius uniform
 * [r_lo, r_hi], inclination uniform [0, max_incl_deg). */
int nbody_scenario_star_system(int64_t n, uint64_t seed, double central_mass,
                               double m_lo, double m_hi, double r_lo,
                               double r_hi, double max_incl_deg, double g,
                               double *masses, double *pos_x, double *pos_y,
                               double *pos_z, double *vel_x, double *vel_y,
                               double *vel_z);
/* Two Plummer spheres of n/2 bodies (masses 1/n each), centres separated by
 * `separation` along x, approaching at `rel_speed`. */
int nbody_scenario_colliding(int64_t n, uint64_t seed, double separation,
                             double rel_speed, double g, double *masses,
                             double *pos_x, double *pos_y, double *pos_z,
                             double *vel_x, double *vel_y, double *vel_z);
/* Disc galaxy after the GUI's generator (gui/generators.py:101-153): n
 * bodies of body_mass, radii radius * U^density_power, a thin layer, each on
 * the softened circular speed of the mass inside its radius plus 10 %
 * dispersion; written in order of radius.  The GUI's defaults are radius 20,
 * body_mass 1e-3, density_power 1.5, softening 0.5. */
int nbody_scenario_disc(int64_t n, uint64_t seed, double radius,
                        double body_mass, double density_power,
                        double softening, double g, double *masses,
                        double *pos_x, double *pos_y, double *pos_z,
                        double *vel_x, double *vel_y, double *vel_z);

/* ---- batched form: many independent systems of uneven N in one arena,
 * stepped by shared launches captured in a HIP graph.  Bodies of system s
 * are rows [offsets[s], offsets[s] + n[s]) of the packed host arrays. */
int nbody_batch_create(nbody_ctx *ctx, int nsys, const int64_t *n,
                       const int *integrator, const double *dt,
                       const double *g, const double *eps2, int precision,
                       const double *masses, const double *pos_x,
                       const double *pos_y, const double *pos_z,
                       const double *vel_x, const double *vel_y,
                       const double *vel_z, nbody_batch **out);
int nbody_batch_step(nbody_ctx *ctx, nbody_batch *bt, int64_t nsteps);
int nbody_batch_download(nbody_ctx *ctx, nbody_batch *bt, double *pos_x,
                         double *pos_y, double *pos_z, double *vel_x,
                         double *vel_y, double *vel_z);
/* out: 10 doubles per system, layout of nbody_energy_momentum. */
int nbody_batch_energy(nbody_ctx *ctx, nbody_batch *bt, double *out);
int nbody_batch_destroy(nbody_batch *bt);

/* ---- served form: continuous batching of simulation requests.
 * Requests are admitted by device-memory footprint as slots free up, packed
 * into shared steps with everything else resident, and stream position
 * snapshots back every snapshot_every steps.  A request's events arrive in
 * step order, its snapshots all before its DONE; DONE is queued as soon as
 * the request's last step has run.  Results are the same bits whatever
 * else is resident. */
typedef struct {
  int device;
  int precision;              /* NBODY_PRECISION_* for every request */
  int64_t capacity_bodies;    /* arena size (bodies) */
  int max_systems;            /* resident systems at once */
  int snapshot_every;         /* steps between position snapshots (>=1) */
  int max_snapshot_bytes;     /* cap on queued snapshot bytes (backpressure) */
} nbody_server_config;

typedef struct {
  int64_t n;
  int integrator;             /* NBODY_INTEGRATOR_* */
  int gravity_method;         /* must be NBODY_GRAVITY_NEWTON_PARALLEL */
  double g, dt, eps2;
  int64_t steps;
  int want_snapshots;         /* 0: final state only */
  const double *masses, *pos_x, *pos_y, *pos_z, *vel_x, *vel_y, *vel_z;
} nbody_request;

#define NBODY_EVENT_SNAPSHOT 1  /* positions after `step` steps */
#define NBODY_EVENT_DONE 2      /* final state; positions and velocities */
#define NBODY_EVENT_FAILED 3    /* request failed; see message */
#define NBODY_EVENT_CANCELLED 4

typedef struct {
  int kind;                   /* NBODY_EVENT_* */
  uint64_t request_id;
  int64_t step;               /* steps completed when the event was made */
  int64_t n;
  double t_submit, t_admit, t_event;  /* seconds, server clock */
  char message[128];
} nbody_event;

int nbody_server_create(const nbody_server_config *cfg, nbody_server **out);
/* Copies the request's arrays; returns its id.  NBODY_E_NOSPACE if the
 * request alone can never fit the arena. */
int nbody_server_submit(nbody_server *s, const nbody_request *req,
                        uint64_t *request_id);
/* Next event, waiting up to timeout_ms (<0: forever; NBODY_E_TIMEOUT when
 * none arrived).  For SNAPSHOT and DONE events, positions (and for DONE
 * velocities) are copied into the non-null arrays, which hold `capacity`
 * values each.  If capacity < ev->n the event is left queued, ev->n tells
 * the size needed, and NBODY_E_NOSPACE is returned. */
int nbody_server_poll(nbody_server *s, int timeout_ms, nbody_event *ev,
                      double *pos_x, double *pos_y, double *pos_z,
                      double *vel_x, double *vel_y, double *vel_z,
                      int64_t capacity);
/* Cancels a queued request (its CANCELLED event is queued at once) or one
 * being admitted or resident (retired at the next epoch end with a
 * CANCELLED event, unless its DONE was already queued).  NBODY_E_NOTFOUND
 * for an id that is not queued, being admitted or resident, including a
 * request already retired after its DONE. */
int nbody_server_cancel(nbody_server *s, uint64_t request_id);
/* Counters: out[0] submitted, [1] admitted, [2] completed, [3] resident,
 * [4] queued, [5] shared steps run, [6] pair interactions evaluated,
 * [7] failed. */
int nbody_server_stats(nbody_server *s, double out[8]);
int nbody_server_destroy(nbody_server *s);

/* ---- snapshot files: replaces SimulationDataWriter
 * (src/output.rs:73-139).  No GPU involved.  The format follows the path's
 * extension, compared case-insensitively: ".csv" writes the header
 * "time,id,x,y,z" with the first record, then one "\n"-ended line per body
 * per snapshot, floats in shortest round-trip form in the layout of the
 * reference's csv writer (src/output.rs:106-122); ".nbody" writes the magic
 * "NBODY001" and 40-byte little-endian records f64 time, u64 id, f64 x,
 * f64 y, f64 z (src/output.rs:124-138).  Other extensions return
 * NBODY_E_INVALID (the reference's "Unsupported output file format"). */
typedef struct nbody_writer nbody_writer;
int nbody_writer_open(const char *path, nbody_writer **out);
/* One snapshot: records (time, id, x[id], y[id], z[id]) for id 0..n-1,
 * the SimulationData::extend_from_snapshot order (src/output.rs:46-54). */
int nbody_writer_append(nbody_writer *w, double time, int64_t n,
                        const double *pos_x, const double *pos_y,
                        const double *pos_z);
/* Flushes, closes and frees (also on error). */
int nbody_writer_close(nbody_writer *w);
/* Text of one float as the snapshot CSV holds it (buf: 32 bytes, NUL-ended). */
int nbody_format_f64(double v, char buf[32]);
/* Parses a NUL-ended field with Rust's f64::from_str grammar, which the
 * reference applies to CSV fields and numeric flags: optional sign, then
 * inf / infinity / nan (any case) or decimal digits with an optional '.' and
 * exponent; nothing else, no spaces.  NBODY_E_INVALID when it is not one. */
int nbody_parse_f64(const char *text, double *out);

/* ---- initial conditions: replaces input::load_bodies (src/input.rs:20-33).
 * CSV with a header naming mass,pos_x,pos_y,pos_z,vel_x,vel_y,vel_z in any
 * order (other columns ignored).  With any array null, only *n is set to the
 * number of bodies.  Otherwise *n is the arrays' capacity on entry and the
 * count on return (NBODY_E_NOSPACE, with *n the count, when they are too
 * small).  NBODY_E_NOTFOUND when the file cannot be opened; NBODY_E_INVALID
 * with the line in nbody_last_error() for a malformed file. */
int nbody_load_initial_conditions(const char *path, int64_t *n,
                                  double *masses, double *pos_x,
                                  double *pos_y, double *pos_z,
                                  double *vel_x, double *vel_y,
                                  double *vel_z);

#ifdef __cplusplus
}
#endif

#endif /* NBODY_AMD_H */
