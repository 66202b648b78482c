/*
 * oracle.c -- CPU restatement of the reference's f64 direct-sum gravity and
 * its three integrators.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links, loads or calls
 * this file; only tests/, the smoke check and the CPU-baseline leg of the
 * benchmark do, and there only as the checker (or as the timed CPU baseline).
 *
 * Arithmetic contract: IEEE-754 binary64, no contraction (built with
 * -ffp-contract=off, no -ffast-math), correctly rounded sqrt and division,
 * and the reference's left-to-right expression trees.  Rust neither
 * contracts nor reassociates, so every function below is bit-identical to the
 * Rust source line it cites.
 *
 * Every function is pinned by tests/test_oracle.py (closed forms, brute force
 * on tiny inputs, invariants, an independent numpy evaluation, convergence
 * orders) -- see the header comment of that file for which check pins what.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* src/gravity/cpu/newton.rs:128-141 (compute_acceleration) */
static inline void pair_term(double g, double eps2, double m_j, double dx,
                             double dy, double dz, double *ox, double *oy,
                             double *oz) {
  double r2 = ((dx * dx) + (dy * dy)) + (dz * dz);
  double inv = 1.0 / sqrt(r2 + eps2);
  double inv3 = (inv * inv) * inv;
  double k = (g * m_j) * inv3;
  *ox = k * dx;
  *oy = k * dy;
  *oz = k * dz;
}

/* src/gravity/cpu/newton.rs:91-126 (compute_acceleration_for_body):
 * sequential j = 0..n-1, j == i skipped, one accumulator per component. */
static void accel_for_body(int64_t i, int64_t n, double g, double eps2,
                           const double *m, const double *rx,
                           const double *ry, const double *rz, double *ax,
                           double *ay, double *az) {
  const double xi = rx[i], yi = ry[i], zi = rz[i];
  double sx = 0.0, sy = 0.0, sz = 0.0;
  for (int64_t j = 0; j < n; ++j) {
    if (j == i) continue;
    double tx, ty, tz;
    pair_term(g, eps2, m[j], rx[j] - xi, ry[j] - yi, rz[j] - zi, &tx, &ty,
              &tz);
    sx += tx;
    sy += ty;
    sz += tz;
  }
  *ax = sx;
  *ay = sy;
  *az = sz;
}

/* src/gravity/cpu/newton.rs:69-87 (NewtonGravity::calculate_accelerations);
 * eps2 = softening * softening as at newton.rs:82. */
void oracle_newton(int64_t n, double g, double softening, const double *m,
                   const double *rx, const double *ry, const double *rz,
                   double *ax, double *ay, double *az) {
  const double eps2 = softening * softening;
  for (int64_t i = 0; i < n; ++i)
    accel_for_body(i, n, g, eps2, m, rx, ry, rz, &ax[i], &ay[i], &az[i]);
}

/* ---- newton-parallel: the same per-target arithmetic spread over threads.
 * src/gravity/cpu/newton_parallel.rs:31-53 distributes i with rayon; the
 * per-i arithmetic is identical, so the result is bit-identical to
 * oracle_newton whatever the thread count or chunking. */
typedef struct {
  int64_t n, lo, hi;
  double g, eps2;
  const double *m, *rx, *ry, *rz;
  const int64_t *idx; /* optional target list (NULL: targets lo..hi) */
  double *ax, *ay, *az;
} par_job;

static void *par_worker(void *arg) {
  par_job *jb = (par_job *)arg;
  for (int64_t t = jb->lo; t < jb->hi; ++t) {
    int64_t i = jb->idx ? jb->idx[t] : t;
    accel_for_body(i, jb->n, jb->g, jb->eps2, jb->m, jb->rx, jb->ry, jb->rz,
                   &jb->ax[t], &jb->ay[t], &jb->az[t]);
  }
  return NULL;
}

static void run_parallel(int64_t n, double g, double eps2, const double *m,
                         const double *rx, const double *ry, const double *rz,
                         const int64_t *idx, int64_t ntargets, double *ax,
                         double *ay, double *az, int nthreads) {
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  if ((int64_t)nthreads > ntargets) nthreads = ntargets > 0 ? (int)ntargets : 1;
  pthread_t th[256];
  par_job jobs[256];
  /* static contiguous chunks over the targets; every target costs the same
   * (n - 1 pair terms) and the partition never changes results. */
  for (int t = 0; t < nthreads; ++t) {
    int64_t lo = ntargets * t / nthreads, hi = ntargets * (t + 1) / nthreads;
    jobs[t] = (par_job){n, lo, hi, g, eps2, m, rx, ry, rz, idx, ax, ay, az};
  }
  for (int t = 1; t < nthreads; ++t)
    pthread_create(&th[t], NULL, par_worker, &jobs[t]);
  par_worker(&jobs[0]);
  for (int t = 1; t < nthreads; ++t) pthread_join(th[t], NULL);
}

/* src/gravity/cpu/newton_parallel.rs:31-53 */
void oracle_newton_parallel(int64_t n, double g, double softening,
                            const double *m, const double *rx,
                            const double *ry, const double *rz, double *ax,
                            double *ay, double *az, int nthreads) {
  run_parallel(n, g, softening * softening, m, rx, ry, rz, NULL, n, ax, ay,
               az, nthreads);
}

/* The same per-target arithmetic for a chosen subset of targets (used as a
 * sampled-target oracle where the full N^2 sum is too slow on the CPU).
 * Output t holds target idx[t]. */
void oracle_newton_targets(int64_t n, double g, double softening,
                           const double *m, const double *rx,
                           const double *ry, const double *rz,
                           const int64_t *idx, int64_t ntargets, double *ax,
                           double *ay, double *az, int nthreads) {
  run_parallel(n, g, softening * softening, m, rx, ry, rz, idx, ntargets, ax,
               ay, az, nthreads);
}

/* ---- integrators.  The force evaluation is newton (nthreads <= 1) or
 * newton-parallel (both give identical bits), or barnes-hut (oracle_bh.c). */
typedef struct {
  double g, softening;
  int nthreads;
  int gravity; /* src/args.rs:84-89 order: 0/1 direct sum, 2 barnes-hut */
  double theta;
} grav_cfg;

int oracle_barnes_hut(int64_t n, double g, double softening, double theta,
                      const double *m, const double *rx, const double *ry,
                      const double *rz, double *ax, double *ay, double *az,
                      int nthreads);

/* src/integrators/cpu/integrator.rs:29-44 (compute_acceleration): zero, then
 * evaluate the gravity method. */
static void eval_acc(const grav_cfg *c, int64_t n, const double *m,
                     const double *rx, const double *ry, const double *rz,
                     double *ax, double *ay, double *az) {
  memset(ax, 0, sizeof(double) * n);
  memset(ay, 0, sizeof(double) * n);
  memset(az, 0, sizeof(double) * n);
  if (c->gravity == 2)
    oracle_barnes_hut(n, c->g, c->softening, c->theta, m, rx, ry, rz, ax, ay,
                      az, c->nthreads);
  else if (c->nthreads <= 1)
    oracle_newton(n, c->g, c->softening, m, rx, ry, rz, ax, ay, az);
  else
    oracle_newton_parallel(n, c->g, c->softening, m, rx, ry, rz, ax, ay, az,
                           c->nthreads);
}

/* src/integrators/cpu/euler.rs:57-82: a = F(r); v += a*dt; r += v*dt. */
static void euler_step(const grav_cfg *c, int64_t n, double dt,
                       const double *m, double *rx, double *ry, double *rz,
                       double *vx, double *vy, double *vz, double *w) {
  double *ax = w, *ay = w + n, *az = w + 2 * n;
  eval_acc(c, n, m, rx, ry, rz, ax, ay, az);
  for (int64_t i = 0; i < n; ++i) {
    vx[i] += ax[i] * dt;
    vy[i] += ay[i] * dt;
    vz[i] += az[i] * dt;
  }
  for (int64_t i = 0; i < n; ++i) {
    rx[i] += vx[i] * dt;
    ry[i] += vy[i] * dt;
    rz[i] += vz[i] * dt;
  }
}

/* src/integrators/cpu/velocity_verlet.rs:63-104:
 * a_n = F(r); r += (v*dt) + ((0.5*a_n)*dt2); a_n1 = F(r);
 * v += (0.5*(a_n + a_n1))*dt. */
static void vv_step(const grav_cfg *c, int64_t n, double dt, const double *m,
                    double *rx, double *ry, double *rz, double *vx,
                    double *vy, double *vz, double *w) {
  double *ax = w, *ay = w + n, *az = w + 2 * n;
  double *bx = w + 3 * n, *by = w + 4 * n, *bz = w + 5 * n;
  const double dt2 = dt * dt;
  eval_acc(c, n, m, rx, ry, rz, ax, ay, az);
  for (int64_t i = 0; i < n; ++i) {
    rx[i] += (vx[i] * dt) + ((0.5 * ax[i]) * dt2);
    ry[i] += (vy[i] * dt) + ((0.5 * ay[i]) * dt2);
    rz[i] += (vz[i] * dt) + ((0.5 * az[i]) * dt2);
  }
  eval_acc(c, n, m, rx, ry, rz, bx, by, bz);
  for (int64_t i = 0; i < n; ++i) {
    vx[i] += (0.5 * (ax[i] + bx[i])) * dt;
    vy[i] += (0.5 * (ay[i] + by[i])) * dt;
    vz[i] += (0.5 * (az[i] + bz[i])) * dt;
  }
}

/* src/integrators/cpu/runge_kutta.rs:149-281.  k1_r = v; k1_v = F(r);
 * k2_r = v + (dt/2)*k1_v; k2_v = F(r + (dt/2)*k1_r); k3_r = v + (dt/2)*k2_v;
 * k3_v = F(r + (dt/2)*k2_r); k4_r = v + dt*k3_v; k4_v = F(r + dt*k3_r);
 * r += (dt/6)*(((k1_r + 2*k2_r) + 2*k3_r) + k4_r), same for v. */
static void rk4_step(const grav_cfg *c, int64_t n, double dt, const double *m,
                     double *rx, double *ry, double *rz, double *vx,
                     double *vy, double *vz, double *w) {
  /* k{1..4}_{r,v}{x,y,z} and temp positions: 27 arrays of n */
  double *k[4][6];
  for (int s = 0; s < 4; ++s)
    for (int q = 0; q < 6; ++q) k[s][q] = w + (int64_t)(s * 6 + q) * n;
  double *tx = w + 24 * n, *ty = w + 25 * n, *tz = w + 26 * n;
  const double h = dt / 2.0;
  /* 1. k1_r = v_n */
  memcpy(k[0][0], vx, sizeof(double) * n);
  memcpy(k[0][1], vy, sizeof(double) * n);
  memcpy(k[0][2], vz, sizeof(double) * n);
  /* 2. k1_v = F(r_n) */
  eval_acc(c, n, m, rx, ry, rz, k[0][3], k[0][4], k[0][5]);
  /* 3. k2_r = v_n + (dt/2)*k1_v */
  for (int64_t i = 0; i < n; ++i) {
    k[1][0][i] = vx[i] + h * k[0][3][i];
    k[1][1][i] = vy[i] + h * k[0][4][i];
    k[1][2][i] = vz[i] + h * k[0][5][i];
  }
  /* 4. k2_v = F(r_n + (dt/2)*k1_r) */
  for (int64_t i = 0; i < n; ++i) {
    tx[i] = rx[i] + h * k[0][0][i];
    ty[i] = ry[i] + h * k[0][1][i];
    tz[i] = rz[i] + h * k[0][2][i];
  }
  eval_acc(c, n, m, tx, ty, tz, k[1][3], k[1][4], k[1][5]);
  /* 5. k3_r = v_n + (dt/2)*k2_v */
  for (int64_t i = 0; i < n; ++i) {
    k[2][0][i] = vx[i] + h * k[1][3][i];
    k[2][1][i] = vy[i] + h * k[1][4][i];
    k[2][2][i] = vz[i] + h * k[1][5][i];
  }
  /* 6. k3_v = F(r_n + (dt/2)*k2_r) */
  for (int64_t i = 0; i < n; ++i) {
    tx[i] = rx[i] + h * k[1][0][i];
    ty[i] = ry[i] + h * k[1][1][i];
    tz[i] = rz[i] + h * k[1][2][i];
  }
  eval_acc(c, n, m, tx, ty, tz, k[2][3], k[2][4], k[2][5]);
  /* 7. k4_r = v_n + dt*k3_v */
  for (int64_t i = 0; i < n; ++i) {
    k[3][0][i] = vx[i] + dt * k[2][3][i];
    k[3][1][i] = vy[i] + dt * k[2][4][i];
    k[3][2][i] = vz[i] + dt * k[2][5][i];
  }
  /* 8. k4_v = F(r_n + dt*k3_r) */
  for (int64_t i = 0; i < n; ++i) {
    tx[i] = rx[i] + dt * k[2][0][i];
    ty[i] = ry[i] + dt * k[2][1][i];
    tz[i] = rz[i] + dt * k[2][2][i];
  }
  eval_acc(c, n, m, tx, ty, tz, k[3][3], k[3][4], k[3][5]);
  /* 9./10. combine */
  const double s6 = dt / 6.0;
  double *r[3] = {rx, ry, rz}, *v[3] = {vx, vy, vz};
  for (int q = 0; q < 3; ++q)
    for (int64_t i = 0; i < n; ++i)
      r[q][i] += s6 * (((k[0][q][i] + (2.0 * k[1][q][i])) +
                        (2.0 * k[2][q][i])) +
                       k[3][q][i]);
  for (int q = 0; q < 3; ++q)
    for (int64_t i = 0; i < n; ++i)
      v[q][i] += s6 * (((k[0][3 + q][i] + (2.0 * k[1][3 + q][i])) +
                        (2.0 * k[2][3 + q][i])) +
                       k[3][3 + q][i]);
}

/* Integrator ids follow the reference's IntegratorMethod order
 * (src/args.rs:140-144): 0 euler, 1 velocity-verlet, 2 runge-kutta.
 * Advances the bodies in place by nsteps steps.  Returns 0, or -1 on a bad
 * integrator id or allocation failure. */
int oracle_integrate_gravity(int integrator, int gravity, double theta,
                             int64_t n, double g, double softening, double dt,
                             int64_t nsteps, const double *m, double *rx,
                             double *ry, double *rz, double *vx, double *vy,
                             double *vz, int nthreads) {
  if (integrator < 0 || integrator > 2 || n < 0 || nsteps < 0) return -1;
  if (gravity < 0 || gravity > 2) return -1;
  grav_cfg c = {g, softening, nthreads, gravity, theta};
  double *w = (double *)malloc(sizeof(double) * 27 * (size_t)(n > 0 ? n : 1));
  if (!w) return -1;
  for (int64_t s = 0; s < nsteps; ++s) {
    if (integrator == 0)
      euler_step(&c, n, dt, m, rx, ry, rz, vx, vy, vz, w);
    else if (integrator == 1)
      vv_step(&c, n, dt, m, rx, ry, rz, vx, vy, vz, w);
    else
      rk4_step(&c, n, dt, m, rx, ry, rz, vx, vy, vz, w);
  }
  free(w);
  return 0;
}

int oracle_integrate(int integrator, int64_t n, double g, double softening,
                     double dt, int64_t nsteps, const double *m, double *rx,
                     double *ry, double *rz, double *vx, double *vy,
                     double *vz, int nthreads) {
  return oracle_integrate_gravity(integrator, 1, 0.0, n, g, softening, dt,
                                  nsteps, m, rx, ry, rz, vx, vy, vz, nthreads);
}

/* The run loop of SimulationRunner::run (src/simulation_runner.rs:28-52):
 * one reference step at a time, positions copied after every step into
 * out_x/out_y/out_z[s*n + i] (step s = 0 .. nsteps-1). */
int oracle_integrate_record(int integrator, int gravity, double theta,
                            int64_t n, double g, double softening, double dt,
                            int64_t nsteps, const double *m, double *rx,
                            double *ry, double *rz, double *vx, double *vy,
                            double *vz, double *out_x, double *out_y,
                            double *out_z, int nthreads) {
  for (int64_t s = 0; s < nsteps; ++s) {
    int rc = oracle_integrate_gravity(integrator, gravity, theta, n, g,
                                      softening, dt, 1, m, rx, ry, rz, vx, vy,
                                      vz, nthreads);
    if (rc != 0) return rc;
    for (int64_t i = 0; i < n; ++i) {
      out_x[s * n + i] = rx[i];
      out_y[s * n + i] = ry[i];
      out_z[s * n + i] = rz[i];
    }
  }
  return 0;
}

/* Softened total energy, not present in the reference (it has no energy
 * diagnostics): K = sum_i 0.5*m_i*|v_i|^2,
 * U = -g * sum_{i<j} m_i*m_j / sqrt(|r_ij|^2 + softening^2),
 * summed in plain index order.  out[0] = K, out[1] = U.  Pinned by the
 * two-body closed form and by a numpy evaluation in the tests. */
void oracle_energy(int64_t n, double g, double softening, const double *m,
                   const double *rx, const double *ry, const double *rz,
                   const double *vx, const double *vy, const double *vz,
                   double *out) {
  const double eps2 = softening * softening;
  double kin = 0.0, pot = 0.0;
  for (int64_t i = 0; i < n; ++i)
    kin += 0.5 * m[i] * ((vx[i] * vx[i] + vy[i] * vy[i]) + vz[i] * vz[i]);
  for (int64_t i = 0; i < n; ++i) {
    double row = 0.0;
    for (int64_t j = i + 1; j < n; ++j) {
      double dx = rx[j] - rx[i], dy = ry[j] - ry[i], dz = rz[j] - rz[i];
      row += m[j] / sqrt(((dx * dx + dy * dy) + dz * dz) + eps2);
    }
    pot += m[i] * row;
  }
  out[0] = kin;
  out[1] = -g * pot;
}
