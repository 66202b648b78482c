/*
 * oracle_bh.c -- CPU restatement of the reference's Barnes-Hut gravity
 * (src/gravity/cpu/barnes_hut.rs, src/gravity/cpu/barnes_hut/octree.rs).
 *
 * TEST INFRASTRUCTURE ONLY, like oracle.c: only tests/ and the benchmark's
 * CPU leg use it, as the checker.
 *
 * Same arithmetic contract as oracle.c (binary64, no contraction, correctly
 * rounded sqrt and division, the reference's expression trees), and the same
 * structure as the reference: bounding box with 1.001 padding, 21-bit
 * quantisation, Morton codes sorted with their body index, a recursive build
 * that creates nodes in preorder (one body per leaf, 21 levels), sequential
 * centre-of-mass sums, and a per-target stack traversal that pushes children
 * in order and pops the last first.  Pinned by the reference's own octree
 * tests (octree.rs:581-699), by theta = 0 against the direct sum, and by
 * structural invariants (tests/test_oracle_bh.py).
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

enum { BH_MAX_DEPTH = 21 };
static const double BH_PADDING = 1.001;  /* octree.rs DEFAULT_PADDING */

typedef struct {
  int64_t n, cap, count;
  /* bodies in Morton order */
  double *m, *x, *y, *z;
  uint64_t *codes;
  int64_t *orig_to_morton, *morton_to_orig;
  /* nodes */
  double *mass, *cx, *cy, *cz, *width;
  int64_t *start, *nb;
  int32_t *nchild;
  int64_t *child; /* 8 per node */
} bh_tree;

/* octree.rs:434-468 find_bounding_box */
static void bounding_box(int64_t n, const double *rx, const double *ry,
                         const double *rz, double *mnx, double *mny,
                         double *mnz, double *width) {
  double min_x = rx[0], max_x = rx[0], min_y = ry[0], max_y = ry[0];
  double min_z = rz[0], max_z = rz[0];
  for (int64_t i = 1; i < n; ++i) {
    if (rx[i] < min_x) min_x = rx[i];
    if (rx[i] > max_x) max_x = rx[i];
    if (ry[i] < min_y) min_y = ry[i];
    if (ry[i] > max_y) max_y = ry[i];
    if (rz[i] < min_z) min_z = rz[i];
    if (rz[i] > max_z) max_z = rz[i];
  }
  double dx = max_x - min_x, dy = max_y - min_y, dz = max_z - min_z;
  /* f64::max ignores a NaN operand */
  double w = isnan(dx) ? dy : (isnan(dy) ? dx : (dx > dy ? dx : dy));
  w = isnan(w) ? dz : (isnan(dz) ? w : (w > dz ? w : dz));
  *mnx = min_x;
  *mny = min_y;
  *mnz = min_z;
  *width = w * BH_PADDING;
}

/* octree.rs:470-474 quantize_f64_to_u64; `as u64` saturates (NaN and
 * negatives to 0). */
static uint64_t quantize(double v, double mn, double width) {
  double q = ((v - mn) / width) * (double)(1 << 21);
  if (!(q >= 0.0)) return 0;
  if (q >= 18446744073709551616.0) return UINT64_MAX;
  return (uint64_t)q;
}

/* octree.rs:556-579 morton_encode / morton_spread_bits */
static uint64_t spread(uint64_t x) {
  x &= 0x1fffff;
  x = (x | (x << 32)) & 0x1f00000000ffffULL;
  x = (x | (x << 16)) & 0x1f0000ff0000ffULL;
  x = (x | (x << 8)) & 0x100f00f00f00f00fULL;
  x = (x | (x << 4)) & 0x10c30c30c30c30c3ULL;
  x = (x | (x << 2)) & 0x1249249249249249ULL;
  return x;
}

uint64_t oracle_bh_morton(uint64_t x, uint64_t y, uint64_t z) {
  return spread(x) | (spread(y) << 1) | (spread(z) << 2);
}

typedef struct {
  uint64_t code;
  int64_t idx;
} code_idx;

static int cmp_code_idx(const void *a, const void *b) {
  const code_idx *p = (const code_idx *)a, *q = (const code_idx *)b;
  if (p->code != q->code) return p->code < q->code ? -1 : 1;
  return p->idx < q->idx ? -1 : (p->idx > q->idx);
}

/* octree.rs:381-409 compute_com_of_bodies_range */
static void com_of_bodies(const bh_tree *t, int64_t s, int64_t e, double *M,
                          double *X, double *Y, double *Z) {
  double tm = 0.0, cx = 0.0, cy = 0.0, cz = 0.0;
  for (int64_t i = s; i < e; ++i) {
    tm += t->m[i];
    cx += t->m[i] * t->x[i];
    cy += t->m[i] * t->y[i];
    cz += t->m[i] * t->z[i];
  }
  if (tm > 0.0) {
    cx /= tm;
    cy /= tm;
    cz /= tm;
  }
  *M = tm;
  *X = cx;
  *Y = cy;
  *Z = cz;
}

/* octree.rs:411-432 compute_com_of_node_indices */
static void com_of_children(const bh_tree *t, int64_t node, double *M,
                            double *X, double *Y, double *Z) {
  double tm = 0.0, cx = 0.0, cy = 0.0, cz = 0.0;
  for (int c = 0; c < t->nchild[node]; ++c) {
    int64_t i = t->child[node * 8 + c];
    tm += t->mass[i];
    cx += t->mass[i] * t->cx[i];
    cy += t->mass[i] * t->cy[i];
    cz += t->mass[i] * t->cz[i];
  }
  if (tm > 0.0) {
    cx /= tm;
    cy /= tm;
    cz /= tm;
  }
  *M = tm;
  *X = cx;
  *Y = cy;
  *Z = cz;
}

/* octree.rs:257-352 build_recursive (max_leaf_size 1). */
static int64_t build_rec(bh_tree *t, int64_t s, int64_t e, int level,
                         double width) {
  int64_t node = t->count++;
  t->start[node] = s;
  t->nb[node] = e - s;
  t->width[node] = width;
  t->nchild[node] = 0;
  if (e - s <= 1 || level <= 0) {
    com_of_bodies(t, s, e, &t->mass[node], &t->cx[node], &t->cy[node],
                  &t->cz[node]);
    return node;
  }
  const unsigned shift = (unsigned)((level - 1) * 3);
  int64_t cur = s;
  while (cur < e) {
    const uint64_t id = (t->codes[cur] >> shift) & 7;
    int64_t end = cur;
    while (end < e && ((t->codes[end] >> shift) & 7) == id) ++end;
    if (end > cur) {
      int64_t c = build_rec(t, cur, end, level - 1, width * 0.5);
      t->child[node * 8 + t->nchild[node]++] = c;
    }
    cur = end;
  }
  if (t->nchild[node] > 0)
    com_of_children(t, node, &t->mass[node], &t->cx[node], &t->cy[node],
                    &t->cz[node]);
  else
    t->mass[node] = t->cx[node] = t->cy[node] = t->cz[node] = 0.0;
  return node;
}

static void tree_free(bh_tree *t) {
  free(t->m);
  free(t->x);
  free(t->y);
  free(t->z);
  free(t->codes);
  free(t->orig_to_morton);
  free(t->morton_to_orig);
  free(t->mass);
  free(t->cx);
  free(t->cy);
  free(t->cz);
  free(t->width);
  free(t->start);
  free(t->nb);
  free(t->nchild);
  free(t->child);
}

/* octree.rs:129-157 build.  Returns 0, or -1 on allocation failure. */
static int tree_build(bh_tree *t, int64_t n, const double *m,
                      const double *rx, const double *ry, const double *rz) {
  memset(t, 0, sizeof(*t));
  t->n = n;
  t->cap = n * (1 + BH_MAX_DEPTH) + 1; /* octree.rs:114 with max_leaf_size 1 */
  t->m = malloc(sizeof(double) * n);
  t->x = malloc(sizeof(double) * n);
  t->y = malloc(sizeof(double) * n);
  t->z = malloc(sizeof(double) * n);
  t->codes = malloc(sizeof(uint64_t) * n);
  t->orig_to_morton = malloc(sizeof(int64_t) * n);
  t->morton_to_orig = malloc(sizeof(int64_t) * n);
  t->mass = malloc(sizeof(double) * t->cap);
  t->cx = malloc(sizeof(double) * t->cap);
  t->cy = malloc(sizeof(double) * t->cap);
  t->cz = malloc(sizeof(double) * t->cap);
  t->width = malloc(sizeof(double) * t->cap);
  t->start = malloc(sizeof(int64_t) * t->cap);
  t->nb = malloc(sizeof(int64_t) * t->cap);
  t->nchild = malloc(sizeof(int32_t) * t->cap);
  t->child = malloc(sizeof(int64_t) * 8 * t->cap);
  code_idx *ci = malloc(sizeof(code_idx) * n);
  if (!t->m || !t->x || !t->y || !t->z || !t->codes || !t->orig_to_morton ||
      !t->morton_to_orig || !t->mass || !t->cx || !t->cy || !t->cz ||
      !t->width || !t->start || !t->nb || !t->nchild || !t->child || !ci) {
    free(ci);
    tree_free(t);
    return -1;
  }
  double mnx, mny, mnz, w;
  bounding_box(n, rx, ry, rz, &mnx, &mny, &mnz, &w);
  /* octree.rs:486-498: codes with their index, sorted as (code, index) */
  for (int64_t i = 0; i < n; ++i) {
    ci[i].code = oracle_bh_morton(quantize(rx[i], mnx, w),
                                  quantize(ry[i], mny, w),
                                  quantize(rz[i], mnz, w));
    ci[i].idx = i;
  }
  qsort(ci, (size_t)n, sizeof(code_idx), cmp_code_idx);
  for (int64_t k = 0; k < n; ++k) {
    const int64_t i = ci[k].idx;
    t->codes[k] = ci[k].code;
    t->morton_to_orig[k] = i;
    t->orig_to_morton[i] = k;
    t->m[k] = m[i];
    t->x[k] = rx[i];
    t->y[k] = ry[i];
    t->z[k] = rz[i];
  }
  free(ci);
  build_rec(t, 0, n, BH_MAX_DEPTH, w);
  return 0;
}

/* octree.rs:159-254 compute_acceleration_for_body; the pair term is
 * newton.rs:128-141 with eps2 = softening^2 (barnes_hut.rs:38). */
static void bh_accel_for_body(const bh_tree *t, int64_t i, double g,
                              double eps2, double theta, double *oax,
                              double *oay, double *oaz) {
  double ax = 0.0, ay = 0.0, az = 0.0;
  const int64_t ti = t->orig_to_morton[i];
  const double tx = t->x[ti], ty = t->y[ti], tz = t->z[ti];
  int64_t stack[8 * (BH_MAX_DEPTH + 2)];
  int sp = 0;
  stack[sp++] = 0;
  while (sp > 0) {
    const int64_t node = stack[--sp];
    const double ndx = t->cx[node] - tx;
    const double ndy = t->cy[node] - ty;
    const double ndz = t->cz[node] - tz;
    const double d2 = ((ndx * ndx) + (ndy * ndy)) + (ndz * ndz);
    if (d2 == 0.0) continue;
    if (t->nchild[node] == 0) {
      const int64_t s = t->start[node], e = s + t->nb[node];
      for (int64_t k = s; k < e; ++k) {
        if (k == ti) continue;
        const double dx = t->x[k] - tx, dy = t->y[k] - ty, dz = t->z[k] - tz;
        const double r2 = ((dx * dx) + (dy * dy)) + (dz * dz);
        const double inv = 1.0 / sqrt(r2 + eps2);
        const double inv3 = (inv * inv) * inv;
        const double kk = (g * t->m[k]) * inv3;
        ax += kk * dx;
        ay += kk * dy;
        az += kk * dz;
      }
    } else if (((t->width[node] * t->width[node]) / d2) < (theta * theta)) {
      const double r2 = ((ndx * ndx) + (ndy * ndy)) + (ndz * ndz);
      const double inv = 1.0 / sqrt(r2 + eps2);
      const double inv3 = (inv * inv) * inv;
      const double kk = (g * t->mass[node]) * inv3;
      ax += kk * ndx;
      ay += kk * ndy;
      az += kk * ndz;
    } else {
      for (int c = 0; c < t->nchild[node]; ++c)
        stack[sp++] = t->child[node * 8 + c];
    }
  }
  *oax = ax;
  *oay = ay;
  *oaz = az;
}

/* Tree of one build, exported for the structural tests.  Arrays hold at
 * least n*22+1 entries (child: 8 per node).  Returns the node count, or -1. */
int64_t oracle_bh_tree(int64_t n, const double *m, const double *rx,
                       const double *ry, const double *rz, double *mass,
                       double *cx, double *cy, double *cz, double *width,
                       int64_t *start, int64_t *count, int32_t *nchild,
                       int64_t *child, int64_t *morton_to_orig,
                       uint64_t *codes) {
  if (n <= 0) return -1;
  bh_tree t;
  if (tree_build(&t, n, m, rx, ry, rz) != 0) return -1;
  const int64_t c = t.count;
  memcpy(mass, t.mass, sizeof(double) * c);
  memcpy(cx, t.cx, sizeof(double) * c);
  memcpy(cy, t.cy, sizeof(double) * c);
  memcpy(cz, t.cz, sizeof(double) * c);
  memcpy(width, t.width, sizeof(double) * c);
  memcpy(start, t.start, sizeof(int64_t) * c);
  memcpy(count, t.nb, sizeof(int64_t) * c);
  memcpy(nchild, t.nchild, sizeof(int32_t) * c);
  memcpy(child, t.child, sizeof(int64_t) * 8 * c);
  memcpy(morton_to_orig, t.morton_to_orig, sizeof(int64_t) * n);
  memcpy(codes, t.codes, sizeof(uint64_t) * n);
  tree_free(&t);
  return c;
}

typedef struct {
  const bh_tree *t;
  double g, eps2, theta;
  int64_t lo, hi;
  double *ax, *ay, *az;
} bh_job;

static void *bh_worker(void *p) {
  bh_job *j = (bh_job *)p;
  for (int64_t i = j->lo; i < j->hi; ++i)
    bh_accel_for_body(j->t, i, j->g, j->eps2, j->theta, &j->ax[i], &j->ay[i],
                      &j->az[i]);
  return NULL;
}

typedef struct {
  const bh_tree *t;
  double g, eps2, theta;
  const int64_t *idx;
  int64_t lo, hi;
  double *ax, *ay, *az;
} bh_tjob;

static void *bh_tworker(void *p) {
  bh_tjob *j = (bh_tjob *)p;
  for (int64_t q = j->lo; q < j->hi; ++q)
    bh_accel_for_body(j->t, j->idx[q], j->g, j->eps2, j->theta, &j->ax[q],
                      &j->ay[q], &j->az[q]);
  return NULL;
}

/* The same evaluation for the k targets idx[0..k) only (one full tree
 * build, then those targets): a sampled check of large systems.  Output q
 * is target idx[q].  Returns 0 or -1. */
int oracle_barnes_hut_targets(int64_t n, double g, double softening,
                              double theta, const double *m,
                              const double *rx, const double *ry,
                              const double *rz, const int64_t *idx,
                              int64_t k, double *ax, double *ay, double *az,
                              int nthreads) {
  if (n <= 0 || k < 0) return -1;
  for (int64_t q = 0; q < k; ++q)
    if (idx[q] < 0 || idx[q] >= n) return -1;
  bh_tree t;
  if (tree_build(&t, n, m, rx, ry, rz) != 0) return -1;
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  if (k > 0 && nthreads > k) nthreads = (int)k;
  pthread_t th[256];
  bh_tjob jobs[256];
  for (int w = 0; w < nthreads; ++w)
    jobs[w] = (bh_tjob){&t, g, softening * softening, theta, idx,
                        k * w / nthreads, k * (w + 1) / nthreads, ax, ay, az};
  for (int w = 1; w < nthreads; ++w)
    pthread_create(&th[w], NULL, bh_tworker, &jobs[w]);
  bh_tworker(&jobs[0]);
  for (int w = 1; w < nthreads; ++w) pthread_join(th[w], NULL);
  tree_free(&t);
  return 0;
}

/* barnes_hut.rs:25-51 BarnesHutGravity::calculate_accelerations: build,
 * then every target (rayon over i; per-target arithmetic is independent of
 * the split, so any thread count gives the same bits).  Returns 0 or -1. */
int oracle_barnes_hut(int64_t n, double g, double softening, double theta,
                      const double *m, const double *rx, const double *ry,
                      const double *rz, double *ax, double *ay, double *az,
                      int nthreads) {
  if (n <= 0) return n == 0 ? 0 : -1;
  bh_tree t;
  if (tree_build(&t, n, m, rx, ry, rz) != 0) return -1;
  const double eps2 = softening * softening;
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  if (nthreads > n) nthreads = (int)n;
  pthread_t th[256];
  bh_job jobs[256];
  for (int k = 0; k < nthreads; ++k) {
    jobs[k] = (bh_job){&t, g, eps2, theta, n * k / nthreads,
                       n * (k + 1) / nthreads, ax, ay, az};
  }
  for (int k = 1; k < nthreads; ++k)
    pthread_create(&th[k], NULL, bh_worker, &jobs[k]);
  bh_worker(&jobs[0]);
  for (int k = 1; k < nthreads; ++k) pthread_join(th[k], NULL);
  tree_free(&t);
  return 0;
}
