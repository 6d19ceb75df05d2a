/*
 * d3oracle.c -- FP64 restatement of d3-force-3d's CPU physics.  TEST INFRASTRUCTURE ONLY
 * (the checker of the GPU backend; never shipped in the product path).
 *
 * Follows, function by function:
 *   lcg.js:1-9, jiggle.js:1-3                       -> or_lcg, jiggle()
 *   simulation.js:49-77 (tick), :79-107 (init)       -> or_sim_tick, or_integrate, or_phyllotaxis
 *   manyBody.js:20-115                               -> or_manybody
 *   link.js:34-97                                    -> or_link_seq, or_link_defaults
 *   collide.js:162-236                               -> or_collide_seq
 *   center.js:8-26, radial.js:16-28, x.js:11-16      -> or_center, or_radial, or_position
 *   examples/benchmark.html:86-93                    -> or_make_graph
 * and the tree library the reference imports but does not vendor (yarn.lock:310-328):
 * d3-binarytree 1.0.1, d3-quadtree 3.0.1, d3-octree 1.0.1 -- their published algorithms
 * (addAll: extent, cover(min), cover(max), add each point; add: descend by midpoints,
 * chain coincident points at the head, split a leaf until the two points separate;
 * visit: pre-order, children 0..2^d-1; visitAfter: post-order) restated for d = 1, 2, 3.
 * The Jacobi variants restate the GPU passes (passes/link.js:17-47, passes/collide.js:24-81)
 * in FP64.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "d3oracle.h"

/* ------------------------------------------------------------------------------ */
/* random sources                                                                   */

double or_lcg(uint64_t* state) {
  *state = (1664525ull * *state + 1013904223ull) % 4294967296ull;
  return (double)*state / 4294967296.0;
}

static double jiggle(uint64_t* rng) { return (or_lcg(rng) - 0.5) * 1e-6; }

static uint32_t pcg(uint32_t v) {
  uint32_t s = v * 747796405u + 2891336453u;
  uint32_t w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
  return (w >> 22u) ^ w;
}

float or_gpu_jiggle(uint32_t a, uint32_t b, uint32_t axis, uint32_t step, uint32_t seed) {
  uint32_t h = pcg(a ^ pcg(b ^ pcg(seed ^ pcg(step * 3u + axis))));
  return ((float)h * (1.0f / 4294967296.0f) - 0.5f) * 1e-6f;
}

/* ------------------------------------------------------------------------------ */
/* inputs                                                                           */

void or_phyllotaxis(int32_t n, int32_t ndim, double* x, double* y, double* z) {
  const double initialRadius = 10;
  const double roll = M_PI * (3 - sqrt(5.0));
  const double yaw = M_PI * 20 / (9 + sqrt(221.0));
  for (int32_t i = 0; i < n; ++i) {
    double radius = initialRadius * (ndim > 2 ? cbrt(0.5 + i) : (ndim > 1 ? sqrt(0.5 + i) : i));
    double ra = i * roll, ya = i * yaw;
    if (ndim == 1) {
      x[i] = radius;
      if (y) y[i] = 0;
      if (z) z[i] = 0;
    } else if (ndim == 2) {
      x[i] = radius * cos(ra);
      y[i] = radius * sin(ra);
      if (z) z[i] = 0;
    } else {
      x[i] = radius * sin(ra) * cos(ya);
      y[i] = radius * cos(ra);
      z[i] = radius * sin(ra) * sin(ya);
    }
  }
}

int32_t or_make_graph(int32_t n, uint32_t seed, uint32_t* src, uint32_t* dst) {
  uint64_t s = seed;
  int32_t m = 0;
  for (int32_t i = 1; i < n; ++i) {
    src[m] = (uint32_t)i;
    dst[m] = (uint32_t)floor(or_lcg(&s) * i);
    ++m;
  }
  for (int32_t i = 0; i < n / 2.0; ++i) {
    src[m] = (uint32_t)floor(or_lcg(&s) * n);
    dst[m] = (uint32_t)floor(or_lcg(&s) * n);
    ++m;
  }
  return m;
}

/* ------------------------------------------------------------------------------ */
/* the d3 trees (binary / quad / oct), one implementation over d = 1..3 axes          */

typedef struct {
  int32_t child[8]; /* internal: pool indices, -1 = empty */
  int32_t data;     /* leaf: point index; internal: -1 */
  int32_t next;     /* leaf: next coincident leaf, -1 = end */
  double c[3];      /* many-body: centre of charge (node.x/y/z of accumulate) */
  double value;     /* many-body: strength */
  double r;         /* collide: radius */
} DNode;

typedef struct {
  int d;            /* axes */
  DNode* pool;
  int32_t count, cap;
  int32_t root;     /* -1 = empty */
  double x0[3], x1[3];
  int has_extent;
  const double* p[3];
} DTree;

static int32_t tree_alloc(DTree* t, int32_t data) {
  if (t->count == t->cap) {
    t->cap = t->cap ? 2 * t->cap : 1024;
    t->pool = (DNode*)realloc(t->pool, sizeof(DNode) * (size_t)t->cap);
  }
  DNode* nd = &t->pool[t->count];
  for (int k = 0; k < 8; ++k) nd->child[k] = -1;
  nd->data = data;
  nd->next = -1;
  nd->c[0] = nd->c[1] = nd->c[2] = 0;
  nd->value = 0;
  nd->r = 0;
  return t->count++;
}

static int is_leaf(const DTree* t, int32_t q) { return t->pool[q].data >= 0; }

/* cover(): integer origin on the first point, then double towards the point. */
static void tree_cover(DTree* t, const double* q) {
  int d = t->d;
  for (int a = 0; a < d; ++a)
    if (isnan(q[a])) return;
  if (!t->has_extent) {
    for (int a = 0; a < d; ++a) {
      t->x0[a] = floor(q[a]);
      t->x1[a] = t->x0[a] + 1;
    }
    t->has_extent = 1;
    return;
  }
  double z = t->x1[0] - t->x0[0];
  if (!(z != 0)) z = 1;
  int32_t node = t->root;
  for (;;) {
    int outside = 0;
    for (int a = 0; a < d; ++a)
      if (t->x0[a] > q[a] || q[a] >= t->x1[a]) outside = 1;
    if (!outside) break;
    int i = 0;
    for (int a = 0; a < d; ++a) i |= (q[a] < t->x0[a]) << a;
    int32_t parent = tree_alloc(t, -1);
    t->pool[parent].child[i] = node;
    node = parent;
    z *= 2;
    for (int a = 0; a < d; ++a) {
      if ((i >> a) & 1) t->x0[a] = t->x1[a] - z;
      else t->x1[a] = t->x0[a] + z;
    }
  }
  if (t->root >= 0 && !is_leaf(t, t->root)) t->root = node;
}

static void tree_add(DTree* t, int32_t index) {
  int d = t->d;
  double q[3];
  for (int a = 0; a < d; ++a) {
    q[a] = t->p[a][index];
    if (isnan(q[a])) return;
  }
  int32_t leaf = tree_alloc(t, index);
  if (t->root < 0) {
    t->root = leaf;
    return;
  }
  double x0[3], x1[3], xm[3];
  for (int a = 0; a < d; ++a) x0[a] = t->x0[a], x1[a] = t->x1[a];
  int32_t node = t->root, parent = -1;
  int i = 0;
  while (!is_leaf(t, node)) {
    i = 0;
    for (int a = 0; a < d; ++a) {
      xm[a] = (x0[a] + x1[a]) / 2;
      int right = q[a] >= xm[a];
      if (right) x0[a] = xm[a];
      else x1[a] = xm[a];
      i |= right << a;
    }
    parent = node;
    node = t->pool[parent].child[i];
    if (node < 0) {
      t->pool[parent].child[i] = leaf;
      return;
    }
  }
  /* node is a leaf: coincident? */
  double qp[3];
  int same = 1;
  for (int a = 0; a < d; ++a) {
    qp[a] = t->p[a][t->pool[node].data];
    if (q[a] != qp[a]) same = 0;
  }
  if (same) {
    t->pool[leaf].next = node;
    if (parent >= 0) t->pool[parent].child[i] = leaf;
    else t->root = leaf;
    return;
  }
  int j;
  do {
    int32_t fresh = tree_alloc(t, -1);
    if (parent >= 0) t->pool[parent].child[i] = fresh;
    else t->root = fresh;
    parent = fresh;
    i = 0;
    j = 0;
    for (int a = 0; a < d; ++a) {
      xm[a] = (x0[a] + x1[a]) / 2;
      int right = q[a] >= xm[a];
      if (right) x0[a] = xm[a];
      else x1[a] = xm[a];
      i |= right << a;
      j |= (qp[a] >= xm[a]) << a;
    }
  } while (i == j);
  t->pool[parent].child[j] = node;
  t->pool[parent].child[i] = leaf;
}

static void tree_build(DTree* t, int d, int32_t n, const double* px, const double* py,
                       const double* pz) {
  memset(t, 0, sizeof(*t));
  t->d = d;
  t->root = -1;
  t->p[0] = px;
  t->p[1] = py;
  t->p[2] = pz;
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int32_t k = 0; k < n; ++k) {
    int bad = 0;
    for (int a = 0; a < d; ++a)
      if (isnan(t->p[a][k])) bad = 1;
    if (bad) continue;
    for (int a = 0; a < d; ++a) {
      double v = t->p[a][k];
      if (v < lo[a]) lo[a] = v;
      if (v > hi[a]) hi[a] = v;
    }
  }
  for (int a = 0; a < d; ++a)
    if (lo[a] > hi[a]) return;
  tree_cover(t, lo);
  tree_cover(t, hi);
  for (int32_t k = 0; k < n; ++k) tree_add(t, k);
}

static void tree_free(DTree* t) {
  free(t->pool);
  t->pool = NULL;
}

/* visit(): pre-order, child 0 first; the callback returns nonzero to skip the children */
typedef struct {
  int32_t node;
  double x0[3], x1[3];
} Quad;

typedef int (*visit_fn)(void* ctx, DTree* t, int32_t q, const double* x0, const double* x1);

static void tree_visit(DTree* t, visit_fn fn, void* ctx, Quad** stack, int32_t* cap) {
  if (t->root < 0) return;
  int d = t->d, nch = 1 << d;
  int32_t sp = 0;
  if (*cap < 64) {
    *cap = 64;
    *stack = (Quad*)realloc(*stack, sizeof(Quad) * (size_t)*cap);
  }
  Quad* s = *stack;
  s[sp].node = t->root;
  for (int a = 0; a < 3; ++a) s[sp].x0[a] = t->x0[a], s[sp].x1[a] = t->x1[a];
  ++sp;
  while (sp > 0) {
    Quad q = s[--sp];
    if (!fn(ctx, t, q.node, q.x0, q.x1) && !is_leaf(t, q.node)) {
      double xm[3];
      for (int a = 0; a < d; ++a) xm[a] = (q.x0[a] + q.x1[a]) / 2;
      for (int c = nch - 1; c >= 0; --c) {
        int32_t ch = t->pool[q.node].child[c];
        if (ch < 0) continue;
        if (sp + 1 > *cap) {
          *cap *= 2;
          *stack = (Quad*)realloc(*stack, sizeof(Quad) * (size_t)*cap);
          s = *stack;
        }
        Quad* e = &s[sp++];
        e->node = ch;
        for (int a = 0; a < 3; ++a) e->x0[a] = q.x0[a], e->x1[a] = q.x1[a];
        for (int a = 0; a < d; ++a) {
          if ((c >> a) & 1) e->x0[a] = xm[a];
          else e->x1[a] = xm[a];
        }
      }
    }
  }
}

/* ------------------------------------------------------------------------------ */
/* many-body                                                                        */

/* visitAfter(accumulate), manyBody.js:42-71: children before parents */
static void mb_accumulate(DTree* t, int32_t q, const double* strengths) {
  DNode* nd = &t->pool[q];
  int d = t->d;
  if (!is_leaf(t, q)) {
    int nch = 1 << d;
    for (int c = 0; c < nch; ++c)
      if (nd->child[c] >= 0) mb_accumulate(t, nd->child[c], strengths);
    nd = &t->pool[q];
    double strength = 0, weight = 0, s[3] = {0, 0, 0};
    for (int c = 0; c < nch; ++c) {
      int32_t ch = nd->child[c];
      if (ch < 0) continue;
      DNode* k = &t->pool[ch];
      double cv = fabs(k->value);
      if (!(cv != 0)) continue;
      strength += k->value;
      weight += cv;
      for (int a = 0; a < d; ++a) s[a] += cv * (isnan(k->c[a]) ? 0 : k->c[a]);
    }
    strength *= sqrt(4.0 / nch);
    for (int a = 0; a < d; ++a) nd->c[a] = s[a] / weight;
    nd->value = strength;
  } else {
    double strength = 0;
    for (int a = 0; a < d; ++a) nd->c[a] = t->p[a][nd->data];
    int32_t k = q;
    do {
      strength += strengths[t->pool[k].data];
      k = t->pool[k].next;
    } while (k >= 0);
    nd->value = strength;
  }
}

typedef struct {
  int32_t self;
  int d;
  double me[3];
  double v[3];
  const double* strengths;
  double alpha, theta2, dmin2, dmax2;
  uint64_t* rng;
  int64_t interactions, opened;
} MBCtx;

static int mb_apply(void* vctx, DTree* t, int32_t q, const double* x0, const double* x1) {
  MBCtx* m = (MBCtx*)vctx;
  DNode* nd = &t->pool[q];
  int d = m->d;
  if (!(nd->value != 0)) return 1;
  double dx[3] = {0, 0, 0};
  for (int a = 0; a < d; ++a) dx[a] = nd->c[a] - m->me[a];
  double w = x1[0] - x0[0];
  double l = dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2];
  if (w * w / m->theta2 < l) {
    if (l < m->dmax2) {
      for (int a = 0; a < d; ++a)
        if (dx[a] == 0) dx[a] = jiggle(m->rng), l += dx[a] * dx[a];
      if (l < m->dmin2) l = sqrt(m->dmin2 * l);
      for (int a = 0; a < d; ++a) m->v[a] += dx[a] * nd->value * m->alpha / l;
      m->interactions++;
    }
    return 1;
  } else if (!is_leaf(t, q) || l >= m->dmax2) {
    if (!is_leaf(t, q)) m->opened++;
    return 0;
  }
  if (nd->data != m->self || nd->next >= 0) {
    for (int a = 0; a < d; ++a)
      if (dx[a] == 0) dx[a] = jiggle(m->rng), l += dx[a] * dx[a];
    if (l < m->dmin2) l = sqrt(m->dmin2 * l);
  }
  int32_t k = q;
  do {
    int32_t data = t->pool[k].data;
    if (data != m->self) {
      double ww = m->strengths[data] * m->alpha / l;
      for (int a = 0; a < d; ++a) m->v[a] += dx[a] * ww;
      m->interactions++;
    }
    k = t->pool[k].next;
  } while (k >= 0);
  return 0;
}

void or_manybody(int32_t n, int32_t ndim, const double* x, const double* y, const double* z,
                 double* vx, double* vy, double* vz, const double* strengths, double alpha,
                 double theta2, double dmin2, double dmax2, uint64_t* rng, int64_t* stats) {
  DTree t;
  const double* P[3] = {x, y, z};
  double* V[3] = {vx, vy, vz};
  tree_build(&t, ndim, n, x, y, z);
  if (t.root >= 0) mb_accumulate(&t, t.root, strengths);
  MBCtx m;
  memset(&m, 0, sizeof(m));
  m.d = ndim;
  m.strengths = strengths;
  m.alpha = alpha;
  m.theta2 = theta2;
  m.dmin2 = dmin2;
  m.dmax2 = dmax2;
  m.rng = rng;
  Quad* stack = NULL;
  int32_t cap = 0;
  for (int32_t i = 0; i < n; ++i) {
    m.self = i;
    for (int a = 0; a < ndim; ++a) m.me[a] = P[a][i], m.v[a] = V[a][i];
    tree_visit(&t, mb_apply, &m, &stack, &cap);
    for (int a = 0; a < ndim; ++a) V[a][i] = m.v[a];
  }
  if (stats) {
    stats[0] = m.interactions;
    stats[1] = m.opened;
  }
  free(stack);
  tree_free(&t);
}

void or_manybody_exact(int32_t n, int32_t ndim, const double* x, const double* y, const double* z,
                       double* vx, double* vy, double* vz, const double* strengths, double alpha,
                       double dmin2, double dmax2) {
  const double* P[3] = {x, y, z};
  double* V[3] = {vx, vy, vz};
  for (int32_t i = 0; i < n; ++i) {
    double acc[3] = {0, 0, 0};
    for (int32_t j = 0; j < n; ++j) {
      if (j == i) continue;
      double dx[3] = {0, 0, 0};
      for (int a = 0; a < ndim; ++a) dx[a] = P[a][j] - P[a][i];
      double l = dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2];
      if (l >= dmax2 || l == 0) continue;
      if (l < dmin2) l = sqrt(dmin2 * l);
      double w = strengths[j] * alpha / l;
      for (int a = 0; a < ndim; ++a) acc[a] += dx[a] * w;
    }
    for (int a = 0; a < ndim; ++a) V[a][i] += acc[a];
  }
}

/* ------------------------------------------------------------------------------ */
/* link                                                                             */

void or_link_defaults(int32_t n, int32_t m, const uint32_t* src, const uint32_t* dst,
                      double* bias, double* strengths, double* distances) {
  double* count = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
  for (int32_t i = 0; i < m; ++i) count[src[i]] += 1, count[dst[i]] += 1;
  for (int32_t i = 0; i < m; ++i) {
    double cs = count[src[i]], ct = count[dst[i]];
    if (bias) bias[i] = cs / (cs + ct);
    if (strengths) strengths[i] = 1 / (cs < ct ? cs : ct);
    if (distances) distances[i] = 30;
  }
  free(count);
}

void or_link_seq(int32_t n, int32_t ndim, const double* x, const double* y, const double* z,
                 double* vx, double* vy, double* vz, int32_t m, const uint32_t* src,
                 const uint32_t* dst, const double* strengths, const double* distances,
                 const double* bias, int32_t iterations, double alpha, uint64_t* rng) {
  (void)n;
  const double* P[3] = {x, y, z};
  double* V[3] = {vx, vy, vz};
  for (int32_t k = 0; k < iterations; ++k) {
    for (int32_t i = 0; i < m; ++i) {
      uint32_t s = src[i], t = dst[i];
      double d[3] = {0, 0, 0};
      for (int a = 0; a < ndim; ++a) {
        d[a] = P[a][t] + V[a][t] - P[a][s] - V[a][s];
        if (!(d[a] != 0)) d[a] = jiggle(rng);
      }
      double l = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      l = (l - distances[i]) / l * alpha * strengths[i];
      for (int a = 0; a < ndim; ++a) d[a] *= l;
      double b = bias[i];
      for (int a = 0; a < ndim; ++a) V[a][t] -= d[a] * b;
      b = 1 - b;
      for (int a = 0; a < ndim; ++a) V[a][s] += d[a] * b;
    }
  }
}

void or_link_jacobi(int32_t n, int32_t ndim, const double* x, const double* y, const double* z,
                    double* vx, double* vy, double* vz, int32_t m, const uint32_t* src,
                    const uint32_t* dst, const double* strengths, const double* distances,
                    const double* bias, int32_t iterations, double alpha) {
  const double* P[3] = {x, y, z};
  double* V[3] = {vx, vy, vz};
  double* snap = (double*)malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
  double* acc = (double*)malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
  for (int32_t k = 0; k < iterations; ++k) {
    for (int32_t i = 0; i < n; ++i)
      for (int a = 0; a < 3; ++a) {
        snap[3 * i + a] = a < ndim ? P[a][i] + V[a][i] : 0;
        acc[3 * i + a] = 0;
      }
    for (int32_t e = 0; e < m; ++e) {
      uint32_t s = src[e], t = dst[e];
      double d[3] = {0, 0, 0};
      for (int a = 0; a < ndim; ++a) d[a] = snap[3 * t + a] - snap[3 * s + a];
      double l = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      if (l == 0) continue; /* coincident endpoints: the GPU jiggles; tests avoid them */
      double f = (l - distances[e]) / l * alpha * strengths[e];
      for (int a = 0; a < ndim; ++a) {
        acc[3 * t + a] -= d[a] * f * bias[e];
        acc[3 * s + a] += d[a] * f * (1 - bias[e]);
      }
    }
    for (int32_t i = 0; i < n; ++i)
      for (int a = 0; a < ndim; ++a) V[a][i] += acc[3 * i + a];
  }
  free(snap);
  free(acc);
}

/* ------------------------------------------------------------------------------ */
/* collide                                                                          */

static void cl_prepare(DTree* t, int32_t q, const double* radii) {
  DNode* nd = &t->pool[q];
  if (is_leaf(t, q)) {
    nd->r = radii[nd->data];
    return;
  }
  int nch = 1 << t->d;
  for (int c = 0; c < nch; ++c)
    if (nd->child[c] >= 0) cl_prepare(t, nd->child[c], radii);
  nd = &t->pool[q];
  nd->r = 0;
  for (int c = 0; c < nch; ++c) {
    int32_t ch = nd->child[c];
    if (ch >= 0 && t->pool[ch].r > nd->r) nd->r = t->pool[ch].r;
  }
}

typedef struct {
  int d;
  int32_t self;
  double ri, ri2, xi[3];
  double *P[3], *V[3];
  const double* radii;
  double strength;
  uint64_t* rng;
} CLCtx;

static int cl_apply(void* vctx, DTree* t, int32_t q, const double* x0, const double* x1) {
  CLCtx* c = (CLCtx*)vctx;
  DNode* nd = &t->pool[q];
  int d = c->d;
  double rj = nd->r, r = c->ri + rj;
  if (is_leaf(t, q)) {
    int32_t data = nd->data;
    if (data > c->self) {
      double dx[3] = {0, 0, 0};
      for (int a = 0; a < d; ++a) dx[a] = c->xi[a] - c->P[a][data] - c->V[a][data];
      double l = dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2];
      if (l < r * r) {
        for (int a = 0; a < d; ++a)
          if (dx[a] == 0) dx[a] = jiggle(c->rng), l += dx[a] * dx[a];
        double s = sqrt(l);
        l = (r - s) / s * c->strength;
        rj *= rj;
        r = rj / (c->ri2 + rj);
        for (int a = 0; a < d; ++a) {
          dx[a] *= l;
          c->V[a][c->self] += dx[a] * r;
        }
        r = 1 - r;
        for (int a = 0; a < d; ++a) c->V[a][data] -= dx[a] * r;
      }
    }
    return 0;
  }
  for (int a = 0; a < d; ++a)
    if (x0[a] > c->xi[a] + r || x1[a] < c->xi[a] - r) return 1;
  return 0;
}

void or_collide_seq(int32_t n, int32_t ndim, const double* x, const double* y, const double* z,
                    double* vx, double* vy, double* vz, const double* radii, double strength,
                    int32_t iterations, uint64_t* rng) {
  const double* P[3] = {x, y, z};
  double* V[3] = {vx, vy, vz};
  double* q[3];
  for (int a = 0; a < 3; ++a) q[a] = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  Quad* stack = NULL;
  int32_t cap = 0;
  CLCtx c;
  memset(&c, 0, sizeof(c));
  c.d = ndim;
  for (int a = 0; a < 3; ++a) c.P[a] = (double*)P[a], c.V[a] = V[a];
  c.radii = radii;
  c.strength = strength;
  c.rng = rng;
  for (int32_t k = 0; k < iterations; ++k) {
    for (int32_t i = 0; i < n; ++i)
      for (int a = 0; a < ndim; ++a) q[a][i] = P[a][i] + V[a][i];
    DTree t;
    tree_build(&t, ndim, n, q[0], q[1], q[2]);
    if (t.root >= 0) cl_prepare(&t, t.root, radii);
    for (int32_t i = 0; i < n; ++i) {
      c.self = i;
      c.ri = radii[i];
      c.ri2 = c.ri * c.ri;
      for (int a = 0; a < ndim; ++a) c.xi[a] = P[a][i] + V[a][i];
      tree_visit(&t, cl_apply, &c, &stack, &cap);
    }
    tree_free(&t);
  }
  free(stack);
  for (int a = 0; a < 3; ++a) free(q[a]);
}

/* Partners of one node in the Jacobi collide: every j != i closer than ri + rj to the
   snapshot of i (passes/collide.js:34-80), accumulated in ascending j.  Brute force
   visits every j; the cell list (cells 2 rmax wide, sorted by cell) visits the 3^d cells
   around i, collects the partners and sorts them, so both add the same terms in the same
   order and agree bit for bit. */
typedef struct {
  int64_t c[3];
  int32_t i;
} CellRef;

static int cellref_cmp(const void* pa, const void* pb) {
  const CellRef* a = (const CellRef*)pa;
  const CellRef* b = (const CellRef*)pb;
  for (int k = 2; k >= 0; --k) {
    if (a->c[k] < b->c[k]) return -1;
    if (a->c[k] > b->c[k]) return 1;
  }
  return a->i < b->i ? -1 : a->i > b->i;
}

static int int32_cmp(const void* pa, const void* pb) {
  int32_t a = *(const int32_t*)pa, b = *(const int32_t*)pb;
  return a < b ? -1 : a > b;
}

/* Cell -> first sorted entry, by open addressing over the distinct cells of `refs`. */
typedef struct {
  int64_t c[3];
  int32_t first;  /* -1: empty slot */
} CellSlot;

static uint64_t cell_hash(const int64_t c[3]) {
  uint64_t h = (uint64_t)c[0] * 0x9E3779B97F4A7C15ull;
  h ^= (uint64_t)c[1] * 0xC2B2AE3D27D4EB4Full + (h << 6) + (h >> 2);
  h ^= (uint64_t)c[2] * 0x165667B19E3779F9ull + (h << 6) + (h >> 2);
  return h ^ (h >> 29);
}

static int32_t cell_find(const CellSlot* tab, uint64_t mask, const int64_t c[3]) {
  for (uint64_t k = cell_hash(c) & mask;; k = (k + 1) & mask) {
    if (tab[k].first < 0) return -1;
    if (tab[k].c[0] == c[0] && tab[k].c[1] == c[1] && tab[k].c[2] == c[2]) return tab[k].first;
  }
}

static int collide_hit(const double* snap, const double* radii, int32_t i, int32_t j) {
  double r = radii[i] + radii[j];
  double d[3];
  for (int a = 0; a < 3; ++a) d[a] = snap[3 * i + a] - snap[3 * j + a];
  double l = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  return !(l >= r * r || l == 0);
}

static void collide_pair(const double* snap, const double* radii, int32_t i, int32_t j, int32_t ndim,
                         double strength, double* acc) {
  double ri = radii[i], ri2 = ri * ri;
  double rj = radii[j], r = ri + rj;
  double d[3];
  for (int a = 0; a < 3; ++a) d[a] = snap[3 * i + a] - snap[3 * j + a];
  double l = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  if (l >= r * r || l == 0) return;
  double sl = sqrt(l), rj2 = rj * rj;
  double f = (r - sl) / sl * strength * rj2 / (ri2 + rj2);
  for (int a = 0; a < ndim; ++a) acc[3 * i + a] += d[a] * f;
}

void or_collide_jacobi(int32_t n, int32_t ndim, const double* x, const double* y, const double* z,
                       double* vx, double* vy, double* vz, const double* radii, double strength,
                       int32_t iterations) {
  const double* P[3] = {x, y, z};
  double* V[3] = {vx, vy, vz};
  double* snap = (double*)malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
  double* acc = (double*)malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
  double rmax = 0;
  for (int32_t i = 0; i < n; ++i)
    if (radii[i] > rmax) rmax = radii[i];
  const int grid = n > 4096 && rmax > 0;
  CellRef* refs = grid ? (CellRef*)malloc(sizeof(CellRef) * (size_t)n) : NULL;
  int32_t* part = grid ? (int32_t*)malloc(sizeof(int32_t) * (size_t)n) : NULL;
  for (int32_t k = 0; k < iterations; ++k) {
    for (int32_t i = 0; i < n; ++i)
      for (int a = 0; a < 3; ++a) {
        snap[3 * i + a] = a < ndim ? P[a][i] + V[a][i] : 0;
        acc[3 * i + a] = 0;
      }
    if (!grid) {
      for (int32_t i = 0; i < n; ++i)
        for (int32_t j = 0; j < n; ++j)
          if (j != i) collide_pair(snap, radii, i, j, ndim, strength, acc);
    } else {
      const double cell = 2 * rmax;
      for (int32_t i = 0; i < n; ++i) {
        for (int a = 0; a < 3; ++a) refs[i].c[a] = (int64_t)floor(snap[3 * i + a] / cell);
        refs[i].i = i;
      }
      qsort(refs, (size_t)n, sizeof(CellRef), cellref_cmp);
      uint64_t slots = 1;
      while (slots < 2ull * (uint64_t)n) slots <<= 1;
      CellSlot* tab = (CellSlot*)malloc(sizeof(CellSlot) * slots);
      for (uint64_t k = 0; k < slots; ++k) tab[k].first = -1;
      for (int32_t t = 0; t < n; ++t) {
        if (t > 0 && refs[t].c[0] == refs[t - 1].c[0] && refs[t].c[1] == refs[t - 1].c[1] &&
            refs[t].c[2] == refs[t - 1].c[2])
          continue;
        uint64_t k = cell_hash(refs[t].c) & (slots - 1);
        while (tab[k].first >= 0) k = (k + 1) & (slots - 1);
        for (int a = 0; a < 3; ++a) tab[k].c[a] = refs[t].c[a];
        tab[k].first = t;
      }
      for (int32_t i = 0; i < n; ++i) {
        int64_t h[3];
        for (int a = 0; a < 3; ++a) h[a] = (int64_t)floor(snap[3 * i + a] / cell);
        int32_t np = 0;
        for (int dz = -(ndim > 2); dz <= (ndim > 2); ++dz)
          for (int dy = -(ndim > 1); dy <= (ndim > 1); ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
              int64_t c[3] = {h[0] + dx, h[1] + dy, h[2] + dz};
              int32_t t = cell_find(tab, slots - 1, c);
              if (t < 0) continue;
              for (; t < n; ++t) {
                if (refs[t].c[0] != c[0] || refs[t].c[1] != c[1] || refs[t].c[2] != c[2]) break;
                if (refs[t].i != i && collide_hit(snap, radii, i, refs[t].i)) part[np++] = refs[t].i;
              }
            }
        qsort(part, (size_t)np, sizeof(int32_t), int32_cmp);
        for (int32_t q = 0; q < np; ++q) collide_pair(snap, radii, i, part[q], ndim, strength, acc);
      }
      free(tab);
    }
    for (int32_t i = 0; i < n; ++i)
      for (int a = 0; a < ndim; ++a) V[a][i] += acc[3 * i + a];
  }
  free(refs);
  free(part);
  free(snap);
  free(acc);
}

/* ------------------------------------------------------------------------------ */
/* center / radial / position / integrate                                           */

void or_center(int32_t n, int32_t ndim, double* x, double* y, double* z, double cx, double cy,
               double cz, double strength) {
  double* P[3] = {x, y, z};
  double c[3] = {cx, cy, cz};
  double s[3] = {0, 0, 0};
  for (int32_t i = 0; i < n; ++i)
    for (int a = 0; a < ndim; ++a) s[a] += isnan(P[a][i]) ? 0 : P[a][i];
  for (int a = 0; a < 3; ++a) s[a] = (s[a] / n - c[a]) * strength;
  for (int32_t i = 0; i < n; ++i)
    for (int a = 0; a < ndim; ++a)
      if (s[a] != 0 && !isnan(s[a])) P[a][i] -= s[a];
}

void or_radial(int32_t n, int32_t ndim, const double* x, const double* y, const double* z,
               double* vx, double* vy, double* vz, const double* radiuses,
               const double* strengths, double cx, double cy, double cz, double alpha) {
  for (int32_t i = 0; i < n; ++i) {
    double px = x[i], py = ndim > 1 ? y[i] : 0, pz = ndim > 2 ? z[i] : 0;
    double dx = px - cx, dy = py - cy, dz = pz - cz;
    if (!(dx != 0)) dx = 1e-6;
    if (!(dy != 0)) dy = 1e-6;
    if (!(dz != 0)) dz = 1e-6;
    double r = sqrt(dx * dx + dy * dy + dz * dz);
    double k = (radiuses[i] - r) * strengths[i] * alpha / r;
    vx[i] += dx * k;
    if (ndim > 1) vy[i] += dy * k;
    if (ndim > 2) vz[i] += dz * k;
  }
}

void or_position(int32_t n, const double* p, double* v, const double* targets,
                 const double* strengths, double alpha) {
  for (int32_t i = 0; i < n; ++i) v[i] += (targets[i] - p[i]) * strengths[i] * alpha;
}

void or_integrate(int32_t n, int32_t ndim, double* x, double* y, double* z, double* vx,
                  double* vy, double* vz, const double* fx, const double* fy, const double* fz,
                  double decay) {
  double* P[3] = {x, y, z};
  double* V[3] = {vx, vy, vz};
  const double* F[3] = {fx, fy, fz};
  for (int32_t i = 0; i < n; ++i)
    for (int a = 0; a < ndim; ++a) {
      if (!F[a] || isnan(F[a][i])) P[a][i] += V[a][i] *= decay;
      else P[a][i] = F[a][i], V[a][i] = 0;
    }
}

/* ------------------------------------------------------------------------------ */
/* a whole simulation                                                               */

typedef struct {
  fg_force_desc d;
  uint32_t* src;
  uint32_t* dst;
  double* a; /* strengths */
  double* b; /* radii / distances / targets */
  double* c; /* bias */
} OrForce;

struct or_sim {
  int32_t n, ndim;
  double *x, *y, *z, *vx, *vy, *vz, *fx, *fy, *fz;
  double alpha, alpha_min, alpha_decay, alpha_target, decay;
  uint64_t rng;
  int32_t jacobi;
  int32_t k;
  OrForce f[32];
  int64_t stats[2];
};

static double* dup(const double* p, int32_t n, double fill) {
  double* q = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  for (int32_t i = 0; i < n; ++i) q[i] = p ? p[i] : fill;
  return q;
}

or_sim* or_sim_create(int32_t n, int32_t ndim, const double* x, const double* y, const double* z,
                      const double* vx, const double* vy, const double* vz, const double* fx,
                      const double* fy, const double* fz) {
  or_sim* s = (or_sim*)calloc(1, sizeof(or_sim));
  s->n = n;
  s->ndim = ndim;
  s->x = dup(x, n, 0);
  s->y = dup(y, n, 0);
  s->z = dup(z, n, 0);
  s->vx = dup(vx, n, 0);
  s->vy = dup(vy, n, 0);
  s->vz = dup(vz, n, 0);
  s->fx = dup(fx, n, NAN);
  s->fy = dup(fy, n, NAN);
  s->fz = dup(fz, n, NAN);
  s->alpha = 1;
  s->alpha_min = 0.001;
  s->alpha_decay = 1 - pow(0.001, 1.0 / 300);
  s->alpha_target = 0;
  s->decay = 0.6;
  s->rng = 1;
  return s;
}

static void free_forces(or_sim* s) {
  for (int32_t i = 0; i < s->k; ++i) {
    free(s->f[i].src);
    free(s->f[i].dst);
    free(s->f[i].a);
    free(s->f[i].b);
    free(s->f[i].c);
  }
  s->k = 0;
}

void or_sim_destroy(or_sim* s) {
  if (!s) return;
  free_forces(s);
  free(s->x), free(s->y), free(s->z), free(s->vx), free(s->vy), free(s->vz);
  free(s->fx), free(s->fy), free(s->fz);
  free(s);
}

int32_t or_sim_set_forces(or_sim* s, int32_t k, const fg_force_desc* descs, int32_t jacobi) {
  if (k > 32) return -1;
  free_forces(s);
  s->jacobi = jacobi;
  for (int32_t i = 0; i < k; ++i) {
    OrForce* f = &s->f[i];
    memset(f, 0, sizeof(*f));
    f->d = descs[i];
    int32_t n = s->n, m = (int32_t)descs[i].count;
    switch (descs[i].type) {
      case FG_FORCE_MANYBODY: f->a = dup(descs[i].strengths, n, 0); break;
      case FG_FORCE_LINK:
        f->src = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(m > 0 ? m : 1));
        f->dst = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(m > 0 ? m : 1));
        for (int32_t e = 0; e < m; ++e) f->src[e] = descs[i].links[2 * e], f->dst[e] = descs[i].links[2 * e + 1];
        f->a = dup(descs[i].strengths, m, 0);
        f->b = dup(descs[i].distances, m, 0);
        f->c = dup(descs[i].bias, m, 0);
        break;
      case FG_FORCE_COLLIDE: f->b = dup(descs[i].radii, n, 0); break;
      case FG_FORCE_RADIAL:
        f->a = dup(descs[i].strengths, n, 0);
        f->b = dup(descs[i].radii, n, 0);
        break;
      case FG_FORCE_POSITION:
        f->a = dup(descs[i].strengths, n, 0);
        f->b = dup(descs[i].targets, n, 0);
        break;
      case FG_FORCE_CENTER: break;
      default: return -1;
    }
  }
  s->k = k;
  return 0;
}

void or_sim_set_params(or_sim* s, double alpha, double alpha_min, double alpha_decay,
                       double alpha_target, double decay) {
  s->alpha = alpha;
  s->alpha_min = alpha_min;
  s->alpha_decay = alpha_decay;
  s->alpha_target = alpha_target;
  s->decay = decay;
}

void or_sim_apply_forces(or_sim* s, double alpha) {
  int32_t n = s->n, d = s->ndim;
  for (int32_t i = 0; i < s->k; ++i) {
    OrForce* f = &s->f[i];
    switch (f->d.type) {
      case FG_FORCE_MANYBODY:
        if (f->d.theta2 > 1e-6 || !s->jacobi)
          or_manybody(n, d, s->x, s->y, s->z, s->vx, s->vy, s->vz, f->a, alpha, f->d.theta2,
                      f->d.distance_min2, f->d.distance_max2, &s->rng, s->stats);
        else
          or_manybody_exact(n, d, s->x, s->y, s->z, s->vx, s->vy, s->vz, f->a, alpha,
                            f->d.distance_min2, f->d.distance_max2);
        break;
      case FG_FORCE_LINK:
        if (s->jacobi)
          or_link_jacobi(n, d, s->x, s->y, s->z, s->vx, s->vy, s->vz, (int32_t)f->d.count, f->src,
                         f->dst, f->a, f->b, f->c, f->d.iterations, alpha);
        else
          or_link_seq(n, d, s->x, s->y, s->z, s->vx, s->vy, s->vz, (int32_t)f->d.count, f->src,
                      f->dst, f->a, f->b, f->c, f->d.iterations, alpha, &s->rng);
        break;
      case FG_FORCE_COLLIDE:
        if (s->jacobi)
          or_collide_jacobi(n, d, s->x, s->y, s->z, s->vx, s->vy, s->vz, f->b, f->d.strength,
                            f->d.iterations);
        else
          or_collide_seq(n, d, s->x, s->y, s->z, s->vx, s->vy, s->vz, f->b, f->d.strength,
                         f->d.iterations, &s->rng);
        break;
      case FG_FORCE_CENTER:
        or_center(n, d, s->x, s->y, s->z, f->d.x, f->d.y, f->d.z, f->d.strength);
        break;
      case FG_FORCE_RADIAL:
        or_radial(n, d, s->x, s->y, s->z, s->vx, s->vy, s->vz, f->b, f->a, f->d.x, f->d.y, f->d.z,
                  alpha);
        break;
      case FG_FORCE_POSITION: {
        int a = f->d.axis;
        if (a >= d) break;
        double* P = a == 0 ? s->x : a == 1 ? s->y : s->z;
        double* V = a == 0 ? s->vx : a == 1 ? s->vy : s->vz;
        or_position(n, P, V, f->b, f->a, alpha);
        break;
      }
      default: break;
    }
  }
}

void or_sim_tick(or_sim* s, int32_t iterations) {
  for (int32_t k = 0; k < iterations; ++k) {
    s->alpha += (s->alpha_target - s->alpha) * s->alpha_decay;
    or_sim_apply_forces(s, s->alpha);
    or_integrate(s->n, s->ndim, s->x, s->y, s->z, s->vx, s->vy, s->vz, s->fx, s->fy, s->fz,
                 s->decay);
  }
}

double or_sim_alpha(const or_sim* s) { return s->alpha; }

void or_sim_get(const or_sim* s, double* x, double* y, double* z, double* vx, double* vy,
                double* vz) {
  size_t b = sizeof(double) * (size_t)s->n;
  if (x) memcpy(x, s->x, b);
  if (y) memcpy(y, s->y, b);
  if (z) memcpy(z, s->z, b);
  if (vx) memcpy(vx, s->vx, b);
  if (vy) memcpy(vy, s->vy, b);
  if (vz) memcpy(vz, s->vz, b);
}

void or_sim_set_state(or_sim* s, const double* x, const double* y, const double* z,
                      const double* vx, const double* vy, const double* vz) {
  size_t b = sizeof(double) * (size_t)s->n;
  if (x) memcpy(s->x, x, b);
  if (y) memcpy(s->y, y, b);
  if (z) memcpy(s->z, z, b);
  if (vx) memcpy(s->vx, vx, b);
  if (vy) memcpy(s->vy, vy, b);
  if (vz) memcpy(s->vz, vz, b);
}

int64_t or_sim_stat(const or_sim* s, int32_t which) {
  return which >= 0 && which < 2 ? s->stats[which] : -1;
}
