/*
 * gputree.c -- bit-exact restatement of the reference GPU Barnes-Hut tree.  TEST
 * INFRASTRUCTURE ONLY (the checker of the HIP tree kernels).
 *
 *   or_gt_cube      bbox + cover()          barnesHut.js:131-211
 *   or_gt_keys      cellOf / morton / keys  barnesHut.js:317-366
 *   or_gt_order     the order of sort + fix/longRuns (or rank): stable on (hi, lo, node)
 *                                           sort.js:1-20, barnesHut.js:368-530
 *   or_gt_build     delta / depthOf / build barnesHut.js:84-117, 587-635; prep's leaf
 *                                           depth barnesHut.js:267-268
 *   or_gt_collapse  collapse                barnesHut.js:641-671
 *   or_gt_cells     cell values and centres over that tree in FP64, with d3's recursive
 *                   weighting (src/manyBody.js:42-71) -- what the GPU's f32 pyramid and
 *                   recentre approximate (barnesHut.js:14-19, 673-703)
 * All integer and f32 arithmetic is done in the same precision and order as the WGSL.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "d3oracle.h"

#define LEAF 0x80000000u
#define NONE 0xffffffffu

static const int LEVELS_OF[4] = {0, 24, 24, 20};

int32_t or_gt_cube(int32_t n, int32_t ndim, const float* x, const float* y, const float* z,
                   float* cube) {
  float lo[3] = {3.4e38f, 3.4e38f, 3.4e38f}, hi[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
  const float* P[3] = {x, y, z};
  for (int32_t i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a) {
      float v = (a < ndim && P[a]) ? P[a][i] : 0.0f;
      if (v < lo[a]) lo[a] = v;
      if (v > hi[a]) hi[a] = v;
    }
  float origin[3] = {floorf(lo[0]), floorf(lo[1]), floorf(lo[2])};
  volatile float size = 1.0f;
  int log2size = 0;
  for (int k = 0; k < 128; ++k) {
    volatile float e0 = origin[0] + size, e1 = origin[1] + size, e2 = origin[2] + size;
    if (hi[0] < e0 && hi[1] < e1 && hi[2] < e2) break;
    size = size * 2.0f;
    log2size += 1;
  }
  cube[0] = origin[0];
  cube[1] = origin[1];
  cube[2] = origin[2];
  cube[3] = size;
  return LEVELS_OF[ndim] - log2size;
}

static uint32_t spread3(uint32_t v) {
  uint32_t x = v & 0x3ffu;
  x = (x | (x << 16u)) & 0x030000ffu;
  x = (x | (x << 8u)) & 0x0300f00fu;
  x = (x | (x << 4u)) & 0x030c30c3u;
  x = (x | (x << 2u)) & 0x09249249u;
  return x;
}

static uint32_t spread2(uint32_t v) {
  uint32_t x = v & 0xffffu;
  x = (x | (x << 8u)) & 0x00ff00ffu;
  x = (x | (x << 4u)) & 0x0f0f0f0fu;
  x = (x | (x << 2u)) & 0x33333333u;
  x = (x | (x << 1u)) & 0x55555555u;
  return x;
}

static uint32_t morton(int dims, int half_bits, const uint32_t* q) {
  uint32_t m = q[0];
  if (dims == 3) m = (spread3(q[2]) << 2) | (spread3(q[1]) << 1) | spread3(q[0]);
  if (dims == 2) m = (spread2(q[1]) << 1) | spread2(q[0]);
  return m << (32 - half_bits);
}

static float cell_of(float x, float origin, int k) {
  if (k >= 0) {
    volatile float a = floorf(ldexpf(x, k));
    volatile float b = ldexpf(origin, k);
    return a - b;
  }
  volatile float f = floorf(x) - origin;
  return floorf(ldexpf(f, k));
}

void or_gt_keys(int32_t n, int32_t ndim, const float* x, const float* y, const float* z,
                const float* cube, int32_t shift, uint32_t* hi, uint32_t* lo) {
  int levels = LEVELS_OF[ndim], half = levels / 2, half_bits = half * ndim;
  float top = (float)((1u << levels) - 1u);
  const float* P[3] = {x, y, z};
  for (int32_t i = 0; i < n; ++i) {
    uint32_t q[3], qh[3], ql[3];
    for (int a = 0; a < 3; ++a) {
      float p = (a < ndim && P[a]) ? P[a][i] : 0.0f;
      float c = cell_of(p, cube[a], shift);
      c = c < 0.0f ? 0.0f : (c > top ? top : c);
      q[a] = (uint32_t)c;
      qh[a] = q[a] >> half;
      ql[a] = q[a] & ((1u << half) - 1u);
    }
    hi[i] = morton(ndim, half_bits, qh);
    lo[i] = morton(ndim, half_bits, ql);
  }
}

static const uint32_t* g_hi;
static const uint32_t* g_lo;
static int cmp_code(const void* pa, const void* pb) {
  uint32_t a = *(const uint32_t*)pa, b = *(const uint32_t*)pb;
  if (g_hi[a] != g_hi[b]) return g_hi[a] < g_hi[b] ? -1 : 1;
  if (g_lo[a] != g_lo[b]) return g_lo[a] < g_lo[b] ? -1 : 1;
  return a < b ? -1 : (a > b);
}

void or_gt_order(int32_t n, const uint32_t* hi, const uint32_t* lo, uint32_t* order) {
  for (int32_t i = 0; i < n; ++i) order[i] = (uint32_t)i;
  g_hi = hi;
  g_lo = lo;
  qsort(order, (size_t)n, sizeof(uint32_t), cmp_code);
}

static int clz32(uint32_t v) { return v ? __builtin_clz(v) : 32; }

typedef struct {
  int32_t n, half_bits;
  const uint32_t *hi, *lo;
} Codes;

static int delta(const Codes* c, int32_t a, int32_t b) {
  if (b < 0 || b >= c->n) return -1;
  if (c->hi[a] != c->hi[b]) return clz32(c->hi[a] ^ c->hi[b]);
  if (c->lo[a] != c->lo[b]) return c->half_bits + clz32(c->lo[a] ^ c->lo[b]);
  return 64 + clz32((uint32_t)a ^ (uint32_t)b);
}

void or_gt_build(int32_t n, int32_t ndim, const uint32_t* shi, const uint32_t* slo,
                 uint32_t* left, uint32_t* right, uint32_t* depth, uint32_t* leaf_depth) {
  int levels = LEVELS_OF[ndim];
  Codes c = {n, levels / 2 * ndim, shi, slo};
  for (int32_t i = 0; i + 1 < n; ++i) {
    int before = delta(&c, i, i - 1), after = delta(&c, i, i + 1);
    int d = after > before ? 1 : -1;
    int floor_prefix = before < after ? before : after;
    int32_t span = 2;
    while (span < (1 << 30) && delta(&c, i, i + span * d) > floor_prefix) span *= 2;
    int32_t len = 0;
    for (int32_t t = span >> 1; t >= 1; t >>= 1)
      if (delta(&c, i, i + (len + t) * d) > floor_prefix) len += t;
    int32_t j = i + len * d;
    int prefix = delta(&c, i, j);
    int32_t s = 0, stride = len;
    for (;;) {
      stride = (stride + 1) >> 1;
      if (delta(&c, i, i + (s + stride) * d) > prefix) s += stride;
      if (stride <= 1) break;
    }
    int32_t split = i + s * d + (d < 0 ? d : 0);
    int32_t first = i < j ? i : j, last = i < j ? j : i;
    left[i] = (uint32_t)split | (first == split ? LEAF : 0u);
    right[i] = (uint32_t)(split + 1) | (last == split + 1 ? LEAF : 0u);
    uint32_t dep = (uint32_t)prefix / (uint32_t)ndim;
    depth[i] = dep < (uint32_t)levels ? dep : (uint32_t)levels;
  }
  if (leaf_depth) {
    for (int32_t l = 0; l < n; ++l) {
      uint32_t dd = 0;
      if (n > 1) {
        int a = delta(&c, l, l - 1), b = delta(&c, l, l + 1);
        int m = a > b ? a : b;
        uint32_t dep = (uint32_t)m / (uint32_t)ndim;
        dd = (dep < (uint32_t)levels ? dep : (uint32_t)levels) + 1;
      }
      leaf_depth[l] = dd;
    }
  }
}

void or_gt_collapse(int32_t n, int32_t ndim, const uint32_t* left, const uint32_t* right,
                    const uint32_t* depth, uint32_t* kids) {
  for (int32_t i = 0; i + 1 < n; ++i) {
    uint32_t list[8] = {left[i], right[i], NONE, NONE, NONE, NONE, NONE, NONE};
    int count = 2;
    for (int level = 1; level < ndim; ++level) {
      uint32_t next[8] = {NONE, NONE, NONE, NONE, NONE, NONE, NONE, NONE};
      int m = 0;
      for (int c = 0; c < count; ++c) {
        uint32_t k = list[c];
        if (!(k & LEAF) && depth[k] == depth[i]) {
          next[m] = left[k];
          next[m + 1] = right[k];
          m += 2;
        } else {
          next[m] = k;
          m += 1;
        }
      }
      memcpy(list, next, sizeof(list));
      count = m;
    }
    memcpy(kids + 8 * (size_t)i, list, sizeof(list));
  }
}

static double damp(int dims, int k) { return pow(sqrt(4.0 / (1 << dims)), k); }

void or_gt_cells(int32_t n, int32_t ndim, const uint32_t* order, const double* x,
                 const double* y, const double* z, const double* strengths,
                 const uint32_t* left, const uint32_t* right, const uint32_t* depth,
                 const uint32_t* leaf_depth, double* com, double* value) {
  if (n < 2) return;
  const double* P[3] = {x, y, z};
  int32_t inner = n - 1;
  const int levels = LEVELS_OF[ndim];
  double* A = (double*)malloc(sizeof(double) * (size_t)inner);
  double* G = (double*)malloc(sizeof(double) * 3 * (size_t)inner);
  double* W = (double*)malloc(sizeof(double) * (size_t)inner);
  /* same[i]: all points of node i lie at one position; up[i]: the depth of its parent */
  uint8_t* same = (uint8_t*)calloc((size_t)inner, 1);
  int32_t* up = (int32_t*)malloc(sizeof(int32_t) * (size_t)inner);
  up[0] = -1;
  for (int32_t i = 0; i < inner; ++i) {
    if (!(left[i] & LEAF)) up[left[i]] = (int32_t)depth[i];
    if (!(right[i] & LEAF)) up[right[i]] = (int32_t)depth[i];
  }
  /* iterative post-order from the root (node 0) */
  int32_t* stack = (int32_t*)malloc(sizeof(int32_t) * 2 * (size_t)n + 16);
  int32_t sp = 0;
  stack[sp++] = 0;
  stack[sp++] = 0;
  while (sp > 0) {
    int32_t state = stack[--sp], node = stack[--sp];
    if (state == 0) {
      stack[sp++] = node;
      stack[sp++] = 1;
      uint32_t kids[2] = {left[node], right[node]};
      for (int c = 0; c < 2; ++c)
        if (!(kids[c] & LEAF)) {
          stack[sp++] = (int32_t)kids[c];
          stack[sp++] = 0;
        }
      continue;
    }
    double a = 0, g[3] = {0, 0, 0}, w = 0;
    uint32_t kids[2] = {left[node], right[node]};
    int dp = (int)depth[node];
    for (int c = 0; c < 2; ++c) {
      uint32_t k = kids[c];
      if (k & LEAF) {
        uint32_t s = k & ~LEAF, v = order[s];
        double sv = strengths[v];
        a += sv * damp(ndim, (int)leaf_depth[s]);
        double ww = fabs(sv);
        for (int q = 0; q < ndim; ++q) g[q] += ww * P[q][v];
        w += ww;
      } else if ((int)depth[k] > dp) {
        a += A[k];
        double ww = fabs(A[k] * damp(ndim, -(dp + 1)));
        for (int q = 0; q < ndim; ++q) g[q] += ww * com[3 * k + q];
        w += ww;
      } else {
        a += A[k];
        for (int q = 0; q < ndim; ++q) g[q] += G[3 * k + q];
        w += W[k];
      }
    }
    /* d3 keeps points at one position as one leaf in the cell where they part from the
     * other points, its value their undamped sum (src/manyBody.js:61-68); on the GPU tree
     * that is the root of a run of equal codes (depth LEVELS) whose points coincide, one
     * level below its parent's cell */
    int vd = dp;
    if (dp == levels) {
      uint32_t split = left[node] & ~LEAF;
      int sl = (left[node] & LEAF) || same[left[node]];
      int sr = (right[node] & LEAF) || same[right[node]];
      uint32_t u = order[split], v = order[split + 1];
      int eq = 1;
      for (int q = 0; q < ndim; ++q) eq = eq && P[q][u] == P[q][v];
      same[node] = (uint8_t)(sl && sr && eq);
      if (up[node] != levels && same[node]) {
        int dl = up[node] < 0 ? 0 : up[node] + 1;
        a *= damp(ndim, dl - (levels + 1));
        vd = dl;
      }
    }
    A[node] = a;
    W[node] = w;
    for (int q = 0; q < 3; ++q) {
      G[3 * node + q] = g[q];
      com[3 * node + q] = (q < ndim && w > 0) ? g[q] / w : 0;
    }
    value[node] = a * damp(ndim, -vd);
  }
  free(same);
  free(up);
  free(stack);
  free(A);
  free(G);
  free(W);
}
