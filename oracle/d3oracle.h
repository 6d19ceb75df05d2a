/*
 * d3oracle.h -- CPU oracle for the forceSimulationGPU physics.  TEST INFRASTRUCTURE:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, as the
 * checker; the product never links or calls it.
 *
 * Two parts:
 *   d3oracle.c  FP64 restatement of d3-force-3d's CPU forces and tick (the src/ JS files of the
 *               reference) over its own restatement of d3-octree 1.0.1 /
 *               d3-quadtree 3.0.1 / d3-binarytree 1.0.1 (cover, add, visit, visitAfter).
 *   gputree.c   bit-exact restatement of the reference GPU tree (src/gpu/passes/barnesHut.js):
 *               cube, Morton halves, sorted order, Karras topology, leaf depths, collapse.
 */
#ifndef D3ORACLE_H
#define D3ORACLE_H

#include <stdint.h>
#include "../include/d3fgpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* --- random sources --------------------------------------------------------- */
/* lcg.js:1-9: s = (1664525 s + 1013904223) mod 2^32, returns s / 2^32; state starts at 1 */
double or_lcg(uint64_t* state);
/* the reference GPU jiggle (wgsl.js:28-39), f32 */
float or_gpu_jiggle(uint32_t a, uint32_t b, uint32_t axis, uint32_t step, uint32_t seed);

/* --- inputs ------------------------------------------------------------------ */
/* simulation.js:79-107 phyllotaxis for nodes without positions (all of them here) */
void or_phyllotaxis(int32_t n, int32_t ndim, double* x, double* y, double* z);
/* examples/benchmark.html:86-93: returns m; src/dst need n - 1 + ceil(n/2) entries */
int32_t or_make_graph(int32_t n, uint32_t seed, uint32_t* src, uint32_t* dst);

/* --- single forces, FP64, d3 semantics (one application with `alpha`) -------- */
/* manyBody.js:20-115 over the d3 tree; stats[0] = point/cell interactions, stats[1] = cells opened */
void or_manybody(int32_t n, int32_t ndim, const double* x, const double* y, const double* z,
                 double* vx, double* vy, double* vz, const double* strengths, double alpha,
                 double theta2, double dmin2, double dmax2, uint64_t* rng, int64_t* stats);
/* exact all-pairs (what manyBody.js gives with theta -> 0), no jiggle (coincident pairs skipped) */
void or_manybody_exact(int32_t n, int32_t ndim, const double* x, const double* y, const double* z,
                       double* vx, double* vy, double* vz, const double* strengths, double alpha,
                       double dmin2, double dmax2);
/* link.js:34-54: sequential (Gauss-Seidel) over links */
void or_link_seq(int32_t n, int32_t ndim, const double* x, const double* y, const double* z,
                 double* vx, double* vy, double* vz, int32_t m, const uint32_t* src,
                 const uint32_t* dst, const double* strengths, const double* distances,
                 const double* bias, int32_t iterations, double alpha, uint64_t* rng);
/* passes/link.js:17-47 semantics in FP64: every node against one snapshot (Jacobi) */
void or_link_jacobi(int32_t n, int32_t ndim, const double* x, const double* y, const double* z,
                    double* vx, double* vy, double* vz, int32_t m, const uint32_t* src,
                    const uint32_t* dst, const double* strengths, const double* distances,
                    const double* bias, int32_t iterations, double alpha);
/* link.js:56-97 defaults: bias[m], strengths[m] = 1/min(count), distances[m] = 30 */
void or_link_defaults(int32_t n, int32_t m, const uint32_t* src, const uint32_t* dst,
                      double* bias, double* strengths, double* distances);
/* collide.js:162-227: sequential over nodes, pairs found with the d3 tree */
void or_collide_seq(int32_t n, int32_t ndim, const double* x, const double* y, const double* z,
                    double* vx, double* vy, double* vz, const double* radii, double strength,
                    int32_t iterations, uint64_t* rng);
/* passes/collide.js:24-81 semantics in FP64: every pair against one snapshot (Jacobi) */
void or_collide_jacobi(int32_t n, int32_t ndim, const double* x, const double* y, const double* z,
                       double* vx, double* vy, double* vz, const double* radii, double strength,
                       int32_t iterations);
/* center.js:8-26 */
void or_center(int32_t n, int32_t ndim, double* x, double* y, double* z, double cx, double cy,
               double cz, double strength);
/* radial.js:16-28 */
void or_radial(int32_t n, int32_t ndim, const double* x, const double* y, const double* z,
               double* vx, double* vy, double* vz, const double* radiuses,
               const double* strengths, double cx, double cy, double cz, double alpha);
/* x.js:11-16 (y.js, z.js): v[axis] += (target - p[axis]) * strength * alpha */
void or_position(int32_t n, const double* p, double* v, const double* targets,
                 const double* strengths, double alpha);
/* simulation.js:61-73; fix* NaN = not pinned; decay = the multiplier (0.6 by default) */
void or_integrate(int32_t n, int32_t ndim, double* x, double* y, double* z, double* vx,
                  double* vy, double* vz, const double* fx, const double* fy, const double* fz,
                  double decay);

/* --- a whole simulation (simulation.js:23-199 tick) ---------------------------- */
typedef struct or_sim or_sim;
/* takes copies of the node arrays; fix NaN = not pinned; NULL fix arrays = none */
or_sim* or_sim_create(int32_t n, int32_t ndim, const double* x, const double* y, const double* z,
                      const double* vx, const double* vy, const double* vz, const double* fx,
                      const double* fy, const double* fz);
void or_sim_destroy(or_sim* s);
/* forces in application order; arrays copied.  jacobi != 0 runs link and collide
 * with the GPU's Jacobi semantics instead of d3's sequential ones. */
int32_t or_sim_set_forces(or_sim* s, int32_t k, const fg_force_desc* descs, int32_t jacobi);
/* alpha, alphaMin, alphaDecay, alphaTarget, velocity multiplier (simulation.js:28-32) */
void or_sim_set_params(or_sim* s, double alpha, double alpha_min, double alpha_decay,
                       double alpha_target, double decay);
void or_sim_tick(or_sim* s, int32_t iterations);
/* forces once at `alpha`, no integrate, no alpha update */
void or_sim_apply_forces(or_sim* s, double alpha);
double or_sim_alpha(const or_sim* s);
void or_sim_get(const or_sim* s, double* x, double* y, double* z, double* vx, double* vy,
                double* vz);
void or_sim_set_state(or_sim* s, const double* x, const double* y, const double* z,
                      const double* vx, const double* vy, const double* vz);
int64_t or_sim_stat(const or_sim* s, int32_t which);

/* --- the reference GPU tree, bit-exact (gputree.c) ----------------------------- */
/* bbox + cover (barnesHut.js:131-211): cube[4] = origin xyz, size; returns shift */
int32_t or_gt_cube(int32_t n, int32_t ndim, const float* x, const float* y, const float* z,
                   float* cube);
/* keys (barnesHut.js:317-366): hi/lo Morton halves per node */
void or_gt_keys(int32_t n, int32_t ndim, const float* x, const float* y, const float* z,
                const float* cube, int32_t shift, uint32_t* hi, uint32_t* lo);
/* stable order on (hi, lo, node) (sort.js + fix/longRuns/rank) */
void or_gt_order(int32_t n, const uint32_t* hi, const uint32_t* lo, uint32_t* order);
/* Karras build (barnesHut.js:587-635) over sorted codes: left/right (bit 31 = point),
 * depth per binary node; leaf_depth per sorted point (barnesHut.js:267-268) */
void or_gt_build(int32_t n, int32_t ndim, const uint32_t* shi, const uint32_t* slo,
                 uint32_t* left, uint32_t* right, uint32_t* depth, uint32_t* leaf_depth);
/* collapse (barnesHut.js:641-671): 8 refs per node, bit 31 = point, 0xffffffff = none */
void or_gt_collapse(int32_t n, int32_t ndim, const uint32_t* left, const uint32_t* right,
                    const uint32_t* depth, uint32_t* kids);
/* FP64 cell values over the GPU tree: com[3(n-1)], value[n-1] at each node's depth,
 * with d3's recursive |value|-weighted centres (manyBody.js:42-71) */
void or_gt_cells(int32_t n, int32_t ndim, const uint32_t* order, const double* x,
                 const double* y, const double* z, const double* strengths,
                 const uint32_t* left, const uint32_t* right, const uint32_t* depth,
                 const uint32_t* leaf_depth, double* com, double* value);

#ifdef __cplusplus
}
#endif

#endif
