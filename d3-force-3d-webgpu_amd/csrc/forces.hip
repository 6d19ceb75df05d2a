// forces.hip -- the streaming force passes and the integrator (SoA f32 state, wave64).
//
//   integrate         v *= decay; p += v; pinned axes p = fix, v = 0      (wgsl.js:67-84)
//   link              a snapshot pass (pos + vel as one float4 per node), then a
//                     node-centric gather over a CSR of incident links; the result goes to
//                     the other velocity buffer, so every node sees the same state
//                     (passes/link.js:17-47)
//   center            block sums in FP64, the last block forms the shift; then the shift
//                     (passes/center.js:11-72)
//   radial, position  one thread per node (passes/radial.js:9-27, passes/position.js:7-21)
//   manybody_pairs    exact all-pairs sum with LDS tiles (passes/manyBody.js:23-74)
//   collide_pairs     exact all-pairs collide for small graphs (passes/collide.js:24-81)
//   grid_*            collide through a hashed uniform grid (passes/collideGrid.js:6-22):
//                     cells 2 * rmax wide (and, for heavy-tailed radii, a second class of
//                     cells 2 * r95 wide for the points up to the 95th percentile radius),
//                     a slot per cell from a hashed block offset plus the cell's Morton index
//                     inside its block, radix sort of (slot, node), a table of each slot's
//                     run, then a neighbourhood query per sorted node.
#include "forces.h"
#include "batch.h"

#include <stdlib.h>

#include <algorithm>
#include <stdexcept>
#include <vector>

namespace fg {

// ---- integrate ----------------------------------------------------------------------

// Node i's step: the position minus a center shift (when a center force is the tick's
// last, applied here as center_apply would just before: rounded, then plus the velocity),
// the velocity decayed and added; a pinned axis takes its fixed value and no velocity.
template <int D>
__device__ inline void integrate_node(uint32_t i, float* __restrict__ x, float* __restrict__ y,
                                      float* __restrict__ z, float* __restrict__ vx, float* __restrict__ vy,
                                      float* __restrict__ vz, const uint8_t* __restrict__ mask,
                                      const float* __restrict__ fx, const float* __restrict__ fy,
                                      const float* __restrict__ fz, bool shifted, float sx, float sy, float sz,
                                      float decay) {
  uint32_t m = mask ? mask[i] : 0u;
  float px = x[i], py = D > 1 ? y[i] : 0.f, pz = D > 2 ? z[i] : 0.f;
  if (shifted) {
    if (sx != 0.0f) px -= sx;
    if (D > 1 && sy != 0.0f) py -= sy;
    if (D > 2 && sz != 0.0f) pz -= sz;
  }
  float v = vx[i] * decay, p = px + v;
  if (m & 1u) p = fx[i], v = 0.0f;
  x[i] = p;
  vx[i] = v;
  if (D > 1) {
    v = vy[i] * decay, p = py + v;
    if (m & 2u) p = fy[i], v = 0.0f;
    y[i] = p;
    vy[i] = v;
  }
  if (D > 2) {
    v = vz[i] * decay, p = pz + v;
    if (m & 4u) p = fz[i], v = 0.0f;
    z[i] = p;
    vz[i] = v;
  }
}

template <int D>
__global__ void __launch_bounds__(256) integrate_kernel(uint32_t n, float* __restrict__ x, float* __restrict__ y,
                                                        float* __restrict__ z, float* __restrict__ vx,
                                                        float* __restrict__ vy, float* __restrict__ vz,
                                                        const uint8_t* __restrict__ mask,
                                                        const float* __restrict__ fx,
                                                        const float* __restrict__ fy,
                                                        const float* __restrict__ fz,
                                                        const CenterWork* __restrict__ center,
                                                        StepParams step) {
  uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  float sx = 0.f, sy = 0.f, sz = 0.f;
  if (center) sx = center->shift[0], sy = center->shift[1], sz = center->shift[2];
  integrate_node<D>(i, x, y, z, vx, vy, vz, mask, fx, fy, fz, center != nullptr, sx, sy, sz, step.decay);
}

// A tick whose last force is center, over at most CENTER_FUSE_MAX nodes, in one launch: one
// workgroup sums the positions with center_reduce's arithmetic (node i on thread i, sums
// over each wave, then over each 256 threads as one of center_reduce's workgroups, those
// folded in order), so the shift is the same bit for bit, and then integrates every node.
template <int D>
__global__ void __launch_bounds__(CENTER_FUSE_MAX) center_integrate(
    uint32_t n, float* __restrict__ x, float* __restrict__ y, float* __restrict__ z, float* __restrict__ vx,
    float* __restrict__ vy, float* __restrict__ vz, const uint8_t* __restrict__ mask, const float* __restrict__ fx,
    const float* __restrict__ fy, const float* __restrict__ fz, double cx, double cy, double cz, double strength,
    StepParams step) {
  const uint32_t i = threadIdx.x, lane = i & 63, wave = i >> 6;
  double s[3] = {0, 0, 0};
  if (i < n) {
    s[0] += (double)x[i];
    if (D > 1) s[1] += (double)y[i];
    if (D > 2) s[2] += (double)z[i];
  }
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s[a] += __shfl_xor(s[a], o, WAVE);
  __shared__ double sh[CENTER_FUSE_MAX / 64][3];
  __shared__ float shift[3];
  if (lane == 0)
    for (int a = 0; a < 3; ++a) sh[wave][a] = s[a];
  __syncthreads();
  if (i == 0) {
    const uint32_t groups = (n + 255) / 256;
    double t[3] = {0, 0, 0};
    for (uint32_t b = 0; b < groups; ++b)
      for (int a = 0; a < 3; ++a) {
        double part = 0;
        part += sh[4 * b][a] + sh[4 * b + 1][a] + sh[4 * b + 2][a] + sh[4 * b + 3][a];
        t[a] += part;
      }
    const double c[3] = {cx, cy, cz};
    for (int a = 0; a < 3; ++a) shift[a] = a < D ? (float)((t[a] / n - c[a]) * strength) : 0.0f;
  }
  __syncthreads();
  if (i < n) integrate_node<D>(i, x, y, z, vx, vy, vz, mask, fx, fy, fz, true, shift[0], shift[1], shift[2], step.decay);
}

void launch_integrate(hipStream_t s, uint32_t n, int ndim, float* const* pos, float* const* vel,
                      const uint8_t* mask, const float* const* fix, StepParams step,
                      const CenterWork* center) {
  if (!n) return;
  dim3 g(ceil_div(n, 256)), b(256);
  if (ndim == 1)
    hipLaunchKernelGGL(integrate_kernel<1>, g, b, 0, s, n, pos[0], pos[1], pos[2], vel[0], vel[1], vel[2], mask,
                       fix[0], fix[1], fix[2], center, step);
  else if (ndim == 2)
    hipLaunchKernelGGL(integrate_kernel<2>, g, b, 0, s, n, pos[0], pos[1], pos[2], vel[0], vel[1], vel[2], mask,
                       fix[0], fix[1], fix[2], center, step);
  else
    hipLaunchKernelGGL(integrate_kernel<3>, g, b, 0, s, n, pos[0], pos[1], pos[2], vel[0], vel[1], vel[2], mask,
                       fix[0], fix[1], fix[2], center, step);
  FG_LAUNCH_CHECK();
}

void launch_center_integrate(hipStream_t s, uint32_t n, int ndim, float* const* pos, float* const* vel,
                             const uint8_t* mask, const float* const* fix, StepParams step, double cx, double cy,
                             double cz, double strength) {
  if (!n) return;
  if (n > CENTER_FUSE_MAX) throw std::logic_error("center_integrate: more nodes than one workgroup takes");
  if (ndim == 1)
    hipLaunchKernelGGL(center_integrate<1>, dim3(1), dim3(CENTER_FUSE_MAX), 0, s, n, pos[0], pos[1], pos[2], vel[0],
                       vel[1], vel[2], mask, fix[0], fix[1], fix[2], cx, cy, cz, strength, step);
  else if (ndim == 2)
    hipLaunchKernelGGL(center_integrate<2>, dim3(1), dim3(CENTER_FUSE_MAX), 0, s, n, pos[0], pos[1], pos[2], vel[0],
                       vel[1], vel[2], mask, fix[0], fix[1], fix[2], cx, cy, cz, strength, step);
  else
    hipLaunchKernelGGL(center_integrate<3>, dim3(1), dim3(CENTER_FUSE_MAX), 0, s, n, pos[0], pos[1], pos[2], vel[0],
                       vel[1], vel[2], mask, fix[0], fix[1], fix[2], cx, cy, cz, strength, step);
  FG_LAUNCH_CHECK();
}

// ---- link ---------------------------------------------------------------------------

// The snapshot every link reads: pos + vel of each node as one float4 (wgsl.js:56-65), so
// that a link's other end is one 16-byte gather rather than six.
template <int D>
__global__ void __launch_bounds__(256) snapshot_kernel(uint32_t n, const float* __restrict__ x,
                                                       const float* __restrict__ y,
                                                       const float* __restrict__ z,
                                                       const float* __restrict__ vx,
                                                       const float* __restrict__ vy,
                                                       const float* __restrict__ vz,
                                                       float4* __restrict__ snap) {
  uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  snap[i] = make_float4(x[i] + vx[i], D > 1 ? y[i] + vy[i] : 0.f, D > 2 ? z[i] + vz[i] : 0.f, 0.f);
}

template <int D>
__global__ void __launch_bounds__(256) link_kernel(uint32_t n, const float4* __restrict__ snap,
                                                   const float* __restrict__ vx,
                                                   const float* __restrict__ vy,
                                                   const float* __restrict__ vz,
                                                   float* __restrict__ ox, float* __restrict__ oy,
                                                   float* __restrict__ oz,
                                                   const uint32_t* __restrict__ row,
                                                   const LinkEntry* __restrict__ ent,
                                                   StepParams step) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;  // (every lane stays: long rows take the wave)
  link_wave<D>(i, i < n, SnapArray{snap}, vx, vy, vz, ox, oy, oz, row, ent, step.alpha, step.step, step.seed);
}

// The same with the snapshots summed where they are read: one launch instead of two, for
// graphs small enough that the launch costs more than the extra gathers.
template <int D>
__global__ void __launch_bounds__(256) link_direct_kernel(uint32_t n, const float* __restrict__ x,
                                                          const float* __restrict__ y,
                                                          const float* __restrict__ z,
                                                          const float* __restrict__ vx,
                                                          const float* __restrict__ vy,
                                                          const float* __restrict__ vz,
                                                          float* __restrict__ ox, float* __restrict__ oy,
                                                          float* __restrict__ oz,
                                                          const uint32_t* __restrict__ row,
                                                          const LinkEntry* __restrict__ ent,
                                                          StepParams step) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  link_wave<D>(i, i < n, SnapDirect<D>{x, y, z, vx, vy, vz}, vx, vy, vz, ox, oy, oz, row, ent, step.alpha,
               step.step, step.seed);
}


void launch_link(hipStream_t s, uint32_t n, int ndim, float* const* pos, const float* const* vin,
                 float* const* vout, const uint32_t* row, const LinkEntry* ent, float4* snap,
                 StepParams step) {
  if (!n) return;
  dim3 g(ceil_div(n, 256)), b(256);
  if (n <= LINK_DIRECT_MAX) {
    if (ndim == 1)
      hipLaunchKernelGGL(link_direct_kernel<1>, g, b, 0, s, n, pos[0], pos[1], pos[2], vin[0], vin[1], vin[2],
                         vout[0], vout[1], vout[2], row, ent, step);
    else if (ndim == 2)
      hipLaunchKernelGGL(link_direct_kernel<2>, g, b, 0, s, n, pos[0], pos[1], pos[2], vin[0], vin[1], vin[2],
                         vout[0], vout[1], vout[2], row, ent, step);
    else
      hipLaunchKernelGGL(link_direct_kernel<3>, g, b, 0, s, n, pos[0], pos[1], pos[2], vin[0], vin[1], vin[2],
                         vout[0], vout[1], vout[2], row, ent, step);
    FG_LAUNCH_CHECK();
    return;
  }
  if (ndim == 1) {
    hipLaunchKernelGGL(snapshot_kernel<1>, g, b, 0, s, n, pos[0], pos[1], pos[2], vin[0], vin[1], vin[2], snap);
    hipLaunchKernelGGL(link_kernel<1>, g, b, 0, s, n, (const float4*)snap, vin[0], vin[1], vin[2], vout[0], vout[1], vout[2], row, ent, step);
  } else if (ndim == 2) {
    hipLaunchKernelGGL(snapshot_kernel<2>, g, b, 0, s, n, pos[0], pos[1], pos[2], vin[0], vin[1], vin[2], snap);
    hipLaunchKernelGGL(link_kernel<2>, g, b, 0, s, n, (const float4*)snap, vin[0], vin[1], vin[2], vout[0], vout[1], vout[2], row, ent, step);
  } else {
    hipLaunchKernelGGL(snapshot_kernel<3>, g, b, 0, s, n, pos[0], pos[1], pos[2], vin[0], vin[1], vin[2], snap);
    hipLaunchKernelGGL(link_kernel<3>, g, b, 0, s, n, (const float4*)snap, vin[0], vin[1], vin[2], vout[0], vout[1], vout[2], row, ent, step);
  }
  FG_LAUNCH_CHECK();
}

// ---- center -------------------------------------------------------------------------

__global__ void __launch_bounds__(256) center_reduce(uint32_t n, int ndim, const float* __restrict__ x,
                                                     const float* __restrict__ y,
                                                     const float* __restrict__ z,
                                                     CenterWork* __restrict__ w, double cx, double cy,
                                                     double cz, double strength) {
  double s[3] = {0, 0, 0};
  auto take = [&](uint32_t i) {
    s[0] += (double)x[i];
    if (ndim > 1) s[1] += (double)y[i];
    if (ndim > 2) s[2] += (double)z[i];
  };
  const uint32_t stride = gridDim.x * 256;
  uint32_t i = blockIdx.x * 256 + threadIdx.x;
  for (; i + 3 * stride < n; i += 4 * stride) {  // four independent loads in flight
    take(i);
    take(i + stride);
    take(i + 2 * stride);
    take(i + 3 * stride);
  }
  for (; i < n; i += stride) take(i);
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s[a] += __shfl_xor(s[a], o, WAVE);
  __shared__ double sh[4][3];
  __shared__ uint32_t s_last;
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0)
    for (int a = 0; a < 3; ++a) sh[wave][a] = s[a];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int a = 0; a < 3; ++a) store_coherent(&w->partial[blockIdx.x][a], sh[0][a] + sh[1][a] + sh[2][a] + sh[3][a]);
    s_last = publish_count(&w->done) + 1 == gridDim.x;
  }
  __syncthreads();
  if (!s_last) return;
  // the last block: sum the partials in block order (deterministic), 256 threads wide
  __shared__ double tsum[256][3];
  double t[3] = {0, 0, 0};
  for (uint32_t b = threadIdx.x; b < gridDim.x; b += 256)
    for (int a = 0; a < 3; ++a) t[a] += load_coherent(&w->partial[b][a]);
  for (int a = 0; a < 3; ++a) tsum[threadIdx.x][a] = t[a];
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int a = 0; a < 3; ++a) t[a] = 0;
  for (uint32_t k = 0; k < gridDim.x && k < 256; ++k)  // the rest are +0: no change
    for (int a = 0; a < 3; ++a) t[a] += tsum[k][a];
  double c[3] = {cx, cy, cz};
  for (int a = 0; a < 3; ++a) w->shift[a] = a < ndim ? (float)((t[a] / n - c[a]) * strength) : 0.0f;
  w->done = 0;
}

__global__ void __launch_bounds__(256) center_apply(uint32_t n, int ndim, float* __restrict__ x,
                                                    float* __restrict__ y, float* __restrict__ z,
                                                    const CenterWork* __restrict__ w) {
  uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  float sx = w->shift[0], sy = w->shift[1], sz = w->shift[2];
  if (sx != 0.0f) x[i] -= sx;
  if (ndim > 1 && sy != 0.0f) y[i] -= sy;
  if (ndim > 2 && sz != 0.0f) z[i] -= sz;
}

void launch_center(hipStream_t s, uint32_t n, int ndim, float* const* pos, CenterWork* w, double cx,
                   double cy, double cz, double strength, bool apply, bool reduce) {
  if (!n) return;
  uint32_t blocks = ceil_div(n, 256);
  // few workgroups: each ends on one atomic on a shared counter, and those serialise
  uint32_t rb = blocks < CENTER_BLOCKS ? blocks : CENTER_BLOCKS;
  if (reduce) {  // else the shift is already in w (another kernel took the mean)
    hipLaunchKernelGGL(center_reduce, dim3(rb), dim3(256), 0, s, n, ndim, pos[0], pos[1], pos[2], w,
                       cx, cy, cz, strength);
    FG_LAUNCH_CHECK();
  }
  if (!apply) return;  // the integrator applies the shift
  hipLaunchKernelGGL(center_apply, dim3(blocks), dim3(256), 0, s, n, ndim, pos[0], pos[1], pos[2],
                     (const CenterWork*)w);
  FG_LAUNCH_CHECK();
}

// ---- radial / position --------------------------------------------------------------

__global__ void __launch_bounds__(256) radial_kernel(uint32_t n, int ndim, const float* __restrict__ x,
                                                     const float* __restrict__ y,
                                                     const float* __restrict__ z,
                                                     float* __restrict__ vx, float* __restrict__ vy,
                                                     float* __restrict__ vz,
                                                     const float* __restrict__ radius,
                                                     const float* __restrict__ weight, float cx,
                                                     float cy, float cz,
                                                     StepParams step) {
  uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  float dx = x[i] - cx, dy = ndim > 1 ? y[i] - cy : 0.f, dz = ndim > 2 ? z[i] - cz : 0.f;
  if (dx == 0.f) dx = 1e-6f;
  if (ndim > 1 && dy == 0.f) dy = 1e-6f;
  if (ndim > 2 && dz == 0.f) dz = 1e-6f;
  float r = sqrtf(dx * dx + dy * dy + dz * dz);
  float k = (radius[i] - r) * weight[i] * step.alpha / r;
  vx[i] += dx * k;
  if (ndim > 1) vy[i] += dy * k;
  if (ndim > 2) vz[i] += dz * k;
}

void launch_radial(hipStream_t s, uint32_t n, int ndim, float* const* pos, float* const* vel,
                   const float* radius, const float* weight, float cx, float cy, float cz,
                   StepParams step) {
  if (!n) return;
  hipLaunchKernelGGL(radial_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, ndim, pos[0], pos[1],
                     pos[2], vel[0], vel[1], vel[2], radius, weight, cx, cy, cz, step);
  FG_LAUNCH_CHECK();
}

__global__ void __launch_bounds__(256) position_kernel(uint32_t n, const float* __restrict__ p,
                                                       float* __restrict__ v,
                                                       const float* __restrict__ goal,
                                                       const float* __restrict__ weight,
                                                       StepParams step) {
  uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  v[i] += (goal[i] - p[i]) * weight[i] * step.alpha;
}

void launch_position(hipStream_t s, uint32_t n, const float* p, float* v, const float* goal,
                     const float* weight, StepParams step) {
  if (!n) return;
  hipLaunchKernelGGL(position_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, p, v, goal, weight,
                     step);
  FG_LAUNCH_CHECK();
}

// ---- exact many-body ----------------------------------------------------------------

constexpr int PAIR_TILE = 256;  // collide's all-pairs block (sources staged in LDS)

// The exact sum runs one wave per 64*B bodies (B per lane) and streams the sources through
// registers: each lane loads one source of a 64-source chunk and the wave broadcasts them
// in turn (readlane), so no LDS and no barrier.  The point force (manyBody.js:99-113) with
// distanceMin (l -> sqrt(dmin2 l) below it, manyBody.js:95) is
//     w / l' = w * rsq(l * max(l, dmin2)),
// one reciprocal square root whatever side of distanceMin the pair falls.  A pair at zero
// distance (the body itself, or a coincident point) has a zero displacement and adds
// exactly nothing on the fast path; PAIR_FLOOR keeps its factor finite.  Coincident
// points take the hash jiggle instead (manyBody.js:100-101 via jiggle, wgsl.js:45-52): the
// fast path records the smallest distance it saw, and a wave that saw zero outside its own
// chunks sums again on the careful path, which has the same arithmetic plus the jiggle.
constexpr float PAIR_FLOOR = 1.17549435e-38f;  // smallest normal float
constexpr float PAIR_PAD = 3e38f;              // padding source: l = inf, adds exactly zero

template <int D, bool DMAX, bool CAREFUL>
__device__ inline void pair_accum(float sx, float sy, float sz, float sw, uint32_t j, float mx, float my,
                                  float mz, uint32_t i, float dmin2, float dmax2, uint32_t sn,
                                  uint32_t seed, float& ax, float& ay, float& az, float& mn) {
  float dx = sx - mx, dy = D > 1 ? sy - my : 0.f, dz = D > 2 ? sz - mz : 0.f;
  float l = dx * dx;
  if (D > 1) l = __builtin_fmaf(dy, dy, l);
  if (D > 2) l = __builtin_fmaf(dz, dz, l);
  const bool near = !DMAX || l < dmax2;  // d3 tests distanceMax before the jiggle
  if (CAREFUL) {
    if (l == 0.f && j != i) {
      coincident(i, j, D, sn, seed, dx, dy, dz);
      l = dx * dx;
      if (D > 1) l = __builtin_fmaf(dy, dy, l);
      if (D > 2) l = __builtin_fmaf(dz, dz, l);
    }
  } else {
    mn = fminf(mn, l);
  }
  float f = sw * __builtin_amdgcn_rsqf(__builtin_fmaf(l, fmaxf(l, dmin2), PAIR_FLOOR));
  if (DMAX) f = near ? f : 0.f;
  ax = __builtin_fmaf(dx, f, ax);
  if (D > 1) ay = __builtin_fmaf(dy, f, ay);
  if (D > 2) az = __builtin_fmaf(dz, f, az);
}

__device__ inline float lane_bcast(float v, int k) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k));
}

// Chunks [c0, c1) for the lane's B bodies; chunks [own0, own0 + B) hold the wave's own
// bodies and always take the careful path.  Returns the smallest distance the fast path saw.
template <int D, int B, bool DMAX, bool CAREFUL>
__device__ inline float pairs_sum(uint32_t n, uint32_t c0, uint32_t c1, const float* __restrict__ x,
                                  const float* __restrict__ y, const float* __restrict__ z,
                                  const float* __restrict__ strength, float alpha, const float (&mx)[B],
                                  const float (&my)[B], const float (&mz)[B], const uint32_t (&idx)[B],
                                  uint32_t own0, float dmin2, float dmax2, uint32_t sn, uint32_t seed,
                                  float (&ax)[B], float (&ay)[B], float (&az)[B]) {
  const uint32_t lane = threadIdx.x & 63u;
  float mn = INFINITY;
  auto load = [&](uint32_t c, float& cx, float& cy, float& cz, float& cw) {
    const uint32_t j = (c << 6) + lane;
    const bool live = j < n;
    cx = live ? x[j] : PAIR_PAD;
    cy = live && D > 1 ? y[j] : PAIR_PAD;
    cz = live && D > 2 ? z[j] : PAIR_PAD;
    cw = live ? strength[j] * alpha : 0.f;  // the product d3's sum forms first
  };
  float cx, cy, cz, cw;
  load(c0, cx, cy, cz, cw);
  for (uint32_t c = c0; c < c1; ++c) {
    float nx = 0.f, ny = 0.f, nz = 0.f, nw = 0.f;
    if (c + 1 < c1) load(c + 1, nx, ny, nz, nw);
    const uint32_t j0 = c << 6;
    if (CAREFUL || (c >= own0 && c < own0 + B)) {
      for (int k = 0; k < 64; ++k) {
        const float sx = lane_bcast(cx, k), sy = D > 1 ? lane_bcast(cy, k) : 0.f;
        const float sz = D > 2 ? lane_bcast(cz, k) : 0.f, sw = lane_bcast(cw, k);
#pragma unroll
        for (int b = 0; b < B; ++b)
          pair_accum<D, DMAX, true>(sx, sy, sz, sw, j0 + k, mx[b], my[b], mz[b], idx[b], dmin2, dmax2, sn, seed,
                                    ax[b], ay[b], az[b], mn);
      }
    } else {
#pragma unroll
      for (int k = 0; k < 64; ++k) {
        const float sx = lane_bcast(cx, k), sy = D > 1 ? lane_bcast(cy, k) : 0.f;
        const float sz = D > 2 ? lane_bcast(cz, k) : 0.f, sw = lane_bcast(cw, k);
#pragma unroll
        for (int b = 0; b < B; ++b)
          pair_accum<D, DMAX, false>(sx, sy, sz, sw, j0 + k, mx[b], my[b], mz[b], idx[b], dmin2, dmax2, sn, seed,
                                     ax[b], ay[b], az[b], mn);
      }
    }
    cx = nx, cy = ny, cz = nz, cw = nw;
  }
  return mn;
}

// One wave per work item: body group g (64*B bodies, B per lane: body b of lane t is
// (g*B + b)*64 + t) against source part s of `splits` (chunks [s*C/splits, (s+1)*C/splits)
// of the C = ceil(n/64) chunks); item = s * groups + g.  With one part the sums go straight
// into the velocities; with several, each part's sums go to part[(s*3 + axis)*n + i] and
// pairs_combine adds them in part order.  For a given plan every body sums its sources in
// the same order, so the results do not depend on the launch.
template <int D, int B, bool DMAX>
__global__ void __launch_bounds__(64) manybody_pairs(uint32_t n, uint32_t groups, uint32_t splits,
                                                     const float* __restrict__ x, const float* __restrict__ y,
                                                     const float* __restrict__ z,
                                                     const float* __restrict__ strength,
                                                     float* __restrict__ vx, float* __restrict__ vy,
                                                     float* __restrict__ vz, float* __restrict__ part,
                                                     float dmin2, float dmax2, StepParams step) {
  const uint32_t g = blockIdx.x % groups, sp = blockIdx.x / groups, lane = threadIdx.x;
  const uint32_t chunks = (n + 63u) >> 6;
  const uint32_t c0 = (uint32_t)((uint64_t)sp * chunks / splits), c1 = (uint32_t)((uint64_t)(sp + 1) * chunks / splits);
  const uint32_t own0 = g * B;  // the group's bodies are chunks [own0, own0 + B)
  float mx[B], my[B], mz[B], ax[B], ay[B], az[B];
  uint32_t idx[B];
#pragma unroll
  for (int b = 0; b < B; ++b) {
    const uint32_t i = ((own0 + b) << 6) + lane;
    const bool live = i < n;
    idx[b] = i;
    // a padding body sits at NaN: it never reads as coincident and its sums are dropped
    mx[b] = live ? x[i] : NAN;
    my[b] = live && D > 1 ? y[i] : NAN;
    mz[b] = live && D > 2 ? z[i] : NAN;
    ax[b] = ay[b] = az[b] = 0.f;
  }
  const float alpha = step.alpha;
  const float mn = pairs_sum<D, B, DMAX, false>(n, c0, c1, x, y, z, strength, alpha, mx, my, mz, idx, own0, dmin2,
                                                dmax2, step.step, step.seed, ax, ay, az);
  if (__builtin_amdgcn_ballot_w64(mn == 0.f)) {  // a coincident pair outside the own chunks
#pragma unroll
    for (int b = 0; b < B; ++b) ax[b] = ay[b] = az[b] = 0.f;
    pairs_sum<D, B, DMAX, true>(n, c0, c1, x, y, z, strength, alpha, mx, my, mz, idx, own0, dmin2, dmax2, step.step,
                                step.seed, ax, ay, az);
  }
#pragma unroll
  for (int b = 0; b < B; ++b) {
    const uint32_t i = idx[b];
    if (i >= n) continue;
    if (splits == 1) {
      vx[i] += ax[b];
      if (D > 1) vy[i] += ay[b];
      if (D > 2) vz[i] += az[b];
    } else {
      float* out = part + (size_t)sp * 3 * n + i;
      out[0] = ax[b];
      if (D > 1) out[n] = ay[b];
      if (D > 2) out[2 * (size_t)n] = az[b];
    }
  }
}

template <int D>
__global__ void pairs_combine(uint32_t n, uint32_t splits, const float* __restrict__ part, float* __restrict__ vx,
                              float* __restrict__ vy, float* __restrict__ vz) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float* v[3] = {vx, vy, vz};
#pragma unroll
  for (int a = 0; a < D; ++a) {
    float t = part[(size_t)a * n + i];
    for (uint32_t s = 1; s < splits; ++s) t += part[((size_t)s * 3 + a) * n + i];
    v[a][i] += t;
  }
}

PairsPlan pairs_plan(uint32_t n, int cus) {
  PairsPlan p;
  p.groups = ceil_div(n, 64u * PAIRS_BODIES);
  p.splits = 1;
  // each SIMD runs its waves' VALU work in turn: pick the part count whose items spread
  // most evenly over the SIMDs (items / (SIMDs * ceil(items / SIMDs))), fewest parts on ties
  const uint64_t simds = 4ull * (uint64_t)(cus > 0 ? cus : 1);
  const uint32_t chunks = ceil_div(n, 64u);
  double best = 0;
  for (uint32_t s = 1; s <= PAIRS_MAX_SPLITS && s <= chunks; ++s) {
    const uint64_t items = (uint64_t)p.groups * s;
    const double eff = (double)items / (double)(simds * ((items + simds - 1) / simds));
    if (eff > best + 0.005) best = eff, p.splits = s;
  }
  return p;
}

size_t pairs_scratch_bytes(const PairsPlan& p, uint32_t n) {
  return p.splits > 1 ? (size_t)p.splits * 3 * n * sizeof(float) : 0;
}

void launch_manybody_pairs(hipStream_t s, uint32_t n, int ndim, float* const* pos, const float* strength,
                           float* const* vel, float dmin2, float dmax2, StepParams step, const PairsPlan& plan,
                           float* part) {
  if (n < 2) return;
  if (plan.groups != ceil_div(n, 64u * PAIRS_BODIES) || plan.splits < 1 || plan.splits > PAIRS_MAX_SPLITS ||
      (plan.splits > 1 && !part))
    throw ArgError(-5, "all-pairs: plan or scratch does not fit the node count");
  dim3 g(plan.groups * plan.splits), b(64);
  const bool dmax = dmax2 < 3e38f;  // 3e38 is infinity as the host passes it
#define FG_PAIRS(D, M)                                                                                  \
  hipLaunchKernelGGL((manybody_pairs<D, PAIRS_BODIES, M>), g, b, 0, s, n, plan.groups, plan.splits, pos[0], pos[1], \
                     pos[2], strength, vel[0], vel[1], vel[2], part, dmin2, dmax2, step)
  if (ndim == 1) {
    if (dmax) FG_PAIRS(1, true); else FG_PAIRS(1, false);
  } else if (ndim == 2) {
    if (dmax) FG_PAIRS(2, true); else FG_PAIRS(2, false);
  } else {
    if (dmax) FG_PAIRS(3, true); else FG_PAIRS(3, false);
  }
#undef FG_PAIRS
  FG_LAUNCH_CHECK();
  if (plan.splits > 1) {
    dim3 cg(ceil_div(n, 256u)), cb(256);
    if (ndim == 1)
      hipLaunchKernelGGL(pairs_combine<1>, cg, cb, 0, s, n, plan.splits, part, vel[0], vel[1], vel[2]);
    else if (ndim == 2)
      hipLaunchKernelGGL(pairs_combine<2>, cg, cb, 0, s, n, plan.splits, part, vel[0], vel[1], vel[2]);
    else
      hipLaunchKernelGGL(pairs_combine<3>, cg, cb, 0, s, n, plan.splits, part, vel[0], vel[1], vel[2]);
    FG_LAUNCH_CHECK();
  }
}

// ---- collide, all pairs -------------------------------------------------------------

template <int D>
__global__ void __launch_bounds__(PAIR_TILE) collide_pairs(uint32_t n, const float* __restrict__ x,
                                                           const float* __restrict__ y,
                                                           const float* __restrict__ z,
                                                           const float* __restrict__ vx,
                                                           const float* __restrict__ vy,
                                                           const float* __restrict__ vz,
                                                           float* __restrict__ ox,
                                                           float* __restrict__ oy,
                                                           float* __restrict__ oz,
                                                           const float* __restrict__ radius,
                                                           float strength,
                                                           StepParams step) {
  __shared__ float4 tile[PAIR_TILE];
  const uint32_t i = blockIdx.x * PAIR_TILE + threadIdx.x;
  const bool live = i < n;
  const uint32_t sn = step.step, seed = step.seed;
  float mx = live ? x[i] + vx[i] : 0.f;
  float my = live && D > 1 ? y[i] + vy[i] : 0.f;
  float mz = live && D > 2 ? z[i] + vz[i] : 0.f;
  float ri = live ? radius[i] : 0.f, ri2 = ri * ri;
  float ax = 0.f, ay = 0.f, az = 0.f;
  for (uint32_t base = 0; base < n; base += PAIR_TILE) {
    uint32_t j = base + threadIdx.x;
    tile[threadIdx.x] = j < n ? make_float4(x[j] + vx[j], D > 1 ? y[j] + vy[j] : 0.f,
                                            D > 2 ? z[j] + vz[j] : 0.f, radius[j])
                              : make_float4(0.f, 0.f, 0.f, 0.f);
    __syncthreads();
    uint32_t count = n - base < (uint32_t)PAIR_TILE ? n - base : (uint32_t)PAIR_TILE;
    for (uint32_t k = 0; k < count; ++k) {
      float4 q = tile[k];
      float rj = q.w, r = ri + rj;
      float dx = mx - q.x, dy = my - q.y, dz = mz - q.z;
      float l = dx * dx + dy * dy + dz * dz;
      if (l >= r * r || base + k == i) continue;
      if (l == 0.f) {
        coincident(i, base + k, D, sn, seed, dx, dy, dz);
        l = dx * dx + dy * dy + dz * dz;
      }
      float sl = sqrtf(l), rj2 = rj * rj;
      float f = (r - sl) / sl * strength * rj2 / (ri2 + rj2);
      ax += dx * f;
      ay += dy * f;
      az += dz * f;
    }
    __syncthreads();
  }
  if (live) {
    ox[i] = vx[i] + ax;
    if (D > 1) oy[i] = vy[i] + ay;
    if (D > 2) oz[i] = vz[i] + az;
  }
}

void launch_collide_pairs(hipStream_t s, uint32_t n, int ndim, float* const* pos,
                          const float* const* vin, float* const* vout, const float* radius,
                          float strength, StepParams step) {
  if (!n) return;
  dim3 g(ceil_div(n, PAIR_TILE)), b(PAIR_TILE);
  if (ndim == 1)
    hipLaunchKernelGGL(collide_pairs<1>, g, b, 0, s, n, pos[0], pos[1], pos[2], vin[0], vin[1], vin[2], vout[0], vout[1], vout[2], radius, strength, step);
  else if (ndim == 2)
    hipLaunchKernelGGL(collide_pairs<2>, g, b, 0, s, n, pos[0], pos[1], pos[2], vin[0], vin[1], vin[2], vout[0], vout[1], vout[2], radius, strength, step);
  else
    hipLaunchKernelGGL(collide_pairs<3>, g, b, 0, s, n, pos[0], pos[1], pos[2], vin[0], vin[1], vin[2], vout[0], vout[1], vout[2], radius, strength, step);
  FG_LAUNCH_CHECK();
}

// ---- collide, grid ------------------------------------------------------------------

__device__ inline uint32_t spread3_3(uint32_t v) {  // 2 bits -> every third bit
  v &= 3u;
  return (v & 1u) | ((v & 2u) << 2);
}
__device__ inline uint32_t spread3_10(uint32_t v) {  // 10 bits -> every third bit
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000ffu;
  v = (v | (v << 8)) & 0x0300f00fu;
  v = (v | (v << 4)) & 0x030c30c3u;
  return (v | (v << 2)) & 0x09249249u;
}
__device__ inline uint32_t spread2_16(uint32_t v) {  // 16 bits -> every other bit
  v &= 0xffffu;
  v = (v | (v << 8)) & 0x00ff00ffu;
  v = (v | (v << 4)) & 0x0f0f0f0fu;
  v = (v | (v << 2)) & 0x33333333u;
  return (v | (v << 1)) & 0x55555555u;
}
__device__ inline uint32_t spread2_3(uint32_t v) {  // 3 bits -> every other bit
  v &= 7u;
  return (v & 1u) | ((v & 2u) << 1) | ((v & 4u) << 2);
}

// Slot of a cell: blocks of 4^3 cells (8^2 in 2D, 64 in 1D) take 64 consecutive slots, the
// cells of a block in Morton order, and the blocks are placed by the Morton code of their
// coordinates' low bits plus an offset hashed from the tag (the class and, in a batch, the
// job's segment), all modulo the table: blocks near each other in space get slots near
// each other, so a point's neighbour cells, and the neighbourhoods of the points one XCD
// queries, share cache lines (1M points, 3D: two collide iterations 0.526 -> 0.479 ms
// against blocks placed by a hash of their coordinates).  Distant blocks that share
// slots only add candidates the pair test rejects.
template <int D>
__device__ inline uint32_t grid_slot(int cx, int cy, int cz, uint32_t seg, uint32_t mask) {
  uint32_t ux = (uint32_t)cx, uy = (uint32_t)cy, uz = (uint32_t)cz;
  uint32_t bx, by, bz, m;
  if (D == 3) {
    bx = ux >> 2, by = uy >> 2, bz = uz >> 2;
    m = (spread3_3(uz) << 2) | (spread3_3(uy) << 1) | spread3_3(ux);
  } else if (D == 2) {
    bx = ux >> 3, by = uy >> 3, bz = 0;
    m = (spread2_3(uy) << 1) | spread2_3(ux);
  } else {
    bx = ux >> 6, by = 0, bz = 0;
    m = ux & 63u;
  }
  uint32_t pos;
  if (D == 3) pos = spread3_10(bx) | spread3_10(by) << 1 | spread3_10(bz) << 2;
  else if (D == 2) pos = spread2_16(bx) | spread2_16(by) << 1;
  else pos = bx;
  pos += pcg_hash(seg * 2654435761u);
  return ((pos << 6) + m) & mask;
}

template <int D>
__device__ inline int3 grid_cell(float4 q, float inv) {
  return make_int3((int)floorf(q.x * inv), D > 1 ? (int)floorf(q.y * inv) : 0,
                   D > 2 ? (int)floorf(q.z * inv) : 0);
}

// The grid's class of a radius: 0 (fine grid) up to |r| = r0, 1 (coarse grid) above.
__device__ inline uint32_t grid_class(float r, const GridClasses& gc) { return fabsf(r) > gc.r0 ? 1u : 0u; }

// The pair of sorted point s with candidate t found under the slot of cell (cx, cy, cz) of
// class c: q is where t was filed (its cell is checked when CELL), qc where it is now.
// Within the ring of 27 cells the check can go: the cells of one class there never share a
// slot (two of them differ by at most one block per axis, so the Morton codes of their
// blocks differ in one of the lowest 3 bits, never by a multiple of the table's blocks),
// and a point filed in a cell elsewhere that shares the slot lies farther away than any
// cell within reach, so the pair test rejects it.
template <int D, bool SEG, bool CELL = true>
__device__ inline void grid_pair(uint32_t t, float4 q, float4 qc, uint32_t qseg, int cx, int cy, int cz, uint32_t c,
                                 float inv, uint32_t s, float4 me, float ri2, uint32_t node,
                                 const uint32_t* __restrict__ svals, uint32_t seg, uint32_t node_off,
                                 const GridClasses& gc, float strength, uint32_t sn, uint32_t seed, float& ax,
                                 float& ay, float& az) {
  // the query inlines this at several sites, and which site takes a candidate depends on
  // its place in the list (so on other jobs' points under shared slots): every operation
  // is spelled out, no contraction left to the compiler, so every site rounds alike
#pragma clang fp contract(off)
  if (t == s) return;
  if (SEG && qseg != seg) return;
  if (grid_class(q.w, gc) != c) return;
  if (CELL) {
    const int3 qcell = grid_cell<D>(q, inv);
    if (qcell.x != cx || qcell.y != cy || qcell.z != cz) return;
  }
  const float rj = q.w, r = me.w + rj;
  float ex = me.x - qc.x, ey = me.y - qc.y, ez = me.z - qc.z;
  float l = __builtin_fmaf(ez, ez, __builtin_fmaf(ey, ey, ex * ex));
  if (l >= r * r) return;
  if (l == 0.f) {
    coincident(node, svals[t] - node_off, D, sn, seed, ex, ey, ez);
    l = __builtin_fmaf(ez, ez, __builtin_fmaf(ey, ey, ex * ex));
  }
  const float len = sqrtf(l), rj2 = rj * rj;
  const float f = (r - len) / len * strength * rj2 / (ri2 + rj2);
  ax = __builtin_fmaf(ex, f, ax);
  ay = __builtin_fmaf(ey, f, ay);
  az = __builtin_fmaf(ez, f, az);
}

template <int D, bool SEG, bool REUSE>
__device__ inline void grid_cell_pairs(GridRun run, int cx, int cy, int cz, uint32_t c, float inv,
                                       uint32_t s, float4 me, float ri2, uint32_t node,
                                       const uint32_t* __restrict__ svals,
                                       const float4* __restrict__ sorted, const float4* __restrict__ cur,
                                       const uint32_t* __restrict__ sseg, uint32_t seg,
                                       uint32_t node_off, const GridClasses& gc, float strength,
                                       uint32_t sn, uint32_t seed, float& ax, float& ay, float& az) {
  for (uint32_t t = run.first; t < run.end; ++t) {
    const float4 q = sorted[t];
    grid_pair<D, SEG>(t, q, REUSE ? cur[t] : q, SEG ? sseg[t] : 0u, cx, cy, cz, c, inv, s, me, ri2, node, svals,
                      seg, node_off, gc, strength, sn, seed, ax, ay, az);
  }
}

// The candidates of the cells of one slab (runs rr[k], cell k at (h.x + k % 3 - 1,
// h.y + k / 3 - RY, cz)) as one list in cell order, GRID_BATCH candidates loaded at a time
// so that their latencies overlap; the pairs are taken in the same order as cell by cell.
constexpr uint32_t GRID_BATCH = 2;

template <int D, bool SEG, bool REUSE, int SLAB>
__device__ inline void grid_slab_pairs(const GridRun (&rr)[SLAB], int3 h, int cz, uint32_t c, float inv, uint32_t s,
                                       float4 me, float ri2, uint32_t node, const uint32_t* __restrict__ svals,
                                       const float4* __restrict__ sorted, const float4* __restrict__ cur,
                                       const uint32_t* __restrict__ sseg,
                                       uint32_t seg, uint32_t node_off, const GridClasses& gc, float strength,
                                       uint32_t sn, uint32_t seed, float& ax, float& ay, float& az) {
  constexpr int RY = D > 1 ? 1 : 0;
  uint32_t cum[SLAB + 1];
  cum[0] = 0;
#pragma unroll
  for (int k = 0; k < SLAB; ++k) cum[k + 1] = cum[k] + (rr[k].end - rr[k].first);
  const uint32_t total = cum[SLAB];
  for (uint32_t j = 0; j < total; j += GRID_BATCH) {
    float4 q[GRID_BATCH], qc[GRID_BATCH];
    uint32_t tt[GRID_BATCH], qs[GRID_BATCH];
    int ox[GRID_BATCH], oy[GRID_BATCH];
#pragma unroll
    for (uint32_t u = 0; u < GRID_BATCH; ++u) {
      const uint32_t jj = j + u;
      uint32_t first = rr[0].first, base = 0;
      int dx = -1, dy = -RY;
#pragma unroll
      for (int k = 1; k < SLAB; ++k) {
        if (cum[k] <= jj) {
          first = rr[k].first;
          base = cum[k];
          dx = k % 3 - 1;
          dy = k / 3 - RY;
        }
      }
      tt[u] = first + (jj - base);
      ox[u] = dx;
      oy[u] = dy;
      q[u] = make_float4(0.f, 0.f, 0.f, 0.f);
      qc[u] = q[u];
      qs[u] = 0u;
      if (jj < total) {
        q[u] = REUSE ? cur[tt[u]] : sorted[tt[u]];  // (no cell check here: only the radius)
        qc[u] = q[u];
        if (SEG) qs[u] = sseg[tt[u]];
      }
    }
#pragma unroll
    for (uint32_t u = 0; u < GRID_BATCH; ++u) {
      if (j + u >= total) break;
      grid_pair<D, SEG, false>(tt[u], q[u], qc[u], qs[u], h.x + ox[u], h.y + oy[u], cz, c, inv, s, me, ri2, node, svals, seg,
                        node_off, gc, strength, sn, seed, ax, ay, az);
    }
  }
}

// Whether cell (cx, cy, cz) of a grid may hold a partner of a point (u: the point in cell
// units): the distance from the point to the cell's box, in cell units, against `reach`
// (the point's |radius| plus the largest |radius| of the cell's class, in cell units).  The
// margin covers the rounding of the cell units and of the pair test, so a cell skipped here
// holds no point the pair test would take.
template <int D>
__device__ inline bool grid_cell_near(float ux, float uy, float uz, int cx, int cy, int cz, float reach,
                                      float slack) {
  const float gx = fmaxf(fmaxf((float)cx - ux, ux - (float)(cx + 1)), 0.f);
  const float gy = D > 1 ? fmaxf(fmaxf((float)cy - uy, uy - (float)(cy + 1)), 0.f) : 0.f;
  const float gz = D > 2 ? fmaxf(fmaxf((float)cz - uz, uz - (float)(cz + 1)), 0.f) : 0.f;
  const float t = reach * 1.0001f + slack;
  return gx * gx + gy * gy + gz * gz <= t * t;
}

// Collide sorted point s (snapshot `me` with its radius, node index `node`) with every
// point closer than the sum of their radii (passes/collide.js:34-80 on candidates from
// passes/collideGrid.js).  Class-0 cells are 2 r0 wide, class-1 cells 2 rmax; a partner
// of class c lies within |me.w| + (c ? rmax : r0), so the cells to visit in class c are
// those within ceil((|me.w| + r_c) / cell_c) of the point's cell -- one ring except for a
// coarse point among fine cells -- and of those only the cells whose box comes that close
// to the point are looked up at all.  Classes, cells and, within a cell, sorted points are
// visited in a fixed order; a point found under a cell's slot counts only if it is of that
// class, of the job (a batch, SEG) and -- beyond the ring of 27 -- filed in that cell, so
// slots shared by cells or jobs change the work but never the sum.  The slot's run of sorted points is
// [start[slot], start[slot + 1]); svals[t] - node_off is sorted point t's node.  In one
// ring, the runs of a z-slab of cells are read together so that their latencies overlap.
//
// REUSE: a later iteration on the grid an earlier one built.  Points stay filed under the
// cells of their positions then (sorted, which the cell check reads); the pair test takes
// their positions now (cur, and `me`), none more than `margin` from where it was filed, so
// the cells to visit are those within |me.w| + r_c + margin of `me` -- the ring of 27
// unless the point lies within that excess of a face of its cell.
template <int D, bool SEG, bool REUSE = false>
__device__ inline void grid_collide_point(uint32_t s, float4 me, uint32_t node,
                                          const uint32_t* __restrict__ svals,
                                          const uint32_t* __restrict__ start,
                                          const float4* __restrict__ sorted,
                                          const uint32_t* __restrict__ sseg, uint32_t seg,
                                          uint32_t node_off, const GridClasses& gc, uint32_t mask,
                                          float strength, uint32_t sn, uint32_t seed, float& ax,
                                          float& ay, float& az, const float4* __restrict__ cur = nullptr,
                                          float margin = 0.f) {
  const float ri2 = me.w * me.w, ri = fabsf(me.w);
  const uint32_t mine = grid_class(me.w, gc);
  const uint32_t classes = gc.two ? 2u : 1u;
  constexpr int RY = D > 1 ? 1 : 0, RZ = D > 2 ? 1 : 0;
  constexpr int SLAB = (2 * RY + 1) * 3;
  for (uint32_t c = 0; c < classes; ++c) {
    const float inv = c ? gc.inv_b : gc.inv_s;
    const int3 h = grid_cell<D>(me, inv);
    const float ux = me.x * inv, uy = D > 1 ? me.y * inv : 0.f, uz = D > 2 ? me.z * inv : 0.f;
    const float slack = 1e-4f + 1e-6f * (fabsf(ux) + fabsf(uy) + fabsf(uz));
    const float reach = (ri + (c ? gc.rmax : gc.r0) + margin) * inv;
    const uint32_t tag = 2u * seg + c;
    int R = (c == 0 && mine == 1) ? (int)ceilf((ri + gc.r0 + margin) * inv * 1.0001f) : 1;
    if (REUSE && R == 1) {
      // the reach past one cell's width: cells two away count when the point is within it
      // of the face between (a cell k away has its near face k - 1 + f or k - f away)
      const float far = reach * 1.0001f + slack - 1.f;
      if (far >= 1.f) {
        R = (int)ceilf(far + 1.f);
      } else if (far > 0.f) {
        const float fx = ux - (float)h.x, fy = uy - (float)h.y, fz = uz - (float)h.z;
        if (fx <= far || fx >= 1.f - far || (D > 1 && (fy <= far || fy >= 1.f - far)) ||
            (D > 2 && (fz <= far || fz >= 1.f - far)))
          R = 2;
      }
    }
    if (R == 1) {
      for (int dz = -RZ; dz <= RZ; ++dz) {
        GridRun rr[SLAB];
#pragma unroll
        for (int k = 0; k < SLAB; ++k) {
          const int cx = h.x + k % 3 - 1, cy = h.y + k / 3 - RY, cz = h.z + dz;
          rr[k] = GridRun{0u, 0u};
          if (grid_cell_near<D>(ux, uy, uz, cx, cy, cz, reach, slack)) {
            const uint32_t slot = grid_slot<D>(cx, cy, cz, tag, mask);
            rr[k] = GridRun{start[slot], start[slot + 1]};
          }
        }
        grid_slab_pairs<D, SEG, REUSE, SLAB>(rr, h, h.z + dz, c, inv, s, me, ri2, node, svals, sorted, cur, sseg, seg,
                                             node_off, gc, strength, sn, seed, ax, ay, az);
      }
    } else {
      const int Ry = D > 1 ? R : 0, Rz = D > 2 ? R : 0;
      for (int dz = -Rz; dz <= Rz; ++dz)
        for (int dy = -Ry; dy <= Ry; ++dy)
          for (int dx = -R; dx <= R; ++dx) {
            const int cx = h.x + dx, cy = h.y + dy, cz = h.z + dz;
            if (!grid_cell_near<D>(ux, uy, uz, cx, cy, cz, reach, slack)) continue;
            const uint32_t slot = grid_slot<D>(cx, cy, cz, tag, mask);
            grid_cell_pairs<D, SEG, REUSE>(GridRun{start[slot], start[slot + 1]}, cx, cy, cz, c, inv, s, me, ri2,
                                           node, svals, sorted, cur, sseg, seg, node_off, gc, strength, sn, seed, ax,
                                           ay, az);
          }
    }
  }
}

// Zeroes the slot counts (a kernel, so that a captured tick can do it), and the words of
// `extra` (or null).  The grid's build kernels take a SkipIf: a later collide iteration
// rebuilds the grid only when the points moved too far for the first grid (grid_run).
__global__ void __launch_bounds__(256) grid_clear(uint32_t words, uint32_t* __restrict__ start,
                                                  uint32_t* __restrict__ extra = nullptr, uint32_t extra_words = 0,
                                                  SkipIf skip = {}) {
  if (skip()) return;
  for (uint32_t k = blockIdx.x * 256 + threadIdx.x; k < words; k += gridDim.x * 256) start[k] = 0u;
  if (extra && blockIdx.x == 0 && threadIdx.x < extra_words) extra[threadIdx.x] = 0u;
}

// A point's slot: its cell in its own class's grid.
template <int D>
__device__ inline uint32_t grid_key(float4 q, uint32_t seg, const GridClasses& gc, uint32_t mask) {
  const uint32_t c = grid_class(q.w, gc);
  const int3 cell = grid_cell<D>(q, c ? gc.inv_b : gc.inv_s);
  return grid_slot<D>(cell.x, cell.y, cell.z, 2u * seg + c, mask);
}

// Snapshot, slot and arrival rank of every node (counting the slots).
template <int D>
__global__ void __launch_bounds__(256) grid_keys(uint32_t n, const float* __restrict__ x,
                                                 const float* __restrict__ y,
                                                 const float* __restrict__ z,
                                                 const float* __restrict__ vx,
                                                 const float* __restrict__ vy,
                                                 const float* __restrict__ vz,
                                                 const float* __restrict__ radius, GridClasses gc,
                                                 uint32_t mask, uint32_t* __restrict__ keys,
                                                 uint32_t* __restrict__ rank, float4* __restrict__ snap,
                                                 uint32_t* __restrict__ count, SkipIf skip) {
  if (skip()) return;
  uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float4 q = make_float4(x[i] + vx[i], D > 1 ? y[i] + vy[i] : 0.f, D > 2 ? z[i] + vz[i] : 0.f,
                               radius[i]);
  const uint32_t key = grid_key<D>(q, 0u, gc, mask);
  keys[i] = key;
  snap[i] = q;
  rank[i] = atomicAdd(&count[key], 1u);
}

// Node i into its slot's run at its arrival rank.
__global__ void __launch_bounds__(256) grid_scatter(uint32_t n, const uint32_t* __restrict__ keys,
                                                    const uint32_t* __restrict__ rank,
                                                    const uint32_t* __restrict__ start,
                                                    uint32_t* __restrict__ arrived, SkipIf skip = {}) {
  if (skip()) return;
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  arrived[start[keys[i]] + rank[i]] = i;
}

// Position t of the arrival order: its node moves to its place by node index within the
// run (the order a stable sort of the slots gives), and its snapshot follows; place[node]
// records the sorted position.  A run of L points costs O(L^2) here, as it does in the
// query.
__global__ void __launch_bounds__(256) grid_order(uint32_t n, const uint32_t* __restrict__ keys,
                                                  const uint32_t* __restrict__ start,
                                                  const uint32_t* __restrict__ arrived,
                                                  const float4* __restrict__ snap,
                                                  uint32_t* __restrict__ svals, float4* __restrict__ sorted,
                                                  const uint32_t* __restrict__ seg_of, uint32_t* __restrict__ sseg,
                                                  uint32_t* __restrict__ place, SkipIf skip = {}) {
  if (skip()) return;
  const uint32_t t = blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  const uint32_t i = arrived[t], key = keys[i];
  const uint32_t first = start[key], end = start[key + 1];
  uint32_t r = 0;
  for (uint32_t u = first; u < end; ++u) r += arrived[u] < i ? 1u : 0u;
  svals[first + r] = i;
  sorted[first + r] = snap[i];
  if (sseg) sseg[first + r] = seg_of[i];
  place[i] = first + r;
}

// The query leaves each sorted point's velocity increment in sorted order (dv, coalesced);
// grid_apply then adds them in node order, one scattered 16-byte read per node where adding
// them in the query took three scattered read-modify-writes (0.468 -> 0.412 ms for two
// iterations at 1M points).
//
// reuse() (a later iteration whose points moved little since the grid was built): the
// points' positions now are cur (sorted order), none more than sqrt(*reuse.word) from where
// they were filed.  A later iteration zeroes *clear, the word the next grid_apply takes its
// maximum in.
template <int D>
__global__ void __launch_bounds__(256) grid_query(uint32_t n, const uint32_t* __restrict__ svals,
                                                  const uint32_t* __restrict__ start,
                                                  const float4* __restrict__ sorted,
                                                  float4* __restrict__ dv, GridClasses gc,
                                                  uint32_t mask, float strength,
                                                  StepParams step, const float4* __restrict__ cur,
                                                  SkipIf reuse, uint32_t* __restrict__ clear) {
  if (clear && blockIdx.x == 0 && threadIdx.x == 0) *clear = 0u;
  // each XCD queries one contiguous eighth of the sorted points (0.479 -> 0.468 ms)
  const uint32_t s = xcd_block() * 256 + threadIdx.x;
  if (s >= n) return;
  const uint32_t i = svals[s];
  float ax = 0.f, ay = 0.f, az = 0.f;
  if (reuse()) {
    // (the bound a little above the largest move, for the rounding of its square root)
    const float margin = sqrtf(__uint_as_float(*reuse.word)) * 1.0001f + 1e-30f;
    grid_collide_point<D, false, true>(s, cur[s], i, svals, start, sorted, nullptr, 0u, 0u, gc, mask,
                                       strength, step.step, step.seed, ax, ay, az, cur, margin);
  } else {
    grid_collide_point<D, false, false>(s, sorted[s], i, svals, start, sorted, nullptr, 0u, 0u, gc, mask,
                                        strength, step.step, step.seed, ax, ay, az);
  }
  dv[s] = make_float4(ax, ay, az, 0.f);
}

// Node i's increment, from its sorted position place[i].  With `cur` (another iteration
// follows on this grid): the node's position for that iteration (p + v, as a new grid's
// snapshot would take it) goes to cur[place[i]], and the largest squared distance of
// those positions from the ones the points were filed at (snap) to *margin2 (by workgroup;
// the float's bits order as the floats do, for non-negative floats).
template <int D>
__global__ void __launch_bounds__(256) grid_apply(uint32_t n, const uint32_t* __restrict__ place,
                                                  const float4* __restrict__ dv, float* __restrict__ vx,
                                                  float* __restrict__ vy, float* __restrict__ vz,
                                                  const float* __restrict__ x, const float* __restrict__ y,
                                                  const float* __restrict__ z, const float4* __restrict__ snap,
                                                  float4* __restrict__ cur, uint32_t* __restrict__ margin2) {
  float m = 0.f;
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const uint32_t t = place[i];
    const float4 d = dv[t];
    const float nx = vx[i] + d.x, ny = D > 1 ? vy[i] + d.y : 0.f, nz = D > 2 ? vz[i] + d.z : 0.f;
    vx[i] = nx;
    if (D > 1) vy[i] = ny;
    if (D > 2) vz[i] = nz;
    if (cur) {
      const float4 p1 = snap[i];
      const float4 p = make_float4(x[i] + nx, D > 1 ? y[i] + ny : 0.f, D > 2 ? z[i] + nz : 0.f, p1.w);
      cur[t] = p;
      const float ex = p.x - p1.x, ey = p.y - p1.y, ez = p.z - p1.z;
      m = fmaxf(m, ex * ex + ey * ey + ez * ez);
    }
  }
  if (cur) {
    __shared__ float s_max[256 / WAVE];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, WAVE));
    if ((threadIdx.x & (WAVE - 1)) == 0) s_max[threadIdx.x / WAVE] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
      float w = s_max[0];
      for (uint32_t k = 1; k < 256 / WAVE; ++k) w = fmaxf(w, s_max[k]);
      atomicMax(margin2, __float_as_uint(w));
    }
  }
}

void GridWork::reserve(uint32_t n_) {
  uint32_t b = 8;
  while ((1ull << b) < 2ull * n_) ++b;
  if (n_ <= cap && bits == b && keys) return;
  release();
  uint32_t m = n_ > 1 ? n_ : 1;
  bits = b;
  const uint32_t words = (1u << bits) + 1u;
  FG_HIP(hipMalloc(&keys, sizeof(uint32_t) * m));
  FG_HIP(hipMalloc(&rank, sizeof(uint32_t) * m));
  FG_HIP(hipMalloc(&start, sizeof(uint32_t) * words));
  FG_HIP(hipMalloc(&partial, sizeof(uint32_t) * scan_partials_size(words)));
  FG_HIP(hipMalloc(&arrived, sizeof(uint32_t) * m));
  FG_HIP(hipMalloc(&svals, sizeof(uint32_t) * m));
  FG_HIP(hipMalloc(&sorted, sizeof(float4) * m));
  FG_HIP(hipMalloc(&snap, sizeof(float4) * m));
  FG_HIP(hipMalloc(&dv, sizeof(float4) * m));
  FG_HIP(hipMalloc(&cur, sizeof(float4) * m));
  FG_HIP(hipMalloc(&margin2, sizeof(uint32_t) * 2));
  cap = n_;
}

void GridWork::release() {
  void* ptrs[] = {keys, rank, start, partial, arrived, svals, sorted, snap, dv, cur, margin2};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  keys = rank = start = partial = arrived = svals = margin2 = nullptr;
  sorted = snap = dv = cur = nullptr;
  cap = 0;
}

// Iteration k of a collide force with `iterations` iterations.  The first builds the grid by
// counting: clear the counts, slot + rank per node, scan, scatter by rank, reorder runs by
// node index (with the snapshots; rank then holds each node's sorted position).  A later one
// queries that grid again when no point has moved more than a quarter of the smaller cell
// width since it was filed (the largest move is on the device, so its build kernels are
// launched and return at once; the cells to visit then lie within two of the point's cell):
// the pairs are those of the positions now, only which candidates are looked at and the
// order the sum takes them in depend on the grid.  Then the query, and the increments added
// in node order with the next iteration's positions when one follows.  At 1M points the
// second of two iterations took 208 us with its own grid.
template <int D>
void grid_run(hipStream_t s, GridWork& g, uint32_t n, float* const* pos, float* const* vel,
              const float* radius, const GridClasses& gc, uint32_t mask, float strength, StepParams step,
              int k, int iterations) {
  dim3 gr(ceil_div(n, 256)), b(256);
  const uint32_t words = (1u << g.bits) + 1u;
  const bool next = k + 1 < iterations;
  // iteration k >= 1 reads the largest squared move that iteration k - 1 left in
  // margin2[k & 1]; iteration k's own goes to the other word, zeroed first (by the clear at
  // k = 0, by the query after)
  SkipIf reuse;
  if (k > 0) {
    const float w = 1.0f / std::max(gc.inv_s, gc.inv_b), lim = 0.25f * w;
    reuse = SkipIf{g.margin2 + (k & 1), lim * lim};
  }
  hipLaunchKernelGGL(grid_clear, dim3(std::min(ceil_div(words, 256), 2048u)), b, 0, s, words, g.start,
                     k == 0 ? g.margin2 + 1 : (uint32_t*)nullptr, 1u, reuse);
  hipLaunchKernelGGL(grid_keys<D>, gr, b, 0, s, n, pos[0], pos[1], pos[2], vel[0], vel[1], vel[2],
                     radius, gc, mask, g.keys, g.rank, g.snap, g.start, reuse);
  FG_LAUNCH_CHECK();
  exclusive_scan_u32(s, g.start, words, g.partial, reuse);
  hipLaunchKernelGGL(grid_scatter, gr, b, 0, s, n, (const uint32_t*)g.keys, (const uint32_t*)g.rank,
                     (const uint32_t*)g.start, g.arrived, reuse);
  hipLaunchKernelGGL(grid_order, gr, b, 0, s, n, (const uint32_t*)g.keys, (const uint32_t*)g.start,
                     (const uint32_t*)g.arrived, (const float4*)g.snap, g.svals, g.sorted,
                     (const uint32_t*)nullptr, (uint32_t*)nullptr, g.rank, reuse);
  FG_LAUNCH_CHECK();
  hipLaunchKernelGGL(grid_query<D>, dim3(xcd_blocks(gr.x)), b, 0, s, n, (const uint32_t*)g.svals,
                     (const uint32_t*)g.start, (const float4*)g.sorted, g.dv, gc, mask, strength, step,
                     (const float4*)g.cur, reuse, k > 0 ? g.margin2 + ((k + 1) & 1) : (uint32_t*)nullptr);
  FG_LAUNCH_CHECK();
  hipLaunchKernelGGL(grid_apply<D>, dim3(std::min(gr.x, 1024u)), b, 0, s, n, (const uint32_t*)g.rank,
                     (const float4*)g.dv, vel[0], vel[1], vel[2], (const float*)pos[0], (const float*)pos[1],
                     (const float*)pos[2], (const float4*)g.snap, next ? g.cur : (float4*)nullptr,
                     g.margin2 + ((k + 1) & 1));
  FG_LAUNCH_CHECK();
}

void launch_collide_grid(hipStream_t s, GridWork& g, uint32_t n, int ndim, float* const* pos,
                         float* const* vel, const float* radius, const GridClasses& gc, float strength,
                         int radix_bits, StepParams step, int k, int iterations) {
  if (!n || !(gc.inv_b > 0)) return;
  g.reserve(n);
  uint32_t mask = (uint32_t)((1ull << g.bits) - 1ull);
  (void)radix_bits;  // the grid is built by counting, not sorted
  if (ndim == 1) grid_run<1>(s, g, n, pos, vel, radius, gc, mask, strength, step, k, iterations);
  else if (ndim == 2) grid_run<2>(s, g, n, pos, vel, radius, gc, mask, strength, step, k, iterations);
  else grid_run<3>(s, g, n, pos, vel, radius, gc, mask, strength, step, k, iterations);
}

GridClasses grid_classes(const float* radii, uint32_t n) {
  GridClasses gc{0.0f, 0.0f, 0.0f, 0, 0.0f};
  float rmax = 0;
  for (uint32_t i = 0; i < n; ++i) rmax = std::max(rmax, std::fabs(radii[i]));
  if (!(rmax > 0)) return gc;
  gc.inv_b = gc.inv_s = 1.0f / (2.0f * rmax);
  gc.r0 = rmax;
  gc.rmax = rmax;
  // experiment overrides: one class; the quantile of the fine class and the ratio above which
  // a coarse class is split off
  static const bool one_class = getenv("D3FGPU_GRID_ONE_CLASS") != nullptr;
  static const double quant = getenv("D3FGPU_GRID_QUANTILE") ? atof(getenv("D3FGPU_GRID_QUANTILE")) : 0.95;
  static const double split = getenv("D3FGPU_GRID_SPLIT") ? atof(getenv("D3FGPU_GRID_SPLIT")) : 4.0;
  if (n >= 1024 && !one_class) {
    std::vector<float> r(n);
    for (uint32_t i = 0; i < n; ++i) r[i] = std::fabs(radii[i]);
    const size_t k = (size_t)(quant * (double)(n - 1));
    std::nth_element(r.begin(), r.begin() + k, r.end());
    const float rq = r[k];
    if (rq > 0 && rmax > (float)split * rq) {
      gc.two = 1;
      gc.r0 = rq;
      gc.inv_s = 1.0f / (2.0f * rq);
    }
  }
  return gc;
}

}  // namespace fg

// ---- batched passes -------------------------------------------------------------------
//
// The same per-node work over the node space (or grid space) of a batch: each index finds
// its job through its segment, works on the job's own arrays with job-local indices and
// the job's own alpha, step and seed, and leaves jobs that are not running alone.

namespace fg {

namespace {

// f(B, seg) with the record B of segment seg's job.  When every active lane of the wave is
// in one segment -- nearly always: a segment is a whole job of thousands of nodes -- the
// record is read through a pointer the wave shares (scalar loads, once for the wave), else
// lane by lane.
template <class F>
__device__ inline void with_job(uint32_t seg, const uint32_t* __restrict__ seg_job, const BJob* __restrict__ jobs,
                                F&& f) {
  const uint32_t s0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)seg);
  if (__builtin_amdgcn_ballot_w64(seg != s0) == 0) f(jobs[seg_job[s0]], s0);
  else f(jobs[seg_job[seg]], seg);
}

// Node i of a batch: f(B, v) with its job's record B and its index v within the job.
template <class F>
__device__ inline void with_node(uint32_t i, const uint32_t* __restrict__ seg_of,
                                 const uint32_t* __restrict__ seg_job, const BJob* __restrict__ jobs,
                                 uint32_t BJob::*off, F&& f) {
  with_job(seg_of[i], seg_job, jobs, [&](const BJob& B, uint32_t) { f(B, i - B.*off); });
}

// d3's alpha step (simulation.js:54-73) for every job, in FP64 without contraction so the
// sequence is the host's bit for bit; also this tick's step number and whether it runs.
__global__ void __launch_bounds__(256) badvance(BJob* jobs, uint32_t njobs) {
  const uint32_t j = blockIdx.x * 256 + threadIdx.x;
  if (j >= njobs) return;
  BJob& B = jobs[j];
  if (B.status == BJOB_RUNNING) {
    // two roundings, as JavaScript computes it: no fused multiply-add
#pragma clang fp contract(off)
    const double a = B.alpha + (B.alpha_target - B.alpha) * B.alpha_decay;
    B.alpha = a;
    B.alpha_f = (float)a;
    B.step = B.ticks;
    B.ticks += 1;
    B.active = 1;
    if (a < B.alpha_min || (B.max_ticks && B.ticks >= B.max_ticks)) B.status = BJOB_FINISHING;
  } else {
    if (B.status == BJOB_FINISHING) B.status = BJOB_DONE;
    B.active = 0;
  }
}

template <int D>
__global__ void __launch_bounds__(256) bsnapshot(uint32_t total, const uint32_t* __restrict__ seg_of,
                                                 const uint32_t* __restrict__ seg_job,
                                                 const BJob* __restrict__ jobs, uint32_t iteration,
                                                 float4* __restrict__ snap) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= total) return;
  with_node(i, seg_of, seg_job, jobs, &BJob::node_off, [&](const BJob& B, uint32_t v) {
    if (!B.active || !(B.forces & BF_LINK) || iteration >= B.link_iters) return;
    const uint32_t c = iteration & 1u;
    snap[i] = make_float4(B.p[0][v] + B.v[c][0][v], D > 1 ? B.p[1][v] + B.v[c][1][v] : 0.f,
                          D > 2 ? B.p[2][v] + B.v[c][2][v] : 0.f, 0.f);
  });
}

template <int D>
__global__ void __launch_bounds__(256) blink(uint32_t total, const uint32_t* __restrict__ seg_of,
                                             const uint32_t* __restrict__ seg_job,
                                             const BJob* __restrict__ jobs, uint32_t iteration,
                                             const float4* __restrict__ snap) {
  // every lane stays (long rows take the wave); a lane past the end reads the last node's job
  const uint32_t i = blockIdx.x * 256 + threadIdx.x, k = i < total ? i : total - 1;
  const BJob& B = jobs[seg_job[seg_of[k]]];
  const bool live = i < total && B.active && (B.forces & BF_LINK) && iteration < B.link_iters;
  const uint32_t c = iteration & 1u;
  link_wave<D>(k - B.node_off, live, SnapArray{snap + B.node_off}, B.v[c][0], B.v[c][1], B.v[c][2], B.v[c ^ 1][0],
               B.v[c ^ 1][1], B.v[c ^ 1][2], B.link_row, B.link_ent, B.alpha_f, B.step, B.seed);
}

template <int D>
__global__ void __launch_bounds__(256) bgrid_keys(uint32_t total, const uint32_t* __restrict__ seg_of,
                                                  const uint32_t* __restrict__ seg_job,
                                                  const BJob* __restrict__ jobs, uint32_t mask,
                                                  uint32_t* __restrict__ keys, uint32_t* __restrict__ rank,
                                                  float4* __restrict__ snap, uint32_t* __restrict__ count) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= total) return;
  with_job(seg_of[i], seg_job, jobs, [&](const BJob& B, uint32_t seg) {
    const uint32_t v = i - B.grid_off, c = B.vcur;
    const float4 q = make_float4(B.p[0][v] + B.v[c][0][v], D > 1 ? B.p[1][v] + B.v[c][1][v] : 0.f,
                                 D > 2 ? B.p[2][v] + B.v[c][2][v] : 0.f, B.radius[v]);
    const uint32_t key = grid_key<D>(q, seg, B.coll, mask);
    keys[i] = key;
    snap[i] = q;
    rank[i] = atomicAdd(&count[key], 1u);
  });
}

template <int D>
__global__ void __launch_bounds__(256) bgrid_query(uint32_t total, const uint32_t* __restrict__ svals,
                                                   const uint32_t* __restrict__ start,
                                                   const float4* __restrict__ sorted,
                                                   const uint32_t* __restrict__ sseg,
                                                   const uint32_t* __restrict__ seg_job,
                                                   const BJob* __restrict__ jobs, uint32_t mask,
                                                   uint32_t iteration, float4* __restrict__ dv) {
  const uint32_t s = xcd_block() * 256 + threadIdx.x;
  if (s >= total) return;
  const uint32_t seg = sseg[s];
  const BJob& B = jobs[seg_job[seg]];
  if (!B.active || iteration >= B.coll_iters) return;
  const uint32_t v = svals[s] - B.grid_off;
  float ax = 0.f, ay = 0.f, az = 0.f;
  grid_collide_point<D, true>(s, sorted[s], v, svals, start, sorted, sseg, seg,
                              B.grid_off, B.coll, mask, B.coll_strength, B.step, B.seed, ax, ay,
                              az);
  dv[s] = make_float4(ax, ay, az, 0.f);
}

// Node i of the batch (global index) takes its increment from its sorted position, as
// grid_apply, if its job collides in this iteration.
template <int D>
__global__ void __launch_bounds__(256) bgrid_apply(uint32_t total, const uint32_t* __restrict__ seg_of,
                                                   const uint32_t* __restrict__ seg_job,
                                                   const BJob* __restrict__ jobs, uint32_t iteration,
                                                   const uint32_t* __restrict__ place,
                                                   const float4* __restrict__ dv) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= total) return;
  with_node(i, seg_of, seg_job, jobs, &BJob::grid_off, [&](const BJob& B, uint32_t v) {
    if (!B.active || iteration >= B.coll_iters) return;
    const float4 d = dv[place[i]];
    float* const* vel = B.v[B.vcur];
    vel[0][v] += d.x;
    if (D > 1) vel[1][v] += d.y;
    if (D > 2) vel[2][v] += d.z;
  });
}

// Centre: FP64 sums per tile (job-relative tiles, fixed order inside a tile); the last tile
// of a job adds its job's tile sums in order and forms the shift (passes/center.js:11-72).
__global__ void __launch_bounds__(256) bcenter_reduce(const BTile* __restrict__ tiles,
                                                      const uint32_t* __restrict__ seg_job,
                                                      BJob* jobs, double* partial) {
  const BTile t = tiles[blockIdx.x];
  BJob* B = jobs + seg_job[t.seg];
  if (!B->active || !(B->forces & BF_CENTER)) return;
  const uint32_t nd = B->ndim, off = B->node_off;
  double s[3] = {0, 0, 0};
  for (uint32_t i = t.first + threadIdx.x; i < t.first + t.count; i += 256) {
    const uint32_t v = i - off;
    s[0] += (double)B->p[0][v];
    if (nd > 1) s[1] += (double)B->p[1][v];
    if (nd > 2) s[2] += (double)B->p[2][v];
  }
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s[a] += __shfl_xor(s[a], o, WAVE);
  __shared__ double sh[4][3];
  __shared__ uint32_t s_last;
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0)
    for (int a = 0; a < 3; ++a) sh[wave][a] = s[a];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int a = 0; a < 3; ++a)
      store_coherent(&partial[3 * blockIdx.x + a], sh[0][a] + sh[1][a] + sh[2][a] + sh[3][a]);
    s_last = publish_count(&B->center_done) + 1 == B->node_tiles;
  }
  __syncthreads();
  if (!s_last || threadIdx.x != 0) return;
  double tot[3] = {0, 0, 0};
  for (uint32_t k = 0; k < B->node_tiles; ++k)
    for (int a = 0; a < 3; ++a) tot[a] += load_coherent(&partial[3 * (B->node_tiles_first + k) + a]);
  for (int a = 0; a < 3; ++a)
    B->center_shift[a] = a < (int)nd ? (float)((tot[a] / B->n - B->center[a]) * B->center_strength) : 0.0f;
  B->center_done = 0;
}

__global__ void __launch_bounds__(256) bcenter_apply(uint32_t total, const uint32_t* __restrict__ seg_of,
                                                     const uint32_t* __restrict__ seg_job,
                                                     const BJob* __restrict__ jobs) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= total) return;
  with_node(i, seg_of, seg_job, jobs, &BJob::node_off, [&](const BJob& B, uint32_t v) {
    if (!B.active || !(B.forces & BF_CENTER)) return;
    for (uint32_t a = 0; a < B.ndim; ++a)
      if (B.center_shift[a] != 0.0f) B.p[a][v] -= B.center_shift[a];
  });
}

template <int D>
__global__ void __launch_bounds__(256) bradial(uint32_t total, const uint32_t* __restrict__ seg_of,
                                               const uint32_t* __restrict__ seg_job,
                                               const BJob* __restrict__ jobs) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= total) return;
  with_node(i, seg_of, seg_job, jobs, &BJob::node_off, [&](const BJob& B, uint32_t v) {
  if (!B.active || !(B.forces & BF_RADIAL)) return;
  float dx = B.p[0][v] - B.radial_c[0];
  float dy = D > 1 ? B.p[1][v] - B.radial_c[1] : 0.f;
  float dz = D > 2 ? B.p[2][v] - B.radial_c[2] : 0.f;
  if (dx == 0.f) dx = 1e-6f;
  if (D > 1 && dy == 0.f) dy = 1e-6f;
  if (D > 2 && dz == 0.f) dz = 1e-6f;
  const float rr = sqrtf(dx * dx + dy * dy + dz * dz);
  const float k = (B.radial_r[v] - rr) * B.radial_w[v] * B.alpha_f / rr;
  float* const* vel = B.v[B.vcur];
  vel[0][v] += dx * k;
  if (D > 1) vel[1][v] += dy * k;
  if (D > 2) vel[2][v] += dz * k;
  });
}

__global__ void __launch_bounds__(256) bposition(uint32_t total, const uint32_t* __restrict__ seg_of,
                                                 const uint32_t* __restrict__ seg_job,
                                                 const BJob* __restrict__ jobs, int axis) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= total) return;
  with_node(i, seg_of, seg_job, jobs, &BJob::node_off, [&](const BJob& B, uint32_t v) {
    if (!B.active || !(B.forces & (BF_POS_X << axis)) || axis >= (int)B.ndim) return;
    B.v[B.vcur][axis][v] += (B.pos_goal[axis][v] - B.p[axis][v]) * B.pos_weight[axis][v] * B.alpha_f;
  });
}

// v *= decay; p += v; pinned axes p = fix, v = 0 (wgsl.js:67-84).  Reads the job's current
// velocity set and writes set 0, so every tick of a job starts from set 0.
template <int D>
__global__ void __launch_bounds__(256) bintegrate(uint32_t total, const uint32_t* __restrict__ seg_of,
                                                  const uint32_t* __restrict__ seg_job,
                                                  const BJob* __restrict__ jobs) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= total) return;
  with_node(i, seg_of, seg_job, jobs, &BJob::node_off, [&](const BJob& B, uint32_t v) {
    if (!B.active) return;
    const uint32_t c = B.vcur;
    const uint32_t m = B.mask ? B.mask[v] : 0u;
    const float decay = B.decay;
#pragma unroll
    for (int a = 0; a < D; ++a) {
      float vv = B.v[c][a][v] * decay, p = B.p[a][v] + vv;
      if ((m >> a) & 1u) p = B.fix[a][v], vv = 0.0f;
      B.p[a][v] = p;
      B.v[0][a][v] = vv;
    }
  });
}

}  // namespace

void BatchGrid::reserve(uint32_t n_) {
  uint32_t b = 8;
  while ((1ull << b) < 2ull * n_) ++b;
  if (n_ <= cap && bits == b && keys) return;
  release();
  uint32_t m = n_ > 1 ? n_ : 1;
  bits = b;
  const uint32_t words = (1u << bits) + 1u;
  FG_HIP(hipMalloc(&keys, sizeof(uint32_t) * m));
  FG_HIP(hipMalloc(&rank, sizeof(uint32_t) * m));
  FG_HIP(hipMalloc(&start, sizeof(uint32_t) * words));
  FG_HIP(hipMalloc(&partial, sizeof(uint32_t) * scan_partials_size(words)));
  FG_HIP(hipMalloc(&arrived, sizeof(uint32_t) * m));
  FG_HIP(hipMalloc(&svals, sizeof(uint32_t) * m));
  FG_HIP(hipMalloc(&sorted, sizeof(float4) * m));
  FG_HIP(hipMalloc(&snap, sizeof(float4) * m));
  FG_HIP(hipMalloc(&sseg, sizeof(uint32_t) * m));
  cap = n_;
}

void BatchGrid::release() {
  void* ptrs[] = {keys, rank, start, partial, arrived, svals, sorted, snap, sseg};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  keys = rank = start = partial = arrived = svals = sseg = nullptr;
  sorted = snap = nullptr;
  cap = 0;
}

void BatchWork::reserve(uint32_t tiles) {
  if (tiles <= center_cap && center_partial) return;
  release();
  uint32_t m = tiles > 1 ? tiles : 1;
  FG_HIP(hipMalloc(&center_partial, sizeof(double) * 3 * m));
  center_cap = m;
}

void BatchWork::release() {
  if (center_partial) (void)hipFree(center_partial);
  center_partial = nullptr;
  center_cap = 0;
}

void batch_advance(hipStream_t s, BJob* jobs, uint32_t njobs, uint32_t*) {
  if (!njobs) return;
  hipLaunchKernelGGL(badvance, dim3(ceil_div(njobs, 256)), dim3(256), 0, s, jobs, njobs);
  FG_LAUNCH_CHECK();
}

#define FG_BY_DIM(ndim, KERNEL, ...)                                                   \
  do {                                                                                 \
    if ((ndim) == 1) hipLaunchKernelGGL(KERNEL<1>, __VA_ARGS__);                       \
    else if ((ndim) == 2) hipLaunchKernelGGL(KERNEL<2>, __VA_ARGS__);                  \
    else hipLaunchKernelGGL(KERNEL<3>, __VA_ARGS__);                                   \
    FG_LAUNCH_CHECK();                                                                 \
  } while (0)

void batch_link(hipStream_t s, int ndim, const BSpace& sp, BJob* jobs, uint32_t iteration, float4* snap) {
  if (!sp.total) return;
  FG_BY_DIM(ndim, bsnapshot, dim3(ceil_div(sp.total, 256)), dim3(256), 0, s, sp.total,
            (const uint32_t*)sp.seg_of, (const uint32_t*)sp.seg_job, (const BJob*)jobs, iteration, snap);
  FG_BY_DIM(ndim, blink, dim3(ceil_div(sp.total, 256)), dim3(256), 0, s, sp.total,
            (const uint32_t*)sp.seg_of, (const uint32_t*)sp.seg_job, (const BJob*)jobs, iteration,
            (const float4*)snap);
}

void batch_collide(hipStream_t s, int ndim, BatchGrid& g, const BSpace& sp, BJob* jobs,
                   uint32_t iteration, int radix_bits) {
  (void)radix_bits;  // the grid is built by counting, not sorted
  const uint32_t n = sp.total;
  if (!n) return;
  g.reserve(n);
  const uint32_t mask = (uint32_t)((1ull << g.bits) - 1ull);
  const uint32_t words = (1u << g.bits) + 1u;
  dim3 gr(ceil_div(n, 256)), b(256);
  hipLaunchKernelGGL(grid_clear, dim3(std::min(ceil_div(words, 256), 2048u)), b, 0, s, words, g.start);
  FG_BY_DIM(ndim, bgrid_keys, gr, b, 0, s, n, (const uint32_t*)sp.seg_of, (const uint32_t*)sp.seg_job,
            (const BJob*)jobs, mask, g.keys, g.rank, g.snap, g.start);
  exclusive_scan_u32(s, g.start, words, g.partial);
  hipLaunchKernelGGL(grid_scatter, gr, b, 0, s, n, (const uint32_t*)g.keys, (const uint32_t*)g.rank,
                     (const uint32_t*)g.start, g.arrived);
  hipLaunchKernelGGL(grid_order, gr, b, 0, s, n, (const uint32_t*)g.keys, (const uint32_t*)g.start,
                     (const uint32_t*)g.arrived, (const float4*)g.snap, g.svals, g.sorted,
                     (const uint32_t*)sp.seg_of, g.sseg, g.rank);
  FG_LAUNCH_CHECK();
  FG_BY_DIM(ndim, bgrid_query, dim3(xcd_blocks(gr.x)), b, 0, s, n, (const uint32_t*)g.svals, (const uint32_t*)g.start,
            (const float4*)g.sorted, (const uint32_t*)g.sseg, (const uint32_t*)sp.seg_job, (const BJob*)jobs, mask,
            iteration, g.snap);
  FG_BY_DIM(ndim, bgrid_apply, gr, b, 0, s, n, (const uint32_t*)sp.seg_of, (const uint32_t*)sp.seg_job,
            (const BJob*)jobs, iteration, (const uint32_t*)g.rank, (const float4*)g.snap);
}

void batch_center(hipStream_t s, int, BatchWork& w, const BSpace& sp, BJob* jobs) {
  if (!sp.total || !sp.tiles) return;
  w.reserve(sp.tiles);
  hipLaunchKernelGGL(bcenter_reduce, dim3(sp.tiles), dim3(256), 0, s, (const BTile*)sp.tile,
                     (const uint32_t*)sp.seg_job, jobs, w.center_partial);
  FG_LAUNCH_CHECK();
  hipLaunchKernelGGL(bcenter_apply, dim3(ceil_div(sp.total, 256)), dim3(256), 0, s, sp.total,
                     (const uint32_t*)sp.seg_of, (const uint32_t*)sp.seg_job, (const BJob*)jobs);
  FG_LAUNCH_CHECK();
}

void batch_radial(hipStream_t s, int ndim, const BSpace& sp, BJob* jobs) {
  if (!sp.total) return;
  FG_BY_DIM(ndim, bradial, dim3(ceil_div(sp.total, 256)), dim3(256), 0, s, sp.total,
            (const uint32_t*)sp.seg_of, (const uint32_t*)sp.seg_job, (const BJob*)jobs);
}

void batch_position(hipStream_t s, const BSpace& sp, BJob* jobs, int axis) {
  if (!sp.total) return;
  hipLaunchKernelGGL(bposition, dim3(ceil_div(sp.total, 256)), dim3(256), 0, s, sp.total,
                     (const uint32_t*)sp.seg_of, (const uint32_t*)sp.seg_job, (const BJob*)jobs, axis);
  FG_LAUNCH_CHECK();
}

void batch_integrate(hipStream_t s, int ndim, const BSpace& sp, BJob* jobs) {
  if (!sp.total) return;
  FG_BY_DIM(ndim, bintegrate, dim3(ceil_div(sp.total, 256)), dim3(256), 0, s, sp.total,
            (const uint32_t*)sp.seg_of, (const uint32_t*)sp.seg_job, (const BJob*)jobs);
}

#undef FG_BY_DIM

}  // namespace fg

// ---- link ingestion on the device -------------------------------------------------------
//
// The CSR of incident links built where it is used (passes/link.js:65-101, src/link.js:
// 56-97): entry e = 2k + side of link k belongs to node links[e]; a stable radix sort of
// (node, e) groups each node's entries in link order (what a host build in link order
// gives), row offsets come from a binary search of the sorted nodes, and every entry takes
// the link's other end and its distance, strength and bias -- d3's defaults (30, 1 / min
// of the two ends' link counts, count[source] / (count[source] + count[target])) where the
// descriptor has no array.

namespace fg {

namespace {

__global__ void __launch_bounds__(256) link_check(uint32_t m2, uint32_t n, const uint32_t* __restrict__ links,
                                                  uint32_t* __restrict__ bad) {
  const uint32_t e = blockIdx.x * 256 + threadIdx.x;
  if (e < m2 && links[e] >= n) atomicOr(bad, 1u);
}

__global__ void __launch_bounds__(256) link_pairs(uint32_t m2, const uint32_t* __restrict__ links,
                                                  uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const uint32_t e = blockIdx.x * 256 + threadIdx.x;
  if (e >= m2) return;
  keys[e] = links[e];
  vals[e] = e;
}

// row[i] = the first sorted entry of node i (row[n] = 2m).
__global__ void __launch_bounds__(256) link_rows(uint32_t n, uint32_t m2, const uint32_t* __restrict__ snodes,
                                                 uint32_t* __restrict__ row) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i > n) return;
  uint32_t lo = 0, hi = m2;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (snodes[mid] < i) lo = mid + 1;
    else hi = mid;
  }
  row[i] = lo;
}

__device__ inline float finite_dev(double v) {
  if (!(v == v)) return 0.0f;
  if (v > 3e38) return 3e38f;
  if (v < -3e38) return -3e38f;
  return (float)v;
}

__global__ void __launch_bounds__(256) link_entries(uint32_t m2, const uint32_t* __restrict__ links,
                                                    const uint32_t* __restrict__ scode,
                                                    const uint32_t* __restrict__ row,
                                                    const double* __restrict__ strengths,
                                                    const double* __restrict__ distances,
                                                    const double* __restrict__ bias,
                                                    LinkEntry* __restrict__ ent) {
  const uint32_t s = blockIdx.x * 256 + threadIdx.x;
  if (s >= m2) return;
  const uint32_t code = scode[s], k = code >> 1;
  const uint32_t src = links[2 * k], tgt = links[2 * k + 1];
  const float cs = (float)(row[src + 1] - row[src]), ct = (float)(row[tgt + 1] - row[tgt]);
  LinkEntry L;
  L.other = (code & 1u) ? src : tgt;
  L.code = code;
  L.distance = distances ? finite_dev(distances[k]) : 30.0f;
  L.strength = strengths ? finite_dev(strengths[k]) : 1.0f / fminf(cs, ct);
  L.bias = bias ? finite_dev(bias[k]) : cs / (cs + ct);
  ent[s] = L;
}

}  // namespace

void link_csr_device(hipStream_t s, uint32_t n, uint32_t m, const uint32_t* links, const double* strengths,
                     const double* distances, const double* bias, uint32_t* row, LinkEntry* ent) {
  const uint32_t m2 = 2 * m;
  uint32_t* bad = nullptr;
  uint32_t* keys = nullptr;
  uint32_t* vals = nullptr;
  SortWork w;
  try {
    FG_HIP(hipMalloc(&bad, sizeof(uint32_t)));
    FG_HIP(hipMemsetAsync(bad, 0, sizeof(uint32_t), s));
    if (m2) {
      hipLaunchKernelGGL(link_check, dim3(ceil_div(m2, 256)), dim3(256), 0, s, m2, n, links, bad);
      FG_LAUNCH_CHECK();
    }
    uint32_t h_bad = 0;
    FG_HIP(hipMemcpyAsync(&h_bad, bad, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    FG_HIP(hipStreamSynchronize(s));
    if (h_bad) throw ArgError(-1, "link: node index out of range");
    uint32_t* snodes = nullptr;
    uint32_t* scode = nullptr;
    if (m2) {
      FG_HIP(hipMalloc(&keys, sizeof(uint32_t) * m2));
      FG_HIP(hipMalloc(&vals, sizeof(uint32_t) * m2));
      hipLaunchKernelGGL(link_pairs, dim3(ceil_div(m2, 256)), dim3(256), 0, s, m2, links, keys, vals);
      FG_LAUNCH_CHECK();
      uint32_t bits = 1;
      while (bits < 32 && (1ull << bits) < (uint64_t)n) ++bits;
      radix_sort_u32(s, w, 8, keys, vals, m2, 0, bits, &snodes, &scode);
    }
    hipLaunchKernelGGL(link_rows, dim3(ceil_div((uint64_t)n + 1, 256)), dim3(256), 0, s, n, m2,
                       (const uint32_t*)snodes, row);
    FG_LAUNCH_CHECK();
    if (m2) {
      hipLaunchKernelGGL(link_entries, dim3(ceil_div(m2, 256)), dim3(256), 0, s, m2, links,
                         (const uint32_t*)scode, (const uint32_t*)row, strengths, distances, bias, ent);
      FG_LAUNCH_CHECK();
    }
    FG_HIP(hipStreamSynchronize(s));
    uint32_t e[4] = {0, 0, 0, 0};
    if (w.errors) FG_HIP(hipMemcpy(e, w.errors, sizeof(e), hipMemcpyDeviceToHost));
    if (e[0]) throw HipError(hipErrorUnknown, "link: sort look-back bound reached");
  } catch (...) {
    if (bad) (void)hipFree(bad);
    if (keys) (void)hipFree(keys);
    if (vals) (void)hipFree(vals);
    throw;
  }
  (void)hipFree(bad);
  if (keys) (void)hipFree(keys);
  if (vals) (void)hipFree(vals);
}

}  // namespace fg
