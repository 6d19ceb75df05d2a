// forces.h -- launchers of the streaming force passes (see forces.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.h"
#include "sort.h"

namespace fg {

// One incident link of a node, as the node-centric link pass reads it.
struct LinkEntry {
  uint32_t other;   // the link's other end
  uint32_t code;    // link index * 2 + (this node is the link's target)
  float distance;
  float strength;
  float bias;
};

// A pointer as lane src of the wave holds it.
template <class T>
__device__ inline T* lane_ptr(T* p, int src) {
  const uint64_t u = (uint64_t)p;
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, src);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), src);
  return (T*)(((uint64_t)hi << 32) | lo);
}

// The snapshot p + v of node j: read from the snapshot array, or summed where it is read.
// at(src): the same as lane src of the wave has it.
struct SnapArray {
  const float4* __restrict__ snap;
  __device__ float4 operator()(uint32_t j) const { return snap[j]; }
  __device__ SnapArray at(int src) const { return SnapArray{lane_ptr(snap, src)}; }
};
template <int D>
struct SnapDirect {
  const float *__restrict__ x, *__restrict__ y, *__restrict__ z;
  const float *__restrict__ vx, *__restrict__ vy, *__restrict__ vz;
  __device__ float4 operator()(uint32_t j) const {
    return make_float4(x[j] + vx[j], D > 1 ? y[j] + vy[j] : 0.f, D > 2 ? z[j] + vz[j] : 0.f, 0.f);
  }
  __device__ SnapDirect at(int src) const {
    return SnapDirect{lane_ptr(x, src),  lane_ptr(y, src),  lane_ptr(z, src),
                      lane_ptr(vx, src), lane_ptr(vy, src), lane_ptr(vz, src)};
  }
};

// One link's term at one of its ends (snapshot me, the other end's t): d3's spring along the
// link between the two snapshots, and this end's share of it (bias for the target, 1 - bias
// for the source), added to (ax, ay, az).
template <int D>
__device__ inline void link_term(float4 me, float4 t, const LinkEntry& L, float alpha, uint32_t sn, uint32_t seed,
                                 float& ax, float& ay, float& az) {
  const bool target = L.code & 1u;
  const uint32_t id = L.code >> 1;
  // d = snap[target] - snap[source], the same vector at both ends
  float dx = target ? me.x - t.x : t.x - me.x;
  float dy = target ? me.y - t.y : t.y - me.y;
  float dz = target ? me.z - t.z : t.z - me.z;
  if (dx == 0.f) dx = jiggle_hash(id, 0u, 0u, sn, seed);
  if (D > 1 && dy == 0.f) dy = jiggle_hash(id, 0u, 1u, sn, seed);
  if (D > 2 && dz == 0.f) dz = jiggle_hash(id, 0u, 2u, sn, seed);
  const float l = sqrtf(dx * dx + (D > 1 ? dy * dy : 0.f) + (D > 2 ? dz * dz : 0.f));
  const float f = (l - L.distance) / l * alpha * L.strength;
  const float w = target ? -L.bias : 1.0f - L.bias;
  ax += dx * f * w;
  if (D > 1) ay += dy * f * w;
  if (D > 2) az += dz * f * w;
}

// A node with more incident links than this has them summed by the whole wave (below): one
// lane per node would leave the wave waiting on the longest row (a hub of a scale-free graph
// has hundreds).
constexpr uint32_t LINK_LANE_MAX = 32;

// The link pass for the nodes of one wave (every lane calls it; `live`: lane i has node i):
// each node's incident links, ox = vx + the sum of their terms.  A node with at most
// LINK_LANE_MAX links is summed by its own lane in CSR order; a longer row by all 64 lanes,
// lane k taking its entries k, k + 64, ... in order and the lanes' sums added in a fixed
// tree -- the same sum wherever the node's wave is, so a job still advances alike alone and
// in a batch.  The pointers may differ by lane (a batch's jobs): a long row uses its lane's.
template <int D, class Snap>
__device__ inline void link_wave(uint32_t i, bool live, Snap snap,
                                 const float* __restrict__ vx, const float* __restrict__ vy,
                                 const float* __restrict__ vz, float* __restrict__ ox,
                                 float* __restrict__ oy, float* __restrict__ oz,
                                 const uint32_t* __restrict__ row, const LinkEntry* __restrict__ ent,
                                 float alpha, uint32_t sn, uint32_t seed) {
  const uint32_t e0 = live ? row[i] : 0u, e1 = live ? row[i + 1] : 0u;
  const bool lone = live && e1 - e0 <= LINK_LANE_MAX;
  if (lone) {
    const float4 me = snap(i);
    float ax = 0.f, ay = 0.f, az = 0.f;
    for (uint32_t e = e0; e < e1; ++e) {
      const LinkEntry L = ent[e];
      link_term<D>(me, snap(L.other), L, alpha, sn, seed, ax, ay, az);
    }
    ox[i] = vx[i] + ax;
    if (D > 1) oy[i] = vy[i] + ay;
    if (D > 2) oz[i] = vz[i] + az;
  }
  uint64_t many = __builtin_amdgcn_ballot_w64(live && !lone);
  const uint32_t lane = __lane_id();
  while (many) {
    const int src = (int)__builtin_ctzll(many);
    many &= many - 1ull;
    const uint32_t j = (uint32_t)__builtin_amdgcn_readlane((int)i, src);
    const uint32_t b0 = (uint32_t)__builtin_amdgcn_readlane((int)e0, src);
    const uint32_t b1 = (uint32_t)__builtin_amdgcn_readlane((int)e1, src);
    const Snap sj = snap.at(src);
    const LinkEntry* ej = lane_ptr(ent, src);
    const float aj = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(alpha), src));
    const uint32_t snj = (uint32_t)__builtin_amdgcn_readlane((int)sn, src);
    const uint32_t seedj = (uint32_t)__builtin_amdgcn_readlane((int)seed, src);
    const float4 me = sj(j);
    float ax = 0.f, ay = 0.f, az = 0.f;
    for (uint32_t e = b0 + lane; e < b1; e += WAVE) {
      const LinkEntry L = ej[e];
      link_term<D>(me, sj(L.other), L, aj, snj, seedj, ax, ay, az);
    }
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) {
      ax += __shfl_xor(ax, o, WAVE);
      if (D > 1) ay += __shfl_xor(ay, o, WAVE);
      if (D > 2) az += __shfl_xor(az, o, WAVE);
    }
    if ((int)lane == src) {
      ox[j] = vx[j] + ax;
      if (D > 1) oy[j] = vy[j] + ay;
      if (D > 2) oz[j] = vz[j] + az;
    }
  }
}

// Graphs of at most this many nodes take the one-launch link pass (the snapshots summed where
// they are read).
constexpr uint32_t LINK_DIRECT_MAX = 32768;

// A link pass run by extra workgroups of another kernel (the one-workgroup tree build):
// every node's incident links from the snapshots summed where they are read, velocities
// from (vx, vy, vz) into (ox, oy, oz).
struct LinkFuse {
  const float *vx, *vy, *vz;
  float *ox, *oy, *oz;
  const uint32_t* row;
  const LinkEntry* ent;
  float alpha;
  uint32_t step, seed;
};

struct CenterWork {
  static constexpr uint32_t MAX_BLOCKS = 1024;
  double partial[MAX_BLOCKS][3];
  float shift[4];
  uint32_t done;
};

// Workgroups of center_reduce, at most: the partial sums it folds in workgroup order.
constexpr uint32_t CENTER_BLOCKS = 256;

// A center force whose mean another kernel takes on its way (the many-body tree's box
// kernel reads every position at the same point of the tick), with center_reduce's
// arithmetic: the shift lands in w->shift, bit for bit what center_reduce would leave.
struct CenterFuse {
  CenterWork* w = nullptr;
  double cx = 0, cy = 0, cz = 0, strength = 1;
};

// The collide grid's two classes of points (one when the radii are even): points up to
// r0 live in cells 2 r0 wide, larger ones in cells 2 rmax wide (inv_* = 1 / cell width).
struct GridClasses {
  float inv_s, inv_b, r0;
  int32_t two;
  float rmax;  // the largest |radius| (class 1's bound; class 0's is r0)
};

// The classes for these radii (by |radius|): two when rmax exceeds four times the 95th
// percentile radius (then one huge radius cannot make every cell hold every point).
GridClasses grid_classes(const float* radii, uint32_t n);

// A slot's run of sorted points [first, end) as the query reads it from the grid's start
// table (start[slot], start[slot + 1]).
struct GridRun {
  uint32_t first, end;
};

// The grid of one collide iteration, built by counting: the slot of each node, its rank
// among the nodes that reached the slot first (atomic, so in no fixed order), the slot
// counts turned by an exclusive scan into each slot's first sorted position, and the
// nodes placed by rank and then reordered by node index inside each slot.
struct GridWork {
  uint32_t cap = 0, bits = 0;
  uint32_t* keys = nullptr;     // slot of each node
  uint32_t* rank = nullptr;     // arrival rank of each node in its slot, then its sorted position
  uint32_t* start = nullptr;    // [2^bits + 1]: counts per slot, then first sorted position
  uint32_t* partial = nullptr;  // scan scratch
  uint32_t* arrived = nullptr;  // nodes by slot in arrival order
  uint32_t* svals = nullptr;    // nodes by slot, ascending node index within a slot
  float4* sorted = nullptr;     // snapshots (pos + vel, radius) in sorted order
  float4* snap = nullptr;       // the same per node (the positions the points were filed at)
  float4* dv = nullptr;         // the query's increments in sorted order
  float4* cur = nullptr;        // the positions of the next iteration on this grid, in sorted order
  uint32_t* margin2 = nullptr;  // [2]: the largest squared move from snap, by iteration parity
  void reserve(uint32_t n);
  void release();
  ~GridWork() { release(); }
};

// pos/vel/fix: three device arrays each (x, y, z); unused axes may be null.  center (or
// null): a center pass whose shift was computed but not yet applied.
void launch_integrate(hipStream_t s, uint32_t n, int ndim, float* const* pos, float* const* vel,
                      const uint8_t* mask, const float* const* fix, StepParams step,
                      const CenterWork* center = nullptr);
// The CSR of incident links, built on the device: links (2m node indices, device), the
// per-link FP64 arrays (device, or null for d3's defaults) -> row[n + 1], ent[2m] (device).
// Synchronizes the stream; throws ArgError on a node index out of range.
void link_csr_device(hipStream_t s, uint32_t n, uint32_t m, const uint32_t* links, const double* strengths,
                     const double* distances, const double* bias, uint32_t* row, LinkEntry* ent);

// snap: n float4 of scratch (the snapshot pos + vel the links read)
void launch_link(hipStream_t s, uint32_t n, int ndim, float* const* pos, const float* const* vin,
                 float* const* vout, const uint32_t* row, const LinkEntry* ent, float4* snap,
                 StepParams step);
// apply false: the shift is only computed (w->shift), for the integrator to apply.
// A tick whose last force is center over at most this many nodes: center and integration in
// one launch (launch_center_integrate), with the shift center_reduce would give.
constexpr uint32_t CENTER_FUSE_MAX = 1024;
void launch_center_integrate(hipStream_t s, uint32_t n, int ndim, float* const* pos, float* const* vel,
                             const uint8_t* mask, const float* const* fix, StepParams step, double cx, double cy,
                             double cz, double strength);
void launch_center(hipStream_t s, uint32_t n, int ndim, float* const* pos, CenterWork* w, double cx,
                   double cy, double cz, double strength, bool apply = true, bool reduce = true);
void launch_radial(hipStream_t s, uint32_t n, int ndim, float* const* pos, float* const* vel,
                   const float* radius, const float* weight, float cx, float cy, float cz,
                   StepParams step);
void launch_position(hipStream_t s, uint32_t n, const float* p, float* v, const float* goal,
                     const float* weight, StepParams step);
// The exact many-body sum (theta 0): `groups` waves of 64 * PAIRS_BODIES bodies, each over
// `splits` parts of the sources; with more than one part the sums are staged in
// pairs_scratch_bytes of device memory and added in part order.
constexpr int PAIRS_BODIES = 4;
constexpr uint32_t PAIRS_MAX_SPLITS = 16;
struct PairsPlan {
  uint32_t groups = 0, splits = 1;
};
PairsPlan pairs_plan(uint32_t n, int cus);
size_t pairs_scratch_bytes(const PairsPlan& p, uint32_t n);
void launch_manybody_pairs(hipStream_t s, uint32_t n, int ndim, float* const* pos,
                           const float* strength, float* const* vel, float dmin2, float dmax2,
                           StepParams step, const PairsPlan& plan, float* part);
void launch_collide_pairs(hipStream_t s, uint32_t n, int ndim, float* const* pos,
                          const float* const* vin, float* const* vout, const float* radius,
                          float strength, StepParams step);
// Iteration k of `iterations` of a grid collide: the first builds the grid, the later ones
// query the grid the first built (with each point's move since then as margin).
void launch_collide_grid(hipStream_t s, GridWork& g, uint32_t n, int ndim, float* const* pos,
                         float* const* vel, const float* radius, const GridClasses& gc, float strength,
                         int radix_bits, StepParams step, int k, int iterations);

}  // namespace fg
