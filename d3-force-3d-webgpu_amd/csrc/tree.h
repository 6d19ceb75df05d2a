// tree.h -- the many-body Barnes-Hut tree of one simulation (see tree.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.h"
#include "forces.h"
#include "sort.h"

namespace fg {

constexpr uint32_t REF_POINT = 0x80000000u;  // a child reference that is a sorted point
constexpr uint32_t TREE_COUNTERS = 16;       // words of TreeBuffers::counters
constexpr uint32_t REF_NONE = 0xffffffffu;
// The walk by tree size (measured: link + manyBody + center ticks, 3D): trees of at most
// WALK_WIDE_MAX points walk one point per wave (the wide walk), up to WALK_GROUP_MAX two
// points per wave (the group walk), larger ones 64-point tiles with the tiled walk's
// cheaper accumulation.  The tiled walk's exact mode (forces independent of the tiling)
// covers trees of at most WALK_EXACT_MAX points where the other walks are turned off.
constexpr uint32_t WALK_WIDE_MAX = 16384;
constexpr uint32_t WALK_GROUP_MAX = 262144;
constexpr uint32_t WALK_GROUP = 2;  // points per wave of the group walk
constexpr uint32_t WALK_EXACT_MAX = 163840;
constexpr uint32_t WALK_SLOTS = 4096;         // workgroups of a walk that counts into slots, at most

// What the walk reads of a binary node: its centre of charge (FP64), its strength at the
// depth of its deepest cell, that depth, and its children in the d3 tree.
struct alignas(64) NodeRec {
  double cx, cy, cz;
  float value;
  uint32_t depth;
  uint32_t kids[8];
};

// The first half of a NodeRec: what a parent's walk step reads of each child.
struct alignas(32) NodeHdr {
  double cx, cy, cz;
  float value;
  uint32_t depth;
};

// Sums the bottom-up reduction carries per node (FP64).
struct NodeSum {
  double a;       // sum of strength * c^leafDepth
  double w;       // sum of the d3-children weights inside this node's cell group
  double gx, gy, gz;  // weight * centre, same children
  uint32_t same;  // all of the node's points lie at one position (nodes in equal-code runs)
  uint32_t pad;
};

// Cube of the tree (d3's cover), and the running box used to find it.
constexpr uint32_t BBOX_BLOCKS = 256;  // workgroups of the single-simulation bbox kernel, at most

struct CubeInfo {
  float origin[3];
  float size;
  int32_t shift;
  uint32_t done;        // blocks of the bbox kernel that have finished
  uint32_t box[6];      // order-preserving min x, y, z, max x, y, z
  uint32_t part[BBOX_BLOCKS][6];  // the same per workgroup of the bbox kernel
};

struct TreeBuffers {
  uint32_t cap = 0;
  uint64_t* keys = nullptr;     // compact Morton code per node, then per sorted point
  uint32_t* vals = nullptr;     // node index, then the sorted order
  uint64_t* skeys = nullptr;    // the sorted codes (points into keys or the sorter's buffer)
  uint32_t* order = nullptr;    // node index at each sorted position
  float4* leaves = nullptr;     // sorted: x, y, z, strength
  float4* pts = nullptr;        // per node: x, y, z, strength (what the leaves gather)
  uint32_t* leaf_depth = nullptr;
  uint32_t* leaf_parent = nullptr;
  uint4* links = nullptr;       // per binary node: left, right, depth, parent
  uint32_t* flags = nullptr;
  NodeRec* rec = nullptr;
  NodeSum* sums = nullptr;
  CubeInfo* cube = nullptr;
  // the wide walk's counts: 8 words per workgroup (u64 interactions, u64 opened, errors,
  // deepest stack), for the slot_blocks workgroups of the last walk (0: that walk counted
  // into `counters`)
  uint32_t* slots = nullptr;
  uint32_t slot_blocks = 0;
  uint8_t* cut_start = nullptr;  // per sorted point: 1 where a subtree-aligned reduction workgroup's subtree starts
  uint32_t* counters = nullptr; // TREE_COUNTERS words: u64 interactions [0,1], opened [2,3], walk errors [4], tile queue [5], lane visits [6,7], deepest stack [8], long runs [10]
  uint32_t* long_runs = nullptr;  // first positions of runs of equal high halves too long to order in place
  void* spill = nullptr;        // the walk's stack entries beyond LDS, per resident wave
  size_t spill_bytes = 0;
  SortWork sort;

  void reserve(uint32_t n);
  void reserve_spill(size_t bytes);
  void release();
  ~TreeBuffers() { release(); }
};

constexpr uint32_t WALK_SOLO_BY_DIMS = 0xffffffffu;  // walk_solo_max: the measured default per dims
constexpr uint32_t WALK_PACKED_BY_MODE = 0xffffffffu;  // walk_packed_max: the measured default per mode
constexpr uint32_t WALK_WIDE_BY_DIMS = 0xffffffffu;  // walk_wide_max: the measured default per dims
constexpr uint32_t WALK_POOL_BY_MODE = 0xffffffffu;  // walk_pool_k: the measured default

struct TreeParams {
  double theta2, dmin2, dmax2;
  uint32_t walk_packed_max = WALK_PACKED_BY_MODE;  // openings by at most this many points run packed
  uint32_t walk_tile = 0;         // points per tile of a small (exact-mode) walk; 0: by size
  // trees of at most this many points walk one point per group of lanes (0: never)
  uint32_t walk_solo_max = WALK_SOLO_BY_DIMS;
  // trees of at most this many points walk one point per wave (0: never)
  uint32_t walk_wide_max = WALK_WIDE_BY_DIMS;
  // points per wave of the group walk: 2..16 (any size), 1 (never), 0 (WALK_GROUP up to
  // WALK_GROUP_MAX points)
  uint32_t walk_group = 0;
  uint32_t reduce_block = 0;    // points per workgroup of the reduction (512 or 1024; 0: by size)
  bool reduce_aligned = true;  // reduction workgroups aligned to subtrees
  // the tiled walk's cheaper mode: openings by at most this many points go to the pool of
  // single-point openings (0: no pool)
  uint32_t walk_pool_k = WALK_POOL_BY_MODE;
  uint32_t walk_pool_cap = 0xffffffffu;  // pool entries per wave, at most the compiled size
  // the tiled walk takes its exact mode for trees of at most this many points
  uint32_t walk_exact_max = WALK_EXACT_MAX;
};

// Bytes of walk spill area for `cus` CUs at an LDS stack of `stack` entries.
size_t spill_bytes_for(int cus, int stack);

// Compute units of the current device.
int device_cus();

// Build the tree over the current positions (and strengths): what the walk reads.  center:
// as in tree_manybody.  link (optional, only where tree_link_pass): a link pass that comes
// right before this force, run by the build's launch on the CUs the build leaves idle.
void tree_build(hipStream_t s, TreeBuffers& t, uint32_t n, int ndim, const float* x, const float* y,
                const float* z, const float* strength, const TreeParams& p, int radix_bits,
                PhaseTimer* timer = nullptr, const CenterFuse* center = nullptr, const LinkFuse* link = nullptr);

// Whether the build of a tree of n points can take a link pass in its first launch (the
// one-workgroup build or the box kernel: the graphs that take the one-launch link pass).
bool tree_link_pass(uint32_t n);

// Add the many-body force of the tree tree_build left to the velocities.
void tree_forces(hipStream_t s, TreeBuffers& t, uint32_t n, int ndim, float* vx, float* vy, float* vz,
                 const TreeParams& p, StepParams step, PhaseTimer* timer = nullptr);

// Build the tree over the current positions and add the many-body force to the velocities.
// center (optional): a center force that runs after this one in the tick, with no force
// between them that moves the nodes; its mean is taken by the box kernel (when there is one:
// tree_box_pass) and its shift left in center->w.
void tree_manybody(hipStream_t s, TreeBuffers& t, uint32_t n, int ndim, const float* x,
                   const float* y, const float* z, const float* strength, float* vx, float* vy,
                   float* vz, const TreeParams& p, StepParams step, int radix_bits,
                   PhaseTimer* timer = nullptr, const CenterFuse* center = nullptr);

// Whether the tree of n points is built with the box kernel (larger than one workgroup's).
bool tree_box_pass(uint32_t n);

}  // namespace fg
