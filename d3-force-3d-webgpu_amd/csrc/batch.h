// batch.h -- many independent layouts advanced by one sequence of launches.
//
// Every job owns its device state (positions, two velocity sets, pins, force arrays).  A
// batch is a set of index spaces over the jobs of one dimensionality: the node space
// (every job), the tree space (jobs with a many-body force) and the grid space (jobs with
// a collide force).  A space lists its jobs as segments [off, off + n) in job order; a
// kernel working on space index i finds its segment through seg_of[i] and the job's
// record through seg_job[seg].  Nothing in a space is relocated when jobs come and go:
// only the segment tables are rebuilt.
//
// Every per-job result depends on that job alone (sorting, tree, walk tiles, link sums,
// collide visiting order and centre sums are all relative to the job's own segment), so
// a job advances bit for bit the same alone and in any mix.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.h"
#include "forces.h"
#include "sort.h"

namespace fg {

enum BJobStatus : uint32_t {
  BJOB_FREE = 0,
  BJOB_RUNNING = 1,    // ticks on
  BJOB_FINISHING = 2,  // the tick now running is its last (alpha fell below alpha_min)
  BJOB_DONE = 3,       // finished; its state stays until it is released
};

enum BJobForce : uint32_t {
  BF_LINK = 1u << 0,
  BF_MANYBODY = 1u << 1,
  BF_COLLIDE = 1u << 2,
  BF_CENTER = 1u << 3,
  BF_RADIAL = 1u << 4,
  BF_POS_X = 1u << 5,
  BF_POS_Y = 1u << 6,
  BF_POS_Z = 1u << 7,
};

// One job as the kernels see it (device memory, one record per job slot).
struct alignas(64) BJob {
  // state owned by the job
  float* p[3];
  float* v[2][3];
  float* fix[3];
  const uint8_t* mask;  // null: nothing pinned
  // forces' arrays (job-local indices)
  const float* mb_strength;
  const uint32_t* link_row;
  const LinkEntry* link_ent;
  const float* radius;
  const float* radial_r;
  const float* radial_w;
  const float* pos_goal[3];
  const float* pos_weight[3];
  // segments in the spaces
  uint32_t n;
  uint32_t node_off, tree_off, grid_off;
  uint32_t node_tiles_first, node_tiles;  // reduction tiles in the node space
  uint32_t tree_tiles_first, tree_tiles;  // reduction tiles in the tree space
  // schedule (d3's alpha: simulation.js:54-73), advanced on the device every tick
  double alpha, alpha_min, alpha_decay, alpha_target;
  float decay;       // velocity multiplier, 1 - velocityDecay
  float alpha_f;     // this tick's alpha as the kernels use it
  uint32_t status;   // BJobStatus
  uint32_t active;   // 1 while the current tick runs for this job
  uint32_t step;     // this tick's step number (the jiggle's step input)
  uint32_t ticks;    // ticks run so far
  uint32_t max_ticks;  // 0: until alpha < alpha_min
  uint32_t seed;
  uint32_t forces;   // BJobForce bits
  uint32_t ndim;
  // many-body
  double theta2, dmin2, dmax2;
  // link
  uint32_t link_iters;
  uint32_t vcur;     // velocity set current after the link iterations (link_iters & 1)
  // collide
  uint32_t coll_iters;
  GridClasses coll;  // the grid's cell widths and radius classes
  float coll_strength;
  // center
  double center[3];
  double center_strength;
  float center_shift[3];
  uint32_t center_done;
  // radial
  float radial_c[3];
  // tree cube (d3's cover of the job's box) and the running box
  float cube_origin[3];
  float cube_size;
  int32_t cube_shift;
  uint32_t cube_done;
  uint32_t cube_box[6];
};

// A reduction tile: count points of segment seg from first (space index).
struct BTile {
  uint32_t seg, first, count, pad;
};

// One index space over some of the jobs (device arrays + host sizes).
struct BSpace {
  uint32_t total = 0;       // points in the space
  uint32_t segs = 0;        // jobs in the space
  uint32_t tiles = 0;       // reduction tiles
  uint32_t walk_tiles = 0;  // 64-point walk tiles (tree space)
  uint32_t max_n = 0;       // points of the largest job in the space
  uint32_t* seg_of = nullptr;   // [total] segment of each index
  uint32_t* seg_job = nullptr;  // [segs] job slot of each segment
  BTile* tile = nullptr;        // [tiles]
  uint2* walk = nullptr;        // [walk_tiles] (segment, first index of the tile)
};

constexpr uint32_t BTILE = 2048;  // points per reduction tile

// Tunables of the batched tick (chosen by the autotuner, see fg_batch_autotune).
struct BatchTune {
  int radix_bits = 8;          // 8 or 11 bits per radix digit
  int walk_waves = 4;          // waves per walk workgroup: 2, 4 or 8
  int walk_stack = 48;         // LDS stack entries per wave: 32, 48, 64 or 96
  int walk_blocks_per_cu = 0;  // persistent walk grid: blocks per CU (0: as many as fit)
  uint32_t walk_packed_max = 24;
};

// Device scratch of the tree of one batch group.
struct BatchTree {
  uint32_t cap = 0;
  uint64_t* keys = nullptr;   // code per tree index, then sorted
  uint64_t* codes = nullptr;  // code per tree index (kept for the sorted copy)
  uint64_t* skeys = nullptr;  // sorted codes
  uint32_t* vals = nullptr;
  uint32_t* digit = nullptr;  // segment of each code-sorted entry (the last sort pass)
  uint32_t* order = nullptr;  // tree index at each sorted position
  float4* leaves = nullptr;
  float4* pts = nullptr;        // per node: x, y, z, strength (what the leaves gather)
  uint32_t* leaf_depth = nullptr;
  uint32_t* leaf_parent = nullptr;
  uint4* links = nullptr;
  uint32_t* flags = nullptr;
  void* rec = nullptr;   // NodeRec[cap]
  void* sums = nullptr;  // NodeSum[cap]
  uint32_t* counters = nullptr;  // interactions [0,1], opened [2,3], errors [4], queue [5], lane visits [6,7], long runs [10]
  uint32_t* long_runs = nullptr; // first positions of runs of equal high halves too long to order in place
  void* spill = nullptr;         // the walk's stack entries beyond LDS, per resident wave
  size_t spill_bytes = 0;
  SortWork sort, sort_seg;
  void reserve(uint32_t n);
  void release();
  ~BatchTree() { release(); }
};

struct BatchGrid {  // as GridWork, over a batch's grid space
  uint32_t cap = 0, bits = 0;
  uint32_t* keys = nullptr;
  uint32_t* rank = nullptr;
  uint32_t* start = nullptr;    // [2^bits + 1]
  uint32_t* partial = nullptr;
  uint32_t* arrived = nullptr;
  uint32_t* svals = nullptr;
  float4* sorted = nullptr;
  float4* snap = nullptr;       // per node, then the query's increments in sorted order
  uint32_t* sseg = nullptr;     // segment of each sorted point
  void reserve(uint32_t n);
  void release();
  ~BatchGrid() { release(); }
};

struct BatchWork {
  double* center_partial = nullptr;  // [node tiles][3]
  uint32_t center_cap = 0;
  void reserve(uint32_t tiles);
  void release();
  ~BatchWork() { release(); }
};

// Launchers (one group: every job of dimensionality D).  `jobs` is the device job table.
void batch_advance(hipStream_t s, BJob* jobs, uint32_t njobs, uint32_t* walk_counters);
void batch_manybody(hipStream_t s, int ndim, BatchTree& t, const BSpace& sp, BJob* jobs,
                    const BatchTune& tune, int cus, PhaseTimer* timer);
// snap: float4 scratch for every index of the node space
void batch_link(hipStream_t s, int ndim, const BSpace& sp, BJob* jobs, uint32_t iteration, float4* snap);
void batch_collide(hipStream_t s, int ndim, BatchGrid& g, const BSpace& sp, BJob* jobs,
                   uint32_t iteration, int radix_bits);
void batch_center(hipStream_t s, int ndim, BatchWork& w, const BSpace& sp, BJob* jobs);
void batch_radial(hipStream_t s, int ndim, const BSpace& sp, BJob* jobs);
void batch_position(hipStream_t s, const BSpace& sp, BJob* jobs, int axis);
void batch_integrate(hipStream_t s, int ndim, const BSpace& sp, BJob* jobs);

}  // namespace fg
