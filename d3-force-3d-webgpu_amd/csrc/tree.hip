// tree.hip -- the many-body force with a Barnes-Hut tree, on CDNA4.
//
// The tree is the one the reference GPU path builds (src/gpu/passes/barnesHut.js:6-47):
// d3's cube (cover), LEVELS levels of Morton code in two halves, points sorted on the
// whole code (ties by node index), a Karras 2012 binary radix tree over the sorted codes
// whose nodes each stand for a chain of d3 cells, and d3's children of every node
// (collapse).  Codes, order, topology and children come out bit-identical to that
// algorithm.  The kernels are this backend's own:
//
//   tree_bbox     grid-stride min/max into per-workgroup slots; the last workgroup to
//                 finish folds them and turns the box into the cube (one launch)
//   tree_keys     exact cellOf -> compact code (high half << HALF | low half) per node
//   (sort)        LSD radix sort of the high halves with the node index as value, then
//                 each run of equal high halves ordered on the low halves (order_runs)
//   tree_leaves_build  sorted points' (x, y, z, strength) and leaf depths, and the Karras
//                 ranges and splits of the binary nodes with parent links for the climb
//   tree_reduce   one thread per point climbs to the root; the second of two siblings to
//                 arrive at a node (atomic flag) combines them -- left, then right, so the
//                 sums do not depend on timing -- in FP64, with d3's |value| weighting
//                 (manyBody.js:42-71); then lists its node's d3 children (collapse)
//   tree_walk     one wave per tile of up to 64 consecutive sorted points: the wave walks
//                 the union of its points' traversals from a stack in LDS of (node, lane
//                 mask); every lane takes d3's own decision at every cell
//                 (barnesHut.js:719-746), so each point gets exactly its d3 traversal.
//
// A batch of trees (run_batch) uses the same steps with per-job segments (btree_*).
#include <math.h>

#include "tree.h"
#include "batch.h"

#include <string>
#include <type_traits>

namespace fg {

namespace {

constexpr int LEVELS_OF[4] = {0, 24, 24, 20};

template <int D>
struct Dims {
  static constexpr int LEVELS = LEVELS_OF[D];
  static constexpr int HALF = LEVELS / 2 * D;  // bits of each half of the code
  static constexpr int BITS = 2 * HALF;
};

// c^k for c = sqrt(4 / 2^D), exact up to the rounding of sqrt(2) (barnesHut.js:83-89).
template <int D>
__host__ __device__ inline double damp(int k) {
  if (D == 2) return 1.0;
  int t = D == 3 ? -k : k;  // exponent in halves
  int q = t >> 1;           // floor(t / 2)
  return ldexp((t & 1) ? 1.4142135623730951 : 1.0, q);
}

__device__ inline uint32_t spread3(uint32_t v) {
  uint32_t x = v & 0x3ffu;
  x = (x | (x << 16u)) & 0x030000ffu;
  x = (x | (x << 8u)) & 0x0300f00fu;
  x = (x | (x << 4u)) & 0x030c30c3u;
  x = (x | (x << 2u)) & 0x09249249u;
  return x;
}

__device__ inline uint32_t spread2(uint32_t v) {
  uint32_t x = v & 0xffffu;
  x = (x | (x << 8u)) & 0x00ff00ffu;
  x = (x | (x << 4u)) & 0x0f0f0f0fu;
  x = (x | (x << 2u)) & 0x33333333u;
  x = (x | (x << 1u)) & 0x55555555u;
  return x;
}

// Morton code of one half, right-aligned (the reference left-aligns it in a u32).
template <int D>
__device__ inline uint32_t morton(uint32_t qx, uint32_t qy, uint32_t qz) {
  if (D == 3) return (spread3(qz) << 2) | (spread3(qy) << 1) | spread3(qx);
  if (D == 2) return (spread2(qy) << 1) | spread2(qx);
  return qx;
}

// floor((x - origin) / cell) without rounding (barnesHut.js:343-350).
__device__ inline float cell_of(float x, float origin, int k) {
  if (k >= 0) return floorf(ldexpf(x, k)) - ldexpf(origin, k);
  return floorf(ldexpf(floorf(x) - origin, k));
}

// ---- bbox + cube --------------------------------------------------------------------

// Reduces the 256 threads' boxes and folds the block's box into the u32 box words with
// order-preserving atomics (barnesHut.js:140-167).
// The workgroup's box (256 threads): threads 0..2 get axis a's order-preserving min and max.
__device__ inline void block_box_of(float lo[3], float hi[3], uint32_t& omin, uint32_t& omax) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      lo[a] = fminf(lo[a], __shfl_xor(lo[a], o, WAVE));
      hi[a] = fmaxf(hi[a], __shfl_xor(hi[a], o, WAVE));
    }
  }
  __shared__ float s_lo[4][3], s_hi[4][3];
  const uint32_t lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x / WAVE;
  if (lane == 0)
    for (int a = 0; a < 3; ++a) s_lo[wave][a] = lo[a], s_hi[wave][a] = hi[a];
  __syncthreads();
  if (threadIdx.x < 3) {
    const int a = threadIdx.x;
    omin = float_ordered(fminf(fminf(s_lo[0][a], s_lo[1][a]), fminf(s_lo[2][a], s_lo[3][a])));
    omax = float_ordered(fmaxf(fmaxf(s_hi[0][a], s_hi[1][a]), fmaxf(s_hi[2][a], s_hi[3][a])));
  }
  __syncthreads();
}

// ... folded into a box in global memory with atomics (the batch's per-job boxes).
__device__ inline void block_box(float lo[3], float hi[3], uint32_t* box) {
  uint32_t omin = 0, omax = 0;
  block_box_of(lo, hi, omin, omax);
  if (threadIdx.x < 3) {
    atomicMin(&box[threadIdx.x], omin);
    atomicMax(&box[3 + threadIdx.x], omax);
  }
  __syncthreads();
}

// d3's cover() of the box (barnesHut.js:169-181): origin = floor(min), size = the smallest
// power of two with max < origin + size; shift = LEVELS - log2(size).
template <int D>
__device__ inline void cover_of(const float blo[3], const float bhi[3], float origin[3], float& size,
                                int32_t& shift) {
#pragma unroll
  for (int a = 0; a < 3; ++a) origin[a] = floorf(blo[a]);
  size = 1.0f;
  int log2size = 0;
  for (int k = 0; k < 128; ++k) {
    if (bhi[0] < origin[0] + size && bhi[1] < origin[1] + size && bhi[2] < origin[2] + size) break;
    size *= 2.0f;
    log2size += 1;
  }
  shift = Dims<D>::LEVELS - log2size;
}

// The same of an order-preserving box in global memory, which it resets.
template <int D>
__device__ inline void cover_box(uint32_t* box, float origin[3], float& size, int32_t& shift) {
  float blo[3], bhi[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    blo[a] = float_unordered(__hip_atomic_load(&box[a], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    bhi[a] = float_unordered(__hip_atomic_load(&box[3 + a], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
  }
  cover_of<D>(blo, bhi, origin, size, shift);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    box[a] = 0xffffffffu;
    box[3 + a] = 0u;
  }
}

// Also clears, in its first workgroup, the tree's counters and the high-half sort's
// histograms (work the tick's later kernels need cleared, here without launches of their
// own).
//
// center.w (optional): also the sums of a center force's mean, with center_reduce's
// arithmetic (the same grid, the same points per thread in the same order, the same adds
// over waves, workgroups and, in the last workgroup, the partials), so the shift it leaves
// is bit for bit center_reduce's.
template <int D>
__global__ void __launch_bounds__(256) tree_bbox(const float* __restrict__ x,
                                                 const float* __restrict__ y,
                                                 const float* __restrict__ z, uint32_t n,
                                                 CubeInfo* __restrict__ cube, uint32_t* __restrict__ counters,
                                                 SortClear sort_clear, CenterFuse center, uint32_t boxes,
                                                 LinkFuse link) {
  if (blockIdx.x >= boxes) {  // the workgroups past the box's run the link pass before this force
    // (a whole wave per step: long rows take the wave)
    const uint32_t lane = threadIdx.x & (WAVE - 1);
    for (uint32_t w = (blockIdx.x - boxes) * 256 + (threadIdx.x & ~(WAVE - 1)); w < n; w += (gridDim.x - boxes) * 256)
      link_wave<D>(w + lane, w + lane < n, SnapDirect<D>{x, y, z, link.vx, link.vy, link.vz}, link.vx, link.vy,
                   link.vz, link.ox, link.oy, link.oz, link.row, link.ent, link.alpha, link.step, link.seed);
    return;
  }
  if (blockIdx.x == 0) {
    if (threadIdx.x < TREE_COUNTERS) counters[threadIdx.x] = 0;
    sort_clear_block(sort_clear);
  }
  CenterWork* const cw = center.w;
  float lo[3] = {3.4e38f, 3.4e38f, 3.4e38f}, hi[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
  double sum[3] = {0, 0, 0};
  auto take = [&](uint32_t i) {
    const float p[3] = {x[i], D > 1 ? y[i] : 0.0f, D > 2 ? z[i] : 0.0f};
#pragma unroll
    for (int a = 0; a < 3; ++a) lo[a] = fminf(lo[a], p[a]), hi[a] = fmaxf(hi[a], p[a]);
    if (cw) {
#pragma unroll
      for (int a = 0; a < D; ++a) sum[a] += (double)p[a];
    }
  };
  const uint32_t stride = boxes * 256;
  uint32_t i = blockIdx.x * 256 + threadIdx.x;
  for (; i + 3 * stride < n; i += 4 * stride) {  // four independent loads in flight
    take(i);
    take(i + stride);
    take(i + 2 * stride);
    take(i + 3 * stride);
  }
  for (; i < n; i += stride) take(i);
  // each workgroup leaves its box in its own slot (no atomics on shared words: on this
  // device those serialise across the chip), the last one to finish folds them
  uint32_t omin = 0, omax = 0;
  block_box_of(lo, hi, omin, omax);
  if (threadIdx.x < 3) {
    store_coherent(&cube->part[blockIdx.x][threadIdx.x], omin);
    store_coherent(&cube->part[blockIdx.x][3 + threadIdx.x], omax);
  }
  __shared__ double s_sum[4][3];
  const uint32_t lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x / WAVE;
  if (cw) {  // center_reduce's sums of this workgroup
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) sum[a] += __shfl_xor(sum[a], o, WAVE);
    if (lane == 0)
      for (int a = 0; a < 3; ++a) s_sum[wave][a] = sum[a];
  }
  __shared__ uint32_t s_last;
  __syncthreads();
  if (threadIdx.x == 0) {
    if (cw)
      for (int a = 0; a < 3; ++a)
        store_coherent(&cw->partial[blockIdx.x][a], s_sum[0][a] + s_sum[1][a] + s_sum[2][a] + s_sum[3][a]);
    s_last = publish_count(&cube->done) + 1 == boxes;
  }
  __syncthreads();
  if (!s_last) return;
  uint32_t m[6] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0u, 0u, 0u};
  for (uint32_t b = threadIdx.x; b < boxes; b += 256)
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      m[a] = min(m[a], load_coherent(&cube->part[b][a]));
      m[3 + a] = max(m[3 + a], load_coherent(&cube->part[b][3 + a]));
    }
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      m[a] = min(m[a], (uint32_t)__shfl_xor((int)m[a], o, WAVE));
      m[3 + a] = max(m[3 + a], (uint32_t)__shfl_xor((int)m[3 + a], o, WAVE));
    }
  __shared__ uint32_t s_m[4][6];
  if ((threadIdx.x & (WAVE - 1)) == 0)
    for (int a = 0; a < 6; ++a) s_m[wave][a] = m[a];
  __syncthreads();
  if (threadIdx.x == 0) {  // the last block turns the box into the cube
    for (int a = 0; a < 3; ++a) {
      cube->box[a] = min(min(s_m[0][a], s_m[1][a]), min(s_m[2][a], s_m[3][a]));
      cube->box[3 + a] = max(max(s_m[0][3 + a], s_m[1][3 + a]), max(s_m[2][3 + a], s_m[3][3 + a]));
    }
    float origin[3], size;
    int32_t shift;
    cover_box<D>(cube->box, origin, size, shift);
#pragma unroll
    for (int a = 0; a < 3; ++a) cube->origin[a] = origin[a];
    cube->size = size;
    cube->shift = shift;
    cube->done = 0;
  }
  if (cw) {  // then center_reduce's fold of the partials, in workgroup order
    __shared__ double tsum[256][3];
    double t[3] = {0, 0, 0};
    for (uint32_t b = threadIdx.x; b < boxes; b += 256)
      for (int a = 0; a < 3; ++a) t[a] += load_coherent(&cw->partial[b][a]);
    for (int a = 0; a < 3; ++a) tsum[threadIdx.x][a] = t[a];
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int a = 0; a < 3; ++a) t[a] = 0;
      for (uint32_t k = 0; k < boxes && k < 256; ++k)  // the rest are +0: no change
        for (int a = 0; a < 3; ++a) t[a] += tsum[k][a];
      const double c[3] = {center.cx, center.cy, center.cz};
      for (int a = 0; a < 3; ++a) cw->shift[a] = a < D ? (float)((t[a] / n - c[a]) * center.strength) : 0.0f;
    }
  }
}

// ---- keys ---------------------------------------------------------------------------

// The compact code of a point: exact cellOf per axis, high half << HALF | low half
// (barnesHut.js:343-366).
template <int D>
__device__ inline uint64_t point_key(const float p[3], const float origin[3], int k) {
  constexpr int L = Dims<D>::LEVELS, H = L / 2;
  const float top = (float)((1u << L) - 1u);
  uint32_t q[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    float c = cell_of(p[a], origin[a], k);
    c = fminf(fmaxf(c, 0.0f), top);
    q[a] = (uint32_t)c;
  }
  constexpr uint32_t lm = (1u << H) - 1u;
  uint32_t hi = morton<D>(q[0] >> H, q[1] >> H, q[2] >> H);
  uint32_t lo = morton<D>(q[0] & lm, q[1] & lm, q[2] & lm);
  return ((uint64_t)hi << Dims<D>::HALF) | lo;
}

template <int D>
__global__ void __launch_bounds__(256) tree_keys(const float* __restrict__ x,
                                                 const float* __restrict__ y,
                                                 const float* __restrict__ z,
                                                 const float* __restrict__ strength, uint32_t n,
                                                 const CubeInfo* __restrict__ cube,
                                                 uint64_t* __restrict__ keys,
                                                 uint32_t* __restrict__ vals,
                                                 float4* __restrict__ pts, uint32_t* __restrict__ rank,
                                                 uint32_t* __restrict__ groups_done) {
  uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  if (rank) {  // the counting sort's ranks and its groups' counters
    rank[i] = 0;
    if (i < (n + 63u) / 64u) groups_done[i] = 0;
  }
  const float p[3] = {x[i], D > 1 ? y[i] : 0.0f, D > 2 ? z[i] : 0.0f};
  keys[i] = point_key<D>(p, cube->origin, cube->shift);
  vals[i] = i;
  pts[i] = make_float4(p[0], p[1], p[2], strength[i]);
}

// Length of the common prefix of the codes at sorted positions a and b, counted as the
// reference counts it: code bits, then 64 + common bits of the positions for equal codes
// (barnesHut.js:104-117); -1 when b is out of range.
template <int D>
__device__ inline int delta(const uint64_t* __restrict__ k, int n, int a, int b) {
  if (b < 0 || b >= n) return -1;
  uint64_t ka = k[a], kb = k[b];
  if (ka != kb) return __clzll(ka ^ kb) - (64 - Dims<D>::BITS);
  return 64 + __clz((uint32_t)a ^ (uint32_t)b);
}

template <int D>
__device__ inline uint32_t depth_of(int prefix) {
  uint32_t d = (uint32_t)prefix / (uint32_t)D;
  return d < (uint32_t)Dims<D>::LEVELS ? d : (uint32_t)Dims<D>::LEVELS;
}

// ---- the order of the codes ---------------------------------------------------------
//
// The radix sort orders the points on the high halves of their codes only (half the
// passes).  Within a run of points sharing a high half they stay in node order; ordering
// each run on the low halves, node order breaking ties, gives exactly the order of a
// stable sort on the whole codes (barnesHut.js:299-305 orders its runs the same way).  A run
// of up to ORDER_RUN points is ordered in place by ranks; a longer one is listed and sorted
// by order_long_runs, a workgroup per run.

constexpr uint32_t ORDER_RUN = 32;
constexpr uint32_t LONG_RUNS_COUNTER = 10;  // counters[10]: runs listed for order_long_runs
constexpr uint32_t LONG_RUN_BLOCKS = 64;    // workgroups of order_long_runs (each takes runs in turn)

// Positions a and b in one run: the same high half (and, in a batch, the same job: seg).
template <int D>
__device__ inline bool same_run(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ seg, uint32_t a,
                                uint64_t high, uint32_t sg) {
  return (keys[a] >> Dims<D>::HALF) == high && (!seg || seg[a] == sg);
}

template <int D>
__global__ void __launch_bounds__(256) order_runs(uint32_t n, const uint64_t* __restrict__ keys,
                                                  const uint32_t* __restrict__ vals,
                                                  const uint32_t* __restrict__ seg, uint64_t* __restrict__ okeys,
                                                  uint32_t* __restrict__ ovals, uint32_t* __restrict__ counters,
                                                  uint32_t* __restrict__ long_runs) {
  const uint32_t s = blockIdx.x * 256 + threadIdx.x;
  if (s >= n) return;
  constexpr int HALF = Dims<D>::HALF;
  const uint64_t k = keys[s];
  const uint64_t high = k >> HALF;
  const uint32_t sg = seg ? seg[s] : 0u;
  uint32_t first = s, last = s;
  while (first > 0 && s - first <= ORDER_RUN && same_run<D>(keys, seg, first - 1, high, sg)) --first;
  while (last + 1 < n && last - s <= ORDER_RUN && same_run<D>(keys, seg, last + 1, high, sg)) ++last;
  const uint32_t v = vals[s];
  if (first == last || last - first + 1 > ORDER_RUN) {
    // alone (in place) or in a long run (copied as it is for order_long_runs)
    okeys[s] = k;
    ovals[s] = v;
    if (first == s && first != last) {
      const uint32_t slot = atomicAdd(&counters[LONG_RUNS_COUNTER], 1u);
      long_runs[slot] = s;
    }
    return;
  }
  // the run is in node order: this point goes after the points with a smaller low half and
  // after those with the same one that come before it
  const uint64_t lo = k & ((1ull << HALF) - 1ull);
  uint32_t place = 0;
  for (uint32_t j = first; j <= last; ++j) {
    const uint64_t other = keys[j] & ((1ull << HALF) - 1ull);
    place += (other < lo || (other == lo && j < s)) ? 1u : 0u;
  }
  okeys[first + place] = k;
  ovals[first + place] = v;
}

// Each workgroup sorts listed long runs of (okeys, ovals) on the low halves (seg: the job of
// each position in a batch, runs never crossing jobs, or null): stable 8-bit
// passes from the least significant digit, through the same positions of (tkeys, tvals)
// and back; ranks within a chunk of 256 come from wave ballots and per-wave digit counts.
template <int D>
__global__ void __launch_bounds__(256) order_long_runs(uint32_t n, const uint32_t* __restrict__ counters,
                                                       const uint32_t* __restrict__ long_runs,
                                                       const uint32_t* __restrict__ seg,
                                                       uint64_t* __restrict__ okeys, uint32_t* __restrict__ ovals,
                                                       uint64_t* __restrict__ tkeys, uint32_t* __restrict__ tvals) {
  constexpr int HALF = Dims<D>::HALF;
  constexpr int PASSES = (HALF + 7) / 8;
  __shared__ uint32_t base[256];
  __shared__ uint32_t wcount[4][256];
  __shared__ uint32_t s_end;
  const uint32_t li = threadIdx.x, lane = li & 63u, wave = li >> 6;
  const uint32_t runs = counters[LONG_RUNS_COUNTER];
  for (uint32_t r = blockIdx.x; r < runs; r += gridDim.x) {
    const uint32_t first = long_runs[r];
    const uint64_t high = okeys[first] >> HALF;
    const uint32_t sg = seg ? seg[first] : 0u;
    // the run's end: the first position after it whose high half differs
    if (li == 0) s_end = n;
    __syncthreads();
    uint32_t end = n;
    for (uint32_t b = first; b < n; b += 256) {
      const uint32_t t = b + li;
      if (t < n && !same_run<D>(okeys, seg, t, high, sg)) atomicMin(&s_end, t);
      __syncthreads();
      end = s_end;  // read by every thread between two barriers: the same value, the same exit
      __syncthreads();
      if (end < n) break;
    }
    const uint32_t count = end - first;
    for (int pass = 0; pass < PASSES; ++pass) {
      const uint32_t shift = 8u * (uint32_t)pass;
      uint64_t* sk = (pass & 1) ? tkeys : okeys;
      uint32_t* sv = (pass & 1) ? tvals : ovals;
      uint64_t* dk = (pass & 1) ? okeys : tkeys;
      uint32_t* dv = (pass & 1) ? ovals : tvals;
      // digit counts of the run, then their exclusive sum as each digit's base
      base[li] = 0;
      __syncthreads();
      for (uint32_t i = li; i < count; i += 256) atomicAdd(&base[(uint32_t)(sk[first + i] >> shift) & 255u], 1u);
      __syncthreads();
      if (li == 0) {
        uint32_t sum = 0;
        for (int d = 0; d < 256; ++d) {
          const uint32_t c = base[d];
          base[d] = sum;
          sum += c;
        }
      }
      __syncthreads();
      for (uint32_t b = 0; b < count; b += 256) {
        const uint32_t i = b + li;
        const bool valid = i < count;
        uint64_t k = 0;
        uint32_t v = 0, d = 256u;
        if (valid) {
          k = sk[first + i];
          v = sv[first + i];
          d = (uint32_t)(k >> shift) & 255u;
        }
        // lanes of this wave with the same digit
        uint64_t peers = __builtin_amdgcn_ballot_w64(valid);
#pragma unroll
        for (int bit = 0; bit < 8; ++bit) {
          const uint64_t on = __builtin_amdgcn_ballot_w64(valid && ((d >> bit) & 1u));
          peers &= ((d >> bit) & 1u) ? on : ~on;
        }
        const uint32_t before = (uint32_t)__popcll(peers & ((1ull << lane) - 1ull));
        wcount[0][li] = 0;
        wcount[1][li] = 0;
        wcount[2][li] = 0;
        wcount[3][li] = 0;
        __syncthreads();
        if (valid && before == 0) wcount[wave][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
          uint32_t rank = base[d] + before;
          for (uint32_t w = 0; w < wave; ++w) rank += wcount[w][d];
          dk[first + rank] = k;
          dv[first + rank] = v;
        }
        __syncthreads();
        base[li] += wcount[0][li] + wcount[1][li] + wcount[2][li] + wcount[3][li];
        __syncthreads();
      }
    }
    if (PASSES & 1) {  // the last pass wrote the scratch: back into the output
      for (uint32_t i = li; i < count; i += 256) {
        okeys[first + i] = tkeys[first + i];
        ovals[first + i] = tvals[first + i];
      }
    }
    __syncthreads();
  }
}

// ---- small trees: the whole codes in order by counting ------------------------------

// Up to RANK_MAX points, the codes go in order by counting instead of by the radix sort
// and the run ordering (six launches): the point with code k and index i goes to the
// number of points whose (code, index) is smaller, which is the stable order on the whole
// codes that those give.  Workgroup (g, q) counts for the 64 points of group g over part q
// of the codes (gridDim.y parts, so that the n^2 comparisons spread over the whole device;
// its RANK_WAVES waves split the part again, each code read from LDS as a broadcast) and
// adds its counts to the points' ranks; the last of a group's workgroups to finish places
// the group's points.  The ranks and the groups' counters start at zero (tree_keys).
constexpr uint32_t RANK_MAX = 8192, RANK_WAVES = 16;

__global__ void __launch_bounds__(RANK_WAVES * 64) tree_rank_sort(const uint64_t* __restrict__ keys,
                                                                  const uint32_t* __restrict__ vals,
                                                                  uint32_t n, uint32_t* rank, uint32_t* done,
                                                                  uint64_t* __restrict__ okeys,
                                                                  uint32_t* __restrict__ ovals) {
  extern __shared__ uint64_t sk[];  // this workgroup's part of the codes
  __shared__ uint32_t part[RANK_WAVES][64];
  const uint32_t parts = gridDim.y, q = blockIdx.y, g = blockIdx.x;
  const uint32_t k0 = (uint32_t)((uint64_t)n * q / parts), k1 = (uint32_t)((uint64_t)n * (q + 1) / parts);
  for (uint32_t j = k0 + threadIdx.x; j < k1; j += RANK_WAVES * 64) sk[j - k0] = keys[j];
  const uint32_t w = threadIdx.x / 64, li = threadIdx.x & 63;
  const uint32_t i = g * 64 + li, first = g * 64, last = min(first + 64, n);
  const uint64_t my = i < n ? keys[i] : 0ull;
  __syncthreads();
  const uint32_t cnt = k1 - k0;
  const uint32_t a = k0 + cnt * w / RANK_WAVES, b = k0 + cnt * (w + 1) / RANK_WAVES;
  uint32_t r = 0;
  if (b <= first) {  // every code here has a smaller index: equal codes count
    for (uint32_t j = a; j < b; ++j) r += sk[j - k0] <= my ? 1u : 0u;
  } else if (a >= last) {  // every code here has a larger index
    for (uint32_t j = a; j < b; ++j) r += sk[j - k0] < my ? 1u : 0u;
  } else {
    for (uint32_t j = a; j < b; ++j) {
      const uint64_t k = sk[j - k0];
      r += (k < my || (k == my && j < i)) ? 1u : 0u;
    }
  }
  part[w][li] = r;
  __syncthreads();
  if (w != 0) return;
  uint32_t sum = 0;
#pragma unroll
  for (uint32_t p = 0; p < RANK_WAVES; ++p) sum += part[p][li];
  if (i < n) __hip_atomic_fetch_add(&rank[i], sum, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  uint32_t before = 0;
  if (li == 0) before = publish_count(&done[g]);  // after this wave's additions completed
  before = __shfl(before, 0, WAVE);
  if (before + 1 != parts || i >= n) return;
  const uint32_t at = load_coherent(&rank[i]);
  okeys[at] = my;
  ovals[at] = vals[i];
}

// ---- leaves and Karras build ---------------------------------------------------------

// Karras 2012 node i of the sorted codes k[0, N): the range [first, last] it covers, its
// split and the common prefix of its range (barnesHut.js:587-624).
template <int D>
__device__ inline void karras(const uint64_t* __restrict__ skeys, int N, int i, int& first,
                              int& last, int& split, int& prefix) {
  int before = delta<D>(skeys, N, i, i - 1), after = delta<D>(skeys, N, i, i + 1);
  int d = after > before ? 1 : -1;
  int floor_prefix = before < after ? before : after;
  int span = 2;
  while (span < (1 << 30) && delta<D>(skeys, N, i, i + span * d) > floor_prefix) span *= 2;
  int len = 0;
  for (int t = span >> 1; t >= 1; t >>= 1)
    if (delta<D>(skeys, N, i, i + (len + t) * d) > floor_prefix) len += t;
  int j = i + len * d;
  prefix = delta<D>(skeys, N, i, j);
  int s = 0, stride = len;
  for (;;) {
    stride = (stride + 1) >> 1;
    if (delta<D>(skeys, N, i, i + (s + stride) * d) > prefix) s += stride;
    if (stride <= 1) break;
  }
  split = i + s * d + (d < 0 ? d : 0);
  first = i < j ? i : j;
  last = i < j ? j : i;
}

// The reduction works on blocks of REDUCE_BLOCK consecutive sorted points (in the index
// space of all trees); a node is local to a block when all its points lie in it.  A tree
// of at most REDUCE_ONE points is reduced by one workgroup, all of it in LDS.
constexpr uint32_t REDUCE_BLOCK = 512;
constexpr uint32_t REDUCE_ONE = 1024;
constexpr uint32_t PARENT_LOCAL = 0x80000000u;  // in a parent reference: the parent is local

// Writes binary node `node` (= off + i) of a tree whose sorted points are [off, off + N):
// its links, and itself (with PARENT_LOCAL when local to a reduction block of `block`
// points) as the parent of its two children.
template <int D>
__device__ inline void karras_write(const uint64_t* __restrict__ skeys, int N, int i, uint32_t off,
                                    uint32_t block, uint4* __restrict__ links,
                                    uint32_t* __restrict__ leaf_parent, uint8_t* __restrict__ cut_start = nullptr) {
  int first, last, split, prefix;
  karras<D>(skeys, N, i, first, last, split, prefix);
  // cut_start (subtree-aligned workgroups of the reduction): a node is local when it covers
  // at most `block` points, and point split + 1 starts a maximal such subtree exactly when
  // this node, which parts points split and split + 1, covers more (every gap between two
  // sorted points is parted by one node, so every flag but point 0's is written once)
  if (cut_start) {
    cut_start[split + 1] = (uint8_t)(last - first + 1 > (int)block);
    if (i == 0) cut_start[0] = 1;
  }
  const uint32_t node = off + (uint32_t)i, a = off + (uint32_t)split, b = a + 1u;
  uint32_t left = a | (first == split ? REF_POINT : 0u);
  uint32_t right = b | (last == split + 1 ? REF_POINT : 0u);
  links[node].x = left;
  links[node].y = right;
  links[node].z = depth_of<D>(prefix);
  const bool local = cut_start ? last - first + 1 <= (int)block
                               : (off + (uint32_t)first) / block == (off + (uint32_t)last) / block;
  const uint32_t up = node | (local ? PARENT_LOCAL : 0u);
  if (left & REF_POINT) leaf_parent[a] = up;
  else links[a].w = up;
  if (right & REF_POINT) leaf_parent[b] = up;
  else links[b].w = up;
  if (i == 0) links[node].w = REF_NONE;
}


// One tree's leaves (the sorted points' records and leaf depths) and binary nodes (Karras
// ranges and splits, parent links for the climb) in one launch: thread s does sorted
// point s and node s; neither part reads what the other writes.
template <int D>
__global__ void __launch_bounds__(256) tree_leaves_build(const uint64_t* __restrict__ skeys,
                                                         const uint32_t* __restrict__ order,
                                                         const float4* __restrict__ pts, uint32_t n,
                                                         uint32_t block, float4* __restrict__ leaves,
                                                         uint32_t* __restrict__ leaf_depth,
                                                         uint32_t* __restrict__ flags,
                                                         uint4* __restrict__ links,
                                                         uint32_t* __restrict__ leaf_parent,
                                                         uint8_t* __restrict__ cut_start) {
  const uint32_t s = blockIdx.x * 256 + threadIdx.x;
  if (s >= n) return;
  leaves[s] = pts[order[s]];
  uint32_t depth = 0;
  if (n > 1) {
    int a = delta<D>(skeys, (int)n, (int)s, (int)s - 1), b = delta<D>(skeys, (int)n, (int)s, (int)s + 1);
    depth = depth_of<D>(a > b ? a : b) + 1;
  }
  leaf_depth[s] = depth;
  if (s + 1 < n) {
    flags[s] = 0;
    karras_write<D>(skeys, (int)n, (int)s, 0u, block, links, leaf_parent, cut_start);
  }
}

// ---- bottom-up reduction + collapse ---------------------------------------------------

// A node's sums and centre as the reduction hands them from child to parent.
struct NodeAcc {
  double a, w, gx, gy, gz;  // NodeSum
  double cx, cy, cz;        // centre (0 when the node has no weight)
};

// One workgroup's share of the reduction in LDS: its points and its binary nodes (a local
// node's index lies within its own points, so both are indexed by index - first point of
// the workgroup), and the sums of its local nodes.
template <uint32_t B>
struct ReduceLDS {
  // the one-workgroup reduction also keeps each node's centre (as its record holds it), so
  // a parent need not divide again: LDS to spare there, and its climb is the whole tick's
  // chain; with many workgroups per CU the room is worth more
  static constexpr bool CACHE = B >= REDUCE_ONE;
  static constexpr uint32_t CB = CACHE ? B : 1;
  double a[B], w[B], gx[B], gy[B], gz[B];
  double cx[CB], cy[CB], cz[CB];
  float4 leaf[B];
  uint4 link[B];
  uint32_t flag[B];
  uint8_t same[B];  // NodeSum::same of the local nodes
};

// The factor that turns a node's weighted sums into its |value|-weighted centre: 1 / w (0
// when it has no weight).  One division for the three coordinates, within an ulp of g / w
// each (the size table 1k-200k 0.8-3.6 % faster than three divisions, which the climb's
// chain of nodes waits on).
__device__ inline double centre_factor(double w) { return w > 0 ? 1.0 / w : 0.0; }

// Adds child k of a node at depth parent_depth to `out`: a point (a d3 child one level
// below the parent's cell), a deeper node (a d3 child, weighed by |its strength one level
// below the parent's cell|) or a node in the same cell (its children are the parent's:
// its sums pass through).  LOCAL: the node and so its children lie within this
// workgroup's points and are read from LDS; otherwise from global memory, sums coherently.
template <int D, bool LOCAL, uint32_t B>
__device__ inline void child_sum(uint32_t k, uint32_t parent_depth, uint32_t lo,
                                 const ReduceLDS<B>& S, const float4* leaves, const uint4* links,
                                 const NodeRec* rec, const NodeSum* sums, NodeSum& out) {
  if (k & REF_POINT) {
    const uint32_t j = k & ~REF_POINT;
    const float4 q = LOCAL ? S.leaf[j - lo] : leaves[j];
    double sv = (double)q.w;
    out.a += sv * damp<D>((int)parent_depth + 1);
    double w = fabs(sv);
    out.w += w;
    out.gx += w * (double)q.x;
    if (D > 1) out.gy += w * (double)q.y;
    if (D > 2) out.gz += w * (double)q.z;
    return;
  }
  const uint32_t q = k - lo;
  const uint32_t kd = LOCAL ? S.link[q].z : links[k].z;
  const double ca = LOCAL ? S.a[q] : load_coherent(&sums[k].a);
  out.a += ca;
  if (kd > parent_depth) {
    double w = fabs(ca * damp<D>(-(int)(parent_depth + 1)));
    out.w += w;
    if (LOCAL && ReduceLDS<B>::CACHE) {
      out.gx += w * S.cx[q];
      if (D > 1) out.gy += w * S.cy[q];
      if (D > 2) out.gz += w * S.cz[q];
    } else if (LOCAL) {
      const double rk = centre_factor(S.w[q]);
      out.gx += w * (S.gx[q] * rk);
      if (D > 1) out.gy += w * (S.gy[q] * rk);
      if (D > 2) out.gz += w * (S.gz[q] * rk);
    } else {
      out.gx += w * load_coherent(&rec[k].cx);
      if (D > 1) out.gy += w * load_coherent(&rec[k].cy);
      if (D > 2) out.gz += w * load_coherent(&rec[k].cz);
    }
  } else {
    out.w += LOCAL ? S.w[q] : load_coherent(&sums[k].w);
    out.gx += LOCAL ? S.gx[q] : load_coherent(&sums[k].gx);
    if (D > 1) out.gy += LOCAL ? S.gy[q] : load_coherent(&sums[k].gy);
    if (D > 2) out.gz += LOCAL ? S.gz[q] : load_coherent(&sums[k].gz);
  }
}

// d3's children of binary node p (barnesHut.js:641-671): its two children, with every
// internal child in the same d3 cell replaced by its own two children, DIMS - 1 times.
template <int D>
__device__ inline void collapse_into(uint4 L, const uint4* links, uint32_t* kids) {
  uint32_t m = 0;
  const uint32_t pair[2] = {L.x, L.y};
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    uint32_t a = pair[c];
    if (D > 1 && !(a & REF_POINT) && links[a].z == L.z) {
      uint4 A = links[a];
      const uint32_t pa[2] = {A.x, A.y};
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        uint32_t b = pa[e];
        if (D > 2 && !(b & REF_POINT) && links[b].z == L.z) {
          uint4 B = links[b];
          kids[m++] = B.x;
          kids[m++] = B.y;
        } else {
          kids[m++] = b;
        }
      }
    } else {
      kids[m++] = a;
    }
  }
  for (; m < 8; ++m) kids[m] = REF_NONE;
}

// d3's children of every binary node of a batch's trees (a pass of its own: it depends on
// the topology only; a single tree's reduce threads do it after their climb instead).
template <int D>
__global__ void __launch_bounds__(256) tree_collapse(uint32_t n, const uint32_t* __restrict__ seg_of,
                                                     const uint32_t* __restrict__ seg_job,
                                                     const BJob* __restrict__ jobs,
                                                     const uint4* __restrict__ links, NodeRec* rec) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i + 1 >= n) return;
  if (seg_of) {  // a batch: the last index of each job's segment is not a node
    const BJob& B = jobs[seg_job[seg_of[i]]];
    if (i - B.tree_off + 1 >= B.n) return;
  }
  uint32_t kids[8];
  collapse_into<D>(links[i], links, kids);
  uint4* out = (uint4*)rec[i].kids;
  out[0] = make_uint4(kids[0], kids[1], kids[2], kids[3]);
  out[1] = make_uint4(kids[4], kids[5], kids[6], kids[7]);
}

// One thread per point climbs towards the root; the second of two siblings to reach a
// node combines them.  Nodes whose points all lie within this workgroup's points are
// combined through LDS (the workgroup's points and nodes are staged there first, and the
// flags and sums live there); above them the hand-off is through global memory, coherent
// across the device.  Every node's record goes to global memory for the walk; a node whose
// parent is not local also leaves its sums there, coherently.
//
// A node covers its children's points, so once a climb leaves the workgroup's own nodes
// it never returns: the climb is two loops.  The first touches only LDS for its hand-off
// (a relaxed LDS atomic between LDS-only fences), so it never waits for the node records
// it writes to global memory on the way; the second publishes through the device.

// Combines node p's children (a node whose sibling has arrived) and writes its record;
// returns its parent reference (with the locality bit).
// Whether child k of an equal-code node has all its points at one position (a point has).
template <bool LOCAL, uint32_t B>
__device__ inline bool child_same(uint32_t k, uint32_t lo, const ReduceLDS<B>& S, const NodeSum* sums) {
  if (k & REF_POINT) return true;
  return LOCAL ? S.same[k - lo] != 0 : load_coherent(&sums[k].same) != 0u;
}

// Whether sorted points a and a + 1 lie at one position.
template <int D, bool LOCAL, uint32_t B>
__device__ inline bool same_position(uint32_t a, uint32_t lo, const ReduceLDS<B>& S, const float4* leaves) {
  const float4 u = LOCAL ? S.leaf[a - lo] : leaves[a];
  const float4 v = LOCAL ? S.leaf[a + 1 - lo] : leaves[a + 1];
  return u.x == v.x && (D < 2 || u.y == v.y) && (D < 3 || u.z == v.z);
}

template <int D, bool LOCAL, uint32_t B>
__device__ inline uint32_t reduce_node(uint32_t p, uint32_t lo, ReduceLDS<B>& S, const float4* leaves,
                                       const uint4* links, NodeRec* rec, NodeSum* sums) {
  constexpr uint32_t LEVELS = (uint32_t)Dims<D>::LEVELS;
  const uint4 L = LOCAL ? S.link[p - lo] : links[p];
  NodeSum sum = {};
  child_sum<D, LOCAL, B>(L.x, L.z, lo, S, leaves, links, rec, sums, sum);
  child_sum<D, LOCAL, B>(L.y, L.z, lo, S, leaves, links, rec, sums, sum);
  // A node at the deepest level is a run of equal codes (unequal codes part above it).  The
  // codes put points at one position there, one level below the run; d3 keeps them as one
  // leaf, their chain, in the cell where they part from the other points: one level below
  // the run's parent, its strength the chain's undamped sum (src/manyBody.js:61-68).  So
  // the root of a run whose points all coincide takes that depth, and the nodes inside a
  // run record depth 0: below any parent's cell, so a walk always opens them.  Codes and
  // topology stay as they are.
  uint32_t vdepth = L.z, rdepth = L.z;  // the depth of the value, and of the record
  int lead = (int)L.z + 1;              // the depth of the node's own points
  uint32_t same = 0;
  if (L.z == LEVELS) {
    same = child_same<LOCAL, B>(L.x, lo, S, sums) && child_same<LOCAL, B>(L.y, lo, S, sums) &&
           same_position<D, LOCAL, B>(L.x & ~REF_POINT, lo, S, leaves);
    const bool has_parent = L.w != REF_NONE;
    uint32_t pd = 0;
    if (has_parent) {
      const uint32_t q = L.w & ~PARENT_LOCAL;
      pd = LOCAL && (L.w & PARENT_LOCAL) ? S.link[q - lo].z : links[q].z;
    }
    if (has_parent && pd == LEVELS) {
      rdepth = 0;
    } else if (same) {
      const uint32_t dl = has_parent ? pd + 1 : 0u;
      sum.a *= damp<D>((int)dl - (int)(LEVELS + 1));
      vdepth = rdepth = dl;
      lead = (int)dl;
    }
  }
  const double rw = centre_factor(sum.w);
  const double cx = sum.gx * rw;
  const double cy = D > 1 ? sum.gy * rw : 0.0;
  const double cz = D > 2 ? sum.gz * rw : 0.0;
  double value = sum.a * damp<D>(-(int)vdepth);
  // d3 skips a cell whose strengths add up to exactly 0; sums in another order leave a
  // residue of rounding, so a cancellation to 2^-40 of the cell's weight counts as 0
  double scale = sum.w * damp<D>(lead);
  if (fabs(sum.a) <= fabs(scale) * 0x1p-40 || !(sum.w > 0)) value = 0.0;
  NodeRec& r = rec[p];
  r.value = (float)value;
  r.depth = rdepth;
  // a node's parent is local only if the node is (the parent covers the node's points)
  const bool parent_local = LOCAL && L.w != REF_NONE && (L.w & PARENT_LOCAL);
  if (parent_local) {
    const uint32_t q = p - lo;
    S.same[q] = (uint8_t)same;
    S.a[q] = sum.a;
    S.w[q] = sum.w;
    S.gx[q] = sum.gx;
    S.gy[q] = sum.gy;
    S.gz[q] = sum.gz;
    if (ReduceLDS<B>::CACHE) {
      S.cx[q] = cx;
      if (D > 1) S.cy[q] = cy;
      if (D > 2) S.cz[q] = cz;
    }
    r.cx = cx;
    r.cy = cy;
    r.cz = cz;
  } else {
    store_coherent(&r.cx, cx);
    store_coherent(&r.cy, cy);
    store_coherent(&r.cz, cz);
    store_coherent(&sums[p].a, sum.a);
    store_coherent(&sums[p].w, sum.w);
    store_coherent(&sums[p].gx, sum.gx);
    store_coherent(&sums[p].gy, sum.gy);
    store_coherent(&sums[p].gz, sum.gz);
    store_coherent(&sums[p].same, same);
  }
  return L.w;
}

// The climb of point s (n >= 2).
template <int D, uint32_t B>
__device__ inline void reduce_climb(uint32_t s, uint32_t n, uint32_t lo, ReduceLDS<B>& S,
                                    const float4* __restrict__ leaves,
                                    const uint32_t* __restrict__ leaf_parent,
                                    const uint4* __restrict__ links, uint32_t* flags, NodeRec* rec,
                                    NodeSum* sums) {
  uint32_t up = leaf_parent[s];  // every point has a parent when n >= 2
  uint32_t guard = 0;
  // this workgroup's own nodes: flags, sums, points and links in LDS (REF_NONE, above the
  // root, has the locality bit set too)
  while (up != REF_NONE && (up & PARENT_LOCAL)) {
    const uint32_t p = up & ~PARENT_LOCAL;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    const uint32_t old =
        __hip_atomic_fetch_add(&S.flag[p - lo], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
    if (old == 0 || ++guard >= n) return;  // the sibling is still on its way
    up = reduce_node<D, true, B>(p, lo, S, leaves, links, rec, sums);
  }
  // shared nodes: their children left their sums in global memory (their parent, p, is
  // not local to their writer)
  while (up != REF_NONE) {
    const uint32_t p = up;
    if (publish_count(&flags[p]) == 0 || ++guard >= n) return;
    up = reduce_node<D, false, B>(p, lo, S, leaves, links, rec, sums);
  }
}

// A tree of at most REDUCE_ONE points built and reduced by one workgroup, one point per
// thread, in one launch instead of five (tree_bbox, tree_keys, tree_rank_sort,
// tree_leaves_build, tree_reduce), with the same arithmetic: the box over the waves and
// the cube; the codes; their stable order; the leaves, leaf depths and Karras nodes, the
// nodes' links and parents kept in LDS for the climb; the reduction and d3's children.
//
// The order: each wave sorts its 64 (code, index) pairs in registers (bitonic), then runs of
// 64, 128, ... are merged between two LDS buffers.  A pair's place in the merge of its run
// with the sibling run is its place in its own run plus the number of the sibling's pairs
// below it (a binary search); the pairs are all distinct, so this is the order on (code,
// index), the stable order that the rank sort and the radix sort give.  The counting of the
// rank sort would take n reads per point on the one CU.
template <int D>
__global__ void __launch_bounds__(REDUCE_ONE) tree_small(const float* __restrict__ x, const float* __restrict__ y,
                                                         const float* __restrict__ z,
                                                         const float* __restrict__ strength, uint32_t n,
                                                         CubeInfo* __restrict__ cube, uint32_t* __restrict__ counters,
                                                         uint64_t* __restrict__ skeys, uint32_t* __restrict__ order,
                                                         float4* __restrict__ leaves, uint32_t* __restrict__ leaf_depth,
                                                         uint4* __restrict__ links, uint32_t* flags, NodeRec* rec,
                                                         NodeSum* sums, LinkFuse link) {
  constexpr uint32_t B = REDUCE_ONE, W = B / WAVE;
  if (blockIdx.x > 0) {  // the workgroups past the first run the link pass before this force
    // (a whole wave per step: long rows take the wave)
    const uint32_t lane = threadIdx.x & (WAVE - 1);
    for (uint32_t w = (blockIdx.x - 1) * B + (threadIdx.x & ~(WAVE - 1)); w < n; w += (gridDim.x - 1) * B)
      link_wave<D>(w + lane, w + lane < n, SnapDirect<D>{x, y, z, link.vx, link.vy, link.vz}, link.vx, link.vy,
                   link.vz, link.ox, link.oy, link.oz, link.row, link.ent, link.alpha, link.step, link.seed);
    return;
  }
  __shared__ ReduceLDS<B> S;
  __shared__ uint64_t s_key[2][B];
  __shared__ uint32_t s_idx[2][B];
  __shared__ uint32_t s_parent[B];
  __shared__ float4 s_pts[B];
  __shared__ uint32_t s_m[W][6];
  __shared__ float s_origin[3];
  __shared__ int32_t s_shift;
  const uint32_t i = threadIdx.x, lane = i & (WAVE - 1), wave = i / WAVE;
  if (i < TREE_COUNTERS) counters[i] = 0;
  // the box: per wave, then over the waves in order-preserving words (as tree_bbox)
  float p[3] = {0.f, 0.f, 0.f};
  float lo[3] = {3.4e38f, 3.4e38f, 3.4e38f}, hi[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
  if (i < n) {
    p[0] = x[i];
    if (D > 1) p[1] = y[i];
    if (D > 2) p[2] = z[i];
#pragma unroll
    for (int a = 0; a < 3; ++a) lo[a] = p[a], hi[a] = p[a];
  }
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      lo[a] = fminf(lo[a], __shfl_xor(lo[a], o, WAVE));
      hi[a] = fmaxf(hi[a], __shfl_xor(hi[a], o, WAVE));
    }
  if (lane == 0)
    for (int a = 0; a < 3; ++a) s_m[wave][a] = float_ordered(lo[a]), s_m[wave][3 + a] = float_ordered(hi[a]);
  __syncthreads();
  if (i == 0) {
    uint32_t m[6] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0u, 0u, 0u};
    for (uint32_t w = 0; w < W; ++w)
      for (int a = 0; a < 3; ++a) m[a] = min(m[a], s_m[w][a]), m[3 + a] = max(m[3 + a], s_m[w][3 + a]);
    float blo[3], bhi[3], origin[3], size;
    int32_t shift;
    for (int a = 0; a < 3; ++a) blo[a] = float_unordered(m[a]), bhi[a] = float_unordered(m[3 + a]);
    cover_of<D>(blo, bhi, origin, size, shift);
    for (int a = 0; a < 3; ++a) s_origin[a] = origin[a], cube->origin[a] = origin[a];
    s_shift = shift;
    cube->size = size;
    cube->shift = shift;
  }
  __syncthreads();
  // the codes in order: each wave sorts its 64 (code, index) pairs in registers (a bitonic
  // network over the lanes, padding lanes last), then the waves' runs are merged in LDS;
  // the result lands in buffer 0
  uint32_t levels = 0;  // merge levels after the waves' own sorts
  while ((64u << levels) < n) ++levels;
  uint32_t cur = levels & 1u;
  uint64_t code = ~0ull;
  uint32_t idx = i;
  if (i < n) {
    const float origin[3] = {s_origin[0], s_origin[1], s_origin[2]};
    code = point_key<D>(p, origin, s_shift);
    s_pts[i] = make_float4(p[0], p[1], p[2], strength[i]);
  }
#pragma unroll
  for (uint32_t k = 2; k <= WAVE; k <<= 1)
#pragma unroll
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      const uint64_t pc = __shfl_xor(code, (int)j, WAVE);
      const uint32_t pi = (uint32_t)__shfl_xor((int)idx, (int)j, WAVE);
      const bool p_less = pc < code || (pc == code && pi < idx);
      const bool keep_min = ((lane & j) == 0) == ((lane & k) == 0);  // lower of an ascending pair
      if (keep_min ? p_less : !p_less) code = pc, idx = pi;
    }
  uint32_t pos = i;  // lane l of wave w holds the wave's l-th pair
  if (code != ~0ull) {
    s_key[cur][pos] = code;
    s_idx[cur][pos] = idx;
  }
  __syncthreads();
  for (uint32_t r = WAVE; r < n; r <<= 1, cur ^= 1u) {
    if (code != ~0ull) {
      const uint32_t sib = (pos ^ r) & ~(r - 1u);  // first place of the sibling run
      const uint32_t len = sib < n ? min(r, n - sib) : 0u;
      uint32_t below = 0;  // the sibling's pairs below (code, idx): branch-free lower bound
      for (uint32_t step = r; step > 0; step >>= 1) {
        if (below + step <= len) {
          const uint32_t j = sib + below + step - 1u;
          const uint64_t k = s_key[cur][j];
          if (k < code || (k == code && s_idx[cur][j] < idx)) below += step;
        }
      }
      pos = (pos & ~(2u * r - 1u)) + (pos & (r - 1u)) + below;
      s_key[cur ^ 1u][pos] = code;
      s_idx[cur ^ 1u][pos] = idx;
    }
    __syncthreads();
  }
  // sorted place pos: the leaf, its code and point
  if (code != ~0ull) {
    const float4 leaf = s_pts[idx];
    S.leaf[pos] = leaf;
    leaves[pos] = leaf;
    skeys[pos] = code;
    order[pos] = idx;
  }
  S.flag[i] = 0;
  // sorted point i: its depth; binary node i: its links, and its parent links (in LDS)
  const uint64_t* sorted = s_key[0];
  if (i < n) {
    uint32_t depth = 0;
    if (n > 1) {
      const int a = delta<D>(sorted, (int)n, (int)i, (int)i - 1), b = delta<D>(sorted, (int)n, (int)i, (int)i + 1);
      depth = depth_of<D>(a > b ? a : b) + 1;
    }
    leaf_depth[i] = depth;
    if (i + 1 < n) karras_write<D>(sorted, (int)n, (int)i, 0u, B, S.link, s_parent);
  }
  __syncthreads();
  if (i >= n || n < 2) return;
  // every point climbs; binary nodes are 0 .. n - 2
  if (i + 1 < n) links[i] = S.link[i];
  reduce_climb<D, B>(i, n, 0u, S, leaves, s_parent, S.link, flags, rec, sums);
  if (i + 1 >= n) return;
  uint32_t kids[8];
  collapse_into<D>(S.link[i], S.link, kids);
  uint4* out = (uint4*)rec[i].kids;
  out[0] = make_uint4(kids[0], kids[1], kids[2], kids[3]);
  out[1] = make_uint4(kids[4], kids[5], kids[6], kids[7]);
}

// COLLAPSE (one tree): thread s also lists d3's children of binary node s once its climb
// is over (the topology is complete before the kernel starts; the records' children are
// not what the climb writes), which saves the collapse pass its launch.
//
// cut_start (subtree-aligned workgroups): workgroup b takes the maximal subtrees of at most
// W = B / 2 points that start among points [b W, (b + 1) W) -- fewer than 2W points, from
// the first such start to the next start at or after (b + 1) W -- so that only nodes of
// more than W points are shared, and the chain of device-wide hand-offs is the depth of
// those (8 levels against 16 at 5,000 points, 20 against 30 at 1M).  Otherwise workgroup b
// takes points [b B, (b + 1) B).
template <int D, uint32_t B, bool COLLAPSE>
__global__ void __launch_bounds__(B) tree_reduce(uint32_t n, const float4* __restrict__ leaves,
                                                            const uint32_t* __restrict__ leaf_parent,
                                                            const uint4* __restrict__ links,
                                                            uint32_t* flags, NodeRec* rec, NodeSum* sums,
                                                            const uint8_t* __restrict__ cut_start = nullptr) {
  __shared__ ReduceLDS<B> S;
  __shared__ uint32_t s_span[2];
  uint32_t lo = blockIdx.x * B, count = B;
  if (cut_start) {
    constexpr uint32_t W = B / 2;
    const uint32_t w0 = blockIdx.x * W, i = w0 + threadIdx.x, half = threadIdx.x < W ? 0u : 1u;
    if (threadIdx.x < 2) s_span[threadIdx.x] = 0xffffffffu;
    __syncthreads();
    // the first start in each half: one LDS atomic per wave (a wave lies in one half)
    const uint64_t starts = __ballot(i < n && cut_start[i] != 0);
    if (starts && (threadIdx.x & (WAVE - 1)) == 0) atomicMin(&s_span[half], w0 + (threadIdx.x & ~(WAVE - 1)) + (uint32_t)__ffsll((long long)starts) - 1u);
    __syncthreads();
    lo = s_span[0];
    const uint32_t end = s_span[1] < n ? s_span[1] : n;
    count = lo < end ? end - lo : 0u;  // none: this window only continues the previous one's subtree
  }
  const uint32_t s = lo + threadIdx.x;
  S.flag[threadIdx.x] = 0;
  if (threadIdx.x < count && s < n) {
    S.leaf[threadIdx.x] = leaves[s];
    S.link[threadIdx.x] = links[s];  // index s is a node or a tree's last point (never read then)
  }
  __syncthreads();
  if (threadIdx.x >= count || s >= n || n < 2) return;
  reduce_climb<D, B>(s, n, lo, S, leaves, leaf_parent, links, flags, rec, sums);
  if (COLLAPSE && s + 1 < n) {
    uint32_t kids[8];
    collapse_into<D>(links[s], links, kids);
    uint4* out = (uint4*)rec[s].kids;
    out[0] = make_uint4(kids[0], kids[1], kids[2], kids[3]);
    out[1] = make_uint4(kids[4], kids[5], kids[6], kids[7]);
  }
}

// ---- walk ---------------------------------------------------------------------------

constexpr int WALK_WAVES = 4;
constexpr int WALK_STACK = 48;   // stack entries per wave in LDS (a walk at 1M points peaks near 30)
constexpr int WALK_DEPTH = 192;  // stack entries per wave in all; 7 * 21 + 8 = 155 in 3D for distinct codes

// Per wave, in LDS: the top STACK entries of the stack of nodes to open, each as its
// depth, the lanes that must open it and its d3 children (so that opening it needs no load
// of its own record) -- entry STACK is where an entry popped from the spill area lands;
// a staging area where one coalesced load puts the 64-byte records of all children of the
// node being opened (points: their float4); the wave's points and their force sums from
// the packed mode.  Entries from STACK to WALK_DEPTH live in the wave's spill area in
// global memory, which deep walks (many coincident points) reach.
//
// The cheaper mode adds a pool of single-point openings: a node that at most POOL_K of the
// tile's points must open goes into the pool as one entry per point (node, point, the depth
// its children's chains start at) instead of onto the stack, and a pool step opens up to
// 64 / 2^D entries at once, lane (entry, child) deciding one child for one point -- deep in
// the tree, where a wave's opening served a few points with a few lanes each, every lane
// now has a child to decide.  Entries that find the pool full go onto the stack as
// one-point entries.
constexpr uint32_t POOL_K = 24;     // openings by at most this many points go to the pool
constexpr uint32_t POOL_SIZE = 64;  // pool entries per wave
// Waves per SIMD the pooled walk is compiled for: 6 (at most 80 registers).
constexpr int POOL_WAVES = 6;

template <int STACK, bool POOLED = false>
struct WalkLDS {
  uint32_t depth[STACK + 1];
  uint64_t mask[STACK + 1];
  uint32_t kids[STACK + 1][8];
  uint32_t stage[8][16];
  float4 body[WAVE];      // x, y, z, node index (as bits)
  float acc[3][WAVE];
  uint32_t list[WAVE];    // the points opening the current node, in lane order
  uint32_t cur[8];        // the children of the node being opened
  uint64_t ballots[8];
  uint2 pool[POOLED ? POOL_SIZE : 1];  // node, point << 8 | chain start depth
};

// The jiggled displacement of two coincident bodies: out of line, it is rare.
__device__ __noinline__ float4 coincident_far(uint32_t i, uint32_t j, int ndim, uint32_t step_no,
                                              uint32_t seed) {
  float dx, dy, dz;
  coincident(i, j, ndim, step_no, seed, dx, dy, dz);
  return make_float4(dx, dy, dz, dx * dx + dy * dy + dz * dz);
}

// c^k in f32 for k >= 0 (the damping of a cell k levels above the node's deepest cell).
template <int D>
__device__ inline float dampf(uint32_t k) {
  if (D == 2) return 1.0f;
  const float odd = (k & 1u) ? (D == 3 ? 0.70710678118654752f : 1.41421356237309505f) : 1.0f;
  return ldexpf(odd, D == 3 ? -(int)(k >> 1) : (int)(k >> 1));
}

// How a force term a * b reaches its accumulator: ACC_FUSED acc += a * b (a fused
// multiply-add where the compiler likes), ACC_ADD acc + a * b with the product rounded on
// its own (so a running sum of several terms rounds like a sum of the rounded terms),
// ACC_SET acc = a * b.
enum : int { ACC_FUSED = 0, ACC_ADD = 1, ACC_SET = 2 };
template <int ACC>
__device__ inline void accumulate(float& acc, float a, float b) {
#pragma clang fp contract(off)
  if (ACC == ACC_SET) {
    acc = a * b;
  } else if (ACC == ACC_ADD) {
    acc += a * b;
  }
}
template <>
__device__ inline void accumulate<ACC_FUSED>(float& acc, float a, float b) {
  acc += a * b;
}

// dx^2 + dy^2 + dz^2 as one fixed chain of fused multiply-adds, for the exact-mode terms
// (ACC_SET, ACC_ADD): left to the compiler, the sum is contracted differently in different
// kernels, and the exact walks (tiled, one point per lane group, the batch's) must give the
// same terms bit for bit.  The cheaper mode (ACC_FUSED) leaves it to the compiler.
template <int D, typename T>
__device__ inline T sq_len(T dx, T dy, T dz) {
  T l = dx * dx;
  if (D > 1) l = __builtin_fma(dy, dy, l);
  if (D > 2) l = __builtin_fma(dz, dz, l);
  return l;
}

// 1 / l' for d3's squared distance l, where below distanceMin l' = sqrt(dmin2 l)
// (manyBody.js:95): rsq(l * max(l, dmin2)) either way, one reciprocal square root.
__device__ inline float inv_dist2(float l, float dmin2) {
  return __builtin_amdgcn_rsqf(l * (l < dmin2 ? dmin2 : l));  // (a select: fmaxf quiets NaNs first)
}

// d3's opening test w * w / theta2 < l (manyBody.js:85) for a cell at depth d: its width is
// size / 2^d (every split halves it exactly), so w * w / theta2 is exactly thr0 / 4^d with
// thr0 = size * size / theta2, the root's.
__device__ inline double open_below(double thr0, uint32_t d) {
  return ldexp(thr0, -2 * (int)d);
}

// The force of one node (its cell chain from depth `top`; thr0 = size * size / theta2 of
// the tree) on one body, with d3's rules (barnesHut.js:719-746, manyBody.js:73-99).
// `cell` is the node's index within its tree (the jiggle's stand-in index is n + cell).
// Returns whether the body is done with the node (false: it must open it).  ACC: how the
// force reaches fx, fy, fz (accumulate).
template <int D, int ACC>
__device__ inline bool cell_force(double mx, double my, double mz, uint32_t self_node, uint32_t n,
                                  uint32_t cell, uint32_t top, double thr0, const NodeHdr& h,
                                  double dmax2, float dmin2f, float alpha,
                                  uint32_t step_no, uint32_t seed, float& fx, float& fy, float& fz,
                                  uint32_t& inter) {
  if (top > h.depth) return false;  // points sharing the smallest cell: open
  if (h.value == 0.0f) return true;  // d3 skips cells with no strength
  const double dx = h.cx - mx, dy = D > 1 ? h.cy - my : 0.0, dz = D > 2 ? h.cz - mz : 0.0;
  const double l = ACC == ACC_FUSED ? dx * dx + dy * dy + dz * dz : sq_len<D>(dx, dy, dz);
  // The thresholds of the chain shrink by 4 (exactly) per level, so some level accepts the
  // cell exactly when its deepest one does: one test decides opening.  The level d3 stops at
  // is the first that accepts, found only for a chain of more than one level.
  if (!(open_below(thr0, h.depth) < l)) return false;
  float vd = h.value;
  if (h.depth != top) {
    uint32_t Dd = top;
    while (!(open_below(thr0, Dd) < l)) ++Dd;
    vd = h.value * dampf<D>(h.depth - Dd);
  }
  if (l >= dmax2) return true;
  float ex = (float)dx, ey = (float)dy, ez = (float)dz, fl = (float)l;
  if (l == 0.0) {
    const float4 j = coincident_far(self_node, n + cell, D, step_no, seed);
    ex = j.x, ey = j.y, ez = j.z, fl = j.w;
  }
  const float f = vd * alpha * inv_dist2(fl, dmin2f);
  accumulate<ACC>(fx, ex, f);
  if (D > 1) accumulate<ACC>(fy, ey, f);
  if (D > 2) accumulate<ACC>(fz, ez, f);
  ++inter;
  return true;
}

// The force of one point (node index `other`) on one body (barnesHut.js:705-717,
// manyBody.js:99-113).
template <int D, int ACC>
__device__ inline void point_force(float mx, float my, float mz, uint32_t self_node, float4 q,
                                   const uint32_t* order, uint32_t j, uint32_t off, float dmax2f,
                                   float dmin2f, float alpha, uint32_t step_no, uint32_t seed,
                                   float& fx, float& fy, float& fz, uint32_t& inter) {
  float dx = q.x - mx, dy = D > 1 ? q.y - my : 0.f, dz = D > 2 ? q.z - mz : 0.f;
  float l = ACC == ACC_FUSED ? dx * dx + dy * dy + dz * dz : sq_len<D>(dx, dy, dz);
  if (!(l < dmax2f)) return;
  if (l == 0.0f) {
    const float4 jj = coincident_far(self_node, order[j] - off, D, step_no, seed);
    dx = jj.x, dy = jj.y, dz = jj.z, l = jj.w;
  }
  const float f = q.w * alpha * inv_dist2(l, dmin2f);
  accumulate<ACC>(fx, dx, f);
  if (D > 1) accumulate<ACC>(fy, dy, f);
  if (D > 2) accumulate<ACC>(fz, dz, f);
  ++inter;
}

// The sum over each aligned group of NK (2, 4 or 8) lanes, in registers (DPP): the two
// halves of a group of 8 mirrored onto each other, then the pairs of a quad, then
// neighbours; every lane of a group ends with the group's sum.  All lanes must be active.
// (bound_ctrl set: with every row and bank enabled and these patterns, no lane is left out,
// and the compiler folds the move into the add; the add is kept from fusing with the
// product that made x, which would need the moved value in a register of its own)
template <int CTRL>
__device__ inline float dpp_plus(float x) {
#pragma clang fp contract(off)
  return x + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, 0xf, 0xf, true));
}
template <int NK>
__device__ inline float group_sum(float x) {
  if (NK >= 8) x = dpp_plus<0x141>(x);  // row_half_mirror: lane i + lane 7 - i
  if (NK >= 4) x = dpp_plus<0x4e>(x);   // quad_perm(2, 3, 0, 1)
  return dpp_plus<0xb1>(x);             // quad_perm(1, 0, 3, 2)
}

// A stack entry beyond the LDS part.
struct alignas(16) SpillEntry {
  uint64_t mask;
  uint32_t depth, pad;
  uint32_t kids[8];
};

// One tree as a walk sees it: a single simulation (off 0) or one job of a batch, whose
// points and nodes are [off, off + n) of the shared arrays.
struct WalkTree {
  uint32_t off, n;
  double size, thr0, dmax2;  // thr0 = size * size / theta2
  float dmin2f, dmax2f, alpha;
  uint32_t step_no, seed;
  float *vx, *vy, *vz;  // the velocities, indexed by the node index within the tree
};

struct WalkCount {
  uint64_t opened = 0, lane_visits = 0;
  uint32_t inter = 0, err = 0, max_sp = 0;
};

// One wave walks the union of the traversals of sorted points [base, base + count) of one
// tree (count <= 64) and adds their forces to the velocities.  POOLED: the cheaper mode
// sends openings by few points to the pool of single-point openings.
template <int D, int STACK, bool EXACT, bool POOLED>
__device__ __forceinline__ void walk_tile(WalkLDS<STACK, POOLED>& L, const float4* __restrict__ leaves,
                                          const uint32_t* __restrict__ order,
                                          const NodeRec* __restrict__ rec, const WalkTree& T,
                                          uint32_t base, uint32_t count, uint32_t packed_max,
                                          uint32_t lane, SpillEntry* spill, WalkCount& C,
                                          uint32_t pool_k = POOL_K, uint32_t pool_cap = POOL_SIZE) {
  constexpr bool POOL = POOLED && !EXACT;
  constexpr int NK = 1 << D;
  const uint32_t off = T.off, n = T.n;
  const uint32_t s = base + lane;
  const bool live = lane < count;
  const float alpha = T.alpha;
  const uint32_t step_no = T.step_no, seed = T.seed;
  const double thr0 = T.thr0, dmax2 = T.dmax2;
  const float dmin2f = T.dmin2f, dmax2f = T.dmax2f;
  const float4 me4 = live ? leaves[s] : make_float4(0.f, 0.f, 0.f, 0.f);
  const uint32_t self_node = live ? order[s] - off : 0u;
  L.body[lane] = make_float4(me4.x, me4.y, me4.z, __uint_as_float(self_node));
  // EXACT: each point's force accumulates in L.acc, one term per opening (its children's
  // forces summed in a fixed tree), so the sum depends neither on the other points of the
  // tile nor on which mode ran an opening.  Otherwise the one-lane mode accumulates in
  // registers (ax, ay, az, fused multiply-adds) and the packed mode in L.acc.
  float ax = 0.f, ay = 0.f, az = 0.f;
  uint32_t inter = 0;

  uint32_t sp = 0, max_sp = 0;
  uint64_t opened = 0, lane_visits = 0;
  uint32_t err = 0;
  {
    const NodeRec& r = rec[off];
    const NodeHdr h = *(const NodeHdr*)&r;
    bool done = live ? cell_force<D, EXACT ? ACC_SET : ACC_FUSED>(me4.x, me4.y, me4.z, self_node, n, 0u, 0u, thr0, h,
                                     dmax2, dmin2f, alpha, step_no, seed, ax, ay, az, inter)
                     : true;
    if (EXACT) {
      L.acc[0][lane] = ax;
      L.acc[1][lane] = ay;
      L.acc[2][lane] = az;
      ax = ay = az = 0.f;
    } else {
      L.acc[0][lane] = 0.f;
      L.acc[1][lane] = 0.f;
      L.acc[2][lane] = 0.f;
    }
    uint64_t open = __ballot(!done);
    if (open) {
      if (lane == 0) {
        L.depth[0] = r.depth;
        L.mask[0] = open;
      }
      if (lane < 8) L.kids[0][lane] = r.kids[lane];
      sp = 1;
    }
  }
  const uint32_t guard_max = 2u * n + 8u;
  // lane t stages dwords 2(t & 7), +1 of child t >> 3
  const uint32_t my_c = lane >> 3, my_w = lane & 7;

  // The pool of single-point openings (POOL only): pn entries, the last pushed opened first.
  // A step takes up to TAKE entries; lane (e, c) decides child c of entry e's node for that
  // entry's point, the terms of an entry summed over its NK lanes and added to the point's
  // sum in LDS; the children to open go back into the pool, those that find it full onto the
  // stack as one-point entries.
  constexpr uint32_t TAKE = (uint32_t)WAVE / NK;
  uint32_t pn = 0;
  // every entry is a (point, node) pair met once, so a tile takes at most 64 n entries
  const uint64_t pool_guard_max = 64ull * n + 8ull;
  uint64_t pool_guard = 0;
  auto pool_step = [&]() {
    if constexpr (POOL) {
      if (++pool_guard > pool_guard_max) {
        err = 1;
        pn = 0;
        return;
      }
      const uint32_t take = pn < TAKE ? pn : TAKE;
      const uint32_t pe = lane / NK, pc = lane % NK;
      const bool act = pe < take;
      const uint2 ent = L.pool[act ? pn - 1u - pe : 0u];
      pn -= take;
      ++opened;
      lane_visits += take * NK;
      __builtin_amdgcn_wave_barrier();
      // Straight-line code: every lane loads a node header and a point (the arrays' first
      // node and the tile's first point where it has none, so every address is valid; the
      // same for every tree of a batch) and d3's rules
      // (barnesHut.js:705-746, manyBody.js:73-113) become selects; the two rare cases --
      // a chain of more than one level that is accepted, a source at the point's own
      // position -- run only when some lane meets them.
      const uint32_t p = act ? ent.y >> 8 : 0u, ptop = ent.y & 255u;
      uint32_t kid = REF_NONE;
      // (the child slot passed through an empty asm: hoisted, the address of slot pc of the
      // first record is a 64-bit loop invariant that the register budget spills)
      uint32_t slot = pc;
      asm volatile("" : "+v"(slot));
      if (act) kid = rec[ent.x].kids[slot];
      // (bitwise & and | on the decisions: no short-circuit branches)
      const bool is_pt = (kid != REF_NONE) & ((kid & REF_POINT) != 0u);
      const bool is_node = (kid != REF_NONE) & ((kid & REF_POINT) == 0u);
      const uint32_t j = is_pt ? kid & ~REF_POINT : base;
      const NodeHdr h = *(const NodeHdr*)&rec[is_node ? kid : 0u];
      const float4 q = leaves[j];
      const float4 bd = L.body[p];
      const double dx = (is_node ? h.cx : (double)q.x) - (double)bd.x;
      const double dy = D > 1 ? (is_node ? h.cy : (double)q.y) - (double)bd.y : 0.0;
      const double dz = D > 2 ? (is_node ? h.cz : (double)q.z) - (double)bd.z : 0.0;
      const double l = dx * dx + dy * dy + dz * dz;
      const int chain = (int)h.depth - (int)ptop;
      // a chain's thresholds shrink by exactly 4 per level: its deepest level decides; a node
      // above its parent's depth (inside a run of equal codes) always opens
      const bool accept = (chain >= 0) & (open_below(thr0, h.depth) < l);
      const bool strong = h.value != 0.0f;
      const bool open = is_node & strong & !accept;
      const bool use = ((is_node & strong & accept) | (is_pt & (j != base + p))) & (l < dmax2);
      float vd = is_node ? h.value : q.w;
      const bool longer = use & is_node & (chain > 0);
      if (__builtin_amdgcn_ballot_w64(longer)) {
        if (longer) {  // the first level of the chain that accepts
          uint32_t Dd = ptop;
          while (!(open_below(thr0, Dd) < l)) ++Dd;
          vd = h.value * dampf<D>(h.depth - Dd);
        }
      }
      float ex = (float)dx, ey = (float)dy, ez = (float)dz, fl = (float)l;
      const bool same = use & (l == 0.0);
      if (__builtin_amdgcn_ballot_w64(same)) {
        if (same) {  // the hash jiggle of a coincident pair
          const float4 jj = coincident_far(__float_as_uint(bd.w), is_node ? n + (kid - off) : order[j] - off, D,
                                           step_no, seed);
          ex = jj.x, ey = jj.y, ez = jj.z, fl = jj.w;
        }
      }
      const float f = use ? vd * alpha * inv_dist2(fl, dmin2f) : 0.f;
      float fx = ex * f, fy = D > 1 ? ey * f : 0.f, fz = D > 2 ? ez * f : 0.f;
      inter += use ? 1u : 0u;
      const uint32_t cdepth = h.depth;
      fx = group_sum<NK>(fx);
      if (D > 1) fy = group_sum<NK>(fy);
      if (D > 2) fz = group_sum<NK>(fz);
      if (act && pc == 0u) {
        __hip_atomic_fetch_add(&L.acc[0][p], fx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (D > 1) __hip_atomic_fetch_add(&L.acc[1][p], fy, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (D > 2) __hip_atomic_fetch_add(&L.acc[2][p], fz, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      }
      const uint64_t ob = __builtin_amdgcn_ballot_w64(open);
      if (ob) {
        const uint32_t room = pool_cap > pn ? pool_cap - pn : 0u, rank = mask_rank(ob);
        if (open && rank < room) L.pool[pn + rank] = make_uint2(kid, p << 8 | (cdepth + 1u));
        const uint32_t cnt = (uint32_t)__popcll(ob);
        uint64_t over = __builtin_amdgcn_ballot_w64(open && rank >= room);
        pn += cnt < room ? cnt : room;
        while (over) {  // the pool is full: one-point stack entries
          const uint32_t src = (uint32_t)__builtin_ctzll(over);
          over &= over - 1ull;
          if (sp >= (uint32_t)WALK_DEPTH) {
            err = 1;
            continue;
          }
          const uint32_t okid = (uint32_t)__builtin_amdgcn_readlane((int)kid, (int)src);
          const uint32_t op = (uint32_t)__builtin_amdgcn_readlane((int)p, (int)src);
          const uint32_t od = (uint32_t)__builtin_amdgcn_readlane((int)cdepth, (int)src);
          const uint32_t k8 = lane < 8u ? rec[okid].kids[lane] : 0u;
          if (sp < (uint32_t)STACK) {
            if (lane == 0) {
              L.depth[sp] = od;
              L.mask[sp] = (uint64_t)1 << op;
            }
            if (lane < 8u) L.kids[sp][lane] = k8;
          } else {
            SpillEntry* E = spill + (sp - STACK);
            if (lane == 0) {
              store_coherent(&E->mask, (uint64_t)1 << op);
              store_coherent(&E->depth, od);
            }
            if (lane < 8u) store_coherent(&E->kids[lane], k8);
          }
          ++sp;
          max_sp = sp > max_sp ? sp : max_sp;
        }
      }
      __builtin_amdgcn_wave_barrier();
    }
  };

  for (uint32_t guard = 0; sp > 0 || (POOL && pn > 0);) {
    // a pool step as soon as the pool fills one, and to drain it once the stack is done
    if (POOL && (sp == 0 || pn >= TAKE)) {
      pool_step();
      continue;
    }
    if (++guard > guard_max) {
      err = 1;
      break;
    }
    --sp;
    uint32_t e = sp;  // the LDS entry to open
    if (sp >= (uint32_t)STACK) {
      // from the spill area: this wave's own earlier stores, completed before the loads
      const SpillEntry* E = spill + (sp - STACK);
      wait_all_memory();
      if (lane == 0) {
        L.mask[STACK] = load_coherent(&E->mask);
        L.depth[STACK] = load_coherent(&E->depth);
      }
      if (lane < 8) L.kids[STACK][lane] = load_coherent(&E->kids[lane]);
      __builtin_amdgcn_wave_barrier();
      e = STACK;
    }
    uint64_t m = L.mask[e];
    m = ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(m >> 32)) << 32) |
        (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)m);
    const bool in = (m >> lane) & 1ull;
    const uint32_t k_active = (uint32_t)__popcll(m);
    ++opened;
    lane_visits += k_active;
    // the popped node's children, and one coalesced load of their records into staging
    uint32_t my_kid = REF_NONE;
    if (lane < (uint32_t)NK) my_kid = L.kids[e][lane];
    {
      const uint32_t k = my_c < (uint32_t)NK ? L.kids[e][my_c] : REF_NONE;
      // one load for point and node children alike (two would wait for each other)
      const bool point = (k & REF_POINT) != 0;
      const uint2* src = point ? (const uint2*)(leaves + (k & ~REF_POINT)) : (const uint2*)(rec + k);
      uint2 v = make_uint2(0u, 0u);
      if (k != REF_NONE && (!point || my_w < 2)) v = src[my_w];
      if (my_c < (uint32_t)NK) {
        L.stage[my_c][2 * my_w] = v.x;
        L.stage[my_c][2 * my_w + 1] = v.y;
      }
    }
    if (lane < (uint32_t)NK) L.cur[lane] = my_kid;
    const uint32_t nk = (uint32_t)__popcll(__ballot(lane < (uint32_t)NK && my_kid != REF_NONE));
    // the children's cell chains start one level below the popped node's cell
    const uint32_t top = (uint32_t)__builtin_amdgcn_readfirstlane(L.depth[e]) + 1;

    uint32_t open_bits = 0;  // bit c: this lane's point must open child c
    if (k_active <= packed_max) {
      // packed: item t of round r is (point list[t / NK], child t % NK), NK items per point
      if (in) L.list[mask_rank(m)] = lane;
      const uint32_t rounds = (k_active * NK + WAVE - 1) / WAVE;
#pragma unroll 1
      for (uint32_t r = 0; r < rounds; ++r) {
        const uint32_t t = r * WAVE + lane;
        const uint32_t i = t / NK, c = t % NK;
        bool open_item = false;
        float fx = 0.f, fy = 0.f, fz = 0.f;
        if (i < k_active && c < nk) {
          const uint32_t b = L.list[i];
          const float4 bd = L.body[b];
          const uint32_t bnode = __float_as_uint(bd.w);
          const uint32_t k = L.cur[c];
          if (k & REF_POINT) {
            const uint32_t j = k & ~REF_POINT;
            if (j != base + b)
              point_force<D, EXACT ? ACC_SET : ACC_FUSED>(bd.x, bd.y, bd.z, bnode, *(const float4*)&L.stage[c][0], order, j, off,
                             dmax2f, dmin2f, alpha, step_no, seed, fx, fy, fz, inter);
          } else {
            const NodeHdr h = *(const NodeHdr*)&L.stage[c][0];
            open_item = !cell_force<D, EXACT ? ACC_SET : ACC_FUSED>(bd.x, bd.y, bd.z, bnode, n, k - off, top, thr0, h,
                                       dmax2, dmin2f, alpha, step_no, seed, fx, fy, fz, inter);
          }
        }
        // sum the NK items of each point (aligned groups of NK lanes; lane c == 0 gets it)
        fx = group_sum<NK>(fx);
        if (D > 1) fy = group_sum<NK>(fy);
        if (D > 2) fz = group_sum<NK>(fz);
        if (c == 0 && i < k_active) {
          const uint32_t b = L.list[i];
          L.acc[0][b] += fx;
          if (D > 1) L.acc[1][b] += fy;
          if (D > 2) L.acc[2][b] += fz;
        }
        const uint64_t bal = __ballot(open_item);
        if (lane == 0) L.ballots[r] = bal;
      }
      if (in) {
        const uint32_t i = mask_rank(m);
        const uint64_t bal = L.ballots[(i * NK) / WAVE];
        open_bits = (uint32_t)(bal >> ((i * NK) % WAVE)) & ((1u << NK) - 1u);
      }
    } else if (!EXACT) {
      // one lane per point, children in turn; the child, its depth and its strength are the
      // same on every lane, so the tests on them are scalar branches
#pragma unroll 1
      for (uint32_t c = 0; c < nk; ++c) {
        const uint32_t k = (uint32_t)__builtin_amdgcn_readlane((int)my_kid, (int)c);
        if (!in) continue;
        if (k & REF_POINT) {
          const uint32_t j = k & ~REF_POINT;
          if (j != s)
            point_force<D, ACC_FUSED>(me4.x, me4.y, me4.z, self_node, *(const float4*)&L.stage[c][0], order, j,
                                  off, dmax2f, dmin2f, alpha, step_no, seed, ax, ay, az, inter);
          continue;
        }
        NodeHdr h = *(const NodeHdr*)&L.stage[c][0];
        h.depth = (uint32_t)__builtin_amdgcn_readfirstlane((int)h.depth);
        h.value = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(h.value)));
        if (!cell_force<D, ACC_FUSED>(me4.x, me4.y, me4.z, self_node, n, k - off, top, thr0, h,
                                  dmax2, dmin2f, alpha, step_no, seed, ax, ay, az, inter))
          open_bits |= 1u << c;
      }
    } else {
      // one lane per point, children in turn -- by pairs, in the order in which group_sum
      // adds them: child c with NK - 1 - c (NK = 8) or c + 2 (NK = 4); the pairs of an
      // 8-child node as (0, 2), then (1, 3)
      float qx = 0.f, qy = 0.f, qz = 0.f, hx = 0.f, hy = 0.f, hz = 0.f;
      float sx = 0.f, sy = 0.f, sz = 0.f;
      constexpr uint32_t PAIRS = NK / 2;
#pragma unroll 1
      for (uint32_t t = 0; t < PAIRS; ++t) {
        const uint32_t ca = NK == 8 ? ((t & 1u) << 1 | t >> 1) : t;
        const uint32_t cb = NK == 8 ? 7u - ca : ca + PAIRS;
        // the pair's sum: the first child's force, plus the second's
        float px = 0.f, py = 0.f, pz = 0.f;
        auto child = [&](uint32_t c, auto acc) {
          constexpr int ACC = decltype(acc)::value;
          if (c >= nk || !in) return;
          const uint32_t k = L.cur[c];
          if (k & REF_POINT) {
            const uint32_t j = k & ~REF_POINT;
            if (j != s)
              point_force<D, ACC>(me4.x, me4.y, me4.z, self_node, *(const float4*)&L.stage[c][0], order, j,
                                  off, dmax2f, dmin2f, alpha, step_no, seed, px, py, pz, inter);
          } else {
            const NodeHdr hd = *(const NodeHdr*)&L.stage[c][0];
            if (!cell_force<D, ACC>(me4.x, me4.y, me4.z, self_node, n, k - off, top, thr0, hd,
                                    dmax2, dmin2f, alpha, step_no, seed, px, py, pz, inter))
              open_bits |= 1u << c;
          }
        };
        child(ca, std::integral_constant<int, ACC_SET>{});
        child(cb, std::integral_constant<int, ACC_ADD>{});
        if (PAIRS == 1 || ((t & 1u) == 0u && PAIRS == 4)) {
          qx = px, qy = py, qz = pz;  // NK = 2: the sum; NK = 8: the first pair of a half
        } else if (PAIRS == 2) {
          if (t == 0) qx = px, qy = py, qz = pz;
          else qx = qx + px, qy = qy + py, qz = qz + pz;
        } else if (t == 1) {
          hx = qx + px, hy = qy + py, hz = qz + pz;
        } else {
          qx = hx + (qx + px), qy = hy + (qy + py), qz = hz + (qz + pz);
        }
      }
      sx = qx, sy = qy, sz = qz;
      if (in) {
        L.acc[0][lane] += sx;
        if (D > 1) L.acc[1][lane] += sy;
        if (D > 2) L.acc[2][lane] += sz;
      }
    }
    if (EXACT) {
      // push the children some lane must open (open_bits is 0 on lanes not in the mask), in
      // order: the k-th opened child goes to entry sp + k (none beyond WALK_DEPTH).  Lane 0
      // writes each entry's lane mask as the ballots come; then lane (c, w) moves word w of
      // child c's children, and lane (c, 0) its depth, all children at once.
      uint32_t pushed = 0, ranks = 0;  // bit c: child c is pushed; the count so far
#pragma unroll
      for (uint32_t c = 0; c < (uint32_t)NK; ++c) {
        const uint64_t open = __builtin_amdgcn_ballot_w64(((open_bits >> c) & 1u) != 0u);
        if (!open) continue;
        const uint32_t slot = sp + ranks;
        ++ranks;
        if (slot >= (uint32_t)WALK_DEPTH) {
          err = 1;
          continue;
        }
        pushed |= 1u << c;
        if (lane == 0) {
          if (slot < (uint32_t)STACK) L.mask[slot] = open;
          else store_coherent(&spill[slot - STACK].mask, open);
        }
      }
      if ((pushed >> my_c) & 1u) {
        const uint32_t slot = sp + (uint32_t)__popc(pushed & ((1u << my_c) - 1u));
        const uint32_t word = L.stage[my_c][8 + my_w];
        const uint32_t cdepth = ((const NodeHdr*)&L.stage[my_c][0])->depth;
        if (slot < (uint32_t)STACK) {
          L.kids[slot][my_w] = word;
          if (my_w == 0) L.depth[slot] = cdepth;
        } else {
          SpillEntry* E = spill + (slot - STACK);
          store_coherent(&E->kids[my_w], word);
          if (my_w == 0) store_coherent(&E->depth, cdepth);
        }
      }
      sp += (uint32_t)__popc(pushed);
    } else {
      // one child at a time (measured cheaper than the above where the walk is bound by
      // issue, 1M points)
#pragma unroll 1
      for (uint32_t c = 0; c < nk; ++c) {
        const uint64_t open = __builtin_amdgcn_ballot_w64(((open_bits >> c) & 1u) != 0u);
        if (!open) continue;
        const uint32_t cdepth = ((const NodeHdr*)&L.stage[c][0])->depth;
        if constexpr (POOL) {
          const uint32_t cnt = (uint32_t)__popcll(open);
          if (cnt <= pool_k && pn + cnt <= pool_cap) {  // one pool entry per point
            const uint32_t k = (uint32_t)__builtin_amdgcn_readlane((int)my_kid, (int)c);
            if ((open >> lane) & 1ull) L.pool[pn + mask_rank(open)] = make_uint2(k, lane << 8 | (cdepth + 1u));
            pn += cnt;
            continue;
          }
        }
        if (sp >= (uint32_t)WALK_DEPTH) {
          err = 1;
          continue;
        }
        if (sp < (uint32_t)STACK) {
          if (lane == 0) {
            L.depth[sp] = cdepth;
            L.mask[sp] = open;
          }
          if (lane < 8) L.kids[sp][lane] = L.stage[c][8 + lane];
        } else {
          SpillEntry* E = spill + (sp - STACK);
          if (lane == 0) {
            store_coherent(&E->mask, open);
            store_coherent(&E->depth, cdepth);
          }
          if (lane < 8) store_coherent(&E->kids[lane], L.stage[c][8 + lane]);
        }
        ++sp;
      }
      if constexpr (POOL) __builtin_amdgcn_wave_barrier();
    }
    max_sp = sp > max_sp ? sp : max_sp;
  }
  if (live) {
    const uint32_t o = self_node;
    if (EXACT) {
      T.vx[o] += L.acc[0][lane];
      if (D > 1) T.vy[o] += L.acc[1][lane];
      if (D > 2) T.vz[o] += L.acc[2][lane];
    } else {
      T.vx[o] += ax + L.acc[0][lane];
      if (D > 1) T.vy[o] += ay + L.acc[1][lane];
      if (D > 2) T.vz[o] += az + L.acc[2][lane];
    }
  }
  C.opened += opened;
  C.lane_visits += lane_visits;
  C.inter += inter;
  C.err |= err;
  C.max_sp = max_sp > C.max_sp ? max_sp : C.max_sp;
}

// The workgroup's counters: its waves add theirs in LDS, then one atomic per counter (the
// waves of a walk finish close together, and atomics on shared words serialise).  Every
// wave of the workgroup must call it.
__device__ inline void walk_flush(const WalkCount& C, uint32_t lane, uint32_t* counters) {
  __shared__ unsigned long long s_sum[3];
  __shared__ uint32_t s_err, s_max;
  if (threadIdx.x == 0) s_sum[0] = s_sum[1] = s_sum[2] = 0, s_err = 0, s_max = 0;
  __syncthreads();
  uint32_t tot = C.inter;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) tot += __shfl_xor(tot, o, WAVE);
  if (lane == 0) {
    atomicAdd(&s_sum[0], (unsigned long long)tot);
    atomicAdd(&s_sum[1], (unsigned long long)C.opened);
    atomicAdd(&s_sum[2], (unsigned long long)C.lane_visits);
    if (C.err) atomicAdd(&s_err, 1u);
    atomicMax(&s_max, C.max_sp);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    atomicAdd((unsigned long long*)&counters[0], s_sum[0]);
    atomicAdd((unsigned long long*)&counters[2], s_sum[1]);
    atomicAdd((unsigned long long*)&counters[6], s_sum[2]);
    if (s_err) atomicAdd(&counters[4], s_err);
    atomicMax(&counters[8], s_max);
  }
}


struct WalkArgs {
  double theta2, dmin2, dmax2;
  uint32_t n;
  uint32_t packed_max;  // openings by at most this many points use the packed mode
  uint32_t tile;        // points per tile (1..64)
  uint32_t pool_k;      // openings by at most this many points go to the pool (cheaper mode)
  uint32_t pool_cap;    // pool entries per wave (at most POOL_SIZE)
};

// Points per tile of an exact walk over n points, on a device of `simds` SIMDs where the
// walk keeps `waves` waves resident.  A small walk is bound by the latency of each wave's
// chain of openings: a smaller tile has a shorter chain but adds waves and total work.
// Measured on 256 CUs (link + manyBody + center ticks, 1k..200k points, tiles of 1..64):
// the tiles of 1, 2 and 4 points pay up to about 2.5 tiles per SIMD, 8- and 32-point ones
// up to about 5, 16-point ones as long as every tile has its own resident wave (no
// queue).  `want` (1..64) overrides.
inline uint32_t walk_tile_for(uint32_t n, uint32_t simds, uint32_t waves, uint32_t want) {
  if (want >= 1u && want <= (uint32_t)WAVE) return want;
  for (uint32_t t = 1; t < (uint32_t)WAVE; t *= 2) {
    const uint64_t tiles = (n + t - 1) / t;
    const uint64_t most = t <= 4u ? (uint64_t)simds * 5 / 2
                          : t == 16u ? (uint64_t)waves : (uint64_t)simds * 5;
    if (tiles <= most && tiles <= waves) return t;
  }
  return (uint32_t)WAVE;
}

// One simulation: every wave takes tiles [tile t, tile (t + 1)) of the sorted points from a
// queue (counters[5]) until none is left; `spill` holds WALK_DEPTH - STACK entries per
// resident wave.  Tiles other than 64 points need EXACT.
template <int D, int STACK, int WAVES, bool EXACT>
__global__ void __launch_bounds__(WAVES * WAVE, EXACT ? 1 : POOL_WAVES)
    tree_walk(const float4* __restrict__ leaves, const uint32_t* __restrict__ order,
              const NodeRec* __restrict__ rec, const CubeInfo* __restrict__ cube,
              float* __restrict__ vx, float* __restrict__ vy, float* __restrict__ vz,
              uint32_t* __restrict__ counters, SpillEntry* __restrict__ spill, StepParams step,
              WalkArgs A) {
  __shared__ WalkLDS<STACK, !EXACT> lds[WAVES];
  const uint32_t lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x / WAVE;
  SpillEntry* my_spill = spill + (size_t)(blockIdx.x * WAVES + wave) * (WALK_DEPTH - STACK);
  WalkTree T;
  T.off = 0;
  T.n = A.n;
  T.size = (double)cube->size;
  T.thr0 = T.size * T.size / A.theta2;
  T.dmax2 = A.dmax2;
  T.dmin2f = (float)A.dmin2;
  T.dmax2f = (float)fmin(A.dmax2, 3.0e38);
  T.alpha = step.alpha;
  T.step_no = step.step;
  T.seed = step.seed;
  T.vx = vx;
  T.vy = vy;
  T.vz = vz;
  const uint32_t tiles = (A.n + A.tile - 1) / A.tile;
  WalkCount C;
  // with a wave for every tile, wave w walks tile w and no queue is touched (atomics on one
  // word serialise across the device: with thousands of waves, taking tiles from a queue
  // cost more than the slowest tile of a 200k-point walk); otherwise the waves that finish
  // first take more tiles from the queue.  Which wave walks which tile changes no result.
  const bool statics = gridDim.x * WAVES >= tiles;
  for (uint32_t round = 0;; ++round) {
    uint32_t t = blockIdx.x * WAVES + wave;
    if (statics) {
      if (round > 0) break;
    } else {
      if (lane == 0) t = atomicAdd(&counters[5], 1u);
      t = (uint32_t)__builtin_amdgcn_readfirstlane(t);
    }
    if (t >= tiles) break;
    const uint32_t base = t * A.tile;
    const uint32_t count = A.n - base < A.tile ? A.n - base : A.tile;
    __builtin_amdgcn_wave_barrier();
    walk_tile<D, STACK, EXACT, !EXACT>(lds[wave], leaves, order, rec, T, base, count, A.packed_max, lane, my_spill,
                               C, A.pool_k, A.pool_cap);
  }
  walk_flush(C, lane, counters);
}

// ---- walk of single points ---------------------------------------------------------
//
// A small tree's walk is a chain of openings per wave, each waiting on its loads and LDS
// round trips with no other wave to hide behind.  Here every point has a group of NK = 2^D
// lanes of its own (64 / NK points per wave, no union of traversals): lane c of a group
// takes child c of each node the point opens, straight from its record into registers, and
// the point's stack (children and depth of each entry to open) lives in LDS.  The terms
// and their order are those of the exact-mode tiled walk: the root's term, then one term
// per opening (the children's forces summed by group_sum), openings in the order of the
// stack, which takes a node's opened children in order and opens the last first.  So the
// forces are bit for bit those of any tiling.
template <int D>
struct SoloLDS {
  static constexpr uint32_t NK = 1u << D, G = WAVE / NK;
  alignas(16) uint32_t kids[G][WALK_DEPTH][NK];  // an entry's children go in as 16-byte stores
  uint32_t depth[G][WALK_DEPTH];
};

// WAVES (1 or 2) waves per workgroup, their stacks 55-74 KB of LDS each: one wave per
// workgroup while every wave can have a CU of its own (2 shared a CU's cache and LDS:
// 1,000 points 0.093 -> 0.096 ms/tick), two beyond that (half the workgroups and their
// counter atomics: 5,000 points 0.181 -> 0.173, 8,000 0.207 -> 0.195).
template <int D, uint32_t WAVES>
__global__ void __launch_bounds__(WAVES * WAVE) tree_walk_solo(const float4* __restrict__ leaves,
                                                       const uint32_t* __restrict__ order,
                                                       const NodeRec* __restrict__ rec,
                                                       const CubeInfo* __restrict__ cube, float* __restrict__ vx,
                                                       float* __restrict__ vy, float* __restrict__ vz,
                                                       uint32_t* __restrict__ counters, StepParams step, WalkArgs A) {
  constexpr uint32_t NK = 1u << D, G = WAVE / NK;
  __shared__ SoloLDS<D> lds[WAVES];
  const uint32_t lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x / WAVE, g = lane / NK, c = lane % NK;
  SoloLDS<D>& L = lds[wave];
  const uint32_t n = A.n, s = (blockIdx.x * WAVES + wave) * G + g;
  const bool live = s < n;
  const double size = (double)cube->size;
  const double thr0 = size * size / A.theta2, dmax2 = A.dmax2;
  const float dmin2f = (float)A.dmin2, dmax2f = (float)fmin(A.dmax2, 3.0e38);
  const float alpha = step.alpha;
  const uint32_t step_no = step.step, seed = step.seed;
  const float4 me = live ? leaves[s] : make_float4(0.f, 0.f, 0.f, 0.f);
  const uint32_t self_node = live ? order[s] : 0u;
  uint32_t inter = 0, sp = 0;
  float ax = 0.f, ay = 0.f, az = 0.f;
  {  // the root: one term (every lane of the group the same; lane 0 counts it)
    const NodeRec& r = rec[0];
    const NodeHdr h = *(const NodeHdr*)&r;
    const bool done = live ? cell_force<D, ACC_SET>(me.x, me.y, me.z, self_node, n, 0u, 0u, thr0, h,
                                                      dmax2, dmin2f, alpha, step_no, seed, ax, ay, az, inter)
                           : true;
    if (c != 0) inter = 0;
    if (!done) {
      if (c == 0) L.depth[g][0] = r.depth;
      L.kids[g][0][c] = r.kids[c];
      sp = 1;
    }
  }
  uint32_t max_sp = sp, opened = 0, guard = 0, err = 0;
  const uint32_t guard_max = 2u * n + 8u;
  while (__builtin_amdgcn_ballot_w64(sp > 0u)) {
    if (sp == 0u) continue;  // the group's walk is over (every lane of a group agrees)
    if (++guard > guard_max) {
      err = 1;
      sp = 0;
      continue;
    }
    --sp;
    const uint32_t k = L.kids[g][sp][c];
    const uint32_t top = L.depth[g][sp] + 1;
    opened += c == 0 ? 1u : 0u;
    float fx = 0.f, fy = 0.f, fz = 0.f;
    bool open = false;
    uint4 k0 = make_uint4(0u, 0u, 0u, 0u), k1 = k0;
    uint32_t cdepth = 0;
    if (k != REF_NONE) {
      if (k & REF_POINT) {
        const uint32_t j = k & ~REF_POINT;
        const float4 q = leaves[j];
        if (j != s)
          point_force<D, ACC_SET>(me.x, me.y, me.z, self_node, q, order, j, 0u, dmax2f, dmin2f, alpha, step_no, seed,
                                  fx, fy, fz, inter);
      } else {
        const uint4* R = (const uint4*)(rec + k);
        const uint4 q0 = R[0], q1 = R[1];
        k0 = R[2];
        k1 = R[3];
        NodeHdr h;
        __builtin_memcpy(&h, &q0, 16);
        __builtin_memcpy((char*)&h + 16, &q1, 16);
        open = !cell_force<D, ACC_SET>(me.x, me.y, me.z, self_node, n, k, top, thr0, h, dmax2, dmin2f,
                                       alpha, step_no, seed, fx, fy, fz, inter);
        cdepth = h.depth;
      }
    }
    fx = group_sum<NK>(fx);
    if (D > 1) fy = group_sum<NK>(fy);
    if (D > 2) fz = group_sum<NK>(fz);
    ax += fx;
    if (D > 1) ay += fy;
    if (D > 2) az += fz;
    // the opened children in order, from the entry just popped upwards
    const uint32_t bits = (uint32_t)(__builtin_amdgcn_ballot_w64(open) >> (g * NK)) & ((1u << NK) - 1u);
    if (open) {
      const uint32_t slot = sp + (uint32_t)__popc(bits & ((1u << c) - 1u));
      if (slot < (uint32_t)WALK_DEPTH) {
        L.depth[g][slot] = cdepth;
        uint32_t* e = L.kids[g][slot];
        if (NK == 8) {
          *(uint4*)e = k0;
          *(uint4*)(e + 4) = k1;
        } else if (NK == 4) {
          *(uint4*)e = k0;
        } else {
          *(uint2*)e = make_uint2(k0.x, k0.y);
        }
      } else {
        err = 1;
      }
    }
    sp += (uint32_t)__popc(bits);
    if (sp > (uint32_t)WALK_DEPTH) sp = WALK_DEPTH;
    max_sp = sp > max_sp ? sp : max_sp;
  }
  if (live && c == 0) {
    vx[self_node] += ax;
    if (D > 1) vy[self_node] += ay;
    if (D > 2) vz[self_node] += az;
  }
  // the wave's counts into lane 0 for the workgroup's flush
  WalkCount C;
  uint32_t op = opened, e = err, m = max_sp;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    op += __shfl_xor(op, o, WAVE);
    e |= __shfl_xor(e, o, WAVE);
    const uint32_t mo = __shfl_xor(m, o, WAVE);
    m = mo > m ? mo : m;
  }
  C.opened = op;
  C.lane_visits = op;
  C.inter = inter;
  C.err = e;
  C.max_sp = m;
  walk_flush(C, lane, counters);
}

// ---- walk of one point per wave, several openings per step ------------------------
//
// Below a few hundred thousand points the walk is bound by the chain of openings each wave
// goes through, one dependent record load after another, with too few waves on the device
// to hide it.  Here a wave walks one point at a time and opens up to P = 64 / 2^D stack
// entries per step: lane (e, c) takes child c of the e-th entry from the top, loads its
// record straight into registers, decides it with d3's rules and adds its term to the
// lane's own sums; the children it must open replace the opened entries on the stack.  The
// chain is about P times shorter and every wave of the device has points to walk.  The
// point's force is the sum of the lanes' sums, lane 0's fixed tree of adds (deterministic
// for a given tree; not the order of the tiled walk: the two differ by rounding only).
//
// The stack: WIDE_DEPTH entries, the first WIDE_LDS in LDS and the rest in the wave's spill
// area in global memory.  Entries open P at a time while the stack stays within WIDE_BROAD,
// one at a time beyond (d3's depth-first order), so the stack exceeds WIDE_BROAD by at most
// what a depth-first walk of one entry adds: 7 * 21 + 8 in 3D for distinct codes, as in the
// other walks; deeper (long runs of coincident points) counts as a walk error.  Walks at
// theta 0.9 peak near 20 entries.
//
// The grid is persistent (each wave takes points wave, wave + waves, ...), and each
// workgroup writes its counts to a slot of its own (plain stores; the host adds the slots):
// tens of thousands of workgroups adding to the same counter words serialise on them.
constexpr uint32_t WIDE_DEPTH = 384;
constexpr uint32_t WIDE_LDS = 64;
constexpr uint32_t WIDE_BROAD = WIDE_DEPTH - (uint32_t)WALK_DEPTH;
constexpr uint32_t WIDE_WAVES = 4;

template <int D>
struct WideLDS {
  static constexpr uint32_t NK = 1u << D;
  alignas(16) uint32_t kids[WIDE_LDS][NK];  // an entry's children go in as 16-byte stores
  uint32_t depth[WIDE_LDS];
};

// Words of spill area per wave of the wide walk: the children and depth of each entry past
// the LDS part.
template <int D>
constexpr uint32_t wide_spill_words() {
  return (WIDE_DEPTH - WIDE_LDS) * ((1u << D) + 1u);
}

// Counts of a workgroup of a walk into its slot (every wave of the workgroup calls it).
__device__ inline void walk_flush_slot(const WalkCount& C, uint32_t lane, uint32_t* slot) {
  __shared__ unsigned long long s_sum[2];
  __shared__ uint32_t s_err, s_max;
  if (threadIdx.x == 0) s_sum[0] = s_sum[1] = 0, s_err = 0, s_max = 0;
  __syncthreads();
  uint32_t tot = C.inter;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) tot += __shfl_xor(tot, o, WAVE);
  if (lane == 0) {
    atomicAdd(&s_sum[0], (unsigned long long)tot);
    atomicAdd(&s_sum[1], (unsigned long long)C.opened);
    if (C.err) atomicAdd(&s_err, 1u);
    atomicMax(&s_max, C.max_sp);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    *(uint4*)slot = make_uint4((uint32_t)s_sum[0], (uint32_t)(s_sum[0] >> 32), (uint32_t)s_sum[1],
                               (uint32_t)(s_sum[1] >> 32));
    *(uint4*)(slot + 4) = make_uint4(s_err, s_max, 0u, 0u);
  }
}

// The walk of sorted point s (tree-space index) of tree T by one wave of the wide walk.
// inter (per lane), opened, max_sp and err accumulate the wave's counts; err set: the wave
// stops.
template <int D>
__device__ __forceinline__ void wide_point(WideLDS<D>& L, uint32_t* sk, uint32_t* sd,
                                           const float4* __restrict__ leaves,
                                           const uint32_t* __restrict__ order,
                                           const NodeRec* __restrict__ rec, const WalkTree& T, uint32_t s,
                                           uint32_t lane, uint32_t& inter, uint64_t& opened, uint32_t& max_sp,
                                           uint32_t& err) {
  constexpr uint32_t NK = 1u << D, P = WAVE / NK;
  const uint32_t e = lane / NK, c = lane % NK;
  const uint32_t off = T.off, n = T.n;
  const double thr0 = T.thr0, dmax2 = T.dmax2;
  const float dmin2f = T.dmin2f, dmax2f = T.dmax2f, alpha = T.alpha;
  const uint32_t step_no = T.step_no, seed = T.seed;
  const uint32_t guard_max = 2u * n + 8u;
  const float4 me = leaves[s];
  const uint32_t self_node = order[s] - off;
  uint32_t sp = 0;
  float ax = 0.f, ay = 0.f, az = 0.f;
  {  // the root: lane 0's term
    const NodeRec& r = rec[off];
    const NodeHdr h = *(const NodeHdr*)&r;
    float rx = 0.f, ry = 0.f, rz = 0.f;
    uint32_t ri = 0;
    const bool done = cell_force<D, ACC_FUSED>(me.x, me.y, me.z, self_node, n, 0u, 0u, thr0, h,
                                               dmax2, dmin2f, alpha, step_no, seed, rx, ry, rz, ri);
    if (lane == 0) ax = rx, ay = ry, az = rz, inter += ri;
    if (!done) {
      if (lane == 0) L.depth[0] = r.depth;
      if (lane < NK) L.kids[0][lane] = r.kids[lane];
      sp = 1;
    }
    sp = (uint32_t)__builtin_amdgcn_readfirstlane(sp);  // every lane decided the root alike
  }
  for (uint32_t guard = 0; sp > 0u; ++guard) {  // sp is the same on every lane (kept in SGPRs)
    if (guard > guard_max) {
      err = 1;
      break;
    }
    // entries to open: P (or all the stack has) while P entries' children fit within
    // WIDE_BROAD, one beyond
    const uint32_t p = sp + (NK - 1) * P <= WIDE_BROAD ? (sp < P ? sp : P) : 1u;
    const uint32_t base = sp - p;  // entries [base, sp) open now; their children go from base up
    opened += p;
    if (sp > WIDE_LDS) wait_all_memory();  // the wave's own spill stores, complete before these loads
    uint32_t k = REF_NONE, top = 0;
    if (e < p) {
      const uint32_t ent = sp - 1u - e;
      if (ent < WIDE_LDS) {
        k = L.kids[ent][c];
        top = L.depth[ent] + 1u;
      } else {
        k = load_coherent(&sk[(ent - WIDE_LDS) * NK + c]);
        top = load_coherent(&sd[ent - WIDE_LDS]) + 1u;
      }
    }
    bool open = false;
    uint4 k0, k1;     // set, with cdepth, wherever open is
    uint32_t cdepth;
    if (k != REF_NONE) {
      if (k & REF_POINT) {
        const uint32_t j = k & ~REF_POINT;
        const float4 q = leaves[j];
        if (j != s)
          point_force<D, ACC_FUSED>(me.x, me.y, me.z, self_node, q, order, j, off, dmax2f, dmin2f, alpha, step_no,
                                    seed, ax, ay, az, inter);
      } else {
        const uint4* R = (const uint4*)(rec + k);
        const uint4 q0 = R[0], q1 = R[1];
        k0 = R[2];
        k1 = R[3];
        NodeHdr h;
        __builtin_memcpy(&h, &q0, 16);
        __builtin_memcpy((char*)&h + 16, &q1, 16);
        open = !cell_force<D, ACC_FUSED>(me.x, me.y, me.z, self_node, n, k - off, top, thr0, h, dmax2,
                                         dmin2f, alpha, step_no, seed, ax, ay, az, inter);
        cdepth = h.depth;
      }
    }
    // the children to open, in lane order, over the entries just opened
    const uint64_t bits = __builtin_amdgcn_ballot_w64(open);
    const uint32_t nsp = (uint32_t)__builtin_amdgcn_readfirstlane(base + (uint32_t)__popcll(bits));
    if (nsp > WIDE_DEPTH) {  // no entry past the stack is written: stop here
      err = 1;
      break;
    }
    if (open) {
      const uint32_t slot = base + mask_rank(bits);
      const uint32_t kk[8] = {k0.x, k0.y, k0.z, k0.w, k1.x, k1.y, k1.z, k1.w};
      if (slot < WIDE_LDS) {
        L.depth[slot] = cdepth;
        uint32_t* en = L.kids[slot];
        if (NK == 8) {
          *(uint4*)en = k0;
          *(uint4*)(en + 4) = k1;
        } else if (NK == 4) {
          *(uint4*)en = k0;
        } else {
          *(uint2*)en = make_uint2(k0.x, k0.y);
        }
      } else {
#pragma unroll
        for (uint32_t q = 0; q < NK; ++q) store_coherent(&sk[(slot - WIDE_LDS) * NK + q], kk[q]);
        store_coherent(&sd[slot - WIDE_LDS], cdepth);
      }
    }
    sp = nsp;
    max_sp = sp > max_sp ? sp : max_sp;
  }
  // the point's force: the lanes' sums added in lane 0's fixed tree
#pragma unroll
  for (int o = 1; o < WAVE; o <<= 1) {
    ax += __shfl_xor(ax, o, WAVE);
    if (D > 1) ay += __shfl_xor(ay, o, WAVE);
    if (D > 2) az += __shfl_xor(az, o, WAVE);
  }
  if (lane == 0 && !err) {
    T.vx[self_node] += ax;
    if (D > 1) T.vy[self_node] += ay;
    if (D > 2) T.vz[self_node] += az;
  }
}

template <int D, uint32_t WAVES>
__global__ void __launch_bounds__(WAVES * WAVE) tree_walk_wide(const float4* __restrict__ leaves,
                                                       const uint32_t* __restrict__ order,
                                                       const NodeRec* __restrict__ rec,
                                                       const CubeInfo* __restrict__ cube, float* __restrict__ vx,
                                                       float* __restrict__ vy, float* __restrict__ vz,
                                                       uint32_t* __restrict__ slots, uint32_t* __restrict__ spill,
                                                       StepParams step, WalkArgs A) {
  __shared__ WideLDS<D> lds[WAVES];
  const uint32_t lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x / WAVE;
  const uint32_t gw = blockIdx.x * WAVES + wave, waves = gridDim.x * WAVES;
  uint32_t* sk = spill + (size_t)gw * wide_spill_words<D>();  // entry i >= WIDE_LDS: kids, then depths
  uint32_t* sd = sk + (WIDE_DEPTH - WIDE_LDS) * (1u << D);
  WalkTree T;
  T.off = 0;
  T.n = A.n;
  T.size = (double)cube->size;
  T.thr0 = T.size * T.size / A.theta2;
  T.dmax2 = A.dmax2;
  T.dmin2f = (float)A.dmin2;
  T.dmax2f = (float)fmin(A.dmax2, 3.0e38);
  T.alpha = step.alpha;
  T.step_no = step.step;
  T.seed = step.seed;
  T.vx = vx;
  T.vy = vy;
  T.vz = vz;
  uint32_t inter = 0, max_sp = 0, err = 0;
  uint64_t opened = 0;
  for (uint32_t s = gw; s < A.n && !err; s += waves)
    wide_point<D>(lds[wave], sk, sd, leaves, order, rec, T, s, lane, inter, opened, max_sp, err);
  WalkCount C;
  C.opened = opened;
  C.inter = inter;
  C.err = err;
  C.max_sp = max_sp;
  walk_flush_slot(C, lane, slots + (size_t)blockIdx.x * 8);
}

// The work items [first, end) and stride of wave `wave` of a persistent walk grid, XCD by
// XCD: workgroups are dealt to the device's 8 XCDs in turn (blockIdx.x % 8), each with an
// L2 of its own, so the grid's workgroups on one XCD take one contiguous eighth of the
// items (points in Morton order: neighbouring points open the same nodes, whose records
// then stay in that XCD's L2).  Any dealing of workgroups covers every item once.
struct WalkRange {
  uint32_t first, end, stride;
};
__device__ inline WalkRange xcd_range(uint32_t items, uint32_t waves_per_block, uint32_t wave) {
  const uint32_t parts = gridDim.x >= 8u && gridDim.x % 8u == 0u ? 8u : 1u;
  const uint32_t x = blockIdx.x % parts, lb = blockIdx.x / parts, nbx = gridDim.x / parts;
  const uint32_t lo = (uint32_t)((uint64_t)items * x / parts), hi = (uint32_t)((uint64_t)items * (x + 1) / parts);
  return WalkRange{lo + lb * waves_per_block + wave, hi, nbx * waves_per_block};
}

// ---- walk of a group of points per wave ----------------------------------------------
//
// As the wide walk, but a wave walks a group of T consecutive sorted points at once: a
// stack entry is a node to open together with the mask of the group's points that must
// open it, and lane (e, c) loads child c of the e-th entry once and decides it for each
// point of that mask in turn (its terms into that point's sums on the lane).  The record
// loads and the stack's pops and pushes are shared by the group, the lanes stay as full as
// in the wide walk, and a point's force is its lanes' sums added in lane 0's fixed tree.
// The root opens as the single child of an entry above it (depth ~0u: its chain starts at
// depth 0).
template <int D, uint32_t T>
struct GroupLDS {
  static constexpr uint32_t NK = 1u << D;
  alignas(16) uint32_t kids[WIDE_LDS][NK];
  uint32_t depth[WIDE_LDS];
  uint32_t mask[WIDE_LDS];
};

// Words of spill area per wave of the group walk: children, depth and mask of each entry
// past the LDS part.
template <int D>
constexpr uint32_t group_spill_words() {
  return (WIDE_DEPTH - WIDE_LDS) * ((1u << D) + 2u);
}

template <int D, uint32_t T>
__device__ __forceinline__ void group_walk(GroupLDS<D, T>& L, uint32_t* sk, const float4* __restrict__ leaves,
                                           const uint32_t* __restrict__ order,
                                           const NodeRec* __restrict__ rec, const WalkTree& T_, uint32_t s0,
                                           uint32_t count, uint32_t lane, uint32_t& inter, uint64_t& opened,
                                           uint64_t& visits, uint32_t& max_sp, uint32_t& err) {
  constexpr uint32_t NK = 1u << D, P = WAVE / NK, SPW = WIDE_DEPTH - WIDE_LDS;
  uint32_t* sd = sk + SPW * NK;  // spilled depths, then masks
  uint32_t* sm = sd + SPW;
  const uint32_t e = lane / NK, c = lane % NK;
  const uint32_t off = T_.off, n = T_.n;
  const double thr0 = T_.thr0, dmax2 = T_.dmax2;
  const float dmin2f = T_.dmin2f, dmax2f = T_.dmax2f, alpha = T_.alpha;
  const uint32_t step_no = T_.step_no, seed = T_.seed;
  const uint32_t guard_max = 2u * n + 8u;
  float4 me[T];  // the group's points, the same on every lane (SGPRs)
  uint32_t self[T];
#pragma unroll
  for (uint32_t b = 0; b < T; ++b) {
    const float4 q = b < count ? leaves[s0 + b] : make_float4(0.f, 0.f, 0.f, 0.f);
    me[b].x = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(q.x)));
    me[b].y = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(q.y)));
    me[b].z = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(q.z)));
    me[b].w = 0.f;
    self[b] = (uint32_t)__builtin_amdgcn_readfirstlane(b < count ? order[s0 + b] - off : 0u);
  }
  float ax[T], ay[T], az[T];
#pragma unroll
  for (uint32_t b = 0; b < T; ++b) ax[b] = ay[b] = az[b] = 0.f;
  // the entry above the root
  if (lane < NK) L.kids[0][lane] = lane == 0 ? off : REF_NONE;
  if (lane == 0) L.depth[0] = ~0u, L.mask[0] = count >= 32u ? ~0u : (1u << count) - 1u;
  uint32_t sp = 1;
  for (uint32_t guard = 0; sp > 0u; ++guard) {  // sp is the same on every lane (kept in SGPRs)
    if (guard > guard_max) {
      err = 1;
      break;
    }
    const uint32_t p = sp + (NK - 1) * P <= WIDE_BROAD ? (sp < P ? sp : P) : 1u;
    const uint32_t base = sp - p;
    opened += p;
    if (sp > WIDE_LDS) wait_all_memory();  // the wave's own spill stores, complete before these loads
    uint32_t k = REF_NONE, top = 0, msk = 0;
    if (e < p) {
      const uint32_t ent = sp - 1u - e;
      if (ent < WIDE_LDS) {
        k = L.kids[ent][c];
        top = L.depth[ent] + 1u;
        msk = L.mask[ent];
      } else {
        k = load_coherent(&sk[(ent - WIDE_LDS) * NK + c]);
        top = load_coherent(&sd[ent - WIDE_LDS]) + 1u;
        msk = load_coherent(&sm[ent - WIDE_LDS]);
      }
    }
    uint32_t omask = 0;  // the points that must open this lane's child
    uint4 k0, k1;        // set, with cdepth, wherever omask is
    uint32_t cdepth;
    if (k != REF_NONE) {
      if (c == 0) visits += (uint32_t)__popc(msk);
      if (k & REF_POINT) {
        const uint32_t j = k & ~REF_POINT;
        const float4 q = leaves[j];
#pragma unroll
        for (uint32_t b = 0; b < T; ++b)
          if (((msk >> b) & 1u) && j != s0 + b)
            point_force<D, ACC_FUSED>(me[b].x, me[b].y, me[b].z, self[b], q, order, j, off, dmax2f, dmin2f, alpha,
                                      step_no, seed, ax[b], ay[b], az[b], inter);
      } else {
        const uint4* R = (const uint4*)(rec + k);
        const uint4 q0 = R[0], q1 = R[1];
        k0 = R[2];
        k1 = R[3];
        NodeHdr h;
        __builtin_memcpy(&h, &q0, 16);
        __builtin_memcpy((char*)&h + 16, &q1, 16);
#pragma unroll
        for (uint32_t b = 0; b < T; ++b)
          if ((msk >> b) & 1u)
            if (!cell_force<D, ACC_FUSED>(me[b].x, me[b].y, me[b].z, self[b], n, k - off, top, thr0, h, dmax2,
                                          dmin2f, alpha, step_no, seed, ax[b], ay[b], az[b], inter))
              omask |= 1u << b;
        cdepth = h.depth;
      }
    }
    const bool open = omask != 0u;
    const uint64_t bits = __builtin_amdgcn_ballot_w64(open);
    const uint32_t nsp = (uint32_t)__builtin_amdgcn_readfirstlane(base + (uint32_t)__popcll(bits));
    if (nsp > WIDE_DEPTH) {  // no entry past the stack is written: stop here
      err = 1;
      break;
    }
    if (open) {
      const uint32_t slot = base + mask_rank(bits);
      const uint32_t kk[8] = {k0.x, k0.y, k0.z, k0.w, k1.x, k1.y, k1.z, k1.w};
      if (slot < WIDE_LDS) {
        L.depth[slot] = cdepth;
        L.mask[slot] = omask;
        uint32_t* en = L.kids[slot];
        if (NK == 8) {
          *(uint4*)en = k0;
          *(uint4*)(en + 4) = k1;
        } else if (NK == 4) {
          *(uint4*)en = k0;
        } else {
          *(uint2*)en = make_uint2(k0.x, k0.y);
        }
      } else {
#pragma unroll
        for (uint32_t q = 0; q < NK; ++q) store_coherent(&sk[(slot - WIDE_LDS) * NK + q], kk[q]);
        store_coherent(&sd[slot - WIDE_LDS], cdepth);
        store_coherent(&sm[slot - WIDE_LDS], omask);
      }
    }
    sp = nsp;
    max_sp = sp > max_sp ? sp : max_sp;
  }
  // each point's force: its lanes' sums added in lane 0's fixed tree
#pragma unroll
  for (uint32_t b = 0; b < T; ++b) {
    float sx = ax[b], sy = ay[b], sz = az[b];
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
      sx += __shfl_xor(sx, o, WAVE);
      if (D > 1) sy += __shfl_xor(sy, o, WAVE);
      if (D > 2) sz += __shfl_xor(sz, o, WAVE);
    }
    if (lane == 0 && b < count && !err) {
      T_.vx[self[b]] += sx;
      if (D > 1) T_.vy[self[b]] += sy;
      if (D > 2) T_.vz[self[b]] += sz;
    }
  }
}

template <int D, uint32_t T, uint32_t WAVES>
__global__ void __launch_bounds__(WAVES * WAVE) tree_walk_group(const float4* __restrict__ leaves,
                                                        const uint32_t* __restrict__ order,
                                                        const NodeRec* __restrict__ rec,
                                                        const CubeInfo* __restrict__ cube, float* __restrict__ vx,
                                                        float* __restrict__ vy, float* __restrict__ vz,
                                                        uint32_t* __restrict__ slots, uint32_t* __restrict__ spill,
                                                        StepParams step, WalkArgs A) {
  __shared__ GroupLDS<D, T> lds[WAVES];
  const uint32_t lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x / WAVE;
  const uint32_t gw = blockIdx.x * WAVES + wave;
  uint32_t* sk = spill + (size_t)gw * group_spill_words<D>();
  WalkTree W;
  W.off = 0;
  W.n = A.n;
  W.size = (double)cube->size;
  W.thr0 = W.size * W.size / A.theta2;
  W.dmax2 = A.dmax2;
  W.dmin2f = (float)A.dmin2;
  W.dmax2f = (float)fmin(A.dmax2, 3.0e38);
  W.alpha = step.alpha;
  W.step_no = step.step;
  W.seed = step.seed;
  W.vx = vx;
  W.vy = vy;
  W.vz = vz;
  uint32_t inter = 0, max_sp = 0, err = 0;
  uint64_t opened = 0, visits = 0;
  const uint32_t groups = (A.n + T - 1) / T;
  const WalkRange wr = xcd_range(groups, WAVES, wave);
  for (uint32_t g = wr.first; g < wr.end && !err; g += wr.stride) {
    const uint32_t s0 = g * T, count = A.n - s0 < T ? A.n - s0 : T;
    group_walk<D, T>(lds[wave], sk, leaves, order, rec, W, s0, count, lane, inter, opened, visits, max_sp, err);
  }
  WalkCount C;
  C.opened = opened;
  C.lane_visits = visits;
  C.inter = inter;
  C.err = err;
  C.max_sp = max_sp;
  walk_flush_slot(C, lane, slots + (size_t)blockIdx.x * 8);
}

// Resident workgroups of a persistent walk kernel on this device (all CUs).
template <typename K>
uint32_t resident_blocks(K kern, int threads, int cus) {
  int per_cu = 0;
  FG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, threads, 0));
  if (per_cu < 1) per_cu = 1;
  return (uint32_t)per_cu * (uint32_t)(cus > 0 ? cus : 1);
}

// Bytes of spill area a persistent walk needs: WALK_DEPTH - STACK entries for each of the
// at most 32 waves per CU.
inline size_t spill_bytes(int cus, int stack) {
  return (size_t)(cus > 0 ? cus : 1) * 32 * (size_t)(WALK_DEPTH - stack) * sizeof(SpillEntry);
}

__global__ void tree_counters_clear(uint32_t* c) {
  if (threadIdx.x < TREE_COUNTERS) c[threadIdx.x] = 0;
}

// ---- batched trees ------------------------------------------------------------------
//
// The trees of many jobs at once over the tree space of a batch (jobs with a many-body
// force, each a segment [off, off + n)).  The codes of all jobs are sorted together, then
// stably by segment, which leaves every job's points in its own segment in the order of a
// sort of that job alone; build, reduction and walk then work on job-relative indices, so
// each tree is the one the single-simulation path builds for that job.

template <int D>
__global__ void __launch_bounds__(256) btree_bbox(const BTile* __restrict__ tiles,
                                                  const uint32_t* __restrict__ seg_job, BJob* jobs) {
  const BTile t = tiles[blockIdx.x];
  BJob* B = jobs + seg_job[t.seg];
  if (!B->active) return;
  const uint32_t off = B->tree_off;
  const float* x = B->p[0];
  const float* y = B->p[1];
  const float* z = B->p[2];
  float lo[3] = {3.4e38f, 3.4e38f, 3.4e38f}, hi[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
  for (uint32_t i = t.first + threadIdx.x; i < t.first + t.count; i += 256) {
    const uint32_t v = i - off;
    float p[3] = {x[v], D > 1 ? y[v] : 0.0f, D > 2 ? z[v] : 0.0f};
#pragma unroll
    for (int a = 0; a < 3; ++a) lo[a] = fminf(lo[a], p[a]), hi[a] = fmaxf(hi[a], p[a]);
  }
  block_box(lo, hi, B->cube_box);
  __shared__ uint32_t s_last;
  if (threadIdx.x == 0) s_last = publish_count(&B->cube_done) + 1 == B->tree_tiles;
  __syncthreads();
  if (!s_last || threadIdx.x != 0) return;
  float origin[3], size;
  int32_t shift;
  cover_box<D>(B->cube_box, origin, size, shift);
#pragma unroll
  for (int a = 0; a < 3; ++a) B->cube_origin[a] = origin[a];
  B->cube_size = size;
  B->cube_shift = shift;
  B->cube_done = 0;
}

template <int D>
__global__ void __launch_bounds__(256) btree_keys(uint32_t total, const uint32_t* __restrict__ seg_of,
                                                  const uint32_t* __restrict__ seg_job,
                                                  const BJob* __restrict__ jobs,
                                                  uint64_t* __restrict__ keys,
                                                  uint64_t* __restrict__ codes,
                                                  uint32_t* __restrict__ vals,
                                                  float4* __restrict__ pts) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= total) return;
  const BJob& B = jobs[seg_job[seg_of[i]]];
  const uint32_t v = i - B.tree_off;
  const float p[3] = {B.p[0][v], D > 1 ? B.p[1][v] : 0.0f, D > 2 ? B.p[2][v] : 0.0f};
  const uint64_t key = point_key<D>(p, B.cube_origin, B.cube_shift);
  keys[i] = key;
  codes[i] = key;
  vals[i] = i;
  pts[i] = make_float4(p[0], p[1], p[2], B.mb_strength[v]);
}

// The segment of each code-sorted entry: the key of the last, stable, sort pass.
__global__ void __launch_bounds__(256) btree_digit(uint32_t total, const uint32_t* __restrict__ sv,
                                                   const uint32_t* __restrict__ seg_of,
                                                   uint32_t* __restrict__ digit) {
  const uint32_t s = blockIdx.x * 256 + threadIdx.x;
  if (s < total) digit[s] = seg_of[sv[s]];
}

__global__ void __launch_bounds__(256) btree_gather(uint32_t total, const uint32_t* __restrict__ order,
                                                    const uint64_t* __restrict__ codes,
                                                    uint64_t* __restrict__ skeys) {
  const uint32_t s = blockIdx.x * 256 + threadIdx.x;
  if (s < total) skeys[s] = codes[order[s]];
}

template <int D>
__global__ void __launch_bounds__(256) btree_leaves(uint32_t total, const uint64_t* __restrict__ skeys,
                                                    const uint32_t* __restrict__ order,
                                                    const uint32_t* __restrict__ seg_of,
                                                    const uint32_t* __restrict__ seg_job,
                                                    const BJob* __restrict__ jobs,
                                                    const float4* __restrict__ pts,
                                                    float4* __restrict__ leaves,
                                                    uint32_t* __restrict__ leaf_depth,
                                                    uint32_t* __restrict__ flags) {
  const uint32_t s = blockIdx.x * 256 + threadIdx.x;
  if (s >= total) return;
  const BJob& B = jobs[seg_job[seg_of[s]]];
  const uint32_t off = B.tree_off, n = B.n;
  leaves[s] = pts[order[s]];
  const int ls = (int)(s - off);
  int a = delta<D>(skeys + off, (int)n, ls, ls - 1), b = delta<D>(skeys + off, (int)n, ls, ls + 1);
  leaf_depth[s] = depth_of<D>(a > b ? a : b) + 1;
  flags[s] = 0;
}

template <int D>
__global__ void __launch_bounds__(256) btree_build(uint32_t total, const uint64_t* __restrict__ skeys,
                                                   const uint32_t* __restrict__ seg_of,
                                                   const uint32_t* __restrict__ seg_job,
                                                   const BJob* __restrict__ jobs,
                                                   uint4* __restrict__ links, uint32_t* __restrict__ leaf_parent) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= total) return;
  const BJob& B = jobs[seg_job[seg_of[i]]];
  const uint32_t off = B.tree_off, k = i - off;
  if (k + 1 >= B.n) return;
  karras_write<D>(skeys + off, (int)B.n, (int)k, off, REDUCE_BLOCK, links, leaf_parent);
}

// Persistent walk: every wave takes 64-point tiles of any job from one queue until it is
// empty; the per-tile work is the single-simulation tiled walk's in its cheaper mode with
// the pool of single-point openings, for jobs of every size.  A job's tiles are its own 64
// points at a time from its first, so its forces are the same alone and in any mix.
// (Config D's 256 jobs of 500-50,000 nodes: 131-142 -> 166 layouts/s against the exact
// mode below 163,840 points.)
template <int D, int STACK, int WAVES>
__global__ void __launch_bounds__(WAVES * WAVE, STACK <= 64 ? POOL_WAVES : 1)
    btree_walk(const float4* __restrict__ leaves, const uint32_t* __restrict__ order,
               const NodeRec* __restrict__ rec, const uint2* __restrict__ tiles, uint32_t ntiles,
               const uint32_t* __restrict__ seg_job, const BJob* __restrict__ jobs,
               uint32_t* __restrict__ counters, SpillEntry* __restrict__ spill, uint32_t packed_max) {
  __shared__ WalkLDS<STACK, true> lds[WAVES];
  const uint32_t lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x / WAVE;
  SpillEntry* my_spill = spill + (size_t)(blockIdx.x * WAVES + wave) * (WALK_DEPTH - STACK);
  WalkCount C;
  for (;;) {
    uint32_t t = 0;
    if (lane == 0) t = atomicAdd(&counters[5], 1u);
    t = (uint32_t)__builtin_amdgcn_readfirstlane(t);
    if (t >= ntiles) break;
    const uint2 tile = tiles[t];
    const BJob& B = jobs[seg_job[tile.x]];
    if (!B.active) continue;
    WalkTree T;
    T.off = B.tree_off;
    T.n = B.n;
    T.size = (double)B.cube_size;
    T.thr0 = T.size * T.size / B.theta2;
    T.dmax2 = B.dmax2;
    T.dmin2f = (float)B.dmin2;
    T.dmax2f = (float)fmin(B.dmax2, 3.0e38);
    T.alpha = B.alpha_f;
    T.step_no = B.step;
    T.seed = B.seed;
    T.vx = B.v[B.vcur][0];
    T.vy = B.v[B.vcur][1];
    T.vz = B.v[B.vcur][2];
    const uint32_t end = T.off + T.n;
    const uint32_t count = end - tile.y < (uint32_t)WAVE ? end - tile.y : (uint32_t)WAVE;
    __builtin_amdgcn_wave_barrier();
    walk_tile<D, STACK, false, true>(lds[wave], leaves, order, rec, T, tile.y, count, packed_max, lane, my_spill, C);
  }
  walk_flush(C, lane, counters);
}

template <int D, int STACK, int WAVES>
void launch_bwalk(hipStream_t st, BatchTree& t, const BSpace& sp, BJob* jobs, const BatchTune& tune,
                  int cus) {
  auto kern = btree_walk<D, STACK, WAVES>;
  int per_cu = 0;
  FG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, WAVES * WAVE, 0));
  if (per_cu < 1) per_cu = 1;
  if (per_cu > 32 / WAVES) per_cu = 32 / WAVES;  // the spill area covers 32 waves per CU
  if (tune.walk_blocks_per_cu > 0 && tune.walk_blocks_per_cu < per_cu) per_cu = tune.walk_blocks_per_cu;
  uint32_t blocks = (uint32_t)(per_cu * (cus > 0 ? cus : 1));
  const uint32_t need = ceil_div(sp.walk_tiles, WAVES);
  if (blocks > need) blocks = need;
  if (!blocks) return;
  if (t.spill_bytes < spill_bytes(cus, STACK)) throw ArgError(-5, "walk: spill area too small");
  hipLaunchKernelGGL(kern, dim3(blocks), dim3(WAVES * WAVE), 0, st, (const float4*)t.leaves,
                     (const uint32_t*)t.order, (const NodeRec*)t.rec, (const uint2*)sp.walk,
                     sp.walk_tiles, (const uint32_t*)sp.seg_job, (const BJob*)jobs, t.counters,
                     (SpillEntry*)t.spill, tune.walk_packed_max);
  FG_LAUNCH_CHECK();
}

// The walk kernel for the tuned workgroup size and LDS stack depth (16, 32, 48, 64 or 96
// entries; deeper entries spill to global memory either way).
template <int D>
void launch_bwalk_tuned(hipStream_t st, BatchTree& t, const BSpace& sp, BJob* jobs,
                        const BatchTune& tune, int cus) {
  const int w = tune.walk_waves == 2 || tune.walk_waves == 8 ? tune.walk_waves : 4;
  const int k = tune.walk_stack;
#define FG_BWALK_W(S)                                                     \
  do {                                                                    \
    if (w == 2) launch_bwalk<D, S, 2>(st, t, sp, jobs, tune, cus);        \
    else if (w == 4) launch_bwalk<D, S, 4>(st, t, sp, jobs, tune, cus);   \
    else launch_bwalk<D, S, 8>(st, t, sp, jobs, tune, cus);               \
  } while (0)
  if (k <= 16) FG_BWALK_W(16);
  else if (k <= 32) FG_BWALK_W(32);
  else if (k <= 48) FG_BWALK_W(48);
  else if (k <= 64) FG_BWALK_W(64);
  else FG_BWALK_W(96);
#undef FG_BWALK_W
}

template <int D>
void run_batch(hipStream_t st, BatchTree& t, const BSpace& sp, BJob* jobs, const BatchTune& tune,
               int cus, PhaseTimer* timer) {
  const uint32_t total = sp.total;
  if (!total) return;
  t.reserve(total);
  const uint32_t blocks = ceil_div(total, 256);
  if (timer) timer->begin(1);
  hipLaunchKernelGGL(tree_counters_clear, dim3(1), dim3(64), 0, st, t.counters);
  hipLaunchKernelGGL((btree_bbox<D>), dim3(sp.tiles), dim3(256), 0, st, (const BTile*)sp.tile,
                     (const uint32_t*)sp.seg_job, jobs);
  FG_LAUNCH_CHECK();
  hipLaunchKernelGGL((btree_keys<D>), dim3(blocks), dim3(256), 0, st, total, (const uint32_t*)sp.seg_of,
                     (const uint32_t*)sp.seg_job, (const BJob*)jobs, t.keys, t.codes, t.vals, t.pts);
  FG_LAUNCH_CHECK();
  if (timer) timer->end(1), timer->begin(2);
  // high halves, then jobs (stable: within a job the high halves stay in order), then the
  // codes gathered in that order and each run of equal (job, high half) ordered on the low
  // halves: the order of (job, code, tree index)
  uint64_t* hk;
  uint32_t* hv;
  radix_sort_u64(st, t.sort, tune.radix_bits, t.keys, t.vals, total, Dims<D>::HALF, Dims<D>::HALF, &hk, &hv);
  uint32_t* order1 = hv;
  uint32_t* sd = nullptr;
  if (sp.segs > 1) {
    hipLaunchKernelGGL(btree_digit, dim3(blocks), dim3(256), 0, st, total, (const uint32_t*)hv,
                       (const uint32_t*)sp.seg_of, t.digit);
    FG_LAUNCH_CHECK();
    uint32_t bits = 1;
    while ((1u << bits) < sp.segs) ++bits;
    radix_sort_u32(st, t.sort_seg, tune.radix_bits, t.digit, hv, total, 0, bits, &sd, &order1);
  }
  hipLaunchKernelGGL(btree_gather, dim3(blocks), dim3(256), 0, st, total, (const uint32_t*)order1,
                     (const uint64_t*)t.codes, hk);
  FG_LAUNCH_CHECK();
  uint32_t* order = hv == t.vals ? t.sort.vals_alt : t.vals;  // aliases none of hv, order1, sd
  hipLaunchKernelGGL((order_runs<D>), dim3(blocks), dim3(256), 0, st, total, (const uint64_t*)hk,
                     (const uint32_t*)order1, (const uint32_t*)sd, t.skeys, order, t.counters, t.long_runs);
  FG_LAUNCH_CHECK();
  hipLaunchKernelGGL((order_long_runs<D>), dim3(LONG_RUN_BLOCKS), dim3(256), 0, st, total,
                     (const uint32_t*)t.counters, (const uint32_t*)t.long_runs, (const uint32_t*)sd, t.skeys, order,
                     hk, order1);
  FG_LAUNCH_CHECK();
  if (timer) timer->end(2), timer->begin(1);
  t.order = order;
  hipLaunchKernelGGL((btree_leaves<D>), dim3(blocks), dim3(256), 0, st, total, (const uint64_t*)t.skeys,
                     (const uint32_t*)order, (const uint32_t*)sp.seg_of, (const uint32_t*)sp.seg_job,
                     (const BJob*)jobs, (const float4*)t.pts, t.leaves, t.leaf_depth, t.flags);
  FG_LAUNCH_CHECK();
  hipLaunchKernelGGL((btree_build<D>), dim3(blocks), dim3(256), 0, st, total, (const uint64_t*)t.skeys,
                     (const uint32_t*)sp.seg_of, (const uint32_t*)sp.seg_job, (const BJob*)jobs, t.links,
                     t.leaf_parent);
  FG_LAUNCH_CHECK();
  hipLaunchKernelGGL((tree_reduce<D, REDUCE_BLOCK, false>), dim3(ceil_div(total, REDUCE_BLOCK)), dim3(REDUCE_BLOCK), 0, st,
                     total, (const float4*)t.leaves, (const uint32_t*)t.leaf_parent, (const uint4*)t.links,
                     t.flags, (NodeRec*)t.rec, (NodeSum*)t.sums);
  FG_LAUNCH_CHECK();
  hipLaunchKernelGGL((tree_collapse<D>), dim3(blocks), dim3(256), 0, st, total, (const uint32_t*)sp.seg_of,
                     (const uint32_t*)sp.seg_job, (const BJob*)jobs, (const uint4*)t.links, (NodeRec*)t.rec);
  FG_LAUNCH_CHECK();
  if (timer) timer->end(1), timer->begin(3);
  launch_bwalk_tuned<D>(st, t, sp, jobs, tune, cus);
  if (timer) timer->end(3);
}

// The walk over a built tree: forces into v (barnesHut.js:434-470).
template <int D>
void walk(hipStream_t st, TreeBuffers& t, uint32_t n, float* vx, float* vy, float* vz, const TreeParams& p,
          StepParams step, PhaseTimer* timer) {
  if (timer) timer->begin(3);
  WalkArgs A;
  A.theta2 = p.theta2;
  A.dmin2 = p.dmin2;
  A.dmax2 = p.dmax2;
  A.n = n;
  const int cus = device_cus();
  const bool exact = n <= p.walk_exact_max;
  // measured (link + manyBody + center ticks): 32 for the exact mode (100,000 points 0.511
  // -> 0.495 ms/tick against 24, 20,000-50,000 flat); 24 for the cheaper mode (1M: 32 within
  // the noise, 40 and 16 slower)
  A.packed_max = p.walk_packed_max != WALK_PACKED_BY_MODE ? p.walk_packed_max : exact ? 32u : 24u;
  A.pool_k = p.walk_pool_k != WALK_POOL_BY_MODE ? p.walk_pool_k : POOL_K;
  A.pool_cap = p.walk_pool_cap < POOL_SIZE ? p.walk_pool_cap : POOL_SIZE;
  // measured on 256 CUs (link + manyBody + center ticks): a group of lanes per point is the
  // faster walk up to 12,000 points in 3D (0.254 against 0.258 ms/tick) but not at 16,000
  // (0.296 against 0.264), and still at 24,000 in 2D (0.234 against 0.237) and 32,000 in 1D
  // (0.259 against 0.291)
  const uint32_t solo_max = p.walk_solo_max != WALK_SOLO_BY_DIMS ? p.walk_solo_max
                            : D == 3 ? 12288u : D == 2 ? 24576u : 32768u;
  t.slot_blocks = 0;
  const uint32_t wide_max = p.walk_wide_max != WALK_WIDE_BY_DIMS ? p.walk_wide_max : WALK_WIDE_MAX;
  const bool forced = p.walk_group >= 2u && p.walk_group <= 16u && (p.walk_group & (p.walk_group - 1u)) == 0u;
  const bool by_size = p.walk_group == 0u && n > wide_max && n <= WALK_GROUP_MAX;
  if (forced || by_size) {  // a group of points per wave
    const uint32_t T = forced ? p.walk_group : WALK_GROUP;
    auto launch = [&](auto kern) {
      uint32_t blocks = resident_blocks(kern, WIDE_WAVES * WAVE, cus);
      if (blocks > WALK_SLOTS) blocks = WALK_SLOTS;
      const uint32_t groups = ceil_div(n, T);
      if (blocks > ceil_div(groups, WIDE_WAVES)) blocks = ceil_div(groups, WIDE_WAVES);
      if (blocks > 8) blocks -= blocks % 8;  // a whole number of workgroups per XCD
      const size_t sbytes = sizeof(uint32_t) * (size_t)blocks * WIDE_WAVES * group_spill_words<D>();
      if (t.spill_bytes < sbytes) t.reserve_spill(sbytes);
      hipLaunchKernelGGL(kern, dim3(blocks), dim3(WIDE_WAVES * WAVE), 0, st, (const float4*)t.leaves,
                         (const uint32_t*)t.order, (const NodeRec*)t.rec, (const CubeInfo*)t.cube, vx, vy, vz,
                         t.slots, (uint32_t*)t.spill, step, A);
      FG_LAUNCH_CHECK();
      t.slot_blocks = blocks;
    };
    if (T == 2) launch(tree_walk_group<D, 2, WIDE_WAVES>);
    else if (T == 4) launch(tree_walk_group<D, 4, WIDE_WAVES>);
    else if (T == 8) launch(tree_walk_group<D, 8, WIDE_WAVES>);
    else launch(tree_walk_group<D, 16, WIDE_WAVES>);
    if (timer) timer->end(3);
    return;
  }
  // measured (link + manyBody + center ticks, 3D): a wave per point up to about 10,000
  // points (5,000: 0.095 against 0.104 ms/tick for two per wave), two per wave from there
  // (50,000 0.219 against 0.237, 100,000 0.321 against 0.373, 200,000 0.536 against the
  // tiled walk's 0.656), the tiled walk from 300,000 (500,000: 1.160 against 1.193)
  if (n <= wide_max) {  // a wave per point
    auto wk = tree_walk_wide<D, WIDE_WAVES>;
    uint32_t blocks = resident_blocks(wk, WIDE_WAVES * WAVE, cus);
    if (blocks > WALK_SLOTS) blocks = WALK_SLOTS;
    if (blocks > ceil_div(n, WIDE_WAVES)) blocks = ceil_div(n, WIDE_WAVES);
    const size_t sbytes = sizeof(uint32_t) * (size_t)blocks * WIDE_WAVES * wide_spill_words<D>();
    if (t.spill_bytes < sbytes) t.reserve_spill(sbytes);
    hipLaunchKernelGGL(wk, dim3(blocks), dim3(WIDE_WAVES * WAVE), 0, st, (const float4*)t.leaves,
                       (const uint32_t*)t.order, (const NodeRec*)t.rec, (const CubeInfo*)t.cube, vx, vy, vz, t.slots,
                       (uint32_t*)t.spill, step, A);
    FG_LAUNCH_CHECK();
    t.slot_blocks = blocks;
    if (timer) timer->end(3);
    return;
  }
  if (exact && n <= solo_max) {  // a group of lanes per point
    constexpr uint32_t G = WAVE >> D;
    if (ceil_div(n, G) <= (uint32_t)cus)
      hipLaunchKernelGGL((tree_walk_solo<D, 1>), dim3(ceil_div(n, G)), dim3(WAVE), 0, st, (const float4*)t.leaves,
                         (const uint32_t*)t.order, (const NodeRec*)t.rec, (const CubeInfo*)t.cube, vx, vy, vz,
                         t.counters, step, A);
    else
      hipLaunchKernelGGL((tree_walk_solo<D, 2>), dim3(ceil_div(n, 2 * G)), dim3(2 * WAVE), 0, st,
                         (const float4*)t.leaves, (const uint32_t*)t.order, (const NodeRec*)t.rec,
                         (const CubeInfo*)t.cube, vx, vy, vz, t.counters, step, A);
    FG_LAUNCH_CHECK();
    if (timer) timer->end(3);
    return;
  }
  auto kern = exact ? tree_walk<D, WALK_STACK, WALK_WAVES, true> : tree_walk<D, WALK_STACK, WALK_WAVES, false>;
  uint32_t wblocks = resident_blocks(kern, WALK_WAVES * WAVE, cus);
  if (wblocks > (uint32_t)cus * (32 / WALK_WAVES)) wblocks = (uint32_t)cus * (32 / WALK_WAVES);
  A.tile = exact ? walk_tile_for(n, 4u * (uint32_t)cus, wblocks * WALK_WAVES, p.walk_tile) : (uint32_t)WAVE;
  const uint32_t need = ceil_div(ceil_div(n, A.tile), WALK_WAVES);
  if (wblocks > need) wblocks = need;
  if (t.spill_bytes < spill_bytes(cus, WALK_STACK)) t.reserve_spill(spill_bytes(cus, WALK_STACK));
  hipLaunchKernelGGL(kern, dim3(wblocks), dim3(WALK_WAVES * WAVE), 0, st, (const float4*)t.leaves,
                     (const uint32_t*)t.order, (const NodeRec*)t.rec, (const CubeInfo*)t.cube, vx, vy, vz,
                     t.counters, (SpillEntry*)t.spill, step, A);
  FG_LAUNCH_CHECK();
  if (timer) timer->end(3);
}

template <int D>
void build(hipStream_t st, TreeBuffers& t, uint32_t n, const float* x, const float* y, const float* z,
           const float* strength, const TreeParams& p, int radix_bits, PhaseTimer* timer,
           const CenterFuse* center, const LinkFuse* link) {
  if (timer) timer->begin(1);
  if (n <= REDUCE_ONE) {  // one workgroup builds and reduces the whole tree (more run a link pass)
    const uint32_t blocks = link ? 1 + ceil_div(n, REDUCE_ONE) : 1;
    hipLaunchKernelGGL((tree_small<D>), dim3(blocks), dim3(REDUCE_ONE), 0, st, x, y, z, strength, n, t.cube,
                       t.counters, t.keys, t.vals, t.leaves, t.leaf_depth, t.links, t.flags, t.rec, t.sums,
                       link ? *link : LinkFuse{});
    FG_LAUNCH_CHECK();
    t.skeys = t.keys;
    t.order = t.vals;
    if (timer) timer->end(1);
    return;
  }
  const uint32_t blocks = ceil_div(n, 256);
  const SortClear sc = sort_prepare(t.sort, radix_bits, n, Dims<D>::HALF);
  const uint32_t bb = blocks < BBOX_BLOCKS ? blocks : BBOX_BLOCKS;
  static_assert(BBOX_BLOCKS == CENTER_BLOCKS, "the box kernel takes center_reduce's grid");
  const uint32_t link_blocks = link ? blocks : 0u;  // the link pass: one node per thread
  hipLaunchKernelGGL((tree_bbox<D>), dim3(bb + link_blocks), dim3(256), 0, st, x, y, z, n, t.cube, t.counters, sc,
                     center ? *center : CenterFuse{}, bb, link ? *link : LinkFuse{});
  FG_LAUNCH_CHECK();
  // the counting sort's ranks and counters live in leaf_parent and flags until the build
  const bool counting = n <= RANK_MAX;
  hipLaunchKernelGGL((tree_keys<D>), dim3(blocks), dim3(256), 0, st, x, y, z, strength, n,
                     (const CubeInfo*)t.cube, t.keys, t.vals, t.pts, counting ? t.leaf_parent : nullptr,
                     t.flags);
  FG_LAUNCH_CHECK();
  if (timer) timer->end(1), timer->begin(2);
  uint64_t* sk;
  uint32_t* sv;
  if (counting) {
    sk = (uint64_t*)t.sort.keys_alt;
    sv = t.sort.vals_alt;
    const uint32_t groups = ceil_div(n, 64), want = ceil_div(2u * (uint32_t)device_cus(), groups);
    const uint32_t parts = want < 1 ? 1 : want > RANK_WAVES ? RANK_WAVES : want;
    hipLaunchKernelGGL(tree_rank_sort, dim3(groups, parts), dim3(RANK_WAVES * 64),
                       sizeof(uint64_t) * ceil_div(n, parts), st, (const uint64_t*)t.keys, (const uint32_t*)t.vals,
                       n, t.leaf_parent, t.flags, sk, sv);
    FG_LAUNCH_CHECK();
  } else {
    uint64_t* hk;
    uint32_t* hv;
    // the high halves, then every run of equal high halves in order (into the other buffers)
    radix_sort_u64(st, t.sort, radix_bits, t.keys, t.vals, n, Dims<D>::HALF, Dims<D>::HALF, &hk, &hv, true);
    sk = hk == t.keys ? (uint64_t*)t.sort.keys_alt : t.keys;
    sv = hv == t.vals ? t.sort.vals_alt : t.vals;
    hipLaunchKernelGGL((order_runs<D>), dim3(blocks), dim3(256), 0, st, n, (const uint64_t*)hk, (const uint32_t*)hv,
                       (const uint32_t*)nullptr, sk, sv, t.counters, t.long_runs);
    FG_LAUNCH_CHECK();
    hipLaunchKernelGGL((order_long_runs<D>), dim3(LONG_RUN_BLOCKS), dim3(256), 0, st, n, (const uint32_t*)t.counters,
                       (const uint32_t*)t.long_runs, (const uint32_t*)nullptr, sk, sv, hk, hv);
    FG_LAUNCH_CHECK();
  }
  if (timer) timer->end(2), timer->begin(1);
  t.skeys = sk;
  t.order = sv;
  // points per reduction workgroup: 1,024 (windows of 512: fewer shared nodes, centres
  // cached in LDS, one workgroup per CU) measured faster up to 100,000 points (0.383 -> 0.376
  // ms/tick there, 50,000 even), 512 from 200,000 (0.663 against 0.669)
  const uint32_t rb = p.reduce_block == 0 ? (n <= 131072u ? REDUCE_ONE : REDUCE_BLOCK)
                      : p.reduce_block == REDUCE_ONE ? REDUCE_ONE : REDUCE_BLOCK;
  // subtree-aligned workgroups (windows of rb / 2 points) unless turned off
  const bool aligned = p.reduce_aligned;
  const uint32_t window = aligned ? rb / 2 : rb;
  hipLaunchKernelGGL((tree_leaves_build<D>), dim3(blocks), dim3(256), 0, st, (const uint64_t*)sk,
                     (const uint32_t*)sv, (const float4*)t.pts, n, window, t.leaves, t.leaf_depth, t.flags,
                     t.links, t.leaf_parent, aligned ? t.cut_start : nullptr);
  FG_LAUNCH_CHECK();
  const uint8_t* cuts = aligned ? (const uint8_t*)t.cut_start : nullptr;
  if (rb == REDUCE_ONE)
    hipLaunchKernelGGL((tree_reduce<D, REDUCE_ONE, true>), dim3(ceil_div(n, window)), dim3(REDUCE_ONE), 0, st,
                       n, (const float4*)t.leaves, (const uint32_t*)t.leaf_parent, (const uint4*)t.links, t.flags,
                       t.rec, t.sums, cuts);
  else
    hipLaunchKernelGGL((tree_reduce<D, REDUCE_BLOCK, true>), dim3(ceil_div(n, window)), dim3(REDUCE_BLOCK), 0,
                       st, n, (const float4*)t.leaves, (const uint32_t*)t.leaf_parent,
                       (const uint4*)t.links, t.flags, t.rec, t.sums, cuts);
  FG_LAUNCH_CHECK();
  if (timer) timer->end(1);
}

}  // namespace

int device_cus() {
  int dev = 0, cus = 0;
  FG_HIP(hipGetDevice(&dev));
  FG_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  return cus;
}

void TreeBuffers::reserve(uint32_t n) {
  if (n <= cap && keys) return;
  release();
  uint32_t m = n > 2 ? n : 2;
  FG_HIP(hipMalloc(&keys, sizeof(uint64_t) * m));
  FG_HIP(hipMalloc(&vals, sizeof(uint32_t) * m));
  FG_HIP(hipMalloc(&leaves, sizeof(float4) * m));
  FG_HIP(hipMalloc(&pts, sizeof(float4) * m));
  FG_HIP(hipMalloc(&leaf_depth, sizeof(uint32_t) * m));
  FG_HIP(hipMalloc(&leaf_parent, sizeof(uint32_t) * m));
  FG_HIP(hipMalloc(&links, sizeof(uint4) * m));
  FG_HIP(hipMalloc(&flags, sizeof(uint32_t) * m));
  FG_HIP(hipMalloc(&rec, sizeof(NodeRec) * m));
  FG_HIP(hipMalloc(&sums, sizeof(NodeSum) * m));
  FG_HIP(hipMalloc(&cube, sizeof(CubeInfo)));
  FG_HIP(hipMalloc(&counters, sizeof(uint32_t) * TREE_COUNTERS));
  FG_HIP(hipMalloc(&slots, sizeof(uint32_t) * 8 * WALK_SLOTS));
  FG_HIP(hipMalloc(&cut_start, m));
  FG_HIP(hipMalloc(&long_runs, sizeof(uint32_t) * (m / (ORDER_RUN + 1) + 1)));
  CubeInfo c0 = {};
  for (int a = 0; a < 3; ++a) c0.box[a] = 0xffffffffu, c0.box[3 + a] = 0u;
  upload_device(cube, &c0, sizeof(c0));
  clear_device(counters, sizeof(uint32_t) * TREE_COUNTERS);
  sort.reserve(m, sizeof(uint64_t));
  cap = n;
}

void TreeBuffers::reserve_spill(size_t bytes) {
  if (spill && spill_bytes >= bytes) return;
  if (spill) (void)hipFree(spill);
  spill = nullptr;
  spill_bytes = 0;
  FG_HIP(hipMalloc(&spill, bytes));
  spill_bytes = bytes;
}

void TreeBuffers::release() {
  if (spill) (void)hipFree(spill);
  spill = nullptr;
  spill_bytes = 0;
  void* ptrs[] = {keys, vals, leaves, pts, leaf_depth, leaf_parent, links, flags, rec, sums, cube, counters, slots,
                  cut_start, long_runs};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  keys = nullptr;
  vals = nullptr;
  leaves = pts = nullptr;
  leaf_depth = leaf_parent = flags = counters = slots = long_runs = nullptr;
  cut_start = nullptr;
  slot_blocks = 0;
  links = nullptr;
  rec = nullptr;
  sums = nullptr;
  cube = nullptr;
  skeys = nullptr;
  order = nullptr;
  sort.release();
  cap = 0;
}

void tree_build(hipStream_t s, TreeBuffers& t, uint32_t n, int ndim, const float* x, const float* y,
                const float* z, const float* strength, const TreeParams& p, int radix_bits, PhaseTimer* timer,
                const CenterFuse* center, const LinkFuse* link) {
  if (n < 1) return;
  if (center && !tree_box_pass(n)) throw std::logic_error("tree_build: no box kernel to take the center's mean");
  if (link && !tree_link_pass(n)) throw std::logic_error("tree_build: no one-workgroup build to take a link pass");
  t.reserve(n);
  if (ndim == 1) build<1>(s, t, n, x, y, z, strength, p, radix_bits, timer, center, link);
  else if (ndim == 2) build<2>(s, t, n, x, y, z, strength, p, radix_bits, timer, center, link);
  else build<3>(s, t, n, x, y, z, strength, p, radix_bits, timer, center, link);
}

void tree_forces(hipStream_t s, TreeBuffers& t, uint32_t n, int ndim, float* vx, float* vy, float* vz,
                 const TreeParams& p, StepParams step, PhaseTimer* timer) {
  if (n < 2) return;
  if (ndim == 1) walk<1>(s, t, n, vx, vy, vz, p, step, timer);
  else if (ndim == 2) walk<2>(s, t, n, vx, vy, vz, p, step, timer);
  else walk<3>(s, t, n, vx, vy, vz, p, step, timer);
}

void tree_manybody(hipStream_t s, TreeBuffers& t, uint32_t n, int ndim, const float* x,
                   const float* y, const float* z, const float* strength, float* vx, float* vy,
                   float* vz, const TreeParams& p, StepParams step, int radix_bits,
                   PhaseTimer* timer, const CenterFuse* center) {
  tree_build(s, t, n, ndim, x, y, z, strength, p, radix_bits, timer, center);
  tree_forces(s, t, n, ndim, vx, vy, vz, p, step, timer);
}

bool tree_box_pass(uint32_t n) { return n > REDUCE_ONE; }

bool tree_link_pass(uint32_t n) { return n >= 1 && n <= LINK_DIRECT_MAX; }

}  // namespace fg

namespace fg {

size_t spill_bytes_for(int cus, int stack) { return spill_bytes(cus, stack); }

void BatchTree::reserve(uint32_t n) {
  if (n <= cap && keys) return;
  release();
  uint32_t m = n > 2 ? n : 2;
  FG_HIP(hipMalloc(&keys, sizeof(uint64_t) * m));
  FG_HIP(hipMalloc(&codes, sizeof(uint64_t) * m));
  FG_HIP(hipMalloc(&skeys, sizeof(uint64_t) * m));
  FG_HIP(hipMalloc(&vals, sizeof(uint32_t) * m));
  FG_HIP(hipMalloc(&digit, sizeof(uint32_t) * m));
  FG_HIP(hipMalloc(&leaves, sizeof(float4) * m));
  FG_HIP(hipMalloc(&pts, sizeof(float4) * m));
  FG_HIP(hipMalloc(&leaf_depth, sizeof(uint32_t) * m));
  FG_HIP(hipMalloc(&leaf_parent, sizeof(uint32_t) * m));
  FG_HIP(hipMalloc(&links, sizeof(uint4) * m));
  FG_HIP(hipMalloc(&flags, sizeof(uint32_t) * m));
  FG_HIP(hipMalloc(&rec, sizeof(NodeRec) * m));
  FG_HIP(hipMalloc(&sums, sizeof(NodeSum) * m));
  FG_HIP(hipMalloc(&counters, sizeof(uint32_t) * TREE_COUNTERS));
  clear_device(counters, sizeof(uint32_t) * TREE_COUNTERS);
  FG_HIP(hipMalloc(&long_runs, sizeof(uint32_t) * (m / (ORDER_RUN + 1) + 1)));
  sort.reserve(m, sizeof(uint64_t));
  sort_seg.reserve(m, sizeof(uint32_t));
  if (!spill) {
    spill_bytes = spill_bytes_for(device_cus(), 16);
    FG_HIP(hipMalloc(&spill, spill_bytes));
  }
  cap = n;
}

void BatchTree::release() {
  void* ptrs[] = {keys, codes, skeys, vals, digit, leaves, pts, leaf_depth, leaf_parent, links, flags, rec, sums, counters, spill,
                  long_runs};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  keys = codes = skeys = nullptr;
  vals = digit = order = leaf_depth = leaf_parent = flags = counters = long_runs = nullptr;
  leaves = pts = nullptr;
  links = nullptr;
  rec = sums = spill = nullptr;
  spill_bytes = 0;
  sort.release();
  sort_seg.release();
  cap = 0;
}

void batch_manybody(hipStream_t s, int ndim, BatchTree& t, const BSpace& sp, BJob* jobs,
                    const BatchTune& tune, int cus, PhaseTimer* timer) {
  if (ndim == 1) run_batch<1>(s, t, sp, jobs, tune, cus, timer);
  else if (ndim == 2) run_batch<2>(s, t, sp, jobs, tune, cus, timer);
  else run_batch<3>(s, t, sp, jobs, tune, cus, timer);
}

}  // namespace fg
