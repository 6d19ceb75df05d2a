// common.h -- shared device helpers and host error plumbing of the physics backend.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <stdexcept>
#include <string>

namespace fg {

constexpr int WAVE = 64;

// A failed HIP call inside the library: caught at the C-ABI boundary and turned into
// FG_ERR_HIP with the message kept for fg_last_error.
struct HipError : std::runtime_error {
  hipError_t code;
  HipError(hipError_t c, const std::string& what) : std::runtime_error(what), code(c) {}
};

struct ArgError : std::runtime_error {
  int status;
  ArgError(int s, const std::string& what) : std::runtime_error(what), status(s) {}
};

#define FG_HIP(call)                                                                        \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess)                                                                  \
      throw ::fg::HipError(e_, std::string(#call) + ": " + hipGetErrorString(e_));          \
  } while (0)

#define FG_LAUNCH_CHECK() FG_HIP(hipGetLastError())

// Set-up writes to freshly allocated device memory (zeroed counters, initial records).  A
// plain hipMemset or hipMemcpy goes to the null stream, which the engines' non-blocking
// streams are not ordered after, and may return before it lands; memory reused from an
// earlier allocation would then still hold the old contents when a kernel first reads it.
// So these wait for the write before returning (they run at allocation time only).
inline void clear_device(void* p, size_t bytes) {
  FG_HIP(hipMemset(p, 0, bytes));
  FG_HIP(hipStreamSynchronize(nullptr));
}
inline void upload_device(void* dst, const void* src, size_t bytes) {
  FG_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  FG_HIP(hipStreamSynchronize(nullptr));
}

inline uint32_t ceil_div(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }
// Workgroups are dealt to the device's 8 XCDs in turn (workgroup b runs on XCD b % 8), each
// XCD with an L2 of its own.  A grid of xcd_blocks(blocks) workgroups numbers its logical
// blocks with xcd_block(): XCD x takes the contiguous logical blocks [x per, (x + 1) per),
// so neighbouring blocks -- neighbouring data -- share one L2.  Logical blocks past `blocks`
// (fewer than 8) have no work; grids under 64 blocks keep their order.
constexpr uint32_t XCD_COUNT = 8;
inline uint32_t xcd_blocks(uint32_t blocks) { return blocks < 64u ? blocks : ceil_div(blocks, XCD_COUNT) * XCD_COUNT; }
__device__ inline uint32_t xcd_block() {
  if (gridDim.x < 64u) return blockIdx.x;
  const uint32_t per = gridDim.x / XCD_COUNT;
  return (blockIdx.x % XCD_COUNT) * per + blockIdx.x / XCD_COUNT;
}

// A launch that a word on the device can cancel: skip() when the word (a non-negative
// float's bits) is at most `lim` -- every workgroup reads the same word, so all of a launch
// returns at once, before any barrier.  No word: never skips.
struct SkipIf {
  const uint32_t* word = nullptr;
  float lim = 0.f;
  __device__ bool operator()() const { return word && __uint_as_float(*word) <= lim; }
};

// ---- device helpers ------------------------------------------------------------------

// The reference jiggle hash (wgsl.js:28-39): pcg, then a value in [-5e-7, 5e-7).
__host__ __device__ inline uint32_t pcg_hash(uint32_t v) {
  uint32_t s = v * 747796405u + 2891336453u;
  uint32_t w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
  return (w >> 22u) ^ w;
}

__host__ __device__ inline float jiggle_hash(uint32_t a, uint32_t b, uint32_t axis, uint32_t step,
                                             uint32_t seed) {
  uint32_t h = pcg_hash(a ^ pcg_hash(b ^ pcg_hash(seed ^ pcg_hash(step * 3u + axis))));
  return ((float)h * (1.0f / 4294967296.0f) - 0.5f) * 1e-6f;
}

// Displacement between two coincident points, antisymmetric in (i, j) (wgsl.js:45-52).
__device__ inline void coincident(uint32_t i, uint32_t j, int ndim, uint32_t step, uint32_t seed,
                                  float& dx, float& dy, float& dz) {
  uint32_t lo = i < j ? i : j, hi = i < j ? j : i;
  float rx = jiggle_hash(lo, hi, 0u, step, seed);
  float ry = jiggle_hash(lo, hi, 1u, step, seed);
  float rz = jiggle_hash(lo, hi, 2u, step, seed);
  if (i > j) rx = -rx, ry = -ry, rz = -rz;
  dx = rx;
  dy = ndim > 1 ? ry : 0.0f;
  dz = ndim > 2 ? rz : 0.0f;
}

// Floats as u32 with the same order (for atomicMin / atomicMax), barnesHut.js:132-138.
__device__ inline uint32_t float_ordered(float f) {
  uint32_t b = __float_as_uint(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ inline float float_unordered(uint32_t u) {
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

__device__ inline uint64_t wave_ballot(bool p) { return __ballot(p); }

__device__ inline uint32_t lane_id() { return __lane_id(); }

// Number of set bits below this lane in a 64-bit wave mask.
__device__ inline uint32_t mask_rank(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Cross-workgroup hand-off without cache maintenance: data goes through coherent (sc1)
// loads and stores, the writer waits for its stores to complete before the flag atomic
// that publishes them, and the reader issues its loads only after that atomic returned.
// An acquire/release atomic at agent scope would instead write back and invalidate the
// whole L2 at every use.
template <typename T>
__device__ inline void store_coherent(T* p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ inline T load_coherent(const T* p) {
  return __hip_atomic_load(const_cast<T*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Every earlier memory operation of this wave has completed; later ones stay after it.
__device__ inline void wait_all_memory() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_s_waitcnt(0);
}
// Publish: counts completed stores; returns the previous count.
__device__ inline uint32_t publish_count(uint32_t* flag) {
  wait_all_memory();
  uint32_t old = __hip_atomic_fetch_add(flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  return old;
}

// Per-step values every kernel of a tick reads (Sim uniform, wgsl.js:13-22).
struct StepParams {
  float alpha;
  float decay;
  uint32_t step;
  uint32_t seed;
};

// Records HIP events around phases of a tick on one stream (off when null).
struct PhaseTimer {
  virtual void begin(int phase) = 0;
  virtual void end(int phase) = 0;
  virtual ~PhaseTimer() {}
};

struct PhaseScope {
  PhaseTimer* t;
  int phase;
  PhaseScope(PhaseTimer* t_, int p) : t(t_), phase(p) {
    if (t) t->begin(phase);
  }
  ~PhaseScope() {
    if (t) t->end(phase);
  }
};

}  // namespace fg
