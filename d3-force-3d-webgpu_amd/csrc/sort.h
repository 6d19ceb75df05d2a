// sort.h -- device radix sort of (key, u32 value) pairs and an exclusive scan (see sort.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fg {

// Scratch of one sorter: the alternate key/value buffers, digit histograms, tile tickets,
// look-back words and the epoch counter that keeps stale look-back words from matching.
struct SortWork {
  void* keys_alt = nullptr;
  uint32_t* vals_alt = nullptr;
  uint32_t* ghist = nullptr;
  uint64_t* lookback = nullptr;
  uint32_t* tickets = nullptr;
  uint32_t* epoch = nullptr;
  uint32_t* errors = nullptr;  // [0]: look-back waits that hit their bound
  uint32_t cap_n = 0;
  size_t cap_key_bytes = 0;
  uint64_t lookback_words = 0;
  // false: the pointers are lent by the caller, sized for cap_n keys of cap_key_bytes and
  // lookback_words words; reserve() then only checks a sort fits and release() frees nothing
  bool owned = true;

  void reserve(uint32_t n, size_t key_bytes);
  void release();
  ~SortWork() { release(); }
};

// Look-back words that a sort of n u32 keys needs at the widest radix (11 bits): one per
// (tile of SORT_U32_TILE pairs, digit).
constexpr uint32_t SORT_U32_TILE = 3072;
constexpr uint64_t sort_u32_lookback_words(uint32_t n) {
  return (uint64_t)((n + SORT_U32_TILE - 1) / SORT_U32_TILE) * 2048u;
}

// Sorts bits [first_bit, first_bit + bits) of the keys, 8 or 11 bits per pass.  The sorted
// pairs end up in (keys, vals) or in the work's alternate buffers: *out_* says which.
// prepared: the caller's own kernel has already done the sort's clear (sort_clear_block
// with sort_prepare's arguments), so the sort launches one kernel fewer.
void radix_sort_u32(hipStream_t stream, SortWork& w, int radix_bits, uint32_t* keys,
                    uint32_t* vals, uint32_t n, uint32_t first_bit, uint32_t bits,
                    uint32_t** out_keys, uint32_t** out_vals);
void radix_sort_u64(hipStream_t stream, SortWork& w, int radix_bits, uint64_t* keys,
                    uint32_t* vals, uint32_t n, uint32_t first_bit, uint32_t bits,
                    uint64_t** out_keys, uint32_t** out_vals, bool prepared = false);

// What a sort clears before its passes: the digit histograms, the tile tickets, and the
// look-back epoch to bump.
struct SortClear {
  uint32_t* ghist;
  uint32_t hist_words;
  uint32_t* tickets;
  uint32_t* epoch;
};
// Reserves the work for a sort of n u64 keys over `bits` bits and returns its clear.
SortClear sort_prepare(SortWork& w, int radix_bits, uint32_t n, uint32_t bits);
// The clear, by one workgroup of any size (every thread calls it).
__device__ inline void sort_clear_block(const SortClear& c) {
  for (uint32_t i = threadIdx.x; i < c.hist_words; i += blockDim.x) c.ghist[i] = 0;
  if (threadIdx.x < 16) c.tickets[threadIdx.x] = 0;
  if (threadIdx.x == 0) c.epoch[1] = (c.epoch[0] += 16u);
}

// In-place exclusive prefix sum of n u32 values; `partial` holds scan_partials_size(n)
// words of scratch.
uint32_t scan_partials_size(uint32_t n);
// (skip: a launch-time condition on the device under which the scan does nothing)
void exclusive_scan_u32(hipStream_t stream, uint32_t* data, uint32_t n, uint32_t* partial, SkipIf skip = {});

}  // namespace fg
