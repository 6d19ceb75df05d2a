// sort.hip -- stable LSD radix sort of (key, u32 value) pairs, wave64 / CDNA4.
//
// Same contract as the reference sorter (src/gpu/sort.js:1-20, 223-273): ascending on a
// chosen bit range of the key, equal keys in input order.  Built differently:
//
//   sort_clear      zeroes the digit histograms and tile tickets, bumps the look-back epoch
//   sort_histogram  one read of the keys: LDS histograms of every pass's digit at once
//   sort_pass       one launch per digit ("onesweep"): a tile of TILE pairs is ranked with
//                   wave64 ballots (stable within the wave, then across waves through LDS
//                   per-wave digit counters), its digit totals are chained to the tiles
//                   before it by decoupled look-back on 64-bit (epoch, status, count)
//                   words, and the pairs are staged in LDS in digit order so that the
//                   global writes go out as coalesced runs.
//
// Tiles take their index from an atomic ticket, so a tile only ever waits for tiles that
// started before it: every wait ends.  The waits are bounded anyway; a bound that trips
// is counted in `errors` (and the caller reports it) instead of hanging the device.
#include "common.h"
#include "sort.h"

namespace fg {

namespace {

constexpr int THREADS = 256;
constexpr int WAVES = THREADS / WAVE;
constexpr uint32_t ST_AGG = 1u, ST_INC = 2u;
constexpr uint32_t SPIN_LIMIT = 1u << 26;

template <typename K>
__device__ inline uint32_t digit_of(K key, uint32_t shift, uint32_t mask) {
  return (uint32_t)(key >> shift) & mask;
}

// Exclusive scan over the block of one value per thread; returns the block total too.
__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t* total, uint32_t* s_tmp) {
  const uint32_t lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x / WAVE;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < WAVE; o <<= 1) {
    uint32_t y = __shfl_up(x, o, WAVE);
    if (lane >= (uint32_t)o) x += y;
  }
  if (lane == WAVE - 1) s_tmp[wave] = x;
  __syncthreads();
  uint32_t before = 0, all = 0;
#pragma unroll
  for (int w = 0; w < WAVES; ++w) {
    uint32_t t = s_tmp[w];
    if (w < (int)wave) before += t;
    all += t;
  }
  __syncthreads();
  *total = all;
  return before + x - v;
}

__global__ void sort_clear(SortClear c) { sort_clear_block(c); }

template <typename K, int RB>
__global__ void __launch_bounds__(THREADS) sort_histogram(const K* __restrict__ keys, uint32_t n,
                                                          uint32_t first_bit, uint32_t bits,
                                                          uint32_t passes,
                                                          uint32_t* __restrict__ ghist) {
  constexpr uint32_t R = 1u << RB;
  constexpr int MAXP = RB >= 10 ? 6 : 8;
  __shared__ uint32_t h[MAXP * R];
  for (uint32_t i = threadIdx.x; i < passes * R; i += THREADS) h[i] = 0;
  __syncthreads();
  for (uint32_t i = blockIdx.x * THREADS + threadIdx.x; i < n; i += gridDim.x * THREADS) {
    K key = keys[i];
    for (uint32_t p = 0; p < passes; ++p) {
      uint32_t shift = first_bit + p * RB;
      uint32_t pb = bits - p * RB < (uint32_t)RB ? bits - p * RB : (uint32_t)RB;
      atomicAdd(&h[p * R + digit_of(key, shift, (1u << pb) - 1u)], 1u);
    }
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < passes * R; i += THREADS) {
    uint32_t c = h[i];
    if (c) atomicAdd(&ghist[i], c);
  }
}

template <typename K, int RB, int ITEMS>
__global__ void __launch_bounds__(THREADS)
    sort_pass(const K* __restrict__ kin, const uint32_t* __restrict__ vin, K* __restrict__ kout,
              uint32_t* __restrict__ vout, uint32_t n, uint32_t shift, uint32_t pbits,
              const uint32_t* __restrict__ ghist, uint64_t* __restrict__ lookback,
              uint32_t* __restrict__ ticket, const uint32_t* __restrict__ epoch_ptr,
              uint32_t pass, uint32_t* __restrict__ errors) {
  constexpr uint32_t R = 1u << RB;
  constexpr uint32_t TILE = THREADS * ITEMS;
  constexpr uint32_t PER = R / THREADS;  // digits per thread
  static_assert(R % THREADS == 0, "radix must be a multiple of the block");

  __shared__ uint32_t s_whist[WAVES][R];   // per-wave digit counts, then wave offsets
  __shared__ uint32_t s_start[R];          // global start of each digit for this tile
  __shared__ uint32_t s_tbase[R];          // start of each digit inside the staged tile
  __shared__ K s_keys[TILE];
  __shared__ uint32_t s_vals[TILE];
  __shared__ uint32_t s_tmp[WAVES];
  __shared__ uint32_t s_tile;

  const uint32_t tid = threadIdx.x, lane = tid & (WAVE - 1), wave = tid / WAVE;
  const uint32_t mask = (1u << pbits) - 1u;
  if (tid == 0) s_tile = atomicAdd(ticket, 1u);
  for (uint32_t i = tid; i < WAVES * R; i += THREADS) (&s_whist[0][0])[i] = 0;
  __syncthreads();
  const uint32_t tile = s_tile;
  const uint32_t epoch = epoch_ptr[1] + pass;

  // load: wave w owns [tile*TILE + w*ITEMS*64, +ITEMS*64), striped over its lanes
  const uint32_t base = tile * TILE + wave * (ITEMS * WAVE);
  K key[ITEMS];
  uint32_t val[ITEMS], dig[ITEMS], rank[ITEMS];
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    uint32_t i = base + k * WAVE + lane;
    bool ok = i < n;
    key[k] = ok ? kin[i] : (K)0;
    val[k] = ok ? vin[i] : 0u;
    dig[k] = ok ? digit_of(key[k], shift, mask) : R;
  }
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    uint32_t d = dig[k];
    uint64_t match = __ballot(d < R);
#pragma unroll
    for (int b = 0; b < RB; ++b) {
      uint64_t bal = __ballot((d >> b) & 1u);
      match &= ((d >> b) & 1u) ? bal : ~bal;
    }
    uint32_t r = __popcll(match & lt);
    uint32_t before = d < R ? s_whist[wave][d] : 0u;
    rank[k] = before + r;
    bool leader = d < R && r == 0;
    // every lane of the wave has read its counter before the leaders write (LDS ops of a
    // wave complete in order)
    if (leader) s_whist[wave][d] = before + (uint32_t)__popcll(match);
  }
  __syncthreads();

  // per digit: offsets of the waves inside the tile, the tile total, and its look-back
  uint32_t tot[PER];
#pragma unroll
  for (uint32_t q = 0; q < PER; ++q) {
    uint32_t d = tid * PER + q;
    uint32_t run = 0;
#pragma unroll
    for (int w = 0; w < WAVES; ++w) {
      uint32_t c = s_whist[w][d];
      s_whist[w][d] = run;
      run += c;
    }
    tot[q] = run;
  }
  // publish this tile's totals (inclusive right away for the first tile)
#pragma unroll
  for (uint32_t q = 0; q < PER; ++q) {
    uint32_t d = tid * PER + q;
    uint64_t word = ((uint64_t)epoch << 32) | ((uint64_t)(tile == 0 ? ST_INC : ST_AGG) << 30) | tot[q];
    __hip_atomic_store(&lookback[(uint64_t)tile * R + d], word, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  }
  // global digit starts: exclusive scan of this pass's histogram
  uint32_t gsum = 0, gloc[PER];
#pragma unroll
  for (uint32_t q = 0; q < PER; ++q) {
    gloc[q] = gsum;
    gsum += ghist[tid * PER + q];
  }
  uint32_t gtotal;
  uint32_t gbase = block_exclusive_scan(gsum, &gtotal, s_tmp);
  // staged positions: exclusive scan of the tile totals
  uint32_t tsum = 0, tloc[PER];
#pragma unroll
  for (uint32_t q = 0; q < PER; ++q) {
    tloc[q] = tsum;
    tsum += tot[q];
  }
  uint32_t ttotal;
  uint32_t tbase = block_exclusive_scan(tsum, &ttotal, s_tmp);
#pragma unroll
  for (uint32_t q = 0; q < PER; ++q) {
    uint32_t d = tid * PER + q;
    uint32_t prefix = 0;
    if (tile > 0) {
      int32_t j = (int32_t)tile - 1;
      uint32_t spins = 0;
      while (j >= 0) {
        uint64_t w = __hip_atomic_load(&lookback[(uint64_t)j * R + d], __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
        uint32_t st = (uint32_t)(w >> 30) & 3u;
        if ((uint32_t)(w >> 32) != epoch || st == 0) {
          if (++spins > SPIN_LIMIT) {
            atomicAdd(errors, 1u);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
          continue;
        }
        prefix += (uint32_t)w & 0x3fffffffu;
        if (st == ST_INC) break;
        --j;
      }
      uint64_t word = ((uint64_t)epoch << 32) | ((uint64_t)ST_INC << 30) | (prefix + tot[q]);
      __hip_atomic_store(&lookback[(uint64_t)tile * R + d], word, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    }
    s_start[d] = gbase + gloc[q] + prefix;
    s_tbase[d] = tbase + tloc[q];
  }
  __syncthreads();

  // stage in digit order
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    uint32_t d = dig[k];
    if (d < R) {
      uint32_t at = s_tbase[d] + s_whist[wave][d] + rank[k];
      s_keys[at] = key[k];
      s_vals[at] = val[k];
    }
  }
  __syncthreads();
  // write out: consecutive staged entries of one digit go to consecutive addresses
  for (uint32_t i = tid; i < ttotal; i += THREADS) {
    K kk = s_keys[i];
    uint32_t d = digit_of(kk, shift, mask);
    uint32_t at = s_start[d] + (i - s_tbase[d]);
    kout[at] = kk;
    vout[at] = s_vals[i];
  }
}

template <typename K, int RB>
void sort_launch(hipStream_t stream, SortWork& w, K* keys, uint32_t* vals, uint32_t n,
                 uint32_t first_bit, uint32_t bits, K** out_keys, uint32_t** out_vals,
                 bool prepared = false) {
  constexpr uint32_t R = 1u << RB;
  // u64 keys: 8 per thread, 4 for a short sort (twice the tiles: more waves to hide each
  // wave's latency where the device is mostly idle, against a longer look-back), 16 for a
  // long one (half the tiles, a shorter look-back: 1M keys 0.139 -> 0.129 ms for the tree's
  // four passes; 200k slower with 16, 1M slower with 24)
  constexpr uint32_t SHORT_SORT = 131072, LONG_SORT = 524288;
  constexpr int ITEMS = sizeof(K) == 8 ? 8 : 12;
  static_assert(sizeof(K) == 8 || THREADS * ITEMS == SORT_U32_TILE, "u32 tile out of step with sort.h");
  const bool short_sort = sizeof(K) == 8 && n <= SHORT_SORT;
  const bool long_sort = sizeof(K) == 8 && n > LONG_SORT;
  const uint32_t TILE = THREADS * (short_sort ? 4 : long_sort ? 16 : ITEMS);
  uint32_t passes = (bits + RB - 1) / RB;
  K* a = keys;
  uint32_t* av = vals;
  K* b = (K*)w.keys_alt;
  uint32_t* bv = w.vals_alt;
  if (n <= 1 || passes == 0) {
    *out_keys = keys;
    *out_vals = vals;
    return;
  }
  uint32_t tiles = ceil_div(n, TILE);
  if ((uint64_t)tiles * R > w.lookback_words) throw ArgError(-1, "sort: look-back buffer too small");
  if (!prepared) {
    hipLaunchKernelGGL(sort_clear, dim3(1), dim3(256), 0, stream, SortClear{w.ghist, passes * R, w.tickets, w.epoch});
    FG_LAUNCH_CHECK();
  }
  uint32_t hblocks = tiles < 1024 ? tiles : 1024;
  hipLaunchKernelGGL((sort_histogram<K, RB>), dim3(hblocks), dim3(THREADS), 0, stream,
                     (const K*)keys, n, first_bit, bits, passes, w.ghist);
  FG_LAUNCH_CHECK();
  for (uint32_t p = 0; p < passes; ++p) {
    uint32_t pb = bits - p * RB < (uint32_t)RB ? bits - p * RB : (uint32_t)RB;
    if (short_sort)
      hipLaunchKernelGGL((sort_pass<K, RB, 4>), dim3(tiles), dim3(THREADS), 0, stream,
                         (const K*)a, (const uint32_t*)av, b, bv, n, first_bit + p * RB, pb,
                         (const uint32_t*)(w.ghist + p * R), w.lookback, w.tickets + p,
                         (const uint32_t*)w.epoch, p, w.errors);
    else if (long_sort) {
      if constexpr (sizeof(K) == 8)
        hipLaunchKernelGGL((sort_pass<K, RB, 16>), dim3(tiles), dim3(THREADS), 0, stream,
                           (const K*)a, (const uint32_t*)av, b, bv, n, first_bit + p * RB, pb,
                           (const uint32_t*)(w.ghist + p * R), w.lookback, w.tickets + p,
                           (const uint32_t*)w.epoch, p, w.errors);
    } else
      hipLaunchKernelGGL((sort_pass<K, RB, ITEMS>), dim3(tiles), dim3(THREADS), 0, stream,
                         (const K*)a, (const uint32_t*)av, b, bv, n, first_bit + p * RB, pb,
                         (const uint32_t*)(w.ghist + p * R), w.lookback, w.tickets + p,
                         (const uint32_t*)w.epoch, p, w.errors);
    FG_LAUNCH_CHECK();
    K* t = a;
    a = b;
    b = t;
    uint32_t* tv = av;
    av = bv;
    bv = tv;
  }
  *out_keys = a;
  *out_vals = av;
}

// ---- exclusive scan ------------------------------------------------------------------

constexpr uint32_t SCAN_ITEMS = 8, SCAN_TILE = THREADS * SCAN_ITEMS;

// partial[b] = the sum of tile b.
__global__ void __launch_bounds__(THREADS) scan_tile_sums(const uint32_t* __restrict__ data, uint32_t n,
                                                          uint32_t* __restrict__ partial, SkipIf skip) {
  if (skip()) return;
  __shared__ uint32_t s_tmp[WAVES];
  const uint32_t base = blockIdx.x * SCAN_TILE;
  uint32_t sum = 0;
#pragma unroll
  for (uint32_t k = 0; k < SCAN_ITEMS; ++k) {
    const uint32_t i = base + k * THREADS + threadIdx.x;
    if (i < n) sum += data[i];
  }
  uint32_t total;
  block_exclusive_scan(sum, &total, s_tmp);
  if (threadIdx.x == 0) partial[blockIdx.x] = total;
}

// One block: the tile sums become each tile's exclusive offset.
__global__ void __launch_bounds__(THREADS) scan_offsets(uint32_t* __restrict__ partial, uint32_t tiles,
                                                        SkipIf skip) {
  if (skip()) return;
  __shared__ uint32_t s_tmp[WAVES];
  uint32_t carry = 0;
  for (uint32_t b = 0; b < tiles; b += THREADS) {
    const uint32_t i = b + threadIdx.x;
    const uint32_t v = i < tiles ? partial[i] : 0u;
    uint32_t total;
    const uint32_t ex = block_exclusive_scan(v, &total, s_tmp);
    if (i < tiles) partial[i] = carry + ex;
    carry += total;
  }
}

// Tile b scanned in LDS (each thread takes SCAN_ITEMS consecutive values) plus its offset.
__global__ void __launch_bounds__(THREADS) scan_apply(uint32_t* __restrict__ data, uint32_t n,
                                                      const uint32_t* __restrict__ partial, SkipIf skip) {
  if (skip()) return;
  __shared__ uint32_t s[SCAN_TILE];
  __shared__ uint32_t s_tmp[WAVES];
  const uint32_t base = blockIdx.x * SCAN_TILE;
#pragma unroll
  for (uint32_t k = 0; k < SCAN_ITEMS; ++k) {
    const uint32_t j = k * THREADS + threadIdx.x;
    s[j] = base + j < n ? data[base + j] : 0u;
  }
  __syncthreads();
  uint32_t v[SCAN_ITEMS], sum = 0;
#pragma unroll
  for (uint32_t k = 0; k < SCAN_ITEMS; ++k) {
    v[k] = s[threadIdx.x * SCAN_ITEMS + k];
    sum += v[k];
  }
  uint32_t total;
  uint32_t ex = block_exclusive_scan(sum, &total, s_tmp) + partial[blockIdx.x];
#pragma unroll
  for (uint32_t k = 0; k < SCAN_ITEMS; ++k) {
    s[threadIdx.x * SCAN_ITEMS + k] = ex;
    ex += v[k];
  }
  __syncthreads();
#pragma unroll
  for (uint32_t k = 0; k < SCAN_ITEMS; ++k) {
    const uint32_t j = k * THREADS + threadIdx.x;
    if (base + j < n) data[base + j] = s[j];
  }
}

}  // namespace

uint32_t scan_partials_size(uint32_t n) { return ceil_div(n > 0 ? n : 1, SCAN_TILE); }

void exclusive_scan_u32(hipStream_t stream, uint32_t* data, uint32_t n, uint32_t* partial, SkipIf skip) {
  if (!n) return;
  const uint32_t tiles = scan_partials_size(n);
  hipLaunchKernelGGL(scan_tile_sums, dim3(tiles), dim3(THREADS), 0, stream, (const uint32_t*)data, n, partial, skip);
  FG_LAUNCH_CHECK();
  hipLaunchKernelGGL(scan_offsets, dim3(1), dim3(THREADS), 0, stream, partial, tiles, skip);
  FG_LAUNCH_CHECK();
  hipLaunchKernelGGL(scan_apply, dim3(tiles), dim3(THREADS), 0, stream, data, n, (const uint32_t*)partial, skip);
  FG_LAUNCH_CHECK();
}

void SortWork::reserve(uint32_t n, size_t key_bytes) {
  if (!owned) {
    // lent work: sort_launch checks the look-back words against the tiles it launches
    if (n > cap_n || key_bytes > cap_key_bytes) throw ArgError(-1, "sort: lent work too small");
    return;
  }
  // tiles at the smallest tile (a short u64 sort: 256 * 4) and the widest radix (11 bits)
  uint64_t tiles = ceil_div(n > 0 ? n : 1, THREADS * 4);
  uint64_t need_lb = tiles * 2048;
  size_t need_keys = (size_t)(n > 0 ? n : 1) * key_bytes;
  if (cap_n >= n && cap_key_bytes >= key_bytes && lookback_words >= need_lb) return;
  release();
  FG_HIP(hipMalloc(&keys_alt, need_keys));
  FG_HIP(hipMalloc(&vals_alt, sizeof(uint32_t) * (size_t)(n > 0 ? n : 1)));
  FG_HIP(hipMalloc(&ghist, sizeof(uint32_t) * 8 * 2048));
  FG_HIP(hipMalloc(&lookback, sizeof(uint64_t) * need_lb));
  clear_device(lookback, sizeof(uint64_t) * need_lb);
  FG_HIP(hipMalloc(&tickets, sizeof(uint32_t) * 16));
  FG_HIP(hipMalloc(&epoch, sizeof(uint32_t) * 4));
  clear_device(epoch, sizeof(uint32_t) * 4);
  FG_HIP(hipMalloc(&errors, sizeof(uint32_t) * 4));
  clear_device(errors, sizeof(uint32_t) * 4);
  cap_n = n;
  cap_key_bytes = key_bytes;
  lookback_words = need_lb;
}

void SortWork::release() {
  if (owned) {
    if (keys_alt) (void)hipFree(keys_alt);
    if (vals_alt) (void)hipFree(vals_alt);
    if (ghist) (void)hipFree(ghist);
    if (lookback) (void)hipFree(lookback);
    if (tickets) (void)hipFree(tickets);
    if (epoch) (void)hipFree(epoch);
    if (errors) (void)hipFree(errors);
  }
  keys_alt = nullptr;
  vals_alt = nullptr;
  ghist = tickets = epoch = errors = nullptr;
  lookback = nullptr;
  cap_n = 0;
  cap_key_bytes = 0;
  lookback_words = 0;
}

void radix_sort_u32(hipStream_t stream, SortWork& w, int radix_bits, uint32_t* keys,
                    uint32_t* vals, uint32_t n, uint32_t first_bit, uint32_t bits,
                    uint32_t** out_keys, uint32_t** out_vals) {
  w.reserve(n, sizeof(uint32_t));
  if (radix_bits == 11) sort_launch<uint32_t, 11>(stream, w, keys, vals, n, first_bit, bits, out_keys, out_vals);
  else sort_launch<uint32_t, 8>(stream, w, keys, vals, n, first_bit, bits, out_keys, out_vals);
}

void radix_sort_u64(hipStream_t stream, SortWork& w, int radix_bits, uint64_t* keys,
                    uint32_t* vals, uint32_t n, uint32_t first_bit, uint32_t bits,
                    uint64_t** out_keys, uint32_t** out_vals, bool prepared) {
  w.reserve(n, sizeof(uint64_t));
  if (radix_bits == 11)
    sort_launch<uint64_t, 11>(stream, w, keys, vals, n, first_bit, bits, out_keys, out_vals, prepared);
  else
    sort_launch<uint64_t, 8>(stream, w, keys, vals, n, first_bit, bits, out_keys, out_vals, prepared);
}

SortClear sort_prepare(SortWork& w, int radix_bits, uint32_t n, uint32_t bits) {
  w.reserve(n, sizeof(uint64_t));
  const uint32_t rb = radix_bits == 11 ? 11u : 8u;
  return SortClear{w.ghist, ((bits + rb - 1) / rb) << rb, w.tickets, w.epoch};
}

}  // namespace fg
