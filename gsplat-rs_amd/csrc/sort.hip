// sort.hip -- onesweep LSD radix sort of (u32 key, u32 value) pairs for gfx950.
//
// One kernel builds the digit histograms of every pass in a single read of
// the keys; then one kernel per 8-bit digit pass ranks each 4096-key tile
// with wave64 ballots (match-any over the 8 digit bits + popcount), publishes
// the tile's digit counts, resolves its global digit offsets by decoupled
// look-back over the preceding tiles, stages the tile in LDS in digit order
// and writes it out coalesced. The sort is stable, so with keys complemented
// (descending) and values in ascending index order it yields the reference's
// back-to-front order: key descending, ties by ascending index
// (CpuSortBackend::sort_values_by_keys, sort/src/lib.rs:36-61, radix.rs:21-58).
//
// Forward progress of the look-back: tiles take their index from a per-pass
// atomic counter when they start, so every tile a block waits on has already
// started and finishes without waiting on later tiles.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "common.h"
#include "sort.h"

namespace gs {

#define GS_HIP(call)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      throw Error(Code::Internal, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

namespace {

constexpr uint32_t T = kSortThreads;
constexpr uint32_t IPT = kSortItems;
constexpr uint32_t TILE = kSortTile;
constexpr uint32_t WAVES = T / 64;
constexpr uint32_t FLAG_AGG = 1u << 30;  // tile count published
constexpr uint32_t FLAG_INC = 2u << 30;  // inclusive prefix published
constexpr uint32_t VALUE_MASK = (1u << 30) - 1;

__device__ __forceinline__ uint64_t lanemask_lt() {
  const uint32_t lane = __lane_id();
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

__device__ __forceinline__ uint64_t match_digit(uint32_t d, bool active) {
  uint64_t m = __ballot(active);
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const bool bit = (d >> b) & 1u;
    const uint64_t bb = __ballot(active && bit);
    m &= bit ? bb : ~bb;
  }
  return m;
}

__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v) {
  const uint32_t lane = __lane_id();
#pragma unroll
  for (uint32_t d = 1; d < 64; d <<= 1) {
    const uint32_t t = __shfl_up(v, d, 64);
    if (lane >= d) v += t;
  }
  return v;
}

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t *lds4) {
  const uint32_t inc = wave_inclusive_scan(v);
  const uint32_t w = threadIdx.x / 64;
  if (__lane_id() == 63) lds4[w] = inc;
  __syncthreads();
  uint32_t off = 0;
#pragma unroll
  for (uint32_t i = 0; i < WAVES; ++i) off += (i < w) ? lds4[i] : 0u;
  __syncthreads();
  return off + inc - v;
}

__device__ __forceinline__ uint32_t load_count(const uint32_t *n_dev, uint32_t n_cap) {
  return n_dev ? min(*n_dev, n_cap) : n_cap;
}

__global__ __launch_bounds__(T) void k_os_zero(uint32_t *__restrict__ p, size_t n) {
  const size_t i = (size_t)blockIdx.x * T + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// Histograms of all passes in one read: LDS bins per pass, ballot-aggregated
// LDS atomics, one global atomic per (block, pass, digit).
__global__ __launch_bounds__(T) void k_os_hist(const uint32_t *__restrict__ keys, const uint32_t *__restrict__ n_dev,
                                               uint32_t n_cap, uint32_t passes, uint32_t flip,
                                               uint32_t *__restrict__ hist) {
  __shared__ uint32_t h[4][256];
  for (uint32_t p = 0; p < 4; ++p) h[p][threadIdx.x] = 0;
  __syncthreads();
  const uint32_t n = load_count(n_dev, n_cap);
  const size_t stride = (size_t)gridDim.x * T;
  for (size_t i = (size_t)blockIdx.x * T + threadIdx.x; i - threadIdx.x < n; i += stride) {
    const bool active = i < n;
    const uint32_t k = active ? (keys[i] ^ flip) : 0u;
    for (uint32_t p = 0; p < passes; ++p) {
      const uint32_t d = (k >> (8 * p)) & 0xFFu;
      const uint64_t peers = match_digit(d, active);
      if (active && (peers & lanemask_lt()) == 0) atomicAdd(&h[p][d], (uint32_t)__popcll(peers));
    }
  }
  __syncthreads();
  for (uint32_t p = 0; p < passes; ++p) {
    const uint32_t c = h[p][threadIdx.x];
    if (c) atomicAdd(&hist[p * 256 + threadIdx.x], c);
  }
}

__global__ __launch_bounds__(T) void k_os_pass(const uint32_t *__restrict__ kin, const uint32_t *__restrict__ vin,
                                               uint32_t *__restrict__ kout, uint32_t *__restrict__ vout,
                                               const uint32_t *__restrict__ n_dev, uint32_t n_cap, uint32_t pass,
                                               uint32_t flip, const uint32_t *__restrict__ hist,
                                               uint32_t *__restrict__ counter, uint32_t *__restrict__ status) {
  __shared__ uint32_t s_tile, s_trivial;
  __shared__ uint32_t lds4[WAVES];
  __shared__ uint32_t gofs[256];    // global start of each digit
  __shared__ uint32_t lstart[256];  // start of each digit inside this tile
  __shared__ uint32_t prefix[256];  // count of each digit in earlier tiles
  __shared__ uint32_t wcnt[WAVES][256];
  __shared__ uint32_t skeys[TILE];
  __shared__ uint32_t svals[TILE];

  const uint32_t t = threadIdx.x, w = t / 64;
  const uint32_t n = load_count(n_dev, n_cap);
  if (t == 0) {
    s_tile = atomicAdd(counter, 1u);
    s_trivial = 0;
  }
  __syncthreads();
  const uint32_t tile = s_tile;
  const size_t base = (size_t)tile * TILE;
  if (base >= n) return;  // uniform: no later tile has work either
  const uint32_t count = (uint32_t)min((size_t)TILE, (size_t)n - base);
  const uint32_t shift = 8 * pass;

  const uint32_t hd = hist[pass * 256 + t];
  if (hd == n) s_trivial = 1;  // every key has this digit: the pass is the identity
  gofs[t] = block_exclusive_scan(hd, lds4);
  __syncthreads();
  if (s_trivial) {
    for (uint32_t j = t; j < count; j += T) {
      kout[base + j] = kin[base + j];
      vout[base + j] = vin[base + j];
    }
    return;
  }

  // Rank: item order is (round, wave, lane) = global position order.
  uint32_t key[IPT], val[IPT], rank[IPT];
  uint32_t run = 0;  // running count of digit t in this tile
#pragma unroll
  for (uint32_t r = 0; r < IPT; ++r) {
#pragma unroll
    for (uint32_t q = 0; q < WAVES; ++q) wcnt[q][t] = 0;
    __syncthreads();
    const uint32_t j = r * T + t;
    const bool active = j < count;
    key[r] = active ? kin[base + j] : 0u;
    val[r] = active ? vin[base + j] : 0u;
    const uint32_t d = ((key[r] ^ flip) >> shift) & 0xFFu;
    const uint64_t peers = match_digit(d, active);
    const uint32_t lrank = (uint32_t)__popcll(peers & lanemask_lt());
    if (active && lrank == 0) wcnt[w][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    uint32_t s = run;
#pragma unroll
    for (uint32_t q = 0; q < WAVES; ++q) {
      const uint32_t c = wcnt[q][t];
      wcnt[q][t] = s;
      s += c;
    }
    run = s;
    __syncthreads();
    rank[r] = active ? wcnt[w][d] + lrank : 0xFFFFFFFFu;
    __syncthreads();  // wcnt is cleared again by the next round
  }

  // Publish this tile's count of digit t, then look back for the prefix.
  uint32_t *st = status + (size_t)tile * 256 + t;
  if (tile == 0) {
    __hip_atomic_store(st, FLAG_INC | run, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    prefix[t] = 0;
  } else {
    __hip_atomic_store(st, FLAG_AGG | run, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t excl = 0;
    int64_t j = (int64_t)tile - 1;
    while (j >= 0) {
      const uint32_t sv =
          __hip_atomic_load(status + (size_t)j * 256 + t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const uint32_t flag = sv & ~VALUE_MASK;
      if (flag == 0) continue;  // predecessor has not published yet
      excl += sv & VALUE_MASK;
      if (flag == FLAG_INC) break;
      --j;
    }
    __hip_atomic_store(st, FLAG_INC | (excl + run), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    prefix[t] = excl;
  }
  lstart[t] = block_exclusive_scan(run, lds4);
  __syncthreads();

  // Stage in digit order, then write coalesced.
#pragma unroll
  for (uint32_t r = 0; r < IPT; ++r) {
    if (rank[r] != 0xFFFFFFFFu) {
      const uint32_t d = ((key[r] ^ flip) >> shift) & 0xFFu;
      const uint32_t pos = lstart[d] + rank[r];
      skeys[pos] = key[r];
      svals[pos] = val[r];
    }
  }
  __syncthreads();
  for (uint32_t j = t; j < count; j += T) {
    const uint32_t k = skeys[j];
    const uint32_t d = ((k ^ flip) >> shift) & 0xFFu;
    const size_t pos = (size_t)gofs[d] + prefix[d] + (j - lstart[d]);
    kout[pos] = k;
    vout[pos] = svals[j];
  }
}

uint32_t div_up(size_t a, size_t b) { return (uint32_t)((a + b - 1) / b); }

}  // namespace

size_t onesweep_status_words(uint32_t n_cap, uint32_t bits) {
  const uint32_t passes = std::max<uint32_t>((bits + 7) / 8, 1);
  return 256 * 4 + 4 + (size_t)passes * std::max<uint32_t>(div_up(n_cap, TILE), 1) * 256;
}

void onesweep_sort(uint32_t **keys, uint32_t **vals, uint32_t **keys_alt, uint32_t **vals_alt, uint32_t n_cap,
                   const uint32_t *n_dev, uint32_t bits, bool descending, uint32_t *scratch, size_t scratch_words,
                   hipStream_t st) {
  if (n_cap <= 1 || bits == 0) return;
  if (n_cap > (1u << 30)) throw Error(Code::Unsupported, "sort length exceeds 2^30");
  const uint32_t passes = std::min<uint32_t>((bits + 7) / 8, 4);
  const uint32_t tiles = div_up(n_cap, TILE);
  const size_t need = onesweep_status_words(n_cap, bits);
  if (need > scratch_words) throw Error(Code::Internal, "onesweep scratch too small");
  uint32_t *hist = scratch;             // [4][256]
  uint32_t *counters = scratch + 1024;  // [4]
  uint32_t *status = scratch + 1028;    // [passes][tiles][256]
  const uint32_t flip = descending ? 0xFFFFFFFFu : 0u;
  k_os_zero<<<div_up(need, T), T, 0, st>>>(scratch, need);
  const uint32_t hist_blocks = std::min<uint32_t>(tiles, 1024);
  k_os_hist<<<hist_blocks, T, 0, st>>>(*keys, n_dev, n_cap, passes, flip, hist);
  for (uint32_t p = 0; p < passes; ++p) {
    k_os_pass<<<tiles, T, 0, st>>>(*keys, *vals, *keys_alt, *vals_alt, n_dev, n_cap, p, flip, hist, counters + p,
                                   status + (size_t)p * tiles * 256);
    std::swap(*keys, *keys_alt);
    std::swap(*vals, *vals_alt);
  }
  GS_HIP(hipGetLastError());
}

}  // namespace gs
