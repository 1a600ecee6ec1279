// sort.hip -- onesweep LSD radix sort of (u32 key, u32 value) pairs for gfx950.
//
//   k_os_clear   clears the histograms and the look-back state
//   k_os_hist    one read of the keys (and values, for a value pass): LDS
//                digit histograms of every pass, added into striped global
//                copies
//   k_os_pass    one launch per 8-bit digit: each 4096-key tile is ranked by
//                its four waves independently (a wave owns 1024 consecutive
//                keys and a private LDS histogram; ballot match-any over the
//                8 digit bits + popcount gives the rank among equal digits),
//                the tile publishes its digit counts, resolves its global
//                offsets by decoupled look-back over the preceding tiles
//                (32 predecessors polled per step), stages the tile in LDS
//                in digit order and writes it out coalesced
//
// The sort is stable, so with keys complemented (descending) and values in
// ascending index order it yields the reference's back-to-front order: key
// descending, ties by ascending index (CpuSortBackend::sort_values_by_keys,
// sort/src/lib.rs:36-61, radix.rs:21-58). An optional last pass orders the
// pairs by a byte of the value (ascending), which groups a batch of views by
// view id while keeping each view's depth order (a segmented sort).
//
// Forward progress of the look-back: tiles take their index from a per-pass
// atomic counter when they start, so every tile a workgroup waits on has
// already started and finishes without waiting on later tiles.

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
constexpr uint32_t WAVE_ITEMS = TILE / WAVES;
constexpr uint32_t MAX_PASSES = 5;
constexpr uint32_t HIST_BLOCKS = 512;
constexpr uint32_t STRIPES = 8;
constexpr uint32_t LOOKBACK = 8;  // predecessors polled per look-back step
constexpr uint32_t FLAG_AGG = 1u << 30;  // tile count published
constexpr uint32_t FLAG_INC = 2u << 30;  // inclusive prefix published
constexpr uint32_t VALUE_MASK = (1u << 30) - 1;

// Digit source of each pass: byte `shift/8` of (key ^ flip), or of the value.
struct PassSet {
  uint32_t count;
  uint32_t shift[MAX_PASSES];
  uint32_t flip[MAX_PASSES];
  uint32_t from_value[MAX_PASSES];
};

__device__ __forceinline__ uint32_t digit(uint32_t key, uint32_t val, uint32_t shift, uint32_t flip,
                                          uint32_t from_value) {
  return (((from_value ? val : key) ^ flip) >> shift) & 0xFFu;
}

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

__global__ __launch_bounds__(T) void k_os_clear(uint32_t *__restrict__ p, size_t n) {
  for (size_t i = (size_t)blockIdx.x * T + threadIdx.x; i < n; i += (size_t)gridDim.x * T) p[i] = 0u;
}

// Workgroups take tiles of 4096 keys (16 per thread, all loaded before any
// is counted) and count every pass's digits with LDS atomics; each
// workgroup then adds its non-zero bins into one of STRIPES global copies
// (copy = workgroup mod STRIPES) to spread the global atomics.
__global__ __launch_bounds__(T) void k_os_hist(const uint32_t *__restrict__ keys, const uint32_t *__restrict__ vals,
                                               const uint32_t *__restrict__ n_dev, uint32_t n_cap, PassSet ps,
                                               uint32_t *__restrict__ hist) {
  __shared__ uint32_t h[MAX_PASSES][256];
  for (uint32_t p = 0; p < MAX_PASSES; ++p) h[p][threadIdx.x] = 0;
  __syncthreads();
  bool need_vals = false;
  for (uint32_t p = 0; p < ps.count; ++p) need_vals = need_vals || ps.from_value[p];
  const uint32_t n = load_count(n_dev, n_cap);
  for (size_t base = (size_t)blockIdx.x * TILE; base < n; base += (size_t)gridDim.x * TILE) {
    uint32_t k[IPT], v[IPT];
#pragma unroll
    for (uint32_t r = 0; r < IPT; ++r) {
      const size_t i = base + (size_t)r * T + threadIdx.x;
      k[r] = i < n ? keys[i] : 0u;
      v[r] = (need_vals && i < n) ? vals[i] : 0u;
    }
#pragma unroll
    for (uint32_t r = 0; r < IPT; ++r) {
      if (base + (size_t)r * T + threadIdx.x < n)
        for (uint32_t p = 0; p < ps.count; ++p)
          atomicAdd(&h[p][digit(k[r], v[r], ps.shift[p], ps.flip[p], ps.from_value[p])], 1u);
    }
  }
  __syncthreads();
  uint32_t *dst = hist + (size_t)(blockIdx.x % STRIPES) * MAX_PASSES * 256;
  for (uint32_t p = 0; p < ps.count; ++p) {
    const uint32_t c = h[p][threadIdx.x];
    if (c) atomicAdd(&dst[p * 256 + threadIdx.x], c);
  }
}

__global__ __launch_bounds__(T) void k_os_pass(const uint32_t *__restrict__ kin, const uint32_t *__restrict__ vin,
                                               uint32_t *__restrict__ kout, uint32_t *__restrict__ vout,
                                               const uint32_t *__restrict__ n_dev, uint32_t n_cap, uint32_t pass,
                                               uint32_t shift, uint32_t flip, uint32_t from_value,
                                               const uint32_t *__restrict__ hist, uint32_t *__restrict__ counter,
                                               uint32_t *__restrict__ status) {
  __shared__ uint32_t s_tile, s_trivial;
  __shared__ uint32_t lds4[WAVES];
  __shared__ uint32_t gofs[256];    // global start of each digit
  __shared__ uint32_t lstart[256];  // start of each digit inside this tile
  __shared__ uint32_t prefix[256];  // count of each digit in earlier tiles
  __shared__ uint32_t whist[WAVES][256];
  __shared__ uint32_t skeys[TILE];
  __shared__ uint32_t svals[TILE];

  const uint32_t t = threadIdx.x, w = t / 64, lane = __lane_id();
  const uint32_t n = load_count(n_dev, n_cap);
  if (t == 0) {
    s_tile = atomicAdd(counter, 1u);
    s_trivial = 0;
  }
#pragma unroll
  for (uint32_t q = 0; q < WAVES; ++q) whist[q][t] = 0;
  __syncthreads();
  const uint32_t tile = s_tile;
  const size_t base = (size_t)tile * TILE;
  if (base >= n) return;  // uniform: no later tile has work either
  const uint32_t count = (uint32_t)min((size_t)TILE, (size_t)n - base);

  uint32_t hd = 0;
#pragma unroll
  for (uint32_t c = 0; c < STRIPES; ++c) hd += hist[(c * MAX_PASSES + pass) * 256 + t];
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

  // Wave-private ranking: wave w owns tile items [w*1024, (w+1)*1024) in
  // order, 64 per round; within a round the lanes of equal digit are ranked
  // by lane, and the running per-digit count lives in whist[w].
  uint32_t key[IPT], val[IPT], rank[IPT];
  uint32_t *wh = whist[w];
#pragma unroll
  for (uint32_t r = 0; r < IPT; ++r) {
    const uint32_t j = w * WAVE_ITEMS + r * 64 + lane;
    const bool active = j < count;
    key[r] = active ? kin[base + j] : 0u;
    val[r] = active ? vin[base + j] : 0u;
    const uint32_t d = digit(key[r], val[r], shift, flip, from_value);
    const uint64_t peers = match_digit(d, active);
    const uint32_t lrank = (uint32_t)__popcll(peers & lanemask_lt());
    const uint32_t before = active ? wh[d] : 0u;
    rank[r] = active ? before + lrank : 0xFFFFFFFFu;
    if (active && lrank == 0) wh[d] = before + (uint32_t)__popcll(peers);
  }
  __syncthreads();
  // Thread t owns digit t: turn the four wave counts into wave offsets.
  uint32_t run = 0;
#pragma unroll
  for (uint32_t q = 0; q < WAVES; ++q) {
    const uint32_t c = whist[q][t];
    whist[q][t] = run;
    run += c;
  }

  // Publish this tile's count of digit t, then look back for the prefix.
  uint32_t *st = status + (size_t)tile * 256 + t;
  if (tile == 0) {
    __hip_atomic_store(st, FLAG_INC | run, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    prefix[t] = 0;
  } else {
    __hip_atomic_store(st, FLAG_AGG | run, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t excl = 0;
    int64_t j = (int64_t)tile - 1;  // next predecessor not yet consumed
    while (j >= 0) {
      uint32_t sv[LOOKBACK];
#pragma unroll
      for (uint32_t k = 0; k < LOOKBACK; ++k)
        sv[k] = (j - (int64_t)k >= 0)
                    ? __hip_atomic_load(status + (size_t)(j - k) * 256 + t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                    : FLAG_INC;
      // Consume predecessors nearest first until an inclusive prefix
      // (done) or an unpublished tile (poll again from there).
      int64_t consumed = 0;
      bool finished = false, stop = false;
#pragma unroll
      for (uint32_t k = 0; k < LOOKBACK; ++k) {
        const uint32_t flag = sv[k] & ~VALUE_MASK;
        const bool take = !stop && flag != 0;
        excl += take ? (sv[k] & VALUE_MASK) : 0u;
        consumed += take ? 1 : 0;
        finished = finished || (take && flag == FLAG_INC);
        stop = stop || !take || flag == FLAG_INC;
      }
      if (finished) break;
      j -= consumed;
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
      const uint32_t d = digit(key[r], val[r], shift, flip, from_value);
      const uint32_t pos = lstart[d] + whist[w][d] + rank[r];
      skeys[pos] = key[r];
      svals[pos] = val[r];
    }
  }
  __syncthreads();
  for (uint32_t j = t; j < count; j += T) {
    const uint32_t k = skeys[j], v = svals[j];
    const uint32_t d = digit(k, v, shift, flip, from_value);
    const size_t pos = (size_t)gofs[d] + prefix[d] + (j - lstart[d]);
    kout[pos] = k;
    vout[pos] = v;
  }
}

uint32_t div_up(size_t a, size_t b) { return (uint32_t)((a + b - 1) / b); }

uint32_t hist_blocks(uint32_t n_cap) { return std::max<uint32_t>(std::min<uint32_t>(div_up(n_cap, TILE), HIST_BLOCKS), 1); }

}  // namespace

// Layout: hist [STRIPES][MAX_PASSES][256] | counters [MAX_PASSES] |
// status [passes][tiles][256], all cleared at the start of every sort.
size_t onesweep_status_words(uint32_t n_cap, uint32_t bits) {
  const uint32_t passes = std::min<uint32_t>(std::max<uint32_t>((bits + 7) / 8, 1), 4) + 1;
  const size_t tiles = std::max<uint32_t>(div_up(n_cap, TILE), 1);
  return (size_t)STRIPES * MAX_PASSES * 256 + MAX_PASSES + (size_t)passes * tiles * 256;
}

void onesweep_sort(uint32_t **keys, uint32_t **vals, uint32_t **keys_alt, uint32_t **vals_alt, uint32_t n_cap,
                   const uint32_t *n_dev, uint32_t bits, bool descending, uint32_t *scratch, size_t scratch_words,
                   hipStream_t st, int value_byte) {
  if (n_cap <= 1 || (bits == 0 && value_byte < 0)) return;
  if (n_cap > (1u << 30)) throw Error(Code::Unsupported, "sort length exceeds 2^30");
  PassSet ps{};
  const uint32_t key_passes = std::min<uint32_t>((bits + 7) / 8, 4);
  for (uint32_t p = 0; p < key_passes; ++p) {
    ps.shift[ps.count] = 8 * p;
    ps.flip[ps.count] = descending ? 0xFFFFFFFFu : 0u;
    ps.from_value[ps.count] = 0;
    ++ps.count;
  }
  if (value_byte >= 0) {
    ps.shift[ps.count] = 8 * (uint32_t)value_byte;
    ps.flip[ps.count] = 0;
    ps.from_value[ps.count] = 1;
    ++ps.count;
  }
  const uint32_t tiles = div_up(n_cap, TILE);
  const size_t words = (size_t)STRIPES * MAX_PASSES * 256 + MAX_PASSES + (size_t)ps.count * tiles * 256;
  if (words > scratch_words) throw Error(Code::Internal, "onesweep scratch too small");
  uint32_t *hist = scratch;
  uint32_t *counters = hist + (size_t)STRIPES * MAX_PASSES * 256;
  uint32_t *status = counters + MAX_PASSES;
  k_os_clear<<<std::min<uint32_t>(div_up(words, T), 1024), T, 0, st>>>(scratch, words);
  k_os_hist<<<hist_blocks(n_cap), T, 0, st>>>(*keys, *vals, n_dev, n_cap, ps, hist);
  for (uint32_t p = 0; p < ps.count; ++p) {
    k_os_pass<<<tiles, T, 0, st>>>(*keys, *vals, *keys_alt, *vals_alt, n_dev, n_cap, p, ps.shift[p], ps.flip[p],
                                   ps.from_value[p], hist, counters + p, status + (size_t)p * tiles * 256);
    std::swap(*keys, *keys_alt);
    std::swap(*vals, *vals_alt);
  }
  GS_HIP(hipGetLastError());
}

}  // namespace gs
