// sort.hip -- LSD radix sort of (u32 key, u32 value) pairs for gfx950, 8-bit
// digits (or one 9-bit pass for 9-bit keys), 4096-item tiles of 512 threads x
// 8 items (eight waves per tile share a CU's four SIMDs; thread t keeps the
// books of digit t, for the 256 or 512 digits of a pass).
//
// Every pass ranks a tile the same way: its eight waves work independently (a
// wave owns 512 consecutive items and a private LDS histogram; ballot
// match-any over the digit bits + popcount gives the rank among equal digits
// of a 64-item round, and one returning LDS atomic per digit group the count
// of earlier rounds), then the tile is staged in LDS in digit order and
// written out coalesced. Ranking is VALU work (about five instructions per
// digit bit and round), the rest memory latency. The passes differ in how a
// tile learns where its digits go:
//
// Look-back mode (onesweep), for sorts of fewer tiles:
//   k_os_clear   clears the histograms and the look-back state
//   k_os_hist    one read of the keys (and values, for a value pass): LDS
//                digit histograms of every pass, added into striped global
//                copies
//   k_os_pass    one launch per digit: the tile publishes its digit counts
//                and resolves its prefix over the preceding tiles by
//                decoupled look-back, 8 predecessors polled per step
// Reduce-then-scan mode, for long sorts (the look-back front advances a
// bounded number of tiles per coherent round trip, which caps its rate):
//   k_rts_up     per-tile digit counts, digit-major
//   k_rts_scan   one exclusive scan of all counts = every (digit, tile) start
//                (chained over 4096-count chunks by decoupled look-back)
//   k_rts_down   rank and scatter
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
// already started and finishes without waiting on later tiles. The
// reduce-then-scan scan takes its chunks by ticket the same way; its up- and
// down-sweeps have no inter-workgroup waits.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <string>

#include "common.h"
#include "kernels.h"
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
constexpr uint32_t MAX_DIGIT_BITS = 9;  // 8-bit digits, or one 9-bit pass for 9-bit keys
static_assert(T >= (1u << MAX_DIGIT_BITS) && T % 64 == 0 && TILE % T == 0, "sort workgroup shape");
constexpr uint32_t WAVE_ITEMS = TILE / WAVES;
constexpr uint32_t MAX_PASSES = kSortMaxPasses;
constexpr uint32_t HIST_BLOCKS = 1024;
constexpr uint32_t HIST_IPT = 2;                   // keys per thread of a histogram chunk
constexpr uint32_t HIST_CHUNK = T * HIST_IPT;      // keys per histogram workgroup step
constexpr uint32_t STRIPES = kSortStripes;
constexpr uint32_t LOOKBACK = 8;  // predecessors polled per look-back step
constexpr uint32_t kPassGrid = 768;   // look-back passes: resident 42 KB-LDS workgroups, 3 per CU
constexpr uint32_t kMinGrid = 1024;   // floor of a hinted grid
constexpr uint32_t FLAG_AGG = 1u << 30;  // tile count published
constexpr uint32_t FLAG_INC = 2u << 30;  // inclusive prefix published
constexpr uint32_t VALUE_MASK = (1u << 30) - 1;
constexpr uint32_t XCDS = 8;  // MI355X: workgroup b is dispatched to XCD b % 8
constexpr uint32_t SCAN_IPT = 16;               // counts per thread of the reduce-then-scan scan
constexpr uint32_t SCAN_CHUNK = T * SCAN_IPT;   // counts per scan chunk (one workgroup step)
constexpr uint32_t kScanGrid = 1024;            // bound of the scan grid (workgroups loop over chunks)
constexpr uint32_t kFineBins = 8192;            // bucket sort: fine bins of the key range

using PassSet = RadixPasses;

// from_value of a bucket pass: the digit of a non-zero key is the bucket
// tab[(kmax - key) >> shift] of its fine bin, with kmax passed as `flip`, and
// the last digit for key 0 (see bucket_sort).
constexpr uint32_t kRangeDigit = 2;

// Digits of DB bits: thread t < 2^DB keeps the books of digit t.
template <uint32_t DB>
__device__ __forceinline__ uint32_t digit(uint32_t key, uint32_t val, uint32_t shift, uint32_t flip,
                                          uint32_t from_value, const uint16_t *tab = nullptr) {
  if (from_value == kRangeDigit) return key ? tab[min((flip - key) >> shift, kFineBins - 1)] : (1u << DB) - 1u;
  return (((from_value ? val : key) ^ flip) >> shift) & ((1u << DB) - 1u);
}

__device__ __forceinline__ uint64_t lanemask_lt() {
  const uint32_t lane = __lane_id();
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Lanes of the wave's active lanes whose digit equals this lane's: per bit,
// the ballot of the bit, kept where it agrees with this lane's bit (x is the
// bit spread to 0 or ~0, so agreement is ~(ballot ^ x)); the active mask is
// applied once at the end.
template <uint32_t DB>
__device__ __forceinline__ uint64_t match_digit(uint32_t d, bool active) {
  uint32_t lo = 0xFFFFFFFFu, hi = 0xFFFFFFFFu;
#pragma unroll
  for (uint32_t b = 0; b < DB; ++b) {
    const uint32_t x = (uint32_t)((int32_t)(d << (31 - b)) >> 31);
    const uint64_t bb = __builtin_amdgcn_ballot_w64(x != 0u);
    lo &= ~((uint32_t)bb ^ x);
    hi &= ~((uint32_t)(bb >> 32) ^ x);
  }
  return ((((uint64_t)hi) << 32) | lo) & __builtin_amdgcn_ballot_w64(active);
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

// Adds one count of digit d for every active lane to the LDS histogram h.
// The lanes sharing the first active lane's digit add once, together: keys
// that share a digit across a wave (a near-constant high byte, a view id)
// would otherwise serialise on one LDS address.
__device__ __forceinline__ void wave_count(uint32_t *h, uint32_t d) {
  const uint32_t lead = __builtin_amdgcn_readfirstlane(d);
  const uint64_t same = __ballot(d == lead);
  if (d != lead)
    atomicAdd(&h[d], 1u);
  else if (__lane_id() == (uint32_t)__builtin_ctzll(same))
    atomicAdd(&h[d], (uint32_t)__popcll(same));
}

__device__ __forceinline__ uint32_t load_count(const uint32_t *n_dev, uint32_t n_cap) {
  return n_dev ? min(*n_dev, n_cap) : n_cap;
}

// Clears fixed + per_tile * (live tiles) words.
__global__ __launch_bounds__(T) void k_os_clear(uint32_t *__restrict__ p, size_t fixed, size_t per_tile,
                                                const uint32_t *__restrict__ n_dev, uint32_t n_cap) {
  const size_t tiles_n = (load_count(n_dev, n_cap) + TILE - 1) / TILE;
  const size_t n = fixed + per_tile * tiles_n;
  for (size_t i = (size_t)blockIdx.x * T + threadIdx.x; i < n; i += (size_t)gridDim.x * T) p[i] = 0u;
}

// Workgroups take chunks of 1024 keys (2 per thread, loaded before any is
// counted: small chunks spread a short sort over the whole chip) and count
// every pass's digits with LDS atomics; each workgroup then adds its
// non-zero bins into one of STRIPES global copies (copy = workgroup mod
// STRIPES) to spread the global atomics.
template <uint32_t DB>
__global__ __launch_bounds__(T) void k_os_hist(const uint32_t *__restrict__ keys, const uint32_t *__restrict__ vals,
                                               const uint32_t *__restrict__ n_dev, uint32_t n_cap, PassSet ps,
                                               uint32_t *__restrict__ hist) {
  constexpr uint32_t D = 1u << DB;
  __shared__ uint32_t h[MAX_PASSES][D];
  for (uint32_t i = threadIdx.x; i < MAX_PASSES * D; i += T) h[i / D][i % D] = 0;
  __syncthreads();
  bool need_vals = false;
  for (uint32_t p = 0; p < ps.count; ++p) need_vals = need_vals || ps.from_value[p];
  const uint32_t n = load_count(n_dev, n_cap);
  for (size_t base = (size_t)blockIdx.x * HIST_CHUNK; base < n; base += (size_t)gridDim.x * HIST_CHUNK) {
    uint32_t k[HIST_IPT], v[HIST_IPT];
#pragma unroll
    for (uint32_t r = 0; r < HIST_IPT; ++r) {
      const size_t i = base + (size_t)r * T + threadIdx.x;
      k[r] = i < n ? keys[i] : 0u;
      v[r] = (need_vals && i < n) ? vals[i] : 0u;
    }
#pragma unroll
    for (uint32_t r = 0; r < HIST_IPT; ++r) {
      if (base + (size_t)r * T + threadIdx.x < n)
        for (uint32_t p = 0; p < ps.count; ++p)
          wave_count(h[p], digit<DB>(k[r], v[r], ps.shift[p], ps.flip[p], ps.from_value[p]));
    }
  }
  __syncthreads();
  uint32_t *dst = hist + (size_t)(blockIdx.x % STRIPES) * MAX_PASSES * D;
  for (uint32_t i = threadIdx.x; i < ps.count * D; i += T) {
    const uint32_t c = h[i / D][i % D];
    if (c) atomicAdd(&dst[i], c);
  }
}

template <uint32_t DB>
__device__ __forceinline__ void clear_whist(uint32_t (*whist)[1u << DB]) {
  constexpr uint32_t D = 1u << DB;
#pragma unroll
  for (uint32_t i = threadIdx.x; i < WAVES * D; i += T) whist[i / D][i % D] = 0;
}

// Loads tile items [base, base+count) and ranks each among the items of
// equal digit in its wave: wave w owns items [w*512, (w+1)*512) in order, 64
// per round; within a round the lanes of equal digit are ranked by lane and
// the running per-digit count lives in whist[w]. On return whist[q][d] is the
// start of wave q's digit-d items among the tile's digit-d items, and the
// result is the tile's count of digit threadIdx.x (0 for threads past the
// digits). whist has WAVES + 1 rows; rows 0..WAVES-1 must be zero.
// Tile items [base, base+count) in registers: item r of a thread is tile
// item w*512 + r*64 + lane (w its wave). All loads are issued before any is
// used: a round's LDS update would otherwise hold back the next round's
// loads, one memory latency per round.
// `per` (a multiple of 64, at most WAVE_ITEMS) spreads a short run over all
// waves: wave w owns items [w * per, (w + 1) * per).
__device__ __forceinline__ void load_tile(const uint32_t *__restrict__ kin, const uint32_t *__restrict__ vin,
                                          size_t base, uint32_t count, uint32_t (&key)[IPT], uint32_t (&val)[IPT],
                                          uint32_t per = WAVE_ITEMS) {
  const uint32_t w = threadIdx.x / 64, lane = __lane_id();
#pragma unroll
  for (uint32_t r = 0; r < IPT; ++r) {
    const uint32_t j = w * per + r * 64 + lane;
    const bool active = j < count;
    key[r] = active ? kin[base + j] : 0u;
    val[r] = active ? vin[base + j] : 0u;
  }
}

// rank_tile of items already loaded by load_tile.
template <uint32_t DB>
__device__ __forceinline__ uint32_t rank_loaded(uint32_t count, uint32_t shift, uint32_t flip, uint32_t from_value,
                                                const uint32_t (&key)[IPT], const uint32_t (&val)[IPT],
                                                uint32_t (&rank)[IPT], uint32_t (*whist)[1u << DB],
                                                const uint16_t *tab = nullptr, uint32_t per = WAVE_ITEMS) {
  const uint32_t t = threadIdx.x, w = t / 64, lane = __lane_id();
  uint32_t *wh = whist[w];
  // Each round's leader (lowest lane of a digit group) adds the group's size
  // to the wave's running count with an LDS atomic that returns the count
  // before it; the group then takes that value from its leader. A wave's LDS
  // operations complete in order, so the rounds' atomics issue back to back
  // instead of each waiting on the last round's read and write. The other
  // lanes add 0 to a word of their own in row WAVES of whist (whatever it
  // holds), so the atomic needs no branch.
  uint32_t old[IPT], leader[IPT];
#pragma unroll
  for (uint32_t r = 0; r < IPT; ++r) {
    if (r * 64 >= per) {  // a round past the wave's items (short runs): nothing to rank
      old[r] = 0;
      leader[r] = lane;
      rank[r] = 0xFFFFFFFFu;
      continue;
    }
    const uint32_t j = w * per + r * 64 + lane;
    const bool active = j < count;
    const uint32_t d = digit<DB>(key[r], val[r], shift, flip, from_value, tab);
    const uint64_t peers = match_digit<DB>(d, active);
    const uint32_t lrank = (uint32_t)__popcll(peers & lanemask_lt());
    const bool lead = active && lrank == 0;
    old[r] = atomicAdd(lead ? &wh[d] : &whist[WAVES][lane], lead ? (uint32_t)__popcll(peers) : 0u);
    leader[r] = active ? (uint32_t)__builtin_ctzll(peers) : lane;
    rank[r] = active ? lrank : 0xFFFFFFFFu;
  }
#pragma unroll
  for (uint32_t r = 0; r < IPT; ++r) {
    const uint32_t before = (uint32_t)__shfl((int)old[r], (int)leader[r], 64);
    if (rank[r] != 0xFFFFFFFFu) rank[r] += before;
  }
  __syncthreads();
  // Thread t < 2^DB owns digit t: turn the wave counts into wave offsets.
  uint32_t run = 0;
  if (t < (1u << DB)) {
#pragma unroll
    for (uint32_t q = 0; q < WAVES; ++q) {
      const uint32_t c = whist[q][t];
      whist[q][t] = run;
      run += c;
    }
  }
  return run;
}

template <uint32_t DB>
__device__ __forceinline__ uint32_t rank_tile(const uint32_t *__restrict__ kin, const uint32_t *__restrict__ vin,
                                              size_t base, uint32_t count, uint32_t shift, uint32_t flip,
                                              uint32_t from_value, uint32_t (&key)[IPT], uint32_t (&val)[IPT],
                                              uint32_t (&rank)[IPT], uint32_t (*whist)[1u << DB]) {
  load_tile(kin, vin, base, count, key, val);
  return rank_loaded<DB>(count, shift, flip, from_value, key, val, rank, whist);
}

// Stages the ranked tile in LDS in digit order (lstart[d] = start of digit d
// in the tile) and writes it out coalesced; digit d's items go to
// dst[d] + (position among the tile's digit-d items). Positions at or past
// the element count n are not written: with correct offsets every position
// is below n (the offsets are a permutation), and offsets broken by a fault
// elsewhere then lose items (a parity failure) instead of writing past the
// buffers.
template <uint32_t DB>
__device__ __forceinline__ void scatter_tile(const uint32_t (&key)[IPT], const uint32_t (&val)[IPT],
                                             const uint32_t (&rank)[IPT], uint32_t (*whist)[1u << DB],
                                             const uint32_t *lstart, const uint32_t *dst, uint32_t count,
                                             uint32_t shift, uint32_t flip, uint32_t from_value, uint32_t *skeys,
                                             uint32_t *svals, uint32_t n, uint32_t *__restrict__ kout,
                                             uint32_t *__restrict__ vout, const uint16_t *tab = nullptr) {
  const uint32_t t = threadIdx.x, w = t / 64;
#pragma unroll
  for (uint32_t r = 0; r < IPT; ++r) {
    if (rank[r] != 0xFFFFFFFFu) {
      const uint32_t d = digit<DB>(key[r], val[r], shift, flip, from_value, tab);
      const uint32_t pos = lstart[d] + whist[w][d] + rank[r];
      if (pos < TILE) {  // always, for ranks of one tile; bounds the LDS stage regardless
        skeys[pos] = key[r];
        svals[pos] = val[r];
      }
    }
  }
  __syncthreads();
  for (uint32_t j = t; j < count; j += T) {
    const uint32_t k = skeys[j], v = svals[j];
    const uint32_t d = digit<DB>(k, v, shift, flip, from_value, tab);
    const size_t pos = (size_t)dst[d] + (j - lstart[d]);
    if (pos < n) {
      kout[pos] = k;
      vout[pos] = v;
    }
  }
}

// Bucket sort. The visible keys' range [kmin, kmax] is cut into kFineBins
// fine bins, fine bin f = (kmax - key) >> s (largest keys first); the plan
// groups consecutive fine bins into buckets of about `fill` keys (a new bucket
// opens every fill keys of the running count, fill = max(1024, visible /
// 254)), so one stable 8-bit pass by bucket leaves runs of at most fill + (one
// fine bin) keys, and each run is sorted in LDS. (An 8-bit pass costs about
// two thirds of a 9-bit one; at 2^19 keys fill is 2065, still well inside the
// 4096 keys of a local sort.) Float depth keys are
// logarithmic in depth, so a few far or near splats stretch the range only
// mildly.
constexpr uint32_t BUCKET_BITS = 8;
constexpr uint32_t BUCKETS = 1u << BUCKET_BITS;  // buckets of the visible keys, the last for key 0
constexpr uint32_t FINE_PER = kFineBins / T;     // fine bins per thread of the plan
static_assert(FINE_PER % 4 == 0, "fine bins are read four at a time");

// Scratch words of a bucket sort of n keys, in order: the look-back region of
// one 9-bit pass (digit histograms, unused here, tile counter, status words)
// and the fine histogram, all zero on entry; then the range words: kmax, s,
// the bucket starts [BUCKETS + 1], and the first and last fine bin of each
// bucket (at kBucketFirst and kBucketLast from the range words).
struct BucketLayout {
  size_t fine, range, zero_words, words;
};
constexpr uint32_t kBucketFirst = 3 + BUCKETS;
constexpr uint32_t kBucketLast = kBucketFirst + BUCKETS;
constexpr uint32_t kBucketRangeWords = kBucketLast + BUCKETS;

// Bucket sort, step 2 (the prologue of the bucket pass, in every workgroup):
// from the fine histogram, fine bin f (running count P of the bins before it)
// goes to bucket P / fill: tab[f] (LDS). Returns, for thread t < BUCKETS, the
// count of bucket t (t = BUCKETS - 1: the n - visible zero keys). With
// first/last, also stores each bucket's first and last non-empty fine bin.
// plan: LDS words [3 * BUCKETS] (counts, first, last).
__device__ __forceinline__ uint32_t bucket_plan(const uint32_t *__restrict__ fine, uint32_t n, uint16_t *tab,
                                                uint32_t *plan, uint32_t *lds4, uint32_t *__restrict__ first,
                                                uint32_t *__restrict__ last) {
  const uint32_t t = threadIdx.x;
  uint32_t *cnt = plan, *lo = plan + BUCKETS, *hi = plan + 2 * BUCKETS;
  for (uint32_t j = t; j < BUCKETS; j += T) {
    cnt[j] = 0;
    lo[j] = 0xFFFFFFFFu;
    hi[j] = 0;
  }
  uint32_t c[FINE_PER], sum = 0;
#pragma unroll
  for (uint32_t q = 0; q < FINE_PER; q += 4) {
    const uint4 x = reinterpret_cast<const uint4 *>(fine + t * FINE_PER)[q / 4];
    c[q] = x.x;
    c[q + 1] = x.y;
    c[q + 2] = x.z;
    c[q + 3] = x.w;
    sum += x.x + x.y + x.z + x.w;
  }
  uint32_t p = block_exclusive_scan(sum, lds4);  // its barrier orders the clears above
  uint32_t visible = 0;
#pragma unroll
  for (uint32_t q = 0; q < WAVES; ++q) visible += lds4[q];
  const uint32_t fill = max(1024u, (visible + BUCKETS - 3) / (BUCKETS - 2));
#pragma unroll
  for (uint32_t q = 0; q < FINE_PER; ++q) {
    const uint32_t b = min(p / fill, BUCKETS - 2), f = t * FINE_PER + q;
    tab[f] = (uint16_t)b;
    p += c[q];
    if (c[q]) {
      atomicAdd(&cnt[b], c[q]);
      if (first) {
        atomicMin(&lo[b], f);
        atomicMax(&hi[b], f);
      }
    }
  }
  __syncthreads();
  if (first && t < BUCKETS) {
    first[t] = lo[t];
    last[t] = hi[t];
  }
  return t < BUCKETS - 1 ? cnt[t] : t == BUCKETS - 1 ? n - visible : 0u;
}

// Onesweep pass: one launch per digit, tile offsets by decoupled look-back.
// Persistent: each workgroup takes tiles from the pass counter until they run
// past the live element count, so the grid is bounded whatever the capacity.
// Pass p's status words start at p * (live tiles) * 2^DB.
// A bucket pass (BUCKET, from_value kRangeDigit, BUCKET_BITS-bit digits)
// reads kmax and the fine shift from range[0..1], plans the fine-bin ->
// bucket table into LDS from the fine histogram, and the workgroup of tile 0
// stores every digit's start at range[2 + d] (range[2 + 2^DB] = n).
template <uint32_t DB, bool BUCKET = false>
__global__ __launch_bounds__(T) void k_os_pass(const uint32_t *__restrict__ kin, const uint32_t *__restrict__ vin,
                                               uint32_t *__restrict__ kout, uint32_t *__restrict__ vout,
                                               const uint32_t *__restrict__ n_dev, uint32_t n_cap, uint32_t pass,
                                               uint32_t shift, uint32_t flip, uint32_t from_value,
                                               const uint32_t *__restrict__ hist, uint32_t *__restrict__ counter,
                                               uint32_t *__restrict__ status_all, uint32_t *__restrict__ range,
                                               const uint32_t *__restrict__ fine, TileRangesOut tro) {
  constexpr bool kBucket = BUCKET;  // only bucket passes carry the bucket plan's LDS
  static_assert(!BUCKET || DB == BUCKET_BITS, "a bucket pass has one digit per bucket");
  __shared__ uint16_t tab[kBucket ? kFineBins : 1];
  __shared__ uint32_t plan[kBucket ? 3 * BUCKETS : 1];
  if (kBucket && from_value == kRangeDigit) {
    flip = range[0];
    shift = range[1];
  }
  __shared__ uint32_t s_tile;
  __shared__ uint32_t lds4[WAVES];
  constexpr uint32_t D = 1u << DB;
  __shared__ uint32_t gofs[D];    // global start of this tile's items of each digit
  __shared__ uint32_t lstart[D];  // start of each digit inside this tile
  __shared__ uint32_t whist[WAVES + 1][D];  // row WAVES: rank_loaded's spare words
  __shared__ uint32_t skeys[TILE];
  __shared__ uint32_t svals[TILE];

  const uint32_t t = threadIdx.x;
  const uint32_t n = load_count(n_dev, n_cap);
  const uint32_t tiles_n = (n + TILE - 1) / TILE;
  uint32_t *status = status_all + (size_t)pass * tiles_n * D;
  // The first tile is claimed before anything else and its items are loaded
  // together with the digit histogram: the claim, the histogram and the tile
  // cost one memory round trip after another instead of three.
  if (t == 0) s_tile = atomicAdd(counter, 1u);
  clear_whist<DB>(whist);
  __syncthreads();
  uint32_t tile = s_tile;
  const uint32_t first_tile = tile;
  size_t base = (size_t)tile * TILE;
  uint32_t count = base < n ? (uint32_t)min((size_t)TILE, (size_t)n - base) : 0u;
  uint32_t key[IPT], val[IPT], rank[IPT];
  load_tile(kin, vin, base, count, key, val);
  uint32_t hd = 0;
  if (kBucket && from_value == kRangeDigit) {
    // every workgroup plans the buckets (fine bin -> bucket table in LDS);
    // tile 0's also stores each bucket's fine bin span for the local sort
    hd = bucket_plan(fine, n, tab, plan, lds4, tile == 0 ? range + kBucketFirst : nullptr,
                     tile == 0 ? range + kBucketLast : nullptr);
  } else if (t < D) {
#pragma unroll
    for (uint32_t c = 0; c < STRIPES; ++c) hd += hist[(c * MAX_PASSES + pass) * D + t];
  }
  const uint32_t g = block_exclusive_scan(hd, lds4);
  if (kBucket && from_value == kRangeDigit && tile == 0) {
    if (t < D) range[2 + t] = g;
    if (t == 0) range[2 + D] = n;
  }
  // every key has the same digit: the pass is the identity
  const bool trivial = __syncthreads_or(n > 0 && hd == n) != 0;

  while (base < n) {  // uniform: no later tile has work either
    if (trivial) {
      const uint32_t w = t / 64, lane = __lane_id();
#pragma unroll
      for (uint32_t r = 0; r < IPT; ++r) {
        const uint32_t j = w * WAVE_ITEMS + r * 64 + lane;
        if (j < count) {
          kout[base + j] = key[r];
          vout[base + j] = val[r];
        }
      }
    } else {
      const uint32_t run = rank_loaded<DB>(count, shift, flip, from_value, key, val, rank, whist, tab);

      // Publish this tile's count of digit t, then look back for the prefix.
      uint32_t excl = 0;
      uint32_t *st = status + (size_t)tile * D + t;
      if (t >= D) {
      } else if (tile == 0) {
        __hip_atomic_store(st, FLAG_INC | run, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      } else {
        __hip_atomic_store(st, FLAG_AGG | run, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        int64_t j = (int64_t)tile - 1;  // next predecessor not yet consumed
        while (j >= 0) {
          uint32_t sv[LOOKBACK];
#pragma unroll
          for (uint32_t k = 0; k < LOOKBACK; ++k)
            sv[k] = (j - (int64_t)k >= 0) ? __hip_atomic_load(status + (size_t)(j - k) * D + t, __ATOMIC_RELAXED,
                                                              __HIP_MEMORY_SCOPE_AGENT)
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
      }
      const uint32_t ls = block_exclusive_scan(run, lds4);
      if (t < D) {
        gofs[t] = g + excl;
        lstart[t] = ls;
      }
      __syncthreads();
      scatter_tile<DB>(key, val, rank, whist, lstart, gofs, count, shift, flip, from_value, skeys, svals, n, kout,
                       vout, tab);
    }
    __syncthreads();
    if (t == 0) s_tile = atomicAdd(counter, 1u);
    clear_whist<DB>(whist);
    __syncthreads();
    tile = s_tile;
    base = (size_t)tile * TILE;
    count = base < n ? (uint32_t)min((size_t)TILE, (size_t)n - base) : 0u;
    if (base < n) load_tile(kin, vin, base, count, key, val);
  }
  if (tro.ranges && first_tile == 0) {  // uniform: the workgroup that sorted tile 0, off the look-back chain
    __shared__ uint32_t olocal[kOrderClasses];
    if (t < kOrderClasses) olocal[t] = 0;
    __syncthreads();
    uint32_t cls = 0, place = 0;
    if (t < tro.count) {  // count <= 2^DB: digit t is tile t
      const uint32_t b0 = min(g, n), b1 = min(g + hd, n);
      tro.ranges[t] = make_uint2(b0, b1);
      cls = raster_order_class(b1 - b0);
      place = atomicAdd(&olocal[cls], 1u);
    }
    __syncthreads();
    if (t < tro.count) tro.order[(size_t)cls * tro.count + place] = t;
    if (t < kOrderClasses) tro.order_counts[t] = olocal[t];
  }
}

// First tile of a reduce-then-scan workgroup (grid a multiple of XCDS):
// each XCD takes one run of consecutive tiles, so the partial cache lines
// that neighbouring tiles write into one digit's output (or into one
// digit's row of counts) meet in the same L2.
__device__ __forceinline__ uint32_t xcd_first_tile(uint32_t b, uint32_t grid) {
  return (b % XCDS) * (grid / XCDS) + b / XCDS;
}

// Index of the (digit, tile) count for D digits. Unsegmented: digit-major
// over all tiles (d * tiles + tile). Segmented: segment by segment,
// digit-major inside each (first * D + d * seg_tiles + tile - first, `first`
// the segment's first tile), so one flat exclusive scan sorts every segment
// in its own range.
struct CountIndex {
  uint32_t first, span, digits;
  __device__ __forceinline__ size_t operator()(uint32_t d, uint32_t tile) const {
    return (size_t)first * digits + (size_t)d * span + (tile - first);
  }
};

template <uint32_t DB>
__device__ __forceinline__ CountIndex count_index(uint32_t tile, uint32_t tiles, const uint32_t *seg_start,
                                                  uint32_t segs) {
  if (segs == 0) return CountIndex{0, tiles, 1u << DB};
  uint32_t lo = 0, hi = segs;  // last segment whose start <= the tile's first element
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (seg_start[mid] / TILE <= tile) lo = mid; else hi = mid;
  }
  const uint32_t first = seg_start[lo] / TILE;
  const uint32_t end = lo + 1 < segs ? seg_start[lo + 1] / TILE : tiles;
  return CountIndex{first, end - first, 1u << DB};
}

// Reduce-then-scan pass, step 1: digit counts of every live tile laid out
// by count_index, so one exclusive scan of the 2^DB * tiles counts (length
// published in *entries) gives every (digit, tile) its output start. Block 0
// also clears the scan state of the pass (ticket + one word per chunk).
template <uint32_t DB>
__global__ __launch_bounds__(T) void k_rts_up(const uint32_t *__restrict__ kin, const uint32_t *__restrict__ vin,
                                              const uint32_t *__restrict__ n_dev, uint32_t n_cap, uint32_t shift,
                                              uint32_t flip, uint32_t from_value, uint32_t *__restrict__ counts,
                                              uint32_t *__restrict__ entries, uint32_t *__restrict__ scan_state,
                                              const uint32_t *__restrict__ seg_start, uint32_t segs) {
  constexpr uint32_t D = 1u << DB;
  __shared__ uint32_t h[D];  // one table for the workgroup (wave_count folds equal digits)
  const uint32_t t = threadIdx.x;
  const uint32_t n = load_count(n_dev, n_cap);
  const uint32_t tiles = (n + TILE - 1) / TILE;
  if (blockIdx.x == 0) {
    if (t == 0) *entries = D * tiles;
    // the scan's ticket and chunk status words start every pass at zero
    const uint32_t chunks = (D * tiles + SCAN_CHUNK - 1) / SCAN_CHUNK;
    for (uint32_t i = t; i <= chunks; i += T) scan_state[i] = 0;
  }
  const uint32_t *src = from_value ? vin : kin;
  for (uint32_t tile = xcd_first_tile(blockIdx.x, gridDim.x); tile < tiles; tile += gridDim.x) {
    for (uint32_t i = t; i < D; i += T) h[i] = 0;
    __syncthreads();
    const size_t base = (size_t)tile * TILE;
    const uint32_t count = (uint32_t)min((size_t)TILE, (size_t)n - base);
    uint32_t x[IPT];
#pragma unroll
    for (uint32_t r = 0; r < IPT; ++r) {
      const uint32_t j = r * T + t;
      x[r] = j < count ? src[base + j] : 0u;
    }
#pragma unroll
    for (uint32_t r = 0; r < IPT; ++r)
      if (r * T + t < count) wave_count(h, ((x[r] ^ flip) >> shift) & (D - 1u));
    __syncthreads();
    if (t < D) counts[count_index<DB>(tile, tiles, seg_start, segs)(t, tile)] = h[t];
    __syncthreads();
  }
}

// Reduce-then-scan pass, step 2: exclusive scan of the live counts in
// place, one launch. Workgroups take chunks of SCAN_CHUNK counts by ticket
// (so every chunk a workgroup waits on has started), publish the chunk total
// and resolve the prefix over the preceding chunks by decoupled look-back:
// wave 0 polls 64 predecessors per step and adds published totals up to the
// nearest inclusive prefix. Count totals are at most the element count, so
// they fit the 30 value bits of a status word.
__global__ __launch_bounds__(T) void k_rts_scan(uint32_t *__restrict__ counts, const uint32_t *__restrict__ entries,
                                                uint32_t *__restrict__ scan_state) {
  __shared__ uint32_t s_chunk, s_prefix;
  __shared__ uint32_t lds4[WAVES];
  const uint32_t t = threadIdx.x, lane = __lane_id();
  const uint32_t n = *entries;  // a multiple of 256: a thread's 16 counts are all live or all past the end
  uint32_t *status = scan_state + 1;
  for (;;) {
    if (t == 0) s_chunk = atomicAdd(scan_state, 1u);
    __syncthreads();
    const uint32_t c = s_chunk;
    const size_t base = (size_t)c * SCAN_CHUNK;
    if (base >= n) break;  // uniform: later tickets are past the end too
    const size_t mine = base + (size_t)t * SCAN_IPT;
    const bool live = mine < n;
    uint4 q[SCAN_IPT / 4];
#pragma unroll
    for (uint32_t k = 0; k < SCAN_IPT / 4; ++k)
      q[k] = live ? reinterpret_cast<const uint4 *>(counts + mine)[k] : make_uint4(0u, 0u, 0u, 0u);
    uint32_t sum = 0;
#pragma unroll
    for (uint32_t k = 0; k < SCAN_IPT / 4; ++k) sum += q[k].x + q[k].y + q[k].z + q[k].w;
    uint32_t run = block_exclusive_scan(sum, lds4);
    if (t < 64) {
      uint32_t total = 0;
#pragma unroll
      for (uint32_t q = 0; q < WAVES; ++q) total += lds4[q];
      uint32_t prefix = 0;
      if (c == 0) {
        if (t == 0) __hip_atomic_store(status, FLAG_INC | total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      } else {
        if (t == 0) __hip_atomic_store(status + c, FLAG_AGG | total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        int64_t j = (int64_t)c - 1;  // nearest predecessor not yet added
        for (;;) {
          const int64_t pred = j - (int64_t)lane;
          const uint32_t sv =
              pred >= 0 ? __hip_atomic_load(status + pred, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : FLAG_INC;
          const uint32_t flag = sv & ~VALUE_MASK;
          const uint64_t unpublished = __ballot(flag == 0);
          const uint64_t inclusive = __ballot(flag == FLAG_INC);
          const uint32_t first_gap = unpublished ? (uint32_t)__builtin_ctzll(unpublished) : 64u;
          const uint32_t first_inc = inclusive ? (uint32_t)__builtin_ctzll(inclusive) : 64u;
          const bool done = first_inc < first_gap;
          const uint32_t take = done ? first_inc + 1 : first_gap;
          const uint32_t part = wave_inclusive_scan(lane < take ? (sv & VALUE_MASK) : 0u);
          prefix += __shfl(part, 63, 64);
          if (done) break;
          j -= take;
        }
        if (t == 0)
          __hip_atomic_store(status + c, FLAG_INC | (prefix + total), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      if (t == 0) s_prefix = prefix;
    }
    __syncthreads();
    run += s_prefix;
    if (live) {
#pragma unroll
      for (uint32_t k = 0; k < SCAN_IPT / 4; ++k) {
        uint4 o;
        o.x = run; run += q[k].x;
        o.y = run; run += q[k].y;
        o.z = run; run += q[k].z;
        o.w = run; run += q[k].w;
        reinterpret_cast<uint4 *>(counts + mine)[k] = o;
      }
    }
    __syncthreads();  // s_chunk, s_prefix and lds4 are reused by the next chunk
  }
}

// Reduce-then-scan pass, step 3: rank and scatter with the scanned offsets.
template <uint32_t DB>
__global__ __launch_bounds__(T) void k_rts_down(const uint32_t *__restrict__ kin, const uint32_t *__restrict__ vin,
                                                uint32_t *__restrict__ kout, uint32_t *__restrict__ vout,
                                                const uint32_t *__restrict__ n_dev, uint32_t n_cap, uint32_t shift,
                                                uint32_t flip, uint32_t from_value,
                                                const uint32_t *__restrict__ offsets,
                                                const uint32_t *__restrict__ seg_start, uint32_t segs) {
  __shared__ uint32_t lds4[WAVES];
  constexpr uint32_t D = 1u << DB;
  __shared__ uint32_t dst[D];
  __shared__ uint32_t lstart[D];
  __shared__ uint32_t whist[WAVES + 1][D];  // row WAVES: rank_loaded's spare words
  __shared__ uint32_t skeys[TILE];
  __shared__ uint32_t svals[TILE];
  const uint32_t t = threadIdx.x;
  const uint32_t n = load_count(n_dev, n_cap);
  const uint32_t tiles = (n + TILE - 1) / TILE;
  for (uint32_t tile = xcd_first_tile(blockIdx.x, gridDim.x); tile < tiles; tile += gridDim.x) {
    const size_t base = (size_t)tile * TILE;
    const uint32_t count = (uint32_t)min((size_t)TILE, (size_t)n - base);
    clear_whist<DB>(whist);
    if (t < D) dst[t] = offsets[count_index<DB>(tile, tiles, seg_start, segs)(t, tile)];
    __syncthreads();
    uint32_t key[IPT], val[IPT], rank[IPT];
    const uint32_t run = rank_tile<DB>(kin, vin, base, count, shift, flip, from_value, key, val, rank, whist);
    const uint32_t ls = block_exclusive_scan(run, lds4);
    if (t < D) lstart[t] = ls;
    __syncthreads();
    scatter_tile<DB>(key, val, rank, whist, lstart, dst, count, shift, flip, from_value, skeys, svals, n, kout,
                     vout);
    __syncthreads();
  }
}

// Bucket sort, step 1 (grid-stride over 4096-key chunks): every workgroup
// reduces the producer's per-workgroup key ranges to kmin (smallest non-zero
// key) and kmax, takes the smallest shift s with (kmax - kmin) >> s <
// kFineBins, and adds its non-zero keys' fine bins into the fine histogram;
// workgroup 0 stores range[0] = kmax, range[1] = s.
__global__ __launch_bounds__(T) void k_bucket_hist(const uint32_t *__restrict__ keys, uint32_t n,
                                                   const uint2 *__restrict__ parts, uint32_t nparts,
                                                   uint32_t *__restrict__ fine, uint32_t *__restrict__ range) {
  __shared__ uint32_t h[kFineBins];
  __shared__ uint32_t red[2][WAVES];
  const uint32_t t = threadIdx.x, lane = __lane_id(), w = t / 64;
  for (uint32_t i = t; i < kFineBins; i += T) h[i] = 0;
  uint32_t mn = 0xFFFFFFFFu, mx = 0;
  for (uint32_t i = t; i < nparts; i += T) {
    const uint2 p = parts[i];
    mn = min(mn, p.x);
    mx = max(mx, p.y);
  }
#pragma unroll
  for (uint32_t d = 1; d < 64; d <<= 1) {
    mn = min(mn, (uint32_t)__shfl_xor((int)mn, (int)d, 64));
    mx = max(mx, (uint32_t)__shfl_xor((int)mx, (int)d, 64));
  }
  if (lane == 0) {
    red[0][w] = mn;
    red[1][w] = mx;
  }
  __syncthreads();
#pragma unroll
  for (uint32_t q = 0; q < WAVES; ++q) {
    mn = min(mn, red[0][q]);
    mx = max(mx, red[1][q]);
  }
  const uint32_t span = mn == 0xFFFFFFFFu ? 0u : mx - mn;  // no non-zero key: every key is 0
  uint32_t s = 0;
  while ((span >> s) >= kFineBins) ++s;
  if (blockIdx.x == 0 && t == 0) {
    range[0] = mx;
    range[1] = s;
  }
  for (size_t base = (size_t)blockIdx.x * TILE; base < n; base += (size_t)gridDim.x * TILE) {
    uint32_t k[IPT];
#pragma unroll
    for (uint32_t r = 0; r < IPT; ++r) {
      const size_t i = base + (size_t)r * T + t;
      k[r] = i < n ? keys[i] : 0u;
    }
#pragma unroll
    for (uint32_t r = 0; r < IPT; ++r)
      if (k[r] != 0u) atomicAdd(&h[min((mx - k[r]) >> s, kFineBins - 1)], 1u);
  }
  __syncthreads();
  for (uint32_t i = t; i < kFineBins; i += T) {
    const uint32_t c = h[i];
    if (c) atomicAdd(&fine[i], c);
  }
}

// Bucket sort, step 4: workgroup b sorts bucket b's run [range[2 + b],
// range[3 + b]) in place. Its keys lie in fine bins first[b]..last[b], so
// w = kmax - key - (first[b] << s) spans fewer than (last - first + 1) << s
// values; stable 8-bit LDS passes over those bits order the run by w
// ascending (key descending), equal keys keeping the order the stable
// bucket pass gave them, and keys are restored on the way out. The run is
// spread over all waves (per items each). A run longer than one tile sets
// *overflow (a fine bin of more than about `fill` keys).
__global__ __launch_bounds__(T) void k_bucket_local(uint32_t *__restrict__ keys, uint32_t *__restrict__ vals,
                                                    const uint32_t *__restrict__ range,
                                                    const uint32_t *__restrict__ first,
                                                    const uint32_t *__restrict__ last,
                                                    uint32_t *__restrict__ overflow) {
  constexpr uint32_t LB = 8, LD = 1u << LB;  // digits of a local pass (9-bit passes measured slower)
  __shared__ uint32_t lds4[WAVES];
  __shared__ uint32_t lstart[LD];
  __shared__ uint32_t whist[WAVES + 1][LD];  // row WAVES: rank_loaded's spare words
  __shared__ uint32_t skeys[TILE];
  __shared__ uint32_t svals[TILE];
  const uint32_t b = blockIdx.x, t = threadIdx.x, w = t / 64, lane = __lane_id();
  const uint32_t kmax = range[0], s = range[1];
  const uint32_t off = range[2 + b], cnt = range[3 + b] - off;
  if (b >= BUCKETS - 1 || cnt <= 1) return;
  if (cnt > TILE) {
    if (t == 0) atomicOr(overflow, 1u);
    return;
  }
  const uint32_t f0 = first[b], f1 = last[b];
  const uint64_t wspan = (((uint64_t)(f1 - f0) + 1) << s) - 1;  // largest w
  const uint32_t bits = wspan ? 64 - __builtin_clzll(wspan) : 0;
  if (bits == 0) return;  // one key value: already in order
  const uint32_t base = kmax - (f0 << s);
  const uint32_t per = min(WAVE_ITEMS, (cnt + WAVES * 64 - 1) / (WAVES * 64) * 64);
  uint32_t key[IPT], val[IPT], rank[IPT];
  load_tile(keys, vals, off, cnt, key, val, per);
#pragma unroll
  for (uint32_t r = 0; r < IPT; ++r) key[r] = base - key[r];
  const uint32_t passes = (bits + LB - 1) / LB;
  for (uint32_t p = 0; p < passes; ++p) {
    clear_whist<LB>(whist);
    __syncthreads();
    const uint32_t run = rank_loaded<LB>(cnt, LB * p, 0u, 0u, key, val, rank, whist, nullptr, per);
    const uint32_t ls = block_exclusive_scan(run, lds4);
    if (t < LD) lstart[t] = ls;
    __syncthreads();
#pragma unroll
    for (uint32_t r = 0; r < IPT; ++r) {
      if (rank[r] != 0xFFFFFFFFu) {
        const uint32_t d = (key[r] >> (LB * p)) & (LD - 1u);
        const uint32_t pos = lstart[d] + whist[w][d] + rank[r];
        if (pos < TILE) {  // always (ranks of one tile); bounds the LDS stage regardless
          skeys[pos] = key[r];
          svals[pos] = val[r];
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (uint32_t r = 0; r < IPT; ++r) {
      const uint32_t j = w * per + r * 64 + lane;
      if (r * 64 < per && j < cnt) {
        key[r] = skeys[j];
        val[r] = svals[j];
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (uint32_t r = 0; r < IPT; ++r) {
    const uint32_t j = w * per + r * 64 + lane;
    if (r * 64 < per && j < cnt) {
      keys[off + j] = base - key[r];
      vals[off + j] = val[r];
    }
  }
}

uint32_t div_up(size_t a, size_t b) { return (uint32_t)((a + b - 1) / b); }

uint32_t hist_blocks(uint32_t n_cap) {
  return std::max<uint32_t>(std::min<uint32_t>(div_up(n_cap, HIST_CHUNK), HIST_BLOCKS), 1);
}

}  // namespace

// Tile count from which the reduce-then-scan passes replace the look-back
// passes: the look-back front advances a bounded number of tiles per
// coherent round trip, so long sorts are faster with three launches per pass
// and no chain (GSPLAT_SORT_RTS_MIN_TILES overrides).
// Bound of the pass grids (GSPLAT_SORT_MAX_GRID overrides, default none):
// the passes loop over tiles, so any grid covers any count; one workgroup
// per tile measured fastest for long sorts (no per-workgroup serial tiles).
static uint32_t pass_grid_bound() {
  const char *e = std::getenv("GSPLAT_SORT_MAX_GRID");
  return e ? std::max<uint32_t>((uint32_t)std::strtoul(e, nullptr, 10), 1u) : 0xFFFFFFFFu;
}

static uint32_t rts_min_tiles() {
  const char *e = std::getenv("GSPLAT_SORT_RTS_MIN_TILES");
  return e ? (uint32_t)std::strtoul(e, nullptr, 10) : 512u;
}

// Layout for D = 2^digit_bits digits, look-back mode: hist
// [STRIPES][MAX_PASSES][D] | counters [MAX_PASSES] | status
// [passes][tiles][D], all cleared at the start of every sort.
// Reduce-then-scan mode: counts [D][tiles] | live count length | scan
// ticket | one status word per scan chunk (cleared by k_rts_up).
static size_t lookback_words(uint32_t n_cap, uint32_t passes, uint32_t digits) {
  return (size_t)STRIPES * MAX_PASSES * digits + MAX_PASSES +
         (size_t)passes * std::max<uint32_t>(div_up(n_cap, TILE), 1) * digits;
}

size_t onesweep_status_words(uint32_t n_cap, uint32_t bits) {
  // the larger of the two layouts, with a value pass and the widest digit
  const uint32_t digits = bits == MAX_DIGIT_BITS ? 1u << MAX_DIGIT_BITS : 256u;
  const uint32_t passes = std::min<uint32_t>(std::max<uint32_t>((bits + 7) / 8, 1), 4) + 1;
  const size_t tiles = std::max<uint32_t>(div_up(n_cap, TILE), 1);
  const size_t rts = digits * tiles + 2 + div_up(digits * tiles, SCAN_CHUNK);
  return std::max(lookback_words(n_cap, passes, digits), rts);
}

RadixPasses radix_passes(uint32_t bits, bool descending, int value_byte) {
  RadixPasses ps{};
  // 9-bit keys (a view of up to 512 binning tiles) sort in one 9-bit pass
  // instead of two 8-bit ones; a value pass keeps 8-bit digits (its byte).
  ps.digit_bits = (bits == MAX_DIGIT_BITS && value_byte < 0) ? MAX_DIGIT_BITS : 8u;
  const uint32_t key_passes = std::min<uint32_t>((bits + ps.digit_bits - 1) / ps.digit_bits, 4);
  for (uint32_t p = 0; p < key_passes; ++p) {
    ps.shift[ps.count] = ps.digit_bits * p;
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
  return ps;
}

bool radix_uses_lookback(uint32_t n_cap, uint32_t n_hint, uint32_t min_tiles) {
  const uint32_t n = n_hint ? std::min(n_hint, n_cap) : n_cap;
  return div_up(n, TILE) < (min_tiles ? min_tiles : rts_min_tiles());
}

size_t radix_lookback_words(uint32_t n_cap, const RadixPasses &ps) {
  return lookback_words(n_cap, ps.count, 1u << ps.digit_bits);
}

size_t radix_rts_words(uint32_t n_cap, const RadixPasses &ps) {
  const size_t entries = (size_t)(1u << ps.digit_bits) * std::max<uint32_t>(div_up(n_cap, TILE), 1);
  return entries + 2 + div_up(entries, SCAN_CHUNK);  // counts | live count | ticket | chunk status words
}

namespace {

template <uint32_t DB>
void sort_passes(uint32_t **keys, uint32_t **vals, uint32_t **keys_alt, uint32_t **vals_alt, uint32_t n_cap,
                 const uint32_t *n_dev, const PassSet &ps, uint32_t *scratch, size_t scratch_words, hipStream_t st,
                 int value_byte, uint32_t n_hint, const SortPrep &prep) {
  constexpr uint32_t D = 1u << DB;
  const uint32_t tiles = div_up(n_cap, TILE);
  // Grid: one workgroup per expected tile (the hint, for a device-side count,
  // with the capacity as the cap), bounded by GSPLAT_SORT_MAX_GRID.
  const uint32_t expect = n_dev && n_hint ? std::min(div_up(n_hint, TILE), tiles) : tiles;
  const uint32_t grid_tiles = std::min(std::max(expect, std::min(tiles, kMinGrid)), pass_grid_bound());
  if (!radix_uses_lookback(n_cap, n_dev ? n_hint : 0, prep.rts_min_tiles)) {
    const size_t entries = (size_t)D * tiles;
    const uint32_t chunks = div_up(entries, SCAN_CHUNK);
    if (entries + 2 + chunks > scratch_words) throw Error(Code::Internal, "radix sort scratch too small");
    uint32_t *counts = scratch;
    uint32_t *live_entries = counts + entries;
    uint32_t *scan_state = live_entries + 1;
    const uint32_t scan_grid = std::min(chunks, kScanGrid);
    // the tile order of xcd_first_tile needs a grid of whole XCD runs
    const uint32_t grid = (grid_tiles + XCDS - 1) / XCDS * XCDS;
    // Segments replace a value pass that orders by segment index.
    const bool segmented = prep.seg_start && prep.segs > 1 && value_byte >= 0 && !n_dev;
    const uint32_t passes = segmented ? ps.count - 1 : ps.count;
    const uint32_t *seg = segmented ? prep.seg_start : nullptr;
    const uint32_t segs = segmented ? prep.segs : 0;
    for (uint32_t p = 0; p < passes; ++p) {
      k_rts_up<DB><<<grid, T, 0, st>>>(*keys, *vals, n_dev, n_cap, ps.shift[p], ps.flip[p], ps.from_value[p],
                                       counts, live_entries, scan_state, seg, segs);
      k_rts_scan<<<scan_grid, T, 0, st>>>(counts, live_entries, scan_state);
      k_rts_down<DB><<<grid, T, 0, st>>>(*keys, *vals, *keys_alt, *vals_alt, n_dev, n_cap, ps.shift[p], ps.flip[p],
                                         ps.from_value[p], counts, seg, segs);
      std::swap(*keys, *keys_alt);
      std::swap(*vals, *vals_alt);
    }
    GS_HIP(hipGetLastError());
    return;
  }
  const size_t fixed = (size_t)STRIPES * MAX_PASSES * D + MAX_PASSES;
  const size_t words = fixed + (size_t)ps.count * tiles * D;
  if (words > scratch_words) throw Error(Code::Internal, "onesweep scratch too small");
  uint32_t *hist = scratch;
  uint32_t *counters = hist + (size_t)STRIPES * MAX_PASSES * D;
  uint32_t *status = counters + MAX_PASSES;
  if (!prep.scratch_clean)
    k_os_clear<<<std::min<uint32_t>(div_up(words, T), 1024), T, 0, st>>>(scratch, fixed, (size_t)ps.count * D,
                                                                          n_dev, n_cap);
  if (!prep.hist_ready) k_os_hist<DB><<<hist_blocks(n_cap), T, 0, st>>>(*keys, *vals, n_dev, n_cap, ps, hist);
  const uint32_t grid = std::min(grid_tiles, kPassGrid);
  for (uint32_t p = 0; p < ps.count; ++p) {
    const bool tiles_out = ps.count == 1 && prep.hist_ready && prep.tiles.ranges;
    k_os_pass<DB><<<grid, T, 0, st>>>(*keys, *vals, *keys_alt, *vals_alt, n_dev, n_cap, p, ps.shift[p], ps.flip[p],
                                      ps.from_value[p], hist, counters + p, status, nullptr, nullptr,
                                      tiles_out ? prep.tiles : TileRangesOut());
    std::swap(*keys, *keys_alt);
    std::swap(*vals, *vals_alt);
  }
  GS_HIP(hipGetLastError());
}

}  // namespace

static BucketLayout bucket_layout(uint32_t n) {
  BucketLayout l{};
  l.fine = lookback_words(n, 1, BUCKETS);
  l.zero_words = l.fine + kFineBins;
  l.range = l.zero_words;
  l.words = l.range + kBucketRangeWords;
  return l;
}

size_t bucket_sort_lookback_words(uint32_t n) { return bucket_layout(n).zero_words; }

size_t bucket_sort_words(uint32_t n) { return bucket_layout(n).words; }

void bucket_sort(uint32_t **keys, uint32_t **vals, uint32_t **keys_alt, uint32_t **vals_alt, uint32_t n,
                 const uint2 *key_range, uint32_t parts, uint32_t *scratch, size_t scratch_words, bool scratch_clean,
                 uint32_t *overflow, hipStream_t st) {
  if (n <= 1) return;
  if (n > kBucketMaxKeys) throw Error(Code::Internal, "bucket sort of more than 2^19 keys");
  const BucketLayout l = bucket_layout(n);
  if (l.words > scratch_words) throw Error(Code::Internal, "bucket sort scratch too small");
  const uint32_t tiles = div_up(n, TILE);
  uint32_t *hist = scratch;  // unused by a bucket pass (it plans its digit counts itself)
  uint32_t *counter = hist + (size_t)STRIPES * MAX_PASSES * BUCKETS;
  uint32_t *status = counter + MAX_PASSES;
  uint32_t *range = scratch + l.range;
  if (!scratch_clean)
    k_os_clear<<<std::min<uint32_t>(div_up(l.zero_words, T), 1024), T, 0, st>>>(scratch, l.zero_words, 0, nullptr,
                                                                                  n);
  k_bucket_hist<<<tiles, T, 0, st>>>(*keys, n, key_range, parts, scratch + l.fine, range);
  k_os_pass<BUCKET_BITS, true><<<std::min(tiles, kPassGrid), T, 0, st>>>(*keys, *vals, *keys_alt, *vals_alt, nullptr, n,
                                                                   0, 0, 0, kRangeDigit, hist, counter, status,
                                                                   range, scratch + l.fine, TileRangesOut());
  std::swap(*keys, *keys_alt);
  std::swap(*vals, *vals_alt);
  k_bucket_local<<<BUCKETS - 1, T, 0, st>>>(*keys, *vals, range, range + kBucketFirst, range + kBucketLast,
                                            overflow);
  GS_HIP(hipGetLastError());
}

void onesweep_sort(uint32_t **keys, uint32_t **vals, uint32_t **keys_alt, uint32_t **vals_alt, uint32_t n_cap,
                   const uint32_t *n_dev, uint32_t bits, bool descending, uint32_t *scratch, size_t scratch_words,
                   hipStream_t st, int value_byte, uint32_t n_hint, SortPrep prep) {
  if (n_cap <= 1 || (bits == 0 && value_byte < 0)) return;
  if (n_cap > (1u << 30)) throw Error(Code::Unsupported, "sort length exceeds 2^30");
  const PassSet ps = radix_passes(bits, descending, value_byte);
  if (ps.digit_bits == MAX_DIGIT_BITS)
    sort_passes<MAX_DIGIT_BITS>(keys, vals, keys_alt, vals_alt, n_cap, n_dev, ps, scratch, scratch_words, st,
                                value_byte, n_hint, prep);
  else
    sort_passes<8>(keys, vals, keys_alt, vals_alt, n_cap, n_dev, ps, scratch, scratch_words, st, value_byte, n_hint,
                   prep);
}

}  // namespace gs
