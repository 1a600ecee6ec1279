// sort.h -- onesweep radix sort launcher.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace gs {

constexpr uint32_t kSortThreads = 512;
constexpr uint32_t kSortItems = 8;
constexpr uint32_t kSortTile = kSortThreads * kSortItems;
constexpr uint32_t kSortMaxPasses = 5;
constexpr uint32_t kSortStripes = 8;  // striped copies of the digit histograms

// The digit passes of one sort: pass p orders by the digit_bits-bit digit
// at bit shift[p] of (key ^ flip[p]), or of the value when from_value[p].
// Digits are 8 bits, except that 9-bit keys without a value pass sort in
// one 9-bit pass.
struct RadixPasses {
  uint32_t count;
  uint32_t digit_bits;
  uint32_t shift[kSortMaxPasses];
  uint32_t flip[kSortMaxPasses];
  uint32_t from_value[kSortMaxPasses];
};
RadixPasses radix_passes(uint32_t bits, bool descending, int value_byte);

// Digit histograms of a look-back sort, as k_os_hist accumulates them: a
// kernel that produces the keys may add them instead (sort with hist_ready):
// hist[((stripe * kSortMaxPasses) + p) * 2^digit_bits + digit], stripe =
// workgroup mod kSortStripes.

// True when a sort of up to n_cap elements, n_hint expected (0: n_cap),
// runs look-back passes (which need clean scratch and digit histograms)
// rather than reduce-then-scan passes; onesweep_sort decides the same way.
// rts_min_tiles: sorts of at least this many 4096-item tiles use
// reduce-then-scan (0: GSPLAT_SORT_RTS_MIN_TILES or 256).
bool radix_uses_lookback(uint32_t n_cap, uint32_t n_hint = 0, uint32_t rts_min_tiles = 0);

// u32 words at the start of the scratch that a look-back sort of n_cap
// elements with passes `ps` needs zeroed (histograms, tile counters, status
// words); they are left dirty after the sort.
size_t radix_lookback_words(uint32_t n_cap, const RadixPasses &ps);

// u32 words at the start of the scratch that a reduce-then-scan sort of up
// to n_cap elements with passes `ps` may write (digit counts, the live count,
// the scan ticket and its chunk status words, which carry flag bits).
size_t radix_rts_words(uint32_t n_cap, const RadixPasses &ps);

// Binning-tile output of a one-pass look-back sort whose keys are tile
// indices below `count` (the pair sort of one view, digit histograms from
// the emission): the workgroup that sorts the first tile also writes tile
// t's range [begin, end) of the sorted keys into ranges[t] (clamped to the
// key count) and the raster's heaviest-first order, order[class * count +
// place] = t and order_counts[class] = class size, from the histogram, so
// that no kernel has to find the range boundaries in the sorted keys.
struct TileRangesOut {
  uint2 *ranges = nullptr;
  uint32_t count = 0;
  uint32_t *order = nullptr;
  uint32_t *order_counts = nullptr;
};

struct SortPrep {
  bool scratch_clean = false;  // the look-back words are zero: skip the clear
  bool hist_ready = false;     // the digit histograms are accumulated: skip k_os_hist
  uint32_t rts_min_tiles = 0;  // mode threshold, as in radix_uses_lookback
  // Segments (device array of `segs` element offsets, ascending, each a
  // multiple of kSortTile, the first 0; the last segment ends at n_cap).
  // When the value pass orders by the segment index (the element's value
  // byte names its segment, as for a batch of views), reduce-then-scan
  // passes sort every segment in place instead and leave the value pass
  // out: the same result in one pass fewer. Look-back passes ignore them.
  const uint32_t *seg_start = nullptr;
  uint32_t segs = 0;
  // With hist_ready, a sort of one look-back pass writes these (else unused).
  TileRangesOut tiles;
};

// u32 words of scratch needed to sort up to n_cap keys on `bits` key bits
// (plus an optional value pass).
size_t onesweep_status_words(uint32_t n_cap, uint32_t bits);

// Stable LSD radix sort of (keys, vals) on the low `bits` bits, 8 bits per
// pass, descending when `descending`; when value_byte >= 0 a final stable
// pass orders the pairs ascending by that byte of the value. The element
// count is *n_dev when n_dev is non-null (read on the device, clamped to
// n_cap) else n_cap; buffers hold n_cap elements. The result ends in
// keys[0]/vals[0] (pointers swapped with the alternates as needed).
// n_hint (with n_dev) is the expected count, which sizes the launch grids;
// any count up to n_cap is still sorted.
void onesweep_sort(uint32_t **keys, uint32_t **vals, uint32_t **keys_alt, uint32_t **vals_alt, uint32_t n_cap,
                   const uint32_t *n_dev, uint32_t bits, bool descending, uint32_t *scratch, size_t scratch_words,
                   hipStream_t stream, int value_byte = -1, uint32_t n_hint = 0, SortPrep prep = SortPrep());

// Bucket sort of short key arrays (n <= kBucketMaxKeys): the same result as
// onesweep_sort(bits 32, descending) -- keys descending, equal keys in their
// input order -- in three launches instead of five, one pass over the keys
// instead of four:
//   k_bucket_hist   reduces the key producer's per-workgroup ranges (smallest
//                   non-zero key, largest key) to [kmin, kmax], picks the
//                   shift s with (kmax - kmin) >> s < 8192 and counts the
//                   fine bins (kmax - key) >> s of the non-zero keys;
//   k_os_pass<8>    each workgroup first plans the buckets (consecutive fine
//                   bins, about max(1024, visible / 254) keys each; key 0 in
//                   the last of 256), then one stable look-back pass by
//                   bucket;
//   k_bucket_local  one workgroup per bucket sorts its run in LDS by
//                   kmax - key (stable 8-bit radix passes over the bits its
//                   fine bins span).
// A bucket of more than kSortTile keys (a fine bin holding thousands of
// equal or nearly equal keys) is left unsorted and *overflow gets bit 0: the
// caller must then sort again with onesweep_sort.
// The result ends in keys[0]/vals[0] (swapped with the alternates).
// key_range[0..parts) are the producer's ranges (launch_project). The
// scratch's first bucket_sort_lookback_words(n) words must be zero unless
// scratch_clean is false (then they are cleared first); the sort leaves
// bucket_sort_words(n) words dirty.
constexpr uint32_t kBucketMaxKeys = 1u << 19;
size_t bucket_sort_lookback_words(uint32_t n);
size_t bucket_sort_words(uint32_t n);
void bucket_sort(uint32_t **keys, uint32_t **vals, uint32_t **keys_alt, uint32_t **vals_alt, uint32_t n,
                 const uint2 *key_range, uint32_t parts, uint32_t *scratch, size_t scratch_words, bool scratch_clean,
                 uint32_t *overflow, hipStream_t stream);

}  // namespace gs
