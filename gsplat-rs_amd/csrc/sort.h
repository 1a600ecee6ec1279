// sort.h -- onesweep radix sort launcher.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace gs {

constexpr uint32_t kSortThreads = 256;
constexpr uint32_t kSortItems = 16;
constexpr uint32_t kSortTile = kSortThreads * kSortItems;

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
                   hipStream_t stream, int value_byte = -1, uint32_t n_hint = 0);

}  // namespace gs
