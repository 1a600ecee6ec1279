// sort_noalloc.h -- the standalone sort over a fixed device array (no
// device allocation).
#pragma once

#include <cstdint>

namespace gs {

// Longest sort the array holds.
constexpr uint32_t kNoAllocSortMaxLen = 1u << 20;

// The stable descending sort of `values` by `keys` (host buffers, n values),
// as gsplat_sort_values_by_keys, with every device buffer in the array of
// sort_noalloc.hip: no hipMalloc, hipFree or stream creation. Throws
// Unsupported for n > kNoAllocSortMaxLen.
void sort_values_by_keys_noalloc(const uint32_t *keys, uint32_t *values, uint32_t n);

}  // namespace gs
