// sort_noalloc.hip -- the standalone stable descending sort with no device
// allocation, for callers that cannot allocate (tight serving loops, code
// that must not call hipMalloc).
//
// Keys, values, their radix ping-pong alternates and the sort's status words
// live in one file-scope device array, `scratch`, of kScratchWords words
// (18 MiB), which each GPU the library runs on holds until the process ends.
// Host code never passes `scratch` itself to a kernel or copy (its name in
// host code is a stand-in, not a device address): the first call on a device
// resolves the array's device address with hipGetSymbolAddress, once per
// device, and every later use goes through that pointer.
//
// Calls share their GPU's array. A process-wide lock serialises the calls of
// this process, and each call has finished on the device when it returns; two
// processes have arrays of their own.

#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <string>

#include "common.h"
#include "sort.h"
#include "sort_noalloc.h"

namespace gs {
namespace {

constexpr size_t kStatusWords = size_t(1) << 19;  // look-back or reduce-then-scan state of a 2^20-key sort
constexpr size_t kScratchWords = 4 * (size_t)kNoAllocSortMaxLen + kStatusWords;

}  // namespace
}  // namespace gs

// keys | values | keys_alt | values_alt (kNoAllocSortMaxLen words each) | status words
static __device__ __align__(256) uint32_t scratch[gs::kScratchWords];

namespace gs {
namespace {

#define GS_HIP_NA(call)                                                                \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess) {                                                            \
      (void)hipGetLastError();                                                         \
      throw Error(Code::Internal, std::string(#call) + ": " + hipGetErrorString(e_)); \
    }                                                                                  \
  } while (0)

std::mutex g_lock;                     // serialises calls (they share the array)
std::map<int, uint32_t *> g_scratch;   // device -> address of `scratch` on it

// Device address of `scratch` on the current device, resolved on the first
// call there (caller holds g_lock).
uint32_t *device_scratch() {
  int dev = 0;
  GS_HIP_NA(hipGetDevice(&dev));
  auto it = g_scratch.find(dev);
  if (it != g_scratch.end()) return it->second;
  void *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) {
    (void)hipGetLastError();
    throw Error(Code::Internal, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  }
  g_scratch[dev] = static_cast<uint32_t *>(ptr);
  return static_cast<uint32_t *>(ptr);
}

}  // namespace

void sort_values_by_keys_noalloc(const uint32_t *keys, uint32_t *values, uint32_t n) {
  if (n <= 1) return;
  if (n > kNoAllocSortMaxLen)
    throw Error(Code::Unsupported, "length " + std::to_string(n) + " exceeds the no-allocation limit of " +
                                       std::to_string(kNoAllocSortMaxLen));
  if (onesweep_status_words(n, 32) > kStatusWords) throw Error(Code::Internal, "status words exceed the reserve");
  std::lock_guard<std::mutex> g(g_lock);
  uint32_t *base = device_scratch();
  uint32_t *k = base, *v = base + kNoAllocSortMaxLen;
  uint32_t *ka = base + 2 * (size_t)kNoAllocSortMaxLen, *va = base + 3 * (size_t)kNoAllocSortMaxLen;
  uint32_t *status = base + 4 * (size_t)kNoAllocSortMaxLen;
  const hipStream_t st = nullptr;  // the null stream: no stream is created
  GS_HIP_NA(hipMemcpyAsync(k, keys, (size_t)n * 4, hipMemcpyHostToDevice, st));
  GS_HIP_NA(hipMemcpyAsync(v, values, (size_t)n * 4, hipMemcpyHostToDevice, st));
  // The status words keep the last call's contents: the sort clears what it
  // needs with a kernel of its own (scratch_clean is false).
  onesweep_sort(&k, &v, &ka, &va, n, nullptr, 32, /*descending=*/true, status, kStatusWords, st);
  GS_HIP_NA(hipMemcpyAsync(values, v, (size_t)n * 4, hipMemcpyDeviceToHost, st));
  GS_HIP_NA(hipStreamSynchronize(st));
}

}  // namespace gs
