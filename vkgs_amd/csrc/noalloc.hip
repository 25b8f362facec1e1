// vkgs_sort_kv_noalloc: the stable key/value radix sort of vkgs_sort_kv with
// its scratch (alternate key and value buffers, per-block digit histograms,
// digit totals) in a static device buffer instead of caller storage, for
// callers that cannot allocate device memory (code under hipGraph capture,
// serving loops, harnesses that forbid hipMalloc).
//
// The buffer holds the scratch of a VKGS_SORT_NOALLOC_MAX_N-pair sort.  Every
// GPU the library runs on holds its own copy until the process ends; its
// address is looked up once per device on the first call made with that
// device current and kept per device.  Every call on one GPU uses that GPU's
// copy, so calls that overlap on one GPU (two streams, two threads) race: the
// caller serializes them.  Nothing in the scratch has to start at zero: every
// pass writes each histogram word it later reads, and the alternate buffers
// are written before they are read.
#include <mutex>

#include "common.h"
#include "sort.h"

namespace vkgs {
namespace {

// sort_plan(VKGS_SORT_NOALLOC_MAX_N, 4): alternate keys and values (4 B each
// per pair), 256 histogram words per block for at most 1024 blocks, 256 digit
// totals, one word per block, each region aligned to 256 bytes.
constexpr size_t kScratchBytes = (size_t)VKGS_SORT_NOALLOC_MAX_N * 8 + (size_t)256 * 1024 * 4 + (size_t)256 * 4 +
                                 (size_t)1024 * 4 + (size_t)5 * 256;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
void* g_scratch[kMaxDevices] = {};

// Device address of the current device's copy of `scratch` (nullptr on error).
void* device_scratch() {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
    g_scratch[dev] = p;
  }
  return g_scratch[dev];
}

}  // namespace
}  // namespace vkgs

using namespace vkgs;

extern "C" int vkgs_sort_kv_noalloc(uint32_t max_n, const uint32_t* d_count, uint32_t* d_key, uint32_t* d_value,
                                    uint32_t key_bits, void* stream) {
  if (max_n == 0) return VKGS_OK;
  if (!d_count || !d_key || !d_value || key_bits == 0 || key_bits > 32) return VKGS_ERR_INVALID_ARG;
  if (max_n > VKGS_SORT_NOALLOC_MAX_N) return VKGS_ERR_CAPACITY;
  const SortPlan p = sort_plan(max_n);
  if (p.bytes > kScratchBytes) return VKGS_ERR_CAPACITY;
  void* storage = device_scratch();
  if (!storage) return VKGS_ERR_HIP;
  VKGS_HIP_CHECK(sort_kv_launch(p, d_count, d_key, d_value, key_bits, storage, static_cast<hipStream_t>(stream)));
  return VKGS_OK;
}
