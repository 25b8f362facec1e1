// LSD radix sort of (u32 key, u32 or u64 value) pairs and an exclusive u32 scan,
// both with the element count read from device memory, so a hipGraph can be
// captured once for a capacity and replayed for any live count.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace vkgs {

struct SortPlan {
  uint32_t max_n = 0;
  uint32_t blocks = 1;  // fixed grid; each block sorts a contiguous chunk
  size_t off_alt_key = 0, off_alt_val = 0, off_hist = 0, off_digit_total = 0, off_block_max = 0;
  size_t bytes = 0;
};

// Storage plan for up to max_n pairs with val_bytes-byte values (4 or 8).
SortPlan sort_plan(uint32_t max_n, uint32_t val_bytes = 4);

// Stable ascending sort of min(*d_count, max_n) pairs (d_count == nullptr:
// max_n) on key bits [0, key_bits), result in (keys, vals).  Passes take 8
// bits each from bit 0 up, except pass narrow_pass, which takes the
// remainder when key_bits is not a multiple of 8.  d_key_max (optional, a
// device word; used by four-pass sorts): the first two passes store the
// largest key there, and the last two, as a pair, return at once when that
// maximum has no bit at or above the third pass's shift, so such a sort
// spends passes on the bits its keys use (the result lands in place either
// way).  first_bits (1-7, optional): the first pass takes that many bits, the
// middle ones 8 and the last the remainder (narrow_pass then unused).
hipError_t sort_kv_launch(const SortPlan& plan, const uint32_t* d_count, uint32_t* keys, uint32_t* vals,
                          uint32_t key_bits, void* storage, hipStream_t stream, uint32_t narrow_pass = 0,
                          uint32_t* d_key_max = nullptr, uint32_t first_bits = 0);

// Depth keys are the bits of fl(1 - z), z in [0, 1]: every such value is a
// multiple of 2^-24, so a distant splat's low mantissa bits are zero and
// digits below bit 16 crowd into few values, which a pass writes in long
// runs; bits 16-21 are spread evenly.  Their 30-bit sort puts the 6-bit
// digit there (passes of 8, 8, 6, 8 bits), which keeps that pass's runs 4x
// longer (6M splats: 0.168 -> 0.161 ms).
constexpr uint32_t kDepthKeyBits = 30;
constexpr uint32_t kDepthNarrowPass = 2;

// The same order as an integer: k = fl(1 - z) * 2^24, exact because every
// fl(1 - z) with z in [0, 1] is a multiple of 2^-24 (so k in [0, 2^24] orders
// and ties as the float bits do), in 25 bits, which leaves room above it for
// a view index in a 32-bit key.
constexpr uint32_t kDepthRankBits = 25;
constexpr uint32_t kDepthRankNarrowPass = 0;
__host__ __device__ inline uint32_t depth_rank_key(float one_minus_z) {
  return (uint32_t)(one_minus_z * 16777216.f);
}
// The float bits of fl(1 - z) back from k (exact: k <= 2^24).
__host__ __device__ inline uint32_t depth_key_bits(uint32_t k) {
  const float f = (float)k * (1.f / 16777216.f);
  uint32_t b;
  __builtin_memcpy(&b, &f, 4);
  return b;
}
// The renderer's depth sort on k alone: passes of 7, 8, 8, 2 bits with a
// d_key_max word, so that a frame whose k stay below 2^15 (every visible
// splat farther than about 512 near-plane distances) sorts in two passes
// instead of four.  k's low bits are evenly spread: the 7-bit first digit
// writes runs twice as long as an 8-bit one.
constexpr uint32_t kDepthRankFirstBits = 7;
// Radix passes of a key_bits-bit sort (8 bits per pass).
uint32_t sort_passes(uint32_t key_bits);
// The storage's alternate (key, value) buffers of a 4-byte-value plan.
void sort_alt_buffers(const SortPlan& plan, void* storage, uint32_t** keys, uint32_t** vals);
// Same sort with the unsorted pairs in the alternate buffers; the result is
// in (keys, vals).  With an odd pass count this needs no final copy, so a
// producer that writes its pairs to the alternate buffers saves one.
hipError_t sort_kv_from_alt_launch(const SortPlan& plan, const uint32_t* d_count, uint32_t* keys, uint32_t* vals,
                                   uint32_t key_bits, void* storage, hipStream_t stream);
// Same with 64-bit values (plan from sort_plan(max_n, 8)).
hipError_t sort_kv64_launch(const SortPlan& plan, const uint32_t* d_count, uint32_t* keys, uint64_t* vals,
                            uint32_t key_bits, void* storage, hipStream_t stream, uint32_t narrow_pass = 0,
                            uint32_t* d_key_max = nullptr, uint32_t first_bits = 0);

struct ScanPlan {
  uint32_t max_n = 0;
  uint32_t blocks = 1;
  size_t off_partial = 0;
  size_t bytes = 0;
};

ScanPlan scan_plan(uint32_t max_n);

// out[i] = sum(in[0..i)) for i < min(*d_count, max_n) (d_count == nullptr:
// i < max_n); *d_total = sum of all.  in and out may alias.
hipError_t scan_exclusive_launch(const ScanPlan& plan, const uint32_t* in, uint32_t* out, const uint32_t* d_count,
                                 uint32_t* d_total, void* storage, hipStream_t stream);

// Same over the tile counts of the packed rectangles in the high words of
// 64-bit sort values (slot | rect << 32).
struct RectBits;
hipError_t scan_rect_counts_launch(const ScanPlan& plan, const uint64_t* val64, const RectBits& rb, uint32_t* out,
                                   const uint32_t* d_count, uint32_t* d_total, void* storage, hipStream_t stream);


// Fills n words with a value (graph-safe replacement for hipMemsetAsync).
hipError_t fill_u32_launch(uint32_t* dst, uint32_t value, size_t n, hipStream_t stream);

// Writes up to 16 words passed by value (graph-safe replacement for a small
// host-to-device copy).
struct Words16 {
  uint32_t w[16];
};
hipError_t set_words_launch(uint32_t* dst, const Words16& words, uint32_t n, hipStream_t stream);

// Stream-ordered host-to-device copy of `bytes` (a multiple of 4) carried in
// the arguments of kUploadChunkWords-word kernel launches: the host buffer can
// change as soon as the call returns, no host sync and no memcpy node is
// needed, and a captured hipGraph replays the same bytes.  For small tables.
constexpr uint32_t kUploadChunkWords = 768;
hipError_t upload_launch(void* dst, const void* src, size_t bytes, hipStream_t stream);

// Sticky overflow record: if *status (a list's overflow flag) is set,
// *sticky = 1 and *sticky_count = max(*sticky_count, *count) (the list's full length).
hipError_t fold_status_launch(const uint32_t* status, const uint32_t* count, uint32_t* sticky,
                              uint32_t* sticky_count, hipStream_t stream);

}  // namespace vkgs
