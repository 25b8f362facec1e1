// Device-only helpers shared by the kernels of the plain and the batched
// paths: the Gaussian falloff of the raster, the wave digit match of the
// radix ranking, and the pair emission of binning.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.h"

namespace vkgs {

// exp(-q / 2) as v_exp_f32 (exp2) of q * (-log2(e) / 2).  Halving the float
// log2(e) is exact, so this equals exp2((-q / 2) * log2(e)) bit for bit while
// saving a multiply.
__device__ __forceinline__ float gauss_falloff(float q) {
  constexpr float kNegHalfLog2e = -0.5f * 1.44269504088896340736f;
  return __builtin_amdgcn_exp2f(q * kNegHalfLog2e);
}

// Load and store through a pointer known to address device (global) memory:
// for a pointer read from memory (a field of a table in HBM) the compiler
// cannot always infer the address space and falls back to flat accesses,
// which also count against the LDS wait counter.
__device__ __forceinline__ float ld_global(const float* p) { return *(const __attribute__((address_space(1))) float*)p; }
__device__ __forceinline__ float4 ld_global(const float4* p) {
  typedef float v4 __attribute__((ext_vector_type(4)));
  const v4 v = *(const __attribute__((address_space(1))) v4*)p;
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void st_global(float* p, float v) { *(__attribute__((address_space(1))) float*)p = v; }
__device__ __forceinline__ void st_global(float4* p, const float4& v) {
  typedef float v4 __attribute__((ext_vector_type(4)));
  *(__attribute__((address_space(1))) v4*)p = v4{v.x, v.y, v.z, v.w};
}

// Number of set bits of m below this lane.
__device__ __forceinline__ uint32_t mbcnt(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Mask of the lanes whose kBits-bit digit equals this lane's (valid lanes
// only): per bit one ballot and m &= ~(ballot ^ this lane's bit broadcast).
template <int kBits>
__device__ __forceinline__ uint64_t match_digit(uint32_t d, bool valid) {
  const uint64_t m0 = __ballot(valid);
  uint32_t lo = (uint32_t)m0, hi = (uint32_t)(m0 >> 32);
#pragma unroll
  for (int b = 0; b < kBits; ++b) {
    const uint32_t sel = (uint32_t)((int32_t)(d << (31 - b)) >> 31);  // bit b of d, all ones or zero
    const uint64_t bb = __ballot(sel != 0u);
    // m &= ~(bb ^ sel): v_bitop3 with m as the middle operand and the
    // symmetric bb / sel on either side, truth table 0x84
    lo = __builtin_amdgcn_bitop3_b32((uint32_t)bb, lo, sel, 0x84);
    hi = __builtin_amdgcn_bitop3_b32((uint32_t)(bb >> 32), hi, sel, 0x84);
  }
  return ((uint64_t)hi << 32) | lo;
}

// Splats covering more tiles than this are emitted by a whole wave.
constexpr uint32_t kEmitSerial = 8;

__device__ __forceinline__ uint32_t lane_u32(uint32_t x, uint32_t lane) {
  return (uint32_t)__builtin_amdgcn_readlane((int)x, (int)lane);
}

// Emits the pairs of one splat per lane: c tiles of rectangle r (tile x in
// r.x's low/high 16 bits, tile y in r.y's) as keys key_base + ty * tiles_x + tx
// at o, o + 1, ... in row-major tile order; pairs at or beyond max_pairs are
// dropped (the caller flags the overflow).  The value of the i-th pair is val,
// or, with kIndexValues, q0 + i: a second numbering of the pairs (the
// caller's q0 per splat), which a later pass uses to find each pair after the
// sort (clamped to max_pairs - 1 on overflow, so readers stay in bounds).  Lanes with up to kEmitSerial tiles write
// their own pairs; each larger splat is then written by all 64 lanes together,
// so one near-camera splat does not serialize a wave.  Every lane of the wave
// must call this (c = 0 for lanes without a splat).
template <bool kIndexValues = false>
__device__ __forceinline__ void emit_pairs_wave(uint32_t c, uint64_t o, TileRect r, uint32_t val, uint32_t key_base,
                                                uint32_t tiles_x, uint64_t max_pairs, uint32_t* __restrict__ keys,
                                                uint32_t* __restrict__ vals, uint64_t q0 = 0) {
  const bool big = c > kEmitSerial;
  if (c != 0 && !big) {
    const uint32_t tx0 = r.x & 0xffffu, tx1 = r.x >> 16, ty0 = r.y & 0xffffu, ty1 = r.y >> 16;
    uint64_t j = o, q = q0;
    for (uint32_t ty = ty0; ty <= ty1; ++ty)
      for (uint32_t tx = tx0; tx <= tx1; ++tx, ++j, ++q) {
        if (j < max_pairs) {
          keys[j] = key_base + ty * tiles_x + tx;
          vals[j] = kIndexValues ? (uint32_t)(q < max_pairs ? q : max_pairs - 1) : val;
        }
      }
  }
  uint64_t pending = __ballot(big);
  const uint32_t lane = __lane_id();
  while (pending) {
    const uint32_t src = (uint32_t)__builtin_ctzll(pending);
    pending &= pending - 1;
    const uint32_t bc = lane_u32(c, src), bval = lane_u32(val, src), bx = lane_u32(r.x, src);
    const uint32_t by = lane_u32(r.y, src), btx = lane_u32(tiles_x, src), bkey = lane_u32(key_base, src);
    const uint64_t bo = ((uint64_t)lane_u32((uint32_t)(o >> 32), src) << 32) | lane_u32((uint32_t)o, src);
    const uint64_t bq = ((uint64_t)lane_u32((uint32_t)(q0 >> 32), src) << 32) | lane_u32((uint32_t)q0, src);
    const uint32_t tx0 = bx & 0xffffu, w = (bx >> 16) - tx0 + 1, ty0 = by & 0xffffu;
    const uint32_t first = bkey + ty0 * btx + tx0;
    for (uint32_t i = lane; i < bc; i += kWave) {
      const uint32_t row = i / w, col = i - row * w;
      const uint64_t j = bo + i;
      if (j < max_pairs) {
        keys[j] = first + row * btx + col;
        vals[j] = kIndexValues ? (uint32_t)(bq + i < max_pairs ? bq + i : max_pairs - 1) : bval;
      }
    }
  }
}

}  // namespace vkgs
