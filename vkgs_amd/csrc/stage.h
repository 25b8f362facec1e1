// Per-tile restaging of a raster record, shared by the plain and the
// differentiable rasters (device code).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.h"

namespace vkgs {

__device__ __forceinline__ uint32_t pack16(int lo, int hi) { return (uint32_t)lo | ((uint32_t)hi << 16); }

// A raster record (cx, cy, i00, i01), (i10, i11, opacity, x rect), (r, g, b,
// y rect) restaged for one tile (see kFrameScale): the scaled inverse
// rot_scale rows with their offsets at the corner (X0, Y0) of the tile's
// 64 x 64 pixel block, -log2(opacity), the colour and the pixel rectangle:
//   s0 = (i'00, i'01, o'x, -log2 opacity), s1 = (i'10, i'11, o'y, x rect),
//   s2 = (r, g, b, y rect).
// A lane with no record stages one that covers no pixel: offsets `inert`
// (+inf, or a large finite value where a zero weight multiplies the frame
// coordinates, which must stay finite) and an empty rectangle.
__device__ __forceinline__ void stage_record(const float4& p0, const float4& p1, const float4& p2, bool valid,
                                             float X0, float Y0, float4* s0, float4* s1, float4* s2,
                                             float inert = __builtin_inff()) {
  if (valid) {
    const float i00 = p0.z * kFrameScale, i01 = p0.w * kFrameScale;
    const float i10 = p1.x * kFrameScale, i11 = p1.y * kFrameScale;
    const float dx = X0 - p0.x, dy = Y0 - p0.y;
    *s0 = make_float4(i00, i01, fmaf(i00, dx, i01 * dy), -__builtin_amdgcn_logf(p1.z));
    *s1 = make_float4(i10, i11, fmaf(i10, dx, i11 * dy), p1.w);
    *s2 = p2;
  } else {
    const float empty = __uint_as_float(pack16(1, 0));
    *s0 = make_float4(0.f, 0.f, inert, inert);
    *s1 = make_float4(0.f, 0.f, inert, empty);
    *s2 = make_float4(0.f, 0.f, 0.f, empty);
  }
}

}  // namespace vkgs
