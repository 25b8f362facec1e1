// Shared device/host helpers for the gfx950 splat pipeline.
//
// Every translation unit that includes this file is compiled with
// -ffp-contract=off: a*b+c rounds twice unless written as fmaf().  The matrix
// helpers sum left to right, ((a0*b0 + a1*b1) + a2*b2) + a3*b3, which is the
// order the CPU oracle uses; that is what makes keys, visible counts, tile
// rectangles and sorted order bit-exact between the two.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include "vkgs_hip.h"

#define VKHD __host__ __device__ __forceinline__

namespace vkgs {

constexpr int kWave = 64;

// Per-frame constants shared by rank, projection and the fused frame.
// Column-major 4x4 matrices as in the Camera UBO (uniforms.h:10-15).
struct FrameConst {
  float P[16];        // projection (Vulkan clip space)
  float V[16];        // view
  float M[16];        // model (push constant, engine.cc:1185-1186)
  float PVM[16];      // (P*V)*M, the rank.comp:32 product
  float eye_model[4]; // inverse(model) * (eye, 1), divided by w (projection.comp:85-86)
  float inv_w2, inv_h2; // (1/W)/W and (1/H)/H, the low-pass terms (projection.comp:115-116)
  uint32_t width, height;
  uint32_t tile_w, tile_h, tiles_x, tiles_y;
  uint32_t triangles;  // records in the triangle-raster layout (VKGS_RASTER_TRIANGLES)
};

VKHD void m4_mul(const float* a, const float* b, float* c) {
  for (int col = 0; col < 4; ++col)
    for (int row = 0; row < 4; ++row) {
      float s = a[0 * 4 + row] * b[col * 4 + 0];
      s = s + a[1 * 4 + row] * b[col * 4 + 1];
      s = s + a[2 * 4 + row] * b[col * 4 + 2];
      s = s + a[3 * 4 + row] * b[col * 4 + 3];
      c[col * 4 + row] = s;
    }
}

VKHD void m4_mulv(const float* m, const float* v, float* o) {
  for (int row = 0; row < 4; ++row) {
    float s = m[0 * 4 + row] * v[0];
    s = s + m[1 * 4 + row] * v[1];
    s = s + m[2 * 4 + row] * v[2];
    s = s + m[3 * 4 + row] * v[3];
    o[row] = s;
  }
}

VKHD void m3_mul(const float* a, const float* b, float* c) {
  float t[9];
  for (int col = 0; col < 3; ++col)
    for (int row = 0; row < 3; ++row) {
      float s = a[0 * 3 + row] * b[col * 3 + 0];
      s = s + a[1 * 3 + row] * b[col * 3 + 1];
      s = s + a[2 * 3 + row] * b[col * 3 + 2];
      t[col * 3 + row] = s;
    }
  for (int i = 0; i < 9; ++i) c[i] = t[i];
}

VKHD void m3_transpose(const float* a, float* t) {
  for (int col = 0; col < 3; ++col)
    for (int row = 0; row < 3; ++row) t[col * 3 + row] = a[row * 3 + col];
}

VKHD void m3_from_m4(const float* m4, float* m3) {
  for (int col = 0; col < 3; ++col)
    for (int row = 0; row < 3; ++row) m3[col * 3 + row] = m4[col * 4 + row];
}

// S <- (A*S)*A^T, the shader's `cov3d = A * cov3d * transpose(A)`.
VKHD void m3_sandwich(const float* a, float* s) {
  float as[9], at[9];
  m3_mul(a, s, as);
  m3_transpose(a, at);
  m3_mul(as, at, s);
}

VKHD void m2_mul(const float* a, const float* b, float* c) {
  float t[4];
  for (int col = 0; col < 2; ++col)
    for (int row = 0; row < 2; ++row) {
      float s = a[0 * 2 + row] * b[col * 2 + 0];
      s = s + a[1 * 2 + row] * b[col * 2 + 1];
      t[col * 2 + row] = s;
    }
  for (int i = 0; i < 4; ++i) c[i] = t[i];
}

// rank.comp:31-37: NDC of the centre through the combined PVM and the
// centre-only frustum test.  *one_minus_z receives 1 - z_ndc (rank.comp:39).
VKHD bool rank_visible(const float* pvm, float x, float y, float z, float* one_minus_z) {
  float v[4] = {x, y, z, 1.f};
  float p[4];
  m4_mulv(pvm, v, p);
  const float nx = p[0] / p[3], ny = p[1] / p[3], nz = p[2] / p[3];
  *one_minus_z = 1.f - nz;
  return fabsf(nx) <= 1.f && fabsf(ny) <= 1.f && nz >= 0.f && nz <= 1.f;
}

VKHD float glsl_max(float x, float y) { return x < y ? y : x; }

// Packed tile rectangle: tx0 | tx1 << 16, ty0 | ty1 << 16.
struct TileRect {
  uint32_t x, y;
};

// A tile rectangle packed into 32 bits: x0, x1 in bx bits each, then y0, y1
// in by bits each.  Usable when 2 * (bx + by) <= 32 (e.g. 200 x 113 tiles of
// 8x8 pixels at 1600x900: 8 + 7 bits).  The empty rectangle is x0 = 1,
// x1 = 0 (count 0).
struct RectBits {
  uint32_t bx, by;
};

VKHD uint32_t coord_bits(uint32_t tiles) {  // bits for values in [0, tiles), at least 1
  uint32_t b = 1;
  while (b < 32 && (1u << b) < tiles) ++b;
  return b;
}

VKHD bool rect_bits(uint32_t tiles_x, uint32_t tiles_y, RectBits* rb) {
  rb->bx = coord_bits(tiles_x);
  rb->by = coord_bits(tiles_y);
  return 2 * (rb->bx + rb->by) <= 32;
}

VKHD uint32_t pack_rect(uint32_t x0, uint32_t x1, uint32_t y0, uint32_t y1, RectBits rb) {
  return x0 | (x1 << rb.bx) | (y0 << (2 * rb.bx)) | (y1 << (2 * rb.bx + rb.by));
}

// (x0 | x1 << 16, y0 | y1 << 16) and the tile count of a packed rectangle.
VKHD TileRect unpack_rect(uint32_t p, RectBits rb, uint32_t* count) {
  const uint32_t mx = (1u << rb.bx) - 1u, my = (1u << rb.by) - 1u;
  const uint32_t x0 = p & mx, x1 = (p >> rb.bx) & mx;
  const uint32_t y0 = (p >> (2 * rb.bx)) & my, y1 = (p >> (2 * rb.bx + rb.by)) & my;
  *count = (x1 >= x0 && y1 >= y0) ? (x1 - x0 + 1) * (y1 - y0 + 1) : 0u;
  TileRect r;
  r.x = x0 | (x1 << 16);
  r.y = y0 | (y1 << 16);
  return r;
}

// The raster tests coverage and evaluates the falloff in the eigenframe
// scaled by kFrameScale = sqrt(log2(e) / 2): with u' = kFrameScale * u the
// reference's alpha = opacity * exp(-|u|^2 / 2) (splat.frag:9-10) is
// exp2(log2(opacity) - |u'|^2), and the square |u_x|, |u_y| <= 3 of the quad
// (splat.vert:21-25) is max(|u'_x|, |u'_y|) <= kCoverScaled.  u' of a pixel is
// formed from the corner (X0, Y0) of its 64 x 64 pixel block:
//   i' = kFrameScale * (row of the inverse rot_scale),
//   o' = fma(i'_0, X0 - cx, i'_1 * (Y0 - cy)),
//   u' = fma(i'_0, px + 0.5 - X0, fma(i'_1, py + 0.5 - Y0, o')),
// so that per pixel it is two FMAs, the terms stay below 64 |i'| in magnitude,
// and the result does not depend on the tile shape (every tile lies in one
// block).  The oracle evaluates the same operations.
constexpr float kFrameScale = 0.849321800288019f;
constexpr float kCoverScaled = 3.0f * kFrameScale;
constexpr int kRasterBlock = 64;

// Raster set-up of one instance (splat.vert:10-26 restated in framebuffer
// pixels): centre, inverse of the pixel-space rot_scale, colour, and the
// conservative pixel rectangle of the +-3 sigma quad.  Returns false when the
// quad produces no fragment (depth test, non-finite geometry, off screen).
struct Splat2D {
  float cx, cy, i00, i01, i10, i11, r, g, b, o;
  int x0, y0, x1, y1;
};

VKHD bool setup_splat(const float* inst, uint32_t width, uint32_t height, Splat2D* s) {
  const float cx = inst[0], cy = inst[1], cz = inst[2];
  const float m00 = inst[4], m10 = inst[5], m01 = inst[6], m11 = inst[7];
  s->r = inst[8];
  s->g = inst[9];
  s->b = inst[10];
  s->o = inst[11];
  // depth test LESS against the cleared 1.0, depth range [0,1]
  // (engine.cc:298-299, graphics_pipeline.cc:77-81)
  if (!(cz >= 0.f && cz < 1.f)) return false;
  const float hw = 0.5f * (float)width, hh = 0.5f * (float)height;
  s->cx = (cx + 1.f) * hw;
  s->cy = (cy + 1.f) * hh;
  const float a00 = hw * m00, a01 = hw * m01, a10 = hh * m10, a11 = hh * m11;
  const float det = a00 * a11 - a01 * a10;
  const float inv = 1.f / det;
  s->i00 = a11 * inv;
  s->i01 = -(a01 * inv);
  s->i10 = -(a10 * inv);
  s->i11 = a00 * inv;
  const float hx = 3.f * (fabsf(a00) + fabsf(a01));
  const float hy = 3.f * (fabsf(a10) + fabsf(a11));
  const float ex = hx * 1.00390625f + 0.015625f;
  const float ey = hy * 1.00390625f + 0.015625f;
  if (!(isfinite(s->cx) && isfinite(s->cy) && isfinite(s->i00) && isfinite(s->i01) && isfinite(s->i10) &&
        isfinite(s->i11) && isfinite(ex) && isfinite(ey)))
    return false;
  float lox = ceilf(s->cx - ex - 0.5f), hix = floorf(s->cx + ex - 0.5f);
  float loy = ceilf(s->cy - ey - 0.5f), hiy = floorf(s->cy + ey - 0.5f);
  if (lox < 0.f) lox = 0.f;
  if (loy < 0.f) loy = 0.f;
  if (hix > (float)(width - 1)) hix = (float)(width - 1);
  if (hiy > (float)(height - 1)) hiy = (float)(height - 1);
  if (!(lox <= hix && loy <= hiy)) return false;
  s->x0 = (int)lox;
  s->x1 = (int)hix;
  s->y0 = (int)loy;
  s->y1 = (int)hiy;
  return true;
}

// The 48-byte raster record of one instance (f = its 12 floats), set up as s
// (setup_splat; an empty rectangle when nothing is drawn).  Analytic layout:
// (cx, cy, i00, i01), (i10, i11, opacity, x0 | x1 << 16), (r, g, b, y0 | y1 << 16)
// in framebuffer pixels.  Triangle layout (the quad is rebuilt per tile from
// the instance's own numbers): (ndc x, ndc y, rot_scale m00, m10),
// (m01, m11, opacity, x rect), (r, g, b, y rect).
VKHD void raster_record(const float* f, const Splat2D& s, bool triangles, float4* out) {
  const uint32_t rx = (uint32_t)s.x0 | ((uint32_t)s.x1 << 16), ry = (uint32_t)s.y0 | ((uint32_t)s.y1 << 16);
  float fx, fy;
  memcpy(&fx, &rx, 4);
  memcpy(&fy, &ry, 4);
  if (triangles) {
    out[0] = make_float4(f[0], f[1], f[4], f[5]);
    out[1] = make_float4(f[6], f[7], s.o, fx);
  } else {
    out[0] = make_float4(s.cx, s.cy, s.i00, s.i01);
    out[1] = make_float4(s.i10, s.i11, s.o, fx);
  }
  out[2] = make_float4(s.r, s.g, s.b, fy);
}

inline uint32_t ceil_div(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }
inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Fills a FrameConst from the Camera UBO and the model matrix (host).
void make_frame_const(const vkgs_camera* cam, const float* model, FrameConst* fc);

}  // namespace vkgs

#define VKGS_HIP_CHECK(expr)                       \
  do {                                             \
    hipError_t e_ = (expr);                        \
    if (e_ != hipSuccess) return VKGS_ERR_HIP;     \
  } while (0)
