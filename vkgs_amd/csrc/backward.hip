// Differentiable, batched splat path for gfx950 (forward and backward).
//
// Forward: per (view, splat) item the activation of parse_ply.comp:46-96
// (fp32 SH, no fp16 rounding) and the projection of projection.comp:80-179,
// ordered compaction with the raster record written at the item's slot, depth
// sort, binning into global tiles, tile sort, and a raster that also stores,
// per pixel, the transmittance after the last processed splat and how many
// list entries it processed.
//
// Backward: the raster backward re-walks each tile list back to front from
// that stored state, recomputing alpha per entry and the transmittance before
// it as T_i = T_{i+1} / (1 - alpha_i); per entry the 9 gradient words (colour
// 3, opacity 1, conic 3, centre 2) are summed over the tile's pixels (DPP
// half-row sums, then one LDS slot per half row) and stored as that (tile,
// splat) pair's partials at the pair's slot-order index q.  The pairs of a
// slot have consecutive q, so a reduction sums each slot's run into its 9
// words, streaming in slot order; no global atomics are used and the result
// does not depend on scheduling.  The projection
// backward then runs one lane per scene splat, chains the words of every view
// that saw the splat to its raw parameters and writes each gradient once.
//
// alpha is clamped to kAlphaMax so the division above stays bounded; this is
// the only departure of this path from the reference's blend.
#pragma clang fp contract(off)

#include "common.h"
#include "diff.h"
#include "device.h"
#include "kernels.h"
#include "stage.h"

namespace vkgs {
namespace {

constexpr float kAlphaMax = 0.999f;


// ------------------------------------------------------ forward projection ---

// Rotation (row-major) of the normalized quaternion q = (x, y, z, w), as built
// by parse_ply.comp:54-72 (the shader stores it column-major).
__device__ __forceinline__ void quat_rot(const float q[4], float R[9]) {
  const float x = q[0], y = q[1], z = q[2], w = q[3];
  const float xx = x * x, yy = y * y, zz = z * z, xy = x * y, xz = x * z, yz = y * z, wx = w * x, wy = w * y,
              wz = w * z;
  R[0] = 1.f - 2.f * (yy + zz);
  R[1] = 2.f * (xy - wz);
  R[2] = 2.f * (xz + wy);
  R[3] = 2.f * (xy + wz);
  R[4] = 1.f - 2.f * (xx + zz);
  R[5] = 2.f * (yz - wx);
  R[6] = 2.f * (xz - wy);
  R[7] = 2.f * (yz + wx);
  R[8] = 1.f - 2.f * (xx + yy);
}

// Normalized (x, y, z, w) from the PLY order (w, x, y, z); returns |q|.
__device__ __forceinline__ float load_quat(const DiffScene& sc, uint32_t id, float q[4]) {
  const float4 q4 = ld_global(reinterpret_cast<const float4*>(sc.quats) + id);
  const float w = q4.x, x = q4.y, y = q4.z, z = q4.w;
  const float len = sqrtf(((x * x + y * y) + z * z) + w * w);
  q[0] = x / len;
  q[1] = y / len;
  q[2] = z / len;
  q[3] = w / len;
  return len;
}

// cov3d in the shader's column-major layout (parse_ply.comp:46-78).
__device__ __forceinline__ void activate_cov(const DiffScene& sc, uint32_t id, float S[9]) {
  float q[4];
  load_quat(sc, id, q);
  const float s0 = expf(sc.log_scales[3 * (size_t)id]), s1 = expf(sc.log_scales[3 * (size_t)id + 1]);
  const float s2 = expf(sc.log_scales[3 * (size_t)id + 2]);
  float Rm[9];
  quat_rot(q, Rm);
  float rot[9];  // column-major
  m3_transpose(Rm, rot);
  const float ss[9] = {s0 * s0, 0.f, 0.f, 0.f, s1 * s1, 0.f, 0.f, 0.f, s2 * s2};
  float rs[9], rt[9];
  m3_mul(rot, ss, rs);
  m3_transpose(rot, rt);
  m3_mul(rs, rt, S);
}

__device__ __forceinline__ void sh_basis(float x, float y, float z, float b[16]) {
  const float C0 = 0.28209479177387814f, C1 = 0.4886025119029199f, C20 = 1.0925484305920792f,
              C21 = 0.31539156525252005f, C22 = 0.5462742152960396f, C30 = 0.5900435899266435f,
              C31 = 2.890611442640554f, C32 = 0.4570457994644658f, C33 = 0.3731763325901154f,
              C34 = 1.445305721320277f;
  const float xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
  b[0] = C0;
  b[1] = -C1 * y;
  b[2] = C1 * z;
  b[3] = -C1 * x;
  b[4] = C20 * xy;
  b[5] = -C20 * yz;
  b[6] = C21 * ((2.f * zz - xx) - yy);
  b[7] = -C20 * xz;
  b[8] = C22 * (xx - yy);
  b[9] = (-C30 * y) * (3.f * xx - yy);
  b[10] = (C31 * xy) * z;
  b[11] = (-C32 * y) * ((4.f * zz - xx) - yy);
  b[12] = (C33 * z) * ((2.f * zz - 3.f * xx) - 3.f * yy);
  b[13] = (-C32 * x) * ((4.f * zz - xx) - yy);
  b[14] = (C34 * z) * (xx - yy);
  b[15] = (-C30 * x) * (xx - 3.f * yy);
}

// View-independent inputs of the projection of one splat: mean, activated
// cov3d (parse_ply.comp:46-78, column-major), fp32 SH and opacity.
struct SplatInputs {
  float p[3];
  float S[9];
  float sh[48];
  float opacity;
};

__device__ __forceinline__ void load_inputs(const DiffScene& sc, uint32_t id, SplatInputs& in) {
  activate_cov(sc, id, in.S);
  for (int i = 0; i < 3; ++i) in.p[i] = sc.means[3 * (size_t)id + i];
  const float4* sh4 = reinterpret_cast<const float4*>(sc.sh + 48 * (size_t)id);
  for (int j = 0; j < 12; ++j) {
    const float4 u = sh4[j];
    in.sh[4 * j] = u.x;
    in.sh[4 * j + 1] = u.y;
    in.sh[4 * j + 2] = u.z;
    in.sh[4 * j + 3] = u.w;
  }
  in.opacity = 1.f / (1.f + expf(-sc.opacity[id]));
}

// projection.comp:80-179 for one view; fills the 48-byte instance.  The
// differentiable path is held to float64, not to the reference's rounding,
// so this form contracts multiply-adds into FMAs, uses the hardware
// reciprocal, square root and reciprocal square root (about 1 ulp), maps the
// mean and the covariance to view space with the one product V M formed on
// the host, and forms only the 2x2 block of J S' J^T that the shader reads
// (its third row of J never reaches the output).  Same quantities as the
// reference's op order to float rounding; NaN for the same degenerate splats.
__device__ void diff_project(const DiffView& view, const SplatInputs& in, float4 out[3]) {
#pragma clang fp contract(fast)
  const FrameConst& fc = view.fc;
  const float* W = view.vm;  // column-major
  float pv[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) pv[r] = W[r] * in.p[0] + W[4 + r] * in.p[1] + W[8 + r] * in.p[2] + W[12 + r];
  const float x = pv[0], y = pv[1], z = pv[2];
  // S' = W3 S W3^T; S symmetric, so S[3k + c] = S(k, c); WS(r, c) = sum_k W3(r, k) S(k, c)
  float WS[9];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) WS[3 * r + c] = W[r] * in.S[c] + W[4 + r] * in.S[3 + c] + W[8 + r] * in.S[6 + c];
  auto sp = [&](int r, int c) { return WS[3 * r] * W[c] + WS[3 * r + 1] * W[4 + c] + WS[3 * r + 2] * W[8 + c]; };
  const float s00 = sp(0, 0), s01 = sp(0, 1), s02 = sp(0, 2), s11 = sp(1, 1), s12 = sp(1, 2), s22 = sp(2, 2);
  // rows 0 and 1 of J: (-1/z, 0, x/z^2) and (0, -1/z, y/z^2)
  const float iz = __builtin_amdgcn_rcpf(z);
  const float ja = -iz, jb = (x * iz) * iz, jc = (y * iz) * iz;
  const float c00 = ja * (ja * s00 + 2.f * jb * s02) + jb * jb * s22;
  const float c01 = ja * (ja * s01 + jc * s02 + jb * s12) + jb * jc * s22;
  const float c11 = ja * (ja * s11 + 2.f * jc * s12) + jc * jc * s22;
  // cov2 = Ps S2 Ps with Ps = [[P0, P4], [P1, P5]]
  const float P0 = fc.P[0], P1 = fc.P[1], P4 = fc.P[4], P5 = fc.P[5];
  const float T00 = P0 * c00 + P4 * c01, T01 = P0 * c01 + P4 * c11;
  const float T10 = P1 * c00 + P5 * c01, T11 = P1 * c01 + P5 * c11;
  const float a = (T00 * P0 + T01 * P1) + fc.inv_w2;
  const float b = (T10 * P4 + T11 * P5) + fc.inv_h2;
  const float c = T00 * P4 + T01 * P5;
  const float D = __builtin_amdgcn_sqrtf((a - b) * (a - b) + (4.f * c) * c);
  const float s0 = __builtin_amdgcn_sqrtf(0.5f * ((a + b) + D));
  const float s1 = __builtin_amdgcn_sqrtf(0.5f * ((a + b) - D));
  const float iD = __builtin_amdgcn_rcpf(D);
  const float sin2t = (2.f * c) * iD;
  const float cos2t = (a - b) * iD;
  float ct = (1.f + cos2t) * 0.5f;
  float st = (1.f - cos2t) * 0.5f;
  ct = ct < 0.f ? 0.f : ct;
  st = st < 0.f ? 0.f : st;
  float cos_t = __builtin_amdgcn_sqrtf(ct);
  float sin_t = copysignf(__builtin_amdgcn_sqrtf(st), sin2t);
  if (sin2t != sin2t || cos2t != cos2t) {
    cos_t = __builtin_nanf("");
    sin_t = __builtin_nanf("");
  }
  float t4[4];
  m4_mulv(fc.P, pv, t4);
  const float iw = __builtin_amdgcn_rcpf(t4[3]);
  const float nx = t4[0] * iw, ny = t4[1] * iw, nz = t4[2] * iw;
  const float d0 = in.p[0] - fc.eye_model[0], d1 = in.p[1] - fc.eye_model[1], d2 = in.p[2] - fc.eye_model[2];
  const float idl = __builtin_amdgcn_rsqf(d0 * d0 + d1 * d1 + d2 * d2);
  float basis[16];
  sh_basis(d0 * idl, d1 * idl, d2 * idl, basis);
  float color[3];
#pragma unroll
  for (int ch = 0; ch < 3; ++ch) {
    float acc = 0.f;
#pragma unroll
    for (int j = 0; j < 16; ++j) acc += basis[j] * in.sh[ch * 16 + j];
    color[ch] = glsl_max(acc + 0.5f, 0.f);
  }
  out[0] = make_float4(nx, ny, nz, 0.f);
  out[1] = make_float4(s0 * cos_t, s0 * sin_t, (-s1) * sin_t, s1 * cos_t);
  out[2] = make_float4(color[0], color[1], color[2], in.opacity);
}

// Largest i < n with bases[i] <= x (bases ascending, x < bases[n]); empty
// ranges (equal bases) resolve to the last of them, the one holding x.
__device__ __forceinline__ uint32_t range_of(const uint32_t* __restrict__ bases, uint32_t n, uint32_t x) {
  uint32_t lo = 0, hi = n;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (bases[mid] <= x)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// ------------------------------------------------------ processing order ---

// Floats as unsigned words that order like the floats (for min / max).
__device__ __forceinline__ uint32_t float_order(float f) {
  const uint32_t b = __float_as_uint(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float order_float(uint32_t o) {
  return __uint_as_float((o & 0x80000000u) ? (o & 0x7fffffffu) : ~o);
}

// bbox_part[6 sb ..]: min x, y, z and max x, y, z (order words) of scene block sb.
__global__ __launch_bounds__(kCompactBlock) void diff_bbox_kernel(DiffLaunch L, uint32_t* __restrict__ bbox_part) {
  __shared__ uint32_t s_b[6][kCompactBlock / kWave];
  const uint32_t sb = blockIdx.x;
  const DiffScene& sc = L.scenes[range_of(L.scene_blocks, L.n_scenes, sb)];
  const uint32_t li = (sb - sc.block_base) * kCompactBlock + threadIdx.x;
  uint32_t v[6] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0u, 0u, 0u};
  if (li < sc.n)
    for (int a = 0; a < 3; ++a) v[a] = v[3 + a] = float_order(sc.means[3 * (size_t)li + a]);
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1)
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      v[a] = min(v[a], (uint32_t)__shfl_xor((int)v[a], d, kWave));
      v[3 + a] = max(v[3 + a], (uint32_t)__shfl_xor((int)v[3 + a], d, kWave));
    }
  if (__lane_id() == 0)
    for (int a = 0; a < 6; ++a) s_b[a][threadIdx.x / kWave] = v[a];
  __syncthreads();
  if (threadIdx.x < 6) {
    const uint32_t a = threadIdx.x;
    uint32_t r = s_b[a][0];
    for (uint32_t i = 1; i < kCompactBlock / kWave; ++i) r = a < 3 ? min(r, s_b[a][i]) : max(r, s_b[a][i]);
    bbox_part[6 * (size_t)sb + a] = r;
  }
}

// bbox[6 s ..] from the scene's block partials; one workgroup per scene.
__global__ __launch_bounds__(kCompactBlock) void diff_bbox_reduce_kernel(DiffLaunch L,
                                                                         const uint32_t* __restrict__ bbox_part,
                                                                         uint32_t* __restrict__ bbox) {
  __shared__ uint32_t s_r[6][kCompactBlock];
  const uint32_t s = blockIdx.x;
  const uint32_t b0 = L.scene_blocks[s], b1 = L.scene_blocks[s + 1];
  uint32_t v[6] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0u, 0u, 0u};
  for (uint32_t b = b0 + threadIdx.x; b < b1; b += kCompactBlock)
    for (int a = 0; a < 6; ++a) v[a] = a < 3 ? min(v[a], bbox_part[6 * (size_t)b + a]) : max(v[a], bbox_part[6 * (size_t)b + a]);
  for (int a = 0; a < 6; ++a) s_r[a][threadIdx.x] = v[a];
  __syncthreads();
  if (threadIdx.x < 6) {
    const uint32_t a = threadIdx.x;
    uint32_t r = s_r[a][0];
    for (uint32_t i = 1; i < kCompactBlock; ++i) r = a < 3 ? min(r, s_r[a][i]) : max(r, s_r[a][i]);
    bbox[6 * s + a] = r;
  }
}

__device__ __forceinline__ uint32_t spread3(uint32_t x) {  // bit i of x to bit 3 i (x < 1024)
  x = (x | (x << 16)) & 0x030000ffu;
  x = (x | (x << 8)) & 0x0300f00fu;
  x = (x | (x << 4)) & 0x030c30c3u;
  x = (x | (x << 2)) & 0x09249249u;
  return x;
}

__global__ __launch_bounds__(kCompactBlock) void diff_morton_kernel(DiffLaunch L, const uint32_t* __restrict__ bbox,
                                                                    uint32_t scene_shift, uint32_t axis_bits,
                                                                    uint32_t* __restrict__ keys,
                                                                    uint32_t* __restrict__ vals) {
  const uint32_t sb = blockIdx.x;
  const uint32_t si = range_of(L.scene_blocks, L.n_scenes, sb);
  const DiffScene& sc = L.scenes[si];
  const uint32_t li = (sb - sc.block_base) * kCompactBlock + threadIdx.x;
  const size_t g = (size_t)sb * kCompactBlock + threadIdx.x;
  uint32_t key = 0xffffffffu;
  if (li < sc.n) {
    const float top = (float)((1u << axis_bits) - 1u);
    uint32_t code = 0;
    for (int a = 0; a < 3; ++a) {
      const float lo = order_float(bbox[6 * si + a]), hi = order_float(bbox[6 * si + 3 + a]);
      const float t = hi > lo ? (sc.means[3 * (size_t)li + a] - lo) / (hi - lo) : 0.f;
      const float qf = fminf(fmaxf(t * top, 0.f), top);  // NaN positions go to 0
      code |= spread3((uint32_t)qf) << a;
    }
    key = (si << scene_shift) | code;
  }
  keys[g] = key;
  vals[g] = li;
}

// --------------------------------------------------------------- forward ---

// Visibility of every (view, splat) item, one lane per scene splat visiting
// the views of its scene: masks[w] = the ballot of wave w (the 64 splats of a
// wave, for one view) and counts[w] = its popcount, where w = kMaskWords *
// (view.block_base + block) + wave.  Wave counts are scanned into slot offsets.
__global__ __launch_bounds__(kCompactBlock) void diff_count_kernel(DiffLaunch L, uint32_t* __restrict__ counts,
                                                                   uint64_t* __restrict__ masks) {
  const uint32_t sb = blockIdx.x;
  const DiffScene& sc = L.scenes[range_of(L.scene_blocks, L.n_scenes, sb)];
  const uint32_t lb = sb - sc.block_base;
  const uint32_t li = lb * kCompactBlock + threadIdx.x;
  const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const bool valid = li < sc.n;
  const uint32_t id = valid && sc.perm ? sc.perm[li] : li;
  float x = 0.f, y = 0.f, z = 0.f;
  if (valid) {
    x = sc.means[3 * (size_t)id];
    y = sc.means[3 * (size_t)id + 1];
    z = sc.means[3 * (size_t)id + 2];
  }
  for (uint32_t vi = sc.view_begin; vi < sc.view_end; ++vi) {
    const DiffView& view = L.views[L.scene_views[vi]];
    float omz;
    const bool vis = valid && rank_visible(view.fc.PVM, x, y, z, &omz);
    const uint64_t m = __ballot(vis);
    if (__lane_id() == 0) {
      const size_t wi = (size_t)(view.block_base + lb) * kMaskWords + w;
      masks[wi] = m;
      counts[wi] = (uint32_t)__popcll(m);
    }
  }
}

// Projection of the visible items into raster records at their slots (as the
// plain fused path does), one lane per scene splat: the activated covariance,
// SH and opacity are formed once and reused for every view of the scene.
// The tile rectangle goes either to rc (with the count and the view, val =
// slot) or, with kPacked, into the depth sort's 64-bit value slot | packed
// rectangle << 32.
template <bool kPacked>
__global__ __launch_bounds__(kCompactBlock) void diff_write_kernel(DiffLaunch L, const uint32_t* __restrict__ offsets,
                                                                   const uint64_t* __restrict__ masks,
                                                                   float4* __restrict__ rec, uint4* __restrict__ rc,
                                                                   uint32_t* __restrict__ scnt,
                                                                   uint32_t* __restrict__ key,
                                                                   uint32_t* __restrict__ val,
                                                                   uint64_t* __restrict__ val64, RectBits rb) {
  const uint32_t sb = blockIdx.x;
  const DiffScene& sc = L.scenes[range_of(L.scene_blocks, L.n_scenes, sb)];
  const uint32_t lb = sb - sc.block_base;
  const uint32_t li = lb * kCompactBlock + threadIdx.x;
  const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const uint32_t lane = __lane_id();
  // a wave with no visible item in any view skips the parameter loads
  bool any = false;
  for (uint32_t vi = sc.view_begin; vi < sc.view_end && !any; ++vi)
    any = masks[(size_t)(L.views[L.scene_views[vi]].block_base + lb) * kMaskWords + w] != 0;
  if (!any) return;
  SplatInputs in;
  if (li < sc.n) load_inputs(sc, sc.perm ? sc.perm[li] : li, in);
  for (uint32_t vi = sc.view_begin; vi < sc.view_end; ++vi) {
    const uint32_t v = L.scene_views[vi];
    const DiffView& view = L.views[v];
    const size_t wi = (size_t)(view.block_base + lb) * kMaskWords + w;
    const uint64_t m = masks[wi];
    if (!((m >> lane) & 1ull)) continue;
    const uint32_t slot = offsets[wi] + mbcnt(m);
    float omz;
    (void)rank_visible(view.fc.PVM, in.p[0], in.p[1], in.p[2], &omz);
    float4 o[3];
    diff_project(view, in, o);
    const float f[12] = {o[0].x, o[0].y, o[0].z, o[0].w, o[1].x, o[1].y, o[1].z, o[1].w, o[2].x, o[2].y, o[2].z, o[2].w};
    Splat2D sp = {};
    uint4 r = make_uint4(0u, 0u, 0u, v);
    uint32_t packed = 0;
    if (kPacked) packed = pack_rect(1u, 0u, 0u, 0u, rb);  // empty
    if (setup_splat(f, view.fc.width, view.fc.height, &sp)) {
      const uint32_t tx0 = (uint32_t)sp.x0 >> L.tile_shift_x, tx1 = (uint32_t)sp.x1 >> L.tile_shift_x;
      const uint32_t ty0 = (uint32_t)sp.y0 >> L.tile_shift_y, ty1 = (uint32_t)sp.y1 >> L.tile_shift_y;
      r.x = tx0 | (tx1 << 16);
      r.y = ty0 | (ty1 << 16);
      r.z = (tx1 - tx0 + 1) * (ty1 - ty0 + 1);
      if (kPacked) packed = pack_rect(tx0, tx1, ty0, ty1, rb);
    } else {  // empty pixel box: every raster tile test rejects it
      sp.x0 = 1;
      sp.x1 = 0;
      sp.y0 = 1;
      sp.y1 = 0;
    }
    rec[3 * (size_t)slot + 0] = make_float4(sp.cx, sp.cy, sp.i00, sp.i01);
    rec[3 * (size_t)slot + 1] = make_float4(sp.i10, sp.i11, sp.o, __uint_as_float(pack16(sp.x0, sp.x1)));
    rec[3 * (size_t)slot + 2] = make_float4(sp.r, sp.g, sp.b, __uint_as_float(pack16(sp.y0, sp.y1)));
    if (kPacked) {
      val64[slot] = (uint64_t)slot | ((uint64_t)packed << 32);
    } else {
      rc[slot] = r;
      val[slot] = slot;
    }
    scnt[slot] = r.z;
    key[slot] = L.local_tile_bits ? (v << kDepthRankBits) | depth_rank_key(omz) : __float_as_uint(omz);
  }
}

// First slot of every view (slots are view-major): vslot[v] = the scanned
// count of the view's first compaction wave; vslot[n_views] = all visible.
__global__ void diff_view_slots_kernel(DiffLaunch L, const uint32_t* __restrict__ offsets,
                                       const uint32_t* __restrict__ d_count, uint32_t* __restrict__ vslot) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < L.n_views)
    vslot[v] = L.view_blocks[v] < L.n_blocks ? offsets[(size_t)L.view_blocks[v] * kMaskWords] : *d_count;
  else if (v == L.n_views)
    vslot[v] = *d_count;
}

__global__ __launch_bounds__(256) void diff_bin_gather_kernel(const uint4* __restrict__ rc,
                                                              const uint32_t* __restrict__ order,
                                                              const uint32_t* __restrict__ d_count, uint32_t max_n,
                                                              TileRect* __restrict__ rect, uint32_t* __restrict__ cnt,
                                                              uint32_t* __restrict__ kview) {
  const uint32_t n = min(*d_count, max_n);
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    const uint4 u = rc[order[k]];
    rect[k] = TileRect{u.x, u.y};
    cnt[k] = u.z;
    kview[k] = u.w;
  }
}

// Per-slot runs of pairs up to this length are handled by one lane in the
// slot fill and the gradient reduction; longer runs by the whole wave.
constexpr uint32_t kRunSerial = 16;

// pslot[q] = slot for q in soffs[slot] .. soffs[slot] + scnt[slot] - 1, in
// slot order (so the writes stream): runs of up to kRunSerial words per lane,
// longer runs by the whole wave.
__global__ __launch_bounds__(256) void diff_slot_fill_kernel(const uint32_t* __restrict__ soffs,
                                                             const uint32_t* __restrict__ scnt,
                                                             const uint32_t* __restrict__ d_count, uint32_t max_n,
                                                             uint64_t max_pairs, uint32_t* __restrict__ pslot) {
  const uint32_t n = min(*d_count, max_n);
  const uint32_t lane = __lane_id();
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t base = wave * kWave; base < n; base += stride) {  // wave-uniform
    const uint32_t k = base + lane;
    uint32_t c = 0;
    uint64_t o = 0;
    if (k < n) {
      o = soffs[k];
      c = scnt[k];
      if (o + c > max_pairs) c = o < max_pairs ? (uint32_t)(max_pairs - o) : 0u;
    }
    const bool big = c > kRunSerial;
    if (!big)
      for (uint32_t i = 0; i < c; ++i) pslot[o + i] = k;
    uint64_t pending = __ballot(big);
    while (pending) {
      const uint32_t src = (uint32_t)__builtin_ctzll(pending);
      pending &= pending - 1;
      const uint32_t bc = lane_u32(c, src);
      const uint64_t bo = ((uint64_t)lane_u32((uint32_t)(o >> 32), src) << 32) | lane_u32((uint32_t)o, src);
      for (uint32_t i = lane; i < bc; i += kWave) pslot[bo + i] = base + src;
    }
  }
}

// Pairs (global tile, q), where q numbers the pairs in slot order (soffs =
// exclusive scan of the per-slot tile counts): the sort carries q, so the
// backward can place each sorted pair's partials where the slot's pairs sit
// contiguously, in slot order.
__global__ __launch_bounds__(256) void diff_bin_emit_kernel(DiffLaunch L, const TileRect* __restrict__ rect,
                                                            const uint32_t* __restrict__ cnt,
                                                            const uint32_t* __restrict__ offs,
                                                            const uint32_t* __restrict__ kview,
                                                            const uint32_t* __restrict__ order,
                                                            const uint32_t* __restrict__ d_count, uint32_t max_n,
                                                            uint64_t max_pairs, uint32_t* __restrict__ keys,
                                                            uint32_t* __restrict__ vals,
                                                            const uint32_t* __restrict__ soffs,
                                                            uint32_t* __restrict__ status) {
  const uint32_t n = min(*d_count, max_n);
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const uint32_t stride = gridDim.x * blockDim.x;
  // wave-uniform loop: emit_pairs_wave needs every lane
  for (uint32_t base = wave * kWave; base < n; base += stride) {
    const uint32_t k = base + __lane_id();
    uint32_t c = 0, slot = 0, key_base = 0, tiles_x = 0;
    uint64_t o = 0, q0 = 0;
    TileRect r = {0u, 0u};
    if (k < n) {
      c = cnt[k];
      if (c != 0) {
        o = offs[k];
        r = rect[k];
        slot = order[k];
        q0 = soffs[slot];
        const uint32_t v = kview[k];
        const DiffView& view = L.views[v];
        key_base = L.local_tile_bits ? v << L.local_tile_bits : view.tile_base;
        tiles_x = view.fc.tiles_x;
        if (o + c > max_pairs) atomicOr(status, 1u);
      }
    }
    emit_pairs_wave<true>(c, o, r, slot, key_base, tiles_x, max_pairs, keys, vals, q0);
  }
}

// Same from the depth-sorted values slot | packed rect << 32: the view of a
// slot from the view-major slot bases vslot, the pair offsets from the scan
// of the packed rectangles' tile counts.
__global__ __launch_bounds__(256) void diff_bin_emit_packed_kernel(DiffLaunch L, const uint64_t* __restrict__ val64,
                                                                   const uint32_t* __restrict__ offs,
                                                                   const uint32_t* __restrict__ vslot, RectBits rb,
                                                                   const uint32_t* __restrict__ d_count,
                                                                   uint32_t max_n, uint64_t max_pairs,
                                                                   uint32_t* __restrict__ keys,
                                                                   uint32_t* __restrict__ vals,
                                                                   const uint32_t* __restrict__ soffs,
                                                                   uint32_t* __restrict__ status) {
  const uint32_t n = min(*d_count, max_n);
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const uint32_t stride = gridDim.x * blockDim.x;
  // wave-uniform loop: emit_pairs_wave needs every lane
  for (uint32_t base = wave * kWave; base < n; base += stride) {
    const uint32_t k = base + __lane_id();
    uint32_t c = 0, slot = 0, key_base = 0, tiles_x = 0;
    uint64_t o = 0, q0 = 0;
    TileRect r = {0u, 0u};
    if (k < n) {
      const uint64_t w = val64[k];
      r = unpack_rect((uint32_t)(w >> 32), rb, &c);
      if (c != 0) {
        slot = (uint32_t)w;
        o = offs[k];
        q0 = soffs[slot];
        const uint32_t v = range_of(vslot, L.n_views, slot);
        const DiffView& view = L.views[v];
        key_base = L.local_tile_bits ? v << L.local_tile_bits : view.tile_base;
        tiles_x = view.fc.tiles_x;
        if (o + c > max_pairs) atomicOr(status, 1u);
      }
    }
    emit_pairs_wave<true>(c, o, r, slot, key_base, tiles_x, max_pairs, keys, vals, q0);
  }
}

// ------------------------------------------------- differentiable raster ---
//
// Records are restaged per tile as in the plain raster (stage_record: the
// inverse rot_scale rows scaled by kFrameScale with their offsets at the
// tile's 64 x 64 block corner, -log2(opacity)), so a pixel's frame
// coordinates u' are two FMAs each and alpha = min(exp2(-|u'|^2 - log2(1 /
// opacity)), kAlphaMax) over the square max(|u'_x|, |u'_y|) <= kCoverScaled.
// The forward and the backward stage the same records at the same corner, so
// the backward recomputes every alpha bit for bit.  Each pixel stops on its
// own once T < t_min.

// One forward (pixel, entry) step; idx = the entry's 1-based place in the
// tile list.  A pixel composites while T >= t_min (T only decreases, so once
// below it stays below); while it does, last = idx, so a pixel that stops
// keeps the entry that stopped it.  The state is re-derived from T on every
// step rather than carried as a flag, which keeps it in a lane mask.
__device__ __forceinline__ void diff_step_at(float ux, float uy, const float4& a, const float4& c, uint32_t idx,
                                             float t_min, float& T, float& cr, float& cg, float& cb, float& aa,
                                             uint32_t& last) {
  const bool alive = T >= t_min;
  const float m = fmaxf(fabsf(ux), fabsf(uy));
  const float e = __builtin_amdgcn_exp2f(-fmaf(ux, ux, fmaf(uy, uy, a.w)));
  const float alpha = ((m <= kCoverScaled) & alive) ? fminf(e, kAlphaMax) : 0.f;
  last = alive ? idx : last;
  const float ta = T * alpha;
  cr = fmaf(c.x, ta, cr);
  cg = fmaf(c.y, ta, cg);
  cb = fmaf(c.z, ta, cb);
  aa = fmaf(alpha, ta, aa);
  T = T - ta;
}

__device__ __forceinline__ void diff_step(const float4& a, const float4& b, const float4& c, float lx, float ly,
                                          uint32_t idx, float t_min, float& T, float& cr, float& cg, float& cb,
                                          float& aa, uint32_t& last) {
  const bool alive = T >= t_min;
  const float ux = fmaf(a.x, lx, fmaf(a.y, ly, a.z));
  const float uy = fmaf(b.x, lx, fmaf(b.y, ly, b.z));
  const float m = fmaxf(fabsf(ux), fabsf(uy));
  const float e = __builtin_amdgcn_exp2f(-fmaf(ux, ux, fmaf(uy, uy, a.w)));
  const float alpha = ((m <= kCoverScaled) & alive) ? fminf(e, kAlphaMax) : 0.f;
  last = alive ? idx : last;
  const float ta = T * alpha;
  cr = fmaf(c.x, ta, cr);
  cg = fmaf(c.y, ta, cg);
  cb = fmaf(c.z, ta, cb);
  aa = fmaf(alpha, ta, aa);
  T = T - ta;
}

// Forward raster of one global tile; optionally keeps the per-pixel state the
// backward needs (final T and the entries each pixel composited).
template <int TW, int TH>
__global__ __launch_bounds__(TW* TH) void diff_raster_kernel(DiffLaunch L, const float4* __restrict__ rec,
                                                             const uint32_t* __restrict__ vals,
                                                             const uint32_t* __restrict__ ranges,
                                                             const uint32_t* __restrict__ pslot,
                                                             const uint32_t* __restrict__ tile_order, int keep_state) {
  constexpr int kN = TW * TH;
  constexpr int kRowsPerWave = kWave / TW;
  constexpr bool kOneWave = kN == kWave;  // the wave covers the tile: every binned record touches it
  __shared__ float4 s_rec[2][3][kN];

  const uint32_t gtile = tile_order ? tile_order[blockIdx.x] : blockIdx.x;
  const DiffView& view = L.views[range_of(L.view_tiles, L.n_views, gtile)];
  const uint32_t tile = gtile - view.tile_base;
  const uint32_t W = view.fc.width, H = view.fc.height;
  const uint32_t tx = tile % view.fc.tiles_x, ty = tile / view.fc.tiles_x;
  const int t = threadIdx.x;
  const int px = (int)tx * TW + t % TW, py = (int)ty * TH + t / TW;
  const bool inside = px < (int)W && py < (int)H;
  const int col0 = __builtin_amdgcn_readfirstlane((int)tx * TW), col1 = col0 + TW - 1;
  const int row0 = __builtin_amdgcn_readfirstlane((int)ty * TH + (t / kWave) * kRowsPerWave);
  const int row1 = row0 + kRowsPerWave - 1;
  const int bx0 = col0 & ~(kRasterBlock - 1), by0 = __builtin_amdgcn_readfirstlane(((int)ty * TH) & ~(kRasterBlock - 1));
  const float X0 = (float)bx0, Y0 = (float)by0;
  const float lx = (float)(px - bx0) + 0.5f, ly = (float)(py - by0) + 0.5f;
  const uint32_t start = ranges[2 * gtile], end = ranges[2 * gtile + 1];
  const uint32_t items = end - start;
  const uint32_t batches = (items + kN - 1) / kN;
  const float t_min = L.t_min;

  float4 p0 = make_float4(0.f, 0.f, 0.f, 0.f), p1 = p0, p2 = p0;
  if ((uint32_t)t < items) {
    const uint32_t k = pslot[vals[end - 1 - t]];
    p0 = rec[3 * (size_t)k];
    p1 = rec[3 * (size_t)k + 1];
    p2 = rec[3 * (size_t)k + 2];
  }
  // alive = "still compositing" (T only decreases; pixels outside the image
  // start at T = 0 and store nothing)
  float T = inside ? 1.f : 0.f, cr = 0.f, cg = 0.f, cb = 0.f, aa = 0.f;
  uint32_t last = 0, walked = 0;
  for (uint32_t b = 0; b < batches; ++b) {
    const int buf = b & 1;
    stage_record(p0, p1, p2, b * kN + t < items, X0, Y0, &s_rec[buf][0][t], &s_rec[buf][1][t], &s_rec[buf][2][t]);
    if (__syncthreads_count(T >= t_min) == 0) break;
    const uint32_t next = (b + 1) * kN + t;
    if (next < items) {
      const uint32_t k = pslot[vals[end - 1 - next]];
      p0 = rec[3 * (size_t)k];
      p1 = rec[3 * (size_t)k + 1];
      p2 = rec[3 * (size_t)k + 2];
    }
    const int cnt = __builtin_amdgcn_readfirstlane((int)min((uint32_t)kN, items - b * kN));
    if (__ballot(T >= t_min) == 0) continue;
    if (kOneWave) {
      // groups of 4 entries between exit tests (the batch's tail is staged as
      // records that cover nothing; a stopped pixel composites nothing more
      // and keeps its last)
      static_assert(kN % 4 == 0, "batches hold whole groups");
      for (int e = 0; e < cnt;) {
#pragma unroll
        for (int j = 0; j < 4; ++j)
          diff_step(s_rec[buf][0][e + j], s_rec[buf][1][e + j], s_rec[buf][2][e + j], lx, ly, b * kN + e + j + 1,
                    t_min, T, cr, cg, cb, aa, last);
        walked += (uint32_t)min(4, cnt - e);
        e += 4;
        if (__ballot(T >= t_min) == 0) break;
      }
      continue;
    }
    for (int e = 0; e < cnt; ++e) {
      const float4 a = s_rec[buf][0][e], bb = s_rec[buf][1][e], c = s_rec[buf][2][e];
      if (!kOneWave) {
        const uint32_t bx = __builtin_amdgcn_readfirstlane(__float_as_uint(bb.w));
        const uint32_t by = __builtin_amdgcn_readfirstlane(__float_as_uint(c.w));
        const int bx0r = (int)(bx & 0xffffu), bx1r = (int)(bx >> 16);
        const int by0r = (int)(by & 0xffffu), by1r = (int)(by >> 16);
        if (by1r < row0 || by0r > row1 || bx1r < col0 || bx0r > col1) continue;
      }
      ++walked;
      diff_step(a, bb, c, lx, ly, b * kN + e + 1, t_min, T, cr, cg, cb, aa, last);
      if (__ballot(T >= t_min) == 0) break;
    }
  }
  if (T >= t_min) last = items;
  if (inside) {
    const size_t pix = (size_t)py * W + px;
    view.out[pix] = make_float4(cr, cg, cb, aa + T);
    if (keep_state) {
      view.final_t[pix] = T;
      view.last[pix] = last;
    }
  }
  if (L.work && t % kWave == 0) L.work[(size_t)blockIdx.x * (kN / kWave) + t / kWave] = walked;
}

// -------------------------------------------------------------- backward ---

template <int kCtrl>
__device__ __forceinline__ float dpp_add(float x) {
  return x + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), kCtrl, 0xf, 0xf, false));
}

// Sum of x over each row of 16 lanes (DPP row); every lane must be active
// and ends holding its row's sum.
__device__ __forceinline__ float row_sum16(float x) {
  x = dpp_add<0xb1>(x);   // quad_perm [1,0,3,2]: pairs
  x = dpp_add<0x4e>(x);   // quad_perm [2,3,0,1]: quads
  x = dpp_add<0x124>(x);  // row_ror:4
  return dpp_add<0x128>(x);  // row_ror:8
}

// Half-row sums: after the quad steps and row_ror:4 every lane holds its quad
// plus a neighbouring quad; lanes 7 and 15 of a row then hold two disjoint
// halves whose sum is the row's (whichever way the rotation runs).
__device__ __forceinline__ float half_row_sum(float x) {
  x = dpp_add<0xb1>(x);
  x = dpp_add<0x4e>(x);
  return dpp_add<0x124>(x);
}

// Finite offsets for a backward lane with no record: its coverage test fails
// and every gradient term it forms is an exact zero times a finite value.
constexpr float kInertFinite = 1e18f;
// 1 / kFrameScale^2 = 2 ln 2: converts a sum over the scaled frame back to
// the pixel frame.
constexpr float kInvScale2 = 1.3862943611198906f;

// The 9 gradient words of one (tile, entry) pair from its per-pixel sums:
//   u[0..2] = sum g_rgb * T alpha (colour),
//   u[3] = sum r (opacity, r = dL/dalpha * raw alpha, 0 where alpha is clamped),
//   u[4..6] = sum r u'x^2, r u'x u'y, r u'y^2,  u[7..8] = sum r u'x, r u'y,
// with u' = M' d the scaled frame coordinates of the pixel offset d from the
// centre (M' the staged rows (a.x, a.y), (b.x, b.y)).  alpha = opacity *
// exp(-q / 2) with q = d^T C d (C the conic), so dL/dq = -r / 2 and
//   dL/d(opacity) = sum r / opacity,
//   dL/dC = sum dL/dq d d^T = M'^-1 S M'^-T,  S = -1/2 [[u4, u5], [u5, u6]],
//   dL/dcentre = sum dL/dq (-2 C d) = -2 / kFrameScale^2 M'^T s,  s = -1/2 (u7, u8),
// written as (colour 3, opacity, C00, 2 C01, C11, centre x, centre y).  The
// opacity's reciprocal is recovered from the staged -log2(opacity) in a.w.
__device__ __forceinline__ void pair_words(const float u[9], const float4& a, const float4& b,
                                           float* __restrict__ out) {
  // 1 / opacity = exp2(log2(1 / opacity)), the staged a.w; reciprocals are the
  // hardware's (1 ulp), well inside the gradients' tolerance
  const float inv_opacity = __builtin_amdgcn_exp2f(a.w);
  out[0] = u[0];
  out[1] = u[1];
  out[2] = u[2];
  out[3] = inv_opacity < __builtin_inff() ? u[3] * inv_opacity : 0.f;
  // M'^-1 = adj(M') / det; det = ab(c^2 + s^2)-like products of one sign, no cancellation
  const float det = a.x * b.y - a.y * b.x;
  const float id = __builtin_amdgcn_rcpf(det);
  const float A00 = b.y * id, A01 = -a.y * id, A10 = -b.x * id, A11 = a.x * id;
  const float p = -0.5f * u[4], r = -0.5f * u[5], q = -0.5f * u[6];
  const float t00 = A00 * p + A01 * r, t01 = A00 * r + A01 * q;  // row 0 of A S
  const float t10 = A10 * p + A11 * r, t11 = A10 * r + A11 * q;  // row 1 of A S
  out[4] = t00 * A00 + t01 * A01;
  out[5] = 2.f * (t00 * A10 + t01 * A11);
  out[6] = t10 * A10 + t11 * A11;
  out[7] = kInvScale2 * (a.x * u[7] + b.x * u[8]);
  out[8] = kInvScale2 * (a.y * u[7] + b.y * u[8]);
}

// Re-walks one tile back to front from the stored per-pixel state.  Per list
// entry 9 sums over the tile's pixels are formed (pair_words): within each
// half of a 16-lane row with DPP, then stored by that half row's last lane
// into a slot of its own (8 per wave per word).  With one wave per tile every
// slot has one writer and plain LDS stores suffice; LDS float atomics, used
// only when several waves share a tile, cost most of the kernel's time in a
// measured variant.  Each batch is flushed to the pairs' 9 words of gpart at
// their slot-order index q; the tile's entries past every pixel's last one
// get zeros, so each pair is written once and gpart needs no fill.  A
// reduction then sums each slot's contiguous run of pairs.
template <int TW, int TH, int KB>
__global__ __launch_bounds__(TW* TH) void diff_raster_backward_kernel(DiffLaunch L, const float4* __restrict__ rec,
                                                                      const uint32_t* __restrict__ vals,
                                                                      const uint32_t* __restrict__ ranges,
                                                                      const uint32_t* __restrict__ pslot,
                                                                      const uint32_t* __restrict__ tile_order,
                                                                      float* __restrict__ gpart) {
  constexpr int kN = TW * TH;
  constexpr int kRowsPerWave = kWave / TW;
  constexpr bool kOneWave = kN == kWave;  // the wave covers the tile: every listed record touches it
  // list entries per batch: one per thread for multi-wave tiles; a one-wave
  // tile takes KB (8, 16 or 32; 0: 16), fewer than its lanes, so the LDS
  // footprint (records + partial sums) admits more resident waves
  constexpr int kB = kOneWave ? (KB ? KB : 16) : kN;
  static_assert(kB <= kN, "batch larger than the workgroup");
  __shared__ float4 s_rec[3][kB];
  __shared__ float s_g[9][8][kB + 1];  // per word, one partial per half DPP row; +1 spreads them over banks
  __shared__ uint32_t s_j[kB];         // slot-order index q of each batch entry
  __shared__ uint32_t s_max;

  const uint32_t gtile = tile_order ? tile_order[blockIdx.x] : blockIdx.x;
  const DiffView& view = L.views[range_of(L.view_tiles, L.n_views, gtile)];
  const uint32_t tile = gtile - view.tile_base;
  const uint32_t W = view.fc.width, H = view.fc.height;
  const uint32_t tx = tile % view.fc.tiles_x, ty = tile / view.fc.tiles_x;
  const int t = threadIdx.x;
  const int px = (int)tx * TW + t % TW, py = (int)ty * TH + t / TW;
  const bool inside = px < (int)W && py < (int)H;
  const int col0 = __builtin_amdgcn_readfirstlane((int)tx * TW), col1 = col0 + TW - 1;
  const int row0 = __builtin_amdgcn_readfirstlane((int)ty * TH + (t / kWave) * kRowsPerWave);
  const int row1 = row0 + kRowsPerWave - 1;
  const int bx0 = col0 & ~(kRasterBlock - 1), by0 = __builtin_amdgcn_readfirstlane(((int)ty * TH) & ~(kRasterBlock - 1));
  const float X0 = (float)bx0, Y0 = (float)by0;
  const float lx = (float)(px - bx0) + 0.5f, ly = (float)(py - by0) + 0.5f;
  const uint32_t end = ranges[2 * gtile + 1];
  const bool half_last = (__lane_id() & 7u) == 7u;  // lanes 7 and 15 of each DPP row
  const uint32_t row = __lane_id() >> 3;              // half-row slot 0..7

  uint32_t last = 0;
  float T = 1.f;
  float4 g = make_float4(0.f, 0.f, 0.f, 0.f);
  if (inside) {
    const size_t pix = (size_t)py * W + px;
    last = view.last[pix];
    T = view.final_t[pix];
    g = view.grad_out[pix];
  }
  if (t == 0) s_max = 0;
  __syncthreads();
  if (last) atomicMax(&s_max, last);
  __syncthreads();
  const uint32_t max_last = s_max;
  // g . (colour, alpha) composited behind the current entry: the colour
  // behind starts at 0 and the alpha behind at 1, and both follow the blend
  // r <- alpha * x + (1 - alpha) * r, so their g-weighted sum does too
  float acc = g.w;
  uint32_t walked = 0;  // entries this wave stepped through (work counter)

  // traversal index i (0 = nearest) maps to list position end - 1 - i; walk i downwards
  for (uint32_t hi = max_last; hi > 0;) {
    const uint32_t lo = hi > (uint32_t)kB ? hi - kB : 0u;
    const int cnt = (int)(hi - lo);
    __syncthreads();
    if (t < cnt) {
      const uint32_t i = hi - 1 - t;
      const uint32_t j = vals[end - 1 - i];
      const uint32_t k = pslot[j];
      s_j[t] = j;
      const float4 r0 = rec[3 * (size_t)k], r1 = rec[3 * (size_t)k + 1], r2 = rec[3 * (size_t)k + 2];
      stage_record(r0, r1, r2, true, X0, Y0, &s_rec[0][t], &s_rec[1][t], &s_rec[2][t]);
#pragma unroll
      for (int w = 0; w < 9; ++w)
#pragma unroll
        for (int r = 0; r < 8; ++r) s_g[w][r][t] = 0.f;
    }
    __syncthreads();
    for (int e = 0; e < cnt; ++e) {
      const uint32_t i = hi - 1 - e;
      const float4 a = s_rec[0][e], bb = s_rec[1][e], c = s_rec[2][e];
      if (!kOneWave) {
        const uint32_t bx = __builtin_amdgcn_readfirstlane(__float_as_uint(bb.w));
        const uint32_t by = __builtin_amdgcn_readfirstlane(__float_as_uint(c.w));
        const int bx0r = (int)(bx & 0xffffu), bx1r = (int)(bx >> 16);
        const int by0r = (int)(by & 0xffffu), by1r = (int)(by >> 16);
        if (by1r < row0 || by0r > row1 || bx1r < col0 || bx0r > col1) continue;
      }
      const float ux = fmaf(a.x, lx, fmaf(a.y, ly, a.z));
      const float uy = fmaf(bb.x, lx, fmaf(bb.y, ly, bb.z));
      const bool in = (i < last) & (fmaxf(fabsf(ux), fabsf(uy)) <= kCoverScaled);
      if (__ballot(in) == 0) continue;
      ++walked;
      // every lane takes part from here (DPP reads all lanes); lanes with
      // !in get raw alpha 0: their T (rcp(1) = 1), acc and every term stay
      // unchanged or zero
      const float raw = in ? __builtin_amdgcn_exp2f(-fmaf(ux, ux, fmaf(uy, uy, a.w))) : 0.f;
      const float alpha = fminf(raw, kAlphaMax);
      const float om = 1.f - alpha;
      T = T * __builtin_amdgcn_rcpf(om);  // transmittance in front of entry i
      const float ta = T * alpha;
      // dL/dalpha = T * (g . (colour, 2 alpha) - g . (behind)) and the blend of
      // g . (colour, alpha) into acc, sharing y = g.w alpha + g . colour - acc
      const float gc = fmaf(g.z, c.z, fmaf(g.y, c.y, g.x * c.x));
      const float y = fmaf(g.w, alpha, gc - acc);
      const float dalpha = T * fmaf(g.w, alpha, y);
      acc = fmaf(alpha, y, acc);
      const float rr = raw > kAlphaMax ? 0.f : dalpha * raw;
      const float dpx = rr * ux, dpy = rr * uy;
      float s0 = half_row_sum(g.x * ta);
      float s1 = half_row_sum(g.y * ta);
      float s2 = half_row_sum(g.z * ta);
      float s3 = half_row_sum(rr);
      float s4 = half_row_sum(dpx * ux);
      float s5 = half_row_sum(dpx * uy);
      float s6 = half_row_sum(dpy * uy);
      float s7 = half_row_sum(dpx);
      float s8 = half_row_sum(dpy);
      // keep the last DPP add ahead of the branch (every lane computes it), so
      // it folds into v_add_f32_dpp instead of a v_mov_b32_dpp + a later add
      __asm__ volatile("" : "+v"(s0), "+v"(s1), "+v"(s2), "+v"(s3), "+v"(s4), "+v"(s5), "+v"(s6), "+v"(s7), "+v"(s8));
      if (half_last) {
        if (kOneWave) {  // one writer per slot: plain LDS stores
          s_g[0][row][e] = s0;
          s_g[1][row][e] = s1;
          s_g[2][row][e] = s2;
          s_g[3][row][e] = s3;
          s_g[4][row][e] = s4;
          s_g[5][row][e] = s5;
          s_g[6][row][e] = s6;
          s_g[7][row][e] = s7;
          s_g[8][row][e] = s8;
        } else {  // half rows of the tile's other waves share the slot
          atomicAdd(&s_g[0][row][e], s0);
          atomicAdd(&s_g[1][row][e], s1);
          atomicAdd(&s_g[2][row][e], s2);
          atomicAdd(&s_g[3][row][e], s3);
          atomicAdd(&s_g[4][row][e], s4);
          atomicAdd(&s_g[5][row][e], s5);
          atomicAdd(&s_g[6][row][e], s6);
          atomicAdd(&s_g[7][row][e], s7);
          atomicAdd(&s_g[8][row][e], s8);
        }
      }
    }
    __syncthreads();
    if (t < cnt) {
      float u[9];
#pragma unroll
      for (int w = 0; w < 9; ++w)
        u[w] = ((s_g[w][0][t] + s_g[w][1][t]) + (s_g[w][2][t] + s_g[w][3][t])) +
               ((s_g[w][4][t] + s_g[w][5][t]) + (s_g[w][6][t] + s_g[w][7][t]));
      pair_words(u, s_rec[0][t], s_rec[1][t], gpart + 9 * (size_t)s_j[t]);  // at the pair's q
    }
    hi = lo;
  }
  // entries beyond every pixel's last one were never composited: their
  // partials are zero (each pair is written exactly once, here or above)
  const uint32_t items = end - ranges[2 * gtile];
  for (uint32_t i = max_last + t; i < items; i += kN) {
    float* out = gpart + 9 * (size_t)vals[end - 1 - i];
#pragma unroll
    for (int w = 0; w < 9; ++w) out[w] = 0.f;
  }
  if (L.work && t % kWave == 0) L.work[(size_t)blockIdx.x * (kN / kWave) + t / kWave] = walked;
}

// ------------------------------------------- two 8x8 tiles per wave ---
//
// The 8x8 forward and backward rasters also run with two tiles per wave:
// lanes 0-31 take tile order[2i], lanes 32-63 tile order[2i + 1], each lane
// two pixels of one row of its tile (columns c and c + 4, which share the
// row's term of each frame coordinate).  A record read from LDS then serves
// two pixels, and in the backward each of the nine per-entry sums first adds
// the lane's two pixels and then runs the same three DPP steps over groups of
// 8 lanes (two pixel rows), so the reduction costs half as much per pixel.
// Consecutive tiles of the longest-first schedule have list lengths within a
// bucket of each other; a half whose list is shorter walks records that
// touch none of its pixels.  The per-pixel arithmetic is that of the one-tile
// kernels, so images are the same bit for bit; the backward adds a lane's two
// pixels' gradient words with FMAs before the half-row sums, so its partials
// agree with the one-tile kernel's to float rounding.

constexpr int kPairHalf = kWave / 2;  // lanes per tile

// Maximum of v over the 32 lanes of each half wave (every lane gets its half's).
__device__ __forceinline__ uint32_t half_max(uint32_t v) {
#pragma unroll
  for (int d = 1; d < kPairHalf; d <<= 1) v = max(v, (uint32_t)__shfl_xor((int)v, d, kWave));
  return v;
}

// The forward of diff_raster_kernel<8, 8> for two tiles per wave.
__global__ __launch_bounds__(kWave) void diff_raster_pair_kernel(DiffLaunch L, const float4* __restrict__ rec,
                                                                 const uint32_t* __restrict__ vals,
                                                                 const uint32_t* __restrict__ ranges,
                                                                 const uint32_t* __restrict__ pslot,
                                                                 const uint32_t* __restrict__ tile_order,
                                                                 int keep_state) {
  constexpr int kStride = kPairHalf + 1;  // the second tile's records start one float4 later: other banks
  __shared__ float4 s_rec[2][3][2 * kStride];
  const int t = threadIdx.x, h = t / kPairHalf, l = t % kPairHalf;
  const uint32_t slot = 2 * blockIdx.x + h;
  const bool has_tile = slot < L.n_tiles;
  const uint32_t gtile = has_tile ? (tile_order ? tile_order[slot] : slot) : 0u;
  const DiffView& view = L.views[range_of(L.view_tiles, L.n_views, gtile)];
  const uint32_t tile = gtile - view.tile_base;
  const uint32_t W = view.fc.width, H = view.fc.height;
  const uint32_t tx = tile % view.fc.tiles_x, ty = tile / view.fc.tiles_x;
  // a lane's two pixels: columns c and c + 4 of one row, sharing the row's
  // term of each frame coordinate (each pixel's arithmetic is diff_step's)
  const int px0 = (int)tx * 8 + l % 4, px1 = px0 + 4, py = (int)ty * 8 + l / 4;
  const bool in0 = has_tile && px0 < (int)W && py < (int)H, in1 = has_tile && px1 < (int)W && py < (int)H;
  const int bx0 = ((int)tx * 8) & ~(kRasterBlock - 1), by0 = ((int)ty * 8) & ~(kRasterBlock - 1);
  const float X0 = (float)bx0, Y0 = (float)by0;
  const float lx0 = (float)(px0 - bx0) + 0.5f, lx1 = lx0 + 4.f, ly = (float)(py - by0) + 0.5f;
  uint32_t end = 0, items = 0;
  if (has_tile) {
    end = ranges[2 * gtile + 1];
    items = end - ranges[2 * gtile];
  }
  const uint32_t items_a = __builtin_amdgcn_readlane(items, 0), items_b = __builtin_amdgcn_readlane(items, kPairHalf);
  const uint32_t batches = (max(items_a, items_b) + kPairHalf - 1) / kPairHalf;
  const float t_min = L.t_min;

  float4 p0 = make_float4(0.f, 0.f, 0.f, 0.f), p1 = p0, p2 = p0;
  if ((uint32_t)l < items) {
    const uint32_t k = pslot[vals[end - 1 - l]];
    p0 = rec[3 * (size_t)k];
    p1 = rec[3 * (size_t)k + 1];
    p2 = rec[3 * (size_t)k + 2];
  }
  float T0 = in0 ? 1.f : 0.f, r0 = 0.f, g0 = 0.f, b0 = 0.f, a0 = 0.f;
  float T1 = in1 ? 1.f : 0.f, r1 = 0.f, g1 = 0.f, b1 = 0.f, a1 = 0.f;
  uint32_t last0 = 0, last1 = 0;
  uint32_t walked_a = 0, walked_b = 0;  // each tile's own list entries stepped through (work counter)
  for (uint32_t b = 0; b < batches; ++b) {
    const int buf = b & 1;
    // entries past a half's list (the other tile's is longer) are staged as
    // records that cover nothing, so they change no pixel
    const int at = h * kStride + l;
    stage_record(p0, p1, p2, b * kPairHalf + l < items, X0, Y0, &s_rec[buf][0][at], &s_rec[buf][1][at],
                 &s_rec[buf][2][at]);
    if (__syncthreads_count((T0 >= t_min) | (T1 >= t_min)) == 0) break;
    const uint32_t next = (b + 1) * kPairHalf + l;
    if (next < items) {
      const uint32_t k = pslot[vals[end - 1 - next]];
      p0 = rec[3 * (size_t)k];
      p1 = rec[3 * (size_t)k + 1];
      p2 = rec[3 * (size_t)k + 2];
    }
    const uint32_t done = b * kPairHalf;
    const int cnt_a = items_a > done ? (int)min((uint32_t)kPairHalf, items_a - done) : 0;
    const int cnt_b = items_b > done ? (int)min((uint32_t)kPairHalf, items_b - done) : 0;
    const int cnt = max(cnt_a, cnt_b);
    if (__ballot((T0 >= t_min) | (T1 >= t_min)) == 0) continue;
    const float4* s0 = s_rec[buf][0] + h * kStride;
    const float4* s1 = s_rec[buf][1] + h * kStride;
    const float4* s2 = s_rec[buf][2] + h * kStride;
    // groups of kGroup entries between exit tests: a pixel that has stopped
    // (T < t_min) composites nothing more and keeps its last, and entries
    // past both lists are staged as records that cover nothing, so the extra
    // steps change no pixel state
    constexpr int kGroup = 4;
    static_assert(kPairHalf % kGroup == 0, "batches hold whole groups");
    int e = 0;
    for (; e < cnt;) {
#pragma unroll
      for (int j = 0; j < kGroup; ++j) {
        const float4 a = s0[e + j], bb = s1[e + j], c = s2[e + j];
        const uint32_t idx = done + e + j + 1;
        const float jx = fmaf(a.y, ly, a.z), jy = fmaf(bb.y, ly, bb.z);
        diff_step_at(fmaf(a.x, lx0, jx), fmaf(bb.x, lx0, jy), a, c, idx, t_min, T0, r0, g0, b0, a0, last0);
        diff_step_at(fmaf(a.x, lx1, jx), fmaf(bb.x, lx1, jy), a, c, idx, t_min, T1, r1, g1, b1, a1, last1);
      }
      e += kGroup;
      if (__ballot((T0 >= t_min) | (T1 >= t_min)) == 0) break;
    }
    walked_a += (uint32_t)min(e, cnt_a);
    walked_b += (uint32_t)min(e, cnt_b);
  }
  // a pixel still compositing went through its whole list (idx also ran over
  // the padding of a shorter half)
  if (T0 >= t_min) last0 = items;
  if (T1 >= t_min) last1 = items;
  if (in0) {
    const size_t pix = (size_t)py * W + px0;
    view.out[pix] = make_float4(r0, g0, b0, a0 + T0);
    if (keep_state) {
      view.final_t[pix] = T0;
      view.last[pix] = last0;
    }
  }
  if (in1) {
    const size_t pix = (size_t)py * W + px1;
    view.out[pix] = make_float4(r1, g1, b1, a1 + T1);
    if (keep_state) {
      view.final_t[pix] = T1;
      view.last[pix] = last1;
    }
  }
  if (L.work && l == 0 && has_tile) L.work[slot] = h ? walked_b : walked_a;
}

// Backward of one (pixel, entry) with the arithmetic of
// diff_raster_backward_kernel: a pixel outside the square or past its last
// entry gets raw alpha 0, so its terms are zero and T and acc stay.
struct PixState {
  float T, acc;
  float4 g;
  uint32_t last;
};

struct PixCover {
  float ux, uy;
  bool in;
};

// (ux, uy) = the pixel's frame coordinates, formed as diff_step forms them:
// fma(a.x, x, fma(a.y, y, a.z)) and the same with bb.
__device__ __forceinline__ PixCover bwd_cover(const PixState& ps, float ux, float uy, uint32_t i) {
  PixCover pc;
  pc.ux = ux;
  pc.uy = uy;
  pc.in = (i < ps.last) & (fmaxf(fabsf(pc.ux), fabsf(pc.uy)) <= kCoverScaled);
  return pc;
}

// The per-pixel terms of one entry: ta = T alpha (the colour words' weight),
// rr = dL/dalpha * raw alpha (0 where alpha is clamped) and rr u'x, rr u'y;
// updates T and acc.
struct PixTerms {
  float ta, rr, dpx, dpy;
};

__device__ __forceinline__ PixTerms bwd_pixel(PixState& ps, const PixCover& pc, const float4& a, const float4& c) {
  const float raw = pc.in ? __builtin_amdgcn_exp2f(-fmaf(pc.ux, pc.ux, fmaf(pc.uy, pc.uy, a.w))) : 0.f;
  const float alpha = fminf(raw, kAlphaMax);
  const float om = 1.f - alpha;
  ps.T = ps.T * __builtin_amdgcn_rcpf(om);
  PixTerms r;
  r.ta = ps.T * alpha;
  const float gc = fmaf(ps.g.z, c.z, fmaf(ps.g.y, c.y, ps.g.x * c.x));
  const float y = fmaf(ps.g.w, alpha, gc - ps.acc);
  const float dalpha = ps.T * fmaf(ps.g.w, alpha, y);
  ps.acc = fmaf(alpha, y, ps.acc);
  r.rr = raw > kAlphaMax ? 0.f : dalpha * raw;
  r.dpx = r.rr * pc.ux;
  r.dpy = r.rr * pc.uy;
  return r;
}

// The backward of diff_raster_backward_kernel<8, 8, KB> for two tiles per wave.
template <int KB>
__global__ __launch_bounds__(kWave) void diff_raster_backward_pair_kernel(
    DiffLaunch L, const float4* __restrict__ rec, const uint32_t* __restrict__ vals,
    const uint32_t* __restrict__ ranges, const uint32_t* __restrict__ pslot, const uint32_t* __restrict__ tile_order,
    float* __restrict__ gpart) {
  constexpr int kB = KB;  // list entries per batch of each tile
  constexpr int kSlots = kPairHalf / 8;  // groups of 8 lanes (two pixel rows) per tile
  __shared__ float4 s_rec[2][3][kB];
  __shared__ float s_g[2][9][kSlots][kB + 1];
  __shared__ uint32_t s_j[2][kB];

  const int t = threadIdx.x, h = t / kPairHalf, l = t % kPairHalf;
  const uint32_t slot = 2 * blockIdx.x + h;
  const bool has_tile = slot < L.n_tiles;
  const uint32_t gtile = has_tile ? (tile_order ? tile_order[slot] : slot) : 0u;
  const DiffView& view = L.views[range_of(L.view_tiles, L.n_views, gtile)];
  const uint32_t tile = gtile - view.tile_base;
  const uint32_t W = view.fc.width, H = view.fc.height;
  const uint32_t tx = tile % view.fc.tiles_x, ty = tile / view.fc.tiles_x;
  // the forward pair kernel's pixels: columns c and c + 4 of one row
  const int px0 = (int)tx * 8 + l % 4, px1 = px0 + 4, py = (int)ty * 8 + l / 4;
  const bool in0 = has_tile && px0 < (int)W && py < (int)H, in1 = has_tile && px1 < (int)W && py < (int)H;
  const int bx0 = ((int)tx * 8) & ~(kRasterBlock - 1), by0 = ((int)ty * 8) & ~(kRasterBlock - 1);
  const float X0 = (float)bx0, Y0 = (float)by0;
  const float lx0 = (float)(px0 - bx0) + 0.5f, lx1 = lx0 + 4.f, ly = (float)(py - by0) + 0.5f;
  uint32_t end = 0, items = 0;
  if (has_tile) {
    end = ranges[2 * gtile + 1];
    items = end - ranges[2 * gtile];
  }
  const bool half_last = (__lane_id() & 7u) == 7u;  // the last lane of each group of 8
  const uint32_t row = (uint32_t)l >> 3;           // the group's slot 0..3 of the tile

  PixState s0 = {1.f, 0.f, make_float4(0.f, 0.f, 0.f, 0.f), 0u}, s1 = s0;
  if (in0) {
    const size_t pix = (size_t)py * W + px0;
    s0.last = view.last[pix];
    s0.T = view.final_t[pix];
    s0.g = view.grad_out[pix];
  }
  if (in1) {
    const size_t pix = (size_t)py * W + px1;
    s1.last = view.last[pix];
    s1.T = view.final_t[pix];
    s1.g = view.grad_out[pix];
  }
  // g . (colour, alpha) composited behind the current entry (as in the one-tile kernel)
  s0.acc = s0.g.w;
  s1.acc = s1.g.w;
  const uint32_t max_half = half_max(max(s0.last, s1.last));
  const uint32_t max_a = __builtin_amdgcn_readlane(max_half, 0), max_b = __builtin_amdgcn_readlane(max_half, kPairHalf);
  const uint32_t max_all = max(max_a, max_b);
  uint32_t walked_a = 0, walked_b = 0;  // each tile's own entries stepped through (work counter)

  // traversal index i (0 = nearest) maps to list position end - 1 - i; walk i downwards
  for (uint32_t hi = max_all; hi > 0;) {
    const uint32_t lo = hi > (uint32_t)kB ? hi - kB : 0u;
    const int cnt = (int)(hi - lo);
    __syncthreads();
    if (l < cnt) {
      const uint32_t i = hi - 1 - l;
      float4 r0 = make_float4(0.f, 0.f, 0.f, 0.f), r1 = r0, r2 = r0;
      uint32_t j = 0;
      const bool mine = i < max_half;  // an entry some pixel of this tile composited
      if (mine) {
        j = vals[end - 1 - i];
        const uint32_t k = pslot[j];
        r0 = rec[3 * (size_t)k];
        r1 = rec[3 * (size_t)k + 1];
        r2 = rec[3 * (size_t)k + 2];
      }
      s_j[h][l] = j;
      stage_record(r0, r1, r2, mine, X0, Y0, &s_rec[h][0][l], &s_rec[h][1][l], &s_rec[h][2][l], kInertFinite);
#pragma unroll
      for (int w = 0; w < 9; ++w)
#pragma unroll
        for (int r = 0; r < kSlots; ++r) s_g[h][w][r][l] = 0.f;
    }
    __syncthreads();
    for (int e = 0; e < cnt; ++e) {
      const uint32_t i = hi - 1 - e;
      const float4 a = s_rec[h][0][e], bb = s_rec[h][1][e], c = s_rec[h][2][e];
      const float jx = fmaf(a.y, ly, a.z), jy = fmaf(bb.y, ly, bb.z);
      const PixCover c0 = bwd_cover(s0, fmaf(a.x, lx0, jx), fmaf(bb.x, lx0, jy), i);
      const PixCover c1 = bwd_cover(s1, fmaf(a.x, lx1, jx), fmaf(bb.x, lx1, jy), i);
      if (__ballot(c0.in | c1.in) == 0) continue;  // no pixel of either tile: the slots stay zero
      walked_a += i < max_a ? 1u : 0u;
      walked_b += i < max_b ? 1u : 0u;
      const PixTerms t0 = bwd_pixel(s0, c0, a, c), t1 = bwd_pixel(s1, c1, a, c);
      // the lane's two pixels summed (FMAs), then over its group of 8 lanes
      float w[9];
      w[0] = fmaf(s1.g.x, t1.ta, s0.g.x * t0.ta);
      w[1] = fmaf(s1.g.y, t1.ta, s0.g.y * t0.ta);
      w[2] = fmaf(s1.g.z, t1.ta, s0.g.z * t0.ta);
      w[3] = t0.rr + t1.rr;
      w[4] = fmaf(t1.dpx, c1.ux, t0.dpx * c0.ux);
      w[5] = fmaf(t1.dpx, c1.uy, t0.dpx * c0.uy);
      w[6] = fmaf(t1.dpy, c1.uy, t0.dpy * c0.uy);
      w[7] = t0.dpx + t1.dpx;
      w[8] = t0.dpy + t1.dpy;
#pragma unroll
      for (int k = 0; k < 9; ++k) w[k] = half_row_sum(w[k]);
      // keep the last DPP add ahead of the branch (every lane computes it)
      __asm__ volatile("" : "+v"(w[0]), "+v"(w[1]), "+v"(w[2]), "+v"(w[3]), "+v"(w[4]), "+v"(w[5]), "+v"(w[6]),
                       "+v"(w[7]), "+v"(w[8]));
      if (half_last) {
#pragma unroll
        for (int k = 0; k < 9; ++k) s_g[h][k][row][e] = w[k];
      }
    }
    __syncthreads();
    if (l < cnt && hi - 1 - l < max_half) {
      float u[9];
#pragma unroll
      for (int k = 0; k < 9; ++k) u[k] = (s_g[h][k][0][l] + s_g[h][k][1][l]) + (s_g[h][k][2][l] + s_g[h][k][3][l]);
      pair_words(u, s_rec[h][0][l], s_rec[h][1][l], gpart + 9 * (size_t)s_j[h][l]);
    }
    hi = lo;
  }
  // entries past every pixel's last one in this tile: zero partials (each pair written once)
  for (uint32_t i = max_half + l; i < items; i += kPairHalf) {
    float* out = gpart + 9 * (size_t)vals[end - 1 - i];
#pragma unroll
    for (int k = 0; k < 9; ++k) out[k] = 0.f;
  }
  if (L.work && l == 0 && has_tile) L.work[slot] = h ? walked_b : walked_a;
}

// Sum of x over the 64 lanes of a wave (all lanes active); lane 63 gets it.
__device__ __forceinline__ float wave_total_last(float x) {
  x = row_sum16(x);
  // rows outside the row mask read 0 (the `old` operand)
  x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x142, 0xa, 0xf, false));  // row_bcast:15
  x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x143, 0xc, 0xf, false));  // row_bcast:31
  return x;
}

// Sum of the partials of one slot's pairs, which sit at o .. o + c - 1 of
// gpart (9 words each; a slot's pairs are numbered contiguously in slot
// order), for every lane of a wave at once; lanes whose slots are consecutive
// have adjacent runs, which together span [r0, r1) (r0 even).  Runs of up to
// kRunSerial pairs: the span is staged through LDS 64 pairs at a time with
// coalesced loads (lane i reads words i, i + 64, ...) and each lane adds the
// part of its run in the chunk (chunks that only long runs need are skipped).
// Longer runs are summed by the whole wave.  Chunks start at even pair
// indices, so the pairs summed two at a time are (even q, q + 1): a slot's sum
// depends on its run alone, whatever the wave's other runs.
__device__ __forceinline__ void sum_slot_runs(const float* __restrict__ gpart, uint64_t o, uint32_t c, uint64_t r0,
                                              uint64_t r1, float* stage, float acc[9]) {
  const uint32_t lane = __lane_id();
  const bool big = c > kRunSerial;
  const uint64_t e = o + c;
#pragma unroll
  for (int w = 0; w < 9; ++w) acc[w] = 0.f;
  // the next needed chunk's words are loaded while the current one is
  // summed, so a wave has one chunk of loads in flight behind its adds
  auto needed = [&](uint64_t cs) {
    const uint64_t lo = o > cs ? o : cs, hi = e < cs + kWave ? e : cs + kWave;
    return __ballot(!big && lo < hi) != 0;
  };
  auto load_chunk = [&](uint64_t cs, float* buf) {
    const uint32_t words = 9u * (uint32_t)min((uint64_t)kWave, r1 - cs);
    const float* src = gpart + 9 * cs;
#pragma unroll
    for (uint32_t q = 0; q < 9; ++q) {
      const uint32_t i = lane + q * kWave;
      buf[q] = i < words ? src[i] : 0.f;
    }
  };
  uint64_t cs = r0;
  while (cs < r1 && !needed(cs)) cs += kWave;
  float cur[9];
  if (cs < r1) load_chunk(cs, cur);
  while (cs < r1) {
    uint64_t nx = cs + kWave;
    while (nx < r1 && !needed(nx)) nx += kWave;
    const uint32_t words = 9u * (uint32_t)min((uint64_t)kWave, r1 - cs);
#pragma unroll
    for (uint32_t q = 0; q < 9; ++q) {
      const uint32_t i = lane + q * kWave;
      if (i < words) stage[i] = cur[q];
    }
    if (nx < r1) load_chunk(nx, cur);
    __builtin_amdgcn_wave_barrier();  // a wave's LDS operations complete in order
    const uint64_t lo = o > cs ? o : cs, hi = e < cs + kWave ? e : cs + kWave;
    if (!big && lo < hi) {
      uint32_t p = (uint32_t)(lo - cs);
      const uint32_t pe = (uint32_t)(hi - cs);
      if (p & 1u) {  // an odd first pair alone (cs is even)
        const float* v = stage + 9 * p;
#pragma unroll
        for (int w = 0; w < 9; ++w) acc[w] += v[w];
        ++p;
      }
      for (; p + 2 <= pe; p += 2) {  // two pairs per trip: 18 independent LDS reads in flight
        const float* v = stage + 9 * p;
        float x[18];
#pragma unroll
        for (int w = 0; w < 18; ++w) x[w] = v[w];
#pragma unroll
        for (int w = 0; w < 9; ++w) acc[w] += x[w] + x[9 + w];
      }
      if (p < pe) {
        const float* v = stage + 9 * p;
#pragma unroll
        for (int w = 0; w < 9; ++w) acc[w] += v[w];
      }
    }
    __builtin_amdgcn_wave_barrier();
    cs = nx;
  }
  uint64_t pending = __ballot(big);
  while (pending) {
    const uint32_t src_lane = (uint32_t)__builtin_ctzll(pending);
    pending &= pending - 1;
    const uint32_t bc = lane_u32(c, src_lane);
    const uint64_t bo = ((uint64_t)lane_u32((uint32_t)(o >> 32), src_lane) << 32) | lane_u32((uint32_t)o, src_lane);
    float t[9] = {};
    for (uint32_t i = lane; i < bc; i += kWave) {
      const float* src = gpart + 9 * (bo + i);
#pragma unroll
      for (int w = 0; w < 9; ++w) t[w] += src[w];
    }
#pragma unroll
    for (int w = 0; w < 9; ++w) {
      const float tot = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(wave_total_last(t[w])), kWave - 1));
      if (lane == src_lane) acc[w] = tot;
    }
  }
}

// Row-major 3x3 helpers (of the projection backward: FMA-contracted).
__device__ __forceinline__ void mm3(const float* A, const float* B, float* C) {
#pragma clang fp contract(fast)
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) C[3 * r + c] = A[3 * r] * B[c] + A[3 * r + 1] * B[3 + c] + A[3 * r + 2] * B[6 + c];
}
__device__ __forceinline__ void mm3_tn(const float* A, const float* B, float* C) {  // A^T B
#pragma clang fp contract(fast)
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) C[3 * r + c] = A[r] * B[c] + A[3 + r] * B[3 + c] + A[6 + r] * B[6 + c];
}
__device__ __forceinline__ void mm3_nt(const float* A, const float* B, float* C) {  // A B^T
#pragma clang fp contract(fast)
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) C[3 * r + c] = A[3 * r] * B[3 * c] + A[3 * r + 1] * B[3 * c + 1] + A[3 * r + 2] * B[3 * c + 2];
}

// View-independent state of one splat for the projection backward.
struct SplatState {
  float m[3];    // mean
  float q[4];    // normalized (x, y, z, w)
  float qlen;    // |q| before normalization
  float s[3];    // scales
  float R[9];    // rotation, row-major
  float Sig[9];  // R diag(s^2) R^T
  float sh[48];  // channel-major SH
};

// Chains the 9 gradient words gv of one view (colour 3, opacity 1, conic 3,
// centre 2) through that view's projection and adds the result to the
// view-summed dL/dSigma (GSig), dL/dmean (gm) and dL/dSH (gsh).  The forward
// quantities are recomputed in the forward's arithmetic (FMA contraction,
// hardware reciprocals, the view's V M product).
__device__ __forceinline__ void chain_view(const DiffView& view, const SplatState& P, const float gv[9],
                                           float GSig[9], float gm[3], float gsh[48]) {
#pragma clang fp contract(fast)
  const FrameConst& fc = view.fc;
  // ---- recompute the forward (row-major) ----
  float Wm[9];  // W = (V M)_3x3, row-major
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) Wm[3 * r + c] = view.vm[4 * c + r];
  float pv[4];
  for (int r = 0; r < 4; ++r)
    pv[r] = view.vm[r] * P.m[0] + view.vm[4 + r] * P.m[1] + view.vm[8 + r] * P.m[2] + view.vm[12 + r];
  const float x = pv[0], y = pv[1], z = pv[2];
  float WS[9], Sp[9];
  mm3(Wm, P.Sig, WS);
  mm3_nt(WS, Wm, Sp);
  const float iz = __builtin_amdgcn_rcpf(z), iz2 = iz * iz;
  const float J2[6] = {-iz, 0.f, x * iz2, 0.f, -iz, y * iz2};
  float JS[6];  // J2 Sp (2x3)
  for (int r = 0; r < 2; ++r)
    for (int c = 0; c < 3; ++c) JS[3 * r + c] = J2[3 * r] * Sp[c] + J2[3 * r + 1] * Sp[3 + c] + J2[3 * r + 2] * Sp[6 + c];
  float S2[4];  // J2 Sp J2^T
  for (int r = 0; r < 2; ++r)
    for (int c = 0; c < 2; ++c)
      S2[2 * r + c] = JS[3 * r] * J2[3 * c] + JS[3 * r + 1] * J2[3 * c + 1] + JS[3 * r + 2] * J2[3 * c + 2];
  const float Pm[4] = {fc.P[0], fc.P[4], fc.P[1], fc.P[5]};  // mat2(projection), row-major
  float PS[4], cov[4];
  for (int r = 0; r < 2; ++r)
    for (int c = 0; c < 2; ++c) PS[2 * r + c] = Pm[2 * r] * S2[c] + Pm[2 * r + 1] * S2[2 + c];
  for (int r = 0; r < 2; ++r)
    for (int c = 0; c < 2; ++c) cov[2 * r + c] = PS[2 * r] * Pm[c] + PS[2 * r + 1] * Pm[2 + c];
  cov[0] += fc.inv_w2;
  cov[3] += fc.inv_h2;
  const float hw = 0.5f * (float)fc.width, hh = 0.5f * (float)fc.height;
  const float cp00 = cov[0] * hw * hw, cp01 = cov[1] * hw * hh, cp10 = cov[2] * hh * hw, cp11 = cov[3] * hh * hh;
  const float det = cp00 * cp11 - cp01 * cp10;
  const float idet = __builtin_amdgcn_rcpf(det);
  const float Q00 = cp11 * idet, Q01 = -cp01 * idet, Q10 = -cp10 * idet, Q11 = cp00 * idet;

  // ---- conic -> cov2d (pixels) -> cov2d (NDC) -> S2 ----
  const float GQ00 = gv[4], GQ01 = 0.5f * gv[5], GQ11 = gv[6];
  // G_px = -Q GQ Q
  const float A00 = Q00 * GQ00 + Q01 * GQ01, A01 = Q00 * GQ01 + Q01 * GQ11;
  const float A10 = Q10 * GQ00 + Q11 * GQ01, A11 = Q10 * GQ01 + Q11 * GQ11;
  const float Gp00 = -(A00 * Q00 + A01 * Q10), Gp01 = -(A00 * Q01 + A01 * Q11);
  const float Gp10 = -(A10 * Q00 + A11 * Q10), Gp11 = -(A10 * Q01 + A11 * Q11);
  const float Gn[4] = {Gp00 * hw * hw, Gp01 * hw * hh, Gp10 * hh * hw, Gp11 * hh * hh};
  // cov = Pm S2 Pm  ->  dS2 = Pm^T Gn Pm^T
  float T1[4], GS2[4];
  for (int r = 0; r < 2; ++r)
    for (int c = 0; c < 2; ++c) T1[2 * r + c] = Pm[r] * Gn[c] + Pm[2 + r] * Gn[2 + c];
  for (int r = 0; r < 2; ++r)
    for (int c = 0; c < 2; ++c) GS2[2 * r + c] = T1[2 * r] * Pm[c * 2] + T1[2 * r + 1] * Pm[c * 2 + 1];
  // S2 = J2 Sp J2^T -> dSp = J2^T GS2 J2, dJ2 = (GS2 + GS2^T) J2 Sp
  float GSp[9];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      float acc = 0.f;
      for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) acc += J2[3 * i + r] * GS2[2 * i + j] * J2[3 * j + c];
      GSp[3 * r + c] = acc;
    }
  float GJ2[6];
  for (int r = 0; r < 2; ++r)
    for (int c = 0; c < 3; ++c) {
      float acc = 0.f;
      for (int i = 0; i < 2; ++i) acc += (GS2[2 * r + i] + GS2[2 * i + r]) * JS[3 * i + c];
      GJ2[3 * r + c] = acc;
    }
  // Sp = W Sig W^T -> dSig = W^T GSp W
  float TT[9], GSv[9];
  mm3_tn(Wm, GSp, TT);
  mm3(TT, Wm, GSv);
  for (int i = 0; i < 9; ++i) GSig[i] += GSv[i];

  // ---- view-space position: J2 and the projected centre ----
  float gpv[3];
  gpv[0] = GJ2[2] * iz2;
  gpv[1] = GJ2[5] * iz2;
  gpv[2] = (GJ2[0] + GJ2[4]) * iz2 - 2.f * (x * GJ2[2] + y * GJ2[5]) * iz2 * iz;
  {
    float clip[4];
    float pv4[4] = {x, y, z, 1.f};
    m4_mulv(fc.P, pv4, clip);
    const float iw = __builtin_amdgcn_rcpf(clip[3]);
    const float ndx = clip[0] * iw, ndy = clip[1] * iw;
    const float gx = gv[7] * hw, gy = gv[8] * hh;  // d/d ndc
    for (int j = 0; j < 3; ++j) {
      gpv[j] += gx * (fc.P[4 * j + 0] - ndx * fc.P[4 * j + 3]) * iw;
      gpv[j] += gy * (fc.P[4 * j + 1] - ndy * fc.P[4 * j + 3]) * iw;
    }
  }
  // mean <- W^T gpv
  for (int j = 0; j < 3; ++j) gm[j] += Wm[j] * gpv[0] + Wm[3 + j] * gpv[1] + Wm[6 + j] * gpv[2];

  // ---- colour: SH coefficients and view direction ----
  const float d0 = P.m[0] - fc.eye_model[0], d1 = P.m[1] - fc.eye_model[1], d2 = P.m[2] - fc.eye_model[2];
  const float idl = __builtin_amdgcn_rsqf(d0 * d0 + d1 * d1 + d2 * d2);
  const float dx = d0 * idl, dy = d1 * idl, dz = d2 * idl;
  float basis[16];
  sh_basis(dx, dy, dz, basis);
  float gcol[3];
  for (int ch = 0; ch < 3; ++ch) {
    float acc = 0.f;
    for (int j = 0; j < 16; ++j) acc += basis[j] * P.sh[ch * 16 + j];
    gcol[ch] = (acc + 0.5f < 0.f) ? 0.f : gv[ch];
  }
  float gb[16];
  for (int j = 0; j < 16; ++j) {
    gb[j] = gcol[0] * P.sh[j] + gcol[1] * P.sh[16 + j] + gcol[2] * P.sh[32 + j];
    for (int ch = 0; ch < 3; ++ch) gsh[ch * 16 + j] += gcol[ch] * basis[j];
  }
  {
    const float C1 = 0.4886025119029199f, C20 = 1.0925484305920792f, C21 = 0.31539156525252005f,
                C22 = 0.5462742152960396f, C30 = 0.5900435899266435f, C31 = 2.890611442640554f,
                C32 = 0.4570457994644658f, C33 = 0.3731763325901154f, C34 = 1.445305721320277f;
    const float X = dx, Y = dy, Z = dz, XX = X * X, YY = Y * Y, ZZ = Z * Z;
    float gdx = 0.f, gdy = 0.f, gdz = 0.f;
    gdy += gb[1] * -C1;
    gdz += gb[2] * C1;
    gdx += gb[3] * -C1;
    gdx += gb[4] * C20 * Y;
    gdy += gb[4] * C20 * X;
    gdy += gb[5] * -C20 * Z;
    gdz += gb[5] * -C20 * Y;
    gdx += gb[6] * -2.f * C21 * X;
    gdy += gb[6] * -2.f * C21 * Y;
    gdz += gb[6] * 4.f * C21 * Z;
    gdx += gb[7] * -C20 * Z;
    gdz += gb[7] * -C20 * X;
    gdx += gb[8] * 2.f * C22 * X;
    gdy += gb[8] * -2.f * C22 * Y;
    gdx += gb[9] * -6.f * C30 * X * Y;
    gdy += gb[9] * -3.f * C30 * (XX - YY);
    gdx += gb[10] * C31 * Y * Z;
    gdy += gb[10] * C31 * X * Z;
    gdz += gb[10] * C31 * X * Y;
    gdx += gb[11] * 2.f * C32 * X * Y;
    gdy += gb[11] * -C32 * (4.f * ZZ - XX - 3.f * YY);
    gdz += gb[11] * -8.f * C32 * Y * Z;
    gdx += gb[12] * -6.f * C33 * X * Z;
    gdy += gb[12] * -6.f * C33 * Y * Z;
    gdz += gb[12] * C33 * (6.f * ZZ - 3.f * XX - 3.f * YY);
    gdx += gb[13] * -C32 * (4.f * ZZ - 3.f * XX - YY);
    gdy += gb[13] * 2.f * C32 * X * Y;
    gdz += gb[13] * -8.f * C32 * X * Z;
    gdx += gb[14] * 2.f * C34 * X * Z;
    gdy += gb[14] * -2.f * C34 * Y * Z;
    gdz += gb[14] * C34 * (XX - YY);
    gdx += gb[15] * -3.f * C30 * (XX - YY);
    gdy += gb[15] * 6.f * C30 * X * Y;
    const float dd = gdx * X + gdy * Y + gdz * Z;
    gm[0] += (gdx - X * dd) * idl;
    gm[1] += (gdy - Y * dd) * idl;
    gm[2] += (gdz - Z * dd) * idl;
  }
}

// One lane per scene splat: visits every view of the scene, sums the view's
// 9 gradient words at the splat's slot (found from the forward's visibility
// ballots and per-wave offsets) from the partials of the slot's pairs and
// chains them; then writes all 59 parameter gradients of the splat once
// (zeros for a splat no view touched).  The slots of a wave's visible lanes
// in one view are consecutive, so their pair runs are adjacent and the wave
// streams them (sum_slot_runs).
__global__ __launch_bounds__(kCompactBlock) void diff_project_backward_kernel(
    DiffLaunch L, const uint32_t* __restrict__ offsets, const uint64_t* __restrict__ masks,
    const uint32_t* __restrict__ soffs, const uint32_t* __restrict__ scnt, const float* __restrict__ gpart,
    uint64_t max_pairs) {
  __shared__ float s_stage[kCompactBlock / kWave][9 * kWave];  // per wave: the partials of 64 consecutive pairs
  float* stage = s_stage[threadIdx.x / kWave];
  const uint32_t sb = blockIdx.x;
  const DiffScene& sc = L.scenes[range_of(L.scene_blocks, L.n_scenes, sb)];
  const uint32_t lb = sb - sc.block_base;
  const uint32_t li = lb * kCompactBlock + threadIdx.x;
  const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const uint32_t lane = __lane_id();
  const bool valid = li < sc.n;
  const uint32_t id = valid && sc.perm ? sc.perm[li] : li;

  SplatState P;
  float op = 0.5f;
  if (valid) {
    P.qlen = load_quat(sc, id, P.q);
    for (int i = 0; i < 3; ++i) {
      P.m[i] = ld_global(sc.means + 3 * (size_t)id + i);
      P.s[i] = expf(ld_global(sc.log_scales + 3 * (size_t)id + i));
    }
    const float4* sh4 = reinterpret_cast<const float4*>(sc.sh + 48 * (size_t)id);
    for (int j = 0; j < 12; ++j) {
      const float4 u = ld_global(sh4 + j);
      P.sh[4 * j] = u.x;
      P.sh[4 * j + 1] = u.y;
      P.sh[4 * j + 2] = u.z;
      P.sh[4 * j + 3] = u.w;
    }
    op = 1.f / (1.f + expf(-ld_global(sc.opacity + id)));
  } else {
    for (int i = 0; i < 3; ++i) {
      P.m[i] = 0.f;
      P.s[i] = 1.f;
    }
    P.q[0] = P.q[1] = P.q[2] = 0.f;
    P.q[3] = 1.f;
    P.qlen = 1.f;
    for (int j = 0; j < 48; ++j) P.sh[j] = 0.f;
  }
  quat_rot(P.q, P.R);
  {
    float RS2[9];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) RS2[3 * r + c] = P.R[3 * r + c] * P.s[c] * P.s[c];
    mm3_nt(RS2, P.R, P.Sig);
  }

  float GSig[9] = {}, gm[3] = {}, gsh[48] = {};
  float gop = 0.f;
  bool any = false;
  // The scene's views 64 at a time: lane j reads view j's visibility ballot
  // and compaction offset for this wave at once; the views this wave has
  // visible splats in are then visited in order, the next one's run bounds
  // (soffs, scnt) loaded while the current one is summed and chained.
  for (uint32_t g0 = sc.view_begin; g0 < sc.view_end; g0 += kWave) {
    uint64_t vm_lane = 0;
    uint32_t voff_lane = 0;
    if (g0 + lane < sc.view_end) {
      const size_t wi = (size_t)(L.views[L.scene_views[g0 + lane]].block_base + lb) * kMaskWords + w;
      vm_lane = masks[wi];
      voff_lane = offsets[wi];
    }
    uint64_t todo = __ballot(vm_lane != 0);
    if (todo == 0) continue;
    // run bounds of this lane's slot in view j of the group (c = 0 where not visible)
    auto bounds = [&](uint32_t j, uint64_t& m, uint64_t& o, uint32_t& c) {
      m = ((uint64_t)lane_u32((uint32_t)(vm_lane >> 32), j) << 32) | lane_u32((uint32_t)vm_lane, j);
      const uint32_t base = lane_u32(voff_lane, j);
      o = 0;
      c = 0;
      if ((m >> lane) & 1ull) {
        const uint32_t slot = base + mbcnt(m);
        o = soffs[slot];
        c = scnt[slot];
      }
    };
    uint32_t j = (uint32_t)__builtin_ctzll(todo);
    todo &= todo - 1;
    uint64_t m, o;
    uint32_t c;
    bounds(j, m, o, c);
    while (true) {
      const bool more = todo != 0;
      const uint32_t jn = more ? (uint32_t)__builtin_ctzll(todo) : 0u;
      uint64_t mn = 0, on = 0;
      uint32_t cn = 0;
      if (more) {
        todo &= todo - 1;
        bounds(jn, mn, on, cn);
      }
      const bool vis = (m >> lane) & 1ull;
      if (o + c > max_pairs) c = o < max_pairs ? (uint32_t)(max_pairs - o) : 0u;
      // the span of the wave's runs: from its first visible lane's run start
      // (rounded down to even) to its last visible lane's run end
      const uint32_t first = (uint32_t)__builtin_ctzll(m), lastl = 63u - (uint32_t)__builtin_clzll(m);
      const uint64_t r0 = (((uint64_t)lane_u32((uint32_t)(o >> 32), first) << 32) | lane_u32((uint32_t)o, first)) & ~1ull;
      const uint64_t e = o + c;
      uint64_t r1 = ((uint64_t)lane_u32((uint32_t)(e >> 32), lastl) << 32) | lane_u32((uint32_t)e, lastl);
      r1 = r1 < max_pairs ? r1 : max_pairs;  // runs clipped at max_pairs may start beyond it
      float gv[9];
      sum_slot_runs(gpart, o, c, r0, r1, stage, gv);
      bool nz = false;
      for (int k = 0; k < 9; ++k) nz = nz || gv[k] != 0.f;
      if (vis && nz) {
        any = true;
        gop += gv[3];
        chain_view(L.views[L.scene_views[g0 + j]], P, gv, GSig, gm, gsh);
      }
      if (!more) break;
      j = jn;
      m = mn;
      o = on;
      c = cn;
    }
  }
  if (!valid) return;

  float gls[3] = {0.f, 0.f, 0.f}, gq[4] = {0.f, 0.f, 0.f, 0.f};  // gq in PLY order (w, x, y, z)
  float gopl = 0.f;
  if (any) {
    gopl = gop * op * (1.f - op);
    // ---- Sig = R diag(s^2) R^T -> scales and quaternion ----
    float Gs[9], GSR[9], GR[9];
    for (int i = 0; i < 9; ++i) Gs[i] = 0.5f * (GSig[i] + GSig[3 * (i % 3) + i / 3]);
    mm3(Gs, P.R, GSR);
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) GR[3 * r + c] = 2.f * GSR[3 * r + c] * P.s[c] * P.s[c];
    for (int c = 0; c < 3; ++c) {
      const float rgr = P.R[c] * GSR[c] + P.R[3 + c] * GSR[3 + c] + P.R[6 + c] * GSR[6 + c];  // (R^T Gs R)_cc
      gls[c] = 2.f * rgr * P.s[c] * P.s[c];
    }
    const float qx = P.q[0], qy = P.q[1], qz = P.q[2], qw = P.q[3];
    const float gqx = 2.f * (qy * (GR[1] + GR[3]) + qz * (GR[2] + GR[6]) - 2.f * qx * (GR[4] + GR[8]) +
                             qw * (GR[7] - GR[5]));
    const float gqy = 2.f * (qx * (GR[1] + GR[3]) + qz * (GR[5] + GR[7]) - 2.f * qy * (GR[0] + GR[8]) +
                             qw * (GR[2] - GR[6]));
    const float gqz = 2.f * (qx * (GR[2] + GR[6]) + qy * (GR[5] + GR[7]) - 2.f * qz * (GR[0] + GR[4]) +
                             qw * (GR[3] - GR[1]));
    const float gqw = 2.f * (qx * (GR[7] - GR[5]) + qy * (GR[2] - GR[6]) + qz * (GR[3] - GR[1]));
    const float dot = gqx * qx + gqy * qy + gqz * qz + gqw * qw;
    gq[0] = (gqw - qw * dot) / P.qlen;
    gq[1] = (gqx - qx * dot) / P.qlen;
    gq[2] = (gqy - qy * dot) / P.qlen;
    gq[3] = (gqz - qz * dot) / P.qlen;
  }
  for (int i = 0; i < 3; ++i) {
    st_global(sc.g_means + 3 * (size_t)id + i, gm[i]);
    st_global(sc.g_log_scales + 3 * (size_t)id + i, gls[i]);
  }
  st_global(reinterpret_cast<float4*>(sc.g_quats) + id, make_float4(gq[0], gq[1], gq[2], gq[3]));
  st_global(sc.g_opacity + id, gopl);
  float4* gsh4 = reinterpret_cast<float4*>(sc.g_sh + 48 * (size_t)id);
  for (int j = 0; j < 12; ++j)
    st_global(gsh4 + j, make_float4(gsh[4 * j], gsh[4 * j + 1], gsh[4 * j + 2], gsh[4 * j + 3]));
}

uint32_t grid_for(uint64_t n, uint32_t block, uint32_t cap) {
  uint64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  return (uint32_t)(g < cap ? g : cap);
}

}  // namespace

hipError_t launch_diff_morton_keys(const DiffLaunch& L, uint32_t* bbox_part, uint32_t* bbox, uint32_t* keys,
                                   uint32_t* vals, hipStream_t s) {
  if (L.n_scene_blocks == 0) return hipSuccess;
  uint32_t sbits = 0;
  while ((1u << sbits) < L.n_scenes) ++sbits;
  uint32_t axis_bits = (31u - sbits) / 3u;  // keeps every key below the padding key 0xffffffff
  if (axis_bits > 10u) axis_bits = 10u;
  diff_bbox_kernel<<<L.n_scene_blocks, kCompactBlock, 0, s>>>(L, bbox_part);
  diff_bbox_reduce_kernel<<<L.n_scenes, kCompactBlock, 0, s>>>(L, bbox_part, bbox);
  diff_morton_kernel<<<L.n_scene_blocks, kCompactBlock, 0, s>>>(L, bbox, 3u * axis_bits, axis_bits, keys, vals);
  return hipGetLastError();
}

hipError_t launch_diff_count(const DiffLaunch& L, uint32_t* counts, uint64_t* masks, uint32_t* d_visible,
                             void* scan_storage, hipStream_t s) {
  const uint32_t waves = L.n_blocks * kMaskWords;
  if (waves == 0) return fill_u32_launch(d_visible, 0u, 1, s);
  diff_count_kernel<<<L.n_scene_blocks, kCompactBlock, 0, s>>>(L, counts, masks);
  return scan_exclusive_launch(scan_plan(waves), counts, counts, nullptr, d_visible, scan_storage, s);
}

size_t diff_count_scan_bytes(uint32_t n_blocks) { return scan_plan(n_blocks * kMaskWords).bytes; }

hipError_t launch_diff_write(const DiffLaunch& L, const uint32_t* offsets, const uint64_t* masks, float4* rec,
                             uint4* rc, uint32_t* scnt, uint32_t* key, uint32_t* val, hipStream_t s) {
  if (L.n_blocks == 0) return hipSuccess;
  diff_write_kernel<false><<<L.n_scene_blocks, kCompactBlock, 0, s>>>(L, offsets, masks, rec, rc, scnt, key, val,
                                                                      nullptr, RectBits{0u, 0u});
  return hipGetLastError();
}

hipError_t launch_diff_write_packed(const DiffLaunch& L, const uint32_t* offsets, const uint64_t* masks,
                                    float4* rec, uint32_t* scnt, uint32_t* key, uint64_t* val64, const RectBits& rb,
                                    hipStream_t s) {
  if (L.n_blocks == 0) return hipSuccess;
  diff_write_kernel<true><<<L.n_scene_blocks, kCompactBlock, 0, s>>>(L, offsets, masks, rec, nullptr, scnt, key,
                                                                     nullptr, val64, rb);
  return hipGetLastError();
}

hipError_t launch_diff_view_slots(const DiffLaunch& L, const uint32_t* offsets, const uint32_t* d_count,
                                  uint32_t* vslot, hipStream_t s) {
  diff_view_slots_kernel<<<ceil_div(L.n_views + 1, 256), 256, 0, s>>>(L, offsets, d_count, vslot);
  return hipGetLastError();
}

hipError_t launch_diff_bin_emit_packed(const DiffLaunch& L, const uint64_t* val64, const uint32_t* offs,
                                       const uint32_t* vslot, const RectBits& rb, const uint32_t* d_count,
                                       uint32_t max_n, uint64_t max_pairs, uint32_t* keys, uint32_t* vals,
                                       uint32_t* pslot, const uint32_t* soffs, const uint32_t* scnt, uint32_t* status,
                                       hipStream_t s) {
  if (max_n == 0) return hipSuccess;
  diff_slot_fill_kernel<<<grid_for(max_n, 256, 8192), 256, 0, s>>>(soffs, scnt, d_count, max_n, max_pairs, pslot);
  diff_bin_emit_packed_kernel<<<grid_for(max_n, 256, 8192), 256, 0, s>>>(L, val64, offs, vslot, rb, d_count, max_n,
                                                                         max_pairs, keys, vals, soffs, status);
  return hipGetLastError();
}

hipError_t launch_diff_bin_gather(const uint4* rc, const uint32_t* order, const uint32_t* d_count, uint32_t max_n,
                                  TileRect* rect, uint32_t* cnt, uint32_t* kview, hipStream_t s) {
  if (max_n == 0) return hipSuccess;
  diff_bin_gather_kernel<<<grid_for(max_n, 256, 8192), 256, 0, s>>>(rc, order, d_count, max_n, rect, cnt, kview);
  return hipGetLastError();
}

hipError_t launch_diff_bin_emit(const DiffLaunch& L, const TileRect* rect, const uint32_t* cnt, const uint32_t* offs,
                                const uint32_t* kview, const uint32_t* order, const uint32_t* d_count,
                                uint32_t max_n, uint64_t max_pairs, uint32_t* keys, uint32_t* vals, uint32_t* pslot,
                                const uint32_t* soffs, const uint32_t* scnt, uint32_t* status, hipStream_t s) {
  if (max_n == 0) return hipSuccess;
  diff_slot_fill_kernel<<<grid_for(max_n, 256, 8192), 256, 0, s>>>(soffs, scnt, d_count, max_n, max_pairs, pslot);
  diff_bin_emit_kernel<<<grid_for(max_n, 256, 8192), 256, 0, s>>>(L, rect, cnt, offs, kview, order, d_count, max_n,
                                                                  max_pairs, keys, vals, soffs, status);
  return hipGetLastError();
}

#define VKGS_DIFF_CASES(MACRO) \
  MACRO(16, 16)                \
  MACRO(8, 8)                  \
  MACRO(32, 8)                 \
  MACRO(16, 8)                 \
  MACRO(8, 16)                 \
  MACRO(16, 4)                 \
  MACRO(32, 4)                 \
  MACRO(64, 4)

hipError_t launch_diff_raster(const DiffLaunch& L, const float4* rec, const uint32_t* vals, const uint32_t* ranges,
                              const uint32_t* pslot, const uint32_t* tile_order, bool keep_state, hipStream_t s) {
  if (L.n_tiles == 0) return hipSuccess;
  const uint32_t cap = L.waves_per_cu;
  if (L.tile_w == 8 && L.tile_h == 8 && use_pairs(L.pairs, L.n_tiles)) {
    const size_t pad = occupancy_lds_pad((const void*)diff_raster_pair_kernel, 1, cap);
    diff_raster_pair_kernel<<<(L.n_tiles + 1) / 2, kWave, pad, s>>>(L, rec, vals, ranges, pslot, tile_order,
                                                                    keep_state ? 1 : 0);
    return hipGetLastError();
  }
#define VKGS_CASE(W_, H_)                                                                                      \
  if (L.tile_w == W_ && L.tile_h == H_) {                                                                      \
    const size_t pad = occupancy_lds_pad((const void*)diff_raster_kernel<W_, H_>, W_ * H_ / kWave, cap);       \
    diff_raster_kernel<W_, H_><<<L.n_tiles, W_ * H_, pad, s>>>(L, rec, vals, ranges, pslot, tile_order,        \
                                                                keep_state ? 1 : 0);                           \
    return hipGetLastError();                                                                                  \
  }
  VKGS_DIFF_CASES(VKGS_CASE)
#undef VKGS_CASE
  return hipErrorInvalidValue;
}

namespace {

template <int KB>
void launch_backward_8x8(const DiffLaunch& L, const float4* rec, const uint32_t* vals, const uint32_t* ranges,
                         const uint32_t* pslot, const uint32_t* tile_order, float* gpart, hipStream_t s) {
  if (use_pairs(L.pairs, L.n_tiles)) {
    const size_t pad = occupancy_lds_pad((const void*)diff_raster_backward_pair_kernel<KB>, 1, L.waves_per_cu);
    diff_raster_backward_pair_kernel<KB><<<(L.n_tiles + 1) / 2, kWave, pad, s>>>(L, rec, vals, ranges, pslot,
                                                                                tile_order, gpart);
  } else {
    const size_t pad = occupancy_lds_pad((const void*)diff_raster_backward_kernel<8, 8, KB>, 1, L.waves_per_cu);
    diff_raster_backward_kernel<8, 8, KB><<<L.n_tiles, kWave, pad, s>>>(L, rec, vals, ranges, pslot, tile_order,
                                                                        gpart);
  }
}

}  // namespace

hipError_t launch_diff_raster_backward(const DiffLaunch& L, const float4* rec, const uint32_t* vals,
                                       const uint32_t* ranges, const uint32_t* pslot, const uint32_t* tile_order,
                                       float* gpart, hipStream_t s) {
  // partials of pairs no pixel reached stay zero
  if (L.n_tiles > 0) {
    bool launched = false;
    if (L.tile_w == 8 && L.tile_h == 8) {
      if (L.lds_batch == 8)
        launch_backward_8x8<8>(L, rec, vals, ranges, pslot, tile_order, gpart, s);
      else if (L.lds_batch == 32)
        launch_backward_8x8<32>(L, rec, vals, ranges, pslot, tile_order, gpart, s);
      else
        launch_backward_8x8<16>(L, rec, vals, ranges, pslot, tile_order, gpart, s);
      launched = true;
    }
#define VKGS_CASE(W_, H_)                                                                                      \
  if (!launched && L.tile_w == W_ && L.tile_h == H_) {                                                         \
    const size_t pad =                                                                                         \
        occupancy_lds_pad((const void*)diff_raster_backward_kernel<W_, H_, 0>, W_ * H_ / kWave, L.waves_per_cu); \
    diff_raster_backward_kernel<W_, H_, 0><<<L.n_tiles, W_ * H_, pad, s>>>(L, rec, vals, ranges, pslot,        \
                                                                            tile_order, gpart);                \
    launched = true;                                                                                           \
  }
    VKGS_DIFF_CASES(VKGS_CASE)
#undef VKGS_CASE
    if (!launched) return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_diff_project_backward(const DiffLaunch& L, const uint32_t* offsets, const uint64_t* masks,
                                        const uint32_t* soffs, const uint32_t* scnt, const float* gpart,
                                        uint64_t max_pairs, hipStream_t s) {
  if (L.n_scene_blocks == 0) return hipSuccess;
  diff_project_backward_kernel<<<L.n_scene_blocks, kCompactBlock, 0, s>>>(L, offsets, masks, soffs, scnt, gpart,
                                                                          max_pairs);
  return hipGetLastError();
}

}  // namespace vkgs
