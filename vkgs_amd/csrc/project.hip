// Activation, rank, inverse and projection kernels for gfx950.
//
// One lane per splat.  The splat storage keeps the reference layout
// (engine.cc:485-492): position f32 (N,3), cov3d f32 (N,6), opacity f32 (N),
// sh f16 (N,48); a wave reads 64 consecutive splats, so every load instruction
// touches one contiguous span of each array.
//
// Visible splats are compacted in ascending id order with a wave ballot and a
// workgroup prefix instead of rank.comp:38's atomicAdd: a count kernel writes
// one word per 256 splats, a single-workgroup scan turns them into offsets
// and the write kernel recomputes visibility and stores at
// offset + (visible splats before it in its workgroup).
#pragma clang fp contract(off)

#include "common.h"
#include "device.h"
#include "kernels.h"
#include "sort.h"

namespace vkgs {
namespace {

__device__ __forceinline__ float h2f(uint32_t bits16) {
  return (float)__builtin_bit_cast(_Float16, (uint16_t)bits16);
}

__device__ __forceinline__ uint32_t f2h(float f) {
  return (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)f);
}

__device__ __forceinline__ const FrameConst& fcref(const FrameConst& f) { return f; }
__device__ __forceinline__ const FrameConst& fcref(const FrameConst* f) { return *f; }

// ------------------------------------------------------------ parse_ply ---

__global__ __launch_bounds__(256) void parse_ply_kernel(const uint32_t* __restrict__ ply, uint32_t n,
                                                        float* __restrict__ position, float* __restrict__ cov3d,
                                                        float* __restrict__ opacity, uint16_t* __restrict__ sh) {
  __shared__ uint32_t off[60];
  if (threadIdx.x < 60) off[threadIdx.x] = ply[threadIdx.x];
  __syncthreads();
  const uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= n) return;
  const float* row = reinterpret_cast<const float*>(ply + 60) + (size_t)off[59] * id;
  // parse_ply.comp:46-52
  const float s0 = expf(row[off[3]]), s1 = expf(row[off[4]]), s2 = expf(row[off[5]]);
  float q[4] = {row[off[6]], row[off[7]], row[off[8]], row[off[9]]};
  const float len = sqrtf(((q[0] * q[0] + q[1] * q[1]) + q[2] * q[2]) + q[3] * q[3]);
  for (int k = 0; k < 4; ++k) q[k] = q[k] / len;
  // :54-72
  const float xx = q[0] * q[0], yy = q[1] * q[1], zz = q[2] * q[2];
  const float xy = q[0] * q[1], xz = q[0] * q[2], yz = q[1] * q[2];
  const float wx = q[3] * q[0], wy = q[3] * q[1], wz = q[3] * q[2];
  float rot[9];
  rot[0] = 1.f - 2.f * (yy + zz);
  rot[1] = 2.f * (xy + wz);
  rot[2] = 2.f * (xz - wy);
  rot[3] = 2.f * (xy - wz);
  rot[4] = 1.f - 2.f * (xx + zz);
  rot[5] = 2.f * (yz + wx);
  rot[6] = 2.f * (xz + wy);
  rot[7] = 2.f * (yz - wx);
  rot[8] = 1.f - 2.f * (xx + yy);
  // :134-138
  float ss[9] = {s0 * s0, 0.f, 0.f, 0.f, s1 * s1, 0.f, 0.f, 0.f, s2 * s2};
  float rs[9], rt[9], cov[9];
  m3_mul(rot, ss, rs);
  m3_transpose(rot, rt);
  m3_mul(rs, rt, cov);
  // :140-149
  float* c = cov3d + 6 * (size_t)id;
  c[0] = cov[0];
  c[1] = cov[3];
  c[2] = cov[6];
  c[3] = cov[4];
  c[4] = cov[7];
  c[5] = cov[8];
  position[3 * (size_t)id + 0] = row[off[0]];
  position[3 * (size_t)id + 1] = row[off[1]];
  position[3 * (size_t)id + 2] = row[off[2]];
  // :151-154 fp16, round to nearest even
  uint4* dst = reinterpret_cast<uint4*>(sh + 48 * (size_t)id);
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    uint32_t w[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int i = 8 * j + 2 * k;
      w[k] = f2h(row[off[10 + i]]) | (f2h(row[off[10 + i + 1]]) << 16);
    }
    dst[j] = make_uint4(w[0], w[1], w[2], w[3]);
  }
  // :90, :156
  opacity[id] = 1.f / (1.f + expf(-row[off[58]]));
}

// ----------------------------------------------------------- projection ---

// projection.comp:80-179 for splat id; fills the 48-byte instance.
__device__ __forceinline__ void project_splat(const FrameConst& fc, const float* __restrict__ position,
                                              const float* __restrict__ cov3d, const float* __restrict__ opacity,
                                              const uint16_t* __restrict__ sh, uint32_t id, float4 out[3]) {
  const float2* c2 = reinterpret_cast<const float2*>(cov3d + 6 * (size_t)id);
  const float2 ca = c2[0], cb = c2[1], cc = c2[2];
  // :92 mat3(v0, v0.y, v1.xy, v0.z, v1.yz)
  float S[9] = {ca.x, ca.y, cb.x, ca.y, cb.y, cc.x, cb.x, cc.x, cc.y};
  float p4[4] = {position[3 * (size_t)id + 0], position[3 * (size_t)id + 1], position[3 * (size_t)id + 2], 1.f};
  // :85-87
  const float d0 = p4[0] - fc.eye_model[0], d1 = p4[1] - fc.eye_model[1], d2 = p4[2] - fc.eye_model[2];
  const float dl = sqrtf((d0 * d0 + d1 * d1) + d2 * d2);
  const float dx = d0 / dl, dy = d1 / dl, dz = d2 / dl;
  // :95-102
  float m3[9], t4[4];
  m3_from_m4(fc.M, m3);
  m3_sandwich(m3, S);
  m4_mulv(fc.M, p4, t4);
  m3_from_m4(fc.V, m3);
  m3_sandwich(m3, S);
  m4_mulv(fc.V, t4, p4);
  // :105-108
  const float x = p4[0], y = p4[1], z = p4[2];
  const float r = sqrtf((x * x + y * y) + z * z);
  float J[9] = {-1.f / z, 0.f, (-2.f * x) / r, 0.f, -1.f / z, (-2.f * y) / r, (x / z) / z, (y / z) / z,
                (-2.f * z) / r};
  m3_sandwich(J, S);
  // :111-116
  const float Ps[4] = {fc.P[0], fc.P[1], fc.P[4], fc.P[5]};
  const float S2[4] = {S[0], S[1], S[3], S[4]};
  float cov2[4];
  m2_mul(Ps, S2, cov2);
  m2_mul(cov2, Ps, cov2);
  cov2[0] = cov2[0] + fc.inv_w2;
  cov2[3] = cov2[3] + fc.inv_h2;
  // :121-134, cos/sin(theta) by the half-angle identities
  const float a = cov2[0], b = cov2[3], c = cov2[2];
  const float D = sqrtf((a - b) * (a - b) + (4.f * c) * c);
  const float s0 = sqrtf(0.5f * ((a + b) + D));
  const float s1 = sqrtf(0.5f * ((a + b) - D));
  const float sin2t = (2.f * c) / D;
  const float cos2t = (a - b) / D;
  float ct = (1.f + cos2t) * 0.5f;
  float st = (1.f - cos2t) * 0.5f;
  ct = ct < 0.f ? 0.f : ct;
  st = st < 0.f ? 0.f : st;
  float cos_t = sqrtf(ct);
  float sin_t = copysignf(sqrtf(st), sin2t);
  if (sin2t != sin2t || cos2t != cos2t) {
    cos_t = __builtin_nanf("");
    sin_t = __builtin_nanf("");
  }
  // :136-137
  m4_mulv(fc.P, p4, t4);
  const float nx = t4[0] / t4[3], ny = t4[1] / t4[3], nz = t4[2] / t4[3];
  // :140-171
  const float C0 = 0.28209479177387814f, C1 = 0.4886025119029199f, C20 = 1.0925484305920792f,
              C21 = 0.31539156525252005f, C22 = 0.5462742152960396f, C30 = 0.5900435899266435f,
              C31 = 2.890611442640554f, C32 = 0.4570457994644658f, C33 = 0.3731763325901154f,
              C34 = 1.445305721320277f;
  const float xx = dx * dx, yy = dy * dy, zz = dz * dz, xy = dx * dy, yz = dy * dz, xz = dx * dz;
  float basis[16];
  basis[0] = C0;
  basis[1] = -C1 * dy;
  basis[2] = C1 * dz;
  basis[3] = -C1 * dx;
  basis[4] = C20 * xy;
  basis[5] = -C20 * yz;
  basis[6] = C21 * ((2.f * zz - xx) - yy);
  basis[7] = -C20 * xz;
  basis[8] = C22 * (xx - yy);
  basis[9] = (-C30 * dy) * (3.f * xx - yy);
  basis[10] = (C31 * xy) * dz;
  basis[11] = (-C32 * dy) * ((4.f * zz - xx) - yy);
  basis[12] = (C33 * dz) * ((2.f * zz - 3.f * xx) - 3.f * yy);
  basis[13] = (-C32 * dx) * ((4.f * zz - xx) - yy);
  basis[14] = (C34 * dz) * (xx - yy);
  basis[15] = (-C30 * dx) * (xx - 3.f * yy);
  const uint4* hp = reinterpret_cast<const uint4*>(sh + 48 * (size_t)id);
  float color[3];
#pragma unroll
  for (int ch = 0; ch < 3; ++ch) {
    const uint4 u0 = hp[2 * ch + 0], u1 = hp[2 * ch + 1];
    const uint32_t w[8] = {u0.x, u0.y, u0.z, u0.w, u1.x, u1.y, u1.z, u1.w};
    float acc = 0.f;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint32_t lo = w[2 * j], hi = w[2 * j + 1];
      float dot = basis[4 * j + 0] * h2f(lo & 0xffffu);
      dot = dot + basis[4 * j + 1] * h2f(lo >> 16);
      dot = dot + basis[4 * j + 2] * h2f(hi & 0xffffu);
      dot = dot + basis[4 * j + 3] * h2f(hi >> 16);
      acc = (j == 0) ? dot : acc + dot;
    }
    color[ch] = glsl_max(acc + 0.5f, 0.f);  // :174
  }
  // :177-179
  out[0] = make_float4(nx, ny, nz, 0.f);
  out[1] = make_float4(s0 * cos_t, s0 * sin_t, (-s1) * sin_t, s1 * cos_t);
  // a NaN anywhere in the eigendecomposition (D = 0 for an isotropic cov2d, or
  // a + b < D by rounding for a needle) makes the whole rot_scale the canonical
  // quiet NaN, whose bits do not depend on how NaN signs propagate; such a
  // splat draws nothing either way
  if (!(out[1].x == out[1].x && out[1].y == out[1].y && out[1].z == out[1].z && out[1].w == out[1].w)) {
    const float q = __uint_as_float(0x7fc00000u);
    out[1] = make_float4(q, q, q, q);
  }
  out[2] = make_float4(color[0], color[1], color[2], opacity[id]);
}

// ------------------------------------------------------ ordered compaction ---

template <typename FC>
__global__ __launch_bounds__(kCompactBlock) void visible_count_kernel(FC fcarg, const float* __restrict__ position,
                                                                      uint32_t n, uint32_t* __restrict__ counts) {
  __shared__ uint32_t s_w[kCompactBlock / kWave];
  const FrameConst& fc = fcref(fcarg);
  const uint32_t id = blockIdx.x * kCompactBlock + threadIdx.x;
  bool vis = false;
  if (id < n) {
    float omz;
    vis = rank_visible(fc.PVM, position[3 * (size_t)id], position[3 * (size_t)id + 1], position[3 * (size_t)id + 2],
                       &omz);
  }
  const uint64_t m = __ballot(vis);
  if (__lane_id() == 0) s_w[threadIdx.x / kWave] = (uint32_t)__popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t s = 0;
    for (uint32_t i = 0; i < kCompactBlock / kWave; ++i) s += s_w[i];
    counts[blockIdx.x] = s;
  }
}

// The fused path's counts without a scan launch: one workgroup per group of
// kOffsetGroup compaction workgroups (kOffsetGroup * kCompactBlock splats)
// writes, for each compaction workgroup b of its group, the visible splats of
// the group's workgroups before b (counts[b]), and the group's total
// (group_sums[group]).  Each splat's slot is then the group sums before its
// group + counts[b] + its rank inside b.
constexpr uint32_t kGroupThreads = 512;
constexpr uint32_t kGroupIters = kOffsetGroup * kCompactBlock / kGroupThreads;
static_assert(kOffsetGroup <= kWave && kGroupThreads % kCompactBlock == 0, "group count layout");

template <typename FC>
__global__ __launch_bounds__(kGroupThreads) void visible_group_count_kernel(FC fcarg,
                                                                            const float* __restrict__ position,
                                                                            uint32_t n, uint32_t nb,
                                                                            uint32_t* __restrict__ counts,
                                                                            uint32_t* __restrict__ group_sums) {
  constexpr uint32_t kWaves = kGroupThreads / kWave;
  constexpr uint32_t kPerIter = kGroupThreads / kCompactBlock;  // compaction workgroups per iteration
  constexpr uint32_t kWavesPer = kCompactBlock / kWave;         // waves per compaction workgroup
  __shared__ uint32_t s_c[kGroupIters * kWaves];                // visible splats of wave w in iteration it
  const FrameConst& fc = fcref(fcarg);
  const uint32_t w = threadIdx.x / kWave;
  const uint32_t base = blockIdx.x * kOffsetGroup * kCompactBlock + threadIdx.x;
  float px[kGroupIters], py[kGroupIters], pz[kGroupIters];
#pragma unroll
  for (uint32_t it = 0; it < kGroupIters; ++it) {
    const uint32_t id = base + it * kGroupThreads;
    const uint32_t j = id < n ? id : 0u;
    px[it] = position[3 * (size_t)j];
    py[it] = position[3 * (size_t)j + 1];
    pz[it] = position[3 * (size_t)j + 2];
  }
#pragma unroll
  for (uint32_t it = 0; it < kGroupIters; ++it) {
    float omz;
    const bool vis = base + it * kGroupThreads < n && rank_visible(fc.PVM, px[it], py[it], pz[it], &omz);
    const uint64_t m = __ballot(vis);
    if (__lane_id() == 0) s_c[it * kWaves + w] = (uint32_t)__popcll(m);
  }
  __syncthreads();
  if (threadIdx.x < kWave) {
    const uint32_t j = threadIdx.x;  // compaction workgroup j of the group
    uint32_t v = 0;
    if (j < kOffsetGroup) {
      const uint32_t* c = s_c + (j / kPerIter) * kWaves + (j % kPerIter) * kWavesPer;
#pragma unroll
      for (uint32_t i = 0; i < kWavesPer; ++i) v += c[i];
    }
    uint32_t inc = v;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
      const uint32_t t = (uint32_t)__shfl_up((int)inc, d, kWave);
      if ((int)j >= d) inc += t;
    }
    const uint32_t b = blockIdx.x * kOffsetGroup + j;
    if (j < kOffsetGroup && b < nb) counts[b] = inc - v;
    if (j == kOffsetGroup - 1) group_sums[blockIdx.x] = inc;
  }
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) v += (uint32_t)__shfl_xor((int)v, d, kWave);
  return v;
}

// Slot of this lane's splat if visible, with the workgroup's offset formed
// from visible_group_count_kernel's output: the group sums before this
// workgroup's group (summed by the whole workgroup) plus counts[blockIdx.x].
__device__ __forceinline__ uint32_t compact_slot_grouped(bool vis, const uint32_t* __restrict__ counts,
                                                         const uint32_t* __restrict__ group_sums, uint32_t* s_w,
                                                         uint32_t* s_off) {
  const uint32_t b = blockIdx.x, g = b / kOffsetGroup;
  uint32_t acc = threadIdx.x == 0 ? counts[b] : 0u;
  for (uint32_t i = threadIdx.x; i < g; i += kCompactBlock) acc += group_sums[i];
  acc = wave_sum_u32(acc);
  const uint64_t m = __ballot(vis);
  const uint32_t w = threadIdx.x / kWave;
  if (__lane_id() == 0) {
    s_w[w] = (uint32_t)__popcll(m);
    s_off[w] = acc;
  }
  __syncthreads();
  uint32_t before = 0;
#pragma unroll
  for (uint32_t i = 0; i < kCompactBlock / kWave; ++i) before += s_off[i];
  for (uint32_t i = 0; i < w; ++i) before += s_w[i];
  return before + mbcnt(m);
}

// The total of the group sums (every visible splat) into *d_count: workgroup
// 0 of the fused write, which runs after the count kernel.
__device__ __forceinline__ void write_visible_total(const uint32_t* __restrict__ group_sums, uint32_t groups,
                                                    uint32_t* s_tot, uint32_t* __restrict__ d_count) {
  uint32_t acc = 0;
  for (uint32_t i = threadIdx.x; i < groups; i += kCompactBlock) acc += group_sums[i];
  acc = wave_sum_u32(acc);
  if (__lane_id() == 0) s_tot[threadIdx.x / kWave] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (uint32_t i = 0; i < kCompactBlock / kWave; ++i) t += s_tot[i];
    *d_count = t;
  }
}

// Slot of this lane's splat if visible; offsets[] is the exclusive scan of the
// per-workgroup counts.
__device__ __forceinline__ uint32_t compact_slot(bool vis, const uint32_t* __restrict__ offsets, uint32_t* s_w) {
  const uint64_t m = __ballot(vis);
  const uint32_t w = threadIdx.x / kWave;
  if (__lane_id() == 0) s_w[w] = (uint32_t)__popcll(m);
  __syncthreads();
  uint32_t before = offsets[blockIdx.x];
  for (uint32_t i = 0; i < w; ++i) before += s_w[i];
  return before + mbcnt(m);
}

// key (optional): the rank key per slot; index: the splat id per slot.
// group_sums (optional): offsets and group_sums are visible_group_count_kernel's
// output instead of the scanned offsets.
template <typename FC>
__global__ __launch_bounds__(kCompactBlock) void rank_write_kernel(FC fcarg, const float* __restrict__ position,
                                                                   uint32_t n, const uint32_t* __restrict__ offsets,
                                                                   uint32_t* __restrict__ key,
                                                                   uint32_t* __restrict__ index,
                                                                   const uint32_t* __restrict__ group_sums) {
  __shared__ uint32_t s_w[kCompactBlock / kWave];
  __shared__ uint32_t s_off[kCompactBlock / kWave];
  const FrameConst& fc = fcref(fcarg);
  const uint32_t id = blockIdx.x * kCompactBlock + threadIdx.x;
  bool vis = false;
  float omz = 0.f;
  if (id < n)
    vis = rank_visible(fc.PVM, position[3 * (size_t)id], position[3 * (size_t)id + 1], position[3 * (size_t)id + 2],
                       &omz);
  const uint32_t slot = group_sums ? compact_slot_grouped(vis, offsets, group_sums, s_w, s_off)
                                   : compact_slot(vis, offsets, s_w);
  if (vis) {
    if (key) key[slot] = __float_as_uint(omz);  // rank.comp:39
    index[slot] = id;                           // rank.comp:40
  }
}

__device__ __forceinline__ uint32_t pack16(int lo, int hi) { return (uint32_t)lo | ((uint32_t)hi << 16); }

// Cull + project + compact, and the raster set-up of the projected quad: the
// raster record (48 B) is written at the compacted slot with coalesced stores,
// with the tile rectangle either as (rect, count) words in rc (16 B; binning
// in depth order then gathers them) or, with kPacked, packed into the high
// word of the 64-bit sort value val64 = slot | rect << 32, so that the depth
// sort carries it into depth order and binning reads it sequentially.
template <bool kPacked>
__global__ __launch_bounds__(kCompactBlock) void fused_write_kernel(
    const FrameConst* __restrict__ fcp, const float* __restrict__ position, const float* __restrict__ cov3d,
    const float* __restrict__ opacity, const uint16_t* __restrict__ sh, uint32_t n,
    const uint32_t* __restrict__ offsets, float4* __restrict__ rec, uint4* __restrict__ rc,
    uint32_t* __restrict__ key, uint32_t* __restrict__ val, uint64_t* __restrict__ val64,
    uint32_t* __restrict__ ids, const uint32_t* __restrict__ group_sums, uint32_t groups,
    uint32_t* __restrict__ d_count) {
  __shared__ uint32_t s_w[kCompactBlock / kWave];
  __shared__ uint32_t s_off[kCompactBlock / kWave];
  __shared__ uint32_t s_tot[kCompactBlock / kWave];
  const FrameConst& fc = *fcp;
  if (blockIdx.x == 0) write_visible_total(group_sums, groups, s_tot, d_count);
  const uint32_t id = blockIdx.x * kCompactBlock + threadIdx.x;
  bool vis = false;
  float omz = 0.f;
  if (id < n)
    vis = rank_visible(fc.PVM, position[3 * (size_t)id], position[3 * (size_t)id + 1], position[3 * (size_t)id + 2],
                       &omz);
  const uint32_t slot = compact_slot_grouped(vis, offsets, group_sums, s_w, s_off);
  if (vis) {
    float4 o[3];
    project_splat(fc, position, cov3d, opacity, sh, id, o);
    const float f[12] = {o[0].x, o[0].y, o[0].z, o[0].w, o[1].x, o[1].y, o[1].z, o[1].w,
                         o[2].x, o[2].y, o[2].z, o[2].w};
    Splat2D s = {};
    uint4 r = make_uint4(0u, 0u, 0u, 0u);
    RectBits rb;
    rect_bits(fc.tiles_x, fc.tiles_y, &rb);
    uint32_t packed = pack_rect(1u, 0u, 0u, 0u, rb);  // empty
    if (setup_splat(f, fc.width, fc.height, &s)) {
      const uint32_t shx = 31u - __clz(fc.tile_w), shy = 31u - __clz(fc.tile_h);
      const uint32_t tx0 = (uint32_t)s.x0 >> shx, tx1 = (uint32_t)s.x1 >> shx;
      const uint32_t ty0 = (uint32_t)s.y0 >> shy, ty1 = (uint32_t)s.y1 >> shy;
      r = make_uint4(tx0 | (tx1 << 16), ty0 | (ty1 << 16), (tx1 - tx0 + 1) * (ty1 - ty0 + 1), 0u);
      packed = pack_rect(tx0, tx1, ty0, ty1, rb);
    } else {
      s.x0 = 1;
      s.x1 = 0;
      s.y0 = 1;
      s.y1 = 0;
    }
    raster_record(f, s, fc.triangles != 0, rec + 3 * (size_t)slot);
    if (kPacked) {
      val64[slot] = (uint64_t)slot | ((uint64_t)packed << 32);
    } else {
      rc[slot] = r;
      val[slot] = slot;
    }
    // k = fl(1 - z) * 2^24: an integer in the order and ties of rank.comp:39's float key
    key[slot] = depth_rank_key(omz);
    if (ids) ids[slot] = id;
  }
}

// -------------------------------------------------------------- inverse ---

__global__ void inverse_fill_kernel(int32_t* __restrict__ inverse, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) inverse[i] = -1;
}

__global__ void inverse_scatter_kernel(const uint32_t* __restrict__ d_count, const uint32_t* __restrict__ index,
                                       uint32_t n, int32_t* __restrict__ inverse) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t v = *d_count;
  if (k < v && k < n) {
    const uint32_t id = index[k];
    if (id < n) inverse[id] = (int32_t)k;  // inverse_index.comp:17
  }
}

__global__ __launch_bounds__(256) void projection_kernel(FrameConst fc, const float* __restrict__ position,
                                                         const float* __restrict__ cov3d,
                                                         const float* __restrict__ opacity,
                                                         const uint16_t* __restrict__ sh,
                                                         const int32_t* __restrict__ inverse, uint32_t n,
                                                         const uint32_t* __restrict__ d_count,
                                                         float4* __restrict__ inst,
                                                         uint32_t* __restrict__ draw_indirect) {
  const uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= n) return;
  if (id == 0) {  // projection.comp:64-75
    const uint32_t v = *d_count;
    draw_indirect[0] = 6u * v;
    draw_indirect[1] = 1u;
    draw_indirect[2] = 0u;
    draw_indirect[3] = 0u;
    draw_indirect[4] = 0u;
    draw_indirect[5] = 0u;
    draw_indirect[6] = 0u;
    draw_indirect[7] = 0u;
    draw_indirect[8] = v;
    draw_indirect[9] = 1u;
    draw_indirect[10] = 0u;
    draw_indirect[11] = 0u;
  }
  const int32_t slot = inverse[id];  // :77-78
  if (slot == -1) return;
  float4 o[3];
  project_splat(fc, position, cov3d, opacity, sh, id, o);
  inst[3 * (size_t)slot + 0] = o[0];
  inst[3 * (size_t)slot + 1] = o[1];
  inst[3 * (size_t)slot + 2] = o[2];
}

__global__ void set_frame_const_kernel(FrameConst* __restrict__ dst, FrameConst fc, uint32_t* __restrict__ clear,
                                       uint32_t n_clear) {
  constexpr uint32_t kWords = sizeof(FrameConst) / 4;
  const uint32_t* src = reinterpret_cast<const uint32_t*>(&fc);
  uint32_t* d = reinterpret_cast<uint32_t*>(dst);
  for (uint32_t i = threadIdx.x; i < kWords; i += blockDim.x) d[i] = src[i];
  for (uint32_t i = threadIdx.x; i < n_clear; i += blockDim.x) clear[i] = 0u;
}

}  // namespace

hipError_t launch_set_frame_const(FrameConst* d_fc, const FrameConst& fc, hipStream_t s, uint32_t* clear,
                                  uint32_t n_clear) {
  set_frame_const_kernel<<<1, 128, 0, s>>>(d_fc, fc, clear, clear ? n_clear : 0u);
  return hipGetLastError();
}

hipError_t launch_parse_ply(const void* d_ply, uint32_t n, float* pos, float* cov3d, float* opacity, uint16_t* sh,
                            hipStream_t s) {
  if (n == 0) return hipSuccess;
  parse_ply_kernel<<<ceil_div(n, 256), 256, 0, s>>>(static_cast<const uint32_t*>(d_ply), n, pos, cov3d, opacity,
                                                    sh);
  return hipGetLastError();
}

size_t compact_storage_bytes(uint32_t n) {
  const uint32_t nb = ceil_div(n ? n : 1, kCompactBlock);
  return align_up((size_t)nb * 4, 256) + scan_plan(nb).bytes +
         align_up((size_t)ceil_div(nb, kOffsetGroup) * 4, 256);
}

uint32_t* compact_group_sums(uint32_t* counts, uint32_t n, uint32_t* groups) {
  const uint32_t nb = ceil_div(n ? n : 1, kCompactBlock);
  if (groups) *groups = ceil_div(nb, kOffsetGroup);
  return reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(counts) + align_up((size_t)nb * 4, 256) +
                                     scan_plan(nb).bytes);
}

// Exclusive scan of the per-workgroup visible counts in place; the scan's
// partials live right after the counts (see compact_storage_bytes).
static hipError_t compact_scan(uint32_t* counts, uint32_t nb, uint32_t* d_count, hipStream_t s) {
  const ScanPlan p = scan_plan(nb);
  void* partials = reinterpret_cast<char*>(counts) + align_up((size_t)nb * 4, 256);
  return scan_exclusive_launch(p, counts, counts, nullptr, d_count, partials, s);
}

hipError_t launch_rank(const FrameConst& fc, const float* pos, uint32_t n, uint32_t* d_count, uint32_t* key,
                       uint32_t* index, uint32_t* counts, hipStream_t s) {
  if (n == 0) return fill_u32_launch(d_count, 0u, 1, s);
  const uint32_t nb = ceil_div(n, kCompactBlock);
  visible_count_kernel<FrameConst><<<nb, kCompactBlock, 0, s>>>(fc, pos, n, counts);
  hipError_t e = compact_scan(counts, nb, d_count, s);
  if (e != hipSuccess) return e;
  rank_write_kernel<FrameConst><<<nb, kCompactBlock, 0, s>>>(fc, pos, n, counts, key, index, nullptr);
  return hipGetLastError();
}

hipError_t launch_inverse(const uint32_t* d_count, const uint32_t* index, uint32_t n, int32_t* inverse,
                          hipStream_t s) {
  if (n == 0) return hipSuccess;
  inverse_fill_kernel<<<ceil_div(n, 256), 256, 0, s>>>(inverse, n);
  inverse_scatter_kernel<<<ceil_div(n, 256), 256, 0, s>>>(d_count, index, n, inverse);
  return hipGetLastError();
}

hipError_t launch_projection(const FrameConst& fc, const float* pos, const float* cov3d, const float* opacity,
                             const uint16_t* sh, const int32_t* inverse, uint32_t n, const uint32_t* d_count,
                             float4* inst, uint32_t* draw_indirect, hipStream_t s) {
  if (n == 0) return hipSuccess;
  projection_kernel<<<ceil_div(n, 256), 256, 0, s>>>(fc, pos, cov3d, opacity, sh, inverse, n, d_count, inst,
                                                     draw_indirect);
  return hipGetLastError();
}

hipError_t launch_fused_count(const FrameConst* d_fc, const float* pos, uint32_t n, uint32_t* d_count,
                              uint32_t* counts, hipStream_t s) {
  if (n == 0) return fill_u32_launch(d_count, 0u, 1, s);
  const uint32_t nb = ceil_div(n, kCompactBlock);
  uint32_t groups;
  uint32_t* gs = compact_group_sums(counts, n, &groups);
  visible_group_count_kernel<const FrameConst*><<<groups, kGroupThreads, 0, s>>>(d_fc, pos, n, nb, counts, gs);
  return hipGetLastError();
}

hipError_t launch_compact_ids(const FrameConst* d_fc, const float* pos, uint32_t n, const uint32_t* counts,
                              uint32_t* ids, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const uint32_t nb = ceil_div(n, kCompactBlock);
  const uint32_t* gs = compact_group_sums(const_cast<uint32_t*>(counts), n, nullptr);
  rank_write_kernel<const FrameConst*><<<nb, kCompactBlock, 0, s>>>(d_fc, pos, n, counts, nullptr, ids, gs);
  return hipGetLastError();
}

hipError_t launch_fused_write(const FrameConst* d_fc, const float* pos, const float* cov3d, const float* opacity,
                              const uint16_t* sh, uint32_t n, float4* rec, uint4* rc, uint32_t* key, uint32_t* val,
                              uint32_t* ids, const uint32_t* counts, uint32_t* d_count, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const uint32_t nb = ceil_div(n, kCompactBlock);
  uint32_t groups;
  const uint32_t* gs = compact_group_sums(const_cast<uint32_t*>(counts), n, &groups);
  fused_write_kernel<false><<<nb, kCompactBlock, 0, s>>>(d_fc, pos, cov3d, opacity, sh, n, counts, rec, rc, key, val,
                                                         nullptr, ids, gs, groups, d_count);
  return hipGetLastError();
}

hipError_t launch_fused_write_packed(const FrameConst* d_fc, const float* pos, const float* cov3d,
                                     const float* opacity, const uint16_t* sh, uint32_t n, float4* rec,
                                     uint32_t* key, uint64_t* val64, uint32_t* ids, const uint32_t* counts,
                                     uint32_t* d_count, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const uint32_t nb = ceil_div(n, kCompactBlock);
  uint32_t groups;
  const uint32_t* gs = compact_group_sums(const_cast<uint32_t*>(counts), n, &groups);
  fused_write_kernel<true><<<nb, kCompactBlock, 0, s>>>(d_fc, pos, cov3d, opacity, sh, n, counts, rec, nullptr, key,
                                                        nullptr, val64, ids, gs, groups, d_count);
  return hipGetLastError();
}

}  // namespace vkgs
