// Tile binning and tile rasterizer for gfx950.
//
// Replaces the hardware draw of the reference (splat.vert:10-26,
// splat.frag:8-12, "over" blend engine.cc:279-299).  Instances are binned into
// screen tiles by their conservative pixel rectangle (bin_count, exclusive
// scan, bin_emit), the (tile, instance) pairs are stably sorted by tile so each
// tile's list keeps draw order, and one workgroup per tile composites its list.
//
// Compositing runs front to back (nearest first) with transmittance T:
//   C += c * alpha * T;  A += alpha * alpha * T;  T *= 1 - alpha;  A_out = A + T
// which is the closed form of the reference's back-to-front blend with clear
// colour (0,0,0,1).  A pixel stops once T < t_min (t_min = 0 composites every
// splat); the skipped splats contribute at most t_min * max(colour, 1).
// Each batch of up to tile_w*tile_h instances is set up once (one lane per
// instance) into LDS and then read by every lane as a broadcast; a wave skips
// an instance whose rectangle misses its rows with a scalar branch.
#pragma clang fp contract(off)

#include "common.h"
#include "kernels.h"

namespace vkgs {
namespace {

__global__ __launch_bounds__(256) void bin_count_kernel(const float4* __restrict__ inst,
                                                        const uint32_t* __restrict__ order,
                                                        const uint32_t* __restrict__ d_count, uint32_t max_n,
                                                        RasterShape rs, TileRect* __restrict__ rect,
                                                        uint32_t* __restrict__ cnt) {
  const uint32_t n = min(*d_count, max_n);
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    const uint32_t idx = order ? order[k] : k;
    const float4 a = inst[3 * (size_t)idx], b = inst[3 * (size_t)idx + 1], c = inst[3 * (size_t)idx + 2];
    const float f[12] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, c.x, c.y, c.z, c.w};
    Splat2D s;
    uint32_t count = 0;
    TileRect r = {0u, 0u};
    if (setup_splat(f, rs.width, rs.height, &s)) {
      const uint32_t tx0 = (uint32_t)s.x0 >> rs.tile_shift_x, tx1 = (uint32_t)s.x1 >> rs.tile_shift_x;
      const uint32_t ty0 = (uint32_t)s.y0 >> rs.tile_shift_y, ty1 = (uint32_t)s.y1 >> rs.tile_shift_y;
      r.x = tx0 | (tx1 << 16);
      r.y = ty0 | (ty1 << 16);
      count = (tx1 - tx0 + 1) * (ty1 - ty0 + 1);
    }
    rect[k] = r;
    cnt[k] = count;
  }
}

__global__ __launch_bounds__(256) void bin_emit_kernel(const TileRect* __restrict__ rect,
                                                       const uint32_t* __restrict__ cnt,
                                                       const uint32_t* __restrict__ offs,
                                                       const uint32_t* __restrict__ order,
                                                       const uint32_t* __restrict__ d_count, uint32_t max_n,
                                                       uint32_t tiles_x, uint64_t max_pairs,
                                                       uint32_t* __restrict__ keys, uint32_t* __restrict__ vals,
                                                       uint32_t* __restrict__ status) {
  const uint32_t n = min(*d_count, max_n);
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    const uint32_t c = cnt[k];
    if (c == 0) continue;
    const uint64_t o = offs[k];
    const TileRect r = rect[k];
    const uint32_t idx = order ? order[k] : k;
    if (o + c > max_pairs) atomicOr(status, 1u);
    const uint32_t tx0 = r.x & 0xffffu, tx1 = r.x >> 16, ty0 = r.y & 0xffffu, ty1 = r.y >> 16;
    uint64_t j = o;
    for (uint32_t ty = ty0; ty <= ty1; ++ty)
      for (uint32_t tx = tx0; tx <= tx1; ++tx, ++j) {
        if (j < max_pairs) {
          keys[j] = ty * tiles_x + tx;
          vals[j] = idx;
        }
      }
  }
}

__global__ void ranges_fill_kernel(uint32_t* __restrict__ ranges, uint32_t n2) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n2) ranges[i] = 0u;
}

__global__ void ranges_mark_kernel(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ d_pairs,
                                   uint64_t max_pairs, uint32_t* __restrict__ ranges) {
  const uint64_t K = min((uint64_t)*d_pairs, max_pairs);
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < K; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t t = keys[i];
    if (i == 0 || keys[i - 1] != t) ranges[2 * t] = (uint32_t)i;
    if (i + 1 == K || keys[i + 1] != t) ranges[2 * t + 1] = (uint32_t)(i + 1);
  }
}

template <int TW, int TH>
__global__ __launch_bounds__(TW* TH) void rasterize_kernel(const float4* __restrict__ inst,
                                                           const uint32_t* __restrict__ vals,
                                                           const uint32_t* __restrict__ ranges, RasterShape rs,
                                                           float4* __restrict__ out) {
  constexpr int kN = TW * TH;
  constexpr int kRowsPerWave = kWave / TW;
  __shared__ float4 s_a[kN];  // cx, cy, i00, i01
  __shared__ float4 s_b[kN];  // i10, i11, opacity, -
  __shared__ float4 s_c[kN];  // r, g, b, -
  __shared__ int4 s_box[kN];  // x0, x1, y0, y1 (pixels)

  const uint32_t tile = blockIdx.x;
  const uint32_t tx = tile % rs.tiles_x, ty = tile / rs.tiles_x;
  const int lx = threadIdx.x % TW, ly = threadIdx.x / TW;
  const int px = (int)tx * TW + lx, py = (int)ty * TH + ly;
  const bool inside = px < (int)rs.width && py < (int)rs.height;
  const int col0 = (int)tx * TW, col1 = col0 + TW - 1;
  const int row0 = (int)ty * TH + (int)(threadIdx.x / kWave) * kRowsPerWave, row1 = row0 + kRowsPerWave - 1;
  const uint32_t start = ranges[2 * tile], end = ranges[2 * tile + 1];
  const float fx = (float)px + 0.5f, fy = (float)py + 0.5f;
  const float t_min = rs.t_min;

  float T = 1.f, cr = 0.f, cg = 0.f, cb = 0.f, aa = 0.f;
  bool done = !inside;
  for (uint32_t hi = end; hi > start;) {
    const uint32_t lo = hi - start > (uint32_t)kN ? hi - kN : start;
    const int cnt = (int)(hi - lo);
    __syncthreads();
    if ((int)threadIdx.x < cnt) {
      const uint32_t idx = vals[hi - 1 - threadIdx.x];  // nearest first
      const float4 a = inst[3 * (size_t)idx], b = inst[3 * (size_t)idx + 1], c = inst[3 * (size_t)idx + 2];
      const float f[12] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, c.x, c.y, c.z, c.w};
      Splat2D s;
      if (!setup_splat(f, rs.width, rs.height, &s)) {  // never binned; keep the slot inert
        s.x0 = 1;
        s.x1 = 0;
        s.y0 = 1;
        s.y1 = 0;
      }
      s_a[threadIdx.x] = make_float4(s.cx, s.cy, s.i00, s.i01);
      s_b[threadIdx.x] = make_float4(s.i10, s.i11, s.o, 0.f);
      s_c[threadIdx.x] = make_float4(s.r, s.g, s.b, 0.f);
      s_box[threadIdx.x] = make_int4(s.x0, s.x1, s.y0, s.y1);
    }
    __syncthreads();
    if (!done) {
      for (int e = 0; e < cnt; ++e) {
        const int4 box = s_box[e];
        const int bx0 = __builtin_amdgcn_readfirstlane(box.x), bx1 = __builtin_amdgcn_readfirstlane(box.y);
        const int by0 = __builtin_amdgcn_readfirstlane(box.z), by1 = __builtin_amdgcn_readfirstlane(box.w);
        if (by1 < row0 || by0 > row1 || bx1 < col0 || bx0 > col1) continue;
        const float4 a = s_a[e];
        const float4 b = s_b[e];
        const float dx = fx - a.x, dy = fy - a.y;
        const float ux = fmaf(a.z, dx, a.w * dy);
        const float uy = fmaf(b.x, dx, b.y * dy);
        if (fabsf(ux) <= 3.f && fabsf(uy) <= 3.f) {
          const float q = fmaf(ux, ux, uy * uy);
          const float alpha = b.z * __expf(-0.5f * q);
          const float4 c = s_c[e];
          const float ta = T * alpha;
          cr = fmaf(c.x, ta, cr);
          cg = fmaf(c.y, ta, cg);
          cb = fmaf(c.z, ta, cb);
          aa = fmaf(alpha, ta, aa);
          T = T - ta;
          if (T < t_min) {
            done = true;
            break;
          }
        }
      }
    }
    hi = lo;
    if (__syncthreads_count(!done) == 0) break;
  }
  if (inside) out[(size_t)py * rs.width + px] = make_float4(cr, cg, cb, aa + T);
}

__global__ void gather_u32_kernel(const uint32_t* __restrict__ src, const uint32_t* __restrict__ idx,
                                  const uint32_t* __restrict__ d_count, uint32_t max_n, uint32_t* __restrict__ out) {
  const uint32_t n = min(*d_count, max_n);
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) out[k] = src[idx[k]];
}

uint32_t grid_for(uint64_t n, uint32_t block, uint32_t cap) {
  uint64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  return (uint32_t)(g < cap ? g : cap);
}

int ilog2(uint32_t v) {
  int s = 0;
  while ((1u << s) < v) ++s;
  return s;
}

}  // namespace

bool make_raster_shape(uint32_t width, uint32_t height, const vkgs_raster_config* cfg, RasterShape* rs) {
  vkgs_raster_config d;
  if (cfg == nullptr) {
    vkgs_raster_config_default(&d);
    cfg = &d;
  }
  const uint32_t tw = cfg->tile_w, th = cfg->tile_h;
  const bool ok_shape = (tw == 8 && th == 8) || (tw == 16 && th == 16) || (tw == 32 && th == 8) ||
                        (tw == 16 && th == 8) || (tw == 8 && th == 16) || (tw == 16 && th == 4) ||
                        (tw == 32 && th == 4) || (tw == 64 && th == 4);
  if (!ok_shape || width == 0 || height == 0 || width > 32768 || height > 32768) return false;
  if (!(cfg->t_min >= 0.f && cfg->t_min < 1.f)) return false;
  rs->width = width;
  rs->height = height;
  rs->tile_w = tw;
  rs->tile_h = th;
  rs->tiles_x = (width + tw - 1) / tw;
  rs->tiles_y = (height + th - 1) / th;
  rs->tile_shift_x = (uint32_t)ilog2(tw);
  rs->tile_shift_y = (uint32_t)ilog2(th);
  rs->t_min = cfg->t_min;
  return true;
}

hipError_t launch_bin_count(const float4* inst, const uint32_t* order, const uint32_t* d_count, uint32_t max_n,
                            const RasterShape& rs, TileRect* rect, uint32_t* cnt, hipStream_t s) {
  if (max_n == 0) return hipSuccess;
  bin_count_kernel<<<grid_for(max_n, 256, 8192), 256, 0, s>>>(inst, order, d_count, max_n, rs, rect, cnt);
  return hipGetLastError();
}

hipError_t launch_bin_emit(const TileRect* rect, const uint32_t* cnt, const uint32_t* offs, const uint32_t* order,
                           const uint32_t* d_count, uint32_t max_n, const RasterShape& rs, uint64_t max_pairs,
                           uint32_t* keys, uint32_t* vals, uint32_t* status, hipStream_t s) {
  if (max_n == 0) return hipSuccess;
  bin_emit_kernel<<<grid_for(max_n, 256, 8192), 256, 0, s>>>(rect, cnt, offs, order, d_count, max_n, rs.tiles_x,
                                                              max_pairs, keys, vals, status);
  return hipGetLastError();
}

hipError_t launch_ranges(const uint32_t* keys, const uint32_t* d_pairs, uint64_t max_pairs, uint32_t tiles,
                         uint32_t* ranges, hipStream_t s) {
  ranges_fill_kernel<<<grid_for(2ull * tiles, 256, 1u << 20), 256, 0, s>>>(ranges, 2 * tiles);
  if (max_pairs > 0)
    ranges_mark_kernel<<<grid_for(max_pairs, 256, 8192), 256, 0, s>>>(keys, d_pairs, max_pairs, ranges);
  return hipGetLastError();
}

hipError_t launch_rasterize(const float4* inst, const uint32_t* vals, const uint32_t* ranges, const RasterShape& rs,
                            float4* out, hipStream_t s) {
  const uint32_t tiles = rs.tiles_x * rs.tiles_y;
  const uint32_t tw = rs.tile_w, th = rs.tile_h;
#define VKGS_RASTER_CASE(W_, H_)                                                          \
  if (tw == W_ && th == H_) {                                                             \
    rasterize_kernel<W_, H_><<<tiles, W_ * H_, 0, s>>>(inst, vals, ranges, rs, out);      \
    return hipGetLastError();                                                             \
  }
  VKGS_RASTER_CASE(16, 16)
  VKGS_RASTER_CASE(8, 8)
  VKGS_RASTER_CASE(32, 8)
  VKGS_RASTER_CASE(16, 8)
  VKGS_RASTER_CASE(8, 16)
  VKGS_RASTER_CASE(16, 4)
  VKGS_RASTER_CASE(32, 4)
  VKGS_RASTER_CASE(64, 4)
#undef VKGS_RASTER_CASE
  return hipErrorInvalidValue;
}

hipError_t launch_gather_u32(const uint32_t* src, const uint32_t* idx, const uint32_t* d_count, uint32_t max_n,
                             uint32_t* out, hipStream_t s) {
  if (max_n == 0) return hipSuccess;
  gather_u32_kernel<<<grid_for(max_n, 256, 8192), 256, 0, s>>>(src, idx, d_count, max_n, out);
  return hipGetLastError();
}

}  // namespace vkgs
