// Tile binning and tile rasterizer for gfx950.
//
// Replaces the hardware draw of the reference (splat.vert:10-26,
// splat.frag:8-12, "over" blend engine.cc:279-299).
//
// bin_count  : one lane per draw-order position k: raster set-up of instance
//              order[k] (centre, inverse pixel-space rot_scale, colour, pixel
//              rectangle) into a 48-byte raster record rec[k], plus its tile
//              rectangle and tile count (the fused renderer writes the record
//              at projection time instead and carries the packed rectangle in
//              the depth sort's value: bin_group_count + bin_emit_grouped,
//              which form the pair offsets without a scan launch);
// scan       : exclusive scan of the tile counts;
// bin_emit   : (tile, k) pairs, tiles of one splat in row-major order (the
//              launch also zeroes the tile ranges);
// tile sort  : stable radix sort by tile, so each tile's list keeps draw order;
// ranges     : [start, end) of every tile; then the longest-first schedule;
// rasterize  : one workgroup per tile, one lane per pixel.  Records are staged
//              in LDS in batches of one per lane; the next batch is gathered
//              into registers while the current one is composited, so global
//              latency overlaps compositing and there is one barrier per batch.
//              A wave skips a record whose rectangle misses its pixel rows with
//              a scalar branch.
//
// Compositing runs front to back (nearest first) with transmittance T:
//   C += c * alpha * T;  A += alpha * alpha * T;  T *= 1 - alpha;  A_out = A + T
// which is the closed form of the reference's back-to-front blend with clear
// colour (0,0,0,1).  A pixel stops once T < t_min (t_min = 0 composites every
// splat); the skipped splats contribute at most t_min * max(colour, 1).
#pragma clang fp contract(off)

#include "common.h"
#include "device.h"
#include "kernels.h"
#include "stage.h"

#include <mutex>

namespace vkgs {
namespace {

__global__ __launch_bounds__(256) void bin_count_kernel(const float4* __restrict__ inst,
                                                        const uint32_t* __restrict__ order,
                                                        const uint32_t* __restrict__ d_count, uint32_t max_n,
                                                        RasterShape rs, float4* __restrict__ rec,
                                                        TileRect* __restrict__ rect, uint32_t* __restrict__ cnt) {
  const uint32_t n = min(*d_count, max_n);
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    const uint32_t idx = order ? order[k] : k;
    const float4 a = inst[3 * (size_t)idx], b = inst[3 * (size_t)idx + 1], c = inst[3 * (size_t)idx + 2];
    const float f[12] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, c.x, c.y, c.z, c.w};
    Splat2D s = {};
    uint32_t count = 0;
    TileRect r = {0u, 0u};
    if (setup_splat(f, rs.width, rs.height, &s)) {
      const uint32_t tx0 = (uint32_t)s.x0 >> rs.tile_shift_x, tx1 = (uint32_t)s.x1 >> rs.tile_shift_x;
      const uint32_t ty0 = (uint32_t)s.y0 >> rs.tile_shift_y, ty1 = (uint32_t)s.y1 >> rs.tile_shift_y;
      r.x = tx0 | (tx1 << 16);
      r.y = ty0 | (ty1 << 16);
      count = (tx1 - tx0 + 1) * (ty1 - ty0 + 1);
    } else {  // never binned; an empty rectangle keeps the record inert
      s.x0 = 1;
      s.x1 = 0;
      s.y0 = 1;
      s.y1 = 0;
    }
    raster_record(f, s, rs.triangles != 0, rec + 3 * (size_t)k);
    rect[k] = r;
    cnt[k] = count;
  }
}

__global__ __launch_bounds__(256) void bin_gather_kernel(const uint4* __restrict__ rc,
                                                         const uint32_t* __restrict__ order,
                                                         const uint32_t* __restrict__ d_count, uint32_t max_n,
                                                         TileRect* __restrict__ rect, uint32_t* __restrict__ cnt) {
  const uint32_t n = min(*d_count, max_n);
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    const uint4 v = rc[order[k]];
    rect[k] = TileRect{v.x, v.y};
    cnt[k] = v.z;
  }
}

__global__ __launch_bounds__(256) void bin_emit_kernel(const TileRect* __restrict__ rect,
                                                       const uint32_t* __restrict__ cnt,
                                                       const uint32_t* __restrict__ offs,
                                                       const uint32_t* __restrict__ order,
                                                       const uint32_t* __restrict__ d_count, uint32_t max_n,
                                                       uint32_t tiles_x, uint64_t max_pairs,
                                                       uint32_t* __restrict__ keys, uint32_t* __restrict__ vals,
                                                       uint32_t* __restrict__ status, uint32_t* __restrict__ clear,
                                                       uint32_t n_clear) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_clear; i += gridDim.x * blockDim.x) clear[i] = 0u;
  const uint32_t n = min(*d_count, max_n);
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const uint32_t stride = gridDim.x * blockDim.x;
  // wave-uniform loop: emit_pairs_wave needs every lane
  for (uint32_t base = wave * kWave; base < n; base += stride) {
    const uint32_t k = base + __lane_id();
    uint32_t c = 0, v = 0;
    uint64_t o = 0;
    TileRect r = {0u, 0u};
    if (k < n) {
      c = cnt[k];
      if (c != 0) {
        o = offs[k];
        r = rect[k];
        v = order ? order[k] : k;
        if (o + c > max_pairs) atomicOr(status, 1u);
      }
    }
    emit_pairs_wave(c, o, r, v, 0u, tiles_x, max_pairs, keys, vals);
  }
}

// Binning of depth-sorted packed values without a scan launch.  One
// workgroup per kBinGroup emit workgroups (kBinGroup x 256 items) sums each
// emit workgroup's tile counts and writes, per emit workgroup, the pairs of
// the group's earlier ones (boff) and the group's total (gsum); an emit
// workgroup then adds the group totals before its group to its first unit's
// boff and scans its own units' counts.  Integer sums: the offsets equal the
// exclusive scan's.
constexpr uint32_t kBinGroup = 16;
constexpr uint32_t kBinGroupThreads = 512;
constexpr uint32_t kBinGroupIters = kBinGroup * 256 / kBinGroupThreads;
static_assert(kBinGroup <= kWave && kBinGroupThreads % 256 == 0, "bin group layout");

__device__ __forceinline__ uint32_t bin_wave_sum(uint32_t v) {
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) v += (uint32_t)__shfl_xor((int)v, d, kWave);
  return v;
}

__global__ __launch_bounds__(kBinGroupThreads) void bin_group_count_kernel(const uint64_t* __restrict__ val64,
                                                                           const uint32_t* __restrict__ d_count,
                                                                           uint32_t max_n, RectBits rb, uint32_t nb,
                                                                           uint32_t* __restrict__ boff,
                                                                           uint32_t* __restrict__ gsum) {
  constexpr uint32_t kWaves = kBinGroupThreads / kWave;
  constexpr uint32_t kPerIter = kBinGroupThreads / 256;  // emit workgroups per iteration
  __shared__ uint32_t s_c[kBinGroupIters * kWaves];      // tiles of wave w's items in iteration it
  const uint32_t n = min(*d_count, max_n);
  const uint32_t w = threadIdx.x / kWave;
  const uint32_t base = blockIdx.x * kBinGroup * 256 + threadIdx.x;
  uint64_t pv[kBinGroupIters];
#pragma unroll
  for (uint32_t it = 0; it < kBinGroupIters; ++it) {
    const uint32_t k = base + it * kBinGroupThreads;
    pv[it] = k < n ? val64[k] : 0ull;
  }
#pragma unroll
  for (uint32_t it = 0; it < kBinGroupIters; ++it) {
    uint32_t c = 0;
    if (base + it * kBinGroupThreads < n) (void)unpack_rect((uint32_t)(pv[it] >> 32), rb, &c);
    c = bin_wave_sum(c);
    if (__lane_id() == 0) s_c[it * kWaves + w] = c;
  }
  __syncthreads();
  if (threadIdx.x < kWave) {
    const uint32_t j = threadIdx.x;  // emit workgroup j of the group
    uint32_t v = 0;
    if (j < kBinGroup) {
      const uint32_t* c = s_c + (j / kPerIter) * kWaves + (j % kPerIter) * (256 / kWave);
#pragma unroll
      for (uint32_t i = 0; i < 256 / kWave; ++i) v += c[i];
    }
    uint32_t inc = v;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
      const uint32_t t = (uint32_t)__shfl_up((int)inc, d, kWave);
      if ((int)j >= d) inc += t;
    }
    const uint32_t b = blockIdx.x * kBinGroup + j;
    if (j < kBinGroup && b < nb) boff[b] = inc - v;
    if (j == kBinGroup - 1) gsum[blockIdx.x] = inc;
  }
}

// Pairs (tile, position) of the depth-sorted packed values val64[k] = k's
// record slot | packed tile rectangle << 32, one item per lane (its tiles in
// row-major order at consecutive pair slots, emit_pairs_wave), `chunks` (1 or
// kBinChunks) units of 256 consecutive items per workgroup: the first unit's
// offset from bin_group_count_kernel's output, the rest by the workgroup's
// running scan.  Several units per workgroup amortize the offset sums on large
// draws; one keeps enough waves emitting on small ones.
// Workgroup 0 stores the pair total in *d_pairs; the first n_clear words at
// clear are zeroed on the way (the tile ranges of this draw).
constexpr uint32_t kBinChunks = 4;
static_assert(kBinGroup % kBinChunks == 0, "a workgroup's units share a group");

__global__ __launch_bounds__(256) void bin_emit_grouped_kernel(
    const uint64_t* __restrict__ val64, const uint32_t* __restrict__ d_count, uint32_t max_n, uint32_t tiles_x,
    RectBits rb, uint64_t max_pairs, const uint32_t* __restrict__ boff, const uint32_t* __restrict__ gsum,
    uint32_t groups, uint32_t chunks, uint32_t* __restrict__ keys, uint32_t* __restrict__ vals,
    uint32_t* __restrict__ d_pairs, uint32_t* __restrict__ status, uint32_t* __restrict__ clear, uint32_t n_clear) {
  constexpr uint32_t kWaves = 256 / kWave;
  __shared__ uint32_t s_tot[kWaves];
  __shared__ uint32_t s_base[kWaves];
  __shared__ uint32_t s_wsum[2][kWaves];
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_clear; i += gridDim.x * blockDim.x) clear[i] = 0u;
  const uint32_t n = min(*d_count, max_n);
  const uint32_t w = threadIdx.x / kWave, lane = __lane_id();
  if (blockIdx.x == 0) {  // the pair total: every group's sum
    uint32_t t = 0;
    for (uint32_t i = threadIdx.x; i < groups; i += 256) t += gsum[i];
    t = bin_wave_sum(t);
    if (lane == 0) s_tot[w] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t all = 0;
#pragma unroll
      for (uint32_t i = 0; i < kWaves; ++i) all += s_tot[i];
      *d_pairs = all;
    }
  }
  const uint32_t u0 = blockIdx.x * chunks;
  if (u0 * 256 >= n) return;  // workgroup-uniform
  uint64_t pw[kBinChunks];
#pragma unroll
  for (uint32_t q = 0; q < kBinChunks; ++q) {
    const uint32_t k = (u0 + q) * 256 + threadIdx.x;
    pw[q] = q < chunks && k < n ? val64[k] : 0ull;
  }
  // the first unit's first pair: group totals before its group + its offset in the group
  uint32_t acc = threadIdx.x == 0 ? boff[u0] : 0u;
  for (uint32_t i = threadIdx.x; i < u0 / kBinGroup; i += 256) acc += gsum[i];
  acc = bin_wave_sum(acc);
  if (lane == 0) s_base[w] = acc;
  __syncthreads();
  uint32_t running = 0;
#pragma unroll
  for (uint32_t i = 0; i < kWaves; ++i) running += s_base[i];
#pragma unroll
  for (uint32_t q = 0; q < kBinChunks; ++q) {
    if (q >= chunks || (u0 + q) * 256 >= n) break;  // workgroup-uniform
    const uint32_t k = (u0 + q) * 256 + threadIdx.x;
    uint32_t c = 0;
    TileRect r = {0u, 0u};
    if (k < n) r = unpack_rect((uint32_t)(pw[q] >> 32), rb, &c);
    uint32_t inc = c;  // inclusive scan of the wave's counts
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
      const uint32_t t = (uint32_t)__shfl_up((int)inc, d, kWave);
      if ((int)lane >= d) inc += t;
    }
    // two buffers: a wave writes unit q + 2's sums only after every wave has
    // passed unit q + 1's barrier, which follows their reads of unit q's
    if (lane == kWave - 1) s_wsum[q & 1][w] = inc;
    __syncthreads();
    uint32_t o = running, unit = 0;
#pragma unroll
    for (uint32_t i = 0; i < kWaves; ++i) {
      const uint32_t x = s_wsum[q & 1][i];
      o += i < w ? x : 0u;
      unit += x;
    }
    o += inc - c;
    running += unit;
    if (c != 0 && (uint64_t)o + c > max_pairs) atomicOr(status, 1u);
    emit_pairs_wave(c, (uint64_t)o, r, (uint32_t)pw[q], 0u, tiles_x, max_pairs, keys, vals);
  }
}

__global__ void ranges_fill_kernel(uint32_t* __restrict__ ranges, uint32_t n2) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n2) ranges[i] = 0u;
}

// A wave takes kRows x 64 consecutive keys per step, all loads in flight
// before any compare; the neighbours come from the adjacent lanes (and the
// adjacent row's end lanes), so only the two keys around the step are loaded
// twice.  Positions before 0 or at K compare as kNoTile.
__global__ __launch_bounds__(256) void ranges_mark_kernel(const uint32_t* __restrict__ keys,
                                                          const uint32_t* __restrict__ d_pairs, uint64_t max_pairs,
                                                          uint32_t* __restrict__ ranges,
                                                          const uint32_t* __restrict__ view_tiles,
                                                          uint32_t local_bits) {
  constexpr int kRows = 4;
  constexpr uint32_t kNoTile = 0xffffffffu;  // tile ids are below 2^24
  const uint64_t K = min((uint64_t)*d_pairs, max_pairs);
  const uint32_t lane = __lane_id();
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const uint64_t step = (uint64_t)gridDim.x * blockDim.x / kWave * (kRows * kWave);
  for (uint64_t base = wave * (kRows * kWave); base < K; base += step) {
    uint32_t t[kRows];
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
      const uint64_t i = base + r * kWave + lane;
      t[r] = i < K ? keys[i] : kNoTile;
    }
    const uint32_t before = __builtin_amdgcn_readfirstlane(lane == 0 && base > 0 ? keys[base - 1] : kNoTile);
    const uint64_t after_i = base + kRows * kWave;
    const uint32_t after = __builtin_amdgcn_readfirstlane(lane == 0 && after_i < K ? keys[after_i] : kNoTile);
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
      const uint32_t row_prev = r == 0 ? before : __builtin_amdgcn_readlane(t[r > 0 ? r - 1 : 0], kWave - 1);
      const uint32_t row_next = r == kRows - 1 ? after : __builtin_amdgcn_readlane(t[r < kRows - 1 ? r + 1 : r], 0);
      uint32_t prev = __shfl_up(t[r], 1, kWave), next = __shfl_down(t[r], 1, kWave);
      if (lane == 0) prev = row_prev;
      if (lane == kWave - 1) next = row_next;
      const uint64_t i = base + r * kWave + lane;
      if (i < K && (prev != t[r] || next != t[r])) {
        // a view << local_bits | local tile key names tile view_tiles[view] + local tile
        const uint32_t g = view_tiles ? view_tiles[t[r] >> local_bits] + (t[r] & ((1u << local_bits) - 1u)) : t[r];
        if (prev != t[r]) ranges[2 * g] = (uint32_t)i;
        if (next != t[r]) ranges[2 * g + 1] = (uint32_t)(i + 1);
      }
    }
  }
}

// Sort key of each tile for the longest-first schedule: 255 - a log-spaced
// bucket of the list length (8 buckets per octave: 8 e + the 3 bits below the
// leading one of len + 1, e its exponent), value = tile id; an ascending
// stable one-pass sort puts long lists first.  Lengths within 1/8 of an
// octave share a bucket, which is fine for a schedule.
__global__ void tile_order_keys_kernel(const uint32_t* __restrict__ ranges, uint32_t tiles,
                                       uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= tiles) return;
  const uint32_t v = ranges[2 * t + 1] - ranges[2 * t] + 1u;  // >= 1 (a list holds fewer than 2^32 pairs)
  const uint32_t e = 31u - __clz(v);
  const uint32_t m = (e >= 3u ? v >> (e - 3u) : v << (3u - e)) & 7u;
  keys[t] = 255u - (8u * e + m);
  vals[t] = t;
}

// One (pixel, record) step of front-to-back compositing; (lx, ly) is the pixel
// centre relative to its block corner.  Per step: 2 FMAs per eigenframe
// coordinate, one max for the square, 2 FMAs and exp2 for alpha (the opacity
// is inside the exponent), a compare and select for coverage, T * alpha, four
// accumulating FMAs and the transmittance update.
__device__ __forceinline__ void composite_at(float ux, float uy, const float4& a, const float4& c, float& T, float& cr,
                                             float& cg, float& cb, float& aa) {
  const float m = fmaxf(fabsf(ux), fabsf(uy));
  const float e = __builtin_amdgcn_exp2f(-fmaf(ux, ux, fmaf(uy, uy, a.w)));
  const float alpha = m <= kCoverScaled ? e : 0.f;
  const float ta = T * alpha;
  cr = fmaf(c.x, ta, cr);
  cg = fmaf(c.y, ta, cg);
  cb = fmaf(c.z, ta, cb);
  aa = fmaf(alpha, ta, aa);
  T = T - ta;
}

__device__ __forceinline__ void composite(const float4& a, const float4& b, const float4& c, float lx, float ly,
                                          float& T, float& cr, float& cg, float& cb, float& aa) {
  composite_at(fmaf(a.x, lx, fmaf(a.y, ly, a.z)), fmaf(b.x, lx, fmaf(b.y, ly, b.z)), a, c, T, cr, cg, cb, aa);
}

// List entries a one-wave tile composites between two exit tests: a pixel may
// composite up to three entries past T < t_min, which only adds terms of the
// exact composite.
constexpr int kStepGroup = 4;

template <int TW, int TH>
__global__ __launch_bounds__(TW* TH) void rasterize_kernel(const float4* __restrict__ rec,
                                                           const uint32_t* __restrict__ vals,
                                                           const uint32_t* __restrict__ ranges, RasterShape rs,
                                                           const uint32_t* __restrict__ tile_order,
                                                           float4* __restrict__ out, uint32_t* __restrict__ work) {
  constexpr int kN = TW * TH;
  constexpr int kRowsPerWave = kWave / TW;
  constexpr bool kOneWave = kN == kWave;  // the wave covers the tile: every binned record touches it
  static_assert(kN % kStepGroup == 0, "batches hold whole step groups");
  __shared__ float4 s_rec[2][3][kN];

  const uint32_t tile = tile_order ? tile_order[blockIdx.x] : blockIdx.x;
  const uint32_t tx = tile % rs.tiles_x, ty = tile / rs.tiles_x;
  const int t = threadIdx.x;
  const int lx = t % TW, ly = t / TW;
  const int px = (int)tx * TW + lx, py = (int)ty * TH + ly;
  const bool inside = px < (int)rs.width && py < (int)rs.height;
  // pixel bounds of this wave (wave-uniform, held in scalar registers)
  const int col0 = __builtin_amdgcn_readfirstlane((int)tx * TW), col1 = col0 + TW - 1;
  const int row0 = __builtin_amdgcn_readfirstlane((int)ty * TH + (t / kWave) * kRowsPerWave);
  const int row1 = row0 + kRowsPerWave - 1;
  const int bx0 = col0 & ~(kRasterBlock - 1), by0 = __builtin_amdgcn_readfirstlane(((int)ty * TH) & ~(kRasterBlock - 1));
  const float X0 = (float)bx0, Y0 = (float)by0;
  const float fx = (float)(px - bx0) + 0.5f, fy = (float)(py - by0) + 0.5f;
  const uint32_t start = ranges[2 * tile], end = ranges[2 * tile + 1];
  const uint32_t items = end - start;
  const uint32_t batches = (items + kN - 1) / kN;
  const float t_min = rs.t_min;

  // batch b holds list positions end-1-b*kN-t (nearest first)
  float4 p0 = make_float4(0.f, 0.f, 0.f, 0.f), p1 = p0, p2 = p0;
  if ((uint32_t)t < items) {
    const uint32_t k = vals[end - 1 - t];
    p0 = rec[3 * (size_t)k];
    p1 = rec[3 * (size_t)k + 1];
    p2 = rec[3 * (size_t)k + 2];
  }

  // T >= t_min is the pixel's "still compositing" state: T only decreases,
  // and pixels outside the image start at T = 0 (their colour is never
  // stored).
  float T = inside ? 1.f : 0.f, cr = 0.f, cg = 0.f, cb = 0.f, aa = 0.f;
  uint32_t walked = 0;
  for (uint32_t b = 0; b < batches; ++b) {
    const int buf = b & 1;
    stage_record(p0, p1, p2, b * kN + t < items, X0, Y0, &s_rec[buf][0][t], &s_rec[buf][1][t], &s_rec[buf][2][t]);
    if (__syncthreads_count(T >= t_min) == 0) break;
    // gather the next batch while this one is composited
    const uint32_t next = (b + 1) * kN + t;
    if (next < items) {
      const uint32_t k = vals[end - 1 - next];
      p0 = rec[3 * (size_t)k];
      p1 = rec[3 * (size_t)k + 1];
      p2 = rec[3 * (size_t)k + 2];
    }
    const int cnt = __builtin_amdgcn_readfirstlane((int)min((uint32_t)kN, items - b * kN));
    if (__ballot(T >= t_min) == 0) continue;
    const float4* s0 = s_rec[buf][0];
    const float4* s1 = s_rec[buf][1];
    const float4* s2 = s_rec[buf][2];
    if (kOneWave) {
      // step groups of kStepGroup entries (the batch's tail is padded with
      // records that cover nothing), one exit test per group
      for (int e = 0; e < cnt; e += kStepGroup) {
#pragma unroll
        for (int j = 0; j < kStepGroup; ++j) composite(s0[e + j], s1[e + j], s2[e + j], fx, fy, T, cr, cg, cb, aa);
        walked += (uint32_t)min(kStepGroup, cnt - e);
        if (__ballot(T >= t_min) == 0) break;
      }
    } else {
      for (int e = 0; e < cnt; ++e) {
        const uint32_t rx = __builtin_amdgcn_readfirstlane(__float_as_uint(s1[e].w));
        const uint32_t ry = __builtin_amdgcn_readfirstlane(__float_as_uint(s2[e].w));
        const int rx0 = (int)(rx & 0xffffu), rx1 = (int)(rx >> 16);
        const int ry0 = (int)(ry & 0xffffu), ry1 = (int)(ry >> 16);
        if (ry1 < row0 || ry0 > row1 || rx1 < col0 || rx0 > col1) continue;
        composite(s0[e], s1[e], s2[e], fx, fy, T, cr, cg, cb, aa);
        ++walked;
        if (__ballot(T >= t_min) == 0) break;
      }
    }
  }
  if (inside) out[(size_t)py * rs.width + px] = make_float4(cr, cg, cb, aa + T);
  if (work && t % kWave == 0) work[(size_t)blockIdx.x * (kN / kWave) + t / kWave] = walked;
}

// Two 8x8 tiles per wave: lanes 0-31 composite tile order[2i], lanes 32-63
// tile order[2i + 1], each lane two pixels of one row of its tile (columns c
// and c + 4), which share the row's term fma(a.y, y, a.z) of each frame
// coordinate (the arithmetic of composite() for each pixel, bit for bit, at
// 15 VALU instructions per evaluation instead of 16).  A
// record read from LDS then serves two evaluations: an 8x8 tile with one pixel
// per lane reads 40 B of LDS per lane and list entry, which bounds it (the LDS
// returns 128 B per clock to the CU's four SIMDs) before its 16 VALU
// instructions do.  Consecutive tiles of the longest-first schedule have list
// lengths within a bucket of each other, so little of the wave idles.  Both
// halves step through their lists together; a half whose list ended reads
// records that cover nothing.
template <bool kWork>
__global__ __launch_bounds__(kWave) void rasterize_pair_kernel(const float4* __restrict__ rec,
                                                               const uint32_t* __restrict__ vals,
                                                               const uint32_t* __restrict__ ranges, RasterShape rs,
                                                               const uint32_t* __restrict__ tile_order,
                                                               float4* __restrict__ out,
                                                               uint32_t* __restrict__ work) {
  constexpr int kHalf = kWave / 2;   // lanes per tile, list entries per batch
  constexpr int kStride = kHalf + 1; // the second tile's records start one float4 later: other LDS banks
  static_assert(kHalf % kStepGroup == 0, "batches hold whole step groups");
  __shared__ float4 s_rec[2][3][2 * kStride];

  const int t = threadIdx.x, h = t / kHalf, l = t % kHalf;
  const uint32_t tiles = rs.tiles_x * rs.tiles_y;
  const uint32_t slot = 2 * blockIdx.x + h;
  const bool has_tile = slot < tiles;
  const uint32_t tile = has_tile ? (tile_order ? tile_order[slot] : slot) : 0u;
  const uint32_t tx = tile % rs.tiles_x, ty = tile / rs.tiles_x;
  const int px0 = (int)tx * 8 + l % 4, px1 = px0 + 4, py = (int)ty * 8 + l / 4;
  const bool in0 = has_tile && px0 < (int)rs.width && py < (int)rs.height;
  const bool in1 = has_tile && px1 < (int)rs.width && py < (int)rs.height;
  const int bx0 = ((int)tx * 8) & ~(kRasterBlock - 1), by0 = ((int)ty * 8) & ~(kRasterBlock - 1);
  const float X0 = (float)bx0, Y0 = (float)by0;
  const float fx0 = (float)(px0 - bx0) + 0.5f, fx1 = fx0 + 4.f, fy = (float)(py - by0) + 0.5f;
  uint32_t end = 0, items = 0;
  if (has_tile) {
    end = ranges[2 * tile + 1];
    items = end - ranges[2 * tile];
  }
  const uint32_t items_a = __builtin_amdgcn_readlane(items, 0), items_b = __builtin_amdgcn_readlane(items, kHalf);
  const uint32_t longest = max(items_a, items_b);
  const uint32_t batches = (longest + kHalf - 1) / kHalf;
  const float t_min = rs.t_min;

  // batch b holds list positions end-1-b*kHalf-l of each half's tile (nearest first)
  float4 p0 = make_float4(0.f, 0.f, 0.f, 0.f), p1 = p0, p2 = p0;
  if ((uint32_t)l < items) {
    const uint32_t k = vals[end - 1 - l];
    p0 = rec[3 * (size_t)k];
    p1 = rec[3 * (size_t)k + 1];
    p2 = rec[3 * (size_t)k + 2];
  }
  float T0 = in0 ? 1.f : 0.f, r0 = 0.f, g0 = 0.f, b0 = 0.f, a0 = 0.f;
  float T1 = in1 ? 1.f : 0.f, r1 = 0.f, g1 = 0.f, b1 = 0.f, a1 = 0.f;
  uint32_t walked_a = 0, walked_b = 0;
  for (uint32_t b = 0; b < batches; ++b) {
    const int buf = b & 1;
    const int at = h * kStride + l;
    stage_record(p0, p1, p2, b * kHalf + l < items, X0, Y0, &s_rec[buf][0][at], &s_rec[buf][1][at],
                 &s_rec[buf][2][at]);
    if (__syncthreads_count((T0 >= t_min) | (T1 >= t_min)) == 0) break;
    const uint32_t next = (b + 1) * kHalf + l;
    if (next < items) {
      const uint32_t k = vals[end - 1 - next];
      p0 = rec[3 * (size_t)k];
      p1 = rec[3 * (size_t)k + 1];
      p2 = rec[3 * (size_t)k + 2];
    }
    const uint32_t done = b * kHalf;
    const int cnt_a = items_a > done ? (int)min((uint32_t)kHalf, items_a - done) : 0;
    const int cnt_b = items_b > done ? (int)min((uint32_t)kHalf, items_b - done) : 0;
    const int cnt = max(cnt_a, cnt_b);
    if (__ballot((T0 >= t_min) | (T1 >= t_min)) == 0) continue;
    const float4* s0 = s_rec[buf][0] + h * kStride;
    const float4* s1 = s_rec[buf][1] + h * kStride;
    const float4* s2 = s_rec[buf][2] + h * kStride;
    for (int e = 0; e < cnt; e += kStepGroup) {
#pragma unroll
      for (int j = 0; j < kStepGroup; ++j) {
        const float4 ra = s0[e + j], rb = s1[e + j], rc = s2[e + j];
        const float jx = fmaf(ra.y, fy, ra.z), jy = fmaf(rb.y, fy, rb.z);
        composite_at(fmaf(ra.x, fx0, jx), fmaf(rb.x, fx0, jy), ra, rc, T0, r0, g0, b0, a0);
        composite_at(fmaf(ra.x, fx1, jx), fmaf(rb.x, fx1, jy), ra, rc, T1, r1, g1, b1, a1);
      }
      if (kWork) {
        walked_a += (uint32_t)max(0, min(kStepGroup, cnt_a - e));
        walked_b += (uint32_t)max(0, min(kStepGroup, cnt_b - e));
      }
      if (__ballot((T0 >= t_min) | (T1 >= t_min)) == 0) break;
    }
  }
  if (in0) out[(size_t)py * rs.width + px0] = make_float4(r0, g0, b0, a0 + T0);
  if (in1) out[(size_t)py * rs.width + px1] = make_float4(r1, g1, b1, a1 + T1);
  if (kWork && l == 0 && has_tile) work[slot] = h ? walked_b : walked_a;
}

// ------------------------------------------------- triangle-exact raster ---

constexpr int kSnapBits = 8;             // vertices snapped to 1/256 pixel
constexpr float kMaxVertexPixels = 1e6f; // quads with a vertex farther out are not drawn

__device__ __forceinline__ int64_t snap(float v) { return (int64_t)rintf(v * (float)(1 << kSnapBits)); }

// One splat's quad restaged for one tile: for each of its two triangles
// (vertices 0,1,2 and 2,1,3) the three edge functions at the centre of the
// tile's first pixel (exact, in 1/256-pixel fixed point, as doubles: every
// edge value that can decide coverage is an integer below 2^53) and their
// steps per pixel, a tie bit per edge (top or left edge: a centre exactly on
// it is covered), and 1 / area for the barycentric weights.
struct TriStage {
  double e0[2][3], ex[2][3], ey[2][3], inv_area[2];
  uint32_t ties;  // bit 3 t + i: edge i of triangle t takes ties; bit 6 + t: triangle t is front-facing
};

__device__ __forceinline__ void stage_triangles(const float4& p0, const float4& p1, bool valid, int px0, int py0,
                                                float hw, float hh, TriStage* st) {
  st->ties = 0;
  if (!valid) return;
  const float cx = p0.x, cy = p0.y, m00 = p0.z, m10 = p0.w, m01 = p1.x, m11 = p1.y;
  int64_t X[4], Y[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const float qx = (float)(v / 2) * 2.f - 1.f, qy = (float)(v % 2) * 2.f - 1.f;
    const float rx = (m00 * qx + m01 * qy) * 3.f, ry = (m10 * qx + m11 * qy) * 3.f;  // splat.vert:19-23
    const float xf = (cx + rx) * hw + hw, yf = (cy + ry) * hh + hh;                   // viewport
    if (!(fabsf(xf) <= kMaxVertexPixels && fabsf(yf) <= kMaxVertexPixels)) return;
    X[v] = snap(xf);
    Y[v] = snap(yf);
  }
  const int64_t one = (int64_t)1 << kSnapBits, half = one >> 1;
  const int64_t sx = (int64_t)px0 * one + half, sy = (int64_t)py0 * one + half;
  const int tri[2][3] = {{0, 1, 2}, {2, 1, 3}};
  uint32_t ties = 0;
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    const int a = tri[t][0], b = tri[t][1], c = tri[t][2];
    const int64_t cross = (X[b] - X[a]) * (Y[c] - Y[a]) - (Y[b] - Y[a]) * (X[c] - X[a]);
    if (!(-cross > 0)) continue;  // back-facing (culled) or degenerate: Vulkan area -cross / 2
    ties |= 1u << (6 + t);
    st->inv_area[t] = 1.0 / (double)(-cross);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const int va = tri[t][i], vb = tri[t][(i + 1) % 3];
      const int64_t dx = X[vb] - X[va], dy = Y[vb] - Y[va];
      st->e0[t][i] = (double)(-dx * (sy - Y[va]) + dy * (sx - X[va]));
      st->ex[t][i] = (double)(dy * one);
      st->ey[t][i] = (double)(-dx * one);
      if (dy > 0 || (dy == 0 && dx < 0)) ties |= 1u << (3 * t + i);
    }
  }
  st->ties = ties;
}

// Coverage and alpha of pixel (lx, ly) of the tile (offsets from its first
// pixel) by a staged quad: at most one triangle covers a centre (the shared
// edge goes to one of them by the tie rule).  p is interpolated from the
// vertices' 3 q and alpha = opacity * exp(-p.p / 2) (splat.frag:9-11).
__device__ __forceinline__ float triangle_alpha(const TriStage& st, float opacity, double lx, double ly) {
  // varying p = 3 q of the vertices of each triangle, in its vertex order
  const float P[2][3][2] = {{{-3.f, -3.f}, {-3.f, 3.f}, {3.f, -3.f}}, {{3.f, -3.f}, {-3.f, 3.f}, {3.f, 3.f}}};
  float alpha = 0.f;
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    if (!((st.ties >> (6 + t)) & 1u)) continue;
    double e[3];
    bool in = true;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      e[i] = fma(st.ex[t][i], lx, fma(st.ey[t][i], ly, st.e0[t][i]));
      const bool tie = (st.ties >> (3 * t + i)) & 1u;
      in = in && (e[i] > 0.0 || (e[i] == 0.0 && tie));
    }
    if (in) {
      // the weight of a vertex is the edge function of the edge opposite it
      const double l0 = e[1] * st.inv_area[t], l1 = e[2] * st.inv_area[t], l2 = e[0] * st.inv_area[t];
      const double q0 = l0 * P[t][0][0] + l1 * P[t][1][0] + l2 * P[t][2][0];
      const double q1 = l0 * P[t][0][1] + l1 * P[t][1][1] + l2 * P[t][2][1];
      alpha = opacity * gauss_falloff((float)(q0 * q0 + q1 * q1));
    }
  }
  return alpha;
}

// The raster with triangle coverage: one workgroup per tile, one lane per
// pixel, records staged per tile in LDS, front-to-back compositing with the
// same early exit as rasterize_kernel.
template <int TW, int TH>
__global__ __launch_bounds__(TW* TH) void rasterize_tri_kernel(const float4* __restrict__ rec,
                                                               const uint32_t* __restrict__ vals,
                                                               const uint32_t* __restrict__ ranges, RasterShape rs,
                                                               const uint32_t* __restrict__ tile_order,
                                                               float4* __restrict__ out,
                                                               uint32_t* __restrict__ work) {
  constexpr int kN = TW * TH;
  constexpr int kB = 32;  // records per batch
  constexpr int kRowsPerWave = kWave / TW;
  __shared__ TriStage s_tri[kB];
  __shared__ float4 s_col[kB];
  __shared__ uint32_t s_rect[kB][2];

  const uint32_t tile = tile_order ? tile_order[blockIdx.x] : blockIdx.x;
  const uint32_t tx = tile % rs.tiles_x, ty = tile / rs.tiles_x;
  const int t = threadIdx.x;
  const int lx = t % TW, ly = t / TW;
  const int px0 = (int)tx * TW, py0 = (int)ty * TH;
  const int px = px0 + lx, py = py0 + ly;
  const bool inside = px < (int)rs.width && py < (int)rs.height;
  const int col0 = __builtin_amdgcn_readfirstlane(px0), col1 = col0 + TW - 1;
  const int row0 = __builtin_amdgcn_readfirstlane(py0 + (t / kWave) * kRowsPerWave);
  const int row1 = row0 + kRowsPerWave - 1;
  const float hw = 0.5f * (float)rs.width, hh = 0.5f * (float)rs.height;
  const uint32_t start = ranges[2 * tile], end = ranges[2 * tile + 1];
  const uint32_t items = end - start;
  const float t_min = rs.t_min;
  float T = inside ? 1.f : 0.f, cr = 0.f, cg = 0.f, cb = 0.f, aa = 0.f;
  uint32_t walked = 0;
  for (uint32_t b0 = 0; b0 < items; b0 += kB) {
    const int cnt = (int)min((uint32_t)kB, items - b0);
    __syncthreads();
    if (t < cnt) {
      const uint32_t k = vals[end - 1 - (b0 + t)];
      const float4 p0 = rec[3 * (size_t)k], p1 = rec[3 * (size_t)k + 1], p2 = rec[3 * (size_t)k + 2];
      stage_triangles(p0, p1, true, px0, py0, hw, hh, &s_tri[t]);
      s_col[t] = make_float4(p2.x, p2.y, p2.z, p1.z);
      s_rect[t][0] = __float_as_uint(p1.w);
      s_rect[t][1] = __float_as_uint(p2.w);
    }
    if (__syncthreads_count(T >= t_min) == 0) break;
    for (int e = 0; e < cnt; ++e) {
      const uint32_t rx = __builtin_amdgcn_readfirstlane(s_rect[e][0]);
      const uint32_t ry = __builtin_amdgcn_readfirstlane(s_rect[e][1]);
      if ((int)(ry >> 16) < row0 || (int)(ry & 0xffffu) > row1 || (int)(rx >> 16) < col0 ||
          (int)(rx & 0xffffu) > col1)
        continue;
      const float4 c = s_col[e];
      const float alpha = triangle_alpha(s_tri[e], c.w, (double)lx, (double)ly);
      const float ta = T * alpha;
      cr = fmaf(c.x, ta, cr);
      cg = fmaf(c.y, ta, cg);
      cb = fmaf(c.z, ta, cb);
      aa = fmaf(alpha, ta, aa);
      T = T - ta;
      ++walked;
      if (__ballot(T >= t_min) == 0) break;
    }
  }
  if (inside) out[(size_t)py * rs.width + px] = make_float4(cr, cg, cb, aa + T);
  if (work && t % kWave == 0) work[(size_t)blockIdx.x * (kN / kWave) + t / kWave] = walked;
}

// Scalar-broadcast raster: every wave walks its tile's list on its own.  The
// list entry and its 48-byte record have wave-uniform addresses, so they are
// fetched with scalar loads into SGPRs once per wave (no LDS traffic, no
// barrier); the next record is fetched while the current one is composited.
// Waves of one tile share records through the scalar cache.  Each lane
// restages the record itself (the same operations as stage_record).
template <int TW, int TH>
__global__ __launch_bounds__(TW* TH) void rasterize_wave_kernel(const float4* __restrict__ rec,
                                                                const uint32_t* __restrict__ vals,
                                                                const uint32_t* __restrict__ ranges, RasterShape rs,
                                                                float4* __restrict__ out,
                                                                uint32_t* __restrict__ work) {
  constexpr int kN = TW * TH;
  constexpr int kRowsPerWave = kWave / TW;
  const uint32_t tile = blockIdx.x;
  const uint32_t tx = tile % rs.tiles_x, ty = tile / rs.tiles_x;
  const int t = threadIdx.x;
  const int px = (int)tx * TW + t % TW, py = (int)ty * TH + t / TW;
  const bool inside = px < (int)rs.width && py < (int)rs.height;
  const int col0 = __builtin_amdgcn_readfirstlane((int)tx * TW), col1 = col0 + TW - 1;
  const int row0 = __builtin_amdgcn_readfirstlane((int)ty * TH + (t / kWave) * kRowsPerWave);
  const int row1 = row0 + kRowsPerWave - 1;
  const int bx0 = col0 & ~(kRasterBlock - 1), by0 = __builtin_amdgcn_readfirstlane(((int)ty * TH) & ~(kRasterBlock - 1));
  const float X0 = (float)bx0, Y0 = (float)by0;
  const float fx = (float)(px - bx0) + 0.5f, fy = (float)(py - by0) + 0.5f;
  const uint32_t start = ranges[2 * tile], end = ranges[2 * tile + 1];
  const float t_min = rs.t_min;

  float T = inside ? 1.f : 0.f, cr = 0.f, cg = 0.f, cb = 0.f, aa = 0.f;
  uint32_t walked = 0;
  if (end > start && __ballot(T >= t_min) != 0) {
    uint32_t j = end - 1;
    uint32_t k = __builtin_amdgcn_readfirstlane(vals[j]);
    float4 a = rec[3 * (size_t)k], bb = rec[3 * (size_t)k + 1], c = rec[3 * (size_t)k + 2];
    for (;;) {
      const bool more = j > start;
      const uint32_t kn = more ? __builtin_amdgcn_readfirstlane(vals[j - 1]) : k;
      const float4 na = rec[3 * (size_t)kn], nb = rec[3 * (size_t)kn + 1], nc = rec[3 * (size_t)kn + 2];
      const uint32_t rx = __float_as_uint(bb.w), ry = __float_as_uint(c.w);
      const int rx0 = (int)(rx & 0xffffu), rx1 = (int)(rx >> 16);
      const int ry0 = (int)(ry & 0xffffu), ry1 = (int)(ry >> 16);
      if (!(ry1 < row0 || ry0 > row1 || rx1 < col0 || rx0 > col1)) {
        float4 s0, s1, s2;
        stage_record(a, bb, c, true, X0, Y0, &s0, &s1, &s2);
        composite(s0, s1, s2, fx, fy, T, cr, cg, cb, aa);
        ++walked;
        if (__ballot(T >= t_min) == 0) break;
      }
      if (!more) break;
      --j;
      a = na;
      bb = nb;
      c = nc;
    }
  }
  if (inside) out[(size_t)py * rs.width + px] = make_float4(cr, cg, cb, aa + T);
  if (work && t % kWave == 0) work[(size_t)blockIdx.x * (kN / kWave) + t / kWave] = walked;
}

__global__ void gather_u32_kernel(const uint32_t* __restrict__ src, const uint32_t* __restrict__ idx,
                                  const uint32_t* __restrict__ idx2, uint32_t idx_stride,
                                  const uint32_t* __restrict__ d_count, uint32_t max_n, uint32_t* __restrict__ out) {
  const uint32_t n = min(*d_count, max_n);
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    const uint32_t i = idx2 ? idx[(size_t)idx2[k] * idx_stride] : idx[(size_t)k * idx_stride];
    out[k] = src[i];
  }
}

__global__ void depth_key_bits_kernel(const uint32_t* __restrict__ k, const uint32_t* __restrict__ d_count,
                                      uint32_t max_n, uint32_t* __restrict__ out) {
  const uint32_t n = min(*d_count, max_n);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    out[i] = depth_key_bits(k[i]);
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

size_t occupancy_lds_pad(const void* kernel, uint32_t wg_waves, uint32_t waves_per_cu) {
  if (waves_per_cu == 0) return 0;
  // per-device LDS per CU and per-kernel static LDS, looked up once
  struct Known {
    const void* kernel;
    size_t lds;
  };
  static std::mutex mu;
  static Known known[64];
  static int n_known = 0;
  static size_t cu_lds[16] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 16) return 0;
  std::lock_guard<std::mutex> lock(mu);
  if (cu_lds[dev] == 0) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) != hipSuccess || v <= 0)
      return 0;
    cu_lds[dev] = (size_t)v;
  }
  size_t fixed = SIZE_MAX;
  for (int i = 0; i < n_known; ++i)
    if (known[i].kernel == kernel) fixed = known[i].lds;
  if (fixed == SIZE_MAX) {
    hipFuncAttributes a;
    if (hipFuncGetAttributes(&a, kernel) != hipSuccess) return 0;
    fixed = a.sharedSizeBytes;
    if (n_known < 64) known[n_known++] = {kernel, fixed};
  }
  // at least 3 workgroups per CU keeps a workgroup's LDS within the 64 KiB a
  // launch may ask for without opting in to more
  uint32_t wgs = waves_per_cu / (wg_waves ? wg_waves : 1u);
  if (wgs < 3) wgs = 3;
  // rounded down to 512 bytes (an LDS allocation granule): wgs workgroups fit, and no more
  // for wgs up to 18 and the even values up to 24 (160 KiB per CU)
  const size_t per = (cu_lds[dev] / wgs) & ~(size_t)511;
  return per > fixed ? per - fixed : 0;
}

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
  if (cfg->variant > 1) return false;
  const uint32_t pf = cfg->flags & (VKGS_RASTER_PAIRS_ALWAYS | VKGS_RASTER_PAIRS_NEVER);
  if (pf == (VKGS_RASTER_PAIRS_ALWAYS | VKGS_RASTER_PAIRS_NEVER)) return false;
  if (cfg->lds_batch != 0 && cfg->lds_batch != 8 && cfg->lds_batch != 16 && cfg->lds_batch != 32) return false;
  if (cfg->waves_per_cu != 0 && (cfg->waves_per_cu < 4 || cfg->waves_per_cu > 32)) return false;
  rs->variant = cfg->variant;
  rs->triangles = (cfg->flags & VKGS_RASTER_TRIANGLES) ? 1u : 0u;
  rs->pairs = pf == VKGS_RASTER_PAIRS_ALWAYS ? 1u : pf == VKGS_RASTER_PAIRS_NEVER ? 2u : 0u;
  rs->lds_batch = cfg->lds_batch ? cfg->lds_batch : 16u;
  rs->waves_per_cu = cfg->waves_per_cu;
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
                            const RasterShape& rs, float4* rec, TileRect* rect, uint32_t* cnt, hipStream_t s) {
  if (max_n == 0) return hipSuccess;
  bin_count_kernel<<<grid_for(max_n, 256, 8192), 256, 0, s>>>(inst, order, d_count, max_n, rs, rec, rect, cnt);
  return hipGetLastError();
}

hipError_t launch_bin_gather(const uint4* rc, const uint32_t* order, const uint32_t* d_count, uint32_t max_n,
                             TileRect* rect, uint32_t* cnt, hipStream_t s) {
  if (max_n == 0) return hipSuccess;
  bin_gather_kernel<<<grid_for(max_n, 256, 8192), 256, 0, s>>>(rc, order, d_count, max_n, rect, cnt);
  return hipGetLastError();
}

hipError_t launch_bin_emit(const TileRect* rect, const uint32_t* cnt, const uint32_t* offs, const uint32_t* order,
                           const uint32_t* d_count, uint32_t max_n, const RasterShape& rs, uint64_t max_pairs,
                           uint32_t* keys, uint32_t* vals, uint32_t* status, hipStream_t s, uint32_t* clear,
                           uint32_t n_clear) {
  if (max_n == 0) return clear && n_clear ? fill_u32_launch(clear, 0u, n_clear, s) : hipSuccess;
  bin_emit_kernel<<<grid_for(max_n, 256, 8192), 256, 0, s>>>(rect, cnt, offs, order, d_count, max_n, rs.tiles_x,
                                                              max_pairs, keys, vals, status, clear,
                                                              clear ? n_clear : 0u);
  return hipGetLastError();
}

size_t bin_grouped_storage_bytes(uint32_t max_n) {
  const uint32_t nb = ceil_div(max_n ? max_n : 1u, 256u);
  return ((size_t)nb + ceil_div(nb, kBinGroup)) * 4;
}

hipError_t launch_bin_grouped(const uint64_t* val64, const uint32_t* d_count, uint32_t max_n, const RasterShape& rs,
                              const RectBits& rb, uint64_t max_pairs, uint32_t* keys, uint32_t* vals,
                              uint32_t* d_pairs, uint32_t* status, uint32_t* storage, hipStream_t s, uint32_t* clear,
                              uint32_t n_clear) {
  if (max_n == 0) {
    hipError_t e = fill_u32_launch(d_pairs, 0u, 1, s);
    if (e == hipSuccess && clear && n_clear) e = fill_u32_launch(clear, 0u, n_clear, s);
    return e;
  }
  const uint32_t nb = ceil_div(max_n, 256u), groups = ceil_div(nb, kBinGroup);
  uint32_t* boff = storage;
  uint32_t* gsum = storage + nb;
  bin_group_count_kernel<<<groups, kBinGroupThreads, 0, s>>>(val64, d_count, max_n, rb, nb, boff, gsum);
  // several units per emit workgroup from 2M items up (measured: 6M splats
  // gain; 100k-1M lose waves in flight)
  const uint32_t chunks = max_n >= (1u << 21) ? kBinChunks : 1u;
  bin_emit_grouped_kernel<<<ceil_div(nb, chunks), 256, 0, s>>>(val64, d_count, max_n, rs.tiles_x, rb, max_pairs, boff,
                                                               gsum, groups, chunks, keys, vals, d_pairs, status,
                                                               clear, clear ? n_clear : 0u);
  return hipGetLastError();
}

hipError_t launch_ranges(const uint32_t* keys, const uint32_t* d_pairs, uint64_t max_pairs, uint32_t tiles,
                         uint32_t* ranges, hipStream_t s, bool cleared, const uint32_t* view_tiles,
                         uint32_t local_bits) {
  if (!cleared) ranges_fill_kernel<<<grid_for(2ull * tiles, 256, 1u << 20), 256, 0, s>>>(ranges, 2 * tiles);
  if (max_pairs > 0)
    ranges_mark_kernel<<<grid_for(max_pairs, 1024, 2048), 256, 0, s>>>(keys, d_pairs, max_pairs, ranges, view_tiles,
                                                                       local_bits);
  return hipGetLastError();
}

hipError_t launch_tile_order(const uint32_t* ranges, uint32_t tiles, uint32_t* keys, uint32_t* order,
                              const SortPlan& plan, void* sort_storage, hipStream_t s) {
  if (tiles == 0) return hipSuccess;
  // one 8-bit pass: the keys go to the sort's alternate buffers so that the
  // result lands in (keys, order) without a copy
  uint32_t *ak, *av;
  sort_alt_buffers(plan, sort_storage, &ak, &av);
  tile_order_keys_kernel<<<ceil_div(tiles, 256), 256, 0, s>>>(ranges, tiles, ak, av);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return sort_kv_from_alt_launch(plan, nullptr, keys, order, 8u, sort_storage, s);
}

hipError_t launch_rasterize(const float4* rec, const uint32_t* vals, const uint32_t* ranges, const RasterShape& rs,
                            const uint32_t* tile_order, float4* out, hipStream_t s, uint32_t* work) {
  const uint32_t tiles = rs.tiles_x * rs.tiles_y;
  const uint32_t tw = rs.tile_w, th = rs.tile_h;
  const uint32_t cap = rs.waves_per_cu;
#define VKGS_RASTER_CASE(W_, H_)                                                                             \
  if (tw == W_ && th == H_) {                                                                                \
    constexpr uint32_t kWg = W_ * H_ / kWave;                                                                \
    if (rs.variant == 0) {                                                                                   \
      const size_t pad = occupancy_lds_pad((const void*)rasterize_kernel<W_, H_>, kWg, cap);                 \
      rasterize_kernel<W_, H_><<<tiles, W_ * H_, pad, s>>>(rec, vals, ranges, rs, tile_order, out, work);    \
    } else {                                                                                                 \
      const size_t pad = occupancy_lds_pad((const void*)rasterize_wave_kernel<W_, H_>, kWg, cap);            \
      rasterize_wave_kernel<W_, H_><<<tiles, W_ * H_, pad, s>>>(rec, vals, ranges, rs, out, work);           \
    }                                                                                                        \
    return hipGetLastError();                                                                                \
  }
  if (rs.triangles) {
#define VKGS_TRI_CASE(W_, H_)                                                                                \
  if (tw == W_ && th == H_) {                                                                                \
    const size_t pad = occupancy_lds_pad((const void*)rasterize_tri_kernel<W_, H_>, W_ * H_ / kWave, cap);   \
    rasterize_tri_kernel<W_, H_><<<tiles, W_ * H_, pad, s>>>(rec, vals, ranges, rs, tile_order, out, work);  \
    return hipGetLastError();                                                                                \
  }
    VKGS_TRI_CASE(8, 8)
    VKGS_TRI_CASE(16, 16)
    VKGS_TRI_CASE(32, 8)
    VKGS_TRI_CASE(16, 8)
    VKGS_TRI_CASE(8, 16)
    VKGS_TRI_CASE(16, 4)
    VKGS_TRI_CASE(32, 4)
    VKGS_TRI_CASE(64, 4)
#undef VKGS_TRI_CASE
    return hipErrorInvalidValue;
  }
  // two tiles per wave halve the launch's waves: by default only where there are enough tiles to fill the GPU
  if (tw == 8 && th == 8 && rs.variant == 0 && use_pairs(rs.pairs, tiles)) {
    if (work) {  // the work-counting form (timed frames)
      const size_t pad = occupancy_lds_pad((const void*)rasterize_pair_kernel<true>, 1, cap);
      rasterize_pair_kernel<true><<<(tiles + 1) / 2, kWave, pad, s>>>(rec, vals, ranges, rs, tile_order, out, work);
    } else {
      const size_t pad = occupancy_lds_pad((const void*)rasterize_pair_kernel<false>, 1, cap);
      rasterize_pair_kernel<false><<<(tiles + 1) / 2, kWave, pad, s>>>(rec, vals, ranges, rs, tile_order, out,
                                                                       nullptr);
    }
    return hipGetLastError();
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

hipError_t launch_gather_u32(const uint32_t* src, const uint32_t* idx, const uint32_t* idx2, const uint32_t* d_count,
                             uint32_t max_n, uint32_t* out, hipStream_t s, uint32_t idx_stride) {
  if (max_n == 0) return hipSuccess;
  gather_u32_kernel<<<grid_for(max_n, 256, 8192), 256, 0, s>>>(src, idx, idx2, idx_stride, d_count, max_n, out);
  return hipGetLastError();
}

hipError_t launch_depth_key_bits(const uint32_t* k, const uint32_t* d_count, uint32_t max_n, uint32_t* out,
                                 hipStream_t s) {
  depth_key_bits_kernel<<<grid_for(max_n, 256, 8192), 256, 0, s>>>(k, d_count, max_n, out);
  return hipGetLastError();
}

}  // namespace vkgs
