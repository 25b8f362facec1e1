// Stage entry points of the C ABI (one per reference stage, engine.cc:1137-1290).
#include <string.h>

#include "common.h"
#include "kernels.h"
#include "sort.h"

namespace vkgs {
namespace {

hipStream_t as_stream(void* s) { return static_cast<hipStream_t>(s); }

// Layout of the caller-provided vkgs_draw storage.
struct DrawPlan {
  uint32_t max_inst = 0;
  uint64_t max_pairs = 0;
  RasterShape rs{};
  SortPlan sort{};
  ScanPlan scan{};
  SortPlan tile_sort{};
  size_t off_hdr = 0, off_rec = 0, off_rect = 0, off_cnt = 0, off_offs = 0, off_scan = 0, off_keys = 0,
         off_vals = 0, off_sort = 0, off_ranges = 0, off_tile_keys = 0, off_tile_order = 0, off_tile_sort = 0,
         bytes = 0;
};

// Header words at the start of the draw storage.
enum : uint32_t {
  kHdrStatus = 0,
  kHdrPairs = 1,
  kHdrTilesX = 2,
  kHdrTilesY = 3,
  kHdrMagic = 4,
  kHdrOffVals = 5,    // byte offset of the tile-sorted instance list (2 words)
  kHdrOffRanges = 7,  // byte offset of the per-tile ranges (2 words)
  kHdrWords = 64
};
constexpr uint32_t kDrawMagic = 0x564b4753u;  // "VKGS"

bool draw_plan(uint32_t max_inst, uint32_t width, uint32_t height, uint64_t max_pairs, const vkgs_raster_config* cfg,
               DrawPlan* p) {
  if (!make_raster_shape(width, height, cfg, &p->rs)) return false;
  if (max_pairs > 0xffffffffull) max_pairs = 0xffffffffull;
  p->max_inst = max_inst;
  p->max_pairs = max_pairs;
  p->sort = sort_plan((uint32_t)max_pairs);
  p->scan = scan_plan(max_inst);
  size_t off = 0;
  p->off_hdr = off;
  off = align_up(off + kHdrWords * 4, 256);
  p->off_rec = off;
  off = align_up(off + (size_t)max_inst * 48, 256);
  p->off_rect = off;
  off = align_up(off + (size_t)max_inst * sizeof(TileRect), 256);
  p->off_cnt = off;
  off = align_up(off + (size_t)max_inst * 4, 256);
  p->off_offs = off;
  off = align_up(off + (size_t)max_inst * 4, 256);
  p->off_scan = off;
  off = align_up(off + p->scan.bytes, 256);
  p->off_keys = off;
  off = align_up(off + (size_t)max_pairs * 4, 256);
  p->off_vals = off;
  off = align_up(off + (size_t)max_pairs * 4, 256);
  p->off_sort = off;
  off = align_up(off + p->sort.bytes, 256);
  p->off_ranges = off;
  const uint32_t tiles = p->rs.tiles_x * p->rs.tiles_y;
  off = align_up(off + (size_t)tiles * 8, 256);
  p->tile_sort = sort_plan(tiles);
  p->off_tile_keys = off;
  off = align_up(off + (size_t)tiles * 4, 256);
  p->off_tile_order = off;
  off = align_up(off + (size_t)tiles * 4, 256);
  p->off_tile_sort = off;
  off = align_up(off + p->tile_sort.bytes, 256);
  p->bytes = off;
  return true;
}

int bits_for(uint32_t v) {  // bits needed for values in [0, v)
  int b = 0;
  while (b < 32 && (1ull << b) < v) ++b;
  return b;
}

}  // namespace

// Where binning writes its (tile, value) pairs: the tile sort's alternate
// buffers when the sort takes an odd number of passes, so that the sorted
// pairs land in b.keys / b.vals without a final copy; else b.keys / b.vals.
static void pair_targets(const DrawBuffers& b, uint32_t tile_bits, uint32_t** keys, uint32_t** vals) {
  *keys = b.keys;
  *vals = b.vals;
  if (sort_passes(tile_bits) % 2 == 1) sort_alt_buffers(b.sort, b.sort_storage, keys, vals);
}

static hipError_t sort_pairs(const DrawBuffers& b, uint32_t tile_bits, hipStream_t s) {
  if (sort_passes(tile_bits) % 2 == 1)
    return sort_kv_from_alt_launch(b.sort, b.pairs, b.keys, b.vals, tile_bits, b.sort_storage, s);
  return sort_kv_launch(b.sort, b.pairs, b.keys, b.vals, tile_bits, b.sort_storage, s);
}

// Raster with the longest-list-first tile schedule when its buffers are given.
// Longest-list-first tile order (when b.tile_order is set), then the raster;
// `mark` (optional) is recorded between the two.
static hipError_t raster_scheduled(const DrawBuffers& b, const RasterShape& rs, uint32_t tiles, float4* out,
                                   hipStream_t s, hipEvent_t mark) {
  if (b.tile_order) {
    hipError_t e = launch_tile_order(b.ranges, tiles, b.tile_keys, b.tile_order, b.tile_sort, b.tile_sort_storage, s);
    if (e != hipSuccess) return e;
  }
  if (mark) (void)hipEventRecord(mark, s);
  return launch_rasterize(b.rec, b.vals, b.ranges, rs, b.tile_order, out, s, b.work);
}

// Shared by vkgs_draw and the renderer: bin, sort by tile, ranges, raster.
hipError_t run_draw(const float4* inst, const uint32_t* order, const uint32_t* d_count, uint32_t max_inst,
                    const RasterShape& rs, uint64_t max_pairs, const DrawBuffers& b, float4* out, hipStream_t s,
                    hipEvent_t* marks) {
  hipError_t e;
  if (!b.status_cleared && (e = fill_u32_launch(b.status, 0u, 1, s)) != hipSuccess) return e;
  if ((e = launch_bin_count(inst, order, d_count, max_inst, rs, b.rec, b.rect, b.cnt, s)) != hipSuccess) return e;
  if ((e = scan_exclusive_launch(b.scan, b.cnt, b.offs, d_count, b.pairs, b.scan_storage, s)) != hipSuccess) return e;
  const uint32_t tiles = rs.tiles_x * rs.tiles_y;
  const uint32_t tile_bits = (uint32_t)bits_for(tiles);
  uint32_t *pk, *pv;
  pair_targets(b, tile_bits, &pk, &pv);
  if ((e = launch_bin_emit(b.rect, b.cnt, b.offs, nullptr, d_count, max_inst, rs, max_pairs, pk, pv, b.status, s,
                           b.ranges, 2 * tiles)) != hipSuccess)
    return e;
  if (marks) (void)hipEventRecord(marks[0], s);
  if ((e = sort_pairs(b, tile_bits, s)) != hipSuccess) return e;
  if ((e = launch_ranges(b.keys, b.pairs, max_pairs, tiles, b.ranges, s, true)) != hipSuccess) return e;
  return raster_scheduled(b, rs, tiles, out, s, marks ? marks[1] : nullptr);
}

// Fused-path draw: raster records and (rect, count) words were written per
// slot by the projection; order[k] is the slot drawn k-th.  Pair values and
// the raster's record indices are slots.
hipError_t run_draw_slots(const uint4* rc, const uint32_t* order, const uint32_t* d_count, uint32_t max_inst,
                          const RasterShape& rs, uint64_t max_pairs, const DrawBuffers& b, float4* out, hipStream_t s,
                          hipEvent_t* marks) {
  hipError_t e;
  if (!b.status_cleared && (e = fill_u32_launch(b.status, 0u, 1, s)) != hipSuccess) return e;
  if ((e = launch_bin_gather(rc, order, d_count, max_inst, b.rect, b.cnt, s)) != hipSuccess) return e;
  if ((e = scan_exclusive_launch(b.scan, b.cnt, b.offs, d_count, b.pairs, b.scan_storage, s)) != hipSuccess) return e;
  const uint32_t tiles = rs.tiles_x * rs.tiles_y;
  const uint32_t tile_bits = (uint32_t)bits_for(tiles);
  uint32_t *pk, *pv;
  pair_targets(b, tile_bits, &pk, &pv);
  if ((e = launch_bin_emit(b.rect, b.cnt, b.offs, order, d_count, max_inst, rs, max_pairs, pk, pv, b.status, s,
                           b.ranges, 2 * tiles)) != hipSuccess)
    return e;
  if (marks) (void)hipEventRecord(marks[0], s);
  if ((e = sort_pairs(b, tile_bits, s)) != hipSuccess) return e;
  if ((e = launch_ranges(b.keys, b.pairs, max_pairs, tiles, b.ranges, s, true)) != hipSuccess) return e;
  return raster_scheduled(b, rs, tiles, out, s, marks ? marks[1] : nullptr);
}

hipError_t run_draw_packed(const uint64_t* val64, const uint32_t* d_count, uint32_t max_inst, const RasterShape& rs,
                           const RectBits& rb, uint64_t max_pairs, const DrawBuffers& b, float4* out, hipStream_t s,
                           hipEvent_t* marks) {
  hipError_t e;
  if (!b.status_cleared && (e = fill_u32_launch(b.status, 0u, 1, s)) != hipSuccess) return e;
  const uint32_t tiles = rs.tiles_x * rs.tiles_y;
  const uint32_t tile_bits = (uint32_t)bits_for(tiles);
  uint32_t *pk, *pv;
  pair_targets(b, tile_bits, &pk, &pv);
  // offs (max_inst words) holds the emit workgroups' group offsets and sums
  if ((e = launch_bin_grouped(val64, d_count, max_inst, rs, rb, max_pairs, pk, pv, b.pairs, b.status, b.offs, s,
                              b.ranges, 2 * tiles)) != hipSuccess)
    return e;
  if (marks) (void)hipEventRecord(marks[0], s);
  if ((e = sort_pairs(b, tile_bits, s)) != hipSuccess) return e;
  if ((e = launch_ranges(b.keys, b.pairs, max_pairs, tiles, b.ranges, s, true)) != hipSuccess) return e;
  return raster_scheduled(b, rs, tiles, out, s, marks ? marks[1] : nullptr);
}

}  // namespace vkgs

using namespace vkgs;

extern "C" {

int vkgs_parse_ply(const void* d_ply, uint32_t n, float* d_position, float* d_cov3d, float* d_opacity,
                   uint16_t* d_sh, void* stream) {
  if (n == 0) return VKGS_OK;
  if (!d_ply || !d_position || !d_cov3d || !d_opacity || !d_sh) return VKGS_ERR_INVALID_ARG;
  VKGS_HIP_CHECK(launch_parse_ply(d_ply, n, d_position, d_cov3d, d_opacity, d_sh, as_stream(stream)));
  return VKGS_OK;
}

size_t vkgs_rank_storage_bytes(uint32_t max_n) { return compact_storage_bytes(max_n); }

int vkgs_rank(const vkgs_camera* camera, const float model[16], const float* d_position, uint32_t n,
              uint32_t* d_visible_count, uint32_t* d_key, uint32_t* d_index, void* d_storage, size_t storage_bytes,
              void* stream) {
  if (!camera || !d_visible_count) return VKGS_ERR_INVALID_ARG;
  if (n > 0 && (!d_position || !d_key || !d_index || !d_storage)) return VKGS_ERR_INVALID_ARG;
  if (n > (1u << 28)) return VKGS_ERR_CAPACITY;
  if (storage_bytes < vkgs_rank_storage_bytes(n)) return VKGS_ERR_CAPACITY;
  FrameConst fc;
  make_frame_const(camera, model, &fc);
  VKGS_HIP_CHECK(launch_rank(fc, d_position, n, d_visible_count, d_key, d_index, static_cast<uint32_t*>(d_storage),
                             as_stream(stream)));
  return VKGS_OK;
}

size_t vkgs_sort_kv_storage_bytes(uint32_t max_n) { return sort_plan(max_n).bytes; }

int vkgs_sort_kv(uint32_t max_n, const uint32_t* d_count, uint32_t* d_key, uint32_t* d_value, uint32_t key_bits,
                 void* d_storage, size_t storage_bytes, void* stream) {
  if (max_n == 0) return VKGS_OK;
  if (!d_count || !d_key || !d_value || !d_storage || key_bits == 0 || key_bits > 32) return VKGS_ERR_INVALID_ARG;
  const SortPlan p = sort_plan(max_n);
  if (storage_bytes < p.bytes) return VKGS_ERR_CAPACITY;
  VKGS_HIP_CHECK(sort_kv_launch(p, d_count, d_key, d_value, key_bits, d_storage, as_stream(stream)));
  return VKGS_OK;
}

int vkgs_inverse_index(const uint32_t* d_count, const uint32_t* d_index, uint32_t n, int32_t* d_inverse,
                       void* stream) {
  if (n == 0) return VKGS_OK;
  if (!d_count || !d_index || !d_inverse) return VKGS_ERR_INVALID_ARG;
  VKGS_HIP_CHECK(launch_inverse(d_count, d_index, n, d_inverse, as_stream(stream)));
  return VKGS_OK;
}

int vkgs_projection(const vkgs_camera* camera, const float model[16], const float* d_position, const float* d_cov3d,
                    const float* d_opacity, const uint16_t* d_sh, const int32_t* d_inverse, uint32_t n,
                    const uint32_t* d_visible_count, float* d_instances, uint32_t* d_draw_indirect, void* stream) {
  if (n == 0) return VKGS_OK;
  if (!camera || !d_position || !d_cov3d || !d_opacity || !d_sh || !d_inverse || !d_visible_count ||
      !d_instances || !d_draw_indirect)
    return VKGS_ERR_INVALID_ARG;
  FrameConst fc;
  make_frame_const(camera, model, &fc);
  VKGS_HIP_CHECK(launch_projection(fc, d_position, d_cov3d, d_opacity, d_sh, d_inverse, n, d_visible_count,
                                   reinterpret_cast<float4*>(d_instances), d_draw_indirect, as_stream(stream)));
  return VKGS_OK;
}

size_t vkgs_draw_storage_bytes(uint32_t max_instances, uint32_t width, uint32_t height, uint64_t max_pairs,
                               const vkgs_raster_config* cfg) {
  DrawPlan p;
  if (!draw_plan(max_instances, width, height, max_pairs, cfg, &p)) return 0;
  return p.bytes;
}

int vkgs_draw(const float* d_instances, const uint32_t* d_draw_indirect, uint32_t max_instances, uint32_t width,
              uint32_t height, uint64_t max_pairs, const vkgs_raster_config* cfg, float* d_rgba, void* d_storage,
              size_t storage_bytes, void* stream) {
  if (!d_draw_indirect || !d_rgba || !d_storage) return VKGS_ERR_INVALID_ARG;
  if (max_instances > 0 && !d_instances) return VKGS_ERR_INVALID_ARG;
  DrawPlan p;
  if (!draw_plan(max_instances, width, height, max_pairs, cfg, &p)) return VKGS_ERR_INVALID_ARG;
  if (storage_bytes < p.bytes) return VKGS_ERR_CAPACITY;
  char* b = static_cast<char*>(d_storage);
  uint32_t* hdr = reinterpret_cast<uint32_t*>(b + p.off_hdr);
  Words16 hw = {};
  hw.w[kHdrTilesX] = p.rs.tiles_x;
  hw.w[kHdrTilesY] = p.rs.tiles_y;
  hw.w[kHdrMagic] = kDrawMagic;
  hw.w[kHdrOffVals] = (uint32_t)p.off_vals;
  hw.w[kHdrOffVals + 1] = (uint32_t)((uint64_t)p.off_vals >> 32);
  hw.w[kHdrOffRanges] = (uint32_t)p.off_ranges;
  hw.w[kHdrOffRanges + 1] = (uint32_t)((uint64_t)p.off_ranges >> 32);
  hipStream_t s = as_stream(stream);
  VKGS_HIP_CHECK(set_words_launch(hdr, hw, 9, s));
  DrawBuffers db;
  db.rec = reinterpret_cast<float4*>(b + p.off_rec);
  db.rect = reinterpret_cast<TileRect*>(b + p.off_rect);
  db.cnt = reinterpret_cast<uint32_t*>(b + p.off_cnt);
  db.offs = reinterpret_cast<uint32_t*>(b + p.off_offs);
  db.scan = p.scan;
  db.scan_storage = b + p.off_scan;
  db.keys = reinterpret_cast<uint32_t*>(b + p.off_keys);
  db.vals = reinterpret_cast<uint32_t*>(b + p.off_vals);
  db.sort = p.sort;
  db.sort_storage = b + p.off_sort;
  db.ranges = reinterpret_cast<uint32_t*>(b + p.off_ranges);
  db.status = hdr + kHdrStatus;
  db.status_cleared = true;  // the header words above include status = 0
  db.pairs = hdr + kHdrPairs;
  db.tile_keys = reinterpret_cast<uint32_t*>(b + p.off_tile_keys);
  db.tile_order = reinterpret_cast<uint32_t*>(b + p.off_tile_order);
  db.tile_sort = p.tile_sort;
  db.tile_sort_storage = b + p.off_tile_sort;
  VKGS_HIP_CHECK(run_draw(reinterpret_cast<const float4*>(d_instances), nullptr, d_draw_indirect + 8, max_instances,
                          p.rs, p.max_pairs, db, reinterpret_cast<float4*>(d_rgba), s, nullptr));
  return VKGS_OK;
}

int vkgs_draw_status(const void* d_storage, size_t storage_bytes, uint64_t* pairs, void* stream) {
  if (!d_storage || storage_bytes < kHdrWords * 4) return VKGS_ERR_INVALID_ARG;
  uint32_t hdr[kHdrWords];
  VKGS_HIP_CHECK(hipStreamSynchronize(as_stream(stream)));
  VKGS_HIP_CHECK(hipMemcpy(hdr, d_storage, sizeof(hdr), hipMemcpyDeviceToHost));
  if (hdr[kHdrMagic] != kDrawMagic) return VKGS_ERR_INVALID_ARG;
  if (pairs) *pairs = hdr[kHdrPairs];
  return hdr[kHdrStatus] ? VKGS_ERR_OVERFLOW : VKGS_OK;
}

int vkgs_draw_lists(const void* d_storage, size_t storage_bytes, const uint32_t** d_ranges, const uint32_t** d_list,
                    uint32_t* tiles_x, uint32_t* tiles_y) {
  if (!d_storage || !d_ranges || !d_list || !tiles_x || !tiles_y || storage_bytes < kHdrWords * 4)
    return VKGS_ERR_INVALID_ARG;
  uint32_t hdr[kHdrWords];
  VKGS_HIP_CHECK(hipMemcpy(hdr, d_storage, sizeof(hdr), hipMemcpyDeviceToHost));
  if (hdr[kHdrMagic] != kDrawMagic) return VKGS_ERR_INVALID_ARG;
  const uint64_t off_vals = hdr[kHdrOffVals] | ((uint64_t)hdr[kHdrOffVals + 1] << 32);
  const uint64_t off_ranges = hdr[kHdrOffRanges] | ((uint64_t)hdr[kHdrOffRanges + 1] << 32);
  const uint64_t tiles = (uint64_t)hdr[kHdrTilesX] * hdr[kHdrTilesY];
  if (off_ranges + tiles * 8 > storage_bytes || off_vals > storage_bytes) return VKGS_ERR_INVALID_ARG;
  const char* b = static_cast<const char*>(d_storage);
  *d_ranges = reinterpret_cast<const uint32_t*>(b + off_ranges);
  *d_list = reinterpret_cast<const uint32_t*>(b + off_vals);
  *tiles_x = hdr[kHdrTilesX];
  *tiles_y = hdr[kHdrTilesY];
  return VKGS_OK;
}

}  // extern "C"
