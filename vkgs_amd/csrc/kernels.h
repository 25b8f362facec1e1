// Host-side launchers of the gfx950 kernels (implemented in *.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "common.h"
#include "sort.h"

namespace vkgs {

constexpr uint32_t kCompactBlock = 256;  // splats per workgroup in rank/fused compaction
constexpr uint32_t kOffsetGroup = 16;   // compaction workgroups per group sum of the fused count

// parse_ply.comp
hipError_t launch_parse_ply(const void* d_ply, uint32_t n, float* pos, float* cov3d, float* opacity, uint16_t* sh,
                            hipStream_t s);

// Bytes of the `counts` storage of launch_rank / launch_fused_count: one word
// per 256 splats, the partials of their scan, and one group sum per
// kOffsetGroup words.
size_t compact_storage_bytes(uint32_t n);
// The group sums inside counts for n splats (*groups = their number).
uint32_t* compact_group_sums(uint32_t* counts, uint32_t n, uint32_t* groups);

// rank.comp with ordered compaction; counts is compact_storage_bytes(n) bytes.
hipError_t launch_rank(const FrameConst& fc, const float* pos, uint32_t n, uint32_t* d_count, uint32_t* key,
                       uint32_t* index, uint32_t* counts, hipStream_t s);

// inverse_index.comp (+ the -1 fill)
hipError_t launch_inverse(const uint32_t* d_count, const uint32_t* index, uint32_t n, int32_t* inverse,
                          hipStream_t s);

// projection.comp
hipError_t launch_projection(const FrameConst& fc, const float* pos, const float* cov3d, const float* opacity,
                             const uint16_t* sh, const int32_t* inverse, uint32_t n, const uint32_t* d_count,
                             float4* inst, uint32_t* draw_indirect, hipStream_t s);

// Fused cull + project + compaction (frame constants read from device memory,
// including the screen and tile shape), in two launches: visible counts per
// group of kOffsetGroup workgroups with the offsets inside each group, then
// per visible splat at its slot (the group sums before its group + its
// workgroup's offset in the group + its rank; the write's workgroup 0 stores
// the total in *d_count):
// raster record rec[3 slot..], (tile rect x, tile rect y, tile count, 0) in
// rc[slot], key = depth_rank_key(1 - z), val = slot, ids[slot] = splat id
// (ids optional: launch_compact_ids rebuilds them from the frame's offsets).
hipError_t launch_fused_count(const FrameConst* d_fc, const float* pos, uint32_t n, uint32_t* d_count,
                              uint32_t* counts, hipStream_t s);
// ids[slot] = the splat id of each visible slot of the last fused frame
// (its frame constants d_fc and its scanned counts), rebuilt on request.
hipError_t launch_compact_ids(const FrameConst* d_fc, const float* pos, uint32_t n, const uint32_t* counts,
                              uint32_t* ids, hipStream_t s);
hipError_t launch_fused_write(const FrameConst* d_fc, const float* pos, const float* cov3d, const float* opacity,
                              const uint16_t* sh, uint32_t n, float4* rec, uint4* rc, uint32_t* key, uint32_t* val,
                              uint32_t* ids, const uint32_t* counts, uint32_t* d_count, hipStream_t s);

// Same, with the tile rectangle packed into the sort value instead of rc:
// val64[slot] = slot | packed rect << 32 (frame tiles must fit rect_bits).
hipError_t launch_fused_write_packed(const FrameConst* d_fc, const float* pos, const float* cov3d,
                                     const float* opacity, const uint16_t* sh, uint32_t n, float4* rec,
                                     uint32_t* key, uint64_t* val64, uint32_t* ids, const uint32_t* counts,
                                     uint32_t* d_count, hipStream_t s);

// Copies a FrameConst passed by value into device memory (stream-ordered)
// and zeroes n_clear words at clear (optional).
hipError_t launch_set_frame_const(FrameConst* d_fc, const FrameConst& fc, hipStream_t s, uint32_t* clear = nullptr,
                                  uint32_t n_clear = 0);

// Binning and raster -----------------------------------------------------

struct RasterShape {
  uint32_t width, height, tile_w, tile_h, tiles_x, tiles_y;
  uint32_t tile_shift_x, tile_shift_y;
  float t_min;
  uint32_t variant;    // 0: LDS batches, 1: scalar-load waves
  uint32_t triangles;  // 1: triangle coverage (VKGS_RASTER_TRIANGLES)
  uint32_t pairs;      // 8x8 tiles two per wave: 0 from kPairMinTiles tiles up, 1 always, 2 never
  uint32_t lds_batch;  // list entries per LDS batch of the differentiable 8x8 backward
  uint32_t waves_per_cu;  // resident-wave cap of the raster kernels (0: none)
};

// Tile count from which 8x8 tiles are rasterized two per wave by default:
// below it the one-tile kernels' twice as many waves fill the GPU better.
constexpr uint32_t kPairMinTiles = 8192;
inline bool use_pairs(uint32_t pairs, uint32_t tiles) { return pairs == 1 || (pairs == 0 && tiles >= kPairMinTiles); }

// Dynamic LDS bytes that, added to `kernel`'s static LDS, keep at most about
// waves_per_cu of its waves (wg_waves per workgroup) resident on a CU; 0 when
// waves_per_cu is 0 or the static LDS already allows no more.
size_t occupancy_lds_pad(const void* kernel, uint32_t wg_waves, uint32_t waves_per_cu);

bool make_raster_shape(uint32_t width, uint32_t height, const vkgs_raster_config* cfg, RasterShape* rs);

// Per draw-order position k < *d_count: raster record rec[3k..3k+2] of
// instance order[k] (order == nullptr: k), its tile rectangle and tile count.
hipError_t launch_bin_count(const float4* inst, const uint32_t* order, const uint32_t* d_count, uint32_t max_n,
                            const RasterShape& rs, float4* rec, TileRect* rect, uint32_t* cnt, hipStream_t s);
// Depth-ordered tile rectangles and counts from the per-slot words of the
// fused projection: rect[k], cnt[k] of slot order[k].
hipError_t launch_bin_gather(const uint4* rc, const uint32_t* order, const uint32_t* d_count, uint32_t max_n,
                             TileRect* rect, uint32_t* cnt, hipStream_t s);
// (tile, k) pairs at offs[k]; pairs beyond max_pairs set *status.  With
// `order` the pair value is order[k] instead of k.
// clear (optional): n_clear words zeroed by the same launch (the tile ranges).
hipError_t launch_bin_emit(const TileRect* rect, const uint32_t* cnt, const uint32_t* offs, const uint32_t* order,
                           const uint32_t* d_count, uint32_t max_n, const RasterShape& rs, uint64_t max_pairs,
                           uint32_t* keys, uint32_t* vals, uint32_t* status, hipStream_t s,
                           uint32_t* clear = nullptr, uint32_t n_clear = 0);
// Binning of depth-sorted packed values val64[k] = slot | packed rect << 32
// in two launches and no scan: pair
// offsets from per-workgroup group sums (storage: bin_grouped_storage_bytes(
// max_n) bytes), *d_pairs = the pair total; pairs (tile, slot) at keys / vals
// in depth order, each splat's tiles row-major.
size_t bin_grouped_storage_bytes(uint32_t max_n);
hipError_t launch_bin_grouped(const uint64_t* val64, const uint32_t* d_count, uint32_t max_n, const RasterShape& rs,
                              const RectBits& rb, uint64_t max_pairs, uint32_t* keys, uint32_t* vals,
                              uint32_t* d_pairs, uint32_t* status, uint32_t* storage, hipStream_t s,
                              uint32_t* clear = nullptr, uint32_t n_clear = 0);
// [start, end) per tile from the tile-sorted keys; `cleared`: the ranges were
// already zeroed (by the emit launch).  Keys are tile ids, or with view_tiles
// view << local_bits | tile in the view, for tile view_tiles[view] + tile in
// the view (each key's pairs contiguous, the keys in any order).
hipError_t launch_ranges(const uint32_t* keys, const uint32_t* d_pairs, uint64_t max_pairs, uint32_t tiles,
                         uint32_t* ranges, hipStream_t s, bool cleared = false, const uint32_t* view_tiles = nullptr,
                         uint32_t local_bits = 0);
// Tiles ordered longest list first, by a log-spaced bucket of the list
// length (stable by tile id within a bucket); one 8-bit sort pass.
hipError_t launch_tile_order(const uint32_t* ranges, uint32_t tiles, uint32_t* keys, uint32_t* order,
                             const SortPlan& plan, void* sort_storage, hipStream_t s);
// tile_order (optional): workgroup i rasterizes tile tile_order[i].
// work (optional): work[i * waves + w] = list entries wave w of workgroup i
// composited (waves = tile pixels / 64), a count of executed evaluations.
hipError_t launch_rasterize(const float4* rec, const uint32_t* vals, const uint32_t* ranges, const RasterShape& rs,
                            const uint32_t* tile_order, float4* out, hipStream_t s, uint32_t* work = nullptr);

// Device buffers of one draw (bin, tile sort, ranges, raster).
struct DrawBuffers {
  float4* rec;       // 3 x float4 raster record per draw-order position
  TileRect* rect;    // tile rectangle per position
  uint32_t* cnt;     // tiles per position
  uint32_t* offs;    // exclusive scan of cnt
  ScanPlan scan;
  void* scan_storage;
  uint32_t* keys;    // tile id per pair
  uint32_t* vals;    // draw-order position per pair
  SortPlan sort;
  void* sort_storage;
  uint32_t* ranges;  // [start, end) per tile
  uint32_t* status;  // overflow flag
  bool status_cleared = false;  // the caller zeroed *status for this draw
  uint32_t* pairs;   // pair count
  // longest-first tile schedule (optional: tile_order == nullptr disables it)
  uint32_t* tile_keys;
  uint32_t* tile_order;
  SortPlan tile_sort;
  void* tile_sort_storage;
  uint32_t* work = nullptr;  // optional: list entries composited per raster wave
};

// Bins instances order[k] (k < *d_count), sorts the pairs by tile, marks the
// tile ranges and rasterizes; marks[0..1] (optional) are recorded after the
// binning and right before the raster kernel (after the tile sort, the ranges
// and the tile schedule).
hipError_t run_draw(const float4* inst, const uint32_t* order, const uint32_t* d_count, uint32_t max_inst,
                    const RasterShape& rs, uint64_t max_pairs, const DrawBuffers& b, float4* out, hipStream_t s,
                    hipEvent_t* marks);

// Same after the fused projection: b.rec holds per-slot records, rc the
// per-slot (rect, count) words, order[k] the slot drawn k-th.
hipError_t run_draw_slots(const uint4* rc, const uint32_t* order, const uint32_t* d_count, uint32_t max_inst,
                          const RasterShape& rs, uint64_t max_pairs, const DrawBuffers& b, float4* out, hipStream_t s,
                          hipEvent_t* marks);

// Same after the packed fused projection: val64[k] = slot | packed rect << 32
// of the slot drawn k-th.
hipError_t run_draw_packed(const uint64_t* val64, const uint32_t* d_count, uint32_t max_inst, const RasterShape& rs,
                           const RectBits& rb, uint64_t max_pairs, const DrawBuffers& b, float4* out, hipStream_t s,
                           hipEvent_t* marks);

// out[k] = src[idx[k * idx_stride]] (idx2 == nullptr) or
// src[idx[idx2[k] * idx_stride]] for k < *d_count.
hipError_t launch_gather_u32(const uint32_t* src, const uint32_t* idx, const uint32_t* idx2, const uint32_t* d_count,
                             uint32_t max_n, uint32_t* out, hipStream_t s, uint32_t idx_stride = 1);
// out[i] = the float bits of fl(1 - z) from the integer depth key k[i] =
// fl(1 - z) * 2^24 (depth_key_bits) for i < *d_count.
hipError_t launch_depth_key_bits(const uint32_t* k, const uint32_t* d_count, uint32_t max_n, uint32_t* out,
                                 hipStream_t s);

}  // namespace vkgs
