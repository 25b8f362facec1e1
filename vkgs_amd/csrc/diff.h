// Differentiable, batched splat path: many views of many scenes per launch.
//
// Parameters are the raw 3DGS ones (fp32): means (N,3), log-scales (N,3),
// quaternions (N,4) in PLY order (w,x,y,z) = (rot_0..rot_3), opacity logits
// (N), SH (N,48) in the vkgs channel-major order.  Activation follows
// parse_ply.comp:46-96 (without the fp16 rounding of the SH), projection
// follows projection.comp:80-179 and the raster the reference's blend.
//
// Every (view, splat) pair is one item; items are grouped per view in blocks
// of kCompactBlock splats (a block belongs to one view) and waves of 64.
// Visible items get slots in (view, splat id) order; the raster record of a
// slot is written by the projection, which runs one lane per scene splat and
// visits the views of its scene.  Tiles of all views share one global id space (view tiles are
// contiguous), so sorting (tile, slot) pairs by global tile id puts the view id
// in the high key bits.  The projection backward runs per scene splat and
// visits the views of its scene, finding each view's slot from the forward's
// per-wave visibility masks, so every gradient is written once, without
// atomics.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.h"
#include "kernels.h"

namespace vkgs {

constexpr uint32_t kMaskWords = kCompactBlock / kWave;  // visibility ballots per compaction block

struct DiffScene {
  const float* means;
  const float* log_scales;
  const float* quats;
  const float* opacity;
  const float* sh;
  float* g_means;
  float* g_log_scales;
  float* g_quats;
  float* g_opacity;
  float* g_sh;
  const uint32_t* perm;  // processing order: local slot i handles splat perm[i] (nullptr: i)
  uint32_t n;
  uint32_t block_base;  // first projection-backward block of this scene
  uint32_t view_begin;  // views of this scene: scene_views[view_begin .. view_end)
  uint32_t view_end;
};

struct DiffView {
  FrameConst fc;            // camera, model, screen and tile shape
  float vm[16];             // V * M (column-major): model space to view space in one product
  uint32_t scene;
  uint32_t block_base;      // first compaction block of this view
  uint32_t tile_base;       // first global tile of this view
  uint32_t pad0;
  float4* out;              // height x width RGBA
  const float4* grad_out;   // dL/dRGBA (backward only)
  float* final_t;           // per-pixel transmittance after the last processed splat
  uint32_t* last;           // per-pixel number of list entries processed, nearest first
};

struct DiffLaunch {
  const DiffView* views;
  const DiffScene* scenes;
  const uint32_t* view_blocks;  // n_views + 1 ascending compaction-block bases
  const uint32_t* view_tiles;   // n_views + 1 ascending global-tile bases
  const uint32_t* scene_blocks; // n_scenes + 1 ascending projection-backward block bases
  const uint32_t* scene_views;  // view indices grouped by scene
  uint32_t n_views, n_scenes, n_blocks, n_tiles, n_scene_blocks;
  uint32_t tile_w, tile_h, tile_shift_x, tile_shift_y;
  float t_min;
  // launch choices (host side): 8x8 tiles two per wave (RasterShape::pairs),
  // backward LDS batch depth on 8x8 tiles, resident-wave cap of the rasters
  uint32_t pairs, lds_batch, waves_per_cu;
  // view-major local tile keys (0: off).  On: the depth key is view << 25 |
  // depth_rank_key(1 - z), so pairs are emitted view by view in depth order,
  // and a pair's tile key is view << local_tile_bits | the tile in its view;
  // the tile sort then orders the low local_tile_bits bits alone (stable, so
  // each (view, tile) run stays contiguous and in depth order), one radix
  // pass fewer than a sort on the global tile id when views are many.
  uint32_t local_tile_bits;
  // optional work counter of a raster launch: work[i * waves + w] = list
  // entries wave w of workgroup i stepped through (two-tile kernels: one
  // word per tile, its own entries), so 64 x the sum counts the (pixel,
  // entry) evaluations executed
  uint32_t* work;
};

// Spatial processing order of every scene's splats: keys[g] = scene index in
// the high bits | Morton code of the position quantized in the scene's
// bounding box (10 bits per axis at most), vals[g] = local splat index, for g
// = scene block * kCompactBlock + lane (padding lanes get key 0xffffffff).
// Sorted, each scene's indices are contiguous from the sum of the earlier
// scenes' sizes.  bbox_part holds 6 words per scene block, bbox 6 per scene.
hipError_t launch_diff_morton_keys(const DiffLaunch& L, uint32_t* bbox_part, uint32_t* bbox, uint32_t* keys,
                                   uint32_t* vals, hipStream_t s);

// Visibility ballots and per-wave visible counts of every (view, splat) item
// (masks and counts hold kMaskWords words per compaction block), then the
// counts scanned in place into slot offsets; *d_visible = total.  scan_storage
// is diff_count_scan_bytes(n_blocks) bytes.
hipError_t launch_diff_count(const DiffLaunch& L, uint32_t* counts, uint64_t* masks, uint32_t* d_visible,
                             void* scan_storage, hipStream_t s);
size_t diff_count_scan_bytes(uint32_t n_blocks);
// Per visible item at its slot offsets[w] + (visible lanes of wave w before
// it): raster record rec[3 slot..], (tile rect x, tile rect y, tile count,
// view) in rc[slot], the tile count again in scnt[slot], key = bits(1 - z)
// (with local tile keys: view << 25 | depth_rank_key(1 - z)), val = slot.
hipError_t launch_diff_write(const DiffLaunch& L, const uint32_t* offsets, const uint64_t* masks, float4* rec,
                             uint4* rc, uint32_t* scnt, uint32_t* key, uint32_t* val, hipStream_t s);
// Same with the tile rectangle packed into the depth sort's value:
// val64[slot] = slot | packed rect << 32 (rb from the largest view's tile
// grid), no rc or val.
hipError_t launch_diff_write_packed(const DiffLaunch& L, const uint32_t* offsets, const uint64_t* masks,
                                    float4* rec, uint32_t* scnt, uint32_t* key, uint64_t* val64, const RectBits& rb,
                                    hipStream_t s);
// vslot[v] = first slot of view v (n_views + 1 words, the last = *d_count).
hipError_t launch_diff_view_slots(const DiffLaunch& L, const uint32_t* offsets, const uint32_t* d_count,
                                  uint32_t* vslot, hipStream_t s);
// Pairs as launch_diff_bin_emit from the depth-sorted packed values, offs =
// exclusive scan of their tile counts.
hipError_t launch_diff_bin_emit_packed(const DiffLaunch& L, const uint64_t* val64, const uint32_t* offs,
                                       const uint32_t* vslot, const RectBits& rb, const uint32_t* d_count,
                                       uint32_t max_n, uint64_t max_pairs, uint32_t* keys, uint32_t* vals,
                                       uint32_t* pslot, const uint32_t* soffs, const uint32_t* scnt, uint32_t* status,
                                       hipStream_t s);
// rect[k], cnt[k], kview[k] of slot order[k].
hipError_t launch_diff_bin_gather(const uint4* rc, const uint32_t* order, const uint32_t* d_count, uint32_t max_n,
                                  TileRect* rect, uint32_t* cnt, uint32_t* kview, hipStream_t s);
// (global tile, q) pairs at offs[k] for depth position k, with q numbering
// the pairs in slot order from soffs[order[k]], and pslot[q] = slot for the
// slot's scnt[slot] pairs; pairs beyond max_pairs set *status.
hipError_t launch_diff_bin_emit(const DiffLaunch& L, const TileRect* rect, const uint32_t* cnt, const uint32_t* offs,
                                const uint32_t* kview, const uint32_t* order, const uint32_t* d_count,
                                uint32_t max_n, uint64_t max_pairs, uint32_t* keys, uint32_t* vals, uint32_t* pslot,
                                const uint32_t* soffs, const uint32_t* scnt, uint32_t* status, hipStream_t s);
// Raster of the tile-sorted pairs (vals = pair indices, pslot maps them to
// slots); tile_order (optional): workgroup i rasterizes global tile
// tile_order[i].
hipError_t launch_diff_raster(const DiffLaunch& L, const float4* rec, const uint32_t* vals, const uint32_t* ranges,
                              const uint32_t* pslot, const uint32_t* tile_order, bool keep_state, hipStream_t s);
// Blend backward: per sorted pair 9 partial gradient words written at the
// pair's q (its slot-order index, the sorted value) in gpart (9 floats per
// pair; every live pair written once, zeros for entries no pixel reached).
// No global atomics.
hipError_t launch_diff_raster_backward(const DiffLaunch& L, const float4* rec, const uint32_t* vals,
                                       const uint32_t* ranges, const uint32_t* pslot, const uint32_t* tile_order,
                                       float* gpart, hipStream_t s);
// Projection backward: a slot's 9 gradient words are the sum of its pairs'
// partials soffs[slot] .. soffs[slot] + scnt[slot] - 1 (clipped at
// max_pairs), chained through each view's projection.  Writes (overwrites)
// every parameter gradient of every scene.
hipError_t launch_diff_project_backward(const DiffLaunch& L, const uint32_t* offsets, const uint64_t* masks,
                                        const uint32_t* soffs, const uint32_t* scnt, const float* gpart,
                                        uint64_t max_pairs, hipStream_t s);

}  // namespace vkgs
