// Differentiable, batched splat path: many views of many scenes per launch.
//
// Parameters are the raw 3DGS ones (fp32): means (N,3), log-scales (N,3),
// quaternions (N,4) in PLY order (w,x,y,z) = (rot_0..rot_3), opacity logits
// (N), SH (N,48) in the vkgs channel-major order.  Activation follows
// parse_ply.comp:106-156 (without the fp16 rounding of the SH), projection
// follows projection.comp:80-179 and the raster the reference's blend.
//
// Every (view, splat) pair is one item; items are grouped per view in blocks
// of kCompactBlock so a block belongs to one view.  Tiles of all views share
// one global id space (view tiles are contiguous), so sorting (tile, k) pairs
// by global tile id puts the view id in the high key bits.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.h"
#include "kernels.h"

namespace vkgs {

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
  uint32_t n;
  uint32_t pad[3];
};

struct DiffView {
  FrameConst fc;            // camera, model, screen and tile shape
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
  const uint32_t* block_view;  // compaction block -> view
  const uint32_t* tile_view;   // global tile -> view
  uint32_t n_views, n_blocks, n_tiles, tile_w, tile_h, tile_shift_x, tile_shift_y;
  float t_min;
};

// Per-frame device buffers of the batched path.
struct DiffBuffers {
  uint32_t* counts;     // compact_storage_bytes(n_blocks * kCompactBlock)
  float4* inst;         // 3 per slot
  uint32_t* key;        // depth key per slot, sorted in place
  uint32_t* val;        // slot, sorted with the keys -> draw order
  uint32_t* ids;        // splat id per slot
  uint32_t* vid;        // view per slot
  uint32_t* kview;      // view per draw-order position
  float* g9;            // 9 gradient words per draw-order position
  uint32_t* misc;       // [0] visible, [1] pairs, [2] overflow
};

hipError_t launch_diff_count(const DiffLaunch& L, uint32_t* counts, uint32_t* d_visible, hipStream_t s);
hipError_t launch_diff_write(const DiffLaunch& L, const uint32_t* counts, float4* inst, uint32_t* key, uint32_t* val,
                             uint32_t* ids, uint32_t* vid, hipStream_t s);
hipError_t launch_diff_bin_count(const DiffLaunch& L, const float4* inst, const uint32_t* order, const uint32_t* vid,
                                 const uint32_t* d_count, uint32_t max_n, float4* rec, TileRect* rect, uint32_t* cnt,
                                 uint32_t* kview, hipStream_t s);
hipError_t launch_diff_bin_emit(const DiffLaunch& L, const TileRect* rect, const uint32_t* cnt, const uint32_t* offs,
                                const uint32_t* kview, const uint32_t* d_count, uint32_t max_n, uint64_t max_pairs,
                                uint32_t* keys, uint32_t* vals, uint32_t* status, hipStream_t s);
hipError_t launch_diff_raster(const DiffLaunch& L, const float4* rec, const uint32_t* vals, const uint32_t* ranges,
                              bool keep_state, hipStream_t s);
hipError_t launch_diff_raster_backward(const DiffLaunch& L, const float4* rec, const uint32_t* vals,
                                       const uint32_t* ranges, float* g9, hipStream_t s);
hipError_t launch_diff_project_backward(const DiffLaunch& L, const uint32_t* order, const uint32_t* ids,
                                        const uint32_t* vid, const float* g9, const uint32_t* d_count, uint32_t max_n,
                                        hipStream_t s);

}  // namespace vkgs
