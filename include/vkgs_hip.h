/*
 * vkgs_hip.h -- C ABI of the MI355X (gfx950) splat pipeline.
 *
 * The reference (jaesung-cs/vkgs) keeps its splat path private to
 * Engine::Impl (src/vkgs/engine/engine.cc:101-1682); its stages talk only
 * through device buffers bound as descriptor sets (engine.cc:128-214).  This
 * header exports one entry point per reference stage, with the same buffer
 * layouts, plus an Engine-like renderer that runs a whole frame.
 *
 * Rules of the ABI:
 *  - Plain C: pointers, sizes and a hipStream_t passed as `void*` (NULL = the
 *    null stream).  No exceptions cross it; every call returns vkgs_status.
 *  - Pointers named d_* are device memory the caller owns; stage entry points
 *    never allocate.  Scratch ("storage") is sized by *_storage_bytes queries,
 *    as vrdxGetSorterKeyValueStorageRequirements does for the sort
 *    (engine.cc:518-520).
 *  - Counts produced on the device stay on the device (indirect semantics of
 *    vrdxCmdSortKeyValueIndirect engine.cc:1218 and vkCmdDrawIndexedIndirect
 *    engine.cc:1479): stage calls take `max_*` capacities from the host and
 *    read the live count through a device pointer.
 *  - Stream-ordered and non-blocking unless stated otherwise.
 */
#ifndef VKGS_HIP_H
#define VKGS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VKGS_ABI_VERSION 1

typedef enum vkgs_status {
  VKGS_OK = 0,
  VKGS_ERR_INVALID_ARG = 1, /* null pointer, zero size where one is required, bad enum */
  VKGS_ERR_CAPACITY = 2,    /* request exceeds a capacity (N > capacity, storage too small) */
  VKGS_ERR_HIP = 3,         /* a HIP runtime call failed */
  VKGS_ERR_IO = 4,          /* file missing or unreadable */
  VKGS_ERR_FORMAT = 5,      /* malformed or unsupported PLY */
  VKGS_ERR_OVERFLOW = 6,    /* a device-side list outgrew its storage; result incomplete */
  VKGS_ERR_NO_DEVICE = 7    /* no HIP device */
} vkgs_status;

/* Byte-compatible with vk::shader::Camera (src/vkgs/vulkan/shader/uniforms.h:10-15):
 * mat4 projection @0, mat4 view @64, vec3 camera_position @128,
 * alignas(16) uvec2 screen_size @144; alignas(64) -> sizeof 192. */
typedef struct vkgs_camera {
  float projection[16]; /* column-major, already converted to Vulkan clip space */
  float view[16];
  float camera_position[3];
  float pad0;
  uint32_t screen_size[2]; /* (width, height) */
  uint32_t pad1[10];
} vkgs_camera;

const char* vkgs_status_string(int status);
int vkgs_abi_version(void);

/* ---------------------------------------------------------------------------
 * Camera producer.  Replaces vkgs::Camera::{ProjectionMatrix, ViewMatrix, Eye}
 * (src/vkgs/scene/camera.cc:25-45): perspective(fovy, W/H, znear, zfar) with
 * the GL->Vulkan conversion (camera.cc:29-34), lookAt(eye, center, +y)
 * (camera.cc:37), eye = center + r (sin phi sin theta, cos phi, sin phi cos theta)
 * (camera.cc:39-45).  Host-only arithmetic; no device is needed.
 * ------------------------------------------------------------------------- */
int vkgs_camera_orbit(float fovy, uint32_t width, uint32_t height, float znear, float zfar,
                      const float center[3], float r, float phi, float theta, vkgs_camera* out);

/* inverse of a 4x4 column-major matrix (used for the SH view direction,
 * projection.comp:85).  Returns VKGS_ERR_INVALID_ARG for a singular matrix. */
int vkgs_mat4_inverse(const float m[16], float out[16]);

/* ---------------------------------------------------------------------------
 * PLY ingestion.  Replaces the header parse and offset table of
 * SplatLoadThread (src/vkgs/engine/splat_load_thread.cc:53-135) with checked
 * parsing: binary_little_endian only, every used property must be `float`,
 * missing properties and truncated files are errors, N > capacity is an error.
 * ------------------------------------------------------------------------- */
typedef struct vkgs_ply_info {
  uint64_t point_count;   /* element vertex N */
  uint32_t stride;        /* floats per vertex row (62 for standard 3DGS) */
  uint32_t offsets[60];   /* offset table, splat_load_thread.cc:115-135; offsets[59] == stride */
  uint64_t header_bytes;  /* byte offset of the first row */
} vkgs_ply_info;

int vkgs_ply_read_header(const char* path, vkgs_ply_info* info);
/* Host buffer of the reference's ply-buffer format: 60 u32 offsets then
 * N x stride floats (splat_load_thread.cc:92, 157-158).  dst_bytes must be
 * >= vkgs_ply_buffer_bytes(info). */
uint64_t vkgs_ply_buffer_bytes(const vkgs_ply_info* info);
int vkgs_ply_read(const char* path, const vkgs_ply_info* info, void* dst, uint64_t dst_bytes);

/* ---------------------------------------------------------------------------
 * Stage 0: activation + layout (parse_ply.comp:34-98, dispatch
 * engine.cc:1137-1161).  d_ply is the device copy of the ply buffer above.
 * Writes the vkgs splat storage (engine.cc:485-492):
 *   position f32 (N,3), cov3d f32 (N,6) = (xx,xy,xz,yy,yz,zz), opacity f32 (N),
 *   sh f16 (N,48) channel-major [R: dc, rest0..14][G: ...][B: ...].
 * ------------------------------------------------------------------------- */
int vkgs_parse_ply(const void* d_ply, uint32_t n, float* d_position, float* d_cov3d, float* d_opacity,
                   uint16_t* d_sh, void* stream);

/* ---------------------------------------------------------------------------
 * Stage 1: rank (rank.comp:27-42, dispatch engine.cc:1165-1194).
 * Centre-only NDC cull, key = floatBitsToUint(1 - z_ndc), index = splat id.
 * Unlike the reference's atomicAdd append (rank.comp:38), visible splats are
 * compacted in ascending id order, so ties in the stable sort resolve by id.
 * Writes *d_visible_count.  Needs storage of vkgs_rank_storage_bytes(n).
 * ------------------------------------------------------------------------- */
size_t vkgs_rank_storage_bytes(uint32_t max_n);
int vkgs_rank(const vkgs_camera* camera, const float model[16], const float* d_position, uint32_t n,
              uint32_t* d_visible_count, uint32_t* d_key, uint32_t* d_index, void* d_storage,
              size_t storage_bytes, void* stream);

/* ---------------------------------------------------------------------------
 * Stage 2: sort.  Replaces vrdxGetSorterKeyValueStorageRequirements
 * (engine.cc:518-520) and vrdxCmdSortKeyValueIndirect (engine.cc:1218-1219):
 * stable ascending LSD radix sort of (u32 key, u32 value) pairs in place; the
 * live count is read from *d_count (<= max_n).  key_bits limits the sorted key
 * width (32 sorts full keys; 30 suffices for rank keys, which are <= 0x3f800000).
 * ------------------------------------------------------------------------- */
size_t vkgs_sort_kv_storage_bytes(uint32_t max_n);
int vkgs_sort_kv(uint32_t max_n, const uint32_t* d_count, uint32_t* d_key, uint32_t* d_value, uint32_t key_bits,
                 void* d_storage, size_t storage_bytes, void* stream);

/* ---------------------------------------------------------------------------
 * Stage 2 without allocation.  No counterpart in the reference: this is for
 * callers that cannot allocate device memory or hand over a workspace (code
 * under hipGraph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The same sort as vkgs_sort_kv, bit for bit, with its scratch in
 * a static device buffer sized for VKGS_SORT_NOALLOC_MAX_N pairs (the
 * reference's MAX_SPLAT_COUNT, engine.cc:1653; about 65 MiB on every GPU the
 * library runs on, held until the process ends).  max_n above that returns
 * VKGS_ERR_CAPACITY.  No hipMalloc, hipFree or allocating library call; the
 * first call on a device (with that device current) looks up its buffer's
 * address.  Every call on one GPU shares that GPU's buffer: calls that
 * overlap on one GPU (several streams or threads) race, so the caller
 * serializes them.
 * ------------------------------------------------------------------------- */
#define VKGS_SORT_NOALLOC_MAX_N (1u << 23)
int vkgs_sort_kv_noalloc(uint32_t max_n, const uint32_t* d_count, uint32_t* d_key, uint32_t* d_value,
                         uint32_t key_bits, void* stream);

/* ---------------------------------------------------------------------------
 * Stage 3: inverse index (inverse_index.comp:13-18 after the -1 fill of
 * engine.cc:1226): d_inverse[d_index[k]] = k for k < *d_count, -1 elsewhere.
 * ------------------------------------------------------------------------- */
int vkgs_inverse_index(const uint32_t* d_count, const uint32_t* d_index, uint32_t n, int32_t* d_inverse,
                       void* stream);

/* ---------------------------------------------------------------------------
 * Stage 4: projection (projection.comp:60-180, dispatch engine.cc:1255-1274).
 * Writes the 48-byte instance of every visible splat at its sorted slot:
 *   [3s+0] = (ndc.xyz, 0), [3s+1] = rot_scale columns (c0.x, c0.y, c1.x, c1.y),
 *   [3s+2] = (rgb, opacity)
 * and the 12-word DrawIndexedIndirect/DrawIndirect block (projection.comp:35-48).
 * d_instances holds float4[3 * max slots].
 * ------------------------------------------------------------------------- */
int vkgs_projection(const vkgs_camera* camera, const float model[16], const float* d_position,
                    const float* d_cov3d, const float* d_opacity, const uint16_t* d_sh,
                    const int32_t* d_inverse, uint32_t n, const uint32_t* d_visible_count,
                    float* d_instances, uint32_t* d_draw_indirect, void* stream);

/* ---------------------------------------------------------------------------
 * Stage 5: draw.  Replaces the splat graphics pipeline (splat.vert:10-26,
 * splat.frag:8-12, blend engine.cc:279-299, raster state
 * graphics_pipeline.cc:58-85, vkCmdDrawIndexedIndirect engine.cc:1479) with a
 * tile rasterizer: instances [0, V) are composited in order with the
 * reference's "over" blend (SRC_ALPHA / ONE_MINUS_SRC_ALPHA on colour and
 * alpha) into a height x width RGBA fp32 image cleared to (0,0,0,1)
 * (engine.cc:1382-1387).  V is read from d_draw_indirect[8] (vertexCount1).
 * max_pairs bounds the (tile, splat) list; overflow is reported by
 * vkgs_draw_status after the stream completes.
 * ------------------------------------------------------------------------- */
typedef struct vkgs_raster_config {
  uint32_t tile_w, tile_h; /* 8x8, 16x16, 32x8, 16x8, 8x16, 16x4, 32x4, 64x4 */
  float t_min;             /* stop a pixel once transmittance < t_min (0 = composite every splat) */
  uint32_t variant;        /* 0: the workgroup stages batches of records in LDS; 1: each wave
                              fetches records with scalar loads.  Both give the same image. */
  uint32_t flags;          /* VKGS_DIFF_SPATIAL_ORDER (differentiable path only), VKGS_RASTER_TRIANGLES,
                              VKGS_RASTER_PAIRS_ALWAYS or VKGS_RASTER_PAIRS_NEVER, or 0 */
  uint32_t lds_batch;      /* list entries per LDS batch of the differentiable backward on 8x8 tiles:
                              8, 16 or 32 (0 = 16).  Deeper batches take more LDS per wave. */
  uint32_t waves_per_cu;   /* cap on the resident waves per CU of the raster kernels (plain and
                              differentiable, forward and backward): 0 = none (registers and LDS
                              decide), else 4..32.  Set through each launch's LDS size, so a cap
                              below what registers and LDS allow keeps fewer waves in flight. */
  uint32_t reserved[1];
} vkgs_raster_config;

/* vkgs_diff_forward: process the splats of every scene that has four or more
 * views in the call in a spatial (Morton) order of their means, so that a
 * wave's splats mostly share their visibility.  Faster for many views per
 * scene; images and gradients then differ from the default order only in the
 * blend order of splats with equal depth keys (the default order, ascending
 * splat id, is the one single-view calls and the plain renderer use).  The
 * order is cached per context while the scenes' means pointers and sizes stay
 * the same and refreshed every 32 forwards. */
#define VKGS_DIFF_SPATIAL_ORDER 1u

/* vkgs_raster_config.flags bit (differentiable path): when a batch's views
 * hold more tiles together than one view does by a radix pass (for instance
 * five 1600x900 views on 8x8 tiles: 113,000 tiles, 17 bits, against 22,600,
 * 15 bits), the forward sorts view-major keys: the depth key carries the view
 * index above a 25-bit integer depth (views: at most 128), so pairs are
 * emitted view by view in depth order, and the tile sort orders the tile
 * within the view alone, one pass fewer over all pairs.  The tile lists, and
 * so images and gradients, are the same bit for bit as with global tile keys;
 * this bit keeps the global keys (for checking and timing). */
#define VKGS_DIFF_GLOBAL_TILE_KEYS 16u

/* vkgs_raster_config.flags bit: rasterize each splat as the reference's
 * fixed-function pipeline does (two triangles per quad, index pattern
 * 4i + {0,1,2,2,1,3} (engine.cc:607-616), vertices snapped to 1/256 pixel,
 * the top-left tie rule, CCW front faces with back faces culled
 * (graphics_pipeline.cc:67-81), the varying p interpolated from the snapped
 * vertices (splat.vert:17-25)) instead of the analytic +-3 sigma square.
 * Slower; a parity option of the plain renderer and vkgs_draw.  Splats with
 * a quad vertex more than 10^6 pixels from the image origin are not drawn in
 * this mode. */
#define VKGS_RASTER_TRIANGLES 2u

/* vkgs_raster_config.flags bits: 8x8 tiles are rasterized two per wave (one
 * 64-lane wave per pair of tiles, two pixels per lane) when the frame or batch
 * has 8192 tiles or more, one per wave below that, where the pairs' halved
 * wave count fills the GPU less.  These force one form (both: invalid).  In
 * the differentiable path the two forms give the same image bit for bit; the
 * plain raster stops compositing when every pixel of a wave is below t_min,
 * so a pixel may add a few more splats in one form than in the other, which
 * changes it by less than t_min x max(colour, 1). */
#define VKGS_RASTER_PAIRS_ALWAYS 4u
#define VKGS_RASTER_PAIRS_NEVER 8u

/* default: 8x8 tiles, t_min = 1/8192 */
void vkgs_raster_config_default(vkgs_raster_config* cfg);

size_t vkgs_draw_storage_bytes(uint32_t max_instances, uint32_t width, uint32_t height, uint64_t max_pairs,
                               const vkgs_raster_config* cfg);
int vkgs_draw(const float* d_instances, const uint32_t* d_draw_indirect, uint32_t max_instances,
              uint32_t width, uint32_t height, uint64_t max_pairs, const vkgs_raster_config* cfg,
              float* d_rgba, void* d_storage, size_t storage_bytes, void* stream);
/* Synchronous: waits for the stream, then reports VKGS_ERR_OVERFLOW if the
 * last vkgs_draw on this storage overflowed, else VKGS_OK.  *pairs receives the
 * (tile, splat) pair count of that draw. */
int vkgs_draw_status(const void* d_storage, size_t storage_bytes, uint64_t* pairs, void* stream);

/* Tile lists of the last vkgs_draw on this storage (device pointers into the
 * storage): d_ranges = uint32 [tiles][2] = [start, end) ([0, 0) for a tile no
 * splat touches), d_list = instance indices in draw order per tile.
 * Synchronous (reads the storage header).  For parity checks and tools. */
int vkgs_draw_lists(const void* d_storage, size_t storage_bytes, const uint32_t** d_ranges,
                    const uint32_t** d_list, uint32_t* tiles_x, uint32_t* tiles_y);

/* ---------------------------------------------------------------------------
 * Renderer: the Engine analogue (Engine::Impl::Draw, engine.cc:1164-1301).
 * Owns the splat storage (preallocated to a capacity, as engine.cc:483-502
 * does for MAX_SPLAT_COUNT) and the frame workspace, and renders a frame with
 * the fused MI355X path: cull + project + compact, depth sort, tile binning,
 * tile sort, ranges, raster.  Frames are captured into a hipGraph per
 * (N, width, height, raster config) and replayed.
 * ------------------------------------------------------------------------- */
typedef struct vkgs_renderer vkgs_renderer;

int vkgs_renderer_create(uint32_t capacity, vkgs_renderer** out);
int vkgs_renderer_destroy(vkgs_renderer* r);
/* Engine::LoadSplats (engine.cc:541-544) + parse (engine.cc:1137-1161):
 * synchronous load of a PLY file into the splat storage. */
int vkgs_renderer_load_ply(vkgs_renderer* r, const char* path);
/* Asynchronous load (SplatLoadThread, splat_load_thread.cc:33-204, and
 * Engine::LoadSplatsAsync / LoadSplats, engine.cc:541-549): a loader thread
 * reads the file in 65536-row chunks into pinned memory and copies it to the
 * device; the next vkgs_renderer_render after it finishes parses it into the
 * splat storage on the render stream (frames before that show the previous
 * splats).  A new call cancels a load in progress; vkgs_renderer_cancel_load
 * cancels too (a canceled or failed load changes nothing). */
typedef enum vkgs_load_state {
  VKGS_LOAD_IDLE = 0,
  VKGS_LOAD_READING = 1,  /* header, rows or upload in progress */
  VKGS_LOAD_READY = 2,    /* on the device, taken by the next render */
  VKGS_LOAD_DONE = 3,     /* parsed into the splat storage */
  VKGS_LOAD_CANCELED = 4,
  VKGS_LOAD_FAILED = 5    /* vkgs_renderer_load_progress returns the error */
} vkgs_load_state;
int vkgs_renderer_load_ply_async(vkgs_renderer* r, const char* path);
int vkgs_renderer_cancel_load(vkgs_renderer* r);
/* Rows in the file and rows read so far, and the vkgs_load_state. */
int vkgs_renderer_load_progress(vkgs_renderer* r, uint64_t* total, uint64_t* loaded, int* state);
/* Load from a host ply buffer (60 offsets + rows), as produced by vkgs_ply_read. */
int vkgs_renderer_load_ply_buffer(vkgs_renderer* r, const void* host_ply, uint32_t n);
uint32_t vkgs_renderer_splat_count(const vkgs_renderer* r);
/* Device pointers of the splat storage (position, cov3d, opacity, sh). */
int vkgs_renderer_storage(vkgs_renderer* r, float** d_position, float** d_cov3d, float** d_opacity,
                          uint16_t** d_sh);
int vkgs_renderer_set_config(vkgs_renderer* r, const vkgs_raster_config* cfg);
/* Render one frame into d_rgba (height x width x 4 fp32, device).  camera and
 * model (NULL = identity) are host structs copied at call time. */
int vkgs_renderer_render(vkgs_renderer* r, const vkgs_camera* camera, const float model[16], float* d_rgba,
                         void* stream);
/* Per-frame results of the last render (synchronous): visible count, pair
 * count, overflow flag -> VKGS_ERR_OVERFLOW. */
int vkgs_renderer_frame_info(vkgs_renderer* r, uint32_t* visible, uint64_t* pairs, void* stream);
/* Device pointers of the last frame's depth-sorted order: d_sorted_ids[k] is
 * the splat id drawn k-th (ascending key = far to near), d_keys[k] its key. */
int vkgs_renderer_sorted(vkgs_renderer* r, const uint32_t** d_sorted_ids, const uint32_t** d_keys);
int vkgs_renderer_tile_lists(vkgs_renderer* r, const uint32_t** d_ranges, const uint32_t** d_list_ids,
                             uint32_t* tiles_x, uint32_t* tiles_y);
/* Autotuner: renders `frames` graph replays of this camera per candidate over
 * a coordinate search: tile shape (8x8 one or two tiles per wave, 16x8, 8x16,
 * 16x4, 16x16, 32x4), then the resident-wave cap (none, 24, 16, 12, 8 per
 * CU).  Each stage runs three rounds visiting its candidates in turn, times
 * them with HIP events and keeps each one's fastest round (t_min unchanged).
 * Synchronous.  The renderer keeps the chosen config; it and its ms/frame are
 * returned when the pointers are non-NULL. */
int vkgs_renderer_autotune(vkgs_renderer* r, const vkgs_camera* camera, const float model[16], float* d_rgba,
                           uint32_t frames, vkgs_raster_config* chosen, float* best_ms, void* stream);
/* Enable per-stage hipEvent timing (the reference's 12-timestamp breakdown,
 * engine.cc:1031-1045); disables graph replay while on. */
int vkgs_renderer_set_timing(vkgs_renderer* r, int enable);
/* ms of the last timed frame: [visible count + scan, project + compact,
 * depth sort, bin, tile sort + ranges, raster, total, 0] */
int vkgs_renderer_stage_ms(vkgs_renderer* r, float ms[8]);
/* Raster work of the last timed frame (timing on): the list entries the
 * raster's waves composited, summed over waves; each is one 64-lane
 * (pixel, splat) evaluation step, so executed evaluations = 64 x this.
 * Entries a pixel skipped after its early exit are not counted.
 * Synchronous. */
int vkgs_renderer_raster_work(vkgs_renderer* r, uint64_t* wave_entries);

/* ---------------------------------------------------------------------------
 * Differentiable, batched path (no reference counterpart; the reference only
 * renders).  Parameters are the raw 3DGS ones in the PLY's conventions:
 * means (N,3), log-scales (N,3) (scale_0..2), quaternions (N,4) in PLY order
 * (rot_0..rot_3 = w,x,y,z), opacity logits (N), SH (N,48) in the vkgs
 * channel-major order; activation as parse_ply.comp:46-96 (SH kept in fp32),
 * projection as projection.comp:80-179, blend as engine.cc:279-299 with alpha
 * clamped to 0.999.  Many views of many scenes render in one set of launches:
 * tiles of all views share one id space, so the tile sort key carries the view
 * in its high bits.  A context keeps one forward's state until its backward,
 * which re-walks each tile list back to front from the stored per-pixel
 * transmittance and processed-entry count.  The raster stops a pixel once its
 * transmittance is below max(cfg->t_min, 2^-20), so the stored transmittance
 * stays a normal float (t_min = 0 therefore omits at most 2^-20 of a pixel's
 * colour on this path).
 * ------------------------------------------------------------------------- */
/* sh, quats, grad_sh and grad_quats must be 16-byte aligned (VKGS_ERR_INVALID_ARG otherwise). */
typedef struct vkgs_diff_scene {
  const float* means;
  const float* log_scales;
  const float* quats;
  const float* opacity_logits;
  const float* sh;
  float* grad_means; /* every gradient element is written (not accumulated) by vkgs_diff_backward */
  float* grad_log_scales;
  float* grad_quats;
  float* grad_opacity_logits;
  float* grad_sh;
  uint32_t n;
  uint32_t reserved[3];
} vkgs_diff_scene;

typedef struct vkgs_diff_view {
  vkgs_camera camera;   /* screen_size gives the image size */
  float model[16];      /* column-major model matrix */
  uint32_t scene;       /* index into the scene table */
  uint32_t reserved[3];
  float* rgba;          /* height x width x 4 output (device) */
  const float* grad_rgba; /* dL/d rgba, same shape (backward only) */
} vkgs_diff_view;

typedef struct vkgs_diff_ctx vkgs_diff_ctx;

int vkgs_diff_create(vkgs_diff_ctx** out);
int vkgs_diff_destroy(vkgs_diff_ctx* ctx);
int vkgs_diff_forward(vkgs_diff_ctx* ctx, const vkgs_diff_scene* scenes, uint32_t n_scenes,
                      const vkgs_diff_view* views, uint32_t n_views, const vkgs_raster_config* cfg, void* stream);
int vkgs_diff_backward(vkgs_diff_ctx* ctx, const vkgs_diff_scene* scenes, uint32_t n_scenes,
                       const vkgs_diff_view* views, uint32_t n_views, void* stream);
/* Synchronous: visible items and pairs of the last forward; VKGS_ERR_OVERFLOW
 * if its pair list overflowed, in which case *pairs is the whole list's length
 * and the next forward sizes the pair storage for it (a retry then fits). */
int vkgs_diff_info(vkgs_diff_ctx* ctx, uint32_t* visible, uint64_t* pairs, void* stream);
/* Synchronous: VKGS_ERR_OVERFLOW if any forward since the previous check
 * overflowed its pair list (the results of that forward and its backward are
 * incomplete), with *pairs_needed the longest such list; the next forward then
 * sizes the pair storage for it.  Clears the record.  vkgs_diff_forward and
 * vkgs_diff_backward themselves never wait for the device or read results
 * back, so a caller may capture them (with its own work between them) into a
 * hipGraph on one stream, after an eager call has sized every buffer for the
 * same scenes and views, and check this after replays. */
int vkgs_diff_check(vkgs_diff_ctx* ctx, uint64_t* pairs_needed, void* stream);
/* Per-stage timing of the differentiable path (the reference's timestamp
 * breakdown, engine.cc:1031-1045, applied to this path): while enabled, each
 * forward and backward records HIP events around its stages and its raster
 * kernels count the list entries each wave stepped through.  For eager calls
 * (not inside a captured graph). */
int vkgs_diff_set_timing(vkgs_diff_ctx* ctx, int enable);
/* ms of the last timed forward + backward (synchronous): [forward up to the
 * raster (cull, project, sorts, binning), forward raster, forward total,
 * backward raster, gradient reduction, projection backward, backward total,
 * 0]. */
int vkgs_diff_stage_ms(vkgs_diff_ctx* ctx, float ms[8]);
/* Raster work of the last timed forward and backward (synchronous): list
 * entries stepped through, summed over waves (the two-tiles-per-wave kernels
 * count each tile's own entries); executed (pixel, entry) evaluations = 64 x
 * each.  Either pointer may be NULL. */
int vkgs_diff_raster_work(vkgs_diff_ctx* ctx, uint64_t* forward_entries, uint64_t* backward_entries);
/* Autotuner of the differentiable path (no counterpart in the reference,
 * whose renderer has fixed-function rasterization): times steps of
 * vkgs_diff_forward + vkgs_diff_backward on these scenes and views (their
 * images and gradients are overwritten; every view needs grad_rgba) over a
 * coordinate search: tile shape (8x8 one or two tiles per wave, 16x8, 8x16,
 * 16x4, 16x16, 32x4), then the LDS batch depth (8, 16, 32) when 8x8 wins,
 * then the resident-wave cap (none, 24, 16, 12, 8 per CU).  Each stage runs
 * three rounds visiting its candidates in turn and keeps each candidate's
 * fastest round.  base (NULL = default) gives t_min and the other flags.
 * *chosen = the fastest configuration, *best_ms its ms per step.
 * Synchronous. */
int vkgs_diff_autotune(vkgs_diff_ctx* ctx, const vkgs_diff_scene* scenes, uint32_t n_scenes,
                       const vkgs_diff_view* views, uint32_t n_views, const vkgs_raster_config* base,
                       uint32_t steps, vkgs_raster_config* chosen, float* best_ms, void* stream);

#ifdef __cplusplus
}
#endif
#endif
