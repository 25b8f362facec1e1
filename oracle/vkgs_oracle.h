/*
 * vkgs_oracle.h -- CPU restatement of vkgs's splat path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * implementation in vkgs_amd/.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product never links it and has
 * no CPU fallback.
 *
 * Every function restates one stage of the reference with an explicit fp32
 * operation order (built with -ffp-contract=off, so no FMA is formed unless
 * written as fmaf).  The HIP kernels follow the same order, which is what makes
 * keys, visible counts, tile ranges and sorted order bit-exact between the two.
 *
 * Reference stages restated (file:line in jaesung-cs/vkgs):
 *   oracle_parse_ply    src/shader/parse_ply.comp:34-98
 *   oracle_rank         src/shader/rank.comp:27-42
 *   oracle_sort_kv      vrdxCmdSortKeyValueIndirect (engine.cc:1218-1219): stable
 *                       ascending sort of (u32 key, u32 value)
 *   oracle_inverse      src/shader/inverse_index.comp:13-18 + fill -1 (engine.cc:1226)
 *   oracle_projection   src/shader/projection.comp:60-180
 *   oracle_draw         splat.vert:10-26 + splat.frag:8-12 + blend state
 *                       engine.cc:279-299 + clear engine.cc:1382-1387
 *   oracle_bin          tile binning of the drawn quads (no reference counterpart:
 *                       the reference uses the hardware rasterizer)
 *
 * Parity status of each function is recorded in the header comment of
 * vkgs_oracle.c and pinned by tests/test_oracle.py and tests/test_golden.py.
 */
#ifndef VKGS_ORACLE_H
#define VKGS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Byte layout of vk::shader::Camera (src/vkgs/vulkan/shader/uniforms.h:10-15):
 * mat4 projection @0, mat4 view @64, vec3 camera_position @128,
 * alignas(16) uvec2 screen_size @144, struct alignas(64) -> 192 bytes. */
typedef struct oracle_camera {
  float projection[16];
  float view[16];
  float camera_position[3];
  float pad0;
  uint32_t screen_size[2];
  uint32_t pad1[10];
} oracle_camera;

/* parse_ply.comp: ply rows are n x stride floats, offsets[59] == stride. */
void oracle_parse_ply(const float* ply, const uint32_t offsets[60], uint32_t n, float* position,
                      float* cov3d, float* opacity, uint16_t* sh);

/* rank.comp; visible splats appended in ascending id order. */
void oracle_rank(const oracle_camera* cam, const float model[16], const float* position, uint32_t n,
                 uint32_t* visible_count, uint32_t* key, uint32_t* index);

/* Stable ascending sort of count (key, value) pairs, in place. */
void oracle_sort_kv(uint32_t count, uint32_t* key, uint32_t* value);

/* inverse_index.comp: inverse[index[k]] = k for k < count, else -1. */
void oracle_inverse(uint32_t count, const uint32_t* index, uint32_t n, int32_t* inverse);

/* projection.comp: writes instances (float4 x 3 per sorted slot) and the
 * 12-word DrawIndirect block.  model_inv is inverse(model) computed once by the
 * caller (the shader computes it per invocation; for the identity and for
 * exactly invertible affine models this is the same matrix). */
void oracle_projection(const oracle_camera* cam, const float model[16], const float model_inv[16],
                       const float* position, const float* cov3d, const float* opacity,
                       const uint16_t* sh, const int32_t* inverse, uint32_t n, uint32_t visible_count,
                       float* instances, uint32_t draw_indirect[12]);

/* Per-instance raster set-up shared by oracle_draw and oracle_bin.
 * Returns 1 when the quad produces fragments and fills the conservative pixel
 * rectangle [x0,x1]x[y0,y1] (inclusive). */
typedef struct oracle_splat2d {
  float cx, cy;             /* centre in framebuffer pixels */
  float i00, i01, i10, i11; /* inverse of the pixel-space rot_scale */
  float r, g, b, o;
  int32_t x0, y0, x1, y1;
} oracle_splat2d;
int oracle_setup_splat(const float* instance, uint32_t width, uint32_t height, oracle_splat2d* s);

/* Back-to-front "over" compositing of count instances in draw order into a
 * height x width x 4 fp32 image cleared to (0,0,0,1). */
void oracle_draw(const float* instances, uint32_t count, uint32_t width, uint32_t height, float* rgba);

/* Same image by brute force over every pixel and every instance (no bounding
 * rectangle); used to show the rectangle in oracle_setup_splat is conservative. */
void oracle_draw_bruteforce(const float* instances, uint32_t count, uint32_t width, uint32_t height,
                            float* rgba);

/* Tile binning: for tiles of tile_w x tile_h pixels, ranges[2t..2t+1] is the
 * [start,end) of tile t in list[], and list[] holds instance indices in draw
 * order within each tile.  Returns the number of (tile, instance) pairs; list
 * may be NULL to count only. */
uint64_t oracle_bin(const float* instances, uint32_t count, uint32_t width, uint32_t height,
                    uint32_t tile_w, uint32_t tile_h, uint32_t* ranges, uint32_t* list);

/* The draw as the reference's fixed-function pipeline rasterizes it: two
 * triangles per instance (indices 4i + {0,1,2,2,1,3}), vertices snapped to
 * 1/2^subpixel_bits of a pixel, the top-left tie rule, CCW front faces with
 * back faces culled, the varying p interpolated from the snapped vertices;
 * back-to-front "over" compositing as oracle_draw.  For measuring how far the
 * analytic coverage of oracle_draw and the HIP raster sits from it. */
void oracle_draw_triangles(const float* instances, uint32_t count, uint32_t width, uint32_t height,
                           uint32_t subpixel_bits, float* rgba);

/* Per-pixel evaluation used by every raster path: returns 1 and alpha when the
 * pixel centre (px+0.5, py+0.5) lies in the +-3 sigma square. */
int oracle_eval_pixel(const oracle_splat2d* s, int px, int py, float* alpha);

/* fp16 conversion (round to nearest even) and back. */
uint16_t oracle_f32_to_f16(float f);
float oracle_f16_to_f32(uint16_t h);

/* Number of OpenMP threads the timed oracle loops use (1 when built without). */
int oracle_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
