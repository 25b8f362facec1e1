// C ABI of the differentiable, batched path (vkgs_diff_*).
//
// One context holds the state of one forward (records, tile lists, per-pixel
// transmittance and processed counts) until its backward.  Views of several
// scenes render in one set of launches; see diff.h for the layout.
#include <string.h>

#include <vector>

#include "autotune.h"
#include "common.h"
#include "diff.h"
#include "kernels.h"
#include "sort.h"

using namespace vkgs;

namespace {

struct Buf {
  void* p = nullptr;
  size_t bytes = 0;
  hipError_t ensure(size_t want) {
    if (want <= bytes) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    hipError_t e = hipMalloc(&p, want ? want : 1);
    if (e == hipSuccess) bytes = want;
    return e;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <typename T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

int bits_for(uint32_t v) {
  int b = 0;
  while (b < 32 && (1ull << b) < v) ++b;
  return b;
}

}  // namespace

struct vkgs_diff_ctx {
  // device tables (one packed upload: views, scenes, range bases, scene views)
  Buf tables;
  // per-item buffers
  uint32_t cap_items = 0;
  Buf counts, count_scan, rec, rc, key, val, val64, vslot, masks, kview, misc, rect, cnt, offs, scnt, soffs,
      scan_storage, depth_sort;
  // pairs
  uint64_t max_pairs = 0;
  Buf pkeys, pvals, pslot, gpart, pair_sort, ranges;
  // longest-list-first tile schedule
  Buf tile_keys, tile_order, tile_sort;
  // per-pixel state
  Buf final_t, last;
  // spatial processing order of the scenes' splats (Morton order in each
  // scene's bounding box, so a wave's splats tend to share their visibility):
  // reused while every scene's means pointer and size stay the same, and
  // refreshed every kPermRefresh forwards; any permutation is correct, the
  // order only affects speed
  Buf perm_keys, perm_vals, perm_bbox_part, perm_bbox, perm_sort;
  SortPlan perm_plan{};
  std::vector<std::pair<const float*, uint32_t>> perm_for;
  const void* perm_buf = nullptr;  // perm_vals storage the cached order lives in
  uint32_t perm_age = 0;
  // host mirrors of the last forward
  std::vector<DiffView> h_views;
  std::vector<DiffScene> h_scenes;
  std::vector<uint32_t> h_view_blocks, h_view_tiles, h_scene_blocks, h_scene_views;
  std::vector<unsigned char> h_tables;
  DiffLaunch launch{};
  SortPlan depth_plan{}, pair_plan{}, tile_plan{};
  ScanPlan scan{};
  bool have_forward = false;
  bool grow_pairs = false;
  uint64_t need_pairs = 0;  // pair count of the last overflowing forward
  // per-stage events and raster work counters (vkgs_diff_set_timing)
  bool timing = false;
  hipEvent_t ev[7] = {};
  Buf work_f, work_b;
  size_t work_words = 0;
};

namespace {

// misc words: per forward (cleared by it) visible, pairs, status, slot-order
// pairs; across forwards (cleared by vkgs_diff_check) an overflow flag and the
// largest overflowing pair count
enum : uint32_t {
  kMiscVisible = 0,
  kMiscPairs = 1,
  kMiscStatus = 2,
  kMiscSlotPairs = 3,
  kMiscSticky = 4,
  kMiscStickyPairs = 5
};
constexpr uint32_t kPermRefresh = 32;
// Smallest stop threshold of the differentiable raster: the backward recovers
// the transmittance in front of each entry as T_i = T_{i+1} / (1 - alpha_i)
// from the stored final T, which stays >= kDiffTMinFloor * (1 - kAlphaMax), a
// normal float, only if compositing stops there (with t_min = 0, 16 saturated
// splats take T to exactly 0 and every recovered T with it).  The omitted
// contribution is below 2^-20 of the pixel's colour.
constexpr float kDiffTMinFloor = 1.0f / 1048576.0f;
constexpr uint32_t kSpatialViews = 4;  // views of a scene in one call for the spatial order to pay

template <typename T>
size_t pack(std::vector<unsigned char>& out, const std::vector<T>& v) {
  const size_t off = align_up(out.size(), 256);
  out.resize(off + sizeof(T) * v.size());
  if (!v.empty()) memcpy(out.data() + off, v.data(), sizeof(T) * v.size());
  return off;
}

// Packs the host tables into one buffer, uploads it through kernel arguments
// (stream-ordered: no host sync, and a captured graph replays the same
// tables) and points the launch's table pointers into it.
hipError_t upload_tables(vkgs_diff_ctx* c, hipStream_t s) {
  std::vector<unsigned char>& h = c->h_tables;
  h.clear();
  const size_t o_views = pack(h, c->h_views), o_scenes = pack(h, c->h_scenes);
  const size_t o_vb = pack(h, c->h_view_blocks), o_vt = pack(h, c->h_view_tiles);
  const size_t o_sb = pack(h, c->h_scene_blocks), o_sv = pack(h, c->h_scene_views);
  hipError_t e = c->tables.ensure(h.size());
  if (e != hipSuccess) return e;
  h.resize(align_up(h.size(), 4));
  e = upload_launch(c->tables.p, h.data(), h.size(), s);
  if (e != hipSuccess) return e;
  unsigned char* base = c->tables.as<unsigned char>();
  DiffLaunch& L = c->launch;
  L.views = reinterpret_cast<const DiffView*>(base + o_views);
  L.scenes = reinterpret_cast<const DiffScene*>(base + o_scenes);
  L.view_blocks = reinterpret_cast<const uint32_t*>(base + o_vb);
  L.view_tiles = reinterpret_cast<const uint32_t*>(base + o_vt);
  L.scene_blocks = reinterpret_cast<const uint32_t*>(base + o_sb);
  L.scene_views = reinterpret_cast<const uint32_t*>(base + o_sv);
  return hipSuccess;
}

bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

}  // namespace

extern "C" {

int vkgs_diff_create(vkgs_diff_ctx** out) {
  if (!out) return VKGS_ERR_INVALID_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return VKGS_ERR_NO_DEVICE;
  *out = new vkgs_diff_ctx();
  return VKGS_OK;
}

int vkgs_diff_destroy(vkgs_diff_ctx* c) {
  if (!c) return VKGS_ERR_INVALID_ARG;
  (void)hipDeviceSynchronize();
  Buf* bufs[] = {&c->tables,     &c->counts,   &c->count_scan,   &c->rec,        &c->rc,         &c->key,
                 &c->val,        &c->masks,    &c->kview,        &c->misc,       &c->rect,
                 &c->cnt,        &c->offs,     &c->scnt,         &c->soffs,      &c->scan_storage, &c->depth_sort,
                 &c->pkeys,      &c->pvals,    &c->pair_sort,    &c->ranges,     &c->final_t,    &c->last,
                 &c->tile_keys,  &c->tile_order, &c->tile_sort,  &c->pslot,      &c->gpart,
                 &c->val64,      &c->vslot,    &c->perm_keys,    &c->perm_vals,  &c->perm_bbox_part,
                 &c->perm_bbox,  &c->perm_sort};
  for (Buf* b : bufs) b->release();
  c->work_f.release();
  c->work_b.release();
  for (hipEvent_t e : c->ev)
    if (e) (void)hipEventDestroy(e);
  delete c;
  return VKGS_OK;
}

int vkgs_diff_forward(vkgs_diff_ctx* c, const vkgs_diff_scene* scenes, uint32_t n_scenes, const vkgs_diff_view* views,
                      uint32_t n_views, const vkgs_raster_config* cfg, void* stream) {
  if (!c || !scenes || !views || n_scenes == 0 || n_views == 0) return VKGS_ERR_INVALID_ARG;
  hipStream_t s = static_cast<hipStream_t>(stream);
  vkgs_raster_config dcfg;
  if (!cfg) {
    vkgs_raster_config_default(&dcfg);
    cfg = &dcfg;
  }
  // ---- host tables ----
  c->h_scenes.assign(n_scenes, DiffScene{});
  c->h_scene_blocks.assign(n_scenes + 1, 0u);
  uint64_t sblocks = 0;
  for (uint32_t i = 0; i < n_scenes; ++i) {
    const vkgs_diff_scene& a = scenes[i];
    if (a.n > 0 && (!a.means || !a.log_scales || !a.quats || !a.opacity_logits || !a.sh)) return VKGS_ERR_INVALID_ARG;
    if (a.n > 0 && (!aligned16(a.sh) || !aligned16(a.quats))) return VKGS_ERR_INVALID_ARG;
    if (a.n > (1u << 28)) return VKGS_ERR_CAPACITY;
    DiffScene& d = c->h_scenes[i];
    d.means = a.means;
    d.log_scales = a.log_scales;
    d.quats = a.quats;
    d.opacity = a.opacity_logits;
    d.sh = a.sh;
    d.g_means = a.grad_means;
    d.g_log_scales = a.grad_log_scales;
    d.g_quats = a.grad_quats;
    d.g_opacity = a.grad_opacity_logits;
    d.g_sh = a.grad_sh;
    d.n = a.n;
    d.block_base = (uint32_t)sblocks;
    c->h_scene_blocks[i] = (uint32_t)sblocks;
    sblocks += ceil_div(a.n, kCompactBlock);
  }
  c->h_scene_blocks[n_scenes] = (uint32_t)sblocks;
  RasterShape rs0;
  if (!make_raster_shape(16, 16, cfg, &rs0)) return VKGS_ERR_INVALID_ARG;
  if (rs0.triangles) return VKGS_ERR_INVALID_ARG;  // the differentiable raster is analytic only
  c->h_views.assign(n_views, DiffView{});
  c->h_view_blocks.assign(n_views + 1, 0u);
  c->h_view_tiles.assign(n_views + 1, 0u);
  uint64_t blocks = 0, tiles = 0, pixels = 0;
  uint32_t max_tiles_x = 1, max_tiles_y = 1, max_view_tiles = 1;
  for (uint32_t v = 0; v < n_views; ++v) {
    const vkgs_diff_view& a = views[v];
    if (a.scene >= n_scenes || !a.rgba) return VKGS_ERR_INVALID_ARG;
    RasterShape rs;
    if (!make_raster_shape(a.camera.screen_size[0], a.camera.screen_size[1], cfg, &rs)) return VKGS_ERR_INVALID_ARG;
    DiffView& d = c->h_views[v];
    make_frame_const(&a.camera, a.model, &d.fc);
    m4_mul(d.fc.V, d.fc.M, d.vm);
    d.fc.tile_w = rs.tile_w;
    d.fc.tile_h = rs.tile_h;
    d.fc.tiles_x = rs.tiles_x;
    d.fc.tiles_y = rs.tiles_y;
    d.scene = a.scene;
    d.block_base = (uint32_t)blocks;
    d.tile_base = (uint32_t)tiles;
    d.out = reinterpret_cast<float4*>(a.rgba);
    d.grad_out = reinterpret_cast<const float4*>(a.grad_rgba);
    c->h_view_blocks[v] = (uint32_t)blocks;
    c->h_view_tiles[v] = (uint32_t)tiles;
    blocks += ceil_div(c->h_scenes[a.scene].n, kCompactBlock);
    tiles += (uint64_t)rs.tiles_x * rs.tiles_y;
    pixels += (uint64_t)rs.width * rs.height;
    max_tiles_x = rs.tiles_x > max_tiles_x ? rs.tiles_x : max_tiles_x;
    max_tiles_y = rs.tiles_y > max_tiles_y ? rs.tiles_y : max_tiles_y;
    if (rs.tiles_x * rs.tiles_y > max_view_tiles) max_view_tiles = rs.tiles_x * rs.tiles_y;
    if (blocks * kCompactBlock > (1ull << 30) || tiles > (1u << 24)) return VKGS_ERR_CAPACITY;
  }
  c->h_view_blocks[n_views] = (uint32_t)blocks;
  c->h_view_tiles[n_views] = (uint32_t)tiles;
  // views grouped by scene, in view order within a scene
  c->h_scene_views.clear();
  for (uint32_t i = 0; i < n_scenes; ++i) {
    c->h_scenes[i].view_begin = (uint32_t)c->h_scene_views.size();
    for (uint32_t v = 0; v < n_views; ++v)
      if (views[v].scene == i) c->h_scene_views.push_back(v);
    c->h_scenes[i].view_end = (uint32_t)c->h_scene_views.size();
  }
  const uint32_t items = (uint32_t)(blocks * kCompactBlock);
  const uint32_t cap = items ? items : 1;

  // ---- device buffers ----
  hipError_t e = hipSuccess;
  if (c->max_pairs == 0 || c->grow_pairs || c->max_pairs < (uint64_t)cap) {
    const uint64_t want = (uint64_t)cap * 4 + (1u << 20);
    if (c->grow_pairs) {  // room for the overflowing forward's whole list, plus an eighth
      const uint64_t need = c->need_pairs + c->need_pairs / 8 + (1u << 20);
      c->max_pairs = need > c->max_pairs * 2 ? need : c->max_pairs * 2;
    } else {
      c->max_pairs = want > c->max_pairs ? want : c->max_pairs;
    }
    if (c->max_pairs > 0xffffffffull) c->max_pairs = 0xffffffffull;
    c->grow_pairs = false;
  }
  // tile rectangles packed into the depth sort's value when every view's
  // tile grid fits (else (rect, count, view) words gathered in depth order)
  RectBits rb;
  const bool packed = rect_bits(max_tiles_x, max_tiles_y, &rb);
  c->depth_plan = sort_plan(cap, packed ? 8u : 4u);
  c->pair_plan = sort_plan((uint32_t)c->max_pairs);
  c->scan = scan_plan(cap);
  c->tile_plan = sort_plan((uint32_t)(tiles ? tiles : 1));
#define ENSURE(buf, bytes) \
  if (e == hipSuccess) e = c->buf.ensure(bytes)
  ENSURE(counts, (size_t)(blocks ? blocks : 1) * kMaskWords * 4);
  ENSURE(count_scan, diff_count_scan_bytes((uint32_t)(blocks ? blocks : 1)));
  ENSURE(rec, (size_t)cap * 48);
  if (packed) {
    ENSURE(val64, (size_t)cap * 8);
    ENSURE(vslot, ((size_t)n_views + 1) * 4);
  } else {
    ENSURE(rc, (size_t)cap * 16);
    ENSURE(val, (size_t)cap * 4);
    ENSURE(kview, (size_t)cap * 4);
    ENSURE(rect, (size_t)cap * sizeof(TileRect));
    ENSURE(cnt, (size_t)cap * 4);
  }
  ENSURE(key, (size_t)cap * 4);
  ENSURE(masks, (size_t)(blocks ? blocks : 1) * kMaskWords * 8);
  const void* old_misc = c->misc.p;
  ENSURE(misc, 256);
  if (e == hipSuccess && c->misc.p != old_misc) e = fill_u32_launch(c->misc.as<uint32_t>(), 0u, 64, s);
  ENSURE(offs, (size_t)cap * 4);
  ENSURE(scnt, (size_t)cap * 4);
  ENSURE(soffs, (size_t)cap * 4);
  ENSURE(scan_storage, c->scan.bytes);
  ENSURE(depth_sort, c->depth_plan.bytes);
  ENSURE(pkeys, (size_t)c->max_pairs * 4);
  ENSURE(pvals, (size_t)c->max_pairs * 4);
  ENSURE(pslot, (size_t)c->max_pairs * 4);
  ENSURE(gpart, (size_t)c->max_pairs * 36);  // per-pair partials of the backward: the device-side bound
  ENSURE(pair_sort, c->pair_plan.bytes);
  ENSURE(ranges, (size_t)(tiles ? tiles : 1) * 8);
  ENSURE(tile_keys, (size_t)(tiles ? tiles : 1) * 4);
  ENSURE(tile_order, (size_t)(tiles ? tiles : 1) * 4);
  ENSURE(tile_sort, c->tile_plan.bytes);
  ENSURE(final_t, (size_t)(pixels ? pixels : 1) * 4);
  ENSURE(last, (size_t)(pixels ? pixels : 1) * 4);
  const bool spatial = (cfg->flags & VKGS_DIFF_SPATIAL_ORDER) != 0;
  bool use_perm = false;
  const size_t padded = (size_t)(sblocks ? sblocks : 1) * kCompactBlock;
  if (spatial) {
    c->perm_plan = sort_plan((uint32_t)padded);
    ENSURE(perm_keys, padded * 4);
    ENSURE(perm_vals, padded * 4);
    ENSURE(perm_bbox_part, (size_t)(sblocks ? sblocks : 1) * 6 * 4);
    ENSURE(perm_bbox, (size_t)n_scenes * 6 * 4);
    ENSURE(perm_sort, c->perm_plan.bytes);
  }
#undef ENSURE
  if (e != hipSuccess) return VKGS_ERR_HIP;
  {
    std::vector<std::pair<const float*, uint32_t>> want(n_scenes);
    uint64_t base = 0;
    for (uint32_t i = 0; i < n_scenes; ++i) {
      want[i] = {c->h_scenes[i].means, c->h_scenes[i].n};
      const bool many = c->h_scenes[i].view_end - c->h_scenes[i].view_begin >= kSpatialViews;
      if (spatial && many) {
        c->h_scenes[i].perm = c->perm_vals.as<uint32_t>() + base;
        use_perm = true;
      }
      base += c->h_scenes[i].n;
    }
    // recomputed below (after the tables are on the device) when the scenes
    // changed, the storage moved or the order is due for a refresh
    if (use_perm && (want != c->perm_for || c->perm_buf != c->perm_vals.p || c->perm_age >= kPermRefresh)) {
      c->perm_for = want;
      c->perm_buf = c->perm_vals.p;
      c->perm_age = kPermRefresh;  // "compute now"
    }
  }
  uint64_t pix = 0;
  for (uint32_t v = 0; v < n_views; ++v) {
    DiffView& d = c->h_views[v];
    d.final_t = c->final_t.as<float>() + pix;
    d.last = c->last.as<uint32_t>() + pix;
    pix += (uint64_t)d.fc.width * d.fc.height;
  }
  VKGS_HIP_CHECK(upload_tables(c, s));

  DiffLaunch& L = c->launch;
  L.n_views = n_views;
  L.n_scenes = n_scenes;
  L.n_blocks = (uint32_t)blocks;
  L.n_tiles = (uint32_t)tiles;
  L.n_scene_blocks = (uint32_t)sblocks;
  L.tile_w = rs0.tile_w;
  L.tile_h = rs0.tile_h;
  L.tile_shift_x = rs0.tile_shift_x;
  L.tile_shift_y = rs0.tile_shift_y;
  L.t_min = cfg->t_min > kDiffTMinFloor ? cfg->t_min : kDiffTMinFloor;
  L.pairs = rs0.pairs;
  L.lds_batch = rs0.lds_batch;
  L.waves_per_cu = rs0.waves_per_cu;
  c->cap_items = cap;
  // view-major local tile keys when they save a tile-sort pass and the view
  // index fits above the 25-bit depth key (128 views) and the local tile key
  const uint32_t view_bits = (uint32_t)bits_for(n_views);
  const uint32_t global_bits = (uint32_t)bits_for((uint32_t)tiles);
  const uint32_t local_bits = bits_for(max_view_tiles) > 1 ? (uint32_t)bits_for(max_view_tiles) : 1u;
  const bool local = !(cfg->flags & VKGS_DIFF_GLOBAL_TILE_KEYS) && view_bits + kDepthRankBits <= 32 &&
                     view_bits + local_bits <= 31 && sort_passes(local_bits) < sort_passes(global_bits);
  L.local_tile_bits = local ? local_bits : 0u;
  const uint32_t depth_bits = local ? kDepthRankBits + view_bits : kDepthKeyBits;
  const uint32_t depth_narrow = local ? kDepthRankNarrowPass : kDepthNarrowPass;
  const uint32_t tile_bits = local ? local_bits : global_bits;

  // ---- processing order (when due) ----
  if (use_perm && c->perm_age == kPermRefresh) {
    VKGS_HIP_CHECK(launch_diff_morton_keys(L, c->perm_bbox_part.as<uint32_t>(), c->perm_bbox.as<uint32_t>(),
                                           c->perm_keys.as<uint32_t>(), c->perm_vals.as<uint32_t>(), s));
    VKGS_HIP_CHECK(sort_kv_launch(c->perm_plan, nullptr, c->perm_keys.as<uint32_t>(), c->perm_vals.as<uint32_t>(),
                                  32u, c->perm_sort.p, s));
    c->perm_age = 0;
  }
  if (use_perm) ++c->perm_age;

  // ---- forward ----
  uint32_t* misc = c->misc.as<uint32_t>();
  uint32_t* order = c->val.as<uint32_t>();
  L.work = nullptr;
  if (c->timing) {
    const uint32_t wg_waves = rs0.tile_w * rs0.tile_h / kWave;
    c->work_words = (size_t)L.n_tiles * (wg_waves ? wg_waves : 1u);
    VKGS_HIP_CHECK(c->work_f.ensure(c->work_words * 4 + 4));
    VKGS_HIP_CHECK(c->work_b.ensure(c->work_words * 4 + 4));
    VKGS_HIP_CHECK(fill_u32_launch(c->work_f.as<uint32_t>(), 0u, c->work_words, s));
    VKGS_HIP_CHECK(fill_u32_launch(c->work_b.as<uint32_t>(), 0u, c->work_words, s));
    VKGS_HIP_CHECK(hipEventRecord(c->ev[0], s));
  }
  VKGS_HIP_CHECK(fill_u32_launch(misc, 0u, 4, s));
  VKGS_HIP_CHECK(launch_diff_count(L, c->counts.as<uint32_t>(), c->masks.as<uint64_t>(), misc + kMiscVisible,
                                   c->count_scan.p, s));
  if (packed)
    VKGS_HIP_CHECK(launch_diff_write_packed(L, c->counts.as<uint32_t>(), c->masks.as<uint64_t>(),
                                            c->rec.as<float4>(), c->scnt.as<uint32_t>(), c->key.as<uint32_t>(),
                                            c->val64.as<uint64_t>(), rb, s));
  else
    VKGS_HIP_CHECK(launch_diff_write(L, c->counts.as<uint32_t>(), c->masks.as<uint64_t>(), c->rec.as<float4>(),
                                     c->rc.as<uint4>(), c->scnt.as<uint32_t>(), c->key.as<uint32_t>(), order, s));
  // slot-order numbering of the pairs (their place in the backward's partials)
  VKGS_HIP_CHECK(scan_exclusive_launch(c->scan, c->scnt.as<uint32_t>(), c->soffs.as<uint32_t>(), misc + kMiscVisible,
                                       misc + kMiscSlotPairs, c->scan_storage.p, s));
  if (packed) {
    VKGS_HIP_CHECK(sort_kv64_launch(c->depth_plan, misc + kMiscVisible, c->key.as<uint32_t>(),
                                    c->val64.as<uint64_t>(), depth_bits, c->depth_sort.p, s, depth_narrow));
    VKGS_HIP_CHECK(launch_diff_view_slots(L, c->counts.as<uint32_t>(), misc + kMiscVisible, c->vslot.as<uint32_t>(),
                                          s));
    VKGS_HIP_CHECK(scan_rect_counts_launch(c->scan, c->val64.as<uint64_t>(), rb, c->offs.as<uint32_t>(),
                                           misc + kMiscVisible, misc + kMiscPairs, c->scan_storage.p, s));
  } else {
    VKGS_HIP_CHECK(sort_kv_launch(c->depth_plan, misc + kMiscVisible, c->key.as<uint32_t>(), order, depth_bits,
                                  c->depth_sort.p, s, depth_narrow));
    VKGS_HIP_CHECK(launch_diff_bin_gather(c->rc.as<uint4>(), order, misc + kMiscVisible, cap,
                                          c->rect.as<TileRect>(), c->cnt.as<uint32_t>(), c->kview.as<uint32_t>(), s));
    VKGS_HIP_CHECK(scan_exclusive_launch(c->scan, c->cnt.as<uint32_t>(), c->offs.as<uint32_t>(),
                                         misc + kMiscVisible, misc + kMiscPairs, c->scan_storage.p, s));
  }
  // with an odd number of tile-sort passes the pairs are emitted into the
  // sort's alternate buffers, so the sorted result lands in pkeys / pvals
  // without a final copy
  const bool odd = sort_passes(tile_bits) % 2 == 1;
  uint32_t *ek = c->pkeys.as<uint32_t>(), *ev = c->pvals.as<uint32_t>();
  if (odd) sort_alt_buffers(c->pair_plan, c->pair_sort.p, &ek, &ev);
  if (packed)
    VKGS_HIP_CHECK(launch_diff_bin_emit_packed(L, c->val64.as<uint64_t>(), c->offs.as<uint32_t>(),
                                               c->vslot.as<uint32_t>(), rb, misc + kMiscVisible, cap, c->max_pairs,
                                               ek, ev, c->pslot.as<uint32_t>(), c->soffs.as<uint32_t>(),
                                               c->scnt.as<uint32_t>(), misc + kMiscStatus, s));
  else
    VKGS_HIP_CHECK(launch_diff_bin_emit(L, c->rect.as<TileRect>(), c->cnt.as<uint32_t>(), c->offs.as<uint32_t>(),
                                        c->kview.as<uint32_t>(), order, misc + kMiscVisible, cap, c->max_pairs, ek,
                                        ev, c->pslot.as<uint32_t>(), c->soffs.as<uint32_t>(),
                                        c->scnt.as<uint32_t>(), misc + kMiscStatus, s));
  if (odd)
    VKGS_HIP_CHECK(sort_kv_from_alt_launch(c->pair_plan, misc + kMiscPairs, c->pkeys.as<uint32_t>(),
                                           c->pvals.as<uint32_t>(), tile_bits, c->pair_sort.p, s));
  else
    VKGS_HIP_CHECK(sort_kv_launch(c->pair_plan, misc + kMiscPairs, c->pkeys.as<uint32_t>(),
                                  c->pvals.as<uint32_t>(), tile_bits, c->pair_sort.p, s));
  // (status, pairs) -> the sticky record read by vkgs_diff_check
  VKGS_HIP_CHECK(fold_status_launch(misc + kMiscStatus, misc + kMiscPairs, misc + kMiscSticky,
                                    misc + kMiscStickyPairs, s));
  VKGS_HIP_CHECK(launch_ranges(c->pkeys.as<uint32_t>(), misc + kMiscPairs, c->max_pairs, (uint32_t)tiles,
                               c->ranges.as<uint32_t>(), s, false, local ? L.view_tiles : nullptr, local_bits));
  VKGS_HIP_CHECK(launch_tile_order(c->ranges.as<uint32_t>(), (uint32_t)tiles, c->tile_keys.as<uint32_t>(),
                                   c->tile_order.as<uint32_t>(), c->tile_plan, c->tile_sort.p, s));
  if (c->timing) {
    VKGS_HIP_CHECK(hipEventRecord(c->ev[1], s));
    L.work = c->work_f.as<uint32_t>();
  }
  VKGS_HIP_CHECK(launch_diff_raster(L, c->rec.as<float4>(), c->pvals.as<uint32_t>(), c->ranges.as<uint32_t>(),
                                    c->pslot.as<uint32_t>(), c->tile_order.as<uint32_t>(), true, s));
  L.work = nullptr;
  if (c->timing) VKGS_HIP_CHECK(hipEventRecord(c->ev[2], s));
  c->have_forward = true;
  return VKGS_OK;
}

int vkgs_diff_backward(vkgs_diff_ctx* c, const vkgs_diff_scene* scenes, uint32_t n_scenes, const vkgs_diff_view* views,
                       uint32_t n_views, void* stream) {
  if (!c || !views || !scenes || !c->have_forward || n_views != c->h_views.size() || n_scenes != c->h_scenes.size())
    return VKGS_ERR_INVALID_ARG;
  hipStream_t s = static_cast<hipStream_t>(stream);
  for (uint32_t v = 0; v < n_views; ++v) {
    if (!views[v].grad_rgba) return VKGS_ERR_INVALID_ARG;
    c->h_views[v].grad_out = reinterpret_cast<const float4*>(views[v].grad_rgba);
  }
  for (uint32_t i = 0; i < n_scenes; ++i) {
    const vkgs_diff_scene& a = scenes[i];
    DiffScene& d = c->h_scenes[i];
    if (a.n != d.n) return VKGS_ERR_INVALID_ARG;
    if (d.n > 0 && (!a.grad_means || !a.grad_log_scales || !a.grad_quats || !a.grad_opacity_logits || !a.grad_sh))
      return VKGS_ERR_INVALID_ARG;
    if (d.n > 0 && (!aligned16(a.grad_sh) || !aligned16(a.grad_quats))) return VKGS_ERR_INVALID_ARG;
    d.g_means = a.grad_means;
    d.g_log_scales = a.grad_log_scales;
    d.g_quats = a.grad_quats;
    d.g_opacity = a.grad_opacity_logits;
    d.g_sh = a.grad_sh;
  }
  const bool timed = c->timing && c->work_words > 0;
  if (timed) VKGS_HIP_CHECK(hipEventRecord(c->ev[3], s));
  VKGS_HIP_CHECK(upload_tables(c, s));
  DiffLaunch L = c->launch;
  L.work = timed ? c->work_b.as<uint32_t>() : nullptr;
  // per-pair partials: gpart holds max_pairs, the bound of the live count the
  // kernels read on the device (no host read-back, so a graph can capture this)
  VKGS_HIP_CHECK(launch_diff_raster_backward(L, c->rec.as<float4>(), c->pvals.as<uint32_t>(),
                                             c->ranges.as<uint32_t>(), c->pslot.as<uint32_t>(),
                                             c->tile_order.as<uint32_t>(), c->gpart.as<float>(), s));
  // the per-slot reduction of the partials runs inside the projection
  // backward: its stage time (ev[4] .. ev[5]) is empty
  if (timed) VKGS_HIP_CHECK(hipEventRecord(c->ev[4], s));
  if (timed) VKGS_HIP_CHECK(hipEventRecord(c->ev[5], s));
  L.work = nullptr;
  VKGS_HIP_CHECK(launch_diff_project_backward(L, c->counts.as<uint32_t>(), c->masks.as<uint64_t>(),
                                              c->soffs.as<uint32_t>(), c->scnt.as<uint32_t>(),
                                              c->gpart.as<float>(), c->max_pairs, s));
  if (timed) VKGS_HIP_CHECK(hipEventRecord(c->ev[6], s));
  return VKGS_OK;
}

int vkgs_diff_set_timing(vkgs_diff_ctx* c, int enable) {
  if (!c) return VKGS_ERR_INVALID_ARG;
  if (enable && !c->ev[0])
    for (hipEvent_t& e : c->ev) VKGS_HIP_CHECK(hipEventCreate(&e));
  c->timing = enable != 0;
  c->work_words = 0;  // valid again after the next timed forward
  return VKGS_OK;
}

int vkgs_diff_stage_ms(vkgs_diff_ctx* c, float ms[8]) {
  if (!c || !ms || !c->timing || c->work_words == 0) return VKGS_ERR_INVALID_ARG;
  VKGS_HIP_CHECK(hipEventSynchronize(c->ev[6]));
  const int pairs[7][2] = {{0, 1}, {1, 2}, {0, 2}, {3, 4}, {4, 5}, {5, 6}, {3, 6}};
  for (int i = 0; i < 7; ++i) VKGS_HIP_CHECK(hipEventElapsedTime(&ms[i], c->ev[pairs[i][0]], c->ev[pairs[i][1]]));
  ms[7] = 0.f;
  return VKGS_OK;
}

int vkgs_diff_raster_work(vkgs_diff_ctx* c, uint64_t* forward_entries, uint64_t* backward_entries) {
  if (!c || !c->timing || c->work_words == 0) return VKGS_ERR_INVALID_ARG;
  std::vector<uint32_t> w(c->work_words);
  VKGS_HIP_CHECK(hipDeviceSynchronize());
  const Buf* bufs[2] = {&c->work_f, &c->work_b};
  uint64_t* outs[2] = {forward_entries, backward_entries};
  for (int k = 0; k < 2; ++k) {
    if (!outs[k]) continue;
    VKGS_HIP_CHECK(hipMemcpy(w.data(), bufs[k]->p, c->work_words * 4, hipMemcpyDeviceToHost));
    uint64_t sum = 0;
    for (uint32_t x : w) sum += x;
    *outs[k] = sum;
  }
  return VKGS_OK;
}

int vkgs_diff_check(vkgs_diff_ctx* c, uint64_t* pairs_needed, void* stream) {
  if (!c || !c->misc.p) return VKGS_ERR_INVALID_ARG;
  hipStream_t s = static_cast<hipStream_t>(stream);
  uint32_t m[8];
  VKGS_HIP_CHECK(hipStreamSynchronize(s));
  VKGS_HIP_CHECK(hipMemcpy(m, c->misc.p, sizeof(m), hipMemcpyDeviceToHost));
  if (pairs_needed) *pairs_needed = m[kMiscStickyPairs];
  if (!m[kMiscSticky]) return VKGS_OK;
  c->grow_pairs = true;
  c->need_pairs = m[kMiscStickyPairs] > c->need_pairs ? m[kMiscStickyPairs] : c->need_pairs;
  VKGS_HIP_CHECK(fill_u32_launch(c->misc.as<uint32_t>() + kMiscSticky, 0u, 2, s));
  VKGS_HIP_CHECK(hipStreamSynchronize(s));
  return VKGS_ERR_OVERFLOW;
}

int vkgs_diff_autotune(vkgs_diff_ctx* c, const vkgs_diff_scene* scenes, uint32_t n_scenes,
                       const vkgs_diff_view* views, uint32_t n_views, const vkgs_raster_config* base_cfg,
                       uint32_t steps, vkgs_raster_config* chosen, float* best_ms, void* stream) {
  if (!c || !scenes || !views || n_views == 0 || !chosen) return VKGS_ERR_INVALID_ARG;
  for (uint32_t v = 0; v < n_views; ++v)
    if (!views[v].grad_rgba) return VKGS_ERR_INVALID_ARG;
  if (steps == 0) steps = 3;
  vkgs_raster_config base;
  if (base_cfg)
    base = *base_cfg;
  else
    vkgs_raster_config_default(&base);
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  VKGS_HIP_CHECK(hipEventCreate(&e0));
  VKGS_HIP_CHECK(hipEventCreate(&e1));
  auto step = [&](const vkgs_raster_config& cfg) -> int {
    const int st = vkgs_diff_forward(c, scenes, n_scenes, views, n_views, &cfg, stream);
    return st != VKGS_OK ? st : vkgs_diff_backward(c, scenes, n_scenes, views, n_views, stream);
  };
  const TuneTimer time_one = [&](const vkgs_raster_config& cfg, float* ms) -> int {
    // a first step sizes the storage for this tile shape; repeat while its pair list overflows
    int st = VKGS_ERR_OVERFLOW;
    for (int i = 0; i < 4 && st == VKGS_ERR_OVERFLOW; ++i) {
      if ((st = step(cfg)) != VKGS_OK) return st;
      st = vkgs_diff_check(c, nullptr, stream);
    }
    if (st != VKGS_OK) return st;
    if ((st = step(cfg)) != VKGS_OK) return st;
    hipEventRecord(e0, s);
    for (uint32_t i = 0; i < steps && st == VKGS_OK; ++i) st = step(cfg);
    hipEventRecord(e1, s);
    if (st != VKGS_OK) return st;
    if ((st = vkgs_diff_check(c, nullptr, stream)) != VKGS_OK) return st;
    if (hipEventSynchronize(e1) != hipSuccess) return VKGS_ERR_HIP;
    hipEventElapsedTime(ms, e0, e1);
    *ms /= (float)steps;
    return VKGS_OK;
  };
  // tile shape (8x8 one or two tiles per wave), then the backward's LDS batch
  // depth on 8x8 tiles, then the resident-wave cap
  const TuneStage batches = [](const vkgs_raster_config& from) {
    std::vector<vkgs_raster_config> out;
    if (from.tile_w == 8 && from.tile_h == 8)
      for (uint32_t b : {8u, 16u, 32u}) {
        vkgs_raster_config cfg = from;
        cfg.lds_batch = b;
        out.push_back(cfg);
      }
    return out;
  };
  const std::vector<TuneStage> stages = {tune_shapes, batches, tune_waves};
  float t = 0.f;
  const int status = tune_configs(base, stages, 3, time_one, chosen, &t);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (status != VKGS_OK) return status;
  if (best_ms) *best_ms = t;
  return VKGS_OK;
}

int vkgs_diff_info(vkgs_diff_ctx* c, uint32_t* visible, uint64_t* pairs, void* stream) {
  if (!c || !c->misc.p) return VKGS_ERR_INVALID_ARG;
  uint32_t m[4];
  VKGS_HIP_CHECK(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
  VKGS_HIP_CHECK(hipMemcpy(m, c->misc.p, sizeof(m), hipMemcpyDeviceToHost));
  if (visible) *visible = m[kMiscVisible];
  if (pairs) *pairs = m[kMiscPairs];
  if (m[kMiscStatus]) {
    c->grow_pairs = true;
    c->need_pairs = m[kMiscPairs];
    return VKGS_ERR_OVERFLOW;
  }
  return VKGS_OK;
}

}  // extern "C"
