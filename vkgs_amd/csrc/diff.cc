// C ABI of the differentiable, batched path (vkgs_diff_*).
//
// One context holds the state of one forward (records, tile lists, per-pixel
// transmittance and processed counts) until its backward.  Views of several
// scenes render in one set of launches; see diff.h for the layout.
#include <string.h>

#include <vector>

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
  // device tables
  Buf views, scenes, block_view, tile_view;
  // per-item buffers
  uint32_t cap_items = 0;
  Buf counts, inst, key, val, ids, vid, kview, g9, misc, rec, rect, cnt, offs, scan_storage, depth_sort;
  // pairs
  uint64_t max_pairs = 0;
  Buf pkeys, pvals, pair_sort, ranges;
  // per-pixel state
  Buf final_t, last;
  // host mirrors of the last forward
  std::vector<DiffView> h_views;
  std::vector<DiffScene> h_scenes;
  std::vector<uint32_t> h_block_view, h_tile_view;
  DiffLaunch launch{};
  SortPlan depth_plan{}, pair_plan{};
  ScanPlan scan{};
  bool have_forward = false;
  bool grow_pairs = false;
};

namespace {

enum : uint32_t { kMiscVisible = 0, kMiscPairs = 1, kMiscStatus = 2 };

hipError_t upload(Buf& b, const void* src, size_t bytes, hipStream_t s) {
  hipError_t e = b.ensure(bytes);
  if (e != hipSuccess) return e;
  // the host vectors live in the context until the next forward; the copy is
  // made synchronous with respect to the host so they can be reused at once
  e = hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(s);
}

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
  Buf* bufs[] = {&c->views,   &c->scenes, &c->block_view, &c->tile_view, &c->counts,       &c->inst,
                 &c->key,     &c->val,    &c->ids,        &c->vid,       &c->kview,        &c->g9,
                 &c->misc,    &c->rec,    &c->rect,       &c->cnt,       &c->offs,         &c->scan_storage,
                 &c->depth_sort, &c->pkeys, &c->pvals,    &c->pair_sort, &c->ranges,       &c->final_t,
                 &c->last};
  for (Buf* b : bufs) b->release();
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
  for (uint32_t i = 0; i < n_scenes; ++i) {
    const vkgs_diff_scene& a = scenes[i];
    if (a.n > 0 && (!a.means || !a.log_scales || !a.quats || !a.opacity_logits || !a.sh)) return VKGS_ERR_INVALID_ARG;
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
  }
  RasterShape rs0;
  if (!make_raster_shape(16, 16, cfg, &rs0)) return VKGS_ERR_INVALID_ARG;
  c->h_views.assign(n_views, DiffView{});
  c->h_block_view.clear();
  c->h_tile_view.clear();
  uint64_t blocks = 0, tiles = 0, pixels = 0;
  for (uint32_t v = 0; v < n_views; ++v) {
    const vkgs_diff_view& a = views[v];
    if (a.scene >= n_scenes || !a.rgba) return VKGS_ERR_INVALID_ARG;
    RasterShape rs;
    if (!make_raster_shape(a.camera.screen_size[0], a.camera.screen_size[1], cfg, &rs)) return VKGS_ERR_INVALID_ARG;
    DiffView& d = c->h_views[v];
    make_frame_const(&a.camera, a.model, &d.fc);
    d.fc.tile_w = rs.tile_w;
    d.fc.tile_h = rs.tile_h;
    d.fc.tiles_x = rs.tiles_x;
    d.fc.tiles_y = rs.tiles_y;
    d.scene = a.scene;
    d.block_base = (uint32_t)blocks;
    d.tile_base = (uint32_t)tiles;
    d.out = reinterpret_cast<float4*>(a.rgba);
    d.grad_out = reinterpret_cast<const float4*>(a.grad_rgba);
    const uint32_t nb = ceil_div(c->h_scenes[a.scene].n, kCompactBlock);
    const uint32_t nt = rs.tiles_x * rs.tiles_y;
    for (uint32_t i = 0; i < nb; ++i) c->h_block_view.push_back(v);
    for (uint32_t i = 0; i < nt; ++i) c->h_tile_view.push_back(v);
    blocks += nb;
    tiles += nt;
    pixels += (uint64_t)rs.width * rs.height;
  }
  if (blocks * kCompactBlock > (1ull << 30) || tiles > (1u << 24)) return VKGS_ERR_CAPACITY;
  const uint32_t items = (uint32_t)(blocks * kCompactBlock);
  const uint32_t cap = items ? items : 1;

  // ---- device buffers ----
  hipError_t e = hipSuccess;
  if (c->max_pairs == 0 || c->grow_pairs || c->max_pairs < (uint64_t)cap) {
    const uint64_t want = (uint64_t)cap * 4 + (1u << 20);
    c->max_pairs = c->grow_pairs ? c->max_pairs * 2 : (want > c->max_pairs ? want : c->max_pairs);
    if (c->max_pairs > 0xffffffffull) c->max_pairs = 0xffffffffull;
    c->grow_pairs = false;
  }
  c->depth_plan = sort_plan(cap);
  c->pair_plan = sort_plan((uint32_t)c->max_pairs);
  c->scan = scan_plan(cap);
#define ENSURE(buf, bytes) \
  if (e == hipSuccess) e = c->buf.ensure(bytes)
  ENSURE(counts, compact_storage_bytes(cap));
  ENSURE(inst, (size_t)cap * 48);
  ENSURE(key, (size_t)cap * 4);
  ENSURE(val, (size_t)cap * 4);
  ENSURE(ids, (size_t)cap * 4);
  ENSURE(vid, (size_t)cap * 4);
  ENSURE(kview, (size_t)cap * 4);
  ENSURE(g9, (size_t)cap * 36);
  ENSURE(misc, 256);
  ENSURE(rec, (size_t)cap * 48);
  ENSURE(rect, (size_t)cap * sizeof(TileRect));
  ENSURE(cnt, (size_t)cap * 4);
  ENSURE(offs, (size_t)cap * 4);
  ENSURE(scan_storage, c->scan.bytes);
  ENSURE(depth_sort, c->depth_plan.bytes);
  ENSURE(pkeys, (size_t)c->max_pairs * 4);
  ENSURE(pvals, (size_t)c->max_pairs * 4);
  ENSURE(pair_sort, c->pair_plan.bytes);
  ENSURE(ranges, (size_t)(tiles ? tiles : 1) * 8);
  ENSURE(final_t, (size_t)(pixels ? pixels : 1) * 4);
  ENSURE(last, (size_t)(pixels ? pixels : 1) * 4);
#undef ENSURE
  if (e != hipSuccess) return VKGS_ERR_HIP;
  uint64_t pix = 0;
  for (uint32_t v = 0; v < n_views; ++v) {
    DiffView& d = c->h_views[v];
    d.final_t = c->final_t.as<float>() + pix;
    d.last = c->last.as<uint32_t>() + pix;
    pix += (uint64_t)d.fc.width * d.fc.height;
  }
  VKGS_HIP_CHECK(upload(c->views, c->h_views.data(), sizeof(DiffView) * n_views, s));
  VKGS_HIP_CHECK(upload(c->scenes, c->h_scenes.data(), sizeof(DiffScene) * n_scenes, s));
  VKGS_HIP_CHECK(upload(c->block_view, c->h_block_view.data(), 4 * (c->h_block_view.size() + 1), s));
  VKGS_HIP_CHECK(upload(c->tile_view, c->h_tile_view.data(), 4 * (c->h_tile_view.size() + 1), s));

  DiffLaunch& L = c->launch;
  L.views = c->views.as<DiffView>();
  L.scenes = c->scenes.as<DiffScene>();
  L.block_view = c->block_view.as<uint32_t>();
  L.tile_view = c->tile_view.as<uint32_t>();
  L.n_views = n_views;
  L.n_blocks = (uint32_t)blocks;
  L.n_tiles = (uint32_t)tiles;
  L.tile_w = rs0.tile_w;
  L.tile_h = rs0.tile_h;
  L.tile_shift_x = rs0.tile_shift_x;
  L.tile_shift_y = rs0.tile_shift_y;
  L.t_min = cfg->t_min;
  c->cap_items = cap;

  // ---- forward ----
  uint32_t* misc = c->misc.as<uint32_t>();
  VKGS_HIP_CHECK(fill_u32_launch(misc, 0u, 4, s));
  VKGS_HIP_CHECK(launch_diff_count(L, c->counts.as<uint32_t>(), misc + kMiscVisible, s));
  VKGS_HIP_CHECK(launch_diff_write(L, c->counts.as<uint32_t>(), c->inst.as<float4>(), c->key.as<uint32_t>(),
                                   c->val.as<uint32_t>(), c->ids.as<uint32_t>(), c->vid.as<uint32_t>(), s));
  VKGS_HIP_CHECK(sort_kv_launch(c->depth_plan, misc + kMiscVisible, c->key.as<uint32_t>(), c->val.as<uint32_t>(), 30u,
                                c->depth_sort.p, s));
  VKGS_HIP_CHECK(launch_diff_bin_count(L, c->inst.as<float4>(), c->val.as<uint32_t>(), c->vid.as<uint32_t>(),
                                       misc + kMiscVisible, cap, c->rec.as<float4>(), c->rect.as<TileRect>(),
                                       c->cnt.as<uint32_t>(), c->kview.as<uint32_t>(), s));
  VKGS_HIP_CHECK(scan_exclusive_launch(c->scan, c->cnt.as<uint32_t>(), c->offs.as<uint32_t>(), misc + kMiscVisible,
                                       misc + kMiscPairs, c->scan_storage.p, s));
  VKGS_HIP_CHECK(launch_diff_bin_emit(L, c->rect.as<TileRect>(), c->cnt.as<uint32_t>(), c->offs.as<uint32_t>(),
                                      c->kview.as<uint32_t>(), misc + kMiscVisible, cap, c->max_pairs,
                                      c->pkeys.as<uint32_t>(), c->pvals.as<uint32_t>(), misc + kMiscStatus, s));
  VKGS_HIP_CHECK(sort_kv_launch(c->pair_plan, misc + kMiscPairs, c->pkeys.as<uint32_t>(), c->pvals.as<uint32_t>(),
                                (uint32_t)bits_for((uint32_t)tiles), c->pair_sort.p, s));
  VKGS_HIP_CHECK(launch_ranges(c->pkeys.as<uint32_t>(), misc + kMiscPairs, c->max_pairs, (uint32_t)tiles,
                               c->ranges.as<uint32_t>(), s));
  VKGS_HIP_CHECK(launch_diff_raster(L, c->rec.as<float4>(), c->pvals.as<uint32_t>(), c->ranges.as<uint32_t>(), true, s));
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
    d.g_means = a.grad_means;
    d.g_log_scales = a.grad_log_scales;
    d.g_quats = a.grad_quats;
    d.g_opacity = a.grad_opacity_logits;
    d.g_sh = a.grad_sh;
  }
  VKGS_HIP_CHECK(upload(c->views, c->h_views.data(), sizeof(DiffView) * n_views, s));
  VKGS_HIP_CHECK(upload(c->scenes, c->h_scenes.data(), sizeof(DiffScene) * n_scenes, s));
  const DiffLaunch& L = c->launch;
  uint32_t* misc = c->misc.as<uint32_t>();
  VKGS_HIP_CHECK(fill_u32_launch(c->g9.as<uint32_t>(), 0u, (size_t)c->cap_items * 9, s));
  VKGS_HIP_CHECK(launch_diff_raster_backward(L, c->rec.as<float4>(), c->pvals.as<uint32_t>(), c->ranges.as<uint32_t>(),
                                             c->g9.as<float>(), s));
  VKGS_HIP_CHECK(launch_diff_project_backward(L, c->val.as<uint32_t>(), c->ids.as<uint32_t>(), c->vid.as<uint32_t>(),
                                              c->g9.as<float>(), misc + kMiscVisible, c->cap_items, s));
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
    return VKGS_ERR_OVERFLOW;
  }
  return VKGS_OK;
}

}  // extern "C"
