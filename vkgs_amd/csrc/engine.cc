// vkgs_renderer: the Engine analogue (Engine::Impl, engine.cc:101-1682).
//
// Owns the splat storage (preallocated to a capacity like engine.cc:483-502)
// and every per-frame buffer, and records one frame as
//   fused cull+project+compact -> depth sort -> bin -> tile sort -> ranges -> raster
// on a stream.  The recording is captured once per (N, width, height, raster
// config, output pointer) into a hipGraph and replayed; the camera and model
// reach the kernels through a device-resident FrameConst written by a
// one-wave kernel launched before each replay (it also zeroes the frame's
// counters), so no host sync happens in a frame.  When the tile grid allows
// (rect_bits), the tile rectangle travels through the depth sort packed into
// the value's high word instead of being gathered in depth order.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "autotune.h"
#include "common.h"
#include "kernels.h"
#include "sort.h"

using namespace vkgs;

namespace {

// misc words; the frame-constant launch zeroes the first four every frame
// (kMiscKeyMax: the frame's largest depth key, found by the depth sort)
enum : uint32_t { kMiscVisible = 0, kMiscPairs = 1, kMiscStatus = 2, kMiscKeyMax = 3, kMiscWords = 64 };
constexpr int kEvents = 7;
constexpr int kStages = 8;

struct DevBuf {
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

// Background PLY loader (the SplatLoadThread of splat_load_thread.cc:33-204
// and Engine::LoadSplatsAsync, engine.cc:541-549): a thread reads the file in
// 65536-row chunks into pinned host memory, checking for cancellation between
// chunks, copies the ply buffer to the device on its own stream and publishes
// it; the next render parses it into the splat storage on the render stream,
// so the swap is ordered with the frames around it.  A canceled or failed load
// publishes nothing and the renderer keeps its splats.
struct Loader {
  std::thread thread;
  std::mutex mu;
  std::condition_variable cv;
  std::string pending;
  bool has_pending = false;
  bool terminate = false;
  std::atomic<bool> cancel{false};
  hipStream_t stream = nullptr;
  // progress (under mu)
  uint64_t total = 0, loaded = 0;
  int state = 0;  // vkgs_load_state
  int error = 0;
  // the published buffer (under mu): device ply buffer of ready_n rows
  void* ready = nullptr;
  uint32_t ready_n = 0;
};

}  // namespace

struct vkgs_renderer {
  uint32_t capacity = 0;
  uint32_t n = 0;
  DevBuf position, cov3d, opacity, sh;

  // frame buffers sized by n
  uint32_t frame_n = 0;
  DevBuf fc, misc, counts, rc, key, val, val64, ids, sort_storage;
  // tile rectangles packed into the depth sort's values (when they fit)
  bool packed = false;
  RectBits rb{};
  // binning sized by (n, pairs, tiles)
  uint64_t max_pairs = 0;
  DevBuf rec, rect, cnt, offs, scan_storage, pkeys, pvals, psort, ranges, tile_keys, tile_order, tile_sort;
  SortPlan tile_plan{};
  DevBuf sorted_ids, sorted_keys, list_ids;
  DevBuf work;  // list entries composited per raster wave (timed frames)
  SortPlan depth_plan{}, pair_plan{};
  ScanPlan scan{};
  RasterShape rs{};
  vkgs_raster_config cfg{};

  hipStream_t capture_stream = nullptr;
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  struct Key {
    uint32_t n, w, h, tw, th;
    float t_min;
    uint64_t pairs;
    void* out;
  } graph_key{};
  bool graph_valid = false;
  bool grow_pairs = false;
  uint64_t need_pairs = 0;  // pair count of the last overflowing frame

  bool timing = false;
  hipEvent_t ev[kEvents] = {};
  float stage_ms[kStages] = {};

  Loader loader;
  void* ply_in_use = nullptr;  // the last parsed async ply buffer (freed when replaced)
};

namespace {

void drop_graph(vkgs_renderer* r) {
  if (r->exec) hipGraphExecDestroy(r->exec);
  if (r->graph) hipGraphDestroy(r->graph);
  r->exec = nullptr;
  r->graph = nullptr;
  r->graph_valid = false;
}

hipError_t ensure_frame(vkgs_renderer* r, uint32_t width, uint32_t height) {
  hipError_t e;
  const uint32_t n = r->n;
  RasterShape rs;
  if (!make_raster_shape(width, height, &r->cfg, &rs)) return hipErrorInvalidValue;
  // at least 4 pairs per splat (more after a frame overflowed), also when a load made n larger
  const uint64_t base_pairs = (uint64_t)(n ? n : 1) * 4 + (1u << 20);
  if (r->max_pairs < base_pairs) r->max_pairs = base_pairs;
  if (r->grow_pairs) {  // room for the overflowing frame's whole list, plus an eighth
    const uint64_t need = r->need_pairs + r->need_pairs / 8 + (1u << 20);
    r->max_pairs = need > r->max_pairs * 2 ? need : r->max_pairs * 2;
    r->grow_pairs = false;
  }
  if (r->max_pairs > 0xffffffffull) r->max_pairs = 0xffffffffull;
  r->rs = rs;
  const uint32_t cap = n ? n : 1;
  r->packed = rect_bits(rs.tiles_x, rs.tiles_y, &r->rb);
  r->depth_plan = sort_plan(cap, r->packed ? 8u : 4u);
  r->pair_plan = sort_plan((uint32_t)r->max_pairs);
  r->scan = scan_plan(cap);
  const uint32_t tiles = rs.tiles_x * rs.tiles_y;
  if ((e = r->fc.ensure(sizeof(FrameConst))) != hipSuccess) return e;
  if ((e = r->misc.ensure(kMiscWords * 4)) != hipSuccess) return e;
  if ((e = r->counts.ensure(compact_storage_bytes(cap))) != hipSuccess) return e;
  if (r->packed) {
    if ((e = r->val64.ensure((size_t)cap * 8)) != hipSuccess) return e;
  } else {
    if ((e = r->rc.ensure((size_t)cap * 16)) != hipSuccess) return e;
    if ((e = r->val.ensure((size_t)cap * 4)) != hipSuccess) return e;
  }
  if ((e = r->key.ensure((size_t)cap * 4)) != hipSuccess) return e;
  if ((e = r->ids.ensure((size_t)cap * 4)) != hipSuccess) return e;
  if ((e = r->sort_storage.ensure(r->depth_plan.bytes)) != hipSuccess) return e;
  if ((e = r->rec.ensure((size_t)cap * 48)) != hipSuccess) return e;
  if (!r->packed) {  // depth-ordered rectangles and counts of the gather path
    if ((e = r->rect.ensure((size_t)cap * sizeof(TileRect))) != hipSuccess) return e;
    if ((e = r->cnt.ensure((size_t)cap * 4)) != hipSuccess) return e;
  }
  // offs: the gather path's scanned counts, or the packed path's emit group sums
  const size_t offs_bytes = std::max((size_t)cap * 4, bin_grouped_storage_bytes(cap));
  if ((e = r->offs.ensure(offs_bytes)) != hipSuccess) return e;
  if ((e = r->scan_storage.ensure(r->scan.bytes)) != hipSuccess) return e;
  if ((e = r->pkeys.ensure((size_t)r->max_pairs * 4)) != hipSuccess) return e;
  if ((e = r->pvals.ensure((size_t)r->max_pairs * 4)) != hipSuccess) return e;
  if ((e = r->psort.ensure(r->pair_plan.bytes)) != hipSuccess) return e;
  if ((e = r->ranges.ensure((size_t)tiles * 8)) != hipSuccess) return e;
  r->tile_plan = sort_plan(tiles);
  if ((e = r->tile_keys.ensure((size_t)tiles * 4)) != hipSuccess) return e;
  if ((e = r->tile_order.ensure((size_t)tiles * 4)) != hipSuccess) return e;
  if ((e = r->tile_sort.ensure(r->tile_plan.bytes)) != hipSuccess) return e;
  if ((e = r->work.ensure((size_t)tiles * (rs.tile_w * rs.tile_h / kWave) * 4)) != hipSuccess) return e;
  r->frame_n = n;
  return hipSuccess;
}

DrawBuffers draw_buffers(vkgs_renderer* r) {
  DrawBuffers b;
  uint32_t* misc = r->misc.as<uint32_t>();
  b.rec = r->rec.as<float4>();
  b.rect = r->rect.as<TileRect>();
  b.cnt = r->cnt.as<uint32_t>();
  b.offs = r->offs.as<uint32_t>();
  b.scan = r->scan;
  b.scan_storage = r->scan_storage.p;
  b.keys = r->pkeys.as<uint32_t>();
  b.vals = r->pvals.as<uint32_t>();
  b.sort = r->pair_plan;
  b.sort_storage = r->psort.p;
  b.ranges = r->ranges.as<uint32_t>();
  b.status = misc + kMiscStatus;
  b.status_cleared = true;  // the frame-constant launch zeroes the misc words
  b.pairs = misc + kMiscPairs;
  b.tile_keys = r->tile_keys.as<uint32_t>();
  b.tile_order = r->tile_order.as<uint32_t>();
  b.tile_sort = r->tile_plan;
  b.tile_sort_storage = r->tile_sort.p;
  b.work = r->timing ? r->work.as<uint32_t>() : nullptr;
  return b;
}

void loader_main(vkgs_renderer* r) {
  Loader& L = r->loader;
  for (;;) {
    std::string path;
    {
      std::unique_lock<std::mutex> g(L.mu);
      L.cv.wait(g, [&] { return L.terminate || L.has_pending; });
      if (L.terminate) return;
      path = L.pending;
      L.has_pending = false;
      L.cancel = false;
      L.total = L.loaded = 0;
      L.state = VKGS_LOAD_READING;
      L.error = 0;
    }
    // status: a vkgs_status error, or kCanceled
    constexpr int kCanceled = -1;
    auto fail = [&](int status) {
      std::lock_guard<std::mutex> g(L.mu);
      if (L.has_pending) return;  // superseded by a newer request, which the loop takes next
      L.state = status == kCanceled ? VKGS_LOAD_CANCELED : VKGS_LOAD_FAILED;
      L.error = status == kCanceled ? 0 : status;
    };
    vkgs_ply_info info;
    int st = vkgs_ply_read_header(path.c_str(), &info);
    if (st != VKGS_OK) {
      fail(st);
      continue;
    }
    if (info.point_count > r->capacity) {
      fail(VKGS_ERR_CAPACITY);
      continue;
    }
    const uint32_t n = (uint32_t)info.point_count;
    {
      std::lock_guard<std::mutex> g(L.mu);
      L.total = n;
    }
    const size_t row_bytes = (size_t)info.stride * 4, bytes = 240 + (size_t)n * row_bytes;
    void* host = nullptr;
    if (hipHostMalloc(&host, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) {
      fail(VKGS_ERR_HIP);
      continue;
    }
    memcpy(host, info.offsets, 240);
    FILE* f = fopen(path.c_str(), "rb");
    bool ok = f != nullptr && fseek(f, (long)info.header_bytes, SEEK_SET) == 0;
    constexpr uint32_t kChunk = 65536;  // rows per read, as splat_load_thread.cc:142
    for (uint32_t start = 0; ok && start < n; start += kChunk) {
      if (L.cancel || L.terminate) break;
      const uint32_t rows = n - start < kChunk ? n - start : kChunk;
      ok = fread(static_cast<char*>(host) + 240 + (size_t)start * row_bytes, row_bytes, rows, f) == rows;
      std::lock_guard<std::mutex> g(L.mu);
      L.loaded = start + rows;
    }
    if (f) fclose(f);
    if (L.cancel || L.terminate || !ok) {
      (void)hipHostFree(host);
      fail(!ok ? VKGS_ERR_IO : kCanceled);
      if (L.terminate) return;
      continue;
    }
    void* dev = nullptr;
    hipError_t e = hipMalloc(&dev, bytes);
    if (e == hipSuccess) e = hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, L.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(L.stream);
    (void)hipHostFree(host);
    if (e != hipSuccess || L.cancel) {
      if (dev) (void)hipFree(dev);
      fail(e != hipSuccess ? VKGS_ERR_HIP : kCanceled);
      continue;
    }
    std::lock_guard<std::mutex> g(L.mu);
    if (L.has_pending) {  // superseded while uploading
      (void)hipFree(dev);
      continue;
    }
    if (L.ready) (void)hipFree(L.ready);  // an earlier load nobody rendered yet
    L.ready = dev;
    L.ready_n = n;
    L.state = VKGS_LOAD_READY;
  }
}

// A synchronous load supersedes any asynchronous one: cancel it and drop a
// published buffer nobody took.
void supersede_async_load(vkgs_renderer* r) {
  Loader& L = r->loader;
  std::lock_guard<std::mutex> g(L.mu);
  L.cancel = true;
  L.has_pending = false;
  if (L.ready) {
    (void)hipFree(L.ready);
    L.ready = nullptr;
  }
  if (L.state == VKGS_LOAD_READING || L.state == VKGS_LOAD_READY) L.state = VKGS_LOAD_CANCELED;
}

// Takes a published async load: parses it into the splat storage on stream s.
hipError_t take_async_load(vkgs_renderer* r, hipStream_t s) {
  Loader& L = r->loader;
  void* dev = nullptr;
  uint32_t n = 0;
  {
    std::lock_guard<std::mutex> g(L.mu);
    if (!L.ready) return hipSuccess;
    dev = L.ready;
    n = L.ready_n;
    L.ready = nullptr;
    L.state = VKGS_LOAD_DONE;
  }
  if (r->ply_in_use) (void)hipFree(r->ply_in_use);  // parsed by an earlier frame
  r->ply_in_use = dev;
  hipError_t e = launch_parse_ply(dev, n, r->position.as<float>(), r->cov3d.as<float>(), r->opacity.as<float>(),
                                  r->sh.as<uint16_t>(), s);
  r->n = n;
  drop_graph(r);
  return e;
}

// Records the frame on stream s (eager, or under capture).
hipError_t record_frame(vkgs_renderer* r, float* out, hipStream_t s, bool timed) {
  hipError_t e;
  const uint32_t n = r->n;
  uint32_t* misc = r->misc.as<uint32_t>();
  if (timed) hipEventRecord(r->ev[0], s);
  if ((e = launch_fused_count(r->fc.as<FrameConst>(), r->position.as<float>(), n, misc + kMiscVisible,
                              r->counts.as<uint32_t>(), s)) != hipSuccess)
    return e;
  if (timed) hipEventRecord(r->ev[1], s);
  if (r->packed)
    e = launch_fused_write_packed(r->fc.as<FrameConst>(), r->position.as<float>(), r->cov3d.as<float>(),
                                  r->opacity.as<float>(), r->sh.as<uint16_t>(), n, r->rec.as<float4>(),
                                  r->key.as<uint32_t>(), r->val64.as<uint64_t>(), nullptr,
                                  r->counts.as<uint32_t>(), misc + kMiscVisible, s);
  else
    e = launch_fused_write(r->fc.as<FrameConst>(), r->position.as<float>(), r->cov3d.as<float>(),
                           r->opacity.as<float>(), r->sh.as<uint16_t>(), n, r->rec.as<float4>(), r->rc.as<uint4>(),
                           r->key.as<uint32_t>(), r->val.as<uint32_t>(), nullptr,
                           r->counts.as<uint32_t>(), misc + kMiscVisible, s);
  if (e != hipSuccess) return e;
  if (timed) hipEventRecord(r->ev[2], s);
  // keys are k = fl(1 - z) * 2^24 <= 2^24: 25 bits; the sort's first pass
  // finds the frame's largest k and the passes above its bits return at once
  if (r->packed)
    e = sort_kv64_launch(r->depth_plan, misc + kMiscVisible, r->key.as<uint32_t>(), r->val64.as<uint64_t>(),
                         kDepthRankBits, r->sort_storage.p, s, 0, misc + kMiscKeyMax, kDepthRankFirstBits);
  else
    e = sort_kv_launch(r->depth_plan, misc + kMiscVisible, r->key.as<uint32_t>(), r->val.as<uint32_t>(),
                       kDepthRankBits, r->sort_storage.p, s, 0, misc + kMiscKeyMax, kDepthRankFirstBits);
  if (e != hipSuccess) return e;
  if (timed) hipEventRecord(r->ev[3], s);
  hipEvent_t marks[2] = {r->ev[4], r->ev[5]};
  if (r->packed)
    e = run_draw_packed(r->val64.as<uint64_t>(), misc + kMiscVisible, n ? n : 1, r->rs, r->rb, r->max_pairs,
                        draw_buffers(r), reinterpret_cast<float4*>(out), s, timed ? marks : nullptr);
  else
    e = run_draw_slots(r->rc.as<uint4>(), r->val.as<uint32_t>(), misc + kMiscVisible, n ? n : 1, r->rs, r->max_pairs,
                       draw_buffers(r), reinterpret_cast<float4*>(out), s, timed ? marks : nullptr);
  if (e != hipSuccess) return e;
  if (timed) hipEventRecord(r->ev[6], s);
  return hipSuccess;
}

}  // namespace

extern "C" {

int vkgs_renderer_create(uint32_t capacity, vkgs_renderer** out) {
  if (!out || capacity == 0 || capacity > (1u << 28)) return VKGS_ERR_INVALID_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return VKGS_ERR_NO_DEVICE;
  vkgs_renderer* r = new vkgs_renderer();
  r->capacity = capacity;
  vkgs_raster_config_default(&r->cfg);
  hipError_t e = hipSuccess;
  if (e == hipSuccess) e = r->position.ensure((size_t)capacity * 12);
  if (e == hipSuccess) e = r->cov3d.ensure((size_t)capacity * 24);
  if (e == hipSuccess) e = r->opacity.ensure((size_t)capacity * 4);
  if (e == hipSuccess) e = r->sh.ensure((size_t)capacity * 96);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&r->capture_stream, hipStreamNonBlocking);
  for (int i = 0; i < kEvents && e == hipSuccess; ++i) e = hipEventCreate(&r->ev[i]);
  if (e != hipSuccess) {
    vkgs_renderer_destroy(r);
    return VKGS_ERR_HIP;
  }
  *out = r;
  return VKGS_OK;
}

int vkgs_renderer_destroy(vkgs_renderer* r) {
  if (!r) return VKGS_ERR_INVALID_ARG;
  if (r->loader.thread.joinable()) {
    {
      std::lock_guard<std::mutex> g(r->loader.mu);
      r->loader.terminate = true;
      r->loader.cancel = true;
    }
    r->loader.cv.notify_one();
    r->loader.thread.join();
  }
  (void)hipDeviceSynchronize();
  if (r->loader.ready) hipFree(r->loader.ready);
  if (r->ply_in_use) hipFree(r->ply_in_use);
  if (r->loader.stream) hipStreamDestroy(r->loader.stream);
  drop_graph(r);
  DevBuf* bufs[] = {&r->work,       &r->position,   &r->cov3d,    &r->opacity,    &r->sh,        &r->fc,
                    &r->misc,
                    &r->counts,     &r->rc,         &r->key,      &r->val,        &r->val64,     &r->ids,
                    &r->sort_storage, &r->rec,      &r->rect,     &r->cnt,        &r->offs,      &r->scan_storage,
                    &r->pkeys,      &r->pvals,      &r->psort,    &r->ranges,     &r->tile_keys, &r->tile_order,
                    &r->tile_sort,  &r->sorted_ids, &r->sorted_keys, &r->list_ids};
  for (DevBuf* b : bufs) b->release();
  if (r->capture_stream) hipStreamDestroy(r->capture_stream);
  for (int i = 0; i < kEvents; ++i)
    if (r->ev[i]) hipEventDestroy(r->ev[i]);
  delete r;
  return VKGS_OK;
}

int vkgs_renderer_load_ply_buffer(vkgs_renderer* r, const void* host_ply, uint32_t n) {
  if (!r || (!host_ply && n)) return VKGS_ERR_INVALID_ARG;
  if (n > r->capacity) return VKGS_ERR_CAPACITY;
  supersede_async_load(r);
  if (n == 0) {
    r->n = 0;
    drop_graph(r);
    return VKGS_OK;
  }
  const uint32_t* offsets = static_cast<const uint32_t*>(host_ply);
  const uint32_t stride = offsets[59];
  if (stride == 0) return VKGS_ERR_FORMAT;
  for (int i = 0; i < 59; ++i)
    if (offsets[i] >= stride) return VKGS_ERR_FORMAT;
  const size_t bytes = 240 + (size_t)n * stride * 4;
  void* d = nullptr;
  VKGS_HIP_CHECK(hipMalloc(&d, bytes));
  hipError_t e = hipMemcpy(d, host_ply, bytes, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = launch_parse_ply(d, n, r->position.as<float>(), r->cov3d.as<float>(), r->opacity.as<float>(),
                         r->sh.as<uint16_t>(), nullptr);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  hipFree(d);
  if (e != hipSuccess) return VKGS_ERR_HIP;
  r->n = n;
  drop_graph(r);
  return VKGS_OK;
}

int vkgs_renderer_load_ply_async(vkgs_renderer* r, const char* path) {
  if (!r || !path) return VKGS_ERR_INVALID_ARG;
  Loader& L = r->loader;
  if (!L.thread.joinable()) {
    VKGS_HIP_CHECK(hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking));
    L.thread = std::thread(loader_main, r);
  }
  {
    std::lock_guard<std::mutex> g(L.mu);
    L.cancel = true;  // a load in progress stops at its next chunk (Engine::LoadSplats, engine.cc:541-544)
    L.pending = path;
    L.has_pending = true;
    L.state = VKGS_LOAD_READING;
    L.total = L.loaded = 0;
  }
  L.cv.notify_one();
  return VKGS_OK;
}

int vkgs_renderer_cancel_load(vkgs_renderer* r) {
  if (!r) return VKGS_ERR_INVALID_ARG;
  Loader& L = r->loader;
  std::lock_guard<std::mutex> g(L.mu);
  L.cancel = true;
  if (L.has_pending) {
    L.has_pending = false;
    L.state = VKGS_LOAD_CANCELED;
  }
  return VKGS_OK;
}

int vkgs_renderer_load_progress(vkgs_renderer* r, uint64_t* total, uint64_t* loaded, int* state) {
  if (!r) return VKGS_ERR_INVALID_ARG;
  Loader& L = r->loader;
  std::lock_guard<std::mutex> g(L.mu);
  if (total) *total = L.total;
  if (loaded) *loaded = L.loaded;
  if (state) *state = L.state;
  return L.state == VKGS_LOAD_FAILED ? L.error : VKGS_OK;
}

int vkgs_renderer_load_ply(vkgs_renderer* r, const char* path) {
  if (!r || !path) return VKGS_ERR_INVALID_ARG;
  vkgs_ply_info info;
  int st = vkgs_ply_read_header(path, &info);
  if (st != VKGS_OK) return st;
  if (info.point_count > r->capacity) return VKGS_ERR_CAPACITY;
  const uint64_t bytes = vkgs_ply_buffer_bytes(&info);
  std::vector<char> host(bytes);
  st = vkgs_ply_read(path, &info, host.data(), bytes);
  if (st != VKGS_OK) return st;
  return vkgs_renderer_load_ply_buffer(r, host.data(), (uint32_t)info.point_count);
}

uint32_t vkgs_renderer_splat_count(const vkgs_renderer* r) { return r ? r->n : 0u; }

int vkgs_renderer_storage(vkgs_renderer* r, float** d_position, float** d_cov3d, float** d_opacity,
                          uint16_t** d_sh) {
  if (!r) return VKGS_ERR_INVALID_ARG;
  if (d_position) *d_position = r->position.as<float>();
  if (d_cov3d) *d_cov3d = r->cov3d.as<float>();
  if (d_opacity) *d_opacity = r->opacity.as<float>();
  if (d_sh) *d_sh = r->sh.as<uint16_t>();
  return VKGS_OK;
}

int vkgs_renderer_set_config(vkgs_renderer* r, const vkgs_raster_config* cfg) {
  if (!r || !cfg) return VKGS_ERR_INVALID_ARG;
  RasterShape rs;
  if (!make_raster_shape(16, 16, cfg, &rs)) return VKGS_ERR_INVALID_ARG;
  r->cfg = *cfg;
  drop_graph(r);
  return VKGS_OK;
}

int vkgs_renderer_render(vkgs_renderer* r, const vkgs_camera* camera, const float model[16], float* d_rgba,
                         void* stream) {
  if (!r || !camera || !d_rgba) return VKGS_ERR_INVALID_ARG;
  const uint32_t width = camera->screen_size[0], height = camera->screen_size[1];
  RasterShape rs;
  if (!make_raster_shape(width, height, &r->cfg, &rs)) return VKGS_ERR_INVALID_ARG;
  hipStream_t s = static_cast<hipStream_t>(stream);
  VKGS_HIP_CHECK(take_async_load(r, s));
  const vkgs_renderer::Key key = {r->n, width, height, r->cfg.tile_w, r->cfg.tile_h, r->cfg.t_min,
                                  r->max_pairs, d_rgba};
  const bool same = r->graph_valid && memcmp(&key, &r->graph_key, sizeof(key)) == 0 && !r->grow_pairs;
  if (!same || r->timing) {
    drop_graph(r);
    VKGS_HIP_CHECK(hipStreamSynchronize(s));
    VKGS_HIP_CHECK(ensure_frame(r, width, height));
  }
  // per-frame constants: passed by value to a one-wave kernel, so the host
  // never waits for the previous frame before writing them
  FrameConst fc;
  make_frame_const(camera, model, &fc);
  fc.triangles = r->rs.triangles;
  fc.tile_w = r->rs.tile_w;
  fc.tile_h = r->rs.tile_h;
  fc.tiles_x = r->rs.tiles_x;
  fc.tiles_y = r->rs.tiles_y;
  // ... which also zeroes the frame's counters (visible, pairs, status)
  VKGS_HIP_CHECK(launch_set_frame_const(r->fc.as<FrameConst>(), fc, s, r->misc.as<uint32_t>(), 4));
  if (r->timing) {
    VKGS_HIP_CHECK(record_frame(r, d_rgba, s, true));
    VKGS_HIP_CHECK(hipEventSynchronize(r->ev[6]));
    float t[kStages] = {0};
    hipEventElapsedTime(&t[0], r->ev[0], r->ev[1]);  // visible counts per group of workgroups
    hipEventElapsedTime(&t[1], r->ev[1], r->ev[2]);  // project + compact (one kernel)
    hipEventElapsedTime(&t[2], r->ev[2], r->ev[3]);  // depth sort
    hipEventElapsedTime(&t[3], r->ev[3], r->ev[4]);  // bin: count, scan, emit
    hipEventElapsedTime(&t[4], r->ev[4], r->ev[5]);  // tile sort + ranges
    hipEventElapsedTime(&t[5], r->ev[5], r->ev[6]);  // raster (one kernel)
    hipEventElapsedTime(&t[6], r->ev[0], r->ev[6]);  // total
    memcpy(r->stage_ms, t, sizeof(t));
    return VKGS_OK;
  }
  if (!r->graph_valid) {
    VKGS_HIP_CHECK(hipStreamBeginCapture(r->capture_stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = record_frame(r, d_rgba, r->capture_stream, false);
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(r->capture_stream, &g);
    if (e != hipSuccess || e2 != hipSuccess) {
      if (g) hipGraphDestroy(g);
      return VKGS_ERR_HIP;
    }
    r->graph = g;
    VKGS_HIP_CHECK(hipGraphInstantiate(&r->exec, r->graph, nullptr, nullptr, 0));
    r->graph_key = key;
    r->graph_valid = true;
  }
  VKGS_HIP_CHECK(hipGraphLaunch(r->exec, s));
  return VKGS_OK;
}

int vkgs_renderer_frame_info(vkgs_renderer* r, uint32_t* visible, uint64_t* pairs, void* stream) {
  if (!r || !r->misc.p) return VKGS_ERR_INVALID_ARG;
  uint32_t m[4];
  VKGS_HIP_CHECK(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
  VKGS_HIP_CHECK(hipMemcpy(m, r->misc.p, sizeof(m), hipMemcpyDeviceToHost));
  if (visible) *visible = m[kMiscVisible];
  if (pairs) *pairs = m[kMiscPairs];
  if (m[kMiscStatus]) {
    r->grow_pairs = true;  // the next render reallocates with room for the whole list
    r->need_pairs = m[kMiscPairs];
    return VKGS_ERR_OVERFLOW;
  }
  return VKGS_OK;
}

int vkgs_renderer_sorted(vkgs_renderer* r, const uint32_t** d_sorted_ids, const uint32_t** d_keys) {
  if (!r || !r->misc.p || !d_sorted_ids) return VKGS_ERR_INVALID_ARG;
  const uint32_t cap = r->frame_n ? r->frame_n : 1;
  VKGS_HIP_CHECK(r->sorted_ids.ensure((size_t)cap * 4));
  uint32_t* misc = r->misc.as<uint32_t>();
  // the frame writes no slot -> splat id table; rebuild it from its offsets
  VKGS_HIP_CHECK(launch_compact_ids(r->fc.as<FrameConst>(), r->position.as<float>(), r->frame_n,
                                    r->counts.as<uint32_t>(), r->ids.as<uint32_t>(), nullptr));
  // sorted slots: the depth sort's values (low words of the packed values)
  if (r->packed)
    VKGS_HIP_CHECK(launch_gather_u32(r->ids.as<uint32_t>(), r->val64.as<uint32_t>(), nullptr, misc + kMiscVisible,
                                     cap, r->sorted_ids.as<uint32_t>(), nullptr, 2));
  else
    VKGS_HIP_CHECK(launch_gather_u32(r->ids.as<uint32_t>(), r->val.as<uint32_t>(), nullptr, misc + kMiscVisible,
                                     cap, r->sorted_ids.as<uint32_t>(), nullptr));
  // the sort ran on the integer keys fl(1 - z) * 2^24: hand back rank.comp:39's float bits
  if (d_keys) {
    VKGS_HIP_CHECK(r->sorted_keys.ensure((size_t)cap * 4));
    VKGS_HIP_CHECK(launch_depth_key_bits(r->key.as<uint32_t>(), misc + kMiscVisible, cap,
                                         r->sorted_keys.as<uint32_t>(), nullptr));
  }
  VKGS_HIP_CHECK(hipDeviceSynchronize());
  *d_sorted_ids = r->sorted_ids.as<uint32_t>();
  if (d_keys) *d_keys = r->sorted_keys.as<uint32_t>();
  return VKGS_OK;
}

int vkgs_renderer_tile_lists(vkgs_renderer* r, const uint32_t** d_ranges, const uint32_t** d_list_ids,
                             uint32_t* tiles_x, uint32_t* tiles_y) {
  if (!r || !r->misc.p || !d_ranges || !d_list_ids) return VKGS_ERR_INVALID_ARG;
  const uint64_t cap = r->max_pairs ? r->max_pairs : 1;
  VKGS_HIP_CHECK(r->list_ids.ensure((size_t)cap * 4));
  uint32_t* misc = r->misc.as<uint32_t>();
  // pairs hold slots; slot -> splat id (the table rebuilt from the frame's offsets)
  VKGS_HIP_CHECK(launch_compact_ids(r->fc.as<FrameConst>(), r->position.as<float>(), r->frame_n,
                                    r->counts.as<uint32_t>(), r->ids.as<uint32_t>(), nullptr));
  VKGS_HIP_CHECK(launch_gather_u32(r->ids.as<uint32_t>(), r->pvals.as<uint32_t>(), nullptr, misc + kMiscPairs,
                                   (uint32_t)cap, r->list_ids.as<uint32_t>(), nullptr));
  VKGS_HIP_CHECK(hipDeviceSynchronize());
  *d_ranges = r->ranges.as<uint32_t>();
  *d_list_ids = r->list_ids.as<uint32_t>();
  if (tiles_x) *tiles_x = r->rs.tiles_x;
  if (tiles_y) *tiles_y = r->rs.tiles_y;
  return VKGS_OK;
}

int vkgs_renderer_autotune(vkgs_renderer* r, const vkgs_camera* camera, const float model[16], float* d_rgba,
                           uint32_t frames, vkgs_raster_config* chosen, float* best_ms, void* stream) {
  if (!r || !camera || !d_rgba) return VKGS_ERR_INVALID_ARG;
  if (frames == 0) frames = 10;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const vkgs_raster_config base = r->cfg;
  const bool timing = r->timing;
  r->timing = false;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  VKGS_HIP_CHECK(hipEventCreate(&e0));
  VKGS_HIP_CHECK(hipEventCreate(&e1));
  const TuneTimer time_one = [&](const vkgs_raster_config& c, float* ms) -> int {
    int st = vkgs_renderer_set_config(r, &c);
    if (st != VKGS_OK) return st;
    // the first frame sizes the pair storage; repeat while it overflows
    for (int i = 0; i < 4; ++i) {
      if ((st = vkgs_renderer_render(r, camera, model, d_rgba, stream)) != VKGS_OK) return st;
      st = vkgs_renderer_frame_info(r, nullptr, nullptr, stream);
      if (st != VKGS_ERR_OVERFLOW) break;
    }
    if (st != VKGS_OK) return st;
    for (uint32_t i = 0; i < 2 && st == VKGS_OK; ++i) st = vkgs_renderer_render(r, camera, model, d_rgba, stream);
    hipEventRecord(e0, s);
    for (uint32_t i = 0; i < frames && st == VKGS_OK; ++i) st = vkgs_renderer_render(r, camera, model, d_rgba, stream);
    hipEventRecord(e1, s);
    if (st != VKGS_OK) return st;
    if (hipEventSynchronize(e1) != hipSuccess) return VKGS_ERR_HIP;
    hipEventElapsedTime(ms, e0, e1);
    *ms /= (float)frames;
    return VKGS_OK;
  };
  // tile shape (and 8x8 one or two tiles per wave), then the resident-wave cap
  const std::vector<TuneStage> stages = {tune_shapes, tune_waves};
  vkgs_raster_config best = base;
  float best_t = 0.f;
  const int status = tune_configs(base, stages, 3, time_one, &best, &best_t);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  r->timing = timing;
  if (status != VKGS_OK) {
    vkgs_renderer_set_config(r, &base);
    return status;
  }
  vkgs_renderer_set_config(r, &best);
  if (chosen) *chosen = best;
  if (best_ms) *best_ms = best_t;
  return VKGS_OK;
}

int vkgs_renderer_set_timing(vkgs_renderer* r, int enable) {
  if (!r) return VKGS_ERR_INVALID_ARG;
  r->timing = enable != 0;
  drop_graph(r);
  return VKGS_OK;
}

int vkgs_renderer_stage_ms(vkgs_renderer* r, float ms[8]) {
  if (!r || !ms) return VKGS_ERR_INVALID_ARG;
  memcpy(ms, r->stage_ms, sizeof(r->stage_ms));
  return VKGS_OK;
}

int vkgs_renderer_raster_work(vkgs_renderer* r, uint64_t* wave_entries) {
  if (!r || !wave_entries || !r->work.p || !r->timing) return VKGS_ERR_INVALID_ARG;
  const size_t words = (size_t)r->rs.tiles_x * r->rs.tiles_y * (r->rs.tile_w * r->rs.tile_h / kWave);
  std::vector<uint32_t> w(words);
  VKGS_HIP_CHECK(hipDeviceSynchronize());
  VKGS_HIP_CHECK(hipMemcpy(w.data(), r->work.p, words * 4, hipMemcpyDeviceToHost));
  uint64_t sum = 0;
  for (uint32_t x : w) sum += x;
  *wave_entries = sum;
  return VKGS_OK;
}

}  // extern "C"
