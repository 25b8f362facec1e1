// Coordinate-descent search over raster configurations, shared by the plain
// renderer's and the differentiable path's autotuners.
#pragma once

#include <functional>
#include <vector>

#include "vkgs_hip.h"

namespace vkgs {

// One stage of the search: the candidates to time, derived from the best
// configuration found so far.
using TuneStage = std::function<std::vector<vkgs_raster_config>(const vkgs_raster_config&)>;
// Times one candidate (ms per frame or step) or returns a VKGS_ERR_* code.
using TuneTimer = std::function<int(const vkgs_raster_config&, float* ms)>;

// Runs the stages in order.  A stage times its candidates `rounds` times,
// visiting them in turn, and keeps each candidate's fastest round, so clock
// ramp-up or a transient slow period does not favour the ones timed later;
// its fastest candidate seeds the next stage.  *best_ms is the best time of
// the last stage.
inline int tune_configs(const vkgs_raster_config& base, const std::vector<TuneStage>& stages, int rounds,
                        const TuneTimer& time_one, vkgs_raster_config* best, float* best_ms) {
  vkgs_raster_config cur = base;
  float cur_ms = 1e30f;
  for (const TuneStage& stage : stages) {
    const std::vector<vkgs_raster_config> cands = stage(cur);
    if (cands.empty()) continue;
    std::vector<float> ms(cands.size(), 1e30f);
    for (int r = 0; r < rounds; ++r)
      for (size_t i = 0; i < cands.size(); ++i) {
        float t = 0.f;
        const int st = time_one(cands[i], &t);
        if (st != VKGS_OK) return st;
        if (t < ms[i]) ms[i] = t;
      }
    size_t bi = 0;
    for (size_t i = 1; i < cands.size(); ++i)
      if (ms[i] < ms[bi]) bi = i;
    cur = cands[bi];
    cur_ms = ms[bi];
  }
  *best = cur;
  *best_ms = cur_ms;
  return VKGS_OK;
}

// Tile shapes of a search's first stage; 8x8 appears twice, as two tiles per
// wave and as one (VKGS_RASTER_PAIRS_ALWAYS / _NEVER).
inline std::vector<vkgs_raster_config> tune_shapes(const vkgs_raster_config& from) {
  static const uint32_t kShapes[][3] = {{8, 8, VKGS_RASTER_PAIRS_ALWAYS},
                                        {8, 8, VKGS_RASTER_PAIRS_NEVER},
                                        {16, 8, 0},
                                        {8, 16, 0},
                                        {16, 4, 0},
                                        {16, 16, 0},
                                        {32, 4, 0}};
  std::vector<vkgs_raster_config> out;
  for (const auto& sh : kShapes) {
    vkgs_raster_config c = from;
    c.tile_w = sh[0];
    c.tile_h = sh[1];
    c.flags = (c.flags & ~(VKGS_RASTER_PAIRS_ALWAYS | VKGS_RASTER_PAIRS_NEVER)) | sh[2];
    out.push_back(c);
  }
  return out;
}

// Resident-wave caps of a search's last stage (0: uncapped).
inline std::vector<vkgs_raster_config> tune_waves(const vkgs_raster_config& from) {
  static const uint32_t kCaps[] = {0, 24, 16, 12, 8};
  std::vector<vkgs_raster_config> out;
  for (uint32_t w : kCaps) {
    vkgs_raster_config c = from;
    c.waves_per_cu = w;
    out.push_back(c);
  }
  return out;
}

}  // namespace vkgs
