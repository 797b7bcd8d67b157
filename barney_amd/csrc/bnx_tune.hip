// bnx_tune.hip -- launch tuning of the render path (include/bnx_rt.h,
// bnx_tuning) and its autotuner: coordinate descent over the raygen tile,
// the LDS share of the traversal stack, the walk order, the reordering of
// secondary rays, the shade block size, the pooled trace's refill threshold
// and launch size, bnx_render's paths per batch and the largest batch run as
// a captured bounce loop, each candidate timed on
// probe renders of the caller's request shape.  Host code only.
#include "bnx_internal.h"
#include "bnx_render.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

namespace bnx {
namespace {

bnx_tuning tuning_of(const Device *d)
{
  bnx_tuning t{};
  t.raygen_tile_w = d->raygen_tile_w;
  t.raygen_tile_h = d->raygen_tile_h;
  t.trace_stack_lds = d->trace_stack_lds;
  t.trace_blocks_per_cu = d->trace_blocks_per_cu;
  t.trace_refill = d->trace_refill;
  t.trace_walk = d->trace_walk;
  t.sort_rays = d->sort_rays;
  t.shade_block = d->shade_block;
  t.trace_refill_prims = d->trace_refill_prims;
  t.render_paths = d->render_paths;
  t.graph_paths = d->graph_paths;
  return t;
}

const char *tuning_error(const bnx_tuning &t)
{
  const bool row_major = t.raygen_tile_w == 0 && t.raygen_tile_h == 0;
  const bool tiled = t.raygen_tile_w >= 1 && t.raygen_tile_w <= 256 && t.raygen_tile_h >= 1 && t.raygen_tile_h <= 256;
  if (!row_major && !tiled) return "raygen tile: both 0, or both in 1..256";
  if (t.trace_stack_lds != 8 && t.trace_stack_lds != 16 && t.trace_stack_lds != 32)
    return "trace_stack_lds: 8, 16 or 32";
  if (t.trace_blocks_per_cu < 1 || t.trace_blocks_per_cu > 64) return "trace_blocks_per_cu: 1..64";
  if (t.trace_refill < 1 || t.trace_refill > 64) return "trace_refill: 1..64";
  if (t.trace_walk != 0 && t.trace_walk != 1) return "trace_walk: 0 or 1";
  if (t.sort_rays != 0 && t.sort_rays != 1) return "sort_rays: 0 or 1";
  if (t.shade_block != 256 && t.shade_block != 1024) return "shade_block: 256 or 1024";
  if (t.trace_refill_prims < 0) return "trace_refill_prims: >= 0";
  if (t.render_paths < (1LL << 16) || t.render_paths > (1LL << 26)) return "render_paths: 2^16 .. 2^26";
  if (t.graph_paths < 0 || t.graph_paths > (1LL << 26)) return "graph_paths: 0 .. 2^26";
  return nullptr;
}

void apply(Device *d, const bnx_tuning &t)
{
  d->raygen_tile_w = t.raygen_tile_w;
  d->raygen_tile_h = t.raygen_tile_h;
  d->trace_stack_lds = t.trace_stack_lds;
  d->trace_blocks_per_cu = t.trace_blocks_per_cu;
  d->trace_refill = t.trace_refill;
  d->trace_walk = t.trace_walk;
  d->sort_rays = t.sort_rays;
  d->shade_block = t.shade_block;
  d->trace_refill_prims = t.trace_refill_prims;
  d->render_paths = t.render_paths;
  d->graph_paths = t.graph_paths;
}

// one probe: the request rendered into the cleared scratch frame; wall time
// in ms (render_frame returns with the stream drained)
int probe_ms(World *w, const bnx_camera &cam, const bnx_renderer &rr, Frame *f, int spp, double &ms)
{
  int rc = frame_clear(f);
  if (rc != BNX_OK) return rc;
  BNX_HIP(hipStreamSynchronize(f->dev->stream));
  const auto t0 = std::chrono::steady_clock::now();
  if ((rc = render_frame(w, cam, rr, f, spp, nullptr)) != BNX_OK) return rc;
  ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return BNX_OK;
}

struct Tuner {
  Device *dev;
  World *w;
  bnx_camera cam;
  bnx_renderer rr;
  Frame *f;
  int spp;
  int reps = 2;
  int probes = 0;

  // best of `reps` probes with tuning t
  int time(const bnx_tuning &t, double &ms)
  {
    apply(dev, t);
    ms = 1e300;
    for (int k = 0; k < reps; k++) {
      double a = 0.0;
      const int rc = probe_ms(w, cam, rr, f, spp, a);
      if (rc != BNX_OK) return rc;
      ms = std::min(ms, a);
    }
    probes += reps;
    return BNX_OK;
  }
};

} // namespace
} // namespace bnx

using namespace bnx;

extern "C" {

int bnx_device_get_tuning(bnx_device d, bnx_tuning *out)
{
  BNX_CHECK(d && out, BNX_ERR_INVALID, "bnx_device_get_tuning: null argument");
  *out = tuning_of((const Device *)d);
  return BNX_OK;
}

int bnx_device_set_tuning(bnx_device d, const bnx_tuning *t)
{
  BNX_CHECK(d && t, BNX_ERR_INVALID, "bnx_device_set_tuning: null argument");
  const char *err = tuning_error(*t);
  BNX_CHECK(!err, BNX_ERR_INVALID, std::string("bnx_device_set_tuning: ") + (err ? err : ""));
  apply((Device *)d, *t);
  return BNX_OK;
}

int bnx_autotune(bnx_world world, const bnx_camera *camera, const bnx_renderer *renderer, int width, int height,
                 int probe_spp, bnx_tuning *chosen, bnx_autotune_report *report)
{
  GUARD_BEGIN
  World *w = (World *)world;
  BNX_CHECK(w && camera && renderer, BNX_ERR_INVALID, "bnx_autotune: null argument");
  BNX_CHECK(w->instances && w->instances->built, BNX_ERR_INVALID, "bnx_autotune: world not built");
  BNX_CHECK(width > 0 && height > 0 && (long long)width * height <= (1LL << 26), BNX_ERR_INVALID,
            "bnx_autotune: bad probe resolution");
  BNX_CHECK(probe_spp >= 1 && probe_spp <= 1024, BNX_ERR_INVALID, "bnx_autotune: probe_spp not in 1..1024");
  BNX_CHECK(camera->type == BNX_CAMERA_PERSPECTIVE && camera->aperture_radius == 0.f, BNX_ERR_UNSUPPORTED,
            "bnx_autotune: only the pinhole perspective camera is provided");
  Device *dev = w->dev;
  const bnx_tuning start = tuning_of(dev);
  Frame f;
  f.dev = dev;
  f.w = width;
  f.h = height;
  const size_t np = (size_t)width * height;
  hipError_t e;
  if ((e = f.accum.ensure(32 * np)) != hipSuccess || (e = f.depth.ensure(4 * np)) != hipSuccess ||
      (e = f.prim.ensure(4 * np)) != hipSuccess || (e = f.inst.ensure(4 * np)) != hipSuccess ||
      (e = f.obj.ensure(4 * np)) != hipSuccess || (e = f.normal.ensure(16 * np)) != hipSuccess) {
    set_error(std::string("bnx_autotune: scratch frame: ") + hipGetErrorString(e));
    return BNX_ERR_OOM;
  }
  Tuner T{dev, w, *camera, *renderer, &f, probe_spp};
  bnx_tuning best = start;
  double best_ms = 0.0, first_ms = 0.0;
  // the tuning in place, after one untimed render (first-use allocations)
  double warm = 0.0;
  int rc = probe_ms(w, T.cam, T.rr, &f, probe_spp, warm);
  // short probes are timed more often: at least about 10 ms per candidate,
  // sized from a third render (the second of a batch shape captures its
  // bounce loop)
  for (int k = 0; k < 2 && rc == BNX_OK; k++) rc = probe_ms(w, T.cam, T.rr, &f, probe_spp, warm);
  T.reps = std::max(2, std::min(20, (int)std::ceil(10.0 / std::max(1e-3, warm))));
  if (rc == BNX_OK) rc = T.time(best, best_ms);
  first_ms = best_ms;
  // every candidate of one field, the other fields at their best so far
  auto sweep = [&](auto &&set, const auto &values) -> int {
    for (const auto &v : values) {
      bnx_tuning t = best;
      set(t, v);
      if (tuning_error(t)) continue;
      if (std::memcmp(&t, &best, sizeof t) == 0) continue;
      double ms = 0.0;
      int r = T.time(t, ms);
      if (r != BNX_OK) return r;
      if (ms < 0.99 * best_ms) {
        best = t;
        best_ms = ms;
      }
    }
    apply(dev, best);
    return BNX_OK;
  };
  const std::vector<std::pair<int, int>> tiles = {{0, 0}, {8, 8}, {16, 4}, {16, 16}, {32, 8}};
  const std::vector<int> onoff = {0, 1};
  const std::vector<int> shade_blocks = {1024, 256};
  const std::vector<std::pair<int, int>> stack_walks = {{8, 0}, {16, 0}, {32, 0}, {8, 1}, {16, 1}, {32, 1}};
  const std::vector<int> refills = {64, 32, 16, 8, 4};
  const std::vector<int> blocks = {2, 4, 8, 16};
  const std::vector<long long> paths = {2LL << 20, 4LL << 20, 8LL << 20, 16LL << 20};
  const std::vector<long long> graphs = {0LL, 1LL << 20, 4LL << 20, 16LL << 20};
  const long long prims = w->instances->blas_prims;
  // two passes: a field's best can move once the fields after it changed
  for (int pass = 0; pass < 2 && rc == BNX_OK; pass++) {
    rc = sweep([](bnx_tuning &t, const std::pair<int, int> &v) { t.raygen_tile_w = v.first; t.raygen_tile_h = v.second; },
               tiles);
    // the LDS stack depth and the walk order together: both change the
    // trace kernel's register budget, so neither's best holds for all of the
    // other's values
    if (rc == BNX_OK)
      rc = sweep(
          [](bnx_tuning &t, const std::pair<int, int> &v) {
            t.trace_stack_lds = v.first;
            t.trace_walk = v.second;
          },
          stack_walks);
    if (rc == BNX_OK) rc = sweep([](bnx_tuning &t, int v) { t.sort_rays = v; }, onoff);
    if (rc == BNX_OK) rc = sweep([](bnx_tuning &t, int v) { t.shade_block = v; }, shade_blocks);
    // the refill threshold for worlds of this size: 64 keeps one ray per
    // thread for them; below 64 the pooled trace takes their secondary rays
    if (rc == BNX_OK)
      rc = sweep(
          [prims, &start](bnx_tuning &t, int v) {
            if (v >= 64) {
              t.trace_refill = start.trace_refill;
              t.trace_refill_prims = std::max<int64_t>(start.trace_refill_prims, prims + 1);
            } else {
              t.trace_refill = v;
              t.trace_refill_prims = std::min<int64_t>(start.trace_refill_prims, prims);
            }
          },
          refills);
    const bool pooled = best.trace_refill < 64 && prims >= best.trace_refill_prims;
    if (rc == BNX_OK && pooled) rc = sweep([](bnx_tuning &t, int v) { t.trace_blocks_per_cu = v; }, blocks);
    if (rc == BNX_OK && (long long)np * probe_spp > (1LL << 20))
      rc = sweep([](bnx_tuning &t, long long v) { t.render_paths = v; }, paths);
    if (rc == BNX_OK) rc = sweep([](bnx_tuning &t, long long v) { t.graph_paths = v; }, graphs);
  }
  if (rc != BNX_OK) {
    apply(dev, start);
    return rc;
  }
  apply(dev, best);
  if (chosen) *chosen = best;
  if (report) {
    report->ms_before = first_ms;
    report->ms_after = best_ms;
    report->probes = T.probes;
    report->_pad = 0;
  }
  return BNX_OK;
  GUARD_END
}

} // extern "C"
