// autotune.hip — picks the voxelize bin size and the cone-march screen tile for a batch
// by timing the batch itself (host code over the public entry points).
//
// Bin size trades the front end's work (more, smaller bins: more (triangle, bin) pairs
// and slot atomics) against the voxelize kernel's (bigger bins: fewer, longer work items
// and bigger LDS tiles swept per item); the screen tile trades workgroup size against
// tail effects and the G-buffer's spatial coherence, and Morton order (a sort before
// the march) pays off only when the G-buffer is not coherent. None has a closed-form optimum
// across scene mixes, so both are measured. The two sweeps are independent: the trace
// reads only the mip levels, which are the same bits for every bin size, and voxelize and
// mips never read the tile shape.
#include <algorithm>
#include <cfloat>
#include <cstring>

#include "vgi_internal.hpp"

namespace {

constexpr uint32_t kBins[3] = {8, 16, 32};
// pixel schedules: screen tiles of three shapes, and Morton order (256-pixel runs)
constexpr uint32_t kSched[4][3] = {{8, 8, VGI_ORDER_TILES},
                                   {16, 16, VGI_ORDER_TILES},
                                   {32, 8, VGI_ORDER_TILES},
                                   {16, 16, VGI_ORDER_MORTON}};

int time_build(vgi_context* ctx, const vgi_voxelize_item* vox, const vgi_mips_item* mips, uint32_t count,
               uint32_t precision, void* stream, float* ms) {
  int rc = vgi_timer_start(ctx, stream);
  if (rc == VGI_OK) rc = vgi_voxelize_batch(ctx, vox, count, stream);
  if (rc == VGI_OK) rc = vgi_build_mips_batch(ctx, mips, count, precision, stream);
  if (rc == VGI_OK) rc = vgi_timer_stop(ctx, stream, ms);
  return rc;
}

int time_trace(vgi_context* ctx, const vgi_trace_item* trace, uint32_t count, void* stream, float* ms) {
  int rc = vgi_timer_start(ctx, stream);
  if (rc == VGI_OK) rc = vgi_trace_batch(ctx, trace, count, stream);
  if (rc == VGI_OK) rc = vgi_timer_stop(ctx, stream, ms);
  return rc;
}

}  // namespace

extern "C" int vgi_autotune_batch(vgi_context* ctx, const vgi_voxelize_item* vox, const vgi_mips_item* mips,
                                  const vgi_trace_item* trace, uint32_t count, uint32_t precision,
                                  uint32_t reps, vgi_autotune_report* report, void* stream) {
  if (!ctx || !vox || !mips || !trace || !report || count == 0)
    return vgi::set_error(VGI_E_INVALID, "autotune: NULL argument or empty batch");
  for (uint32_t i = 0; i < count; i++)
    if (vox[i].scene != mips[i].scene || (const vgi_scene*)vox[i].scene != trace[i].scene)
      return vgi::set_error(VGI_E_INVALID, "autotune: item %u names different scenes in vox/mips/trace", i);
  reps = std::max(reps, 1u);
  vgi_tuning base;
  int rc = vgi_get_tuning(ctx, &base);
  if (rc != VGI_OK) return rc;
  vgi_autotune_report r;
  std::memset(&r, 0, sizeof(r));
  for (int b = 0; b < 3; b++) {
    vgi_tuning t = base;
    t.bin_size = kBins[b];
    r.bin_sizes[b] = kBins[b];
    if ((rc = vgi_set_tuning(ctx, &t)) != VGI_OK) return rc;
    float best = FLT_MAX, ms = 0.f;
    for (uint32_t k = 0; k <= reps; k++) {
      if ((rc = time_build(ctx, vox, mips, count, precision, stream, &ms)) != VGI_OK) {
        vgi_set_tuning(ctx, &base);
        return rc;
      }
      if (k > 0) best = std::min(best, ms);  // run 0 warms allocations and caches
    }
    r.build_ms[b] = best;
  }
  for (int j = 0; j < 4; j++) {
    vgi_tuning t = base;
    t.tile_w = kSched[j][0];
    t.tile_h = kSched[j][1];
    t.pixel_order = kSched[j][2];
    for (int c = 0; c < 3; c++) r.tiles[j][c] = kSched[j][c];
    if ((rc = vgi_set_tuning(ctx, &t)) != VGI_OK) return rc;
    float best = FLT_MAX, ms = 0.f;
    for (uint32_t k = 0; k <= reps; k++) {
      if ((rc = time_trace(ctx, trace, count, stream, &ms)) != VGI_OK) {
        vgi_set_tuning(ctx, &base);
        return rc;
      }
      if (k > 0) best = std::min(best, ms);
    }
    r.trace_ms[j] = best;
  }
  int bb = 0, bt = 0;
  for (int b = 0; b < 3; b++)
    for (int j = 0; j < 4; j++) {
      r.frame_ms[b][j] = r.build_ms[b] + r.trace_ms[j];
      if (r.frame_ms[b][j] < r.frame_ms[bb][bt]) {
        bb = b;
        bt = j;
      }
    }
  r.best = base;
  r.best.bin_size = kBins[bb];
  r.best.tile_w = kSched[bt][0];
  r.best.tile_h = kSched[bt][1];
  r.best.pixel_order = kSched[bt][2];
  if ((rc = vgi_set_tuning(ctx, &r.best)) != VGI_OK) return rc;
  // leave the scenes built with the chosen bin size
  if ((rc = vgi_voxelize_batch(ctx, vox, count, stream)) != VGI_OK) return rc;
  if ((rc = vgi_build_mips_batch(ctx, mips, count, precision, stream)) != VGI_OK) return rc;
  if ((rc = vgi_trace_batch(ctx, trace, count, stream)) != VGI_OK) return rc;
  if ((rc = vgi_sync(ctx, stream)) != VGI_OK) return rc;
  *report = r;
  return VGI_OK;
}
