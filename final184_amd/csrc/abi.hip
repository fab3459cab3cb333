// abi.hip — the C ABI of include/vgi.h: contexts, scenes (device memory owners),
// the voxelize / mips / trace passes and parity readback.
//
// Ownership follows the reference: CMegaPipeline owns the voxel image
// (Foreground/Renderer/MegaPipeline.cpp:470-490, MegaPipeline.h:60,72) and the passes
// borrow the caller's GPU buffers for the duration of a draw. Here a vgi_scene owns the
// accumulator, the bin/sort scratch and the mip levels; meshes, G-buffers and outputs are
// caller device pointers borrowed on the given stream. Errors never throw across the
// ABI: every entry returns a vgi_status and sets a thread-local message, replacing the
// reference's silent early returns (VoxelizeRenderer.cpp:49-50,87-88).
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "vgi_internal.hpp"

using namespace vgi;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIP_TRY(expr)                                                                        \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess)                                                                    \
      return fail(e_ == hipErrorOutOfMemory ? VGI_E_NOMEM : VGI_E_HIP, "%s failed: %s (%s:%d)", \
                  #expr, hipGetErrorString(e_), __FILE__, __LINE__);                         \
  } while (0)

#define TRY(expr)          \
  do {                     \
    int r_ = (expr);       \
    if (r_ != VGI_OK) return r_; \
  } while (0)

bool is_pow2(uint32_t v) { return v && !(v & (v - 1)); }
uint32_t ilog2(uint32_t v) {
  uint32_t r = 0;
  while ((1u << (r + 1)) <= v) r++;
  return r;
}

}  // namespace

struct vgi_context {
  int device = 0;
  int num_cus = 256;
  hipStream_t stream = nullptr;
  vgi_tuning tuning{8, 64, 16, 16};
  int profiling = 0;
  float* d_lut = nullptr;
  unsigned long long* d_counter = nullptr;
  unsigned long long* h_counter = nullptr;
  hipEvent_t ev[10] = {};
  bool ev_vox = false, ev_mips = false, ev_trace = false;
  hipEvent_t tstart = nullptr, tstop = nullptr;
  vgi_timings timings{};
  std::mutex mu;
};

struct vgi_scene {
  vgi_context* ctx = nullptr;
  vgi_grid grid{};
  bool grid_set = false;
  uint32_t bin_shift = 3;
  uint32_t nbins_axis = 0;
  uint32_t chunk = 64;
  uint32_t* acc = nullptr;   // n^3 x 8 words, brick-major
  uint32_t* mask = nullptr;  // 16 words per 8^3 brick
  size_t acc_bytes = 0;
  TriRec* tris = nullptr;
  uint32_t* tri_slot = nullptr;
  uint32_t* big_list = nullptr;
  size_t tris_cap = 0;
  MeshDev* meshes = nullptr;
  MeshDev* meshes_host = nullptr;
  size_t meshes_cap = 0;
  uint32_t* bin_small = nullptr;  // bin_small | bin_big | big_cursor | zero_list, nbins each
  unsigned long long* bin_packed = nullptr;
  unsigned long long* bin_scan = nullptr;
  size_t bins_cap = 0;
  uint32_t* pairs = nullptr;
  size_t pairs_cap = 0;
  uint32_t* work_bin = nullptr;
  size_t work_cap = 0;
  DevStats* stats = nullptr;
  DevStats* h_stats = nullptr;
  unsigned long long* h_total = nullptr;
  void* scan_tmp = nullptr;
  size_t scan_tmp_bytes = 0;
  hipEvent_t staging_done = nullptr;
  uint32_t prec = 0;
  void* level[VGI_MAX_LEVELS] = {};
  size_t level_bytes[VGI_MAX_LEVELS] = {};
  uint32_t alloc_levels = 0, alloc_n = 0, alloc_prec = 0;
  bool has_voxels = false, has_mips = false;
  uint64_t tris_in = 0, num_pairs = 0, work_items = 0;
};

namespace {

size_t scene_device_bytes(const vgi_scene* s) {
  size_t b = s->acc_bytes + s->acc_bytes / 16 + s->tris_cap * (sizeof(TriRec) + 4 * (kSmallBins + 1)) +
             s->meshes_cap * sizeof(MeshDev) + s->bins_cap * (4 * 4 + 2 * 8) +
             (s->pairs_cap + s->work_cap) * sizeof(uint32_t) + s->scan_tmp_bytes + sizeof(DevStats);
  for (uint32_t k = 0; k < VGI_MAX_LEVELS; k++) b += s->level_bytes[k];
  return b;
}

template <typename T>
int grow(T** p, size_t* cap, size_t need, const char* what) {
  if (*cap >= need && *p) return VGI_OK;
  size_t n = std::max(need, *cap + *cap / 2);
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  *cap = 0;
  hipError_t e = hipMalloc((void**)p, n * sizeof(T));
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return fail(VGI_E_NOMEM, "device allocation of %zu bytes for %s failed: %s", n * sizeof(T), what,
                hipGetErrorString(e));
  }
  *cap = n;
  return VGI_OK;
}

int check_grid(const vgi_grid* g) {
  if (!g) return fail(VGI_E_INVALID, "grid is NULL");
  if (!is_pow2(g->n) || g->n < 8) return fail(VGI_E_INVALID, "grid n=%u must be a power of two >= 8", g->n);
  if (g->n > 1024) return fail(VGI_E_RANGE, "grid n=%u exceeds the maximum 1024", g->n);
  if (g->levels == 0 || g->levels > VGI_MAX_LEVELS || (g->n >> (g->levels - 1)) == 0)
    return fail(VGI_E_INVALID, "levels=%u invalid for n=%u", g->levels, g->n);
  if (!(g->extent > 0.0f) || !std::isfinite(g->extent))
    return fail(VGI_E_INVALID, "grid extent must be positive and finite");
  for (int a = 0; a < 3; a++)
    if (!std::isfinite(g->origin[a])) return fail(VGI_E_INVALID, "grid origin must be finite");
  return VGI_OK;
}

hipStream_t pick(vgi_context* ctx, void* stream) {
  return stream ? (hipStream_t)stream : ctx->stream;
}

// Host-side storage decode (readback only).
float host_half(uint16_t h) {
  uint32_t sign = ((uint32_t)h & 0x8000u) << 16, exp = (h >> 10) & 0x1fu, mant = h & 0x3ffu;
  uint32_t bits;
  if (exp == 0) {
    float v = std::ldexp((float)mant, -24);
    return sign ? -v : v;
  }
  if (exp == 31) bits = sign | 0x7f800000u | (mant << 13);
  else bits = sign | ((exp + 112u) << 23) | (mant << 13);
  float f;
  std::memcpy(&f, &bits, 4);
  return f;
}
float host_bf16(uint16_t b) {
  uint32_t bits = (uint32_t)b << 16;
  float f;
  std::memcpy(&f, &bits, 4);
  return f;
}

}  // namespace

extern "C" {

const char* vgi_version(void) { return "vgi 0.1.0 (gfx950)"; }
uint32_t vgi_abi_version(void) { return VGI_ABI_VERSION; }
const char* vgi_last_error(void) { return g_err.c_str(); }

int vgi_create(vgi_context** out, int device) {
  if (!out) return fail(VGI_E_INVALID, "out is NULL");
  *out = nullptr;
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0) {
    (void)hipGetLastError();
    return fail(VGI_E_NODEVICE, "no HIP device available (%s)", hipGetErrorString(e));
  }
  if (device < 0 || device >= count) return fail(VGI_E_INVALID, "device %d out of range [0,%d)", device, count);
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, device));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(VGI_E_NODEVICE, "device %d is %s; this library carries gfx950 code only", device,
                prop.gcnArchName);
  HIP_TRY(hipSetDevice(device));
  vgi_context* ctx = new vgi_context();
  ctx->device = device;
  ctx->num_cus = prop.multiProcessorCount;
  e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (e != hipSuccess) { delete ctx; return fail(VGI_E_HIP, "stream create: %s", hipGetErrorString(e)); }
  float lut[256];
  for (int i = 0; i < 256; i++) lut[i] = (float)std::pow((double)i / 255.0, 2.2);
  if (hipMalloc((void**)&ctx->d_lut, sizeof(lut)) != hipSuccess ||
      hipMemcpy(ctx->d_lut, lut, sizeof(lut), hipMemcpyHostToDevice) != hipSuccess ||
      hipMalloc((void**)&ctx->d_counter, sizeof(unsigned long long)) != hipSuccess ||
      hipHostMalloc((void**)&ctx->h_counter, sizeof(unsigned long long), hipHostMallocDefault) != hipSuccess) {
    vgi_destroy(ctx);
    return fail(VGI_E_NOMEM, "context allocation failed");
  }
  for (auto& ev : ctx->ev) HIP_TRY(hipEventCreate(&ev));
  HIP_TRY(hipEventCreate(&ctx->tstart));
  HIP_TRY(hipEventCreate(&ctx->tstop));
  *out = ctx;
  return VGI_OK;
}

int vgi_destroy(vgi_context* ctx) {
  if (!ctx) return VGI_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->d_lut) (void)hipFree(ctx->d_lut);
  if (ctx->d_counter) (void)hipFree(ctx->d_counter);
  if (ctx->h_counter) (void)hipHostFree(ctx->h_counter);
  for (auto& ev : ctx->ev)
    if (ev) (void)hipEventDestroy(ev);
  if (ctx->tstart) (void)hipEventDestroy(ctx->tstart);
  if (ctx->tstop) (void)hipEventDestroy(ctx->tstop);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return VGI_OK;
}

int vgi_set_tuning(vgi_context* ctx, const vgi_tuning* t) {
  if (!ctx || !t) return fail(VGI_E_INVALID, "NULL argument");
  if (t->bin_size != 8 && t->bin_size != 16 && t->bin_size != 32)
    return fail(VGI_E_INVALID, "bin_size %u not in {8,16,32}", t->bin_size);
  if (t->chunk_tris < 1 || t->chunk_tris > (uint32_t)kMaxChunk)
    return fail(VGI_E_INVALID, "chunk_tris %u not in [1,%d]", t->chunk_tris, kMaxChunk);
  const bool ok_tile = (t->tile_w == 8 && t->tile_h == 8) || (t->tile_w == 16 && t->tile_h == 16) ||
                       (t->tile_w == 32 && t->tile_h == 8) || (t->tile_w == 16 && t->tile_h == 8) ||
                       (t->tile_w == 8 && t->tile_h == 16);
  if (!ok_tile) return fail(VGI_E_INVALID, "tile %ux%u not supported", t->tile_w, t->tile_h);
  std::lock_guard<std::mutex> lk(ctx->mu);
  ctx->tuning = *t;
  return VGI_OK;
}

int vgi_get_tuning(const vgi_context* ctx, vgi_tuning* t) {
  if (!ctx || !t) return fail(VGI_E_INVALID, "NULL argument");
  *t = ctx->tuning;
  return VGI_OK;
}

int vgi_set_profiling(vgi_context* ctx, int enabled) {
  if (!ctx) return fail(VGI_E_INVALID, "ctx is NULL");
  ctx->profiling = enabled ? 1 : 0;
  return VGI_OK;
}

int vgi_get_timings(const vgi_context* cctx, vgi_timings* out) {
  vgi_context* ctx = const_cast<vgi_context*>(cctx);
  if (!ctx || !out) return fail(VGI_E_INVALID, "NULL argument");
  HIP_TRY(hipSetDevice(ctx->device));
  vgi_timings t = ctx->timings;
  if (ctx->ev_vox) {
    HIP_TRY(hipEventSynchronize(ctx->ev[4]));
    HIP_TRY(hipEventElapsedTime(&t.voxelize_setup_ms, ctx->ev[0], ctx->ev[1]));
    HIP_TRY(hipEventElapsedTime(&t.voxelize_sort_ms, ctx->ev[2], ctx->ev[3]));
    float a = 0.f;
    HIP_TRY(hipEventElapsedTime(&a, ctx->ev[1], ctx->ev[2]));
    t.voxelize_sort_ms += a;  // scan + host read of the pair total
    HIP_TRY(hipEventElapsedTime(&t.voxelize_raster_ms, ctx->ev[3], ctx->ev[4]));
    HIP_TRY(hipEventElapsedTime(&t.voxelize_total_ms, ctx->ev[0], ctx->ev[4]));
    ctx->ev_vox = false;
  }
  if (ctx->ev_mips) {
    HIP_TRY(hipEventSynchronize(ctx->ev[6]));
    HIP_TRY(hipEventElapsedTime(&t.mips_ms, ctx->ev[5], ctx->ev[6]));
    ctx->ev_mips = false;
  }
  if (ctx->ev_trace) {
    HIP_TRY(hipEventSynchronize(ctx->ev[8]));
    HIP_TRY(hipEventElapsedTime(&t.trace_ms, ctx->ev[7], ctx->ev[8]));
    HIP_TRY(hipMemcpy(ctx->h_counter, ctx->d_counter, sizeof(unsigned long long), hipMemcpyDeviceToHost));
    t.trace_trilinear_samples = *ctx->h_counter;
    ctx->ev_trace = false;
  }
  ctx->timings = t;
  *out = t;
  return VGI_OK;
}

int vgi_sync(vgi_context* ctx, void* stream) {
  if (!ctx) return fail(VGI_E_INVALID, "ctx is NULL");
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipStreamSynchronize(pick(ctx, stream)));
  return VGI_OK;
}

int vgi_scene_create(vgi_context* ctx, vgi_scene** out) {
  if (!ctx || !out) return fail(VGI_E_INVALID, "NULL argument");
  HIP_TRY(hipSetDevice(ctx->device));
  vgi_scene* s = new vgi_scene();
  s->ctx = ctx;
  if (hipMalloc((void**)&s->stats, sizeof(DevStats)) != hipSuccess ||
      hipHostMalloc((void**)&s->h_stats, sizeof(DevStats), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&s->h_total, 2 * sizeof(unsigned long long), hipHostMallocDefault) != hipSuccess ||
      hipEventCreateWithFlags(&s->staging_done, hipEventDisableTiming) != hipSuccess) {
    vgi_scene_release(s);
    return fail(VGI_E_NOMEM, "scene allocation failed");
  }
  *out = s;
  return VGI_OK;
}

int vgi_scene_release(vgi_scene* s) {
  if (!s) return VGI_OK;
  (void)hipSetDevice(s->ctx->device);
  (void)hipDeviceSynchronize();
  void* dev[] = {s->acc, s->mask, s->tris, s->tri_slot, s->big_list, s->meshes, s->bin_small,
                 s->bin_packed, s->bin_scan, s->pairs, s->work_bin, s->stats, s->scan_tmp};
  for (void* p : dev)
    if (p) (void)hipFree(p);
  for (uint32_t k = 0; k < VGI_MAX_LEVELS; k++)
    if (s->level[k]) (void)hipFree(s->level[k]);
  if (s->meshes_host) (void)hipHostFree(s->meshes_host);
  if (s->h_stats) (void)hipHostFree(s->h_stats);
  if (s->h_total) (void)hipHostFree(s->h_total);
  if (s->staging_done) (void)hipEventDestroy(s->staging_done);
  delete s;
  return VGI_OK;
}

uint64_t vgi_scene_bytes_estimate(const vgi_grid* g, uint32_t precision, uint64_t ntri) {
  if (!g || !is_pow2(g->n) || g->levels == 0 || g->levels > VGI_MAX_LEVELS) return 0;
  const uint64_t n = g->n;
  uint64_t b = n * n * n * kAccWords * 4;
  for (uint32_t k = 0; k < g->levels; k++) {
    const uint64_t p = (n >> k) + 2;
    b += p * p * p * texel_bytes(precision);
  }
  const uint64_t nb = n / 8;
  b += nb * nb * nb * 12 + ntri * (sizeof(TriRec) + 16) + (1u << 20);
  return b;
}

int vgi_scene_get_stats(const vgi_scene* s, vgi_scene_stats* out) {
  if (!s || !out) return fail(VGI_E_INVALID, "NULL argument");
  std::memset(out, 0, sizeof(*out));
  out->triangles_in = s->tris_in;
  out->bin_pairs = s->num_pairs;
  out->work_items = s->work_items;
  out->device_bytes = scene_device_bytes(s);
  out->n = s->grid.n;
  out->levels = s->grid.levels;
  out->precision = s->prec;
  out->has_voxels = s->has_voxels;
  out->has_mips = s->has_mips;
  out->bin_size = 1u << s->bin_shift;
  if (s->has_voxels) {
    (void)hipSetDevice(s->ctx->device);
    HIP_TRY(hipMemcpy(s->h_stats, s->stats, sizeof(DevStats), hipMemcpyDeviceToHost));
    out->triangles_rejected = s->h_stats->rejected;
    out->triangles_degenerate = s->h_stats->degenerate;
  }
  return VGI_OK;
}

int vgi_voxelize(vgi_context* ctx, vgi_scene* s, const vgi_mesh* meshes, uint32_t num_meshes,
                 const vgi_grid* grid, void* stream) {
  if (!ctx || !s) return fail(VGI_E_INVALID, "NULL context or scene");
  if (s->ctx != ctx) return fail(VGI_E_INVALID, "scene belongs to another context");
  TRY(check_grid(grid));
  if (num_meshes > 0 && !meshes) return fail(VGI_E_INVALID, "meshes is NULL");
  uint64_t ntri = 0;
  for (uint32_t i = 0; i < num_meshes; i++) {
    const vgi_mesh& m = meshes[i];
    if (m.num_triangles == 0) continue;
    if (!m.positions) return fail(VGI_E_INVALID, "mesh %u: positions is NULL", i);
    const uint32_t stride = m.position_stride ? m.position_stride : 12u;
    if (stride < 12 || (stride & 3)) return fail(VGI_E_INVALID, "mesh %u: position_stride %u", i, stride);
    if (m.index_type > VGI_INDEX_U32) return fail(VGI_E_INVALID, "mesh %u: index_type %u", i, m.index_type);
    if (m.index_type != VGI_INDEX_NONE && !m.indices)
      return fail(VGI_E_INVALID, "mesh %u: indices is NULL", i);
    if (m.index_type == VGI_INDEX_NONE && (uint64_t)m.num_triangles * 3 > m.num_vertices)
      return fail(VGI_E_INVALID, "mesh %u: non-indexed mesh needs 3 vertices per triangle", i);
    for (int k = 0; k < 12; k++)
      if (!std::isfinite(m.model[k])) return fail(VGI_E_INVALID, "mesh %u: model matrix not finite", i);
    ntri += m.num_triangles;
  }
  if (ntri >= (1ull << 31)) return fail(VGI_E_RANGE, "%llu triangles exceed 2^31", (unsigned long long)ntri);
  HIP_TRY(hipSetDevice(ctx->device));
  hipStream_t st = pick(ctx, stream);
  vgi_tuning tn;
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    tn = ctx->tuning;
  }
  const uint32_t n = grid->n;
  const uint32_t bin = std::min(tn.bin_size, n);
  s->bin_shift = ilog2(bin);
  s->nbins_axis = n >> s->bin_shift;
  s->chunk = tn.chunk_tris;
  const uint32_t nbins = s->nbins_axis * s->nbins_axis * s->nbins_axis;
  s->has_voxels = false;
  s->has_mips = false;

  // ---- allocations (grown on demand, kept across calls) ----
  const size_t acc_words = (size_t)n * n * n * kAccWords;
  if (s->acc_bytes != acc_words * 4) {
    if (s->acc) (void)hipFree(s->acc);
    if (s->mask) (void)hipFree(s->mask);
    s->acc = nullptr;
    s->mask = nullptr;
    s->acc_bytes = 0;
    const size_t bricks = (size_t)(n / 8) * (n / 8) * (n / 8);
    if (hipMalloc((void**)&s->acc, acc_words * 4) != hipSuccess ||
        hipMalloc((void**)&s->mask, bricks * 16 * 4) != hipSuccess) {
      (void)hipGetLastError();
      return fail(VGI_E_NOMEM, "accumulator allocation (%zu bytes) failed", acc_words * 4);
    }
    s->acc_bytes = acc_words * 4;
  }
  const size_t tneed = std::max<size_t>(ntri, 1);
  if (s->tris_cap < tneed) {
    void* arrs[3] = {s->tris, s->tri_slot, s->big_list};
    for (void* a : arrs)
      if (a) (void)hipFree(a);
    s->tris = nullptr;
    s->tri_slot = s->big_list = nullptr;
    s->tris_cap = 0;
    const size_t cap = std::max(tneed, (size_t)1024);
    if (hipMalloc((void**)&s->tris, cap * sizeof(TriRec)) != hipSuccess ||
        hipMalloc((void**)&s->tri_slot, cap * kSmallBins * 4) != hipSuccess ||
        hipMalloc((void**)&s->big_list, cap * 4) != hipSuccess) {
      (void)hipGetLastError();
      return fail(VGI_E_NOMEM, "triangle record allocation failed");
    }
    s->tris_cap = cap;
  }
  if (s->bins_cap < (size_t)nbins + 1) {
    if (s->bin_small) (void)hipFree(s->bin_small);
    if (s->bin_packed) (void)hipFree(s->bin_packed);
    if (s->bin_scan) (void)hipFree(s->bin_scan);
    s->bin_small = nullptr;
    s->bin_packed = s->bin_scan = nullptr;
    s->bins_cap = 0;
    const size_t cap = (size_t)nbins + 1;
    if (hipMalloc((void**)&s->bin_small, cap * 4 * 4) != hipSuccess ||
        hipMalloc((void**)&s->bin_packed, cap * 8) != hipSuccess ||
        hipMalloc((void**)&s->bin_scan, cap * 8) != hipSuccess) {
      (void)hipGetLastError();
      return fail(VGI_E_NOMEM, "bin array allocation failed");
    }
    s->bins_cap = cap;
  }
  {
    size_t need = 0;
    HIP_TRY(scan_bins(nullptr, &need, s->bin_packed, s->bin_scan, nbins + 1, st));
    if (need > s->scan_tmp_bytes) {
      if (s->scan_tmp) (void)hipFree(s->scan_tmp);
      s->scan_tmp = nullptr;
      s->scan_tmp_bytes = 0;
      HIP_TRY(hipMalloc(&s->scan_tmp, need));
      s->scan_tmp_bytes = need;
    }
  }
  const uint32_t nm = std::max<uint32_t>(num_meshes, 1);
  if (s->meshes_cap < nm) {
    if (s->meshes) (void)hipFree(s->meshes);
    if (s->meshes_host) {
      (void)hipEventSynchronize(s->staging_done);
      (void)hipHostFree(s->meshes_host);
    }
    s->meshes = nullptr;
    s->meshes_host = nullptr;
    s->meshes_cap = 0;
    HIP_TRY(hipMalloc((void**)&s->meshes, nm * sizeof(MeshDev)));
    HIP_TRY(hipHostMalloc((void**)&s->meshes_host, nm * sizeof(MeshDev), hipHostMallocDefault));
    s->meshes_cap = nm;
  }

  // ---- mesh table ----
  HIP_TRY(hipEventSynchronize(s->staging_done));
  {
    uint32_t base = 0;
    for (uint32_t i = 0; i < nm; i++) {
      MeshDev& d = s->meshes_host[i];
      std::memset(&d, 0, sizeof(d));
      if (i >= num_meshes) continue;
      const vgi_mesh& m = meshes[i];
      d.pos = m.positions;
      d.idx = m.indices;
      d.albedo = m.albedo_rgb8;
      d.stride = m.position_stride ? m.position_stride : 12u;
      d.nverts = m.num_vertices;
      d.idx_type = m.index_type;
      d.ntri = m.num_triangles;
      d.tri_base = base;
      d.base_rgb = (uint32_t)m.base_color[0] | ((uint32_t)m.base_color[1] << 8) |
                   ((uint32_t)m.base_color[2] << 16);
      std::memcpy(d.m, m.model, sizeof(d.m));
      base += m.num_triangles;
    }
  }
  HIP_TRY(hipMemcpyAsync(s->meshes, s->meshes_host, nm * sizeof(MeshDev), hipMemcpyHostToDevice, st));
  HIP_TRY(hipEventRecord(s->staging_done, st));

  VoxParams P{};
  P.meshes = s->meshes;
  P.tris = s->tris;
  P.tri_slot = s->tri_slot;
  P.big_list = s->big_list;
  P.bin_small = s->bin_small;
  P.bin_big = s->bin_small + s->bins_cap;
  P.big_cursor = s->bin_small + 2 * s->bins_cap;
  P.zero_list = s->bin_small + 3 * s->bins_cap;
  P.bin_packed = s->bin_packed;
  P.bin_scan = s->bin_scan;
  P.acc = s->acc;
  P.mask = s->mask;
  P.stats = s->stats;
  for (int a = 0; a < 3; a++) P.origin[a] = grid->origin[a];
  P.snap_scale = (float)(n * 256u) / grid->extent;
  P.nmesh = num_meshes;
  P.ntri = num_meshes ? (uint32_t)ntri : 0u;
  P.n = n;
  P.bin_shift = s->bin_shift;
  P.nbins_axis = s->nbins_axis;
  P.chunk = s->chunk;

  const uint32_t persistent = (uint32_t)ctx->num_cus * 4;
  const bool prof = ctx->profiling != 0;
  if (prof) HIP_TRY(hipEventRecord(ctx->ev[0], st));
  // bin_small, bin_big, big_cursor: 3 x bins_cap words; stats
  HIP_TRY(launch_fill_u32(s->bin_small, 0u, 3 * s->bins_cap, st));
  HIP_TRY(launch_fill_u32((uint32_t*)s->stats, 0u, sizeof(DevStats) / 4, st));
  HIP_TRY(launch_setup(P, st));
  HIP_TRY(launch_big(P, false, persistent, st));
  if (prof) HIP_TRY(hipEventRecord(ctx->ev[1], st));
  HIP_TRY(launch_bins(P, st));
  HIP_TRY(scan_bins(s->scan_tmp, &s->scan_tmp_bytes, s->bin_packed, s->bin_scan, nbins + 1, st));
  HIP_TRY(hipMemcpyAsync(s->h_total, s->bin_scan + nbins, sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  const unsigned long long tot = s->h_total[0];
  const uint32_t total_pairs = (uint32_t)tot, total_work = (uint32_t)(tot >> 32);
  if (prof) HIP_TRY(hipEventRecord(ctx->ev[2], st));
  TRY(grow(&s->pairs, &s->pairs_cap, std::max<size_t>(total_pairs, 1), "bin pairs"));
  TRY(grow(&s->work_bin, &s->work_cap, std::max<size_t>(total_work, 1), "work items"));
  P.pairs = s->pairs;
  P.work_bin = s->work_bin;
  P.pairs_cap = (uint32_t)std::min<size_t>(s->pairs_cap, 0xFFFFFFFFu);
  if (total_pairs > 0) {
    HIP_TRY(launch_workmap(P, st));
    HIP_TRY(launch_emit_small(P, st));
    HIP_TRY(launch_big(P, true, persistent, st));
    HIP_TRY(launch_zero(P, persistent, st));
  }
  if (prof) HIP_TRY(hipEventRecord(ctx->ev[3], st));
  const uint32_t tile = std::min(1u << s->bin_shift, 16u);
  const uint32_t grid_blocks = std::min<uint32_t>(total_work, (uint32_t)ctx->num_cus * 8);
  HIP_TRY(launch_voxelize(P, tile, grid_blocks, total_work, st));
  if (prof) {
    HIP_TRY(hipEventRecord(ctx->ev[4], st));
    ctx->ev_vox = true;
  }
  s->grid = *grid;
  s->grid_set = true;
  s->tris_in = ntri;
  s->num_pairs = total_pairs;
  s->work_items = total_work;
  s->has_voxels = true;
  return VGI_OK;
}

int vgi_build_mips(vgi_context* ctx, vgi_scene* s, const vgi_lights* lights, uint32_t precision,
                   void* stream) {
  if (!ctx || !s) return fail(VGI_E_INVALID, "NULL context or scene");
  if (s->ctx != ctx) return fail(VGI_E_INVALID, "scene belongs to another context");
  if (!s->has_voxels) return fail(VGI_E_NOT_READY, "scene has not been voxelized");
  if (precision > VGI_PREC_BF16) return fail(VGI_E_INVALID, "precision %u", precision);
  if (lights && lights->count > VGI_MAX_LIGHTS)
    return fail(VGI_E_RANGE, "%u lights exceed %u", lights->count, VGI_MAX_LIGHTS);
  HIP_TRY(hipSetDevice(ctx->device));
  hipStream_t st = pick(ctx, stream);
  const vgi_grid& g = s->grid;
  if (s->alloc_n != g.n || s->alloc_levels != g.levels || s->alloc_prec != precision) {
    HIP_TRY(hipStreamSynchronize(st));
    for (uint32_t k = 0; k < VGI_MAX_LEVELS; k++) {
      if (s->level[k]) (void)hipFree(s->level[k]);
      s->level[k] = nullptr;
      s->level_bytes[k] = 0;
    }
    s->alloc_n = s->alloc_levels = 0;
    for (uint32_t k = 0; k < g.levels; k++) {
      const size_t p = (size_t)(g.n >> k) + 2;
      const size_t bytes = p * p * p * texel_bytes(precision);
      hipError_t e = hipMalloc(&s->level[k], bytes);
      if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(VGI_E_NOMEM, "mip level %u allocation (%zu bytes) failed", k, bytes);
      }
      s->level_bytes[k] = bytes;
      HIP_TRY(hipMemset(s->level[k], 0, bytes));  // zero border, written once
    }
    s->alloc_n = g.n;
    s->alloc_levels = g.levels;
    s->alloc_prec = precision;
  }
  MipParams P{};
  P.acc = s->acc;
  P.mask = s->mask;
  P.bin_packed = s->bin_packed;
  P.lut = ctx->d_lut;
  for (uint32_t k = 0; k < g.levels; k++) {
    P.level[k] = s->level[k];
    P.nk[k] = g.n >> k;
  }
  P.n = g.n;
  P.levels = g.levels;
  P.bin_shift = s->bin_shift;
  P.nbins_axis = s->nbins_axis;
  P.prec = precision;
  for (int a = 0; a < 3; a++) P.origin[a] = g.origin[a];
  P.h = g.extent / (float)g.n;
  P.nlights = lights ? lights->count : 0;
  for (uint32_t l = 0; l < P.nlights; l++) {
    const vgi_light& L = lights->light[l];
    if (L.type > VGI_LIGHT_POINT) return fail(VGI_E_INVALID, "light %u: type %u", l, L.type);
    P.lights[l].type = L.type;
    for (int a = 0; a < 3; a++) {
      P.lights[l].vec[a] = L.vec[a];
      P.lights[l].color[a] = L.color[a];
    }
    P.lights[l].min_distance = L.min_distance;
  }
  const bool prof = ctx->profiling != 0;
  if (prof) HIP_TRY(hipEventRecord(ctx->ev[5], st));
  uint32_t top = 0;
  HIP_TRY(launch_mips(P, st, &top));
  if (prof) {
    HIP_TRY(hipEventRecord(ctx->ev[6], st));
    ctx->ev_mips = true;
  }
  s->prec = precision;
  s->has_mips = true;
  return VGI_OK;
}

int vgi_trace(vgi_context* ctx, const vgi_scene* s, const vgi_gbuffer* gb, const vgi_cone_params* c,
              float* out_diffuse, float* out_specular, void* stream) {
  if (!ctx || !s || !gb || !c) return fail(VGI_E_INVALID, "NULL argument");
  if (s->ctx != ctx) return fail(VGI_E_INVALID, "scene belongs to another context");
  if (!s->has_mips) return fail(VGI_E_NOT_READY, "scene has no mip chain (call vgi_build_mips)");
  if (!gb->position || !gb->normal || !out_diffuse) return fail(VGI_E_INVALID, "G-buffer/output pointer is NULL");
  if (c->num_diffuse > VGI_MAX_DIFFUSE_CONES)
    return fail(VGI_E_RANGE, "num_diffuse %u exceeds %u", c->num_diffuse, VGI_MAX_DIFFUSE_CONES);
  if (!(c->step_factor > 0.0f) || !std::isfinite(c->max_distance))
    return fail(VGI_E_INVALID, "step_factor must be > 0 and max_distance finite");
  if ((uint64_t)gb->width * gb->height >= (1ull << 32)) return fail(VGI_E_RANGE, "G-buffer too large");
  HIP_TRY(hipSetDevice(ctx->device));
  hipStream_t st = pick(ctx, stream);
  vgi_tuning tn;
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    tn = ctx->tuning;
  }
  const vgi_grid& g = s->grid;
  TraceParams P{};
  for (uint32_t k = 0; k < g.levels; k++) {
    P.level[k] = s->level[k];
    P.level_bytes[k] = s->level_bytes[k];
    P.nk[k] = g.n >> k;
  }
  P.levels = g.levels;
  P.n = g.n;
  P.prec = s->prec;
  P.width = gb->width;
  P.height = gb->height;
  P.tile_w = tn.tile_w;
  P.tile_h = tn.tile_h;
  P.tiles_x = (gb->width + tn.tile_w - 1) / tn.tile_w;
  P.tiles_y = (gb->height + tn.tile_h - 1) / tn.tile_h;
  P.num_diffuse = c->num_diffuse;
  P.specular = c->specular;
  for (int a = 0; a < 3; a++) {
    P.origin[a] = g.origin[a];
    P.eye[a] = c->eye[a];
  }
  P.h = g.extent / (float)g.n;
  P.inv_h = (float)g.n / g.extent;
  const double PI = 3.14159265358979323846;
  P.tan_d = (float)std::tan((double)c->diffuse_aperture_deg * PI / 360.0);
  P.tan_s = (float)std::tan((double)c->specular_aperture_deg * PI / 360.0);
  P.step = c->step_factor;
  P.max_dist = c->max_distance;
  P.cutoff = c->alpha_cutoff;
  P.normal_off = c->normal_offset;
  P.start_off = c->start_offset;
  if (c->num_diffuse > 0) {
    const uint32_t K = c->num_diffuse;
    const double ring = (double)c->diffuse_ring_deg * PI / 180.0;
    const double wsum = 1.0 + (double)(K - 1) * std::cos(ring);
    for (uint32_t k = 0; k < K; k++) {
      if (k == 0) {
        P.dirs[0][0] = 0.f; P.dirs[0][1] = 0.f; P.dirs[0][2] = 1.f;
        P.w[0] = (float)(1.0 / wsum);
      } else {
        const double phi = 2.0 * PI * (double)(k - 1) / (double)(K - 1);
        P.dirs[k][0] = (float)(std::sin(ring) * std::cos(phi));
        P.dirs[k][1] = (float)(std::sin(ring) * std::sin(phi));
        P.dirs[k][2] = (float)std::cos(ring);
        P.w[k] = (float)(std::cos(ring) / wsum);
      }
    }
  }
  P.pos = gb->position;
  P.nrm = gb->normal;
  P.rough = gb->roughness;
  P.out_d = (float4*)out_diffuse;
  P.out_s = (float4*)out_specular;
  const bool prof = ctx->profiling != 0;
  P.counter = prof ? ctx->d_counter : nullptr;
  if (prof) {
    HIP_TRY(launch_fill_u32((uint32_t*)ctx->d_counter, 0u, 2, st));
    HIP_TRY(hipEventRecord(ctx->ev[7], st));
  }
  HIP_TRY(launch_trace(P, st));
  if (prof) {
    HIP_TRY(hipEventRecord(ctx->ev[8], st));
    ctx->ev_trace = true;
  }
  return VGI_OK;
}

int vgi_scene_read_voxels(vgi_context* ctx, const vgi_scene* s, uint32_t* host_out) {
  if (!ctx || !s || !host_out) return fail(VGI_E_INVALID, "NULL argument");
  if (!s->has_voxels) return fail(VGI_E_NOT_READY, "scene has not been voxelized");
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipDeviceSynchronize());
  const uint32_t n = s->grid.n, nb = s->nbins_axis, bs = s->bin_shift, nbr = n >> 3;
  std::vector<uint32_t> acc((size_t)n * n * n * kAccWords), mask((size_t)nbr * nbr * nbr * 16);
  std::vector<unsigned long long> packed((size_t)nb * nb * nb);
  HIP_TRY(hipMemcpy(acc.data(), s->acc, acc.size() * 4, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(mask.data(), s->mask, mask.size() * 4, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(packed.data(), s->bin_packed, packed.size() * 8, hipMemcpyDeviceToHost));
  for (uint32_t z = 0; z < n; z++)
    for (uint32_t y = 0; y < n; y++)
      for (uint32_t x = 0; x < n; x++) {
        uint32_t* o = host_out + ((size_t)x + (size_t)n * ((size_t)y + (size_t)n * z)) * kAccWords;
        const uint32_t bin = (x >> bs) + nb * ((y >> bs) + nb * (z >> bs));
        const uint32_t brick = (x >> 3) + nbr * ((y >> 3) + nbr * (z >> 3));
        const uint32_t local = (x & 7u) + 8u * ((y & 7u) + 8u * (z & 7u));
        const bool occ = (uint32_t)packed[bin] != 0 && ((mask[(size_t)brick * 16 + (local >> 5)] >> (local & 31u)) & 1u);
        if (!occ) {
          std::memset(o, 0, kAccWords * 4);
          continue;
        }
        std::memcpy(o, acc.data() + ((size_t)brick * 512 + local) * kAccWords, kAccWords * 4);
      }
  return VGI_OK;
}

int vgi_scene_read_level(vgi_context* ctx, const vgi_scene* s, uint32_t level, float* host_out) {
  if (!ctx || !s || !host_out) return fail(VGI_E_INVALID, "NULL argument");
  if (!s->has_mips) return fail(VGI_E_NOT_READY, "scene has no mip chain");
  if (level >= s->grid.levels) return fail(VGI_E_INVALID, "level %u >= %u", level, s->grid.levels);
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipDeviceSynchronize());
  const uint32_t nk = s->grid.n >> level;
  const size_t p1 = (size_t)nk + 2;
  std::vector<uint8_t> raw(s->level_bytes[level]);
  HIP_TRY(hipMemcpy(raw.data(), s->level[level], raw.size(), hipMemcpyDeviceToHost));
  for (uint32_t z = 0; z < nk; z++)
    for (uint32_t y = 0; y < nk; y++)
      for (uint32_t x = 0; x < nk; x++) {
        const size_t pi = (size_t)(x + 1) + p1 * ((size_t)(y + 1) + p1 * (size_t)(z + 1));
        float* o = host_out + ((size_t)x + (size_t)nk * ((size_t)y + (size_t)nk * z)) * 4;
        if (s->prec == VGI_PREC_FP32) {
          std::memcpy(o, raw.data() + pi * 16, 16);
        } else {
          uint16_t h[4];
          std::memcpy(h, raw.data() + pi * 8, 8);
          for (int c = 0; c < 3; c++) o[c] = s->prec == VGI_PREC_FP16 ? host_half(h[c]) : host_bf16(h[c]);
          o[3] = host_half(h[3]);
        }
      }
  return VGI_OK;
}

int vgi_malloc(vgi_context* ctx, void** out, size_t bytes) {
  if (!ctx || !out) return fail(VGI_E_INVALID, "NULL argument");
  HIP_TRY(hipSetDevice(ctx->device));
  *out = nullptr;
  hipError_t e = hipMalloc(out, bytes ? bytes : 1);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return fail(VGI_E_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  }
  return VGI_OK;
}

int vgi_free(vgi_context* ctx, void* ptr) {
  if (!ctx) return fail(VGI_E_INVALID, "ctx is NULL");
  if (!ptr) return VGI_OK;
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipFree(ptr));
  return VGI_OK;
}

int vgi_memcpy_h2d(vgi_context* ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx || (bytes && (!dst || !src))) return fail(VGI_E_INVALID, "NULL argument");
  HIP_TRY(hipSetDevice(ctx->device));
  if (bytes) HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return VGI_OK;
}

int vgi_memcpy_d2h(vgi_context* ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx || (bytes && (!dst || !src))) return fail(VGI_E_INVALID, "NULL argument");
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipDeviceSynchronize());
  if (bytes) HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return VGI_OK;
}

int vgi_memset(vgi_context* ctx, void* dst, int value, size_t bytes) {
  if (!ctx || (bytes && !dst)) return fail(VGI_E_INVALID, "NULL argument");
  HIP_TRY(hipSetDevice(ctx->device));
  if (bytes) HIP_TRY(hipMemset(dst, value, bytes));
  return VGI_OK;
}

int vgi_memcpy_d2d(vgi_context* ctx, void* dst, const void* src, size_t bytes, void* stream) {
  if (!ctx || (bytes && (!dst || !src))) return fail(VGI_E_INVALID, "NULL argument");
  if (bytes % 16) return fail(VGI_E_INVALID, "bytes (%zu) must be a multiple of 16", bytes);
  if (((uintptr_t)dst | (uintptr_t)src) & 15) return fail(VGI_E_INVALID, "pointers must be 16-byte aligned");
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(launch_copy(dst, src, bytes, pick(ctx, stream)));
  return VGI_OK;
}

int vgi_stream_create(vgi_context* ctx, void** out) {
  if (!ctx || !out) return fail(VGI_E_INVALID, "NULL argument");
  HIP_TRY(hipSetDevice(ctx->device));
  hipStream_t s = nullptr;
  HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *out = (void*)s;
  return VGI_OK;
}

int vgi_stream_destroy(vgi_context* ctx, void* stream) {
  if (!ctx) return fail(VGI_E_INVALID, "ctx is NULL");
  if (!stream) return VGI_OK;
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
  HIP_TRY(hipStreamDestroy((hipStream_t)stream));
  return VGI_OK;
}

int vgi_timer_start(vgi_context* ctx, void* stream) {
  if (!ctx) return fail(VGI_E_INVALID, "ctx is NULL");
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipEventRecord(ctx->tstart, pick(ctx, stream)));
  return VGI_OK;
}

int vgi_timer_stop(vgi_context* ctx, void* stream, float* out_ms) {
  if (!ctx || !out_ms) return fail(VGI_E_INVALID, "NULL argument");
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipEventRecord(ctx->tstop, pick(ctx, stream)));
  HIP_TRY(hipEventSynchronize(ctx->tstop));
  HIP_TRY(hipEventElapsedTime(out_ms, ctx->tstart, ctx->tstop));
  return VGI_OK;
}

}  // extern "C"
