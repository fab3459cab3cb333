// abi.hip — the C ABI of include/vgi.h: contexts, scenes (device memory owners),
// per-stream workspaces, the voxelize / mips / trace passes (single scenes and batches)
// and parity readback.
//
// Ownership follows the reference: CMegaPipeline owns the voxel image
// (Foreground/Renderer/MegaPipeline.cpp:470-490, MegaPipeline.h:60,72) and the passes
// borrow the caller's GPU buffers for the duration of a draw. Here a vgi_scene owns what
// stays resident between passes (accumulator, occupancy mask, bin table, mip levels);
// everything a voxelize call needs only while it runs (triangle records, bin counters,
// pair lists, the work map, scan scratch) and the staging of per-call descriptor tables
// belong to a workspace of the stream the call runs on, so calls on different streams
// never share scratch. Meshes, G-buffers and outputs are caller device pointers borrowed
// on the given stream. Errors never throw across the ABI: every entry returns a
// vgi_status and sets a thread-local message, replacing the reference's silent early
// returns (VoxelizeRenderer.cpp:49-50,87-88).
//
// A single-scene call is a batch of one: every pass builds its descriptor table(s) on the
// host, uploads them with one copy from a pinned staging slot, and launches the
// batch-native kernels.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "gltf.hpp"
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

#define TRY(expr)                \
  do {                           \
    int r_ = (expr);             \
    if (r_ != VGI_OK) return r_; \
  } while (0)

bool is_pow2(uint32_t v) { return v && !(v & (v - 1)); }
uint32_t ilog2(uint32_t v) {
  uint32_t r = 0;
  while ((1u << (r + 1)) <= v) r++;
  return r;
}
size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }
uint32_t ceil_div(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// Device allocation grown on demand (content is not preserved). The caller has made sure
// no queued work still reads the old block.
int grow_bytes(void** p, size_t* cap, size_t need, const char* what) {
  if (*p && *cap >= need) return VGI_OK;
  const size_t n = std::max(need, *cap + *cap / 2);
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  *cap = 0;
  hipError_t e = hipMalloc(p, std::max<size_t>(n, 256));
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    return fail(VGI_E_NOMEM, "device allocation of %zu bytes for %s failed: %s", n, what, hipGetErrorString(e));
  }
  *cap = n;
  return VGI_OK;
}

// Pinned staging for descriptor tables: a ring of slots per stream. A slot's host half is
// reused once the copy that read it has completed (its event); its device half is reused
// four calls later on the same stream, when the kernels that read it have run.
constexpr int kSlots = 4;
struct Slot {
  uint8_t* h = nullptr;
  uint8_t* d = nullptr;
  size_t cap = 0;
  hipEvent_t done = nullptr;
  bool pending = false;
  uint64_t gen = 0;  // changes whenever the slot is taken for new contents
};

// Call sites whose descriptor uploads are reused while unchanged.
enum UploadSite { kSiteVoxelize = 0, kSiteMips = 1, kSiteTrace = 2, kUploadSites = 3 };
struct UploadCache {
  int slot = -1;
  uint64_t gen = 0;
  size_t bytes = 0;
};

// Carves aligned sub-ranges out of one staging slot.
struct Carve {
  size_t off = 0;
  size_t take(size_t bytes) {
    const size_t o = off;
    off = align_up(off + bytes, 256);
    return o;
  }
};

}  // namespace

int vgi::set_error(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

// Per-stream workspace: voxelize scratch + staging ring. Calls on one stream are
// serialized by its mutex.
struct vgi_workspace {
  hipStream_t stream = nullptr;
  std::mutex mu;
  Slot slot[kSlots];
  uint64_t slot_gen = 0;
  UploadCache cache[kUploadSites];  // each site's last upload
  std::vector<uint8_t> staging;     // candidate descriptor bytes for the reuse check
  uint32_t next = 0;
  void* tris = nullptr;       // TriRec per triangle
  void* tritex = nullptr;     // TriTex per triangle (batches with a textured mesh)
  size_t tritex_bytes = 0;
  size_t tris_bytes = 0;
  void* tri_slot = nullptr;   // kSmallBins words per triangle
  size_t tri_slot_bytes = 0;
  void* big_list = nullptr;   // one word per triangle
  size_t big_list_bytes = 0;
  void* zeroed = nullptr;     // bin_small | bin_big | big_cursor | DevStats[] | totals[]
  size_t zeroed_bytes = 0;
  void* zero_list = nullptr;
  size_t zero_list_bytes = 0;
  void* scan_all = nullptr;
  size_t scan_bytes = 0;
  void* work_map = nullptr;
  size_t work_bytes = 0;
  void* pairs = nullptr;
  size_t pairs_bytes = 0;
  void* scan_tmp = nullptr;
  size_t scan_tmp_bytes = 0;
  void* sort_keys = nullptr;  // Morton order: keys in | keys out (u64)
  size_t sort_keys_bytes = 0;
  void* sort_pix = nullptr;   // pixel indices in | out (u32)
  size_t sort_pix_bytes = 0;
  void* sort_tmp = nullptr;
  size_t sort_tmp_bytes = 0;
  // pinned: the last call's batch pairs, work units, large triangles and split bins, then
  // its call number, which the settle step stores once the four totals are complete
  unsigned long long* hist = nullptr;
  uint32_t seq = 0;        // number of the last voxelize call on this stream (never 0)
  uint32_t hist_read = 0;  // the call whose totals were taken into the maxima below
  uint64_t hist_pairs = 0, hist_units = 0, hist_big = 0, hist_split = 0;  // largest seen

  size_t device_bytes() const {
    size_t b = tris_bytes + tritex_bytes + tri_slot_bytes + big_list_bytes + zeroed_bytes + zero_list_bytes +
               scan_bytes + work_bytes + pairs_bytes + scan_tmp_bytes + sort_keys_bytes + sort_pix_bytes +
               sort_tmp_bytes;
    for (const Slot& s : slot) b += s.cap;
    return b;
  }
};

struct vgi_context {
  int device = 0;
  int num_cus = 256;
  hipStream_t stream = nullptr;
  vgi_tuning tuning{8, 64, 16, 16, VGI_ORDER_TILES, VGI_ADDRESS_AUTO};
  int profiling = 0;
  float* d_lut = nullptr;
  unsigned long long* d_counter = nullptr;
  unsigned long long* h_counter = nullptr;
  hipEvent_t ev[10] = {};
  bool ev_vox = false, ev_mips = false, ev_trace = false;
  hipEvent_t tstart = nullptr, tstop = nullptr;
  hipEvent_t dep = nullptr;  // vgi_stream_wait (a wait captures the record it follows)
  std::atomic<uint32_t> vox_calls{0};  // voxelize call numbers, unique across the context's streams
  uint32_t scan_spin = kScanSpin;      // the bin scan's look-back limit (VGI_SCAN_SPIN, for tests)
  vgi_timings timings{};
  std::mutex mu;  // tuning, the workspace map
  std::map<hipStream_t, std::unique_ptr<vgi_workspace>> ws;
};

struct vgi_scene {
  vgi_context* ctx = nullptr;
  vgi_grid grid{};
  uint32_t bin_shift = 3;
  uint32_t nbins_axis = 0;
  uint32_t chunk = 64;
  uint32_t* acc = nullptr;  // n^3 x 8 words, brick-major
  uint32_t* mask = nullptr; // 16 words per 8^3 brick
  size_t acc_bytes = 0;
  size_t mask_bytes = 0;
  void* bin_packed = nullptr;  // work units << 32 | pairs per bin
  size_t bin_packed_bytes = 0;
  uint32_t prec = 0;
  void* level[VGI_MAX_LEVELS] = {};
  size_t level_bytes[VGI_MAX_LEVELS] = {};
  uint32_t* live = nullptr;  // per 16^3 tile of level 0: a word, then 8 row words (the mip kernel's skip state)
  size_t live_bytes = 0;
  uint32_t alloc_levels = 0, alloc_n = 0, alloc_prec = 0;
  // the injecting launch's tile-list counts are zero once the work queued so far has run
  // (the tile state was just cleared, or the last ordered build's later launch zeroed
  // them); otherwise the next ordered build clears them first
  bool tile_counts_zero = false;
  bool has_voxels = false, has_mips = false;
  uint64_t tris_in = 0;
  // the last voxelize call's statistics: written by the device into the pinned record after
  // the bin scan (`result`), its call number (vox_seq) stored last, so the record is
  // complete once the number matches; no event is recorded between the call's kernels
  SceneVoxResult* result = nullptr;
  uint32_t vox_seq = 0;
};

namespace {

// A voxelize call's number (never 0), unique across the context's streams: a scene's
// record matches its last call's number only once that call's settle step has stored it.
uint32_t next_vox_call(vgi_context* ctx) {
  uint32_t v = ctx->vox_calls.fetch_add(1u, std::memory_order_relaxed) + 1u;
  if (v == 0u) v = ctx->vox_calls.fetch_add(1u, std::memory_order_relaxed) + 1u;
  return v;
}

size_t scene_device_bytes(const vgi_scene* s) {
  size_t b = s->acc_bytes + s->mask_bytes + s->bin_packed_bytes + s->live_bytes;
  for (uint32_t k = 0; k < VGI_MAX_LEVELS; k++) b += s->level_bytes[k];
  return b;
}

void free_workspace(vgi_workspace* w) {
  void* dev[] = {w->tris,     w->tritex,   w->tri_slot, w->big_list, w->zeroed,    w->zero_list, w->scan_all,
                 w->work_map, w->pairs,    w->scan_tmp, w->sort_keys, w->sort_pix, w->sort_tmp};
  for (void* p : dev)
    if (p) (void)hipFree(p);
  for (Slot& s : w->slot) {
    if (s.d) (void)hipFree(s.d);
    if (s.h) (void)hipHostFree(s.h);
    if (s.done) (void)hipEventDestroy(s.done);
  }
  if (w->hist) (void)hipHostFree(w->hist);
}

hipStream_t pick(vgi_context* ctx, void* stream) { return stream ? (hipStream_t)stream : ctx->stream; }

// The workspace of a stream, created on first use.
vgi_workspace* workspace(vgi_context* ctx, hipStream_t st) {
  std::lock_guard<std::mutex> lk(ctx->mu);
  auto it = ctx->ws.find(st);
  if (it != ctx->ws.end()) return it->second.get();
  auto w = std::make_unique<vgi_workspace>();
  w->stream = st;
  for (Slot& s : w->slot)
    if (hipEventCreateWithFlags(&s.done, hipEventDisableTiming) != hipSuccess) {
      (void)hipGetLastError();
      free_workspace(w.get());
      return nullptr;
    }
  vgi_workspace* raw = w.get();
  ctx->ws.emplace(st, std::move(w));
  return raw;
}

vgi_tuning tuning_of(vgi_context* ctx) {
  std::lock_guard<std::mutex> lk(ctx->mu);
  return ctx->tuning;
}

// Takes the next staging slot with room for `bytes`; waits for the copy that last used it.
int acquire_slot(vgi_workspace* w, size_t bytes, Slot** out) {
  Slot& s = w->slot[w->next];
  w->next = (w->next + 1) % kSlots;
  s.gen = ++w->slot_gen;  // invalidates any reuse entry for this slot's old contents
  if (s.pending) {
    HIP_TRY(hipEventSynchronize(s.done));
    s.pending = false;
  }
  if (s.cap < bytes) {
    HIP_TRY(hipStreamSynchronize(w->stream));  // the device half may still be read
    if (s.d) (void)hipFree(s.d);
    if (s.h) (void)hipHostFree(s.h);
    s.d = nullptr;
    s.h = nullptr;
    s.cap = 0;
    const size_t cap = std::max<size_t>(align_up(bytes + bytes / 2, 4096), 65536);
    if (hipMalloc((void**)&s.d, cap) != hipSuccess ||
        hipHostMalloc((void**)&s.h, cap, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      return fail(VGI_E_NOMEM, "staging allocation of %zu bytes failed", cap);
    }
    s.cap = cap;
  }
  *out = &s;
  return VGI_OK;
}

int upload(vgi_workspace* w, Slot* s, size_t offset, size_t bytes) {
  if (bytes) HIP_TRY(hipMemcpyAsync(s->d + offset, s->h + offset, bytes, hipMemcpyHostToDevice, w->stream));
  HIP_TRY(hipEventRecord(s->done, w->stream));
  s->pending = true;
  return VGI_OK;
}

// A call's descriptor tables: fill(h, d) writes them into host memory h for a device copy
// at d (tables may point into themselves). If the site's last upload is still held (its
// slot not taken since) and the new bytes equal it, that device copy is used as it is and
// nothing is copied -- a repeated frame over unchanged scenes, lights and G-buffers uploads
// nothing. Otherwise a slot is taken and the tables are uploaded. Equal bytes mean the
// device already holds exactly what the upload would write, so reuse cannot change results.
template <class F>
int stage_descriptors(vgi_workspace* w, UploadSite site, size_t bytes, F&& fill, Slot** out) {
  UploadCache& c = w->cache[site];
  if (c.slot >= 0 && c.bytes == bytes && w->slot[c.slot].gen == c.gen) {
    Slot& s = w->slot[c.slot];
    w->staging.assign(bytes, 0);
    fill(w->staging.data(), s.d);
    if (std::memcmp(w->staging.data(), s.h, bytes) == 0) {
      *out = &s;
      return VGI_OK;
    }
  }
  Slot* sl = nullptr;
  TRY(acquire_slot(w, bytes, &sl));
  std::memset(sl->h, 0, bytes);
  fill(sl->h, sl->d);
  TRY(upload(w, sl, 0, bytes));
  c.slot = (int)(sl - w->slot);
  c.gen = sl->gen;
  c.bytes = bytes;
  *out = sl;
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

bool mesh_textured(const vgi_mesh& m) { return m.texcoords != nullptr && m.texture_rgba8 != nullptr; }

// The device form of a mesh; tri_base = its first triangle within its scene.
MeshDev mesh_dev(const vgi_mesh& m, uint32_t tri_base) {
  MeshDev d;
  std::memset(&d, 0, sizeof(d));
  d.pos = m.positions;
  d.idx = m.indices;
  d.albedo = m.albedo_rgb8;
  d.stride = m.position_stride ? m.position_stride : 12u;
  d.nverts = m.num_vertices;
  d.idx_type = m.index_type;
  d.ntri = m.num_triangles;
  d.tri_base = tri_base;
  d.base_rgb = (uint32_t)m.base_color[0] | ((uint32_t)m.base_color[1] << 8) | ((uint32_t)m.base_color[2] << 16);
  std::memcpy(d.m, m.model, sizeof(d.m));
  if (mesh_textured(m)) {
    d.uv = m.texcoords;
    d.tex = (const uint32_t*)m.texture_rgba8;
    d.uv_stride = m.texcoord_stride ? m.texcoord_stride : 8u;
    d.tw = m.texture_width;
    d.th = m.texture_height;
    d.textured = 1;
    std::memcpy(d.factor, m.base_color_factor, sizeof(d.factor));
  }
  return d;
}

int check_meshes(const vgi_mesh* meshes, uint32_t num_meshes, uint64_t* ntri_out) {
  if (num_meshes > 0 && !meshes) return fail(VGI_E_INVALID, "meshes is NULL");
  uint64_t ntri = 0;
  for (uint32_t i = 0; i < num_meshes; i++) {
    const vgi_mesh& m = meshes[i];
    if (m.num_triangles == 0) continue;
    if (!m.positions) return fail(VGI_E_INVALID, "mesh %u: positions is NULL", i);
    const uint32_t stride = m.position_stride ? m.position_stride : 12u;
    if (stride < 12 || (stride & 3)) return fail(VGI_E_INVALID, "mesh %u: position_stride %u", i, stride);
    if (m.index_type > VGI_INDEX_U32) return fail(VGI_E_INVALID, "mesh %u: index_type %u", i, m.index_type);
    if (m.index_type != VGI_INDEX_NONE && !m.indices) return fail(VGI_E_INVALID, "mesh %u: indices is NULL", i);
    if (m.index_type == VGI_INDEX_NONE && (uint64_t)m.num_triangles * 3 > m.num_vertices)
      return fail(VGI_E_INVALID, "mesh %u: non-indexed mesh needs 3 vertices per triangle", i);
    for (int k = 0; k < 12; k++)
      if (!std::isfinite(m.model[k])) return fail(VGI_E_INVALID, "mesh %u: model matrix not finite", i);
    if (mesh_textured(m)) {
      const uint32_t us = m.texcoord_stride ? m.texcoord_stride : 8u;
      if (us < 8 || (us & 3)) return fail(VGI_E_INVALID, "mesh %u: texcoord_stride %u", i, us);
      if (m.texture_width == 0 || m.texture_height == 0 || m.texture_width > 16384 || m.texture_height > 16384)
        return fail(VGI_E_INVALID, "mesh %u: texture %ux%u (1..16384 per side)", i, m.texture_width,
                    m.texture_height);
      if (((uintptr_t)m.texture_rgba8 & 3) || ((uintptr_t)m.texcoords & 3))
        return fail(VGI_E_INVALID, "mesh %u: texture and texcoords must be 4-byte aligned", i);
      for (int k = 0; k < 4; k++)
        if (!std::isfinite(m.base_color_factor[k]))
          return fail(VGI_E_INVALID, "mesh %u: base_color_factor not finite", i);
    }
    ntri += m.num_triangles;
  }
  if (ntri >= (1ull << 31)) return fail(VGI_E_RANGE, "%llu triangles exceed 2^31", (unsigned long long)ntri);
  *ntri_out = ntri;
  return VGI_OK;
}

// Resident allocations of a voxelized scene (accumulator, mask, bin table).
int scene_reserve(vgi_scene* s, uint32_t n, uint32_t nbins) {
  const size_t acc_bytes = (size_t)n * n * n * kAccWords * 4;
  const size_t bricks = (size_t)(n / 8) * (n / 8) * (n / 8);
  if (s->acc_bytes != acc_bytes) {
    if (s->acc || s->mask) HIP_TRY(hipDeviceSynchronize());
    if (s->acc) (void)hipFree(s->acc);
    if (s->mask) (void)hipFree(s->mask);
    s->acc = nullptr;
    s->mask = nullptr;
    s->acc_bytes = s->mask_bytes = 0;
    if (hipMalloc((void**)&s->acc, acc_bytes) != hipSuccess ||
        hipMalloc((void**)&s->mask, bricks * 16 * 4) != hipSuccess) {
      (void)hipGetLastError();
      return fail(VGI_E_NOMEM, "accumulator allocation (%zu bytes) failed", acc_bytes);
    }
    s->acc_bytes = acc_bytes;
    s->mask_bytes = bricks * 16 * 4;
  }
  if (s->bin_packed_bytes < (size_t)nbins * 8) {
    if (s->bin_packed) {
      HIP_TRY(hipDeviceSynchronize());
      (void)hipFree(s->bin_packed);
    }
    s->bin_packed = nullptr;
    s->bin_packed_bytes = 0;
    if (hipMalloc(&s->bin_packed, (size_t)nbins * 8) != hipSuccess) {
      (void)hipGetLastError();
      return fail(VGI_E_NOMEM, "bin table allocation failed");
    }
    s->bin_packed_bytes = (size_t)nbins * 8;
  }
  return VGI_OK;
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

bool distinct_scenes(std::vector<const vgi_scene*> v) {
  std::sort(v.begin(), v.end());
  return std::adjacent_find(v.begin(), v.end()) == v.end();
}

// ---------------------------------------------------------------------------
// voxelize batch
// ---------------------------------------------------------------------------

struct VoxPlan {
  vgi_scene* s;
  const vgi_voxelize_item* it;
  uint64_t ntri;
  uint32_t n, bin_shift, nbins_axis, nbins;
  size_t bin_base, tri_off, mesh_off;
  uint32_t tri_blocks, bin_blocks;
};

// The kernels of one voxelize call once its descriptors and scratch are in place: B holds
// their device addresses, the zeroed_bytes at zb (bin counters, per-scene statistics, the
// overflow word) are cleared first; the first wavefront of k_scatter settles the call
// (statistics, overflow word). Everything after the scan reads its sizes on the device:
// k_scatter's role blocks are sized from the host's counts (triangle and bin blocks) and
// persistent for the
// large triangles and the split bins, whose counts only the device knows; k_voxelize takes
// one work unit per workgroup up to grid_blocks and loops over the rest.
int launch_voxelize_kernels(vgi_context* ctx, const VoxBatch& B, uint32_t* zb, size_t zeroed_bytes, void* scan_tmp,
                            size_t scan_tmp_bytes, size_t nblk_t, size_t nblk_b, uint32_t ts, uint32_t grid_blocks,
                            uint32_t big_blocks, uint32_t zero_blocks, vgi_scene* const* scenes, uint32_t count,
                            hipStream_t st, bool prof) {
  const uint32_t cus = (uint32_t)ctx->num_cus;
  const uint32_t persistent = cus * 4;
  if (prof) HIP_TRY(hipEventRecord(ctx->ev[0], st));
  // bin_small, bin_big, big_cursor, the per-scene counters and the overflow word start at zero
  HIP_TRY(launch_fill_u32(zb, 0u, zeroed_bytes / 4, st));
  HIP_TRY(launch_setup(B, (uint32_t)nblk_t, st));
  HIP_TRY(launch_big_count(B, persistent, st));
  if (prof) HIP_TRY(hipEventRecord(ctx->ev[1], st));
  HIP_TRY(launch_bins(B, (uint32_t)nblk_b, st));
  for (uint32_t i = 0; i < count; i++) scenes[i]->vox_seq = B.seq;  // k_scatter stores it
  if (prof) HIP_TRY(hipEventRecord(ctx->ev[2], st));
  HIP_TRY(launch_scatter(B, (uint32_t)nblk_b, (uint32_t)nblk_t, big_blocks, zero_blocks, st));
  if (prof) HIP_TRY(hipEventRecord(ctx->ev[3], st));
  HIP_TRY(launch_voxelize(B, ts, grid_blocks, st));
  HIP_TRY(launch_voxelize_fallback(B, cus, st));
  if (prof) {
    HIP_TRY(hipEventRecord(ctx->ev[4], st));
    ctx->ev_vox = true;
  }
  return VGI_OK;
}

int voxelize_batch(vgi_context* ctx, const vgi_voxelize_item* items, uint32_t count, void* stream) {
  if (!ctx) return fail(VGI_E_INVALID, "ctx is NULL");
  if (count == 0) return VGI_OK;
  if (!items) return fail(VGI_E_INVALID, "items is NULL");
  const vgi_tuning tn = tuning_of(ctx);
  std::vector<VoxPlan> plan(count);
  std::vector<const vgi_scene*> seen(count);
  size_t total_bins = 0, total_tri = 0, total_mesh = 0, nblk_t = 0, nblk_b = 0;
  bool textured = false;  // some mesh is textured: per-triangle texcoord records, sampling kernels
  uint32_t ts = 16;
  for (uint32_t i = 0; i < count; i++) {
    const vgi_voxelize_item& it = items[i];
    if (!it.scene) return fail(VGI_E_INVALID, "item %u: scene is NULL", i);
    if (it.scene->ctx != ctx) return fail(VGI_E_INVALID, "item %u: scene belongs to another context", i);
    TRY(check_grid(&it.grid));
    VoxPlan& p = plan[i];
    p.s = it.scene;
    p.it = &it;
    TRY(check_meshes(it.meshes, it.num_meshes, &p.ntri));
    p.n = it.grid.n;
    p.bin_shift = ilog2(std::min(tn.bin_size, p.n));
    p.nbins_axis = p.n >> p.bin_shift;
    p.nbins = p.nbins_axis * p.nbins_axis * p.nbins_axis;
    p.bin_base = total_bins;
    p.tri_off = total_tri;
    p.mesh_off = total_mesh;
    p.tri_blocks = ceil_div(p.ntri, kBlockTris);
    p.bin_blocks = ceil_div(p.nbins, kBlockBins);
    total_bins += p.nbins;
    total_tri += p.ntri;
    total_mesh += it.num_meshes;
    for (uint32_t k = 0; k < it.num_meshes; k++)
      textured = textured || (it.meshes[k].num_triangles > 0 && mesh_textured(it.meshes[k]));
    nblk_t += p.tri_blocks;
    nblk_b += p.bin_blocks;
    ts = std::min(ts, 1u << p.bin_shift);
    seen[i] = it.scene;
  }
  if (!distinct_scenes(seen)) return fail(VGI_E_INVALID, "a scene appears twice in one batch");
  if (total_bins + 1 >= (1ull << 32) || nblk_t >= (1ull << 31) || nblk_b >= (1ull << 31))
    return fail(VGI_E_RANGE, "batch too large (%zu bins, %zu triangles)", total_bins, total_tri);

  HIP_TRY(hipSetDevice(ctx->device));
  const hipStream_t st = pick(ctx, stream);
  vgi_workspace* w = workspace(ctx, st);
  if (!w) return fail(VGI_E_HIP, "workspace creation failed");
  std::lock_guard<std::mutex> wl(w->mu);

  for (VoxPlan& p : plan) {
    p.s->has_voxels = false;
    p.s->has_mips = false;
    TRY(scene_reserve(p.s, p.n, p.nbins));
  }

  // ---- history: the largest batch totals of earlier calls on this stream (read without
  // waiting, once the last call's settle step has stored them; a call still in flight is read
  // by a later one). No later call is enqueued yet, so nothing rewrites them meanwhile. ----
  if (w->hist && w->hist_read != w->seq && __atomic_load_n(&w->hist[4], __ATOMIC_ACQUIRE) == w->seq) {
    w->hist_pairs = std::max<uint64_t>(w->hist_pairs, w->hist[0]);
    w->hist_units = std::max<uint64_t>(w->hist_units, w->hist[1]);
    w->hist_big = std::max<uint64_t>(w->hist_big, w->hist[2]);
    w->hist_split = std::max<uint64_t>(w->hist_split, w->hist[3]);
    w->hist_read = w->seq;
  }

  // ---- scratch sized for this batch (grown; queued work on this stream is drained
  // first so nothing still reads a block being replaced) ----
  // device: the three per-bin counter arrays, then DevStats[count], then the overflow word
  // (all cleared per call)
  const size_t stats_off = align_up(3 * total_bins * 4, 256);
  const size_t stats_bytes = (size_t)count * sizeof(DevStats);
  const size_t ovf_off = align_up(stats_off + stats_bytes, 256);
  const size_t flags_off = ovf_off + 256;  // the bin scan's per-workgroup flags
  const size_t zeroed_need = flags_off + align_up(std::max<size_t>(nblk_b, 1) * 4, 256);
  // The pair list holds 4 pairs per triangle plus 64 Ki (a triangle in at most 8 bins is
  // the common case; large triangles add more), or 5/4 of the most an earlier call on this
  // stream needed. A call that needs more is still exact: k_scatter sees it and the
  // per-triangle path runs, and the next call is sized from it.
  const uint64_t pair_want = std::max<uint64_t>(4ull * total_tri + 65536, w->hist_pairs + w->hist_pairs / 4 + 4096);
  const uint32_t pair_cap = (uint32_t)std::min<uint64_t>(pair_want, 0xFFFFFFF0ull);
  // Work units: a bin with pairs has ceil(ceil(pairs / chunk) / unit_items) of them, at
  // most 1 + pairs / chunk, so total_bins + pair_cap / chunk bounds every batch that fits.
  const uint64_t work_cap64 = total_bins + pair_cap / std::max(1u, tn.chunk_tris) + 1;
  if (work_cap64 >= (1ull << 32)) return fail(VGI_E_RANGE, "batch too large (%zu bins)", total_bins);
  const uint32_t work_cap = (uint32_t)work_cap64;
  const size_t scan_need = std::max<size_t>(nblk_b, 1) * 2 * sizeof(unsigned long long);  // totals, prefixes
  const size_t tneed = std::max<size_t>(total_tri, 1);
  const bool grow = w->tris_bytes < tneed * sizeof(TriRec) || w->zeroed_bytes < zeroed_need ||
                    (textured && w->tritex_bytes < tneed * sizeof(TriTex)) ||
                    w->scan_bytes < (total_bins + 1) * 8 || w->scan_tmp_bytes < scan_need ||
                    w->pairs_bytes < (size_t)pair_cap * 4 || w->work_bytes < (size_t)work_cap * kWorkEntryBytes;
  if (grow) HIP_TRY(hipStreamSynchronize(st));
  TRY(grow_bytes(&w->pairs, &w->pairs_bytes, (size_t)pair_cap * 4, "bin pairs"));
  TRY(grow_bytes(&w->work_map, &w->work_bytes, (size_t)work_cap * kWorkEntryBytes, "work map"));
  TRY(grow_bytes(&w->tris, &w->tris_bytes, tneed * sizeof(TriRec), "triangle records"));
  if (textured) TRY(grow_bytes(&w->tritex, &w->tritex_bytes, tneed * sizeof(TriTex), "texcoord records"));
  TRY(grow_bytes(&w->tri_slot, &w->tri_slot_bytes, tneed * kSmallBins * 4, "triangle slots"));
  TRY(grow_bytes(&w->big_list, &w->big_list_bytes, tneed * 4, "large-triangle list"));
  TRY(grow_bytes(&w->zeroed, &w->zeroed_bytes, zeroed_need, "bin counters"));
  TRY(grow_bytes(&w->zero_list, &w->zero_list_bytes, std::max<size_t>(total_bins, 1) * 4, "bin list"));
  TRY(grow_bytes(&w->scan_all, &w->scan_bytes, (total_bins + 1) * 8, "bin scan"));
  TRY(grow_bytes(&w->scan_tmp, &w->scan_tmp_bytes, std::max<size_t>(scan_need, 1), "scan scratch"));
  if (!w->hist) {  // the settle step writes it with system-scope stores
    HIP_TRY(hipHostMalloc((void**)&w->hist, kHostHistWords * sizeof(unsigned long long), hipHostMallocCoherent));
    std::memset(w->hist, 0, kHostHistWords * sizeof(unsigned long long));
  }
  w->seq = next_vox_call(ctx);

  // ---- descriptor tables: scenes | meshes | triangle blocks | bin blocks ----
  const bool multi = count > 1;
  Carve cv;
  const size_t o_scenes = cv.take((size_t)count * sizeof(VoxScene));
  const size_t o_meshes = cv.take(std::max<size_t>(total_mesh, 1) * sizeof(MeshDev));
  const size_t o_tblk = cv.take(multi ? nblk_t * sizeof(uint2) : 0);
  const size_t o_bblk = cv.take(multi ? nblk_b * sizeof(uint2) : 0);
  uint32_t* zb = (uint32_t*)w->zeroed;
  DevStats* dstats = (DevStats*)((uint8_t*)w->zeroed + stats_off);
  auto fill = [&](uint8_t* h, const uint8_t* d) {
  VoxScene* hs = (VoxScene*)(h + o_scenes);
  MeshDev* hm = (MeshDev*)(h + o_meshes);
  uint2* htb = (uint2*)(h + o_tblk);
  uint2* hbb = (uint2*)(h + o_bblk);
  size_t tb = 0, bb = 0;
  for (uint32_t i = 0; i < count; i++) {
    const VoxPlan& p = plan[i];
    const vgi_voxelize_item& it = *p.it;
    uint32_t base = 0;
    for (uint32_t k = 0; k < it.num_meshes; k++) {
      hm[p.mesh_off + k] = mesh_dev(it.meshes[k], base);
      base += it.meshes[k].num_triangles;
    }
    VoxScene& S = hs[i];
    std::memset(&S, 0, sizeof(S));
    S.meshes = (const MeshDev*)(d + o_meshes) + p.mesh_off;
    S.tris = (TriRec*)w->tris + p.tri_off;
    S.tritex = textured ? (TriTex*)w->tritex + p.tri_off : nullptr;
    S.tri_slot = (uint32_t*)w->tri_slot + p.tri_off * kSmallBins;
    S.big_list = (uint32_t*)w->big_list + p.tri_off;
    S.bin_small = zb + p.bin_base;
    S.bin_big = zb + total_bins + p.bin_base;
    S.big_cursor = zb + 2 * total_bins + p.bin_base;
    S.zero_list = (uint32_t*)w->zero_list + p.bin_base;
    S.bin_packed = (unsigned long long*)p.s->bin_packed;
    S.acc = p.s->acc;
    S.mask = p.s->mask;
    S.stats = dstats + i;
    S.result = p.s->result;
    for (int a = 0; a < 3; a++) S.origin[a] = it.grid.origin[a];
    S.snap_scale = (float)(p.n * 256u) / it.grid.extent;
    S.nmesh = it.num_meshes;
    S.ntri = (uint32_t)p.ntri;
    S.n = p.n;
    S.bin_shift = p.bin_shift;
    S.nbins_axis = p.nbins_axis;
    S.chunk = tn.chunk_tris;
    S.bin_base = (uint32_t)p.bin_base;
    S.unit_items = (ts == 8 && p.bin_shift == 3) ? kUnitItems : 1u;
    if (multi) {
      for (uint32_t b = 0; b < p.tri_blocks; b++) htb[tb++] = make_uint2(i, b * kBlockTris);
      for (uint32_t b = 0; b < p.bin_blocks; b++) hbb[bb++] = make_uint2(i, b * kBlockBins);
    }
  }
  };
  Slot* sl = nullptr;
  TRY(stage_descriptors(w, kSiteVoxelize, cv.off, fill, &sl));

  VoxBatch B{};
  B.scenes = (const VoxScene*)(sl->d + o_scenes);
  B.tri_blocks = multi ? (const uint2*)(sl->d + o_tblk) : nullptr;
  B.bin_blocks = multi ? (const uint2*)(sl->d + o_bblk) : nullptr;
  B.scan_all = (unsigned long long*)w->scan_all;
  B.scan_flags = (uint32_t*)((uint8_t*)w->zeroed + flags_off);
  B.scan_vals = (unsigned long long*)w->scan_tmp;
  B.scan_spin = ctx->scan_spin;
  B.work_map = (uint2*)w->work_map;
  B.unit_order = (uint32_t*)(B.work_map + work_cap);
  B.pairs_all = (uint32_t*)w->pairs;
  B.pair_cap = pair_cap;
  B.work_cap = work_cap;
  B.overflow = (uint32_t*)((uint8_t*)w->zeroed + ovf_off);
  B.host_hist = w->hist;
  B.seq = w->seq;
  B.nscenes = count;
  B.total_bins = (uint32_t)total_bins;
  B.textured = textured ? 1u : 0u;

  std::vector<vgi_scene*> scs(count);
  for (uint32_t i = 0; i < count; i++) scs[i] = plan[i].s;
  // launch sizes that only the device knows, from the largest this stream has seen:
  // voxelize work units (one per workgroup), large triangles (4 per workgroup, one per
  // wavefront) and split bins (one per workgroup); kernels loop over any excess
  const uint64_t cus = (uint64_t)ctx->num_cus;
  const uint64_t grid_want = w->hist_units ? w->hist_units + w->hist_units / 8 + 64 : cus * 8;
  const uint32_t grid_blocks = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(grid_want, work_cap));
  const uint32_t big_blocks =
      (uint32_t)std::min<uint64_t>(cus * 4, w->hist_big ? (w->hist_big + w->hist_big / 4 + 3) / 4 + 1 : cus);
  const uint32_t zero_blocks = (uint32_t)std::min<uint64_t>(
      std::min<uint64_t>(cus * 4, total_bins), w->hist_split ? w->hist_split + w->hist_split / 4 + 1 : cus / 4);
  TRY(launch_voxelize_kernels(ctx, B, zb, zeroed_need, w->scan_tmp, w->scan_tmp_bytes, nblk_t, nblk_b, ts,
                              grid_blocks, big_blocks, zero_blocks, scs.data(), count, st, ctx->profiling != 0));
  for (uint32_t i = 0; i < count; i++) {
    vgi_scene* s = plan[i].s;
    s->grid = items[i].grid;
    s->bin_shift = plan[i].bin_shift;
    s->nbins_axis = plan[i].nbins_axis;
    s->chunk = tn.chunk_tris;
    s->tris_in = plan[i].ntri;
    s->has_voxels = true;
  }
  return VGI_OK;
}

// ---------------------------------------------------------------------------
// mips batch
// ---------------------------------------------------------------------------

int scene_levels(vgi_scene* s, uint32_t precision, hipStream_t st) {
  const vgi_grid& g = s->grid;
  if (s->alloc_n == g.n && s->alloc_levels == g.levels && s->alloc_prec == precision) return VGI_OK;
  bool held = s->live != nullptr;
  for (uint32_t k = 0; k < VGI_MAX_LEVELS; k++) held = held || s->level[k];
  if (held) HIP_TRY(hipDeviceSynchronize());  // queued work may still read the old levels
  for (uint32_t k = 0; k < VGI_MAX_LEVELS; k++) {
    if (s->level[k]) (void)hipFree(s->level[k]);
    s->level[k] = nullptr;
    s->level_bytes[k] = 0;
  }
  if (s->live) (void)hipFree(s->live);
  s->live = nullptr;
  s->live_bytes = 0;
  s->alloc_n = s->alloc_levels = 0;
  {
    // The levels start all zero and a tile's texels (a level-0 x-row's) are stored only
    // while the tile (the row) or its previous build held occupied voxels, so every texel
    // of the levels stays exact: one word per tile, then eight row words per tile; then
    // the injecting launch's tile list (a word per tile) and its two counts.
    const size_t tpa = g.n / std::min<uint32_t>(g.n, 16u);
    const size_t bytes = tpa * tpa * tpa * 10 * sizeof(uint32_t) + 256;
    hipError_t e = hipMalloc((void**)&s->live, bytes);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      s->live = nullptr;
      return fail(VGI_E_NOMEM, "tile state allocation (%zu bytes) failed", bytes);
    }
    s->live_bytes = bytes;
    HIP_TRY(launch_fill_u32(s->live, 0u, bytes / 4, st));
    s->tile_counts_zero = true;
  }
  for (uint32_t k = 0; k < g.levels; k++) {
    const size_t p = (size_t)(g.n >> k) + 2;
    const size_t bytes = p * p * p * texel_bytes(precision);
    hipError_t e = hipMalloc(&s->level[k], bytes);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      return fail(VGI_E_NOMEM, "mip level %u allocation (%zu bytes) failed", k, bytes);
    }
    s->level_bytes[k] = bytes;
    HIP_TRY(launch_fill_u32((uint32_t*)s->level[k], 0u, bytes / 4, st));  // zero border, written once
  }
  s->alloc_n = g.n;
  s->alloc_levels = g.levels;
  s->alloc_prec = precision;
  return VGI_OK;
}

int mips_batch(vgi_context* ctx, const vgi_mips_item* items, uint32_t count, uint32_t precision, void* stream) {
  if (!ctx) return fail(VGI_E_INVALID, "ctx is NULL");
  if (count == 0) return VGI_OK;
  if (!items) return fail(VGI_E_INVALID, "items is NULL");
  if (precision > VGI_PREC_FP16X2) return fail(VGI_E_INVALID, "precision %u", precision);
  std::vector<const vgi_scene*> seen(count);
  for (uint32_t i = 0; i < count; i++) {
    const vgi_mips_item& it = items[i];
    if (!it.scene) return fail(VGI_E_INVALID, "item %u: scene is NULL", i);
    if (it.scene->ctx != ctx) return fail(VGI_E_INVALID, "item %u: scene belongs to another context", i);
    if (!it.scene->has_voxels) return fail(VGI_E_NOT_READY, "item %u: scene has not been voxelized", i);
    if (it.lights) {
      if (it.lights->count > VGI_MAX_LIGHTS)
        return fail(VGI_E_RANGE, "item %u: %u lights exceed %u", i, it.lights->count, VGI_MAX_LIGHTS);
      for (uint32_t l = 0; l < it.lights->count; l++)
        if ((it.lights->light[l].type & ~VGI_LIGHT_SHADOW) > VGI_LIGHT_POINT)
          return fail(VGI_E_INVALID, "item %u: light %u: type %#x", i, l, it.lights->light[l].type);
    }
    seen[i] = it.scene;
  }
  if (!distinct_scenes(seen)) return fail(VGI_E_INVALID, "a scene appears twice in one batch");
  HIP_TRY(hipSetDevice(ctx->device));
  const hipStream_t st = pick(ctx, stream);
  vgi_workspace* w = workspace(ctx, st);
  if (!w) return fail(VGI_E_HIP, "workspace creation failed");
  std::lock_guard<std::mutex> wl(w->mu);
  for (uint32_t i = 0; i < count; i++) TRY(scene_levels(items[i].scene, precision, st));

  // Stage plan: stage 0 resolves the accumulator and builds up to four levels per 16^3
  // tile; each later stage continues every unfinished chain from its last level.
  std::vector<std::vector<MipBlock>> stages;
  for (uint32_t i = 0; i < count; i++) {
    const vgi_grid& g = items[i].scene->grid;
    uint32_t src = 0;
    for (uint32_t j = 0;; j++) {
      const uint32_t ns = g.n >> src;
      const uint32_t S = std::min(ns, 16u);
      const uint32_t nout = std::min(ilog2(S), g.levels - 1 - src);
      if (j > 0 && nout == 0) break;
      if (stages.size() <= j) stages.emplace_back();
      const uint32_t tpa = ns / S;
      for (uint32_t t = 0; t < tpa * tpa * tpa; t++) stages[j].push_back(MipBlock{i, t, src, nout});
      src += nout;
      if (src >= g.levels - 1 || nout == 0) break;
    }
  }
  // one scene: workgroup b of a stage is tile b, so no block table is uploaded
  const bool table = count > 1;
  size_t nblocks = 0;
  for (const auto& v : stages) nblocks += v.size();
  size_t total_lights = 0;
  bool shadow = false;  // the injecting launch walks toward shadow-casting lights
  for (uint32_t i = 0; i < count; i++)
    if (items[i].lights) {
      total_lights += items[i].lights->count;
      for (uint32_t l = 0; l < items[i].lights->count; l++)
        shadow = shadow || (items[i].lights->light[l].type & VGI_LIGHT_SHADOW) != 0;
    }
  Carve cv;
  const size_t o_scenes = cv.take((size_t)count * sizeof(MipScene));
  const size_t o_blocks = cv.take(table ? nblocks * sizeof(MipBlock) : 0);
  const size_t o_lights = cv.take(std::max<size_t>(total_lights, 1) * sizeof(LightDev));
  auto fill = [&](uint8_t* h, const uint8_t* d) {
  MipScene* hs = (MipScene*)(h + o_scenes);
  LightDev* hl = (LightDev*)(h + o_lights);
  size_t lbase = 0;
  for (uint32_t i = 0; i < count; i++) {
    const vgi_scene* s = items[i].scene;
    const vgi_grid& g = s->grid;
    MipScene& P = hs[i];
    std::memset(&P, 0, sizeof(P));
    P.acc = s->acc;
    P.mask = s->mask;
    P.bin_packed = (const unsigned long long*)s->bin_packed;
    P.live = s->live;
    {
      const size_t tpa = g.n / std::min<uint32_t>(g.n, 16u);
      P.order = s->live + 9 * tpa * tpa * tpa;
    }
    for (uint32_t k = 0; k < g.levels; k++) {
      P.level[k] = s->level[k];
      P.nk[k] = g.n >> k;
    }
    P.n = g.n;
    P.levels = g.levels;
    P.bin_shift = s->bin_shift;
    P.nbins_axis = s->nbins_axis;
    for (int a = 0; a < 3; a++) P.origin[a] = g.origin[a];
    P.h = g.extent / (float)g.n;
    const vgi_lights* L = items[i].lights;
    P.nlights = L ? L->count : 0;
    P.lights = (const LightDev*)(d + o_lights) + lbase;
    for (uint32_t l = 0; l < P.nlights; l++) hl[lbase + l] = light_dev(L->light[l], P.origin, P.h);
    lbase += P.nlights;
  }
  MipBlock* hb = (MipBlock*)(h + o_blocks);
  size_t boff = 0;
  if (table)
    for (const auto& v : stages) {
      std::memcpy(hb + boff, v.data(), v.size() * sizeof(MipBlock));
      boff += v.size();
    }
  };
  Slot* sl = nullptr;
  TRY(stage_descriptors(w, kSiteMips, cv.off, fill, &sl));
  const bool prof = ctx->profiling != 0;
  if (prof) HIP_TRY(hipEventRecord(ctx->ev[5], st));
  size_t off = 0;
  uint32_t* list_counts = nullptr;  // an ordered build's tile-list counts
  for (size_t j = 0; j < stages.size(); j++) {
    MipBatch B{};
    B.scenes = (const MipScene*)(sl->d + o_scenes);
    B.blocks = table ? (const MipBlock*)(sl->d + o_blocks) + off : nullptr;
    B.lut = ctx->d_lut;
    B.src = stages[j][0].src;  // one scene: every block of the stage shares src and nout
    B.nout = stages[j][0].nout;
    B.shadow = shadow;
    // (not for a one-level chain: its launch writes every texel of every tile and keeps no
    // tile words, so every tile is visited; not with visibility walks, whose cost does not
    // follow the pair count: the room's shadowed builds were 3-5% slower in list order)
    B.ordered = j == 0 && !table && stages[0][0].nout > 0 && !shadow;
    if (B.ordered) {  // the injecting launch's tile list, dense tiles first
      vgi_scene* sc = items[0].scene;
      const vgi_grid& g = sc->grid;
      const size_t tpa = g.n / std::min<uint32_t>(g.n, 16u), ntl = tpa * tpa * tpa;
      list_counts = sc->live + 10 * ntl;
      if (!sc->tile_counts_zero) HIP_TRY(launch_fill_u32(list_counts, 0u, 2, st));
      sc->tile_counts_zero = false;
      HIP_TRY(launch_tiles(B, (uint32_t)ntl, st));
    }
    B.clear_counts = j == 1 ? list_counts : nullptr;  // the next build then needs no clearing launch
    HIP_TRY(launch_mips(B, precision, j == 0, (uint32_t)stages[j].size(), st));
    if (B.clear_counts) items[0].scene->tile_counts_zero = true;
    off += stages[j].size();
  }
  if (prof) {
    HIP_TRY(hipEventRecord(ctx->ev[6], st));
    ctx->ev_mips = true;
  }
  for (uint32_t i = 0; i < count; i++) {
    items[i].scene->prec = precision;
    items[i].scene->has_mips = true;
  }
  return VGI_OK;
}

// ---------------------------------------------------------------------------
// trace batch
// ---------------------------------------------------------------------------

int check_frame(const vgi_gbuffer& gb, const vgi_cone_params& c, const float* out_diffuse, uint32_t i) {
  if (!gb.position || !gb.normal || !out_diffuse)
    return fail(VGI_E_INVALID, "item %u: G-buffer/output pointer is NULL", i);
  if (c.num_diffuse > VGI_MAX_DIFFUSE_CONES)
    return fail(VGI_E_RANGE, "item %u: num_diffuse %u exceeds %u", i, c.num_diffuse, VGI_MAX_DIFFUSE_CONES);
  if (!(c.step_factor > 0.0f) || !std::isfinite(c.max_distance))
    return fail(VGI_E_INVALID, "item %u: step_factor must be > 0 and max_distance finite", i);
  if ((uint64_t)gb.width * gb.height >= (1ull << 32)) return fail(VGI_E_RANGE, "item %u: G-buffer too large", i);
  return VGI_OK;
}

int check_trace_item(vgi_context* ctx, const vgi_trace_item& it, uint32_t i) {
  const vgi_scene* s = it.scene;
  if (!s) return fail(VGI_E_INVALID, "item %u: scene is NULL", i);
  if (s->ctx != ctx) return fail(VGI_E_INVALID, "item %u: scene belongs to another context", i);
  if (!s->has_mips) return fail(VGI_E_NOT_READY, "item %u: scene has no mip chain (call vgi_build_mips)", i);
  return check_frame(it.gbuffer, it.cones, it.out_diffuse, i);
}

void fill_trace_params(TraceParams& P, const vgi_trace_item& it, const vgi_tuning& tn) {
  const vgi_scene* s = it.scene;
  const vgi_grid& g = s->grid;
  const vgi_gbuffer& gb = it.gbuffer;
  const vgi_cone_params& c = it.cones;
  std::memset(&P, 0, sizeof(P));
  for (uint32_t k = 0; k < g.levels; k++) {
    P.level[k] = s->level[k];
    P.level_bytes[k] = s->level_bytes[k];
    P.nk[k] = g.n >> k;
    P.lv[k].ptr = s->level[k];
    P.lv[k].bytes32 = (uint32_t)std::min<uint64_t>(s->level_bytes[k], 0xFFFFFFFFull);
    P.lv[k].p1 = (g.n >> k) + 2;
  }
  P.levels = g.levels;
  P.n = g.n;
  P.width = gb.width;
  P.height = gb.height;
  P.tiles_x = (gb.width + tn.tile_w - 1) / tn.tile_w;
  P.tiles_y = (gb.height + tn.tile_h - 1) / tn.tile_h;
  P.num_diffuse = c.num_diffuse;
  P.specular = c.specular;
  for (int a = 0; a < 3; a++) {
    P.origin[a] = g.origin[a];
    P.eye[a] = c.eye[a];
  }
  P.h = g.extent / (float)g.n;
  P.inv_h = (float)g.n / g.extent;
  const double PI = 3.14159265358979323846;
  P.tan_d = (float)std::tan((double)c.diffuse_aperture_deg * PI / 360.0);
  P.tan_s = (float)std::tan((double)c.specular_aperture_deg * PI / 360.0);
  P.step = c.step_factor;
  P.max_dist = c.max_distance;
  P.cutoff = c.alpha_cutoff;
  P.normal_off = c.normal_offset;
  P.start_off = c.start_offset;
  if (c.num_diffuse > 0) {
    const uint32_t K = c.num_diffuse;
    const double ring = (double)c.diffuse_ring_deg * PI / 180.0;
    const double wsum = 1.0 + (double)(K - 1) * std::cos(ring);
    for (uint32_t k = 0; k < K; k++) {
      if (k == 0) {
        P.dirs[0][0] = 0.f;
        P.dirs[0][1] = 0.f;
        P.dirs[0][2] = 1.f;
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
  P.pos = gb.position;
  P.nrm = gb.normal;
  P.rough = gb.roughness;
  P.out_d = (float4*)it.out_diffuse;
  P.out_s = (float4*)it.out_specular;
}

// (w, 0) as packed halves, w rounded to nearest even
static uint32_t half_weight(float w) {
  const _Float16 h = (_Float16)w;
  uint16_t b;
  std::memcpy(&b, &h, 2);
  return (uint32_t)b;
}

// k_trace deals its screen tiles to the XCDs in bands of this many tile rows, one XCD's
// band before the next XCD's (room, 8x8 tiles: FP16 x-pair trace 0.737 ms with one
// contiguous eighth per XCD, 0.710 with one-row bands, 0.715 tile by tile, 0.727 and 0.755
// with bands of 120 and 30 tiles). Runs of sorted pixels (Morton order) keep one
// contiguous eighth per XCD (band 0).
constexpr uint32_t kTraceBandRows = 1;

// March schedule of a cone whose sample distances do not depend on the pixel (a fixed
// aperture): the per-pixel recurrence of the cone march evaluated once, in FP32 -- diameter
// d = max(h, 2 t tan), level max(0, log2(d / h)), next t = t + step d -- while
// t <= max_distance, for at most VGI_MAX_MARCH_STEPS samples.
void march_schedule(const TraceParams& P, float tan_half, std::vector<SchedEntry>& out) {
  out.clear();
  const float top = (float)(P.levels - 1);
  float t = P.start_off * P.h;
  for (int it = 0; it < VGI_MAX_MARCH_STEPS && t <= P.max_dist; it++) {
    const float twice = 2.0f * t;
    const float diam = std::max(P.h, twice * tan_half);
    const float lvl = std::max(0.0f, std::log2(diam * P.inv_h));
    SchedEntry e;
    if (lvl >= top) {
      e.t = t;
      e.fr = 1.0f;
      e.gr = 0.0f;
      e.code = kSchedTop | (P.levels - 1);
    } else {
      const float fl = std::floor(lvl);
      e.t = t;
      e.fr = lvl - fl;
      e.gr = 1.0f - e.fr;
      e.code = (uint32_t)fl;
    }
    e.hfr = half_weight(e.fr);
    e.hgr = half_weight(e.gr);
    e.pad[0] = e.pad[1] = 0u;
    out.push_back(e);
    const float adv = P.step * diam;
    t = t + adv;
  }
}

int trace_batch(vgi_context* ctx, const vgi_trace_item* items, uint32_t count, void* stream) {
  if (!ctx) return fail(VGI_E_INVALID, "ctx is NULL");
  if (count == 0) return VGI_OK;
  if (!items) return fail(VGI_E_INVALID, "items is NULL");
  for (uint32_t i = 0; i < count; i++) TRY(check_trace_item(ctx, items[i], i));
  const vgi_tuning tn = tuning_of(ctx);
  const bool morton = tn.pixel_order == VGI_ORDER_MORTON;
  // one launch per storage precision
  constexpr uint32_t kPrecisions = VGI_PREC_FP16X2 + 1;
  std::vector<uint32_t> group[kPrecisions];
  for (uint32_t i = 0; i < count; i++) group[items[i].scene->prec].push_back(i);
  struct Launch {
    uint32_t prec, nscenes;
    size_t o_params, o_blocks, o_seg, o_sched, nblocks, entries, nsched;
    bool buffer_path;
    uint32_t band;  // work items an XCD takes in a row (k_trace's block order)
  };
  // per item: its parameters and the schedules of its fixed-aperture cones
  std::vector<TraceParams> tp(count);
  std::vector<std::vector<SchedEntry>> sched(2 * (size_t)count);
  for (uint32_t i = 0; i < count; i++) {
    fill_trace_params(tp[i], items[i], tn);
    if (tp[i].num_diffuse > 0) march_schedule(tp[i], tp[i].tan_d, sched[2 * i]);
    if (tp[i].specular && !tp[i].rough) march_schedule(tp[i], tp[i].tan_s, sched[2 * i + 1]);
  }
  std::vector<Launch> launches;
  Carve cv;
  size_t max_entries = 0;
  for (uint32_t p = 0; p < kPrecisions; p++) {
    if (group[p].empty()) continue;
    Launch L{p, (uint32_t)group[p].size(), 0, 0, 0, 0, 0, 0, 0, tn.address_mode == VGI_ADDRESS_AUTO, 0};
    // bands of kTraceBandRows rows of the first item's tiles
    L.band = morton ? 0u : kTraceBandRows * (uint32_t)ceil_div(items[group[p][0]].gbuffer.width, tn.tile_w);
    for (uint32_t i : group[p]) {
      L.nsched += sched[2 * i].size() + sched[2 * i + 1].size();
      const vgi_gbuffer& gb = items[i].gbuffer;
      if (morton) L.entries += align_up((size_t)gb.width * gb.height, kSortBlock);
      else L.nblocks += (size_t)ceil_div(gb.width, tn.tile_w) * ceil_div(gb.height, tn.tile_h);
      for (uint32_t k = 0; k < items[i].scene->grid.levels; k++)
        // buffer gathers: unsigned 32-bit byte offsets and a 32-bit record count
        L.buffer_path = L.buffer_path && items[i].scene->level_bytes[k] < 0xFFFFFFF0ull;
    }
    if (morton) L.nblocks = L.entries / kSortBlock;
    if (L.nblocks >= (1ull << 31) || L.entries >= (1ull << 32))
      return fail(VGI_E_RANGE, "trace batch too large (%zu workgroups)", L.nblocks);
    L.o_params = cv.take((size_t)L.nscenes * sizeof(TraceParams));
    L.o_blocks = cv.take(!morton && L.nscenes > 1 ? L.nblocks * sizeof(uint2) : 0);
    L.o_seg = cv.take(morton ? ((size_t)L.nscenes + 1) * sizeof(uint32_t) : 0);
    L.o_sched = cv.take(L.nsched * sizeof(SchedEntry));
    max_entries = std::max(max_entries, L.entries);
    launches.push_back(L);
  }
  HIP_TRY(hipSetDevice(ctx->device));
  const hipStream_t st = pick(ctx, stream);
  vgi_workspace* w = workspace(ctx, st);
  if (!w) return fail(VGI_E_HIP, "workspace creation failed");
  std::lock_guard<std::mutex> wl(w->mu);
  if (morton) {  // sort buffers for the largest launch (grown; queued work drained first)
    size_t tmp_need = 0;
    HIP_TRY(sort_pixels(nullptr, &tmp_need, nullptr, nullptr, nullptr, nullptr, (uint32_t)max_entries, 64, st));
    if (w->sort_keys_bytes < max_entries * 16 || w->sort_tmp_bytes < tmp_need) HIP_TRY(hipStreamSynchronize(st));
    TRY(grow_bytes(&w->sort_keys, &w->sort_keys_bytes, max_entries * 16, "sort keys"));
    TRY(grow_bytes(&w->sort_pix, &w->sort_pix_bytes, max_entries * 8, "sort pixels"));
    TRY(grow_bytes(&w->sort_tmp, &w->sort_tmp_bytes, std::max<size_t>(tmp_need, 1), "sort scratch"));
  }
  Slot* sl = nullptr;
  auto fill = [&](uint8_t* h, const uint8_t* d) {
  for (const Launch& L : launches) {
    TraceParams* hp = (TraceParams*)(h + L.o_params);
    uint2* hb = (uint2*)(h + L.o_blocks);
    uint32_t* hs = (uint32_t*)(h + L.o_seg);
    SchedEntry* hsch = (SchedEntry*)(h + L.o_sched);
    const SchedEntry* dsch = (const SchedEntry*)(d + L.o_sched);
    size_t b = 0, e = 0, ns = 0;
    for (uint32_t j = 0; j < L.nscenes; j++) {
      const uint32_t i = group[L.prec][j];
      const vgi_trace_item& it = items[i];
      hp[j] = tp[i];
      for (int c = 0; c < 2; c++) {
        const std::vector<SchedEntry>& v = sched[2 * i + c];
        if (!v.empty()) std::memcpy(hsch + ns, v.data(), v.size() * sizeof(SchedEntry));
        hp[j].sched[c] = dsch + ns;
        hp[j].sched_len[c] = (uint32_t)v.size();
        ns += v.size();
      }
      if (morton) {
        hs[j] = (uint32_t)e;
        e += align_up((size_t)it.gbuffer.width * it.gbuffer.height, kSortBlock);
      } else if (L.nscenes > 1) {
        const uint32_t nt = hp[j].tiles_x * hp[j].tiles_y;
        for (uint32_t t = 0; t < nt; t++) hb[b++] = make_uint2(j, t);
      }
    }
    if (morton) hs[L.nscenes] = (uint32_t)e;
  }
  };
  TRY(stage_descriptors(w, kSiteTrace, cv.off, fill, &sl));
  const bool prof = ctx->profiling != 0;
  if (prof) {
    HIP_TRY(launch_fill_u32((uint32_t*)ctx->d_counter, 0u, 2, st));
    HIP_TRY(hipEventRecord(ctx->ev[7], st));
  }
  for (const Launch& L : launches) {
    TraceBatch B{};
    B.scenes = (const TraceParams*)(sl->d + L.o_params);
    B.blocks = !morton && L.nscenes > 1 ? (const uint2*)(sl->d + L.o_blocks) : nullptr;
    B.counter = prof ? ctx->d_counter : nullptr;
    B.nblocks = (uint32_t)L.nblocks;
    B.tile_w = tn.tile_w;
    B.tile_h = tn.tile_h;
    B.prec = L.prec;
    B.nscenes = L.nscenes;
    B.band = L.band;
    if (morton && L.entries > 0) {
      unsigned long long* keys = (unsigned long long*)w->sort_keys;
      uint32_t* pix = (uint32_t*)w->sort_pix;
      B.seg = (const uint32_t*)(sl->d + L.o_seg);
      HIP_TRY(launch_morton_keys(B, keys, pix, (uint32_t)L.entries, st));
      const int end_bit = 32 + (int)ilog2(std::max<uint32_t>(L.nscenes - 1, 1)) + 1;
      HIP_TRY(sort_pixels(w->sort_tmp, &w->sort_tmp_bytes, keys, keys + max_entries, pix, pix + max_entries,
                          (uint32_t)L.entries, std::min(end_bit, 64), st));
      B.skeys = keys + max_entries;
      B.spix = pix + max_entries;
    }
    HIP_TRY(launch_trace(B, L.buffer_path, st));
  }
  if (prof) {
    HIP_TRY(hipEventRecord(ctx->ev[8], st));
    ctx->ev_trace = true;
  }
  return VGI_OK;
}

}  // namespace

int vgi::check_request_input(const vgi_mesh* meshes, uint32_t num_meshes, const vgi_grid* grid,
                             const vgi_lights* lights, const vgi_frame* frames, uint32_t num_frames) {
  TRY(check_grid(grid));
  uint64_t ntri = 0;
  TRY(check_meshes(meshes, num_meshes, &ntri));
  if (lights) {
    if (lights->count > VGI_MAX_LIGHTS) return fail(VGI_E_RANGE, "%u lights exceed %u", lights->count, VGI_MAX_LIGHTS);
    for (uint32_t l = 0; l < lights->count; l++)
      if ((lights->light[l].type & ~VGI_LIGHT_SHADOW) > VGI_LIGHT_POINT)
        return fail(VGI_E_INVALID, "light %u: type %#x", l, lights->light[l].type);
  }
  if (num_frames > 0 && !frames) return fail(VGI_E_INVALID, "frames is NULL");
  for (uint32_t f = 0; f < num_frames; f++) TRY(check_frame(frames[f].gbuffer, frames[f].cones, frames[f].out_diffuse, f));
  return VGI_OK;
}

extern "C" {

const char* vgi_version(void) { return "vgi 0.2.0 (gfx950)"; }
uint32_t vgi_abi_version(void) { return VGI_ABI_VERSION; }
const char* vgi_last_error(void) { return g_err.c_str(); }

// ---------------------------------------------------------------------------
// allocation-free voxelize: one static device array per GPU
// ---------------------------------------------------------------------------

// The scratch of vgi_voxelize_noalloc. Each GPU the library runs on holds its own copy
// (module globals are per device) until the process ends. In host code the array's name
// is a host-side stand-in, not a device address: the per-device address comes from
// hipGetSymbolAddress in noalloc_scratch_setup, and only that address is used.
static __device__ __align__(256) uint8_t scratch[VGI_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint8_t* g_scratch_dev[kMaxDevices] = {};  // scratch's device address, per device

// Once per device, with that device current (vgi_create).
int noalloc_scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return fail(VGI_E_RANGE, "device %d beyond %d", device, kMaxDevices);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_dev[device]) return VGI_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    (void)hipGetLastError();
    return fail(VGI_E_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
  }
  g_scratch_dev[device] = (uint8_t*)p;
  return VGI_OK;
}

uint8_t* noalloc_scratch(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return (device >= 0 && device < kMaxDevices) ? g_scratch_dev[device] : nullptr;
}
}  // namespace

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
  TRY(noalloc_scratch_setup(device));
  vgi_context* ctx = new vgi_context();
  ctx->device = device;
  ctx->num_cus = prop.multiProcessorCount;
  if (const char* e = std::getenv("VGI_SCAN_SPIN")) ctx->scan_spin = (uint32_t)std::strtoul(e, nullptr, 10);
  e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete ctx;
    return fail(VGI_E_HIP, "stream create: %s", hipGetErrorString(e));
  }
  float lut[256];
  for (int i = 0; i < 256; i++) lut[i] = (float)std::pow((double)i / 255.0, 2.2);
  if (hipMalloc((void**)&ctx->d_lut, sizeof(lut)) != hipSuccess ||
      hipMemcpy(ctx->d_lut, lut, sizeof(lut), hipMemcpyHostToDevice) != hipSuccess ||
      hipMalloc((void**)&ctx->d_counter, sizeof(unsigned long long)) != hipSuccess ||
      hipHostMalloc((void**)&ctx->h_counter, sizeof(unsigned long long), hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    vgi_destroy(ctx);
    return fail(VGI_E_NOMEM, "context allocation failed");
  }
  for (auto& ev : ctx->ev) HIP_TRY(hipEventCreate(&ev));
  HIP_TRY(hipEventCreate(&ctx->tstart));
  HIP_TRY(hipEventCreate(&ctx->tstop));
  HIP_TRY(hipEventCreateWithFlags(&ctx->dep, hipEventDisableTiming));
  *out = ctx;
  return VGI_OK;
}

int vgi_destroy(vgi_context* ctx) {
  if (!ctx) return VGI_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipDeviceSynchronize();
  for (auto& kv : ctx->ws) free_workspace(kv.second.get());
  ctx->ws.clear();
  if (ctx->d_lut) (void)hipFree(ctx->d_lut);
  if (ctx->d_counter) (void)hipFree(ctx->d_counter);
  if (ctx->h_counter) (void)hipHostFree(ctx->h_counter);
  for (auto& ev : ctx->ev)
    if (ev) (void)hipEventDestroy(ev);
  if (ctx->tstart) (void)hipEventDestroy(ctx->tstart);
  if (ctx->tstop) (void)hipEventDestroy(ctx->tstop);
  if (ctx->dep) (void)hipEventDestroy(ctx->dep);
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
  if (t->pixel_order > VGI_ORDER_MORTON) return fail(VGI_E_INVALID, "pixel_order %u", t->pixel_order);
  if (t->address_mode > VGI_ADDRESS_WIDE) return fail(VGI_E_INVALID, "address_mode %u", t->address_mode);
  std::lock_guard<std::mutex> lk(ctx->mu);
  ctx->tuning = *t;
  return VGI_OK;
}

int vgi_get_tuning(const vgi_context* ctx, vgi_tuning* t) {
  if (!ctx || !t) return fail(VGI_E_INVALID, "NULL argument");
  *t = tuning_of(const_cast<vgi_context*>(ctx));
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
    t.voxelize_sort_ms += a;  // bins + scan + host read of the pair totals
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
  *out = nullptr;
  HIP_TRY(hipSetDevice(ctx->device));
  vgi_scene* s = new vgi_scene();
  s->ctx = ctx;
  if (hipHostMalloc((void**)&s->result, sizeof(SceneVoxResult), hipHostMallocCoherent) != hipSuccess) {
    (void)hipGetLastError();
    delete s;
    return fail(VGI_E_NOMEM, "scene record allocation failed");
  }
  std::memset(s->result, 0, sizeof(SceneVoxResult));
  *out = s;
  return VGI_OK;
}

int vgi_scene_release(vgi_scene* s) {
  if (!s) return VGI_OK;
  (void)hipSetDevice(s->ctx->device);
  (void)hipDeviceSynchronize();
  void* dev[] = {s->acc, s->mask, s->bin_packed};
  for (void* p : dev)
    if (p) (void)hipFree(p);
  for (uint32_t k = 0; k < VGI_MAX_LEVELS; k++)
    if (s->level[k]) (void)hipFree(s->level[k]);
  if (s->live) (void)hipFree(s->live);
  if (s->result) (void)hipHostFree(s->result);
  delete s;
  return VGI_OK;
}

uint64_t vgi_scene_bytes_estimate(const vgi_grid* g, uint32_t precision, uint64_t ntri) {
  (void)ntri;
  if (!g || !is_pow2(g->n) || g->n < 8 || g->levels == 0 || g->levels > VGI_MAX_LEVELS) return 0;
  const uint64_t n = g->n;
  const uint64_t tpa = n / std::min<uint64_t>(n, 16);  // 16^3 tiles per axis (mip tile state)
  uint64_t b = n * n * n * kAccWords * 4 + (n / 8) * (n / 8) * (n / 8) * 64 + (n / 8) * (n / 8) * (n / 8) * 8 +
               tpa * tpa * tpa * 10 * 4 + 256;
  for (uint32_t k = 0; k < g->levels; k++) {
    const uint64_t p = (n >> k) + 2;
    b += p * p * p * texel_bytes(precision);
  }
  return b;
}

int vgi_scene_get_stats(const vgi_scene* s, vgi_scene_stats* out) {
  if (!s || !out) return fail(VGI_E_INVALID, "NULL argument");
  std::memset(out, 0, sizeof(*out));
  out->triangles_in = s->tris_in;
  if (s->vox_seq != 0) {  // the voxelize call's own statistics, once the device stored them
    const SceneVoxResult* r = s->result;  // the settle step stores its call number after the rest
    if (__atomic_load_n(&r->seq, __ATOMIC_ACQUIRE) != s->vox_seq) {
      // that call's settle step has not run yet: wait for the device's queued work
      HIP_TRY(hipSetDevice(s->ctx->device));
      HIP_TRY(hipDeviceSynchronize());
      if (__atomic_load_n(&r->seq, __ATOMIC_ACQUIRE) != s->vox_seq)
        return fail(VGI_E_HIP, "voxelize statistics of the scene's last call are missing");
    }
    out->triangles_rejected = r->st.rejected;
    out->triangles_degenerate = r->st.degenerate;
    out->bin_pairs = r->st.pairs;
    out->work_items = r->totals >> 32;
    out->voxelize_fallback = r->overflow;
  }
  out->device_bytes = scene_device_bytes(s);
  out->n = s->grid.n;
  out->levels = s->grid.levels;
  out->precision = s->prec;
  out->has_voxels = s->has_voxels;
  out->has_mips = s->has_mips;
  out->bin_size = 1u << s->bin_shift;
  return VGI_OK;
}

int vgi_voxelize(vgi_context* ctx, vgi_scene* s, const vgi_mesh* meshes, uint32_t num_meshes,
                 const vgi_grid* grid, void* stream) {
  if (!ctx || !s) return fail(VGI_E_INVALID, "NULL context or scene");
  TRY(check_grid(grid));
  vgi_voxelize_item it{s, meshes, num_meshes, *grid};
  return voxelize_batch(ctx, &it, 1, stream);
}

int vgi_voxelize_noalloc(vgi_context* ctx, vgi_scene* sc, const vgi_mesh* meshes, uint32_t num_meshes,
                         const vgi_grid* grid, void* stream) {
  if (!ctx || !sc) return fail(VGI_E_INVALID, "NULL context or scene");
  if (sc->ctx != ctx) return fail(VGI_E_INVALID, "scene belongs to another context");
  TRY(check_grid(grid));
  if (num_meshes > VGI_NOALLOC_MAX_MESHES)
    return fail(VGI_E_RANGE, "%u meshes exceed the allocation-free limit of %u", num_meshes, VGI_NOALLOC_MAX_MESHES);
  uint64_t ntri = 0;
  TRY(check_meshes(meshes, num_meshes, &ntri));
  const vgi_tuning tn = tuning_of(ctx);
  const uint32_t n = grid->n;
  const uint32_t bin_shift = ilog2(std::min(tn.bin_size, n));
  const uint32_t nba = n >> bin_shift;
  const uint64_t nbins = (uint64_t)nba * nba * nba;
  // the scene's resident volume must exist already: sizing it would allocate
  if (sc->acc_bytes != (size_t)n * n * n * kAccWords * 4 || sc->bin_packed_bytes < nbins * 8)
    return fail(VGI_E_STATE, "scene has no volume for a %u^3 grid with %u^3 bins: voxelize it once with vgi_voxelize",
                n, 1u << bin_shift);
  uint8_t* base = noalloc_scratch(ctx->device);
  if (!base) return fail(VGI_E_STATE, "no scratch for device %d", ctx->device);

  // ---- the call's scratch, carved from the array ----
  const uint64_t pair_cap64 = 4ull * ntri + 65536;
  const uint64_t work_cap64 = nbins + pair_cap64 / std::max(1u, tn.chunk_tris) + 1;
  const size_t stats_off = align_up(3 * nbins * 4, 256);
  const size_t ovf_off = align_up(stats_off + sizeof(DevStats), 256);
  const size_t nblk_b = ceil_div(nbins, kBlockBins);
  const size_t flags_off = ovf_off + 256;  // the bin scan's per-workgroup flags
  const size_t zeroed_bytes = flags_off + align_up(nblk_b * 4, 256);
  const size_t scan_need = nblk_b * 2 * sizeof(unsigned long long);  // totals, prefixes
  Carve cv;
  const size_t o_desc = cv.take(sizeof(NoAllocDesc));
  const size_t o_hist = cv.take(kHostHistWords * sizeof(unsigned long long));
  const size_t o_zero = cv.take(zeroed_bytes);
  bool textured = false;
  for (uint32_t k = 0; k < num_meshes; k++)
    textured = textured || (meshes[k].num_triangles > 0 && mesh_textured(meshes[k]));
  const size_t o_tris = cv.take(std::max<uint64_t>(ntri, 1) * sizeof(TriRec));
  const size_t o_tritex = cv.take(textured ? ntri * sizeof(TriTex) : 0);
  const size_t o_slot = cv.take(std::max<uint64_t>(ntri, 1) * kSmallBins * 4);
  const size_t o_big = cv.take(std::max<uint64_t>(ntri, 1) * 4);
  const size_t o_zlist = cv.take(nbins * 4);
  const size_t o_scan = cv.take((nbins + 1) * 8);
  const size_t o_stmp = cv.take(std::max<size_t>(scan_need, 1));
  const size_t o_pairs = cv.take(pair_cap64 * 4);
  const size_t o_work = cv.take(work_cap64 * kWorkEntryBytes);
  if (cv.off > (size_t)VGI_NOALLOC_SCRATCH_BYTES)
    return fail(VGI_E_RANGE,
                "voxelize of %llu triangles on a %u^3 grid needs %zu bytes of scratch; the allocation-free limit is %u",
                (unsigned long long)ntri, n, cv.off, (unsigned)VGI_NOALLOC_SCRATCH_BYTES);

  HIP_TRY(hipSetDevice(ctx->device));
  const hipStream_t st = pick(ctx, stream);
  sc->has_voxels = false;
  sc->has_mips = false;

  // ---- descriptors, by value into the scratch ----
  NoAllocDesc d;
  std::memset(&d, 0, sizeof(d));
  NoAllocDesc* dd = (NoAllocDesc*)(base + o_desc);
  uint32_t tbase = 0;
  for (uint32_t k = 0; k < num_meshes; k++) {
    d.meshes[k] = mesh_dev(meshes[k], tbase);
    tbase += meshes[k].num_triangles;
  }
  uint32_t* zb = (uint32_t*)(base + o_zero);
  VoxScene& S = d.scene;
  S.meshes = dd->meshes;
  S.tris = (TriRec*)(base + o_tris);
  S.tritex = textured ? (TriTex*)(base + o_tritex) : nullptr;
  S.tri_slot = (uint32_t*)(base + o_slot);
  S.big_list = (uint32_t*)(base + o_big);
  S.bin_small = zb;
  S.bin_big = zb + nbins;
  S.big_cursor = zb + 2 * nbins;
  S.zero_list = (uint32_t*)(base + o_zlist);
  S.bin_packed = (unsigned long long*)sc->bin_packed;
  S.acc = sc->acc;
  S.mask = sc->mask;
  S.stats = (DevStats*)(base + o_zero + stats_off);
  S.result = sc->result;
  for (int a = 0; a < 3; a++) S.origin[a] = grid->origin[a];
  S.snap_scale = (float)(n * 256u) / grid->extent;
  S.nmesh = num_meshes;
  S.ntri = (uint32_t)ntri;
  S.n = n;
  S.bin_shift = bin_shift;
  S.nbins_axis = nba;
  S.chunk = tn.chunk_tris;
  S.bin_base = 0;
  const uint32_t ts = std::min(16u, 1u << bin_shift);
  S.unit_items = (ts == 8 && bin_shift == 3) ? kUnitItems : 1u;
  HIP_TRY(launch_stage_desc(d, dd, st));

  VoxBatch B{};
  B.scenes = &dd->scene;
  B.tri_blocks = nullptr;
  B.bin_blocks = nullptr;
  B.scan_all = (unsigned long long*)(base + o_scan);
  B.scan_flags = (uint32_t*)(base + o_zero + flags_off);
  B.scan_vals = (unsigned long long*)(base + o_stmp);
  B.scan_spin = ctx->scan_spin;
  B.work_map = (uint2*)(base + o_work);
  B.unit_order = (uint32_t*)(B.work_map + work_cap64);
  B.pairs_all = (uint32_t*)(base + o_pairs);
  B.pair_cap = (uint32_t)pair_cap64;
  B.work_cap = (uint32_t)work_cap64;
  B.overflow = (uint32_t*)(base + o_zero + ovf_off);
  B.host_hist = (unsigned long long*)(base + o_hist);  // device memory: nothing reads it back
  B.seq = next_vox_call(ctx);
  B.nscenes = 1;
  B.total_bins = (uint32_t)nbins;
  B.textured = textured ? 1u : 0u;
  const uint32_t grid_blocks = (uint32_t)std::min<uint64_t>((uint64_t)ctx->num_cus * 8, work_cap64);
  const uint32_t cus = (uint32_t)ctx->num_cus;
  TRY(launch_voxelize_kernels(ctx, B, zb, zeroed_bytes, base + o_stmp, scan_need, ceil_div(ntri, kBlockTris),
                              ceil_div(nbins, kBlockBins), ts, grid_blocks, cus,
                              (uint32_t)std::min<uint64_t>(cus / 4, nbins), &sc, 1, st, false));
  sc->grid = *grid;
  sc->bin_shift = bin_shift;
  sc->nbins_axis = nba;
  sc->chunk = tn.chunk_tris;
  sc->tris_in = ntri;
  sc->has_voxels = true;
  return VGI_OK;
}

int vgi_voxelize_batch(vgi_context* ctx, const vgi_voxelize_item* items, uint32_t count, void* stream) {
  return voxelize_batch(ctx, items, count, stream);
}

int vgi_build_mips(vgi_context* ctx, vgi_scene* s, const vgi_lights* lights, uint32_t precision, void* stream) {
  if (!ctx || !s) return fail(VGI_E_INVALID, "NULL context or scene");
  vgi_mips_item it{s, lights};
  return mips_batch(ctx, &it, 1, precision, stream);
}

int vgi_build_mips_batch(vgi_context* ctx, const vgi_mips_item* items, uint32_t count, uint32_t precision,
                         void* stream) {
  return mips_batch(ctx, items, count, precision, stream);
}

int vgi_trace(vgi_context* ctx, const vgi_scene* s, const vgi_gbuffer* gb, const vgi_cone_params* c,
              float* out_diffuse, float* out_specular, void* stream) {
  if (!ctx || !s || !gb || !c) return fail(VGI_E_INVALID, "NULL argument");
  vgi_trace_item it{s, *gb, *c, out_diffuse, out_specular};
  return trace_batch(ctx, &it, 1, stream);
}

int vgi_trace_batch(vgi_context* ctx, const vgi_trace_item* items, uint32_t count, void* stream) {
  return trace_batch(ctx, items, count, stream);
}

// ---- glTF ingestion -----------------------------------------------------------

struct vgi_gltf {
  vgi::GltfScene scene;
  vgi_context* ctx = nullptr;               // device copies belong to this context's device
  std::vector<void*> dev_pos, dev_idx, dev_uv;  // per geometry (dev_uv null without texcoords)
  struct Texels {
    std::vector<uint8_t> rgba;              // decoded by the caller, RGBA8 row-major
    uint32_t w = 0, h = 0;
    void* dev = nullptr;                    // device copy (on ctx's device)
    bool dirty = false;                     // set since the last upload
  };
  std::vector<Texels> texels;               // per image
};

namespace {

int gltf_finish(int rc, const std::string& err, vgi::GltfScene& scene, vgi_gltf** out) {
  if (rc != VGI_OK) return fail(rc, "glTF: %s", err.c_str());
  auto* g = new (std::nothrow) vgi_gltf;
  if (!g) return fail(VGI_E_NOMEM, "out of host memory");
  g->scene = std::move(scene);
  g->texels.resize(g->scene.images.size());
  *out = g;
  return VGI_OK;
}

void gltf_free_device(vgi_gltf* g) {
  if (!g->ctx) return;
  // best effort on release: a failure here leaves nothing the caller could act on
  (void)hipSetDevice(g->ctx->device);
  for (void* p : g->dev_pos) (void)hipFree(p);
  for (void* p : g->dev_idx) (void)hipFree(p);
  for (void* p : g->dev_uv)
    if (p) (void)hipFree(p);
  for (auto& t : g->texels) {
    if (t.dev) (void)hipFree(t.dev);
    t.dev = nullptr;
    t.dirty = !t.rgba.empty();
  }
  g->dev_pos.clear();
  g->dev_idx.clear();
  g->dev_uv.clear();
  g->ctx = nullptr;
}

}  // namespace

int vgi_gltf_load_file(const char* path, uint32_t albedo_mode, vgi_gltf** out) {
  if (!path || !out) return fail(VGI_E_INVALID, "NULL argument");
  *out = nullptr;
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(VGI_E_INVALID, "glTF: cannot open %s", path);
  std::vector<uint8_t> data;
  uint8_t buf[1 << 16];
  size_t n;
  while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) data.insert(data.end(), buf, buf + n);
  std::fclose(f);
  std::string dir(path);
  const size_t slash = dir.find_last_of('/');
  dir = slash == std::string::npos ? std::string(".") : dir.substr(0, slash == 0 ? 1 : slash);
  vgi::GltfScene scene;
  std::string err;
  const int rc = vgi::gltf_parse(data.data(), data.size(), dir, albedo_mode, scene, err);
  return gltf_finish(rc, err, scene, out);
}

int vgi_gltf_load_memory(const void* data, size_t bytes, const char* base_dir, uint32_t albedo_mode,
                         vgi_gltf** out) {
  if (!data || !out) return fail(VGI_E_INVALID, "NULL argument");
  *out = nullptr;
  vgi::GltfScene scene;
  std::string err;
  const int rc = vgi::gltf_parse((const uint8_t*)data, bytes, base_dir ? base_dir : "", albedo_mode, scene, err);
  return gltf_finish(rc, err, scene, out);
}

int vgi_gltf_get_info(const vgi_gltf* g, vgi_gltf_info* out) {
  if (!g || !out) return fail(VGI_E_INVALID, "NULL argument");
  *out = vgi_gltf_info{};
  out->num_primitives = (uint32_t)g->scene.instances.size();
  out->num_geometries = (uint32_t)g->scene.geometry.size();
  out->skipped_primitives = g->scene.skipped_primitives;
  out->num_images = (uint32_t)g->scene.images.size();
  for (const auto& in : g->scene.instances) out->num_triangles += g->scene.geometry[in.geometry].indices.size() / 3;
  return VGI_OK;
}

int vgi_gltf_get_primitive(const vgi_gltf* g, uint32_t index, vgi_gltf_primitive* out) {
  if (!g || !out) return fail(VGI_E_INVALID, "NULL argument");
  if (index >= g->scene.instances.size())
    return fail(VGI_E_INVALID, "primitive %u >= %zu", index, g->scene.instances.size());
  const vgi::GltfInstance& in = g->scene.instances[index];
  const vgi::GltfGeometry& ge = g->scene.geometry[in.geometry];
  out->positions = ge.positions.data();
  out->indices = ge.indices.data();
  out->num_vertices = (uint32_t)(ge.positions.size() / 3);
  out->num_triangles = (uint32_t)(ge.indices.size() / 3);
  std::memcpy(out->model, in.model, sizeof(in.model));
  std::memcpy(out->base_color, in.base_color, 4);
  out->node = in.node;
  out->mesh = ge.mesh;
  out->primitive = ge.primitive;
  out->material = in.material;
  out->texcoords = ge.texcoords.empty() ? nullptr : ge.texcoords.data();
  out->image = in.image;
  out->reserved = 0;
  std::memcpy(out->base_color_factor, in.factor, sizeof(in.factor));
  return VGI_OK;
}

int vgi_gltf_get_image(vgi_gltf* g, uint32_t index, vgi_gltf_image* out) {
  if (!g || !out) return fail(VGI_E_INVALID, "NULL argument");
  if (index >= g->scene.images.size()) return fail(VGI_E_INVALID, "image %u >= %zu", index, g->scene.images.size());
  vgi::GltfImage& im = g->scene.images[index];
  std::string err;
  const int rc = vgi::gltf_load_image(g->scene, im, err);
  if (rc != VGI_OK) return fail(rc, "glTF: %s", err.c_str());
  out->data = im.data.data();
  out->bytes = im.data.size();
  out->mime_type = im.mime_type.c_str();
  out->uri = im.uri.c_str();
  return VGI_OK;
}

int vgi_gltf_set_image_rgba8(vgi_gltf* g, uint32_t index, const uint8_t* rgba8, uint32_t width, uint32_t height) {
  if (!g || !rgba8) return fail(VGI_E_INVALID, "NULL argument");
  if (index >= g->texels.size()) return fail(VGI_E_INVALID, "image %u >= %zu", index, g->texels.size());
  if (width == 0 || height == 0 || width > 16384 || height > 16384)
    return fail(VGI_E_INVALID, "image %ux%u (1..16384 per side)", width, height);
  auto& t = g->texels[index];
  t.rgba.assign(rgba8, rgba8 + (size_t)width * height * 4);
  t.w = width;
  t.h = height;
  t.dirty = true;
  return VGI_OK;
}

int vgi_gltf_upload(vgi_context* ctx, vgi_gltf* g, vgi_mesh* meshes, uint32_t capacity, uint32_t* count) {
  if (!ctx || !g || !count || (capacity && !meshes)) return fail(VGI_E_INVALID, "NULL argument");
  const auto& sc = g->scene;
  *count = (uint32_t)sc.instances.size();
  if (capacity < sc.instances.size())
    return fail(VGI_E_RANGE, "%zu primitives do not fit in %u entries", sc.instances.size(), capacity);
  if (g->ctx && g->ctx != ctx) gltf_free_device(g);
  HIP_TRY(hipSetDevice(ctx->device));
  if (!g->ctx) {
    g->dev_pos.assign(sc.geometry.size(), nullptr);
    g->dev_idx.assign(sc.geometry.size(), nullptr);
    g->dev_uv.assign(sc.geometry.size(), nullptr);
    g->ctx = ctx;
    for (size_t i = 0; i < sc.geometry.size(); i++) {
      const auto& ge = sc.geometry[i];
      const size_t pb = std::max<size_t>(ge.positions.size() * 4, 4), ib = std::max<size_t>(ge.indices.size() * 4, 4);
      hipError_t e = hipMalloc(&g->dev_pos[i], pb);
      if (e == hipSuccess) e = hipMalloc(&g->dev_idx[i], ib);
      if (e == hipSuccess && !ge.texcoords.empty()) e = hipMalloc(&g->dev_uv[i], ge.texcoords.size() * 4);
      if (e == hipSuccess && !ge.positions.empty())
        e = hipMemcpy(g->dev_pos[i], ge.positions.data(), ge.positions.size() * 4, hipMemcpyHostToDevice);
      if (e == hipSuccess && !ge.indices.empty())
        e = hipMemcpy(g->dev_idx[i], ge.indices.data(), ge.indices.size() * 4, hipMemcpyHostToDevice);
      if (e == hipSuccess && !ge.texcoords.empty())
        e = hipMemcpy(g->dev_uv[i], ge.texcoords.data(), ge.texcoords.size() * 4, hipMemcpyHostToDevice);
      if (e != hipSuccess) {
        gltf_free_device(g);
        return fail(e == hipErrorOutOfMemory ? VGI_E_NOMEM : VGI_E_HIP, "glTF upload: %s", hipGetErrorString(e));
      }
    }
  }
  for (auto& t : g->texels) {  // decoded images set since the last upload
    if (!t.dirty) continue;
    if (t.dev) (void)hipFree(t.dev);
    t.dev = nullptr;
    hipError_t e = hipMalloc(&t.dev, t.rgba.size());
    if (e == hipSuccess) e = hipMemcpy(t.dev, t.rgba.data(), t.rgba.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      if (t.dev) (void)hipFree(t.dev);
      t.dev = nullptr;
      (void)hipGetLastError();
      return fail(e == hipErrorOutOfMemory ? VGI_E_NOMEM : VGI_E_HIP, "glTF image upload: %s", hipGetErrorString(e));
    }
    t.dirty = false;
  }
  for (size_t i = 0; i < sc.instances.size(); i++) {
    const auto& in = sc.instances[i];
    const auto& ge = sc.geometry[in.geometry];
    vgi_mesh m{};
    m.positions = (const float*)g->dev_pos[in.geometry];
    m.position_stride = 12;
    m.num_vertices = (uint32_t)(ge.positions.size() / 3);
    m.indices = g->dev_idx[in.geometry];
    m.index_type = VGI_INDEX_U32;
    m.num_triangles = (uint32_t)(ge.indices.size() / 3);
    m.albedo_rgb8 = nullptr;
    std::memcpy(m.base_color, in.base_color, 4);
    std::memcpy(m.model, in.model, sizeof(m.model));
    if (in.image >= 0 && g->dev_uv[in.geometry] && g->texels[(size_t)in.image].dev) {
      const auto& t = g->texels[(size_t)in.image];
      m.texcoords = (const float*)g->dev_uv[in.geometry];
      m.texcoord_stride = 8;
      m.texture_rgba8 = (const uint8_t*)t.dev;
      m.texture_width = t.w;
      m.texture_height = t.h;
      std::memcpy(m.base_color_factor, in.factor, sizeof(m.base_color_factor));
    }
    meshes[i] = m;
  }
  return VGI_OK;
}

int vgi_gltf_release(vgi_gltf* g) {
  if (!g) return fail(VGI_E_INVALID, "NULL argument");
  gltf_free_device(g);
  delete g;
  return VGI_OK;
}

// ---- screen-space passes ----------------------------------------------------

namespace {

constexpr uint32_t kMaxScreenEdge = 16384;

bool image_format_ok(uint32_t f) { return f == VGI_IMAGE_RGBA32F || f == VGI_IMAGE_RGBA16F; }
size_t image_texel_bytes(uint32_t f) { return f == VGI_IMAGE_RGBA16F ? 8 : 16; }
bool aligned(const void* p, size_t a) { return ((uintptr_t)p % a) == 0; }
bool overlaps(const void* a, size_t na, const void* b, size_t nb) {
  const uintptr_t x = (uintptr_t)a, y = (uintptr_t)b;
  return x < y + nb && y < x + na;
}

int check_screen(uint32_t w, uint32_t h) {
  if (w > kMaxScreenEdge || h > kMaxScreenEdge)
    return fail(VGI_E_RANGE, "image %ux%u exceeds %u per edge", w, h, kMaxScreenEdge);
  return VGI_OK;
}

}  // namespace

int vgi_gbuffer_decode(vgi_context* ctx, const vgi_gbuffer_ref* in, float* out_position, float* out_normal,
                       float* out_view_depth, void* stream) {
  if (!ctx || !in || !out_position || !out_normal) return fail(VGI_E_INVALID, "NULL argument");
  TRY(check_screen(in->width, in->height));
  if (!in->width || !in->height) return VGI_OK;
  if (!in->depth || !in->normals) return fail(VGI_E_INVALID, "NULL depth or normal target");
  if (!aligned(in->depth, 4) || !aligned(in->normals, 8) || !aligned(out_position, 4) ||
      !aligned(out_normal, 4) || !aligned(out_view_depth, 4))
    return fail(VGI_E_INVALID, "misaligned buffer (depth 4 B, normals 8 B, outputs 4 B)");
  const size_t npx = (size_t)in->width * in->height;
  if (overlaps(out_position, npx * 12, out_normal, npx * 12) ||
      (out_view_depth && (overlaps(out_view_depth, npx * 4, out_position, npx * 12) ||
                          overlaps(out_view_depth, npx * 4, out_normal, npx * 12))))
    return fail(VGI_E_INVALID, "output buffers overlap");
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(launch_gbuffer_decode(*in, out_position, out_normal, out_view_depth, pick(ctx, stream)));
  return VGI_OK;
}

int vgi_temporal_resolve(vgi_context* ctx, const vgi_temporal_params* p, void* stream) {
  if (!ctx || !p) return fail(VGI_E_INVALID, "NULL argument");
  TRY(check_screen(p->width, p->height));
  if (!p->width || !p->height) return VGI_OK;
  if (!p->indirect || !p->position || !p->view_depth || !p->history || !p->out)
    return fail(VGI_E_INVALID, "NULL buffer");
  if (!image_format_ok(p->history_format) || !image_format_ok(p->out_format))
    return fail(VGI_E_INVALID, "unknown image format");
  if (!aligned(p->indirect, 16) || !aligned(p->position, 4) || !aligned(p->view_depth, 4) ||
      !aligned(p->history, image_texel_bytes(p->history_format)) || !aligned(p->out, image_texel_bytes(p->out_format)))
    return fail(VGI_E_INVALID, "misaligned buffer (one texel for images, 4 B otherwise)");
  const size_t npx = (size_t)p->width * p->height;
  if (overlaps(p->out, npx * image_texel_bytes(p->out_format), p->history, npx * image_texel_bytes(p->history_format)))
    return fail(VGI_E_INVALID, "out overlaps history");
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(launch_temporal(*p, pick(ctx, stream)));
  return VGI_OK;
}

int vgi_bilateral_blur(vgi_context* ctx, const vgi_blur_params* p, void* stream) {
  if (!ctx || !p) return fail(VGI_E_INVALID, "NULL argument");
  TRY(check_screen(p->width, p->height));
  if (!p->width || !p->height) return VGI_OK;
  if (!p->color || !p->depth || !p->out) return fail(VGI_E_INVALID, "NULL buffer");
  if (!image_format_ok(p->color_format) || !image_format_ok(p->out_format))
    return fail(VGI_E_INVALID, "unknown image format");
  if (p->axis > 1) return fail(VGI_E_INVALID, "axis %u (0: x, 1: y)", p->axis);
  if (!aligned(p->color, image_texel_bytes(p->color_format)) || !aligned(p->out, image_texel_bytes(p->out_format)) ||
      !aligned(p->depth, 4))
    return fail(VGI_E_INVALID, "misaligned buffer (one texel for images, 4 B for depth)");
  const size_t npx = (size_t)p->width * p->height;
  if (overlaps(p->out, npx * image_texel_bytes(p->out_format), p->color, npx * image_texel_bytes(p->color_format)))
    return fail(VGI_E_INVALID, "out overlaps color");
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(launch_blur(*p, pick(ctx, stream)));
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
        const bool occ =
            (uint32_t)packed[bin] != 0 && ((mask[(size_t)brick * 16 + (local >> 5)] >> (local & 31u)) & 1u);
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
        } else if (s->prec == VGI_PREC_FP16X2) {
          // texel from its entry's low halves; the entry before holds it in its high halves
          uint16_t h[8], g[8];
          std::memcpy(h, raw.data() + pi * 16, 16);
          std::memcpy(g, raw.data() + (pi - 1) * 16, 16);
          for (int c = 0; c < 4; c++) {
            if (g[2 * c + 1] != h[2 * c])
              return fail(VGI_E_STATE, "level %u texel (%u, %u, %u): x-pair copies differ", level, x, y, z);
            o[c] = host_half(h[2 * c]);
          }
          if (x + 1 == nk)  // the last texel's entry pairs it with the zero border
            for (int c = 0; c < 4; c++)
              if (h[2 * c + 1] != 0) return fail(VGI_E_STATE, "level %u: border of row (%u, %u) not zero", level, y, z);
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
  HIP_TRY(hipDeviceSynchronize());
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

int vgi_probe(vgi_context* ctx, uint32_t kind, void* buf, size_t buf_bytes, uint32_t blocks, uint32_t iters,
              void* stream) {
  if (!ctx || !buf) return fail(VGI_E_INVALID, "NULL argument");
  if (kind > VGI_PROBE_INT) return fail(VGI_E_INVALID, "probe kind %u", kind);
  if (blocks == 0 || blocks > (1u << 20)) return fail(VGI_E_INVALID, "probe blocks %u", blocks);
  const bool lines = kind == VGI_PROBE_ATOMIC_RETURN || kind == VGI_PROBE_ATOMIC_NORETURN;
  const size_t need = (size_t)blocks * 256u * (lines ? 64u : 4u);
  if (buf_bytes < need) return fail(VGI_E_INVALID, "probe buffer of %zu bytes, %zu needed", buf_bytes, need);
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(launch_probe(kind, buf, blocks, iters, pick(ctx, stream)));
  return VGI_OK;
}

int vgi_stream_query(vgi_context* ctx, void* stream) {
  if (!ctx) return fail(VGI_E_INVALID, "ctx is NULL");
  HIP_TRY(hipSetDevice(ctx->device));
  const hipError_t e = hipStreamQuery(pick(ctx, stream));
  if (e == hipSuccess) return VGI_OK;
  if (e == hipErrorNotReady) {
    (void)hipGetLastError();
    return VGI_E_NOT_READY;
  }
  return fail(VGI_E_HIP, "hipStreamQuery: %s", hipGetErrorString(e));
}

int vgi_stream_wait(vgi_context* ctx, void* waiting, void* signalling) {
  if (!ctx) return fail(VGI_E_INVALID, "ctx is NULL");
  HIP_TRY(hipSetDevice(ctx->device));
  const hipStream_t w = pick(ctx, waiting), sg = pick(ctx, signalling);
  if (w == sg) return VGI_OK;  // stream order already holds
  // one event serves every call: hipStreamWaitEvent waits for the record made just before
  // it, and a later record of the same event does not change an enqueued wait
  std::lock_guard<std::mutex> lk(ctx->mu);
  HIP_TRY(hipEventRecord(ctx->dep, sg));
  HIP_TRY(hipStreamWaitEvent(w, ctx->dep, 0));
  return VGI_OK;
}

int vgi_stream_destroy(vgi_context* ctx, void* stream) {
  if (!ctx) return fail(VGI_E_INVALID, "ctx is NULL");
  if (!stream) return VGI_OK;
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    auto it = ctx->ws.find((hipStream_t)stream);
    if (it != ctx->ws.end()) {
      free_workspace(it->second.get());
      ctx->ws.erase(it);
    }
  }
  HIP_TRY(hipStreamDestroy((hipStream_t)stream));
  return VGI_OK;
}

int vgi_workspace_bytes(vgi_context* ctx, uint64_t* out) {
  if (!ctx || !out) return fail(VGI_E_INVALID, "NULL argument");
  std::lock_guard<std::mutex> lk(ctx->mu);
  uint64_t b = 0;
  for (auto& kv : ctx->ws) b += kv.second->device_bytes();
  *out = b;
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
