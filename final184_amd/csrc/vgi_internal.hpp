// vgi_internal.hpp — device-side structures and launch declarations shared by the
// kernels (voxelize.hip, mips.hip, trace.hip) and the C ABI layer (abi.hip).
//
// Every kernel is batch-native: it reads per-scene descriptors from device memory and
// maps its workgroups to (scene, first item) through a block table built on the host. A
// single scene is a batch of one (its kernels index directly, without the table).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/vgi.h"

namespace vgi {

constexpr int kFixedOne = 256;         // fixed-point units per voxel (1 << VGI_FIXED_BITS)
constexpr int kAccWords = 8;           // uint32 words per accumulator voxel
constexpr int kAccBrick = 8;           // accumulator storage brick edge (voxels)
constexpr int kMaxChunk = 64;          // triangles per voxelize work item (LDS field width)
constexpr int kNormalBias = 1024;      // quantized normal offset in the accumulator
constexpr int kSmallBins = 8;          // a triangle whose bin box has <= 8 bins keeps its slots
constexpr uint32_t kBlockTris = 256;   // triangles per setup / emit workgroup
constexpr uint32_t kBlockBins = 256;   // bins per bins / workmap workgroup
constexpr uint32_t kUnitItems = 16;    // chunks one voxelize workgroup takes in turn (8^3 bins)
constexpr size_t kWorkEntryBytes = 12;  // per work unit: its (scene, bin) and its place in the order

// Per-mesh descriptor, device resident.
struct MeshDev {
  const float* pos;
  const void* idx;
  const uint8_t* albedo;
  uint32_t stride;
  uint32_t nverts;
  uint32_t idx_type;
  uint32_t ntri;
  uint32_t tri_base;  // first triangle of this mesh within its scene
  uint32_t base_rgb;  // r | g << 8 | b << 16
  float m[12];
  const float* uv;      // textured meshes: uv per vertex (uv_stride bytes apart), else null
  const uint32_t* tex;  // RGBA8 texels, one word each, row-major tw x th
  uint32_t uv_stride;
  uint32_t tw, th;
  uint32_t textured;    // 1: albedo from the texture (uv and tex set)
  float factor[4];      // RGBA multiplier of the sample
};
static_assert(sizeof(MeshDev) == 144, "MeshDev layout");

// Texture coordinates of a triangle of a textured batch (written by the setup kernel).
struct __align__(16) TriTex {
  float uv[6];        // uv of vertices 0, 1, 2
  uint32_t mesh;      // index of the triangle's mesh in its scene
  uint32_t textured;  // 1: its mesh is textured
};
static_assert(sizeof(TriTex) == 32, "TriTex layout");

// Triangle record written by the setup kernel (64 bytes).
struct __align__(16) TriRec {
  int32_t q[9];       // snapped vertices, 1/256-voxel units relative to the grid origin
  uint32_t rgb;       // r | g << 8 | b << 16 | status << 24 (0 ok, 1 rejected, 2 degenerate)
  int16_t nq[3];      // unit normal x 1024, rounded
  uint8_t axis;       // dominant axis index 0=x 1=y 2=z (VoxelGS rule)
  uint8_t pad0;
  int16_t lo[3];      // voxel bounding box clipped to the grid (inclusive); lo > hi if empty
  int16_t hi[3];
  uint32_t pad1;
};
static_assert(sizeof(TriRec) == 64, "TriRec layout");

struct DevStats {
  unsigned long long rejected;
  unsigned long long degenerate;
  unsigned long long pairs;  // exact (triangle, bin) pair count of the scene
  uint32_t big_count;        // triangles spanning more than kSmallBins bins
  uint32_t zero_count;       // bins split over several work items
};
static_assert(sizeof(DevStats) == 32, "DevStats layout");

// Words of a call's pinned batch totals (VoxBatch::host_hist).
constexpr uint32_t kHostHistWords = 5;

// Default polls of a predecessor's flag before the bin scan's look-back gives up.
constexpr uint32_t kScanSpin = 1u << 22;

// A scene's result of one voxelize call, stored by the settle step (the first wavefront of
// k_scatter, after the bin scan) into the scene's pinned host record, seq once the other
// words are complete: the record is whole once seq equals the call's number (read by
// vgi_scene_get_stats).
struct SceneVoxResult {
  DevStats st;
  unsigned long long totals;  // work units << 32 | pairs of the scene's bins
  uint32_t overflow;          // the call's pair list did not fit: per-triangle path taken
  uint32_t seq;               // call number
  unsigned long long pad[2];
};
static_assert(sizeof(SceneVoxResult) == 64, "SceneVoxResult layout");

// One scene of a voxelize batch (device memory). acc, mask and bin_packed belong to the
// scene and stay resident; every other array is this scene's slice of the calling
// stream's voxelize scratch, valid for the duration of one call.
struct VoxScene {
  const MeshDev* meshes;
  TriRec* tris;
  TriTex* tritex;                  // per triangle when the batch has a textured mesh, else null
  uint32_t* tri_slot;              // kSmallBins slots per triangle (small triangles)
  uint32_t* big_list;              // ids of large triangles
  // per-bin counters (nbins each), bin b at a hashed index (voxelize.hip: ctr)
  uint32_t* bin_small;             // pairs from small triangles (slot counters)
  uint32_t* bin_big;               // pairs from large triangles
  uint32_t* big_cursor;            // large-triangle pairs placed so far
  uint32_t* zero_list;             // bins with more than one work item
  unsigned long long* bin_packed;  // nbins: work units << 32 | pairs (kept with the scene)
  uint32_t* acc;                   // n^3 x 8 words, brick-major (8^3 bricks)
  uint32_t* mask;                  // 16 words (512 bits) per 8^3 brick: occupied voxels
  DevStats* stats;
  SceneVoxResult* result;          // pinned host record of the scene
  float origin[3];
  float snap_scale;                // float(n * 256) / extent
  uint32_t nmesh;
  uint32_t ntri;
  uint32_t n;
  uint32_t bin_shift;              // log2(bin edge in voxels)
  uint32_t nbins_axis;             // n >> bin_shift
  uint32_t chunk;                  // triangles per work item (<= kMaxChunk)
  uint32_t bin_base;               // first bin of this scene in the batch-wide arrays
  uint32_t unit_items;             // work items (of chunk triangles) per work unit: 1, or
                                   // kUnitItems when the bin is one 8^3 tile
};

// Descriptors of an allocation-free voxelize call (one scene), passed by value to
// k_stage_desc, which writes them into the device scratch the other kernels read.
struct NoAllocDesc {
  VoxScene scene;
  MeshDev meshes[VGI_NOALLOC_MAX_MESHES];
};

struct VoxBatch {
  const VoxScene* scenes;
  const uint2* tri_blocks;          // (scene, first triangle) per setup/emit workgroup
  const uint2* bin_blocks;          // (scene, first bin) per bins/workmap workgroup
  unsigned long long* scan_all;     // exclusive scan of the packed bin entries (k_bins)
  uint32_t* scan_flags;             // per k_bins workgroup: 0, 1 total published, 2 prefix
                                    // published (cleared with the overflow word)
  unsigned long long* scan_vals;    // per k_bins workgroup: its total, then its inclusive prefix
  uint32_t scan_spin;               // polls of a predecessor's flag before the look-back gives up
  uint2* work_map;                  // (scene, bin) of every work unit (work_cap entries)
  uint32_t* unit_order;             // the order k_voxelize takes the units in (work_cap entries)
  uint32_t* pairs_all;              // triangle ids grouped by bin, every scene's bins in turn
  uint32_t pair_cap;                // entries of pairs_all
  uint32_t work_cap;                // entries of work_map
  uint32_t* overflow;               // device word: 1 when the pairs or units exceed the caps;
                                    // the two words after it (cleared with it) are the unit
                                    // order's front and back cursors, the third is set when
                                    // the bin scan's look-back gave up
  unsigned long long* host_hist;    // pinned (kHostHistWords): batch pairs, units, large
                                    // triangles, split bins, then seq (stored after them)
  uint32_t seq;                     // this call's number (never 0)
  uint32_t nscenes;
  uint32_t total_bins;
  uint32_t textured;                // some mesh of the batch is textured (scenes carry tritex)
};

struct LightDev {
  uint32_t type;  // VGI_LIGHT_DIRECTIONAL or VGI_LIGHT_POINT
  float vec[3];
  float color[3];
  float min_distance;
  uint32_t shadow;  // 1: cut where the visibility walk toward the light meets an occupied voxel
  int32_t step[3];  // directional, shadow: the walk's step toward the light, 16.16 fixed point
  float pv[3];      // point, shadow: position in voxel units of level 0, (vec - origin) / h
  uint32_t pad;
};
static_assert(sizeof(LightDev) == 64, "one light is 64 bytes of the descriptor upload");

// The per-scene light record of a vgi_light (host side): kind, shadow flag, and the
// visibility walk's precomputed step (directional) or voxel-unit position (point).
LightDev light_dev(const vgi_light& L, const float origin[3], float h);

// One scene of a mip batch (device memory).
struct MipScene {
  const uint32_t* acc;
  const uint32_t* mask;
  const unsigned long long* bin_packed;
  uint32_t* live;  // per 16^3 tile of level 0: 1 while the tile's stored texels are not all zero;
                   // after the tiles' words, 8 words per tile: bit (ly + 16 lz) while that
                   // x-row's stored level-0 texels are not all zero
  uint32_t* order; // the injecting launch's tiles, dense ones first (one word per tile), then
                   // the dense and other tiles' counts (k_tiles)
  void* level[VGI_MAX_LEVELS];
  uint32_t nk[VGI_MAX_LEVELS];
  uint32_t n;
  uint32_t levels;
  uint32_t bin_shift;
  uint32_t nbins_axis;
  uint32_t nlights;
  float origin[3];
  float h;
  const LightDev* lights;  // nlights entries, in the call's descriptor upload
};

// Mip workgroup: scene, tile of its source level, source level, levels produced.
struct MipBlock {
  uint32_t scene, tile, src, nout;
};

struct MipBatch {
  const MipScene* scenes;
  const MipBlock* blocks;  // per workgroup; null for one scene: then block b is tile b of
                           // scene 0 with the launch's src and nout
  const float* lut;        // 256 entries: (i/255)^2.2
  uint32_t src, nout;
  bool shadow;             // some light of the batch casts shadows (the injecting launch walks)
  bool ordered;            // one scene, injecting launch: block b takes the listed tile b
  uint32_t* clear_counts;  // a later launch of an ordered build: the list's two counts, which
                           // its first workgroup zeroes for the next build (else null)
};

// One scene of a trace batch (device memory).
// One sample of a fixed-aperture cone's march (the same for every pixel): distance t, the
// level fraction fr and 1 - fr, and the lower level l0, or kSchedTop | top level for a
// sample taken at the top level alone (fr 1, gr 0); then fr and gr as the packed halves
// (w, 0), low half first, that the FP16 x-pair gathers scale their x weights with.
struct SchedEntry {
  float t, fr, gr;
  uint32_t code;
  uint32_t hfr, hgr;
  uint32_t pad[2];
};
constexpr uint32_t kSchedTop = 0x80000000u;
static_assert(sizeof(SchedEntry) == 32, "schedule entries are read as 32-byte scalar loads");

// One mip level as the cone march addresses it, read with one 16-byte scalar load: base
// pointer, bytes (as a buffer record size, levels under 4 GiB) and the padded row pitch
// in texels.
struct __align__(16) LevelDesc {
  const void* ptr;
  uint32_t bytes32;
  uint32_t p1;  // n_k + 2
};
static_assert(sizeof(LevelDesc) == 16, "LevelDesc is one 16-byte load");

struct TraceParams {
  LevelDesc lv[VGI_MAX_LEVELS];
  const void* level[VGI_MAX_LEVELS];
  uint64_t level_bytes[VGI_MAX_LEVELS];
  uint32_t nk[VGI_MAX_LEVELS];
  uint32_t levels;
  uint32_t n;
  uint32_t width, height;
  uint32_t tiles_x, tiles_y;
  uint32_t num_diffuse;
  uint32_t specular;
  float origin[3];
  float h, inv_h;
  float tan_d, tan_s;
  float step, max_dist, cutoff, normal_off, start_off;
  float eye[3];
  float dirs[VGI_MAX_DIFFUSE_CONES][3];
  float w[VGI_MAX_DIFFUSE_CONES];
  const SchedEntry* sched[2];  // march schedules: [0] diffuse cones, [1] specular cone
  uint32_t sched_len[2];       // (no roughness buffer); filled by the host
  const float* pos;
  const float* nrm;
  const float* rough;
  float4* out_d;
  float4* out_s;
};

struct TraceBatch {
  const TraceParams* scenes;
  const uint2* blocks;            // (scene, tile) per workgroup (screen-tile order, > 1 scene)
  unsigned long long* counter;    // trilinear gathers (may be null)
  // Morton order (null in screen-tile order): the batch's pixels sorted by
  // (scene << 32 | Morton code of the position); scene s owns entries [seg[s], seg[s+1]),
  // a multiple of 256, padded with pixel index kNoPixel after its sorted pixels.
  const unsigned long long* skeys;
  const uint32_t* spix;
  const uint32_t* seg;
  uint32_t nblocks;
  uint32_t tile_w, tile_h;
  uint32_t prec;
  uint32_t nscenes;
  uint32_t band;  // work items (tiles or sorted runs) an XCD takes in a row before the next XCD's
};

constexpr uint32_t kNoPixel = 0xFFFFFFFFu;
constexpr uint32_t kSortBlock = 256;  // sorted pixels per workgroup

// Texel bytes per precision.
inline uint32_t texel_bytes(uint32_t prec) { return (prec == VGI_PREC_FP32 || prec == VGI_PREC_FP16X2) ? 16u : 8u; }

// ---- global-memory access ----
// Kernels take their descriptor tables as __restrict__ arguments, so the compiler may
// read descriptors with scalar loads. The buffer pointers inside a descriptor are generic
// pointers, and accesses through generic pointers compile to flat instructions (which
// also wait on the LDS counter); every buffer of this library is in global memory, so
// the hot accesses go through these helpers, which access it as such.
#define VGI_GLOBAL __attribute__((address_space(1)))
typedef uint32_t vgi_u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t vgi_u32x4 __attribute__((ext_vector_type(4)));

template <class T>
__device__ __forceinline__ T gld(const T* p) {  // scalar types
  return *(const VGI_GLOBAL T*)p;
}
// base + a 32-bit byte offset: with a wave-uniform base the load takes the base in SGPRs
// and one VGPR of offset (global_load ... v_off, s[base]) instead of a 64-bit address
template <class T>
__device__ __forceinline__ T gld_at(const T* base, uint32_t byte_off) {  // scalar types
  return *(const VGI_GLOBAL T*)((const VGI_GLOBAL char*)base + byte_off);
}
template <class T>
__device__ __forceinline__ void gst(T* p, T v) {  // scalar types
  *(VGI_GLOBAL T*)p = v;
}
__device__ __forceinline__ uint2 gld2(const void* p) {
  const vgi_u32x2 v = *(const VGI_GLOBAL vgi_u32x2*)p;
  return make_uint2(v.x, v.y);
}
__device__ __forceinline__ uint4 gld4(const void* p) {
  const vgi_u32x4 v = *(const VGI_GLOBAL vgi_u32x4*)p;
  return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void gst2(void* p, uint2 v) {
  *(VGI_GLOBAL vgi_u32x2*)p = vgi_u32x2{v.x, v.y};
}
__device__ __forceinline__ void gst4(void* p, uint4 v) {
  *(VGI_GLOBAL vgi_u32x4*)p = vgi_u32x4{v.x, v.y, v.z, v.w};
}
// whole records (size a multiple of 16 bytes, 16-byte aligned) as 16-byte accesses
// A record at a wave-uniform address, read through the scalar data cache (s_load):
// the constant address space tells the compiler the data is read-only for the kernel.
template <class T>
__device__ __forceinline__ T sld_rec(const T* p) {
  static_assert(sizeof(T) % 16 == 0, "record size");
  struct Raw {
    uint4 w[sizeof(T) / 16];
  } r;
  const __attribute__((address_space(4))) vgi_u32x4* q = (const __attribute__((address_space(4))) vgi_u32x4*)p;
#pragma unroll
  for (uint32_t i = 0; i < sizeof(T) / 16; i++) {
    const vgi_u32x4 v = q[i];
    r.w[i] = make_uint4(v.x, v.y, v.z, v.w);
  }
  return __builtin_bit_cast(T, r);
}

template <class T>
__device__ __forceinline__ T gld_rec(const T* p) {
  static_assert(sizeof(T) % 16 == 0, "record size");
  struct Raw {
    uint4 w[sizeof(T) / 16];
  } r;
#pragma unroll
  for (uint32_t i = 0; i < sizeof(T) / 16; i++) r.w[i] = gld4((const uint4*)p + i);
  return __builtin_bit_cast(T, r);
}
template <class T>
__device__ __forceinline__ void gst_rec(T* p, const T& v) {
  static_assert(sizeof(T) % 16 == 0, "record size");
  struct Raw {
    uint4 w[sizeof(T) / 16];
  };
  const Raw r = __builtin_bit_cast(Raw, v);
#pragma unroll
  for (uint32_t i = 0; i < sizeof(T) / 16; i++) gst4((uint4*)p + i, r.w[i]);
}
// relaxed, device scope: the semantics of atomicAdd / atomicOr
__device__ __forceinline__ uint32_t gatomic_add(uint32_t* p, uint32_t v) {
  return __hip_atomic_fetch_add((VGI_GLOBAL uint32_t*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long gatomic_add(unsigned long long* p, unsigned long long v) {
  return __hip_atomic_fetch_add((VGI_GLOBAL unsigned long long*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void gatomic_or(uint32_t* p, uint32_t v) {
  __hip_atomic_fetch_or((VGI_GLOBAL uint32_t*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// FP32 -> FP16 bits, round to nearest even. The value is rounded to FP32 first, then to
// FP16: without the register barrier the backend may fuse a preceding multiply into the
// conversion (v_fma_mixlo_f16), which rounds once and differs from the two roundings by
// an ulp in rare near-ties.
__device__ __forceinline__ uint16_t f32_to_f16_bits(float f) {
  __asm__ volatile("" : "+v"(f));
  _Float16 h = (_Float16)f;
  return __builtin_bit_cast(uint16_t, h);
}
__device__ __forceinline__ float f16_bits_to_f32(uint16_t b) {
  return (float)__builtin_bit_cast(_Float16, b);
}

// Sets the calling thread's vgi_last_error() message and returns code (defined in abi.hip).
int set_error(int code, const char* fmt, ...);

// Validates one serving request's geometry, lights and frames (defined in abi.hip).
int check_request_input(const vgi_mesh* meshes, uint32_t num_meshes, const vgi_grid* grid,
                        const vgi_lights* lights, const vgi_frame* frames, uint32_t num_frames);

// ---- launchers (each returns hipError_t) ----
hipError_t launch_fill_u32(uint32_t* p, uint32_t value, size_t count, hipStream_t s);
hipError_t launch_copy(void* dst, const void* src, size_t bytes, hipStream_t s);
hipError_t launch_setup(const VoxBatch& B, uint32_t blocks, hipStream_t s);
hipError_t launch_big_count(const VoxBatch& B, uint32_t blocks, hipStream_t s);
hipError_t launch_bins(const VoxBatch& B, uint32_t blocks, hipStream_t s);
hipError_t launch_probe(uint32_t kind, void* buf, uint32_t blocks, uint32_t iters, hipStream_t s);
hipError_t launch_stage_desc(const NoAllocDesc& d, NoAllocDesc* dst, hipStream_t s);
hipError_t launch_scatter(const VoxBatch& B, uint32_t nblk_b, uint32_t nblk_t, uint32_t big_blocks,
                          uint32_t zero_blocks, hipStream_t s);
hipError_t launch_voxelize(const VoxBatch& B, uint32_t tile, uint32_t grid_blocks, hipStream_t s);
hipError_t launch_voxelize_fallback(const VoxBatch& B, uint32_t grid_blocks, hipStream_t s);
hipError_t launch_mips(const MipBatch& B, uint32_t prec, bool from_acc, uint32_t blocks, hipStream_t s);
hipError_t launch_tiles(const MipBatch& B, uint32_t ntiles, hipStream_t s);
hipError_t launch_trace(const TraceBatch& B, bool buffer_path, hipStream_t s);
// Morton keys of every pixel of a batch (entries seg[0]..seg[nscenes], padding included).
hipError_t launch_morton_keys(const TraceBatch& B, unsigned long long* keys, uint32_t* pix, uint32_t entries,
                              hipStream_t s);
hipError_t sort_pixels(void* tmp, size_t* tmp_bytes, const unsigned long long* keys_in,
                       unsigned long long* keys_out, const uint32_t* pix_in, uint32_t* pix_out, uint32_t n,
                       int end_bit, hipStream_t s);

// ---- screen-space passes (screen.hip); arguments validated by the caller ----
hipError_t launch_gbuffer_decode(const vgi_gbuffer_ref& in, float* pos, float* nrm, float* view_depth,
                                 hipStream_t s);
hipError_t launch_temporal(const vgi_temporal_params& p, hipStream_t s);
hipError_t launch_blur(const vgi_blur_params& p, hipStream_t s);

}  // namespace vgi
