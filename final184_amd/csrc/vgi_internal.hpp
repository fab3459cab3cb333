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
};

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

// One scene of a voxelize batch (device memory). acc, mask and bin_packed belong to the
// scene and stay resident; every other array is this scene's slice of the calling
// stream's voxelize scratch, valid for the duration of one call.
struct VoxScene {
  const MeshDev* meshes;
  TriRec* tris;
  uint32_t* tri_slot;              // kSmallBins slots per triangle (small triangles)
  uint32_t* big_list;              // ids of large triangles
  uint32_t* bin_small;             // nbins: pairs from small triangles (slot counters)
  uint32_t* bin_big;               // nbins: pairs from large triangles
  uint32_t* big_cursor;            // nbins
  uint32_t* zero_list;             // bins with more than one work item
  unsigned long long* bin_packed;  // nbins: work items << 32 | pairs (kept with the scene)
  uint32_t* pairs;                 // triangle ids grouped by bin
  uint32_t* acc;                   // n^3 x 8 words, brick-major (8^3 bricks)
  uint32_t* mask;                  // 16 words (512 bits) per 8^3 brick: occupied voxels
  DevStats* stats;
  float origin[3];
  float snap_scale;                // float(n * 256) / extent
  uint32_t nmesh;
  uint32_t ntri;
  uint32_t n;
  uint32_t bin_shift;              // log2(bin edge in voxels)
  uint32_t nbins_axis;             // n >> bin_shift
  uint32_t chunk;                  // triangles per work item (<= kMaxChunk)
  uint32_t pairs_cap;
  uint32_t bin_base;               // first bin of this scene in the batch-wide arrays
};

struct VoxBatch {
  const VoxScene* scenes;
  const uint2* tri_blocks;          // (scene, first triangle) per setup/emit workgroup
  const uint2* bin_blocks;          // (scene, first bin) per bins/workmap workgroup
  unsigned long long* packed_all;   // total_bins + 1, scene bins concatenated
  unsigned long long* scan_all;     // exclusive scan of packed_all
  unsigned long long* totals;       // per scene: items << 32 | pairs; [nscenes] = batch total
  uint2* work_map;                  // (scene, bin) of every work item
  uint32_t nscenes;
  uint32_t total_bins;
};

struct LightDev {
  uint32_t type;
  float vec[3];
  float color[3];
  float min_distance;
};

// One scene of a mip batch (device memory).
struct MipScene {
  const uint32_t* acc;
  const uint32_t* mask;
  const unsigned long long* bin_packed;
  void* level[VGI_MAX_LEVELS];
  uint32_t nk[VGI_MAX_LEVELS];
  uint32_t n;
  uint32_t levels;
  uint32_t bin_shift;
  uint32_t nbins_axis;
  uint32_t nlights;
  float origin[3];
  float h;
  LightDev lights[VGI_MAX_LIGHTS];
};

// Mip workgroup: scene, tile of its source level, source level, levels produced.
struct MipBlock {
  uint32_t scene, tile, src, nout;
};

struct MipBatch {
  const MipScene* scenes;
  const MipBlock* blocks;
  const float* lut;  // 256 entries: (i/255)^2.2
};

// One scene of a trace batch (device memory).
struct TraceParams {
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
  const float* pos;
  const float* nrm;
  const float* rough;
  float4* out_d;
  float4* out_s;
};

struct TraceBatch {
  const TraceParams* scenes;
  const uint2* blocks;            // (scene, tile) per workgroup, in XCD-band order
  unsigned long long* counter;    // trilinear gathers (may be null)
  uint32_t nblocks;
  uint32_t tile_w, tile_h;
  uint32_t prec;
  uint32_t nscenes;
};

// Texel bytes per precision.
inline uint32_t texel_bytes(uint32_t prec) { return prec == VGI_PREC_FP32 ? 16u : 8u; }

// ---- launchers (each returns hipError_t) ----
hipError_t launch_fill_u32(uint32_t* p, uint32_t value, size_t count, hipStream_t s);
hipError_t launch_copy(void* dst, const void* src, size_t bytes, hipStream_t s);
hipError_t launch_setup(const VoxBatch& B, uint32_t blocks, hipStream_t s);
hipError_t launch_big(const VoxBatch& B, bool emit, uint32_t blocks, hipStream_t s);
hipError_t launch_bins(const VoxBatch& B, uint32_t blocks, hipStream_t s);
hipError_t scan_bins(void* tmp, size_t* tmp_bytes, const unsigned long long* in, unsigned long long* out,
                     uint32_t count, hipStream_t s);
hipError_t launch_totals(const VoxBatch& B, hipStream_t s);
hipError_t launch_workmap(const VoxBatch& B, uint32_t blocks, hipStream_t s);
hipError_t launch_emit_small(const VoxBatch& B, uint32_t blocks, hipStream_t s);
hipError_t launch_zero(const VoxBatch& B, uint32_t blocks, hipStream_t s);
hipError_t launch_voxelize(const VoxBatch& B, uint32_t tile, uint32_t grid_blocks, uint32_t nwork,
                           hipStream_t s);
hipError_t launch_mips(const MipBatch& B, uint32_t prec, bool from_acc, uint32_t blocks, hipStream_t s);
hipError_t launch_trace(const TraceBatch& B, bool buffer_path, hipStream_t s);

}  // namespace vgi
