// voxelize.hip — conservative triangle voxelization for gfx950.
//
// Replaces the reference's voxelize raster pipeline: the pass-through VS, the VoxelGS
// dominant-axis projection (Pipelang/Internal/main.lua:83-144) and VoxelPS's RG16UI
// imageStore with last-writer-wins (main.lua:242-275), issued one draw per primitive by
// CVoxelizeRenderer::Render (Foreground/Renderer/VoxelizeRenderer.cpp:81-127).
//
// Every kernel serves a batch of scenes (one scene is a batch of one): per-triangle and
// per-bin kernels map workgroups to (scene, first item) through a block table, and the
// scenes' bins are concatenated so that one scan and one persistent voxelize launch
// cover the whole batch. Pipeline (one stream, enqueued without waiting: the host sizes
// the pair list and the work map before the call from the triangle count and what earlier
// calls on the stream needed; after step 4 every k_scatter wavefront checks that the pairs
// and work units fit, and when they do not, steps 5 and 6 give way to a per-triangle path
// (7) that needs neither):
//   1 k_setup       one thread per triangle: fetch, model transform, snap to 1/256 voxel,
//                   exact integer normal, dominant axis (VoxelGS rule), quantized normal,
//                   voxel bounding box. A triangle whose bounding box spans <= 8 bins
//                   (bin = B^3 voxels) tests each bin exactly and takes a slot in every
//                   overlapped bin (ranked per workgroup in LDS, one global atomic per
//                   distinct bin; the slot is kept); larger triangles go to a list.
//   2 k_big_count   one wavefront per listed triangle: its bins, tested cooperatively.
//   3 k_bins        per bin: pairs = small + large, work items = ceil(pairs / chunk),
//                   work units = ceil(items / unit_items), packed as (units << 32 | pairs);
//                   bins with > 1 unit are listed. In the same pass:
//   4 scan          one 64-bit exclusive scan over the batch (a look-back between the
//                   workgroups) gives every bin its pair offset and its first work unit.
//   5 k_scatter     one launch; its first wavefront settles the call (per-scene totals
//                   and statistics to the host records, the overflow word for steps 6
//                   and 7); work unit -> (scene, bin) and the units' order (those of
//                   two or more chunks first); small triangles to their kept
//                   slots (no atomics); large triangles after the small ones of each bin;
//                   clear the accumulator and occupancy mask of bins with several units.
//   6 k_voxelize    one workgroup per work unit of the batch, in that order; per work item
//                   (chunk) of it:
//                   triangles' edge/plane coefficients to LDS, columns of the dominant-axis
//                   projection spread over the workgroup, each column tests its 1-4
//                   candidate voxels with the exact separating-axis test and adds packed
//                   64-bit increments into a per-tile LDS accumulator. An 8^3 bin adds its
//                   chunks' tiles into registers and is flushed once per unit; larger bins
//                   flush every tile. Flush: occupancy mask + the records of occupied
//                   voxels, by plain stores for bins of one unit (most bins: a unit is up
//                   to kUnitItems chunks) and by 64-bit integer atomics otherwise.
//   7 k_voxelize_fallback  only when the overflow word is set (k_scatter then cleared every
//                   bin with pairs instead of scattering): one wavefront per triangle tests
//                   every voxel of its box and adds its increments by 64-bit atomics. The
//                   voxel set and the sums are the same; it is slower, and the host grows the
//                   pair list for the calls that follow.
//
// All coverage decisions are exact integer arithmetic on snapped vertices, so the
// voxel set and the integer sums do not depend on order, rounding or thread mapping.
#include <cstddef>


#include "vgi_internal.hpp"

#pragma clang fp contract(off)

namespace vgi {

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t value, size_t count) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < count; i += stride) p[i] = value;
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t value, size_t count, hipStream_t s) {
  if (count == 0) return hipSuccess;
  size_t blocks = (count + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)blocks), dim3(256), 0, s, p, value, count);
  return hipGetLastError();
}

__global__ void k_copy(uint4* __restrict__ dst, const uint4* __restrict__ src, size_t count) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < count; i += stride) dst[i] = src[i];
}

// Copies an allocation-free call's descriptors from the kernel argument into the device
// scratch (one workgroup, one word per lane in turn).
__global__ __launch_bounds__(256) void k_stage_desc(NoAllocDesc d, NoAllocDesc* __restrict__ dst) {
  static_assert(sizeof(NoAllocDesc) % 4 == 0, "descriptors are copied as words");
  const uint32_t* src = (const uint32_t*)&d;
  uint32_t* out = (uint32_t*)dst;
  for (uint32_t i = threadIdx.x; i < sizeof(NoAllocDesc) / 4; i += blockDim.x) gst(out + i, src[i]);
}

hipError_t launch_stage_desc(const NoAllocDesc& d, NoAllocDesc* dst, hipStream_t s) {
  hipLaunchKernelGGL(k_stage_desc, dim3(1), dim3(256), 0, s, d, dst);
  return hipGetLastError();
}

hipError_t launch_copy(void* dst, const void* src, size_t bytes, hipStream_t s) {
  const size_t count = bytes / 16;
  if (count == 0) return hipSuccess;
  size_t blocks = (count + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_copy, dim3((unsigned)blocks), dim3(256), 0, s, (uint4*)dst, (const uint4*)src,
                     count);
  return hipGetLastError();
}

__device__ __forceinline__ int64_t imin64(int64_t a, int64_t b) { return a < b ? a : b; }
__device__ __forceinline__ int64_t imax64(int64_t a, int64_t b) { return a > b ? a : b; }

static_assert(sizeof(MeshDev) % 16 == 0, "MeshDev is read as 16-byte words");

__device__ __forceinline__ uint32_t find_mesh(const MeshDev* __restrict__ m, uint32_t nmesh,
                                              uint32_t t) {
  uint32_t lo = 0, hi = nmesh - 1;
  while (lo < hi) {
    uint32_t mid = (lo + hi + 1) >> 1;
    if (gld(&m[mid].tri_base) <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Exact closed-box test of a triangle against [b, b + d]^3 (fixed-point units, d = 2^dshift),
// with the bounding boxes already known to overlap. Magnitudes: |q| <= 2^18 (the snap guard),
// 0 <= b <= 2^18, so |b - q| <= 2^19 and |edge| <= 2^19 fit int32; |N| <= 2^39, d <= 2^13.
// Edge-axis products are 32 x 32 -> 64-bit multiplies, products with d are shifts, and the
// normal-axis sums stay below 2^61: every value is the exact integer.
__device__ __forceinline__ int64_t shl64(int64_t v, uint32_t s) { return (int64_t)((uint64_t)v << s); }
__device__ __forceinline__ int64_t mul32x32(int32_t a, int32_t b) { return (int64_t)a * (int64_t)b; }

__device__ __forceinline__ bool sat_box(const int64_t (&q)[3][3], const int64_t (&N)[3],
                                        const int32_t (&b)[3], uint32_t dshift) {
  int32_t qi[3][3];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int a = 0; a < 3; a++) qi[i][a] = (int32_t)q[i][a];
  int64_t s_lo = 0, s_hi = 0;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const int64_t t0 = N[a] * (int64_t)(b[a] - qi[0][a]);
    const int64_t dd = shl64(N[a], dshift);
    s_lo += t0 + imin64(0, dd);
    s_hi += t0 + imax64(0, dd);
  }
  if (s_lo > 0 || s_hi < 0) return false;
#pragma unroll
  for (int u = 0; u < 3; u++) {
    const int A = (u + 1) % 3, B = (u + 2) % 3;
    const bool neg = N[u] < 0;  // edge normal (-dB, dA) * sign(N[u])
#pragma unroll
    for (int i = 0; i < 3; i++) {
      const int j = (i + 1) % 3;
      const int32_t dB = qi[j][B] - qi[i][B], dA = qi[j][A] - qi[i][A];
      const int32_t na = neg ? dB : -dB;
      const int32_t nb = neg ? -dA : dA;
      const int64_t e = mul32x32(na, b[A] - qi[i][A]) + mul32x32(nb, b[B] - qi[i][B]) +
                        imax64(0, shl64(na, dshift)) + imax64(0, shl64(nb, dshift));
      if (e < 0) return false;
    }
  }
  return true;
}

__device__ __forceinline__ void load_q(const TriRec& R, int64_t (&q)[3][3], int64_t (&N)[3]) {
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int a = 0; a < 3; a++) q[i][a] = R.q[i * 3 + a];
  int32_t e0[3], e1[3];  // |edge| <= 2^19
#pragma unroll
  for (int a = 0; a < 3; a++) {
    e0[a] = R.q[3 + a] - R.q[a];
    e1[a] = R.q[6 + a] - R.q[a];
  }
  N[0] = mul32x32(e0[1], e1[2]) - mul32x32(e0[2], e1[1]);
  N[1] = mul32x32(e0[2], e1[0]) - mul32x32(e0[0], e1[2]);
  N[2] = mul32x32(e0[0], e1[1]) - mul32x32(e0[1], e1[0]);
}

__device__ __forceinline__ bool tri_ok(const TriRec& R) {
  return (R.rgb >> 24) == 0 && R.lo[0] <= R.hi[0] && R.lo[1] <= R.hi[1] && R.lo[2] <= R.hi[2];
}

// Bin-space bounding box of a triangle record.
struct BinBox {
  int x0, y0, z0, nx, ny, nz;
  __device__ int count() const { return nx * ny * nz; }
};
__device__ __forceinline__ BinBox bin_box(const TriRec& R, uint32_t bs) {
  BinBox b;
  b.x0 = R.lo[0] >> bs;
  b.y0 = R.lo[1] >> bs;
  b.z0 = R.lo[2] >> bs;
  b.nx = (R.hi[0] >> bs) - b.x0 + 1;
  b.ny = (R.hi[1] >> bs) - b.y0 + 1;
  b.nz = (R.hi[2] >> bs) - b.z0 + 1;
  return b;
}

// j-th bin of a bin box (j in [0, count)), and whether the triangle overlaps it.
__device__ __forceinline__ bool bin_candidate(const VoxScene& S, const int64_t (&q)[3][3],
                                              const int64_t (&N)[3], const BinBox& bb, int j,
                                              uint32_t& bin) {
  const int bx = bb.x0 + j % bb.nx, by = bb.y0 + (j / bb.nx) % bb.ny, bz = bb.z0 + j / (bb.nx * bb.ny);
  const uint32_t dshift = VGI_FIXED_BITS + S.bin_shift;
  const int32_t b[3] = {bx << dshift, by << dshift, bz << dshift};  // <= n * 256 <= 2^18
  bin = (uint32_t)bx + S.nbins_axis * ((uint32_t)by + S.nbins_axis * (uint32_t)bz);
  return sat_box(q, N, b, dshift);
}

// Index of bin b's counter in bin_small / bin_big / big_cursor. Bins are scattered over the
// counter array by an odd multiplier (a bijection modulo the power-of-two bin count), so
// spatially adjacent bins -- a wall's run of bins, each hit by many workgroups -- do not
// share a 64-byte line, on which returning atomics from different workgroups serialize.
__device__ __forceinline__ uint32_t ctr(const VoxScene& S, uint32_t bin) {
  const uint32_t nb = S.nbins_axis * S.nbins_axis * S.nbins_axis;
  return (bin * 0x9E3779B1u) & (nb - 1u);
}

// Workgroup -> (scene, first item) for per-triangle and per-bin kernels.
__device__ __forceinline__ uint2 tri_block(const VoxBatch& B, uint32_t blk) {
  return B.nscenes == 1 ? make_uint2(0u, blk * kBlockTris) : B.tri_blocks[blk];
}
__device__ __forceinline__ uint2 bin_block(const VoxBatch& B, uint32_t blk) {
  return B.nscenes == 1 ? make_uint2(0u, blk * kBlockBins) : B.bin_blocks[blk];
}
__device__ __forceinline__ uint2 tri_block(const VoxBatch& B) { return tri_block(B, blockIdx.x); }
__device__ __forceinline__ uint2 bin_block(const VoxBatch& B) { return bin_block(B, blockIdx.x); }

// Offset of bin b's first pair in the batch's pair list (the scenes' lists follow each
// other in bin order, so this is the low half of the batch-wide exclusive scan), and its
// first work unit within the batch (the high half).
__device__ __forceinline__ uint32_t pair_base(const VoxBatch& B, const VoxScene& S, uint32_t b) {
  return (uint32_t)gld(B.scan_all + S.bin_base + b);
}
__device__ __forceinline__ uint32_t first_item(const VoxBatch& B, const VoxScene& S, uint32_t b) {
  return (uint32_t)(gld(B.scan_all + S.bin_base + b) >> 32);
}

// ---------------------------------------------------------------------------
// 1 setup
// ---------------------------------------------------------------------------

__device__ void make_record(const VoxScene& P, uint32_t t, TriRec& R) {
  // the mesh descriptor: through the scalar cache when the wavefront's triangles share
  // one mesh (the usual case), which shortens the fetch chain ahead of the index loads
  const uint32_t mi = find_mesh(P.meshes, P.nmesh, t);
  const uint32_t mi0 = __builtin_amdgcn_readfirstlane(mi);
  MeshDev M;
  if (__all(mi == mi0)) M = sld_rec(P.meshes + mi0);
  else M = gld_rec(P.meshes + mi);
  const uint32_t lt = t - M.tri_base;
  uint32_t vid[3];
  if (M.idx_type == VGI_INDEX_U32) {
    const uint32_t* ix = (const uint32_t*)M.idx + (size_t)lt * 3;
    vid[0] = gld(ix); vid[1] = gld(ix + 1); vid[2] = gld(ix + 2);
  } else if (M.idx_type == VGI_INDEX_U16) {
    const uint16_t* ix = (const uint16_t*)M.idx + (size_t)lt * 3;
    vid[0] = gld(ix); vid[1] = gld(ix + 1); vid[2] = gld(ix + 2);
  } else {
    vid[0] = lt * 3; vid[1] = lt * 3 + 1; vid[2] = lt * 3 + 2;
  }
  uint32_t status = 0;
  int32_t qi[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
  for (int i = 0; i < 3; i++) {
    if (vid[i] >= M.nverts) { status = 1; continue; }
    const float* p = (const float*)((const char*)M.pos + (size_t)vid[i] * M.stride);
    const float x = gld(p), y = gld(p + 1), z = gld(p + 2);
#pragma unroll
    for (int r = 0; r < 3; r++) {
      float w = M.m[r * 4 + 0] * x;
      w = w + M.m[r * 4 + 1] * y;
      w = w + M.m[r * 4 + 2] * z;
      w = w + M.m[r * 4 + 3];
      float v = (w - P.origin[r]) * P.snap_scale;
      if (!(fabsf(v) <= (float)VGI_GUARD_FIXED)) status = 1;
      else qi[i][r] = (int32_t)__builtin_rintf(v);
    }
  }
  if (P.tritex) {  // a textured batch: the triangle's texcoords (zero when untextured)
    TriTex T{};
    T.mesh = mi;
    T.textured = M.textured;
    if (M.textured) {
#pragma unroll
      for (int i = 0; i < 3; i++) {
        if (vid[i] >= M.nverts) continue;  // rejected (status 1): never sampled
        const float* u = (const float*)((const char*)M.uv + (size_t)vid[i] * M.uv_stride);
        T.uv[2 * i] = gld(u);
        T.uv[2 * i + 1] = gld(u + 1);
      }
    }
    gst_rec(P.tritex + t, T);
  }
#pragma unroll
  for (int i = 0; i < 9; i++) R.q[i] = qi[i / 3][i % 3];
  R.nq[0] = R.nq[1] = R.nq[2] = 0;
  R.axis = 2;
  R.pad0 = 0;
  R.pad1 = 0;
  int64_t N[3] = {0, 0, 0};
  if (status == 0) {
    int64_t q[3][3];
    load_q(R, q, N);
    if (N[0] == 0 && N[1] == 0 && N[2] == 0) status = 2;
  }
  uint32_t rgb;
  if (M.albedo) {
    const uint8_t* c = M.albedo + (size_t)lt * 3;
    rgb = (uint32_t)gld(c) | ((uint32_t)gld(c + 1) << 8) | ((uint32_t)gld(c + 2) << 16);
  } else {
    rgb = M.base_rgb & 0xFFFFFFu;
  }
#pragma unroll
  for (int a = 0; a < 3; a++) { R.lo[a] = 1; R.hi[a] = 0; }
  if (status == 0) {
    const double nx = (double)N[0], ny = (double)N[1], nz = (double)N[2];
    const double len = __builtin_sqrt((nx * nx + ny * ny) + nz * nz);
    R.nq[0] = (int16_t)__builtin_rint(nx * 1024.0 / len);
    R.nq[1] = (int16_t)__builtin_rint(ny * 1024.0 / len);
    R.nq[2] = (int16_t)__builtin_rint(nz * 1024.0 / len);
    const int64_t ax = N[0] < 0 ? -N[0] : N[0];
    const int64_t ay = N[1] < 0 ? -N[1] : N[1];
    const int64_t az = N[2] < 0 ? -N[2] : N[2];
    // VoxelGS (main.lua:102-115): X if x>y && x>z; else Y if y>z; Z otherwise.
    if (ax > ay) R.axis = (ax > az) ? 0 : 2;
    else R.axis = (ay > az) ? 1 : 2;
    bool empty = false;
    int16_t lo[3], hi[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
      const int32_t mn = min(qi[0][a], min(qi[1][a], qi[2][a]));
      const int32_t mx = max(qi[0][a], max(qi[1][a], qi[2][a]));
      const int32_t l = max((mn - 1) >> 8, 0);               // floor((mn-1)/256)
      const int32_t h = min(mx >> 8, (int32_t)P.n - 1);      // floor(mx/256)
      lo[a] = (int16_t)l;
      hi[a] = (int16_t)h;
      empty = empty || l > h;
    }
    if (!empty)
#pragma unroll
      for (int a = 0; a < 3; a++) { R.lo[a] = lo[a]; R.hi[a] = hi[a]; }
  }
  R.rgb = rgb | (status << 24);
}

// Slots of small triangles in their bins, aggregated per workgroup: each (triangle, bin)
// pair takes a rank in an LDS hash table keyed by bin, then one returning global atomic
// per distinct bin reserves that bin's run of slots, so a workgroup waits for one atomic
// round trip however many bins its triangles touch (neighbouring triangles of a mesh
// share most of them). Pairs that find no table entry within kProbe probes take a slot
// with their own global atomic.
constexpr uint32_t kSlotTable = 1024;  // distinct bins per workgroup before overflow
constexpr uint32_t kProbe = 32;
constexpr uint32_t kNoKey = 0xFFFFFFFFu;

__global__ __launch_bounds__(256) void k_setup(VoxBatch B, const VoxScene* __restrict__ scenes) {
  __shared__ uint32_t s_key[kSlotTable], s_cnt[kSlotTable], s_base[kSlotTable];
  __shared__ uint32_t s_big_n, s_big_base;  // the workgroup's large triangles and their list run
  for (uint32_t e = threadIdx.x; e < kSlotTable; e += blockDim.x) {
    s_key[e] = kNoKey;
    s_cnt[e] = 0u;
  }
  if (threadIdx.x == 0) s_big_n = 0u;
  const uint2 blk = tri_block(B);  // a workgroup never spans two scenes
  const VoxScene& P = scenes[blk.x];
  const uint32_t t = blk.y + threadIdx.x;
  const bool in = t < P.ntri;
  TriRec R;
  bool small = false, big = false;
  BinBox bb{0, 0, 0, 0, 0, 0};
  int64_t q[3][3], N[3];
  if (in) {
    make_record(P, t, R);
    gst_rec(P.tris + t, R);
    const uint32_t status = R.rgb >> 24;
    if (status == 1) gatomic_add(&P.stats->rejected, 1ull);
    if (status == 2) gatomic_add(&P.stats->degenerate, 1ull);
    if (tri_ok(R)) {
      bb = bin_box(R, P.bin_shift);
      small = bb.count() <= kSmallBins;
      big = !small;
      if (small) load_q(R, q, N);
    }
  }
  __syncthreads();  // table and large-triangle count cleared
  // large triangles: a rank within the workgroup (one LDS atomic per wavefront); the
  // workgroup's run of the list is reserved with one global atomic beside the bin
  // reservations below, and the ids are written after it
  uint32_t big_rank = 0;
  {
    const uint64_t bm = __ballot(big);
    if (bm) {
      const uint32_t lane = __lane_id();
      const int leader = __ffsll((unsigned long long)bm) - 1;
      uint32_t base = 0;
      if (lane == (uint32_t)leader) base = atomicAdd(&s_big_n, (uint32_t)__popcll(bm));
      big_rank = __shfl(base, leader) + (uint32_t)__popcll(bm & ((1ull << lane) - 1ull));
    }
  }
  // rank each (triangle, bin) pair within the workgroup
  uint32_t ent[kSmallBins], rank[kSmallBins];  // table entry (or kNoKey: own atomic) and rank
  uint32_t hits = 0;                           // bit j: candidate j overlaps its bin
#pragma unroll
  for (int j = 0; j < kSmallBins; j++) {
    ent[j] = kNoKey;
    rank[j] = 0;
    uint32_t bin = 0;
    if (small && j < bb.count() && bin_candidate(P, q, N, bb, j, bin)) {
      hits |= 1u << j;
      uint32_t h = (bin * 0x9E3779B1u) >> 22;  // 10-bit hash
      for (uint32_t probe = 0; probe < kProbe; probe++) {
        const uint32_t k = atomicCAS(&s_key[h], kNoKey, bin);
        if (k == kNoKey || k == bin) {
          ent[j] = h;
          rank[j] = atomicAdd(&s_cnt[h], 1u);
          break;
        }
        h = (h + 1u) & (kSlotTable - 1u);
      }
      if (ent[j] == kNoKey) rank[j] = gatomic_add(P.bin_small + ctr(P, bin), 1u);  // table overflow
    }
  }
  __syncthreads();
  {  // reserve every listed bin's run (all four returning atomics in flight together) and
     // the large triangles' run of the list
    static_assert(kSlotTable == 4 * kBlockTris, "four table entries per thread");
    uint32_t key[4], base[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const uint32_t e = threadIdx.x + (uint32_t)r * kBlockTris;
      key[r] = s_key[e];
      base[r] = key[r] != kNoKey ? gatomic_add(P.bin_small + ctr(P, key[r]), s_cnt[e]) : 0u;
    }
    const uint32_t nbig = s_big_n;
    if (threadIdx.x == 0 && nbig != 0u) s_big_base = gatomic_add(&P.stats->big_count, nbig);
#pragma unroll
    for (int r = 0; r < 4; r++)
      if (key[r] != kNoKey) s_base[threadIdx.x + (uint32_t)r * kBlockTris] = base[r];
  }
  __syncthreads();
  if (big) gst(P.big_list + s_big_base + big_rank, t);
  uint32_t nslot = 0;
#pragma unroll
  for (int j = 0; j < kSmallBins; j++) {
    if (!((hits >> j) & 1u)) continue;
    const uint32_t slot = ent[j] == kNoKey ? rank[j] : s_base[ent[j]] + rank[j];
    gst(P.tri_slot + (size_t)t * kSmallBins + nslot++, slot);
  }
}

hipError_t launch_setup(const VoxBatch& B, uint32_t blocks, hipStream_t s) {
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(k_setup, dim3(blocks), dim3(kBlockTris), 0, s, B, B.scenes);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// 2 large triangles: count (then scatter) their bins, one wavefront per triangle
// ---------------------------------------------------------------------------

// Large triangles with wavefront ids wave, wave + nwaves, ...; EMIT false counts their
// pairs per bin, EMIT true places them after the bin's small-triangle pairs.
template <bool EMIT>
__device__ __forceinline__ void big_pass(const VoxBatch& B, const VoxScene* __restrict__ scenes, uint32_t wave,
                                         uint32_t nwaves) {
  // one wavefront per triangle: most large triangles span a few dozen bins
  const int lane = (int)(threadIdx.x & 63u);
  for (uint32_t sc = 0; sc < B.nscenes; sc++) {
    const VoxScene& P = scenes[sc];
    const uint32_t nbig = gld(&P.stats->big_count);
    for (uint32_t i = wave; i < nbig; i += nwaves) {
      const uint32_t t = gld(P.big_list + i);
      const TriRec R = gld_rec(P.tris + t);
      int64_t q[3][3], N[3];
      load_q(R, q, N);
      const BinBox bb = bin_box(R, P.bin_shift);
      const int nbb = bb.count();
      for (int j = lane; j < nbb; j += 64) {
        uint32_t bin;
        if (!bin_candidate(P, q, N, bb, j, bin)) continue;
        if (!EMIT) {
          gatomic_add(P.bin_big + ctr(P, bin), 1u);
        } else {
          const uint32_t c = ctr(P, bin);
          const uint32_t pos = pair_base(B, P, bin) + gld(P.bin_small + c) + gatomic_add(P.big_cursor + c, 1u);
          if (pos < B.pair_cap) gst(B.pairs_all + pos, t);
        }
      }
    }
  }
}

__global__ __launch_bounds__(256) void k_big_count(VoxBatch B, const VoxScene* __restrict__ scenes) {
  big_pass<false>(B, scenes, blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6), gridDim.x * (blockDim.x >> 6));
}

hipError_t launch_big_count(const VoxBatch& B, uint32_t blocks, hipStream_t s) {
  hipLaunchKernelGGL(k_big_count, dim3(blocks), dim3(256), 0, s, B, B.scenes);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// 3 per-bin totals, 4 scan over every bin of the batch
// ---------------------------------------------------------------------------

// Device-coherent (agent scope) word accesses for the lookback between workgroups, which
// may run on different XCDs: they bypass the XCD's own L2.
__device__ __forceinline__ void agent_store(unsigned long long* p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void agent_store(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long agent_load(const unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t agent_load(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// VoxBatch::scan_spin polls of a predecessor's flag before a wait gives up (then the call
// takes the per-triangle path, which needs no scan): kScanSpin, or what VGI_SCAN_SPIN held
// when the context was created (tests set 0 to take that path). Workgroups are dispatched
// in index order and a workgroup waits only for lower ones, so a wait ends long before.

// Per bin: pairs = small + large, units, the packed (units << 32 | pairs) entry, the list
// of bins with several units, the scene's exact pair count; and the batch-wide exclusive
// scan of the packed entries in one pass (replacing a separate two-kernel scan): each
// workgroup scans its 256 bins, publishes its total (flag 1), finds its prefix by looking
// back over its predecessors 64 at a time (wavefront 0: a predecessor's inclusive prefix,
// flag 2, ends the look-back, totals of the ones after it are added), publishes its
// inclusive prefix (flag 2) and stores its bins' scanned entries. A value is stored and
// complete before its flag. Workgroups follow the batch's bin order (scenes in turn), so
// workgroup j's bins precede workgroup j + 1's.
__global__ __launch_bounds__(256) void k_bins(VoxBatch B, const VoxScene* __restrict__ scenes) {
  __shared__ unsigned long long s_wsum[4], s_prefix;
  const uint32_t j = blockIdx.x, nblk = gridDim.x;
  const uint2 blk = bin_block(B);
  const VoxScene& P = scenes[blk.x];
  const uint32_t nbins = P.nbins_axis * P.nbins_axis * P.nbins_axis;
  const uint32_t b = blk.y + threadIdx.x;
  const uint32_t lane = __lane_id(), wave = threadIdx.x >> 6;
  unsigned long long total = 0, packed = 0;
  if (b < nbins) {
    const uint32_t pairs = gld(P.bin_small + ctr(P, b)) + gld(P.bin_big + ctr(P, b));
    const uint32_t items = (pairs + P.chunk - 1) / P.chunk;
    const uint32_t units = (items + P.unit_items - 1) / P.unit_items;
    packed = ((unsigned long long)units << 32) | pairs;
    gst(P.bin_packed + b, packed);
    if (units > 1) gst(P.zero_list + gatomic_add(&P.stats->zero_count, 1u), b);
    total = pairs;
  }
  // exact 64-bit pair count per scene (the packed scan keeps only 32 bits of it)
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) total += __shfl_xor(total, o);
  if (lane == 0 && total) gatomic_add(&P.stats->pairs, total);
  // the workgroup's inclusive scan
  unsigned long long incl = packed;
#pragma unroll
  for (uint32_t o = 1; o < 64; o <<= 1) {
    const unsigned long long t = __shfl_up(incl, o);
    if (lane >= o) incl += t;
  }
  if (lane == 63u) s_wsum[wave] = incl;
  __syncthreads();
  unsigned long long woff = 0, agg = 0;
#pragma unroll
  for (uint32_t w = 0; w < 4; w++) {
    if (w < wave) woff += s_wsum[w];
    agg += s_wsum[w];
  }
  // the prefix of the bins before this workgroup's
  unsigned long long* agg_v = B.scan_vals;          // per workgroup: its total
  unsigned long long* inc_v = B.scan_vals + nblk;   // per workgroup: its inclusive prefix
  if (wave == 0) {
    unsigned long long prefix = 0;
    if (j == 0) {
      if (lane == 0) {
        agent_store(inc_v, agg);
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        agent_store(B.scan_flags, 2u);
      }
    } else {
      if (lane == 0) {
        agent_store(agg_v + j, agg);
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        agent_store(B.scan_flags + j, 1u);
      }
      bool failed = false;
      for (int top = (int)j - 1;; top -= 64) {  // lane l looks at workgroup top - l
        const int k = top - (int)lane;
        uint32_t f = 2u;  // before workgroup 0: an empty prefix
        unsigned long long v = 0;
        if (k >= 0) {
          f = 0u;
          for (uint32_t spin = 0; spin < B.scan_spin; spin++) {
            f = agent_load(B.scan_flags + k);
            if (f != 0u) break;
            __builtin_amdgcn_s_sleep(1);
          }
          if (f == 0u) {
            failed = true;
            f = 2u;
          } else {
            v = agent_load(f == 2u ? inc_v + k : agg_v + k);
          }
        }
        const uint64_t pm = __ballot(f == 2u);
        const int m = pm ? __ffsll((unsigned long long)pm) - 1 : 63;  // nearest with its prefix
        unsigned long long sv = (int)lane <= m ? v : 0ull;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) sv += __shfl_xor(sv, o);
        prefix += sv;
        if (pm) break;
      }
      failed = __any(failed);
      if (lane == 0) {
        agent_store(inc_v + j, prefix + agg);
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        agent_store(B.scan_flags + j, 2u);
        if (failed) agent_store(B.overflow + 3, 1u);  // the call takes the per-triangle path
      }
    }
    if (lane == 0) s_prefix = prefix;
  }
  __syncthreads();
  const unsigned long long prefix = s_prefix;
  if (b < nbins) gst(B.scan_all + P.bin_base + b, prefix + woff + incl - packed);
  if (j == nblk - 1u && threadIdx.x == 0) gst(B.scan_all + B.total_bins, prefix + agg);  // the batch's totals
}

hipError_t launch_bins(const VoxBatch& B, uint32_t blocks, hipStream_t s) {
  hipLaunchKernelGGL(k_bins, dim3(blocks), dim3(kBlockBins), 0, s, B, B.scenes);
  return hipGetLastError();
}

// The settle step, after the scan (the first wavefront of k_scatter; every wavefront of
// k_scatter takes the overflow decision itself, call_overflow): each scene's totals (work
// units << 32 | pairs of its bins) and statistics go to its pinned host record; the
// batch's exact pair count (a 64-bit sum of the per-scene counts, which the packed 32-bit
// field cannot hold beyond 2^32) and unit count are checked against the pair list and
// work map the host sized, and the overflow word tells steps 6-7 which path to take (it
// was a kernel of its own, k_settle: folding it into k_scatter saved a launch, about
// 3 us a frame). The batch totals (pairs, units,
// large triangles, split bins) go to the stream's pinned history, from which the host
// sizes later calls. Both are coherent host memory, written with system-scope stores; each
// record's call number (and the history's) is stored once the other words are complete,
// and the host reads a record once it finds the call's number there. No event is recorded
// between the call's kernels for it: on the room each event between two kernels left the
// GPU idle about 4 us (the two this replaced, 7.8 us of a 0.89 ms frame).
// a store to coherent host memory performed at system scope (written through the caches)
__device__ __forceinline__ void sys_store(unsigned long long* p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
// the same, issued once this wavefront's earlier stores have completed. Those are
// system-scope stores to coherent host memory, written through the caches, so their
// completion (vmcnt 0) is all the order the host's reader needs; a release store would
// also write back this XCD's whole L2 (buffer_wbl2) for data the host never reads.
__device__ __forceinline__ void sys_store_last(unsigned long long* p, unsigned long long v) {
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  sys_store(p, v);
}

// The call's overflow decision, taken by a whole wavefront (every lane returns it): the
// batch's exact pair count and its unit count against the pair list and the work map.
__device__ __forceinline__ uint32_t call_overflow(const VoxBatch& B, const VoxScene* __restrict__ scenes) {
  const uint32_t lane = __lane_id();
  unsigned long long pairs = 0;
  for (uint32_t sc = lane; sc < B.nscenes; sc += 64) pairs += gld(&scenes[sc].stats->pairs);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) pairs += __shfl_xor(pairs, o);
  const uint32_t units = (uint32_t)(gld(B.scan_all + B.total_bins) >> 32);
  // with pairs <= pair_cap < 2^32 no pair count carried into the unit field, so units is
  // exact; the fourth word is set when the bin scan's look-back gave up (never seen)
  return (pairs > B.pair_cap || units > B.work_cap || gld(B.overflow + 3) != 0u) ? 1u : 0u;
}

// The settle step, by one wavefront (the first of k_scatter's first workgroup).
__device__ void settle_wave(const VoxBatch& B, const VoxScene* __restrict__ scenes, uint32_t ovf) {
  const uint32_t lane = __lane_id();
  unsigned long long pairs = 0, big = 0, split = 0;
  for (uint32_t sc = lane; sc < B.nscenes; sc += 64) {
    pairs += gld(&scenes[sc].stats->pairs);
    big += gld(&scenes[sc].stats->big_count);
    split += gld(&scenes[sc].stats->zero_count);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    pairs += __shfl_xor(pairs, o);
    big += __shfl_xor(big, o);
    split += __shfl_xor(split, o);
  }
  const uint32_t units = (uint32_t)(gld(B.scan_all + B.total_bins) >> 32);
  for (uint32_t sc = lane; sc < B.nscenes; sc += 64) {
    const VoxScene& P = scenes[sc];
    const uint32_t nbins = P.nbins_axis * P.nbins_axis * P.nbins_axis;
    SceneVoxResult r;
    r.st.rejected = gld(&P.stats->rejected);
    r.st.degenerate = gld(&P.stats->degenerate);
    r.st.pairs = gld(&P.stats->pairs);
    r.st.big_count = gld(&P.stats->big_count);
    r.st.zero_count = gld(&P.stats->zero_count);
    r.totals = gld(B.scan_all + P.bin_base + nbins) - gld(B.scan_all + P.bin_base);
    r.overflow = ovf;
    r.seq = B.seq;
    r.pad[0] = r.pad[1] = 0;
    const unsigned long long* src = (const unsigned long long*)&r;
    unsigned long long* dst = (unsigned long long*)P.result;
    // the word holding seq last, after the others have completed: the host takes the record
    // as complete once it reads the call's number there
    constexpr int kSeqWord = (int)(offsetof(SceneVoxResult, seq) / 8);
#pragma unroll
    for (int q = 0; q < 8; q++)
      if (q != kSeqWord) sys_store(dst + q, src[q]);
    sys_store_last(dst + kSeqWord, src[kSeqWord]);
  }
  if (lane == 0) {
    gst(B.overflow, ovf);
    sys_store(B.host_hist + 0, pairs);
    sys_store(B.host_hist + 1, (unsigned long long)units);
    sys_store(B.host_hist + 2, big);
    sys_store(B.host_hist + 3, split);
    static_assert(kHostHistWords == 5, "four totals, then the call number");
    sys_store_last(B.host_hist + 4, (unsigned long long)B.seq);
  }
}


// work unit -> (scene, bin), from the scanned unit offsets; workgroup blk of the bins. Also
// the order in which k_voxelize takes the units: those of two or more chunks first, from
// the front of the order, then the single-chunk ones from the back. A unit's time grows
// with its chunks (on the room a 4-chunk unit runs about 27 us against 5-8 us for one
// chunk), and taken in bin order the long ones that start last set the launch's end.
// Within each class the order is that of the cursors' atomics; every unit's result is the
// same in any order.
__device__ __forceinline__ void workmap_block(const VoxBatch& B, const VoxScene* __restrict__ scenes, uint32_t blk_id) {
  const uint2 blk = bin_block(B, blk_id);
  const VoxScene& P = scenes[blk.x];
  const uint32_t nbins = P.nbins_axis * P.nbins_axis * P.nbins_axis;
  const uint32_t b = blk.y + threadIdx.x;
  uint32_t units = 0, first = 0, heavy = 0;
  if (b < nbins) {
    const unsigned long long packed = gld(P.bin_packed + b);
    units = (uint32_t)(packed >> 32);
    if (units) {
      first = first_item(B, P, b);
      // units 0 .. units-2 hold unit_items chunks each, the last one the rest
      const uint32_t items = ((uint32_t)packed + P.chunk - 1) / P.chunk;
      const uint32_t last = items - (units - 1) * P.unit_items;
      heavy = (P.unit_items >= 2u ? units - 1u : 0u) + (last >= 2u ? 1u : 0u);
      for (uint32_t i = 0; i < units; i++) gst2(B.work_map + first + i, make_uint2(blk.x, b));
    }
  }
  // places in the order: one cursor atomic per wavefront and class
  const uint32_t light = units - heavy;
  const uint32_t lane = __lane_id();
  uint32_t hs = heavy, ls = light;  // inclusive scans over the wavefront
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t h = __shfl_up(hs, o), l = __shfl_up(ls, o);
    if ((int)lane >= o) {
      hs += h;
      ls += l;
    }
  }
  uint32_t hbase = 0, lbase = 0;
  if (lane == 63u) {
    if (hs) hbase = gatomic_add(B.overflow + 1, hs);
    if (ls) lbase = gatomic_add(B.overflow + 2, ls);
  }
  hbase = __shfl(hbase, 63) + hs - heavy;
  lbase = __shfl(lbase, 63) + ls - light;
  if (units == 0u) return;
  const uint32_t nwork = (uint32_t)(gld(B.scan_all + B.total_bins) >> 32);  // the batch's units
  // heavy units (the first `heavy` of the bin, or all but the last) from the front, light
  // ones from the back: positions hbase.. and nwork-1-lbase, nwork-2-lbase, ..
  for (uint32_t i = 0; i < units; i++) {
    const bool h = i < heavy;  // units 0 .. heavy-1 are the heavy ones (the last may be light)
    const uint32_t pos = h ? hbase + i : nwork - 1u - (lbase + (i - heavy));
    gst(B.unit_order + pos, first + i);
  }
}

// ---------------------------------------------------------------------------
// 5 scatter small triangles into their slots
// ---------------------------------------------------------------------------

__device__ __forceinline__ void emit_small_block(const VoxBatch& B, const VoxScene* __restrict__ scenes, uint32_t blk_id) {
  const uint2 blk = tri_block(B, blk_id);
  const VoxScene& P = scenes[blk.x];
  const uint32_t t = blk.y + threadIdx.x;
  if (t >= P.ntri) return;
  const TriRec R = gld_rec(P.tris + t);
  if (!tri_ok(R)) return;
  const BinBox bb = bin_box(R, P.bin_shift);
  if (bb.count() > kSmallBins) return;
  int64_t q[3][3], N[3];
  load_q(R, q, N);
  uint32_t nslot = 0;
  for (int j = 0; j < bb.count(); j++) {
    uint32_t bin;
    if (!bin_candidate(P, q, N, bb, j, bin)) continue;
    const uint32_t pos = pair_base(B, P, bin) + gld(P.tri_slot + (size_t)t * kSmallBins + nslot++);
    if (pos < B.pair_cap) gst(B.pairs_all + pos, t);
  }
}

// Workgroups blk, blk + nblk, ... of the listed bins clear them.
__device__ __forceinline__ void zero_pass(const VoxBatch& Bt, const VoxScene* __restrict__ scenes, uint32_t blk,
                                          uint32_t nblk) {
  for (uint32_t sc = 0; sc < Bt.nscenes; sc++) {
    const VoxScene& P = scenes[sc];
    const uint32_t nz = gld(&P.stats->zero_count);
    const uint32_t nb = P.nbins_axis, B = 1u << P.bin_shift, nbr = P.n >> 3, bpb = B >> 3;
    for (uint32_t i = blk; i < nz; i += nblk) {
      const uint32_t bin = gld(P.zero_list + i);
      const uint32_t bx = bin % nb, by = (bin / nb) % nb, bz = bin / (nb * nb);
      for (uint32_t k = 0; k < bpb * bpb * bpb; k++) {
        const uint32_t bxx = bx * bpb + k % bpb, byy = by * bpb + (k / bpb) % bpb, bzz = bz * bpb + k / (bpb * bpb);
        const uint32_t brick = bxx + nbr * (byy + nbr * bzz);
        uint4* dst = (uint4*)(P.acc + (size_t)brick * 512u * kAccWords);
        for (uint32_t j = threadIdx.x; j < 1024; j += blockDim.x) gst4(dst + j, make_uint4(0, 0, 0, 0));
        if (threadIdx.x < 4) gst4((uint4*)(P.mask + (size_t)brick * 16u) + threadIdx.x, make_uint4(0, 0, 0, 0));
      }
    }
  }
}

// Overflow path: workgroup blk of the bins clears the accumulator records and occupancy
// words of its bins that have pairs, for the per-triangle kernel to add into.
__device__ __forceinline__ void clear_bins_block(const VoxBatch& Bt, const VoxScene* __restrict__ scenes,
                                                 uint32_t blk_id) {
  const uint2 blk = bin_block(Bt, blk_id);
  const VoxScene& P = scenes[blk.x];
  const uint32_t nb = P.nbins_axis, nbins = nb * nb * nb, B = 1u << P.bin_shift, nbr = P.n >> 3, bpb = B >> 3;
  const uint32_t bend = min(blk.y + kBlockBins, nbins);
  for (uint32_t bin = blk.y; bin < bend; bin++) {
    if ((uint32_t)gld(P.bin_packed + bin) == 0u) continue;  // no pairs: the bin stays empty
    const uint32_t bx = bin % nb, by = (bin / nb) % nb, bz = bin / (nb * nb);
    for (uint32_t k = 0; k < bpb * bpb * bpb; k++) {
      const uint32_t bxx = bx * bpb + k % bpb, byy = by * bpb + (k / bpb) % bpb, bzz = bz * bpb + k / (bpb * bpb);
      const uint32_t brick = bxx + nbr * (byy + nbr * bzz);
      uint4* dst = (uint4*)(P.acc + (size_t)brick * 512u * kAccWords);
      for (uint32_t j = threadIdx.x; j < 1024; j += blockDim.x) gst4(dst + j, make_uint4(0, 0, 0, 0));
      if (threadIdx.x < 4) gst4((uint4*)(P.mask + (size_t)brick * 16u) + threadIdx.x, make_uint4(0, 0, 0, 0));
    }
  }
}

// One launch after the scan, before k_voxelize: workgroups [0, nblk_b) map work units, the
// next nblk_t scatter small triangles, the next nblk_big place large triangles (one
// wavefront each), the rest clear the bins that several work units add into (none on most
// scenes). They write disjoint outputs. When the pair list or the work map would overflow,
// workgroups [0, nblk_b) clear every bin with pairs instead and the rest do nothing.
__global__ __launch_bounds__(256) void k_scatter(VoxBatch B, const VoxScene* __restrict__ scenes, uint32_t nblk_b,
                                                 uint32_t nblk_t, uint32_t nblk_big) {
  const uint32_t b = blockIdx.x;
  const uint32_t ovf = call_overflow(B, scenes);  // the same in every wavefront
  if (b == 0 && threadIdx.x < 64) settle_wave(B, scenes, ovf);
  if (ovf != 0u) {
    if (b < nblk_b) clear_bins_block(B, scenes, b);
    return;
  }
  if (b < nblk_b) {
    workmap_block(B, scenes, b);
  } else if (b < nblk_b + nblk_t) {
    emit_small_block(B, scenes, b - nblk_b);
  } else if (b < nblk_b + nblk_t + nblk_big) {
    const uint32_t w0 = b - nblk_b - nblk_t;
    big_pass<true>(B, scenes, w0 * (blockDim.x >> 6) + (threadIdx.x >> 6), nblk_big * (blockDim.x >> 6));
  } else {
    const uint32_t first = nblk_b + nblk_t + nblk_big;
    zero_pass(B, scenes, b - first, gridDim.x - first);
  }
}

hipError_t launch_scatter(const VoxBatch& B, uint32_t nblk_b, uint32_t nblk_t, uint32_t big_blocks,
                          uint32_t zero_blocks, hipStream_t s) {
  const uint32_t blocks = nblk_b + nblk_t + big_blocks + zero_blocks;
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(k_scatter, dim3(blocks), dim3(256), 0, s, B, B.scenes, nblk_b, nblk_t, big_blocks);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// clearing the bins that several work items add into (a k_scatter role)
// ---------------------------------------------------------------------------



// ---------------------------------------------------------------------------
// 7 brick voxelize
// ---------------------------------------------------------------------------

// A textured triangle's texture inputs, beside its coefficients in LDS.
struct TexCoef {
  float uv[6];
  const uint32_t* tex;
  uint32_t tw, th;
  float factor[4];
  uint32_t textured;
  uint32_t pad;
};

// Albedo of a covered voxel from a textured triangle (BasicMaterial's texture path,
// Pipelang/Internal/main.lua:188-205). e0, e1, e2: the barycentric numerators of vertices
// 0, 1, 2 at the voxel centre projected along the dominant axis (exact integers, >= 0
// inside the triangle); negative ones are clamped to zero. Single-precision steps in a
// fixed order (no contraction): uv = ((w0 uv0 + w1 uv1) + w2 uv2) / (w0 + w1 + w2), a
// non-finite uv taken as (0, 0); wrap addressing; bilinear between texel centres at level
// 0 (x = u W - 1/2), texel values c * (1/255), a + t (b - a) along x, then y; times the
// factor. Returns false for a discarded fragment (alpha < 0.05), else RGB8 in rgb
// (r | g << 8 | b << 16), each (uint)(clamp(c, 0, 1) * 255 + 0.5).
__device__ __forceinline__ bool tex_albedo(const TexCoef& T, int64_t e0, int64_t e1, int64_t e2, uint32_t& rgb) {
  e0 = imax64(e0, 0);
  e1 = imax64(e1, 0);
  e2 = imax64(e2, 0);
  const float w0 = (float)e0, w1 = (float)e1, w2 = (float)e2, ws = (float)(e0 + e1 + e2);
  float tu = ((w0 * T.uv[0] + w1 * T.uv[2]) + w2 * T.uv[4]) / ws;
  float tv = ((w0 * T.uv[1] + w1 * T.uv[3]) + w2 * T.uv[5]) / ws;
  if (!__builtin_isfinite(tu) || !__builtin_isfinite(tv)) tu = tv = 0.0f;
  const float fu = tu - floorf(tu), fv = tv - floorf(tv);
  const float x = fu * (float)T.tw - 0.5f, y = fv * (float)T.th - 0.5f;
  const float x0 = floorf(x), y0 = floorf(y);
  const float gx = x - x0, gy = y - y0;
  int ix0 = (int)x0, iy0 = (int)y0, ix1 = ix0 + 1, iy1 = iy0 + 1;  // in [-1, W], [-1, H]
  if (ix0 < 0) ix0 += (int)T.tw;
  if (iy0 < 0) iy0 += (int)T.th;
  if (ix1 >= (int)T.tw) ix1 -= (int)T.tw;
  if (iy1 >= (int)T.th) iy1 -= (int)T.th;
  const uint32_t r0 = (uint32_t)iy0 * T.tw, r1 = (uint32_t)iy1 * T.tw;  // < 2^28 texels
  const uint32_t t00 = gld(T.tex + r0 + (uint32_t)ix0), t10 = gld(T.tex + r0 + (uint32_t)ix1);
  const uint32_t t01 = gld(T.tex + r1 + (uint32_t)ix0), t11 = gld(T.tex + r1 + (uint32_t)ix1);
  const float k255 = 1.0f / 255.0f;
  float col[4];
#pragma unroll
  for (int ch = 0; ch < 4; ch++) {
    const float a = (float)((t00 >> (8 * ch)) & 0xFFu) * k255, b = (float)((t10 >> (8 * ch)) & 0xFFu) * k255;
    const float c = (float)((t01 >> (8 * ch)) & 0xFFu) * k255, d = (float)((t11 >> (8 * ch)) & 0xFFu) * k255;
    const float top = a + gx * (b - a), bot = c + gx * (d - c);
    col[ch] = (top + gy * (bot - top)) * T.factor[ch];
  }
  if (col[3] < 0.05f) return false;
  uint32_t out = 0;
#pragma unroll
  for (int ch = 0; ch < 3; ch++) out |= (uint32_t)(fminf(fmaxf(col[ch], 0.0f), 1.0f) * 255.0f + 0.5f) << (8 * ch);
  rgb = out;
  return true;
}

__device__ __forceinline__ void load_texcoef(const VoxScene& P, uint32_t t, TexCoef& C) {
  const TriTex X = gld_rec(P.tritex + t);
  C.textured = X.textured;
  if (!X.textured) return;
#pragma unroll
  for (int i = 0; i < 6; i++) C.uv[i] = X.uv[i];
  const MeshDev* M = P.meshes + X.mesh;
  C.tex = (const uint32_t*)gld((const unsigned long long*)&M->tex);
  C.tw = gld(&M->tw);
  C.th = gld(&M->th);
#pragma unroll
  for (int i = 0; i < 4; i++) C.factor[i] = gld(&M->factor[i]);
}

// The same numerators from the snapped vertices (the per-triangle path):
// sign(N_u) * cross(q_j - q_i, c - q_i) for the edge (i, j) opposite vertex k.
__device__ __forceinline__ void voxel_bary(const int64_t (&q)[3][3], const int64_t (&N)[3], int u, int x, int y,
                                           int z, int64_t (&e)[3]) {
  const int A = (u + 1) % 3, B = (u + 2) % 3;
  const int v[3] = {x, y, z};
  const int64_t cA = (int64_t)v[A] * kFixedOne + kFixedOne / 2, cB = (int64_t)v[B] * kFixedOne + kFixedOne / 2;
  const int64_t sg = N[u] < 0 ? -1 : 1;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const int i = (k + 1) % 3, j = (k + 2) % 3;
    const int64_t dA = q[j][A] - q[i][A], dB = q[j][B] - q[i][B];
    const int64_t pA = cA - q[i][A], pB = cB - q[i][B];
    e[k] = sg * (dA * pB - dB * pA);
  }
}

__device__ __forceinline__ unsigned long long inc0_of(uint32_t rgb) {
  return (unsigned long long)(rgb & 0xFFu) | ((unsigned long long)((rgb >> 8) & 0xFFu) << 16) |
         ((unsigned long long)((rgb >> 16) & 0xFFu) << 32) | (1ull << 48);
}

// Per-triangle coefficients in LDS. Edge function of projection u (up axis), edge i:
//   E = eA * v_A + eB * v_B + eC >= 0   (v = integer voxel coordinates)
// plane: s = N256 . v;  overlap iff s + pmin <= 0 <= s + pmax.
struct TriCoef {
  int64_t N256[3];
  int64_t pmin, pmax;
  int64_t eC[9];
  int32_t eA[9];
  int32_t eB[9];
  float ra, rb, zq0, q0a, q0b;
  int16_t lo[3], hi[3];    // voxel box clipped to the grid
  int16_t clo[3], chi[3];  // ... and to the current tile
  uint32_t axis;
  unsigned long long inc0, inc1;
};

// Barycentric numerators of vertices 0, 1, 2 at the centre of column (va, vb) of the
// triangle's dominant projection u: edge i (vertex i to i + 1, opposite vertex i + 2)
// evaluated without the box-corner offset of the coverage test,
// eA va + eB vb + eC - max(0, eA) - max(0, eB) + (eA + eB) / 2 (eA, eB are even).
__device__ __forceinline__ void column_bary(const TriCoef& C, int u, int va, int vb, int64_t (&e)[3]) {
  int64_t f[3];
#pragma unroll
  for (int i = 0; i < 3; i++) {
    const int64_t ea = C.eA[u * 3 + i], eb = C.eB[u * 3 + i];
    f[i] = ea * va + eb * vb + C.eC[u * 3 + i] - imax64(0, ea) - imax64(0, eb) + ((ea + eb) >> 1);
  }
  e[0] = f[1];
  e[1] = f[2];
  e[2] = f[0];
}

__device__ __forceinline__ int sel3(int a, int b, int c, int i) { return i == 0 ? a : (i == 1 ? b : c); }

// The three edge functions of projection axis U (compile-time, so every index is static).
template <int U>
__device__ __forceinline__ void make_edges(const TriRec& R, const int64_t (&N)[3], TriCoef& C) {
  constexpr int A = (U + 1) % 3, B = (U + 2) % 3;
  const bool neg = N[U] < 0;  // edge normal (-dB, dA) * sign(N[U]), |dA|, |dB| <= 2^19
#pragma unroll
  for (int i = 0; i < 3; i++) {
    const int j = (i + 1) % 3;
    const int32_t dB = R.q[j * 3 + B] - R.q[i * 3 + B], dA = R.q[j * 3 + A] - R.q[i * 3 + A];
    const int32_t na = neg ? dB : -dB;
    const int32_t nb = neg ? -dA : dA;
    const int64_t ea = shl64(na, VGI_FIXED_BITS), eb = shl64(nb, VGI_FIXED_BITS);
    C.eA[U * 3 + i] = (int32_t)ea;
    C.eB[U * 3 + i] = (int32_t)eb;
    C.eC[U * 3 + i] = -(mul32x32(na, R.q[i * 3 + A]) + mul32x32(nb, R.q[i * 3 + B])) + imax64(0, ea) +
                      imax64(0, eb);
  }
}

// Coefficients of one triangle, built by the four wavefronts of a workgroup together
// (part is wave-uniform): part 0 writes the plane, the dominant-axis depth terms, the box
// and the increments, part 1 + u the three edge functions of projection axis u. The parts
// write disjoint fields of C, each with the same arithmetic as one thread computing all.
__device__ void make_coef(const TriRec& R, TriCoef& C, uint32_t part) {
  int64_t q[3][3], N[3];
  load_q(R, q, N);
  if (part != 0) {
    if (part == 1) make_edges<0>(R, N, C);
    else if (part == 2) make_edges<1>(R, N, C);
    else make_edges<2>(R, N, C);
    return;
  }
  int64_t c0 = -(N[0] * q[0][0] + N[1] * q[0][1] + N[2] * q[0][2]);
  int64_t pmin = c0, pmax = c0;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const int64_t p = N[a] * kFixedOne;
    C.N256[a] = p;
    pmin += imin64(0, p);
    pmax += imax64(0, p);
  }
  C.pmin = pmin;
  C.pmax = pmax;
  const int u = R.axis;
  const int A = (u + 1) % 3, B = (u + 2) % 3;
  const float nu = (float)(u == 0 ? N[0] : (u == 1 ? N[1] : N[2]));
  const float na = (float)(A == 0 ? N[0] : (A == 1 ? N[1] : N[2]));
  const float nbv = (float)(B == 0 ? N[0] : (B == 1 ? N[1] : N[2]));
  C.ra = na / nu;
  C.rb = nbv / nu;
  C.zq0 = (float)sel3(R.q[0], R.q[1], R.q[2], u);
  C.q0a = (float)sel3(R.q[0], R.q[1], R.q[2], A);
  C.q0b = (float)sel3(R.q[0], R.q[1], R.q[2], B);
#pragma unroll
  for (int a = 0; a < 3; a++) {
    C.lo[a] = R.lo[a];
    C.hi[a] = R.hi[a];
  }
  C.axis = (uint32_t)u;
  const uint32_t r = R.rgb & 0xFFu, g = (R.rgb >> 8) & 0xFFu, b = (R.rgb >> 16) & 0xFFu;
  C.inc0 = (unsigned long long)r | ((unsigned long long)g << 16) | ((unsigned long long)b << 32) |
           (1ull << 48);
  C.inc1 = (unsigned long long)(uint32_t)(R.nq[0] + kNormalBias) |
           ((unsigned long long)(uint32_t)(R.nq[1] + kNormalBias) << 20) |
           ((unsigned long long)(uint32_t)(R.nq[2] + kNormalBias) << 40);
}

__device__ __forceinline__ bool edge_ok(const TriCoef& C, int u, int va, int vb) {
  bool ok = true;
#pragma unroll
  for (int i = 0; i < 3; i++) {
    const int64_t e = (int64_t)C.eA[u * 3 + i] * va + (int64_t)C.eB[u * 3 + i] * vb + C.eC[u * 3 + i];
    ok = ok && (e >= 0);
  }
  return ok;
}

// LDS tile accumulator: two arrays of packed 64-bit increments (RGB+count, normals) with
// padded pitches, entry x + PY*y + PZ*z. Unpadded power-of-two pitches put lanes that
// step along y or z (the column direction of X- and Y-dominant triangles) into one LDS
// bank; with PY = TS+1 and PZ = (TS+1)*TS+1 (odd multiples of 8 bytes, 18 banks apart)
// sixteen lanes of an 8-byte access hit sixteen distinct bank pairs along any axis.
// Writes one voxel's record (R, G, B sums, count, three normal sums) into its 8^3 brick,
// and with the wavefront's ballot the brick's occupancy words for these 64 voxels (l is
// the voxel's index in the brick; lanes hold 64 consecutive indices). Alone in its bin
// (multi false) the record and both mask words are stored; otherwise they are added
// (the bin was cleared before): word pairs as 64-bit atomics, which is exact with fewer
// than 2^21 triangles per voxel (no 32-bit word's total reaches 2^32, so no carry
// crosses a word: R, G, B add <= 255 each, normals <= 2048 each).
__device__ __forceinline__ void flush_voxel(const VoxScene& P, uint32_t brick, uint32_t l, const uint32_t (&rec)[7],
                                            bool multi, uint32_t lane) {
  const uint32_t c = rec[3];
  if (c != 0) {
    uint32_t* dst = P.acc + ((size_t)brick * 512u + l) * kAccWords;
    if (!multi) {
      gst4(dst, make_uint4(rec[0], rec[1], rec[2], c));
      gst4(dst + 4, make_uint4(rec[4], rec[5], rec[6], 0u));
    } else {
      unsigned long long* d64 = (unsigned long long*)dst;
      gatomic_add(d64 + 0, (unsigned long long)rec[0] | ((unsigned long long)rec[1] << 32));
      gatomic_add(d64 + 1, (unsigned long long)rec[2] | ((unsigned long long)c << 32));
      gatomic_add(d64 + 2, (unsigned long long)rec[4] | ((unsigned long long)rec[5] << 32));
      gatomic_add(d64 + 3, (unsigned long long)rec[6]);
    }
  }
  const unsigned long long bits = __ballot(c != 0);  // bit j: voxel l - lane + j
  if (lane < 2) {
    const uint32_t word = lane == 0 ? (uint32_t)bits : (uint32_t)(bits >> 32);
    uint32_t* m = P.mask + (size_t)brick * 16u + ((l - lane) >> 5) + lane;
    if (!multi) gst(m, word);  // alone in its bin: every word is written, zeros too
    else if (word) gatomic_or(m, word);
  }
}

template <int TS>
struct TileLayout {
  static constexpr int PY = TS + 1;
  static constexpr int PZ = (TS + 1) * TS + 1;
  static constexpr int ENTRIES = (TS - 1) * (1 + PY + PZ) + 1;
};

// Registers: the 8^3 tile's unit accumulators (14 per thread) are held at 6 waves per SIMD
// (80 VGPRs; 6 resident workgroups per CU, as many as LDS allows at 25.5 KB each), which
// spills one 64-bit value per thread to scratch, reloaded at the flush. The kernel is
// latency-bound (dependent loads and barriers per chunk), so resident workgroups count:
// 6 instead of 5 took the room's launch from 80 to 75 us.
// TEX: the batch has a textured mesh; each triangle's texture inputs sit beside its
// coefficients and a textured triangle samples its albedo per column (kept apart so that
// untextured batches run the kernel without the sampler's registers and LDS).
template <int TS, bool TEX>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(TS == 8 && !TEX ? 6 : 1)))
void k_voxelize(VoxBatch Bt, const VoxScene* __restrict__ scenes) {
  constexpr int TV = TS * TS * TS;
  constexpr int NBK = TS / 8;  // 8^3 accumulator bricks per tile edge
  using L = TileLayout<TS>;
  __shared__ unsigned long long s_acc0[L::ENTRIES];
  __shared__ unsigned long long s_acc1[L::ENTRIES];
  __shared__ TriCoef s_tri[kMaxChunk];
  __shared__ TexCoef s_tex[TEX ? kMaxChunk : 1];
  __shared__ uint32_t s_pref[kMaxChunk + 1];
  static_assert(kMaxChunk == 64, "a chunk's triangles are the 64 lanes of each of the 4 wavefronts");

  const uint32_t tid = threadIdx.x;
  const uint32_t lane = tid & 63u;
  if (gld(Bt.overflow) != 0u) return;  // the per-triangle kernel voxelizes this call
  // the batch's work units, from the scan (the launch is sized from earlier calls: normally
  // one unit per workgroup, and workgroups take further units in turn when it falls short)
  const uint32_t nwork = (uint32_t)(gld(Bt.scan_all + Bt.total_bins) >> 32);

  for (uint32_t w = blockIdx.x; w < nwork; w += gridDim.x) {
    const uint32_t u = gld(Bt.unit_order + w);  // the longest units first
    const uint2 item = gld2(Bt.work_map + u);
    const VoxScene& P = scenes[item.x];
    const uint32_t bin = item.y;
    const uint32_t nb = P.nbins_axis, B = 1u << P.bin_shift, nbr = P.n >> 3;
    const uint32_t sub = B / TS;  // sub-tiles per bin edge
    const unsigned long long packed = gld(P.bin_packed + bin);
    const uint32_t cnt_bin = (uint32_t)packed, units = (uint32_t)(packed >> 32);
    const uint32_t unit = u - first_item(Bt, P, bin);
    const uint32_t base = pair_base(Bt, P, bin);
    const uint32_t items_bin = (cnt_bin + P.chunk - 1) / P.chunk;
    const uint32_t it0 = unit * P.unit_items, it1 = min(it0 + P.unit_items, items_bin);
    const bool multi = units > 1;  // other workgroups add into this bin: atomics
    const uint32_t bx = bin % nb, by = (bin / nb) % nb, bz = bin / (nb * nb);
    // A bin that is one 8^3 tile keeps its sums in registers across the chunks of its unit
    // (each thread owns tile entries tid and tid + 256) and is flushed once.
    const bool regs = TS == 8 && sub == 1;
    uint32_t racc[2][7];
#pragma unroll
    for (int r = 0; r < 2; r++)
#pragma unroll
      for (int k = 0; k < 7; k++) racc[r][k] = 0u;

    for (uint32_t it = it0; it < it1; it++) {
      const uint32_t first = base + it * P.chunk;
      const uint32_t cnt = min(P.chunk, cnt_bin - it * P.chunk);
      if ((tid & 63u) < cnt) {  // triangle tid % 64, coefficient part tid / 64
        const uint32_t t = gld(Bt.pairs_all + first + (tid & 63u));
        const TriRec R = gld_rec(P.tris + t);
        make_coef(R, s_tri[tid & 63u], tid >> 6);
        if constexpr (TEX)
          if ((tid >> 6) == 3u) load_texcoef(P, t, s_tex[tid & 63u]);
      }

      for (uint32_t st = 0; st < sub * sub * sub; st++) {
        const int tx0 = (int)(bx * B + TS * (st % sub));
        const int ty0 = (int)(by * B + TS * ((st / sub) % sub));
        const int tz0 = (int)(bz * B + TS * (st / (sub * sub)));
        for (int i = tid; i < L::ENTRIES; i += 256) {
          s_acc0[i] = 0ull;
          s_acc1[i] = 0ull;
        }
        // clip to the tile, count columns of the dominant projection, scan (wave 0)
        if (tid < 64) {
          uint32_t cols = 0;
          if (tid < cnt) {
            TriCoef& C = s_tri[tid];
            const int t0[3] = {tx0, ty0, tz0};
            bool emp = false;
#pragma unroll
            for (int a = 0; a < 3; a++) {
              const int lo = max((int)C.lo[a], t0[a]), hi = min((int)C.hi[a], t0[a] + TS - 1);
              C.clo[a] = (int16_t)lo;
              C.chi[a] = (int16_t)hi;
              emp = emp || lo > hi;
            }
            if (!emp) {
              const int u = (int)C.axis, A = (u + 1) % 3, Bx = (u + 2) % 3;
              cols = (uint32_t)(C.chi[A] - C.clo[A] + 1) * (uint32_t)(C.chi[Bx] - C.clo[Bx] + 1);
            }
          }
          uint32_t incl = cols;
#pragma unroll
          for (int o = 1; o < 64; o <<= 1) {
            const uint32_t v = __shfl_up(incl, o);
            if ((int)tid >= o) incl += v;
          }
          s_pref[tid + 1] = incl;
          if (tid == 0) s_pref[0] = 0;
        }
        __syncthreads();
        const uint32_t total = s_pref[cnt];
        for (uint32_t wi = tid; wi < total; wi += 256) {
          int lo = 0, hi = (int)cnt - 1;
          while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (s_pref[mid] <= wi) lo = mid;
            else hi = mid - 1;
          }
          const TriCoef& C = s_tri[lo];
          const uint32_t local = wi - s_pref[lo];
          const int u = (int)C.axis, A = (u + 1) % 3, Bx = (u + 2) % 3;
          const int wa = C.chi[A] - C.clo[A] + 1;
          const int va = C.clo[A] + (int)(local % (uint32_t)wa);
          const int vb = C.clo[Bx] + (int)(local / (uint32_t)wa);
          if (!edge_ok(C, u, va, vb)) continue;
          unsigned long long inc0 = C.inc0;
          if constexpr (TEX) {
            const TexCoef& X = s_tex[lo];
            if (X.textured) {  // one sample per column: every voxel of it projects to its centre
              int64_t e[3];
              column_bary(C, u, va, vb, e);
              uint32_t rgb;
              if (!tex_albedo(X, e[0], e[1], e[2], rgb)) continue;  // discarded
              inc0 = inc0_of(rgb);
            }
          }
          const float zc = C.zq0 - C.ra * ((float)(va * kFixedOne + kFixedOne / 2) - C.q0a) -
                           C.rb * ((float)(vb * kFixedOne + kFixedOne / 2) - C.q0b);
          const float wd = 0.5f * (float)kFixedOne * (fabsf(C.ra) + fabsf(C.rb));
          int k0 = (int)floorf((zc - wd - 4.0f) * (1.0f / kFixedOne));
          int k1 = (int)floorf((zc + wd + 4.0f) * (1.0f / kFixedOne));
          k0 = max(k0, (int)C.clo[u]);
          k1 = min(k1, (int)C.chi[u]);
          const int u1 = (u + 1) % 3, u2 = (u + 2) % 3;
          for (int k = k0; k <= k1; k++) {
            const int vx = (u == 0) ? k : (A == 0 ? va : vb);
            const int vy = (u == 1) ? k : (A == 1 ? va : vb);
            const int vz = (u == 2) ? k : (A == 2 ? va : vb);
            const int64_t sv = C.N256[0] * vx + C.N256[1] * vy + C.N256[2] * vz;
            if (sv + C.pmin > 0 || sv + C.pmax < 0) continue;
            if (!edge_ok(C, u1, sel3(vx, vy, vz, (u1 + 1) % 3), sel3(vx, vy, vz, (u1 + 2) % 3))) continue;
            if (!edge_ok(C, u2, sel3(vx, vy, vz, (u2 + 1) % 3), sel3(vx, vy, vz, (u2 + 2) % 3))) continue;
            const int pi = (vx - tx0) + L::PY * (vy - ty0) + L::PZ * (vz - tz0);
            atomicAdd(&s_acc0[pi], inc0);
            atomicAdd(&s_acc1[pi], C.inc1);
          }
        }
        __syncthreads();
        if (regs) {
          // add this chunk's tile into the registers (entries in brick order, as flushed)
#pragma unroll
          for (int r = 0; r < 2; r++) {
            const uint32_t l = tid + 256u * (uint32_t)r;
            const int pi = (int)(l & 7u) + L::PY * (int)((l >> 3) & 7u) + L::PZ * (int)(l >> 6);
            const unsigned long long w0 = s_acc0[pi], w1 = s_acc1[pi];
            racc[r][0] += (uint32_t)(w0 & 0xFFFFu);
            racc[r][1] += (uint32_t)((w0 >> 16) & 0xFFFFu);
            racc[r][2] += (uint32_t)((w0 >> 32) & 0xFFFFu);
            racc[r][3] += (uint32_t)(w0 >> 48);
            racc[r][4] += (uint32_t)(w1 & 0xFFFFFu);
            racc[r][5] += (uint32_t)((w1 >> 20) & 0xFFFFFu);
            racc[r][6] += (uint32_t)((w1 >> 40) & 0xFFFFFu);
          }
          __syncthreads();  // the next chunk clears the tile and overwrites the coefficients
          continue;
        }
        // flush the tile in brick-major order (64 consecutive entries of a wavefront are one
        // 64-voxel slab of one 8^3 brick): records of occupied voxels, and the brick's
        // occupancy mask, two words per wavefront from a ballot
        for (int i = tid; i < TV; i += 256) {  // trip count is the same for every lane
          const uint32_t bk = (uint32_t)i >> 9, l = (uint32_t)i & 511u;
          const uint32_t bkx = bk % NBK, bky = (bk / NBK) % NBK, bkz = bk / (NBK * NBK);
          const uint32_t lx = bkx * 8 + (l & 7u), ly = bky * 8 + ((l >> 3) & 7u), lz = bkz * 8 + (l >> 6);
          const int pi = (int)lx + L::PY * (int)ly + L::PZ * (int)lz;
          const unsigned long long w0 = s_acc0[pi];
          const unsigned long long w1 = s_acc1[pi];
          const uint32_t rec[7] = {(uint32_t)(w0 & 0xFFFFu),  (uint32_t)((w0 >> 16) & 0xFFFFu),
                                   (uint32_t)((w0 >> 32) & 0xFFFFu), (uint32_t)(w0 >> 48),
                                   (uint32_t)(w1 & 0xFFFFFu), (uint32_t)((w1 >> 20) & 0xFFFFFu),
                                   (uint32_t)((w1 >> 40) & 0xFFFFFu)};
          const uint32_t brick = ((uint32_t)tx0 >> 3) + bkx +
                                 nbr * (((uint32_t)ty0 >> 3) + bky + nbr * (((uint32_t)tz0 >> 3) + bkz));
          flush_voxel(P, brick, l, rec, multi, lane);
        }
        __syncthreads();
      }
    }
    if (regs) {  // one flush per unit
      const uint32_t brick = bx + nbr * (by + nbr * bz);  // the bin is one brick
#pragma unroll
      for (int r = 0; r < 2; r++) flush_voxel(P, brick, tid + 256u * (uint32_t)r, racc[r], multi, lane);
    }
  }
}

hipError_t launch_voxelize(const VoxBatch& B, uint32_t tile, uint32_t grid_blocks, hipStream_t s) {
  if (grid_blocks == 0) return hipSuccess;
  static_assert(TileLayout<8>::ENTRIES == 582 && TileLayout<16>::ENTRIES == 4366, "tile layout");
  if (B.textured) {
    if (tile == 8) hipLaunchKernelGGL((k_voxelize<8, true>), dim3(grid_blocks), dim3(256), 0, s, B, B.scenes);
    else hipLaunchKernelGGL((k_voxelize<16, true>), dim3(grid_blocks), dim3(256), 0, s, B, B.scenes);
  } else {
    if (tile == 8) hipLaunchKernelGGL((k_voxelize<8, false>), dim3(grid_blocks), dim3(256), 0, s, B, B.scenes);
    else hipLaunchKernelGGL((k_voxelize<16, false>), dim3(grid_blocks), dim3(256), 0, s, B, B.scenes);
  }
  return hipGetLastError();
}


// ---------------------------------------------------------------------------
// 7 overflow path: per-triangle voxelize with global atomics
// ---------------------------------------------------------------------------

// Runs only when k_scatter set the overflow word (and cleared every bin with pairs). A kernel of its own: inlined into k_voxelize it would spill that kernel's
// registers, called from it the call would cut its occupancy. Wavefront w of the launch
// takes triangles w, w + nwaves, ... of every scene; its
// lanes walk the triangle's voxel box, keep the voxels the exact closed-box test accepts
// (the same test as the bin pass, at voxel size) and add the triangle's increments: R, G,
// B and the count, then the three biased normal components, as four 64-bit atomics (exact
// with fewer than 2^21 triangles per voxel: no 32-bit word's total reaches 2^32), and set
// the voxel's occupancy bit. Integer sums do not depend on order, so the accumulator is
// the same as the binned path's.
// Not inlined into k_voxelize, whose register budget (6 workgroups per CU) it would
// otherwise share.
template <bool TEX>
__global__ __launch_bounds__(256) void k_voxelize_fallback(VoxBatch B, const VoxScene* __restrict__ scenes) {
  if (gld(B.overflow) == 0u) return;
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6), nwaves = gridDim.x * (blockDim.x >> 6);
  for (uint32_t sc = 0; sc < B.nscenes; sc++) {
    const VoxScene& P = scenes[sc];
    const uint32_t nbr = P.n >> 3;
    for (uint32_t t = wave; t < P.ntri; t += nwaves) {
      const TriRec R = gld_rec(P.tris + t);
      if (!tri_ok(R)) continue;
      int64_t q[3][3], N[3];
      load_q(R, q, N);
      const int nx = R.hi[0] - R.lo[0] + 1, ny = R.hi[1] - R.lo[1] + 1, nz = R.hi[2] - R.lo[2] + 1;
      const uint32_t r = R.rgb & 0xFFu, g = (R.rgb >> 8) & 0xFFu, b = (R.rgb >> 16) & 0xFFu;
      unsigned long long i0 = (unsigned long long)r | ((unsigned long long)g << 32);
      unsigned long long i1 = (unsigned long long)b | (1ull << 32);
      TexCoef X;
      X.textured = 0;
      if constexpr (TEX) load_texcoef(P, t, X);
      const unsigned long long i2 = (unsigned long long)(uint32_t)(R.nq[0] + kNormalBias) |
                                    ((unsigned long long)(uint32_t)(R.nq[1] + kNormalBias) << 32);
      const unsigned long long i3 = (unsigned long long)(uint32_t)(R.nq[2] + kNormalBias);
      const uint32_t total = (uint32_t)nx * (uint32_t)ny * (uint32_t)nz;
      for (uint32_t j = lane; j < total; j += 64) {
        const int x = R.lo[0] + (int)(j % (uint32_t)nx), y = R.lo[1] + (int)((j / (uint32_t)nx) % (uint32_t)ny),
                  z = R.lo[2] + (int)(j / ((uint32_t)nx * (uint32_t)ny));
        const int32_t bb[3] = {x << VGI_FIXED_BITS, y << VGI_FIXED_BITS, z << VGI_FIXED_BITS};
        if (!sat_box(q, N, bb, VGI_FIXED_BITS)) continue;
        if constexpr (TEX) {
          if (X.textured) {
            int64_t e[3];
            voxel_bary(q, N, R.axis, x, y, z, e);
            uint32_t rgb;
            if (!tex_albedo(X, e[0], e[1], e[2], rgb)) continue;  // discarded: nothing, not even the bit
            i0 = (unsigned long long)(rgb & 0xFFu) | ((unsigned long long)((rgb >> 8) & 0xFFu) << 32);
            i1 = (unsigned long long)((rgb >> 16) & 0xFFu) | (1ull << 32);
          }
        }
        const uint32_t brick = ((uint32_t)x >> 3) + nbr * (((uint32_t)y >> 3) + nbr * ((uint32_t)z >> 3));
        const uint32_t l = ((uint32_t)x & 7u) + 8u * ((uint32_t)y & 7u) + 64u * ((uint32_t)z & 7u);
        unsigned long long* d64 = (unsigned long long*)(P.acc + ((size_t)brick * 512u + l) * kAccWords);
        gatomic_add(d64 + 0, i0);
        gatomic_add(d64 + 1, i1);
        gatomic_add(d64 + 2, i2);
        gatomic_add(d64 + 3, i3);
        gatomic_or(P.mask + (size_t)brick * 16u + (l >> 5), 1u << (l & 31u));
      }
    }
  }
}

hipError_t launch_voxelize_fallback(const VoxBatch& B, uint32_t grid_blocks, hipStream_t s) {
  if (grid_blocks == 0) return hipSuccess;
  if (B.textured) hipLaunchKernelGGL(k_voxelize_fallback<true>, dim3(grid_blocks), dim3(256), 0, s, B, B.scenes);
  else hipLaunchKernelGGL(k_voxelize_fallback<false>, dim3(grid_blocks), dim3(256), 0, s, B, B.scenes);
  return hipGetLastError();
}

}  // namespace vgi
