// voxelize.hip — conservative triangle voxelization for gfx950.
//
// Replaces the reference's voxelize raster pipeline: the pass-through VS, the VoxelGS
// dominant-axis projection (Pipelang/Internal/main.lua:83-144) and VoxelPS's RG16UI
// imageStore with last-writer-wins (main.lua:242-275), issued one draw per primitive by
// CVoxelizeRenderer::Render (Foreground/Renderer/VoxelizeRenderer.cpp:81-127).
//
// Pipeline (all on one stream):
//   k_setup_count  one thread per triangle: fetch + model transform + snap to 1/256 voxel,
//                  exact integer normal, dominant axis (VoxelGS rule), quantized normal,
//                  voxel bounding box; exact box test against every bin (bin = B^3
//                  voxels) in the box; count (triangle, bin) pairs per bin.
//   scan_bins      exclusive scan of the per-bin counts (rocPRIM via hipCUB).
//   k_emit         the same loop again, scattering triangle ids into per-bin lists.
//   k_worklist     one wave per non-empty bin: split its list into chunks of <= 64
//                  triangles (work items); zero the bin's accumulator when it has >1 chunk.
//   k_voxelize     persistent work loop; per item: the chunk's triangle edge/plane
//                  coefficients go to LDS, columns of the dominant-axis projection are
//                  spread over the workgroup, each column tests its 1-4 candidate voxels
//                  with the exact separating-axis test and adds packed 64-bit increments
//                  into a per-tile LDS accumulator; the tile is then flushed: plain
//                  vector stores when the bin has one chunk, integer atomics otherwise.
//
// All coverage decisions are exact integer arithmetic on snapped vertices, so the
// voxel set and the integer sums do not depend on order, rounding or thread mapping.
#include <hipcub/hipcub.hpp>

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

__device__ __forceinline__ uint32_t find_mesh(const MeshDev* __restrict__ m, uint32_t nmesh,
                                              uint32_t t) {
  uint32_t lo = 0, hi = nmesh - 1;
  while (lo < hi) {
    uint32_t mid = (lo + hi + 1) >> 1;
    if (m[mid].tri_base <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Exact closed-box test of a triangle against [b, b + d]^3 (fixed-point units), with
// the bounding boxes already known to overlap. Magnitudes: |q| <= 2^18, |edge| <= 2^19,
// |N| <= 2^39, d <= 2^13: every product and sum fits in int64.
__device__ __forceinline__ bool sat_box(const int64_t (&q)[3][3], const int64_t (&N)[3],
                                        const int64_t (&b)[3], int64_t d) {
  int64_t s_lo = 0, s_hi = 0;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    int64_t t0 = N[a] * (b[a] - q[0][a]);
    int64_t dd = N[a] * d;
    s_lo += t0 + imin64(0, dd);
    s_hi += t0 + imax64(0, dd);
  }
  if (s_lo > 0 || s_hi < 0) return false;
#pragma unroll
  for (int u = 0; u < 3; u++) {
    const int A = (u + 1) % 3, B = (u + 2) % 3;
    const int64_t sgn = N[u] < 0 ? -1 : 1;
#pragma unroll
    for (int i = 0; i < 3; i++) {
      const int j = (i + 1) % 3;
      int64_t na = -(q[j][B] - q[i][B]) * sgn;
      int64_t nb = (q[j][A] - q[i][A]) * sgn;
      int64_t e = na * (b[A] - q[i][A]) + nb * (b[B] - q[i][B]) + imax64(0, na * d) +
                  imax64(0, nb * d);
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
  int64_t e0[3], e1[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    e0[a] = q[1][a] - q[0][a];
    e1[a] = q[2][a] - q[0][a];
  }
  N[0] = e0[1] * e1[2] - e0[2] * e1[1];
  N[1] = e0[2] * e1[0] - e0[0] * e1[2];
  N[2] = e0[0] * e1[1] - e0[1] * e1[0];
}

// Visit every bin the triangle overlaps (exact test at bin granularity).
template <typename F>
__device__ __forceinline__ void for_each_bin(const VoxParams& P, const TriRec& R, F&& f) {
  int64_t q[3][3], N[3];
  load_q(R, q, N);
  const int bs = (int)P.bin_shift;
  const int64_t d = (int64_t)kFixedOne << bs;
  const int bx0 = R.lo[0] >> bs, bx1 = R.hi[0] >> bs;
  const int by0 = R.lo[1] >> bs, by1 = R.hi[1] >> bs;
  const int bz0 = R.lo[2] >> bs, bz1 = R.hi[2] >> bs;
  const uint32_t nb = P.nbins_axis;
  for (int bz = bz0; bz <= bz1; bz++)
    for (int by = by0; by <= by1; by++)
      for (int bx = bx0; bx <= bx1; bx++) {
        const int64_t b[3] = {(int64_t)bx * d, (int64_t)by * d, (int64_t)bz * d};
        if (sat_box(q, N, b, d)) f((uint32_t)bx + nb * ((uint32_t)by + nb * (uint32_t)bz));
      }
}

// ---------------------------------------------------------------------------
// setup + per-bin counts
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_setup_count(VoxParams P) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= P.ntri) return;
  const MeshDev& M = P.meshes[find_mesh(P.meshes, P.nmesh, t)];
  const uint32_t lt = t - M.tri_base;
  uint32_t vid[3];
  if (M.idx_type == VGI_INDEX_U32) {
    const uint32_t* ix = (const uint32_t*)M.idx + (size_t)lt * 3;
    vid[0] = ix[0]; vid[1] = ix[1]; vid[2] = ix[2];
  } else if (M.idx_type == VGI_INDEX_U16) {
    const uint16_t* ix = (const uint16_t*)M.idx + (size_t)lt * 3;
    vid[0] = ix[0]; vid[1] = ix[1]; vid[2] = ix[2];
  } else {
    vid[0] = lt * 3; vid[1] = lt * 3 + 1; vid[2] = lt * 3 + 2;
  }
  TriRec R;
  uint32_t status = 0;
  int32_t qi[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
  for (int i = 0; i < 3; i++) {
    if (vid[i] >= M.nverts) { status = 1; continue; }
    const float* p = (const float*)((const char*)M.pos + (size_t)vid[i] * M.stride);
    const float x = p[0], y = p[1], z = p[2];
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
    rgb = (uint32_t)c[0] | ((uint32_t)c[1] << 8) | ((uint32_t)c[2] << 16);
  } else {
    rgb = M.base_rgb & 0xFFFFFFu;
  }
  bool empty = true;
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
    empty = false;
#pragma unroll
    for (int a = 0; a < 3; a++) {
      int32_t mn = min(qi[0][a], min(qi[1][a], qi[2][a]));
      int32_t mx = max(qi[0][a], max(qi[1][a], qi[2][a]));
      int32_t lo = max((mn - 1) >> 8, 0);               // floor((mn-1)/256)
      int32_t hi = min(mx >> 8, (int32_t)P.n - 1);      // floor(mx/256)
      R.lo[a] = (int16_t)lo;
      R.hi[a] = (int16_t)hi;
      if (lo > hi) empty = true;
    }
  } else {
#pragma unroll
    for (int a = 0; a < 3; a++) { R.lo[a] = 1; R.hi[a] = 0; }
  }
  if (empty && status == 0) {
#pragma unroll
    for (int a = 0; a < 3; a++) { R.lo[a] = 1; R.hi[a] = 0; }
  }
  R.rgb = rgb | (status << 24);
  P.tris[t] = R;
  if (status == 1) { atomicAdd(&P.stats->rejected, 1ull); return; }
  if (status == 2) { atomicAdd(&P.stats->degenerate, 1ull); return; }
  if (empty) return;
  for_each_bin(P, R, [&](uint32_t bin) { atomicAdd(&P.bin_count[bin], 1u); });
}

hipError_t launch_setup_count(const VoxParams& P, hipStream_t s) {
  if (P.ntri == 0) return hipSuccess;
  hipLaunchKernelGGL(k_setup_count, dim3((P.ntri + 255) / 256), dim3(256), 0, s, P);
  return hipGetLastError();
}

hipError_t scan_bins(void* tmp, size_t* tmp_bytes, const uint32_t* in, uint32_t* out, uint32_t count,
                     hipStream_t s) {
  return hipcub::DeviceScan::ExclusiveSum(tmp, *tmp_bytes, in, out, (int)count, s);
}

// ---------------------------------------------------------------------------
// emit (triangle, bin) pairs
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_emit(VoxParams P) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= P.ntri) return;
  const TriRec R = P.tris[t];
  if ((R.rgb >> 24) != 0) return;
  if (R.lo[0] > R.hi[0] || R.lo[1] > R.hi[1] || R.lo[2] > R.hi[2]) return;
  for_each_bin(P, R, [&](uint32_t bin) {
    uint32_t pos = atomicAdd(&P.bin_cursor[bin], 1u);
    if (pos < P.pairs_cap) P.pairs[pos] = t;
  });
}

hipError_t launch_emit(const VoxParams& P, hipStream_t s) {
  if (P.ntri == 0) return hipSuccess;
  hipLaunchKernelGGL(k_emit, dim3((P.ntri + 255) / 256), dim3(256), 0, s, P);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// work list: one wave per bin
// ---------------------------------------------------------------------------

__device__ __forceinline__ size_t acc_index(uint32_t x, uint32_t y, uint32_t z, uint32_t nbr) {
  const uint32_t brick = (x >> 3) + nbr * ((y >> 3) + nbr * (z >> 3));
  const uint32_t local = (x & 7u) + 8u * ((y & 7u) + 8u * (z & 7u));
  return ((size_t)brick * 512u + local) * kAccWords;
}

__global__ __launch_bounds__(64) void k_worklist(VoxParams P) {
  const uint32_t bin = blockIdx.x;
  const uint32_t cnt = P.bin_count[bin];
  if (cnt == 0) return;
  const uint32_t nch = (cnt + P.chunk - 1) / P.chunk;
  uint32_t base = 0;
  if (threadIdx.x == 0) base = atomicAdd(&P.stats->work_count, nch);
  base = __shfl(base, 0);
  for (uint32_t i = threadIdx.x; i < nch; i += 64) P.work[base + i] = make_uint2(bin, i);
  if (nch > 1) {
    // zero the bin's accumulator bricks (one 8^3 brick = 16 KiB contiguous)
    const uint32_t nb = P.nbins_axis, B = 1u << P.bin_shift, nbr = P.n >> 3;
    const uint32_t bx = bin % nb, by = (bin / nb) % nb, bz = bin / (nb * nb);
    const uint32_t bpb = B >> 3;  // bricks per bin edge
    for (uint32_t k = 0; k < bpb * bpb * bpb; k++) {
      const uint32_t x = bx * B + 8 * (k % bpb), y = by * B + 8 * ((k / bpb) % bpb),
                     z = bz * B + 8 * (k / (bpb * bpb));
      uint4* dst = (uint4*)(P.acc + acc_index(x, y, z, nbr));
      for (uint32_t i = threadIdx.x; i < 1024; i += 64) dst[i] = make_uint4(0, 0, 0, 0);
    }
  }
}

hipError_t launch_worklist(const VoxParams& P, hipStream_t s) {
  const uint32_t nbins = P.nbins_axis * P.nbins_axis * P.nbins_axis;
  hipLaunchKernelGGL(k_worklist, dim3(nbins), dim3(64), 0, s, P);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// brick voxelize
// ---------------------------------------------------------------------------

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

__device__ __forceinline__ int sel3(int a, int b, int c, int i) { return i == 0 ? a : (i == 1 ? b : c); }

__device__ void make_coef(const TriRec& R, TriCoef& C) {
  int64_t q[3][3], N[3];
  load_q(R, q, N);
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
#pragma unroll
  for (int u = 0; u < 3; u++) {
    const int A = (u + 1) % 3, B = (u + 2) % 3;
    const int64_t sgn = N[u] < 0 ? -1 : 1;
#pragma unroll
    for (int i = 0; i < 3; i++) {
      const int j = (i + 1) % 3;
      const int64_t na = -(q[j][B] - q[i][B]) * sgn;
      const int64_t nb = (q[j][A] - q[i][A]) * sgn;
      const int64_t ea = na * kFixedOne, eb = nb * kFixedOne;
      C.eA[u * 3 + i] = (int32_t)ea;
      C.eB[u * 3 + i] = (int32_t)eb;
      C.eC[u * 3 + i] = -(na * q[i][A] + nb * q[i][B]) + imax64(0, ea) + imax64(0, eb);
    }
  }
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

template <int TS>
__global__ __launch_bounds__(256) void k_voxelize(VoxParams P) {
  constexpr int TV = TS * TS * TS;
  __shared__ unsigned long long s_acc[2 * TV];
  __shared__ TriCoef s_tri[kMaxChunk];
  __shared__ uint32_t s_pref[kMaxChunk + 1];

  const uint32_t tid = threadIdx.x;
  const uint32_t nwork = P.stats->work_count;
  const uint32_t nb = P.nbins_axis, B = 1u << P.bin_shift, nbr = P.n >> 3;
  const uint32_t sub = B / TS;  // sub-tiles per bin edge

  for (uint32_t w = blockIdx.x; w < nwork; w += gridDim.x) {
    const uint2 item = P.work[w];
    const uint32_t bin = item.x, chunk = item.y;
    const uint32_t cnt_bin = P.bin_count[bin];
    const uint32_t nch = (cnt_bin + P.chunk - 1) / P.chunk;
    const uint32_t first = P.bin_offset[bin] + chunk * P.chunk;
    const uint32_t cnt = min(P.chunk, cnt_bin - chunk * P.chunk);
    const uint32_t bx = bin % nb, by = (bin / nb) % nb, bz = bin / (nb * nb);

    if (tid < cnt) {
      const uint32_t t = P.pairs[first + tid];
      const TriRec R = P.tris[t];
      make_coef(R, s_tri[tid]);
    }

    for (uint32_t st = 0; st < sub * sub * sub; st++) {
      const int tx0 = (int)(bx * B + TS * (st % sub));
      const int ty0 = (int)(by * B + TS * ((st / sub) % sub));
      const int tz0 = (int)(bz * B + TS * (st / (sub * sub)));
      for (int i = tid; i < 2 * TV; i += 256) s_acc[i] = 0ull;
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
          const int64_t s = C.N256[0] * vx + C.N256[1] * vy + C.N256[2] * vz;
          if (s + C.pmin > 0 || s + C.pmax < 0) continue;
          if (!edge_ok(C, u1, sel3(vx, vy, vz, (u1 + 1) % 3), sel3(vx, vy, vz, (u1 + 2) % 3))) continue;
          if (!edge_ok(C, u2, sel3(vx, vy, vz, (u2 + 1) % 3), sel3(vx, vy, vz, (u2 + 2) % 3))) continue;
          const int li = (vx - tx0) + TS * ((vy - ty0) + TS * (vz - tz0));
          atomicAdd(&s_acc[2 * li], C.inc0);
          atomicAdd(&s_acc[2 * li + 1], C.inc1);
        }
      }
      __syncthreads();
      // flush the tile
      for (int li = tid; li < TV; li += 256) {
        const unsigned long long w0 = s_acc[2 * li], w1 = s_acc[2 * li + 1];
        const uint32_t lx = li % TS, ly = (li / TS) % TS, lz = li / (TS * TS);
        uint32_t* dst = P.acc + acc_index(tx0 + lx, ty0 + ly, tz0 + lz, nbr);
        const uint32_t r = (uint32_t)(w0 & 0xFFFFu), g = (uint32_t)((w0 >> 16) & 0xFFFFu),
                       b = (uint32_t)((w0 >> 32) & 0xFFFFu), c = (uint32_t)(w0 >> 48);
        const uint32_t nx = (uint32_t)(w1 & 0xFFFFFu), ny = (uint32_t)((w1 >> 20) & 0xFFFFFu),
                       nz = (uint32_t)((w1 >> 40) & 0xFFFFFu);
        if (nch == 1) {
          ((uint4*)dst)[0] = make_uint4(r, g, b, c);
          ((uint4*)dst)[1] = make_uint4(nx, ny, nz, 0u);
        } else if (c != 0) {
          atomicAdd(dst + 0, r);
          atomicAdd(dst + 1, g);
          atomicAdd(dst + 2, b);
          atomicAdd(dst + 3, c);
          atomicAdd(dst + 4, nx);
          atomicAdd(dst + 5, ny);
          atomicAdd(dst + 6, nz);
        }
      }
      __syncthreads();
    }
  }
}

hipError_t launch_voxelize(const VoxParams& P, uint32_t tile, uint32_t grid_blocks, hipStream_t s) {
  if (grid_blocks == 0) return hipSuccess;
  if (tile == 8) hipLaunchKernelGGL(k_voxelize<8>, dim3(grid_blocks), dim3(256), 0, s, P);
  else hipLaunchKernelGGL(k_voxelize<16>, dim3(grid_blocks), dim3(256), 0, s, P);
  return hipGetLastError();
}

}  // namespace vgi
