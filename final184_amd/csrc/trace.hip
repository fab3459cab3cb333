// trace.hip — per-pixel diffuse + specular cone march for gfx950.
//
// Replaces the reference's lighting_indirect pass (Foreground/Renderer/MegaPipeline.cpp:
// 252-268): getIndirect's 60-step nearest-voxel ray march, called for 4 cosine-sampled
// rays plus one bounce each (Shader/Lighting/indirect.frag:104-184,200-222). Here each
// pixel marches num_diffuse fixed cones around its normal and one specular cone along
// the view reflection through the prefiltered mip chain: sample diameter
// d(t) = max(h, 2 t tan(aperture/2)), level log2(d/h), quadrilinear gather (trilinear in
// two adjacent levels), front-to-back compositing C += (1-a) c, a += (1-a) alpha,
// step t += step_factor * d, stop at max_distance, alpha >= cutoff or the grid boundary.
//
// Schedule: a workgroup is one screen tile (8x8, 16x16 or 32x8 pixels; each wavefront an
// 8x8 block), tiles are remapped so that each XCD gets a contiguous band of the image
// (neighbouring tiles share that XCD's L2), and every lane of a wavefront marches the
// same cone index at the same time, so directions stay coherent across the wavefront.
#include "vgi_internal.hpp"

namespace vgi {

__device__ __forceinline__ float h2f(uint32_t b) { return (float)__builtin_bit_cast(_Float16, (uint16_t)b); }
__device__ __forceinline__ float bf2f(uint32_t b) { return __builtin_bit_cast(float, b << 16); }

template <int PREC> struct Fetch;
template <> struct Fetch<VGI_PREC_FP32> {
  __device__ static float4 at(const void* base, size_t i) { return ((const float4*)base)[i]; }
};
template <> struct Fetch<VGI_PREC_FP16> {
  __device__ static float4 at(const void* base, size_t i) {
    const uint2 u = ((const uint2*)base)[i];
    return make_float4(h2f(u.x & 0xFFFFu), h2f(u.x >> 16), h2f(u.y & 0xFFFFu), h2f(u.y >> 16));
  }
};
template <> struct Fetch<VGI_PREC_BF16> {
  __device__ static float4 at(const void* base, size_t i) {
    const uint2 u = ((const uint2*)base)[i];
    return make_float4(bf2f(u.x & 0xFFFFu), bf2f(u.x >> 16), bf2f(u.y & 0xFFFFu), h2f(u.y >> 16));
  }
};

__device__ __forceinline__ float4 lerp4(float4 a, float4 b, float f) {
  return make_float4(a.x + (b.x - a.x) * f, a.y + (b.y - a.y) * f, a.z + (b.z - a.z) * f,
                     a.w + (b.w - a.w) * f);
}

// Trilinear sample of level k at level-0 voxel coordinates u (texel centres at i+0.5).
template <int PREC>
__device__ __forceinline__ float4 trilinear(const TraceParams& P, uint32_t k, float ux, float uy,
                                            float uz) {
  const float sc = __builtin_ldexpf(1.0f, -(int)k);
  const float gx = ux * sc - 0.5f, gy = uy * sc - 0.5f, gz = uz * sc - 0.5f;
  const float fx0 = floorf(gx), fy0 = floorf(gy), fz0 = floorf(gz);
  const float fx = gx - fx0, fy = gy - fy0, fz = gz - fz0;
  const uint32_t p1 = P.nk[k] + 2;
  const size_t p2 = (size_t)p1 * p1;
  // +1: padded border
  const size_t base = (size_t)((int)fx0 + 1) + (size_t)p1 * (size_t)((int)fy0 + 1) +
                      p2 * (size_t)((int)fz0 + 1);
  const void* L = P.level[k];
  const float4 c000 = Fetch<PREC>::at(L, base);
  const float4 c100 = Fetch<PREC>::at(L, base + 1);
  const float4 c010 = Fetch<PREC>::at(L, base + p1);
  const float4 c110 = Fetch<PREC>::at(L, base + p1 + 1);
  const float4 c001 = Fetch<PREC>::at(L, base + p2);
  const float4 c101 = Fetch<PREC>::at(L, base + p2 + 1);
  const float4 c011 = Fetch<PREC>::at(L, base + p2 + p1);
  const float4 c111 = Fetch<PREC>::at(L, base + p2 + p1 + 1);
  const float4 y00 = lerp4(c000, c100, fx), y10 = lerp4(c010, c110, fx);
  const float4 y01 = lerp4(c001, c101, fx), y11 = lerp4(c011, c111, fx);
  const float4 z0 = lerp4(y00, y10, fy), z1 = lerp4(y01, y11, fy);
  return lerp4(z0, z1, fz);
}

template <int PREC>
__device__ __forceinline__ float4 cone_march(const TraceParams& P, float ox, float oy, float oz,
                                             float dx, float dy, float dz, float tan_half,
                                             uint32_t& count) {
  float cr = 0.f, cg = 0.f, cb = 0.f, a = 0.f;
  float t = P.start_off * P.h;
  const float nf = (float)P.n;
  const float top = (float)(P.levels - 1);
  // the step cap bounds the loop even when t stops growing in float (zero aperture,
  // zero direction, huge max distance); no workload comes near it
  for (int it = 0; it < VGI_MAX_MARCH_STEPS && t <= P.max_dist; it++) {
    const float ux = ((ox + dx * t) - P.origin[0]) * P.inv_h;
    const float uy = ((oy + dy * t) - P.origin[1]) * P.inv_h;
    const float uz = ((oz + dz * t) - P.origin[2]) * P.inv_h;
    if (!(ux >= 0.f && ux <= nf && uy >= 0.f && uy <= nf && uz >= 0.f && uz <= nf)) break;
    const float diam = fmaxf(P.h, 2.0f * t * tan_half);
    const float lvl = fmaxf(0.0f, __log2f(diam * P.inv_h));
    float4 s;
    if (lvl >= top) {
      s = trilinear<PREC>(P, P.levels - 1, ux, uy, uz);
      count += 1;
    } else {
      const float fl = floorf(lvl);
      const uint32_t l0 = (uint32_t)fl;
      const float4 s0 = trilinear<PREC>(P, l0, ux, uy, uz);
      const float4 s1 = trilinear<PREC>(P, l0 + 1, ux, uy, uz);
      s = lerp4(s0, s1, lvl - fl);
      count += 2;
    }
    const float w = 1.0f - a;
    cr = cr + w * s.x;
    cg = cg + w * s.y;
    cb = cb + w * s.z;
    a = a + w * s.w;
    if (a >= P.cutoff) break;
    t = t + P.step * diam;
  }
  return make_float4(cr, cg, cb, a);
}

template <int PREC>
__global__ __launch_bounds__(256) void k_trace(TraceParams P) {
  // XCD-aware tile order: block b runs on XCD b % 8; give each XCD a contiguous band.
  const uint32_t ntiles = P.tiles_x * P.tiles_y;
  const uint32_t b = blockIdx.x;
  const uint32_t xcd = b & 7u, slot = b >> 3;
  const uint32_t full = ntiles >> 3, rem = ntiles & 7u;
  const uint32_t tile = xcd * full + (xcd < rem ? xcd : rem) + slot;
  const uint32_t tx = tile % P.tiles_x, ty = tile / P.tiles_x;
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  const uint32_t wpr = P.tile_w >> 3;  // 8x8 wave blocks per tile row
  const uint32_t px = tx * P.tile_w + (wave % wpr) * 8 + (lane & 7u);
  const uint32_t py = ty * P.tile_h + (wave / wpr) * 8 + (lane >> 3);
  const bool active = px < P.width && py < P.height;
  uint32_t count = 0;
  if (active) {
    const size_t p = (size_t)py * P.width + px;
    const float Px = P.pos[p * 3], Py = P.pos[p * 3 + 1], Pz = P.pos[p * 3 + 2];
    const float Nx = P.nrm[p * 3], Ny = P.nrm[p * 3 + 1], Nz = P.nrm[p * 3 + 2];
    const float len2 = (Nx * Nx + Ny * Ny) + Nz * Nz;
    float4 od = make_float4(0.f, 0.f, 0.f, 0.f), os = make_float4(0.f, 0.f, 0.f, 0.f);
    if (len2 > 0.f) {
      const float inv = 1.0f / __builtin_sqrtf(len2);
      const float nx = Nx * inv, ny = Ny * inv, nz = Nz * inv;
      const float off = P.normal_off * P.h;
      const float ox = Px + nx * off, oy = Py + ny * off, oz = Pz + nz * off;
      // orthonormal basis (Duff et al. 2017)
      const float sign = copysignf(1.0f, nz);
      const float ab = -1.0f / (sign + nz);
      const float bb = nx * ny * ab;
      const float t0 = 1.0f + sign * nx * nx * ab, t1 = sign * bb, t2 = -sign * nx;
      const float b0 = bb, b1 = sign + ny * ny * ab, b2 = -ny;
      float ar = 0.f, ag = 0.f, abl = 0.f, aa = 0.f;
      for (uint32_t c = 0; c < P.num_diffuse; c++) {
        const float lx = P.dirs[c][0], ly = P.dirs[c][1], lz = P.dirs[c][2];
        const float dx = (t0 * lx + b0 * ly) + nx * lz;
        const float dy = (t1 * lx + b1 * ly) + ny * lz;
        const float dz = (t2 * lx + b2 * ly) + nz * lz;
        const float4 r = cone_march<PREC>(P, ox, oy, oz, dx, dy, dz, P.tan_d, count);
        const float wc = P.w[c];
        ar = ar + wc * r.x;
        ag = ag + wc * r.y;
        abl = abl + wc * r.z;
        aa = aa + wc * r.w;
      }
      od = make_float4(ar, ag, abl, 1.0f - aa);
      if (P.out_s && P.specular) {
        const float vx = P.eye[0] - Px, vy = P.eye[1] - Py, vz = P.eye[2] - Pz;
        const float vl2 = (vx * vx + vy * vy) + vz * vz;
        const float vi = vl2 > 0.f ? 1.0f / __builtin_sqrtf(vl2) : 0.f;
        const float ix = -vx * vi, iy = -vy * vi, iz = -vz * vi;
        const float dn = (nx * ix + ny * iy) + nz * iz;
        const float rx = ix - 2.0f * dn * nx, ry = iy - 2.0f * dn * ny, rz = iz - 2.0f * dn * nz;
        float th = P.tan_s;
        if (P.rough) th = fminf(fmaxf(P.rough[p], 0.0087f), 1.0f);
        os = cone_march<PREC>(P, ox, oy, oz, rx, ry, rz, th, count);
      }
    }
    P.out_d[p] = od;
    if (P.out_s) P.out_s[p] = os;
  }
  if (P.counter) {
    unsigned long long c = count;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
    if (lane == 0) atomicAdd(P.counter, c);
  }
}

hipError_t launch_trace(const TraceParams& P, hipStream_t s) {
  const uint32_t ntiles = P.tiles_x * P.tiles_y;
  if (ntiles == 0) return hipSuccess;
  const dim3 block(P.tile_w * P.tile_h);
  switch (P.prec) {
    case VGI_PREC_FP32: hipLaunchKernelGGL(k_trace<VGI_PREC_FP32>, dim3(ntiles), block, 0, s, P); break;
    case VGI_PREC_FP16: hipLaunchKernelGGL(k_trace<VGI_PREC_FP16>, dim3(ntiles), block, 0, s, P); break;
    case VGI_PREC_BF16: hipLaunchKernelGGL(k_trace<VGI_PREC_BF16>, dim3(ntiles), block, 0, s, P); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace vgi
