// mips.hip — radiance injection + fused 2x2x2 mip chain for gfx950.
//
// No reference code exists for this stage: the reference keeps a single 128^3 RG16UI
// volume and lights voxels at lookup time (Shader/Lighting/indirect.frag:157-169); the
// LOD volumes appear only as a plan (docs/index.md:36).
//
// Each workgroup reads a 16^3 tile of a source level (512 threads, one 2x2x2 group each)
// and produces up to four levels above it, staging each level's tile in LDS:
// 16^3 -> 8^3 -> 4^3 -> 2^3 -> 1. A launch covers every scene of a batch through its
// block table (scene, tile, source level, levels produced). The first launch's source is
// the integer accumulator: each thread resolves its 8 voxels (rounded mean albedo through
// the (i/255)^2.2 table, mean quantized normal, direct light at the voxel centre) into
// level 0 on the way. Further launches start from the last level written until every
// chain is complete (256^3: levels 0-4, then 5-7). Sums run in a fixed order (x fastest, then y,
// then z) and every level is rounded to its storage precision before it is reduced, so
// the result is a pure function of the accumulator.
//
// Level k is stored padded: (n_k + 2)^3 texels with a zero border, x fastest.
#include "vgi_internal.hpp"

#pragma clang fp contract(off)

namespace vgi {

__device__ __forceinline__ uint16_t f32_to_bf16_bits(float f) {
  uint32_t x = __builtin_bit_cast(uint32_t, f);
  x += 0x7fffu + ((x >> 16) & 1u);
  return (uint16_t)(x >> 16);
}
__device__ __forceinline__ float bf16_bits_to_f32(uint16_t b) {
  return __builtin_bit_cast(float, (uint32_t)b << 16);
}

template <int PREC> struct Texel;

template <> struct Texel<VGI_PREC_FP32> {
  __device__ static float4 load(const void* base, size_t i) {
    const uint4 u = gld4((const float4*)base + i);
    return make_float4(__builtin_bit_cast(float, u.x), __builtin_bit_cast(float, u.y),
                       __builtin_bit_cast(float, u.z), __builtin_bit_cast(float, u.w));
  }
  __device__ static void store(void* base, size_t i, float4 v) {
    gst4((float4*)base + i, make_uint4(__builtin_bit_cast(uint32_t, v.x), __builtin_bit_cast(uint32_t, v.y),
                                       __builtin_bit_cast(uint32_t, v.z), __builtin_bit_cast(uint32_t, v.w)));
  }
  __device__ static float4 round(float4 v) { return v; }
};

template <> struct Texel<VGI_PREC_FP16> {
  __device__ static float4 load(const void* base, size_t i) {
    const uint2 u = gld2((const uint2*)base + i);
    return make_float4(f16_bits_to_f32((uint16_t)(u.x & 0xFFFFu)), f16_bits_to_f32((uint16_t)(u.x >> 16)),
                       f16_bits_to_f32((uint16_t)(u.y & 0xFFFFu)), f16_bits_to_f32((uint16_t)(u.y >> 16)));
  }
  __device__ static uint2 pack(float4 v) {
    return make_uint2((uint32_t)f32_to_f16_bits(v.x) | ((uint32_t)f32_to_f16_bits(v.y) << 16),
                      (uint32_t)f32_to_f16_bits(v.z) | ((uint32_t)f32_to_f16_bits(v.w) << 16));
  }
  __device__ static void store(void* base, size_t i, float4 v) { gst2((uint2*)base + i, pack(v)); }
  __device__ static float4 round(float4 v) {
    return make_float4(f16_bits_to_f32(f32_to_f16_bits(v.x)), f16_bits_to_f32(f32_to_f16_bits(v.y)),
                       f16_bits_to_f32(f32_to_f16_bits(v.z)), f16_bits_to_f32(f32_to_f16_bits(v.w)));
  }
};

// RGB in bf16, opacity in fp16.
template <> struct Texel<VGI_PREC_BF16> {
  __device__ static float4 load(const void* base, size_t i) {
    const uint2 u = gld2((const uint2*)base + i);
    return make_float4(bf16_bits_to_f32((uint16_t)(u.x & 0xFFFFu)), bf16_bits_to_f32((uint16_t)(u.x >> 16)),
                       bf16_bits_to_f32((uint16_t)(u.y & 0xFFFFu)), f16_bits_to_f32((uint16_t)(u.y >> 16)));
  }
  __device__ static void store(void* base, size_t i, float4 v) {
    gst2((uint2*)base + i,
         make_uint2((uint32_t)f32_to_bf16_bits(v.x) | ((uint32_t)f32_to_bf16_bits(v.y) << 16),
                    (uint32_t)f32_to_bf16_bits(v.z) | ((uint32_t)f32_to_f16_bits(v.w) << 16)));
  }
  __device__ static float4 round(float4 v) {
    return make_float4(bf16_bits_to_f32(f32_to_bf16_bits(v.x)), bf16_bits_to_f32(f32_to_bf16_bits(v.y)),
                       bf16_bits_to_f32(f32_to_bf16_bits(v.z)), f16_bits_to_f32(f32_to_f16_bits(v.w)));
  }
};

__device__ __forceinline__ size_t padded_index(uint32_t x, uint32_t y, uint32_t z, uint32_t nk) {
  const size_t p1 = (size_t)nk + 2;
  return (size_t)(x + 1) + p1 * ((size_t)(y + 1) + p1 * (size_t)(z + 1));
}

__device__ __forceinline__ float4 add4(float4 a, float4 b) {
  return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
}
__device__ __forceinline__ float4 scale4(float4 a, float s) {
  return make_float4(a.x * s, a.y * s, a.z * s, a.w * s);
}

// Occupancy of level-0 voxel (x, y, z): its bin holds triangles and its mask bit is set
// (records of unoccupied voxels are not written; masks of empty bins are not either).
__device__ __forceinline__ bool occupied(const MipScene& P, uint32_t x, uint32_t y, uint32_t z) {
  const uint32_t bs = P.bin_shift, nb = P.nbins_axis;
  const uint32_t bin = (x >> bs) + nb * ((y >> bs) + nb * (z >> bs));
  if ((uint32_t)gld(P.bin_packed + bin) == 0) return false;
  const uint32_t nbr = P.n >> 3;
  const uint32_t brick = (x >> 3) + nbr * ((y >> 3) + nbr * (z >> 3));
  const uint32_t local = (x & 7u) + 8u * ((y & 7u) + 8u * (z & 7u));
  return (gld(P.mask + (size_t)brick * 16u + (local >> 5)) >> (local & 31u)) & 1u;
}

// Occupancy bits of the 2x2x2 group at even (x, y, z): bit dx + 2 dy + 4 dz. The group lies
// in one bin and one 8^3 brick, and each of its two z-slices in one mask word.
__device__ __forceinline__ uint32_t group_occupancy(const MipScene& P, uint32_t x, uint32_t y, uint32_t z) {
  const uint32_t bs = P.bin_shift, nb = P.nbins_axis;
  const uint32_t bin = (x >> bs) + nb * ((y >> bs) + nb * (z >> bs));
  if ((uint32_t)gld(P.bin_packed + bin) == 0) return 0u;
  const uint32_t nbr = P.n >> 3;
  const uint32_t brick = (x >> 3) + nbr * ((y >> 3) + nbr * (z >> 3));
  const uint32_t l0 = (x & 7u) + 8u * ((y & 7u) + 8u * (z & 7u));  // z-slice 0 of the group
  const uint32_t* m = P.mask + (size_t)brick * 16u;
  const uint32_t sh = l0 & 31u;  // bits sh, sh+1, sh+8, sh+9 of each slice's word
  const uint32_t w0 = gld(m + (l0 >> 5)) >> sh, w1 = gld(m + ((l0 + 64u) >> 5)) >> sh;
  return (w0 & 3u) | ((w0 >> 6) & 12u) | ((w1 & 3u) << 4) | ((w1 >> 2) & 192u);
}

// Radiance of an occupied level-0 voxel (x, y, z) from its accumulator record.
__device__ float4 inject(const MipScene& P, const float* lut, uint32_t x, uint32_t y, uint32_t z) {
  const uint32_t nbr = P.n >> 3;
  const uint32_t brick = (x >> 3) + nbr * ((y >> 3) + nbr * (z >> 3));
  const uint32_t local = (x & 7u) + 8u * ((y & 7u) + 8u * (z & 7u));
  const uint32_t* a = P.acc + ((size_t)brick * 512u + local) * kAccWords;
  const uint4 s0 = gld4(a), s1 = gld4(a + 4);
  const uint32_t c = s0.w;
  const float lr = lut[(s0.x + c / 2) / c];
  const float lg = lut[(s0.y + c / 2) / c];
  const float lb = lut[(s0.z + c / 2) / c];
  const float inv = 1.0f / ((float)c * 1024.0f);
  const int32_t bias = (int32_t)(1024u * c);
  const float nx = (float)((int32_t)s1.x - bias) * inv;
  const float ny = (float)((int32_t)s1.y - bias) * inv;
  const float nz = (float)((int32_t)s1.z - bias) * inv;
  const float px = P.origin[0] + ((float)x + 0.5f) * P.h;
  const float py = P.origin[1] + ((float)y + 0.5f) * P.h;
  const float pz = P.origin[2] + ((float)z + 0.5f) * P.h;
  float er = 0.f, eg = 0.f, eb = 0.f;
  for (uint32_t l = 0; l < P.nlights; l++) {
    const LightDev& L = P.lights[l];
    if (L.type == VGI_LIGHT_DIRECTIONAL) {
      const float cosv = (nx * -L.vec[0] + ny * -L.vec[1]) + nz * -L.vec[2];
      if (cosv > 0.f) {
        er = er + L.color[0] * cosv;
        eg = eg + L.color[1] * cosv;
        eb = eb + L.color[2] * cosv;
      }
    } else {
      const float dx = L.vec[0] - px, dy = L.vec[1] - py, dz = L.vec[2] - pz;
      const float d2 = (dx * dx + dy * dy) + dz * dz;
      if (d2 > 0.f) {
        const float cosv = ((nx * dx + ny * dy) + nz * dz) / __builtin_sqrtf(d2);
        if (cosv > 0.f) {
          const float md = L.min_distance;
          const float att = fmaxf(d2, md * md);
          const float f = cosv / att;
          er = er + L.color[0] * f;
          eg = eg + L.color[1] * f;
          eb = eb + L.color[2] * f;
        }
      }
    }
  }
  return make_float4(lr * er, lg * eg, lb * eb, 1.0f);
}

template <int PREC, bool SRC_ACC>
__global__ __launch_bounds__(512) void k_mips(const MipScene* __restrict__ scenes,
                                               const MipBlock* __restrict__ blocks,
                                               const float* __restrict__ lut) {
  using TX = Texel<PREC>;
  __shared__ float4 s_a[512];
  __shared__ float4 s_b[64];
  const MipBlock blk = blocks[blockIdx.x];
  const MipScene& P = scenes[blk.scene];
  const uint32_t src = blk.src, nout = blk.nout;
  const uint32_t ns = P.nk[src];
  const uint32_t S = ns < 16u ? ns : 16u;  // source tile edge
  const uint32_t tpa = ns / S;             // tiles per axis
  const uint32_t tile = blk.tile;
  const uint32_t tx = tile % tpa, ty = (tile / tpa) % tpa, tz = tile / (tpa * tpa);
  const uint32_t G = S >> 1;  // groups per tile edge (>= 1 when nout >= 1)
  const uint32_t g = threadIdx.x;
  __shared__ float s_lut[256];  // (i/255)^2.2, read once per workgroup
  if (SRC_ACC) {
    if (g < 256) s_lut[g] = lut[g];
    __syncthreads();
  }

  if (SRC_ACC && nout == 0) {
    // single-level chain: resolve level 0 only, one voxel per thread
    for (uint32_t v = g; v < S * S * S; v += 512) {
      const uint32_t x = tx * S + v % S, y = ty * S + (v / S) % S, z = tz * S + v / (S * S);
      const float4 c = occupied(P, x, y, z) ? TX::round(inject(P, s_lut, x, y, z)) : make_float4(0.f, 0.f, 0.f, 0.f);
      TX::store(P.level[0], padded_index(x, y, z, ns), c);
    }
    return;
  }

  if (g < G * G * G) {
    const uint32_t gx = g % G, gy = (g / G) % G, gz = g / (G * G);
    float4 sum = make_float4(0.f, 0.f, 0.f, 0.f);
    const uint32_t x0 = tx * S + 2 * gx, y0 = ty * S + 2 * gy, z0 = tz * S + 2 * gz;
    const uint32_t occ = SRC_ACC ? group_occupancy(P, x0, y0, z0) : 0u;
#pragma unroll
    for (uint32_t dz = 0; dz < 2; dz++)
#pragma unroll
      for (uint32_t dy = 0; dy < 2; dy++)
#pragma unroll
        for (uint32_t dx = 0; dx < 2; dx++) {
          const uint32_t x = x0 + dx, y = y0 + dy, z = z0 + dz;
          float4 c;
          if (SRC_ACC) {
            c = make_float4(0.f, 0.f, 0.f, 0.f);
            if ((occ >> (dx + 2 * dy + 4 * dz)) & 1u) c = TX::round(inject(P, s_lut, x, y, z));
            TX::store(P.level[0], padded_index(x, y, z, ns), c);
          } else {
            c = TX::load(P.level[src], padded_index(x, y, z, ns));
          }
          sum = add4(sum, c);
        }
    const float4 v = TX::round(scale4(sum, 0.125f));
    const uint32_t nd = P.nk[src + 1];
    TX::store(P.level[src + 1], padded_index(tx * G + gx, ty * G + gy, tz * G + gz, nd), v);
    s_a[g] = v;
  }
  __syncthreads();
  float4* in = s_a;
  float4* out = s_b;
  uint32_t Gk = G;
  for (uint32_t j = 2; j <= nout; j++) {
    const uint32_t Gn = Gk >> 1;
    if (g < Gn * Gn * Gn) {
      const uint32_t gx = g % Gn, gy = (g / Gn) % Gn, gz = g / (Gn * Gn);
      float4 sum = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
      for (uint32_t dz = 0; dz < 2; dz++)
#pragma unroll
        for (uint32_t dy = 0; dy < 2; dy++)
#pragma unroll
          for (uint32_t dx = 0; dx < 2; dx++)
            sum = add4(sum, in[(2 * gx + dx) + Gk * ((2 * gy + dy) + Gk * (2 * gz + dz))]);
      const float4 v = TX::round(scale4(sum, 0.125f));
      const uint32_t nd = P.nk[src + j];
      TX::store(P.level[src + j], padded_index(tx * Gn + gx, ty * Gn + gy, tz * Gn + gz, nd), v);
      out[g] = v;
    }
    __syncthreads();
    float4* t = in;
    in = out;
    out = t;
    Gk = Gn;
  }
}

hipError_t launch_mips(const MipBatch& B, uint32_t prec, bool from_acc, uint32_t blocks, hipStream_t s) {
  if (blocks == 0) return hipSuccess;
  const dim3 g(blocks), b(512);
  switch (prec) {
    case VGI_PREC_FP32:
      if (from_acc) hipLaunchKernelGGL((k_mips<VGI_PREC_FP32, true>), g, b, 0, s, B.scenes, B.blocks, B.lut);
      else hipLaunchKernelGGL((k_mips<VGI_PREC_FP32, false>), g, b, 0, s, B.scenes, B.blocks, B.lut);
      break;
    case VGI_PREC_FP16:
      if (from_acc) hipLaunchKernelGGL((k_mips<VGI_PREC_FP16, true>), g, b, 0, s, B.scenes, B.blocks, B.lut);
      else hipLaunchKernelGGL((k_mips<VGI_PREC_FP16, false>), g, b, 0, s, B.scenes, B.blocks, B.lut);
      break;
    case VGI_PREC_BF16:
      if (from_acc) hipLaunchKernelGGL((k_mips<VGI_PREC_BF16, true>), g, b, 0, s, B.scenes, B.blocks, B.lut);
      else hipLaunchKernelGGL((k_mips<VGI_PREC_BF16, false>), g, b, 0, s, B.scenes, B.blocks, B.lut);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace vgi
