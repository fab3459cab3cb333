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
// Level k is stored padded: (n_k + 2)^3 texels with a zero border, x fastest (FP16 x-pairs:
// (n_k + 2)^3 entries of two texels, below). The levels
// are zero-filled when allocated, and the first launch stores a 16^3 tile (its level-0
// texels and its blocks of the four levels above) only while the tile holds occupied
// voxels, plus once, as zeros, when it has lost them; a per-tile word records which, and
// for the 16-byte texel forms a level-0 x-row of a stored tile is stored only while it
// holds voxels or held them at the last build (one bit per row). A surface leaves most
// tiles empty (70% at 256^3 for the room), so most level bytes are not rewritten each
// build, and every texel still equals the full computation.
#include <cmath>

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

// A texel in its storage precision: Packed is the stored form (lossless for a value that
// round() has already rounded), load/store move it to and from a level.
template <int PREC> struct Texel;

template <> struct Texel<VGI_PREC_FP32> {
  using Packed = uint4;
  __device__ static Packed pack(float4 v) {
    return make_uint4(__builtin_bit_cast(uint32_t, v.x), __builtin_bit_cast(uint32_t, v.y),
                      __builtin_bit_cast(uint32_t, v.z), __builtin_bit_cast(uint32_t, v.w));
  }
  __device__ static float4 unpack(Packed u) {
    return make_float4(__builtin_bit_cast(float, u.x), __builtin_bit_cast(float, u.y),
                       __builtin_bit_cast(float, u.z), __builtin_bit_cast(float, u.w));
  }
  __device__ static float4 load(const void* base, size_t i) { return unpack(gld4((const uint4*)base + i)); }
  __device__ static void store_packed(void* base, size_t i, Packed u) { gst4((uint4*)base + i, u); }
  __device__ static void store(void* base, size_t i, float4 v) { store_packed(base, i, pack(v)); }
  __device__ static float4 round(float4 v) { return v; }
};

template <> struct Texel<VGI_PREC_FP16> {
  using Packed = uint2;
  __device__ static Packed pack(float4 v) {
    return make_uint2((uint32_t)f32_to_f16_bits(v.x) | ((uint32_t)f32_to_f16_bits(v.y) << 16),
                      (uint32_t)f32_to_f16_bits(v.z) | ((uint32_t)f32_to_f16_bits(v.w) << 16));
  }
  __device__ static float4 unpack(Packed u) {
    return make_float4(f16_bits_to_f32((uint16_t)(u.x & 0xFFFFu)), f16_bits_to_f32((uint16_t)(u.x >> 16)),
                       f16_bits_to_f32((uint16_t)(u.y & 0xFFFFu)), f16_bits_to_f32((uint16_t)(u.y >> 16)));
  }
  __device__ static float4 load(const void* base, size_t i) { return unpack(gld2((const uint2*)base + i)); }
  __device__ static void store_packed(void* base, size_t i, Packed u) { gst2((uint2*)base + i, u); }
  __device__ static void store(void* base, size_t i, float4 v) { store_packed(base, i, pack(v)); }
  __device__ static float4 round(float4 v) {
    return make_float4(f16_bits_to_f32(f32_to_f16_bits(v.x)), f16_bits_to_f32(f32_to_f16_bits(v.y)),
                       f16_bits_to_f32(f32_to_f16_bits(v.z)), f16_bits_to_f32(f32_to_f16_bits(v.w)));
  }
};

// RGB in bf16, opacity in fp16.
template <> struct Texel<VGI_PREC_BF16> {
  using Packed = uint2;
  __device__ static Packed pack(float4 v) {
    return make_uint2((uint32_t)f32_to_bf16_bits(v.x) | ((uint32_t)f32_to_bf16_bits(v.y) << 16),
                      (uint32_t)f32_to_bf16_bits(v.z) | ((uint32_t)f32_to_f16_bits(v.w) << 16));
  }
  __device__ static float4 unpack(Packed u) {
    return make_float4(bf16_bits_to_f32((uint16_t)(u.x & 0xFFFFu)), bf16_bits_to_f32((uint16_t)(u.x >> 16)),
                       bf16_bits_to_f32((uint16_t)(u.y & 0xFFFFu)), f16_bits_to_f32((uint16_t)(u.y >> 16)));
  }
  __device__ static float4 load(const void* base, size_t i) { return unpack(gld2((const uint2*)base + i)); }
  __device__ static void store_packed(void* base, size_t i, Packed u) { gst2((uint2*)base + i, u); }
  __device__ static void store(void* base, size_t i, float4 v) { store_packed(base, i, pack(v)); }
  __device__ static float4 round(float4 v) {
    return make_float4(bf16_bits_to_f32(f32_to_bf16_bits(v.x)), bf16_bits_to_f32(f32_to_bf16_bits(v.y)),
                       bf16_bits_to_f32(f32_to_bf16_bits(v.z)), f16_bits_to_f32(f32_to_f16_bits(v.w)));
  }
};

// FP16 texels in x-pairs: entry i of a padded level holds texel i in its low halves and
// texel i + 1 in its high halves, channel by channel (dwords R_i R_i+1, G.., B.., A..), so a
// texel is written twice: as the low halves of its own entry and the high halves of the
// entry before it (the border entries' other halves stay zero). Values and rounding are
// VGI_PREC_FP16's; Packed (the LDS form) is the plain FP16 texel.
template <> struct Texel<VGI_PREC_FP16X2> {
  using Packed = uint2;
  __device__ static Packed pack(float4 v) { return Texel<VGI_PREC_FP16>::pack(v); }
  __device__ static float4 unpack(Packed u) { return Texel<VGI_PREC_FP16>::unpack(u); }
  __device__ static float4 load(const void* base, size_t i) {
    const uint4 e = gld4((const uint4*)base + i);
    return unpack(make_uint2((e.x & 0xFFFFu) | (e.y << 16), (e.z & 0xFFFFu) | (e.w << 16)));
  }
  // the halves of texel u: (R, G, B, A) as 16-bit words
  __device__ static void halves(Packed u, uint16_t (&h)[4]) {
    h[0] = (uint16_t)(u.x & 0xFFFFu);
    h[1] = (uint16_t)(u.x >> 16);
    h[2] = (uint16_t)(u.y & 0xFFFFu);
    h[3] = (uint16_t)(u.y >> 16);
  }
  // one whole entry: texel lo in the low halves, texel hi in the high halves
  __device__ static void store_entry(void* base, size_t i, Packed lo, Packed hi) {
    uint16_t a[4], b[4];
    halves(lo, a);
    halves(hi, b);
    gst4((uint4*)base + i, make_uint4((uint32_t)a[0] | ((uint32_t)b[0] << 16), (uint32_t)a[1] | ((uint32_t)b[1] << 16),
                                      (uint32_t)a[2] | ((uint32_t)b[2] << 16), (uint32_t)a[3] | ((uint32_t)b[3] << 16)));
  }
  // one half of an entry: the texel's halves at word offset `hi` (0: low, 1: high) of each dword
  __device__ static void store_half(void* base, size_t i, Packed u, uint32_t hi) {
    uint16_t h[4];
    halves(u, h);
    uint16_t* e = (uint16_t*)((uint4*)base + i);
#pragma unroll
    for (int c = 0; c < 4; c++) gst(e + 2 * c + hi, h[c]);
  }
  __device__ static void store_packed(void* base, size_t i, Packed u) {
    store_half(base, i, u, 0u);      // its own entry
    store_half(base, i - 1, u, 1u);  // the entry before it (i >= 1: padded indices)
  }
  __device__ static void store(void* base, size_t i, float4 v) { store_packed(base, i, pack(v)); }
  __device__ static float4 round(float4 v) { return Texel<VGI_PREC_FP16>::round(v); }
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
  const uint32_t nbr = P.n >> 3;
  const uint32_t brick = (x >> 3) + nbr * ((y >> 3) + nbr * (z >> 3));
  const uint32_t l0 = (x & 7u) + 8u * ((y & 7u) + 8u * (z & 7u));  // z-slice 0 of the group
  const uint32_t* m = P.mask + (size_t)brick * 16u;
  const uint32_t sh = l0 & 31u;  // bits sh, sh+1, sh+8, sh+9 of each slice's word
  // the three loads are issued together: every brick's mask words exist (those of an empty
  // bin hold stale bits, discarded here), so the mask need not wait for the bin's entry
  const uint32_t pairs = (uint32_t)gld(P.bin_packed + bin);
  const uint32_t w0 = gld(m + (l0 >> 5)) >> sh, w1 = gld(m + ((l0 + 64u) >> 5)) >> sh;
  const uint32_t bits = (w0 & 3u) | ((w0 >> 6) & 12u) | ((w1 & 3u) << 4) | ((w1 >> 2) & 192u);
  return pairs != 0 ? bits : 0u;  // a select, not a branch, so the mask loads stay with the first
}

__device__ __forceinline__ int32_t sign_i32(int32_t v) { return (v > 0) - (v < 0); }
__device__ __forceinline__ uint32_t abs_u32(int32_t v) { return v < 0 ? 0u - (uint32_t)v : (uint32_t)v; }

// Visibility of shadow-casting light L from occupied voxel (x, y, z) whose summed normal
// (unbiased integer sums) is (n0, n1, n2): a walk through the level-0 occupancy toward the
// light in 16.16 fixed point. Start: the voxel centre moved one voxel along the dominant
// axis of the normal sum (ties to the lowest axis). Each sample advances exactly one voxel
// along the direction's dominant axis, so a walk leaves the grid within n + 1 samples; a
// point light's walk also ends when it reaches the light's distance along that axis. The
// voxel itself is passed over. Integer steps and IEEE single-precision subtraction,
// division and rounding only, so every build of the walk takes the same decisions.
__device__ bool light_visible(const MipScene& P, const LightDev& L, uint32_t x, uint32_t y, uint32_t z,
                              int32_t n0, int32_t n1, int32_t n2) {
  const uint32_t m0 = abs_u32(n0), m1 = abs_u32(n1), m2 = abs_u32(n2);
  int32_t o0 = 0, o1 = 0, o2 = 0;
  if (m0 >= m1 && m0 >= m2) o0 = sign_i32(n0);
  else if (m1 >= m2) o1 = sign_i32(n1);
  else o2 = sign_i32(n2);
  const int32_t p0 = (int32_t)x + o0, p1 = (int32_t)y + o1, p2 = (int32_t)z + o2;
  int32_t s0, s1, s2;
  uint32_t steps = P.n + 2u;
  if (L.type == VGI_LIGHT_DIRECTIONAL) {
    s0 = L.step[0];
    s1 = L.step[1];
    s2 = L.step[2];
  } else {
    const float l0 = L.pv[0] - ((float)p0 + 0.5f);
    const float l1 = L.pv[1] - ((float)p1 + 0.5f);
    const float l2 = L.pv[2] - ((float)p2 + 0.5f);
    const float a0 = fabsf(l0), a1 = fabsf(l1), a2 = fabsf(l2);
    const float la = (a0 >= a1 && a0 >= a2) ? a0 : (a1 >= a2 ? a1 : a2);
    if (!(la > 0.0f)) return true;  // the light sits at the start
    s0 = (int32_t)__builtin_rintf((l0 / la) * 65536.0f);
    s1 = (int32_t)__builtin_rintf((l1 / la) * 65536.0f);
    s2 = (int32_t)__builtin_rintf((l2 / la) * 65536.0f);
    const float c = __builtin_ceilf(la);
    if (c < (float)steps) steps = (uint32_t)c;
  }
  int32_t q0 = p0 * 65536 + 32768, q1 = p1 * 65536 + 32768, q2 = p2 * 65536 + 32768;
  const uint32_t bs = P.bin_shift, nb = P.nbins_axis, nbr = P.n >> 3;
  // One sample per iteration: issuing the loads of 2, 4 or 8 samples together before
  // testing the first measured no faster (0.27, 0.31, 0.43 ms against 0.27 ms for the
  // room's four shadowed lights), the extra samples' work being wasted when a walk ends.
  for (uint32_t m = 0; m < steps; m++) {
    const int32_t c0 = q0 >> 16, c1 = q1 >> 16, c2 = q2 >> 16;  // floor
    if ((uint32_t)c0 >= P.n || (uint32_t)c1 >= P.n || (uint32_t)c2 >= P.n) return true;
    const uint32_t u0 = (uint32_t)c0, u1 = (uint32_t)c1, u2 = (uint32_t)c2;
    const uint32_t bin = (u0 >> bs) + nb * ((u1 >> bs) + nb * (u2 >> bs));
    const uint32_t brick = (u0 >> 3) + nbr * ((u1 >> 3) + nbr * (u2 >> 3));
    const uint32_t local = (u0 & 7u) + 8u * ((u1 & 7u) + 8u * (u2 & 7u));
    // both loads in flight together (the mask word of an empty bin is stale and unused);
    // byte offsets below 2^32 (bins * 8 B, bricks * 64 B for n <= 1024), a bin entry's pair
    // count is its low word
    const uint32_t pairs = gld_at((const uint32_t*)P.bin_packed, bin * 8u);
    const uint32_t word = gld_at(P.mask, brick * 64u + ((local >> 5) << 2));
    const bool self = u0 == x && u1 == y && u2 == z;
    if (!self && pairs != 0u && ((word >> (local & 31u)) & 1u) != 0u) return false;
    q0 += s0;
    q1 += s1;
    q2 += s2;
  }
  return true;
}

// Radiance of an occupied level-0 voxel (x, y, z) from its accumulator record. SHADOW:
// lights flagged to cast shadows are cut where their visibility walk is blocked.
// (s + c / 2) / c, the rounded mean of a colour channel (s <= 255 c, 1 <= c < 2^24),
// from the correctly rounded reciprocal rc of c: the product's relative error is below
// 2^-21 on a quotient below 256, so its integer part is within one of the exact quotient,
// and one remainder test corrects it. Equal to the integer division for every such input,
// at a fraction of a 32-bit division's instructions, with rc shared by the three channels.
__device__ __forceinline__ uint32_t rounded_mean(uint32_t s, uint32_t c, float rc) {
  const uint32_t x = s + (c >> 1);
  const uint32_t q = (uint32_t)((float)x * rc);
  const int32_t r = (int32_t)(x - __umul24(q, c));
  return r < 0 ? q - 1u : (r >= (int32_t)c ? q + 1u : q);
}

template <bool SHADOW>
__device__ float4 inject(const MipScene& P, const float* lut, uint32_t x, uint32_t y, uint32_t z) {
  const uint32_t nbr = P.n >> 3;
  const uint32_t brick = (x >> 3) + nbr * ((y >> 3) + nbr * (z >> 3));
  const uint32_t local = (x & 7u) + 8u * ((y & 7u) + 8u * (z & 7u));
  const uint32_t* a = P.acc + ((size_t)brick * 512u + local) * kAccWords;
  const uint4 s0 = gld4(a), s1 = gld4(a + 4);
  const uint32_t c = s0.w;
  const float rc = 1.0f / (float)c;
  const float lr = lut[rounded_mean(s0.x, c, rc)];
  const float lg = lut[rounded_mean(s0.y, c, rc)];
  const float lb = lut[rounded_mean(s0.z, c, rc)];
  const float inv = rc * 0x1p-10f;  // 1 / (1024 c): scaling by 2^-10 commutes with the rounding
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
    float f = 0.f;  // the light's factor: Lambert term, with the falloff for a point light
    bool lit;
    if (L.type == VGI_LIGHT_DIRECTIONAL) {
      f = (nx * -L.vec[0] + ny * -L.vec[1]) + nz * -L.vec[2];
      lit = f > 0.f;
    } else {
      const float dx = L.vec[0] - px, dy = L.vec[1] - py, dz = L.vec[2] - pz;
      const float d2 = (dx * dx + dy * dy) + dz * dz;
      lit = false;
      if (d2 > 0.f) {
        const float cosv = ((nx * dx + ny * dy) + nz * dz) / __builtin_sqrtf(d2);
        lit = cosv > 0.f;
        const float md = L.min_distance;
        f = cosv / fmaxf(d2, md * md);
      }
    }
    if constexpr (SHADOW)  // one walk site for both kinds of light
      if (lit && L.shadow) lit = light_visible(P, L, x, y, z, (int32_t)s1.x - bias, (int32_t)s1.y - bias, (int32_t)s1.z - bias);
    if (lit) {
      er = er + L.color[0] * f;
      eg = eg + L.color[1] * f;
      eb = eb + L.color[2] * f;
    }
  }
  return make_float4(lr * er, lg * eg, lb * eb, 1.0f);
}

// Threads per 16^3 tile. The FP32 injecting launch without shadows runs 1024: its 64 KB LDS
// tile holds it to two workgroups per CU either way, and a dense tile's inject (up to 4096
// voxels, a few hundred instructions each with four point lights), which sets the launch's
// length, then takes half the rounds (FP32 inject+mips 68.6 -> 59.6 us on the room). The
// 8-byte forms fit three 512-thread workgroups per CU, which serve the many empty tiles
// better (FP16 x-pair 67.5 us either way). With visibility walks 1024 threads were slower
// (FP32 0.32 -> 0.43 ms, FP16 x-pair 0.28 -> 0.45 ms on the room's four shadowed lights).
// The reducing launch runs one thread per 2x2x2 group of a 16^3 block.
constexpr uint32_t mips_threads(int prec, bool src_acc, bool shadow) {
  return src_acc && !shadow && prec == VGI_PREC_FP32 ? 1024u : 512u;
}

template <int PREC, bool SRC_ACC, bool SHADOW>
__global__ __launch_bounds__(1024) void k_mips(const MipScene* __restrict__ scenes,
                                               const MipBlock* __restrict__ blocks,
                                               const float* __restrict__ lut, uint32_t src1, uint32_t nout1,
                                               uint32_t ordered, uint32_t* __restrict__ clear_counts) {
  using TX = Texel<PREC>;
  constexpr uint32_t kThreads = mips_threads(PREC, SRC_ACC, SHADOW);
  __shared__ float4 s_a[512];
  __shared__ float4 s_b[64];
  const MipBlock blk = blocks ? blocks[blockIdx.x] : MipBlock{0u, blockIdx.x, src1, nout1};
  const MipScene& P = scenes[blk.scene];
  // a later launch of an ordered build: the injecting launch, the list's only reader, has
  // completed, so the counts are zeroed here for the next build's k_tiles
  if (!SRC_ACC && clear_counts && blockIdx.x == 0 && threadIdx.x < 2) gst(clear_counts + threadIdx.x, 0u);
  const uint32_t src = blk.src, nout = blk.nout;
  const uint32_t ns = P.nk[src];
  const uint32_t S = ns < 16u ? ns : 16u;  // source tile edge (a power of two)
  const uint32_t tpa = ns / S;             // tiles per axis
  uint32_t tile = blk.tile;
  if (SRC_ACC && ordered) {  // block b takes the b-th tile of k_tiles' list; unlisted tiles keep their zeros
    const uint32_t ntl = tpa * tpa * tpa, b = blockIdx.x;
    // (clamped: a list never holds more than ntl tiles)
    const uint32_t nd = min(gld(P.order + ntl), ntl), nl = min(gld(P.order + ntl + 1), ntl - nd);
    if (b >= nd + nl) return;  // uniform
    tile = gld(P.order + (b < nd ? b : ntl - 1u - (b - nd)));
  }
  const uint32_t tx = tile % tpa, ty = (tile / tpa) % tpa, tz = tile / (tpa * tpa);
  const uint32_t G = S >> 1;  // groups per tile edge (>= 1 when nout >= 1)
  const uint32_t gsh = G > 1 ? (uint32_t)__builtin_ctz(G) : 0u;
  const uint32_t g = threadIdx.x;
  const uint32_t gx = g & (G - 1u), gy = (g >> gsh) & (G - 1u), gz = g >> (2 * gsh);
  const bool grp = g < G * G * G;

  if constexpr (SRC_ACC) {
    __shared__ float s_lut[256];  // (i/255)^2.2, read once per workgroup
    // loaded now, stored to LDS just before the inject loop's barrier: barriers wait on LDS
    // operations only, so the load's latency overlaps the occupancy pass
    const float lut_v = g < 256 ? lut[g] : 0.0f;
    if (nout == 0) {
      if (g < 256) s_lut[g] = lut_v;
      __syncthreads();
      // single-level chain: resolve level 0 only, one voxel per thread
      for (uint32_t v = g; v < S * S * S; v += kThreads) {
        const uint32_t x = tx * S + v % S, y = ty * S + (v / S) % S, z = tz * S + v / (S * S);
        const float4 c = occupied(P, x, y, z) ? TX::round(inject<SHADOW>(P, s_lut, x, y, z)) : make_float4(0.f, 0.f, 0.f, 0.f);
        TX::store(P.level[0], padded_index(x, y, z, ns), c);
      }
      return;
    }
    // Level 0 of the tile in LDS (x fastest), zero where unoccupied. The occupied voxels
    // are listed first and injected by all threads in turn, so a wavefront's lanes all do
    // useful work however sparse the surface is; the level is then stored row by row and
    // each 2x2x2 group summed from LDS.
    __shared__ typename TX::Packed s_val[4096];
    __shared__ uint32_t s_nocc, s_was;
    __shared__ uint32_t s_rows_now[8], s_rows_prev[8];  // one bit per x-row (ly + S lz) of the tile
    // the tile's state from the previous build, read now so that its latency overlaps the
    // occupancy pass: 1 while its stored texels (level 0 and its blocks of the levels
    // produced here) are not all zero; and per level-0 x-row, 1 while the row's stored
    // texels are not all zero (the rows' words follow the tiles' words)
    // (rows are skipped for the 16-byte texel forms, FP32 and FP16 x-pairs, whose level-0
    // stores cost most; for 8-byte texels the row bookkeeping cost more than it saved)
    constexpr bool ROWS = PREC == VGI_PREC_FP32 || PREC == VGI_PREC_FP16X2;
    const uint32_t ntiles = tpa * tpa * tpa, nrw = (S * S + 31u) >> 5;
    const uint32_t was_v = g == 0 ? gld(P.live + tile) : 0u;
    const uint32_t rows_v = ROWS && g < nrw ? gld(P.live + ntiles + tile * 8u + g) : 0u;
    uint16_t* s_item = (uint16_t*)s_a;  // occupied voxel list (group * 8 + bit), before s_a is used
    static_assert(sizeof(s_a) >= 4096 * sizeof(uint16_t), "item list fits in s_a");
    const uint32_t ssh = gsh + 1u, NV = S * S * S;
    const uint32_t lane = g & 63u;
    for (uint32_t v = g; v < NV; v += kThreads) s_val[v] = TX::pack(make_float4(0.f, 0.f, 0.f, 0.f));
    if (g == 0) s_nocc = 0u;
    if (g < 8) s_rows_now[g] = 0u;
    __syncthreads();
    const uint32_t occ = grp ? group_occupancy(P, tx * S + 2 * gx, ty * S + 2 * gy, tz * S + 2 * gz) : 0u;
    if constexpr (ROWS) {
#pragma unroll
      for (uint32_t q = 0; q < 4; q++)  // x-rows (2 gy + dy, 2 gz + dz) of the group that hold voxels
        if ((occ >> (2 * q)) & 3u) {
          const uint32_t row = (2 * gy + (q & 1u)) + S * (2 * gz + (q >> 1));
          atomicOr(&s_rows_now[row >> 5], 1u << (row & 31u));
        }
    }
    const uint32_t cnt = (uint32_t)__builtin_popcount(occ);
    uint32_t incl = cnt;  // inclusive scan of the counts over the wavefront
#pragma unroll
    for (uint32_t d = 1; d < 64; d <<= 1) {
      const uint32_t o = __shfl_up(incl, d);
      if (lane >= d) incl += o;
    }
    uint32_t wbase = 0;
    if (lane == 63u) wbase = atomicAdd(&s_nocc, incl);
    wbase = __shfl(wbase, 63);
    uint32_t k = wbase + incl - cnt;
    for (uint32_t m = occ; m != 0u; m &= m - 1u) s_item[k++] = (uint16_t)(g * 8u + (uint32_t)__builtin_ctz(m));
    if (g < 256) s_lut[g] = lut_v;
    if (g == 0) s_was = was_v;
    if (g < 8) s_rows_prev[g] = rows_v;
    __syncthreads();
    const uint32_t nocc = s_nocc;
    // A tile without occupied voxels whose stored texels are already zero (never written
    // since the levels were allocated and zero-filled, or cleared by an earlier build) is
    // left as it is; one that held voxels before is stored, as zeros, once.
    const uint32_t was = s_was;
    if (nocc == 0u && was == 0u) return;  // uniform: both values come from LDS
    if (g == 0 && (nocc != 0u) != (was != 0u)) gst(P.live + tile, nocc != 0u ? 1u : 0u);
    if (ROWS && g < nrw && s_rows_now[g] != s_rows_prev[g]) gst(P.live + ntiles + tile * 8u + g, s_rows_now[g]);
    for (uint32_t it = g; it < nocc; it += kThreads) {
      const uint32_t e = s_item[it], gi = e >> 3, bit = e & 7u;
      const uint32_t lx = 2 * (gi & (G - 1u)) + (bit & 1u);
      const uint32_t ly = 2 * ((gi >> gsh) & (G - 1u)) + ((bit >> 1) & 1u);
      const uint32_t lz = 2 * (gi >> (2 * gsh)) + (bit >> 2);
      const float4 c = TX::round(inject<SHADOW>(P, s_lut, tx * S + lx, ty * S + ly, tz * S + lz));
      s_val[lx + (ly << ssh) + (lz << (2 * ssh))] = TX::pack(c);
    }
    __syncthreads();
    if constexpr (PREC == VGI_PREC_FP16X2) {
      // whole entries from the tile's own texels; the entry of the tile's last x gets its
      // low halves here and its high halves from the next tile, the entry before the
      // tile's first x its high halves from here
      for (uint32_t v = g; v < NV; v += kThreads) {
        const uint32_t row = v >> ssh;  // rows that hold voxels now or held them at the last build
        if (!(((s_rows_now[row >> 5] | s_rows_prev[row >> 5]) >> (row & 31u)) & 1u)) continue;
        const uint32_t lx = v & (S - 1u), ly = (v >> ssh) & (S - 1u), lz = v >> (2 * ssh);
        const size_t pi = padded_index(tx * S + lx, ty * S + ly, tz * S + lz, ns);
        if (lx + 1u < S) TX::store_entry(P.level[0], pi, s_val[v], s_val[v + 1]);
        else TX::store_half(P.level[0], pi, s_val[v], 0u);
        if (lx == 0u) TX::store_half(P.level[0], pi - 1, s_val[v], 1u);
      }
    } else {
      for (uint32_t v = g; v < NV; v += kThreads) {
        const uint32_t row = v >> ssh;
        if (ROWS && !(((s_rows_now[row >> 5] | s_rows_prev[row >> 5]) >> (row & 31u)) & 1u)) continue;
        const uint32_t lx = v & (S - 1u), ly = (v >> ssh) & (S - 1u), lz = v >> (2 * ssh);
        TX::store_packed(P.level[0], padded_index(tx * S + lx, ty * S + ly, tz * S + lz, ns), s_val[v]);
      }
    }
    if (grp) {
      float4 sum = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
      for (uint32_t dz = 0; dz < 2; dz++)
#pragma unroll
        for (uint32_t dy = 0; dy < 2; dy++)
#pragma unroll
          for (uint32_t dx = 0; dx < 2; dx++)
            sum = add4(sum, TX::unpack(s_val[(2 * gx + dx) + ((2 * gy + dy) << ssh) + ((2 * gz + dz) << (2 * ssh))]));
      const float4 v = TX::round(scale4(sum, 0.125f));
      TX::store(P.level[src + 1], padded_index(tx * G + gx, ty * G + gy, tz * G + gz, P.nk[src + 1]), v);
      s_a[g] = v;  // the item list in s_a was last read before the barrier above
    }
  } else {
    if (grp) {
      float4 sum = make_float4(0.f, 0.f, 0.f, 0.f);
      const uint32_t x0 = tx * S + 2 * gx, y0 = ty * S + 2 * gy, z0 = tz * S + 2 * gz;
#pragma unroll
      for (uint32_t dz = 0; dz < 2; dz++)
#pragma unroll
        for (uint32_t dy = 0; dy < 2; dy++)
#pragma unroll
          for (uint32_t dx = 0; dx < 2; dx++)
            sum = add4(sum, TX::load(P.level[src], padded_index(x0 + dx, y0 + dy, z0 + dz, ns)));
      const float4 v = TX::round(scale4(sum, 0.125f));
      TX::store(P.level[src + 1], padded_index(tx * G + gx, ty * G + gy, tz * G + gz, P.nk[src + 1]), v);
      s_a[g] = v;
    }
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

template <int PREC>
static void launch_mips_prec(const MipBatch& B, bool from_acc, dim3 g, hipStream_t s) {
  const dim3 b(mips_threads(PREC, from_acc, from_acc && B.shadow));
  if (!from_acc)
    hipLaunchKernelGGL((k_mips<PREC, false, false>), g, b, 0, s, B.scenes, B.blocks, B.lut, B.src, B.nout,
                       (uint32_t)B.ordered, B.clear_counts);
  else if (B.shadow)
    hipLaunchKernelGGL((k_mips<PREC, true, true>), g, b, 0, s, B.scenes, B.blocks, B.lut, B.src, B.nout,
                       (uint32_t)B.ordered, (uint32_t*)nullptr);
  else
    hipLaunchKernelGGL((k_mips<PREC, true, false>), g, b, 0, s, B.scenes, B.blocks, B.lut, B.src, B.nout,
                       (uint32_t)B.ordered, (uint32_t*)nullptr);
}

// The injecting launch's tiles for one scene, densest first: a tile is listed when its bins
// hold pairs (it may hold voxels) or its stored texels are not all zero (it must be
// cleared); tiles whose bins hold at least kDenseTilePairs pairs go to the front of the
// list, the others to the back, with one atomic per wavefront and class on the two counts
// after the list (cleared before the launch). Unlisted tiles hold no voxel and only zeros,
// and are left as they are. A dense tile's inject sets the injecting launch's length (on
// the room, about 3200 occupied voxels against a median of 430), so those start first,
// and empty tiles no longer take a workgroup slot each. A tile's bins' pair count follows
// its occupied voxel count closely (correlation 0.96 on the room); 384 pairs marks about
// the densest third of the room's occupied tiles.
constexpr uint32_t kDenseTilePairs = 384;
__global__ __launch_bounds__(256) void k_tiles(const MipScene* __restrict__ scenes) {
  const MipScene& P = scenes[0];
  const uint32_t ns = P.nk[0], S = ns < 16u ? ns : 16u, tpa = ns / S, ntl = tpa * tpa * tpa;
  const uint32_t t = blockIdx.x * 256u + threadIdx.x;
  uint32_t dense = 0, other = 0;
  if (t < ntl) {
    const uint32_t tx = t % tpa, ty = (t / tpa) % tpa, tz = t / (tpa * tpa);
    const uint32_t bs = P.bin_shift, nb = P.nbins_axis;
    const uint32_t x0 = (tx * S) >> bs, y0 = (ty * S) >> bs, z0 = (tz * S) >> bs;
    const uint32_t bpa = (S >> bs) > 1u ? S >> bs : 1u;  // bins per tile edge
    const uint32_t live = gld(P.live + t);
    unsigned long long pairs = 0;
    if (bpa <= 2u) {  // bins of >= 8 voxels (or one bin for grids under 8): all loads in flight together
      uint32_t v[8];
#pragma unroll
      for (uint32_t i = 0; i < 8u; i++) {
        const uint32_t dx = i & 1u, dy = (i >> 1) & 1u, dz = i >> 2;
        v[i] = dx < bpa && dy < bpa && dz < bpa ? (uint32_t)gld(P.bin_packed + (x0 + dx) + nb * ((y0 + dy) + nb * (z0 + dz)))
                                                : 0u;
      }
#pragma unroll
      for (uint32_t i = 0; i < 8u; i++) pairs += v[i];
    } else {
      for (uint32_t z = z0; z < z0 + bpa; z++)
        for (uint32_t y = y0; y < y0 + bpa; y++)
          for (uint32_t x = x0; x < x0 + bpa; x++) pairs += (uint32_t)gld(P.bin_packed + x + nb * (y + nb * z));
    }
    const bool listed = pairs != 0 || live != 0u;
    dense = listed && pairs >= kDenseTilePairs ? 1u : 0u;
    other = listed && !dense ? 1u : 0u;
  }
  const uint32_t lane = __lane_id();
  const uint64_t dm = __ballot(dense), om = __ballot(other);
  const uint64_t below = (1ull << lane) - 1ull;
  uint32_t dbase = 0, obase = 0;
  if (lane == 0) {
    if (dm) dbase = gatomic_add(P.order + ntl, (uint32_t)__popcll(dm));
    if (om) obase = gatomic_add(P.order + ntl + 1, (uint32_t)__popcll(om));
  }
  dbase = __shfl(dbase, 0);
  obase = __shfl(obase, 0);
  if (dense) gst(P.order + dbase + (uint32_t)__popcll(dm & below), t);
  if (other) gst(P.order + ntl - 1u - (obase + (uint32_t)__popcll(om & below)), t);
}

hipError_t launch_tiles(const MipBatch& B, uint32_t ntiles, hipStream_t s) {
  hipLaunchKernelGGL(k_tiles, dim3((ntiles + 255) / 256), dim3(256), 0, s, B.scenes);
  return hipGetLastError();
}

hipError_t launch_mips(const MipBatch& B, uint32_t prec, bool from_acc, uint32_t blocks, hipStream_t s) {
  if (blocks == 0) return hipSuccess;
  const dim3 g(blocks);
  switch (prec) {
    case VGI_PREC_FP32: launch_mips_prec<VGI_PREC_FP32>(B, from_acc, g, s); break;
    case VGI_PREC_FP16: launch_mips_prec<VGI_PREC_FP16>(B, from_acc, g, s); break;
    case VGI_PREC_BF16: launch_mips_prec<VGI_PREC_BF16>(B, from_acc, g, s); break;
    case VGI_PREC_FP16X2: launch_mips_prec<VGI_PREC_FP16X2>(B, from_acc, g, s); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

LightDev light_dev(const vgi_light& L, const float origin[3], float h) {
  LightDev D{};
  D.type = L.type & ~VGI_LIGHT_SHADOW;
  for (int a = 0; a < 3; a++) {
    D.vec[a] = L.vec[a];
    D.color[a] = L.color[a];
  }
  D.min_distance = L.min_distance;
  D.shadow = (L.type & VGI_LIGHT_SHADOW) ? 1u : 0u;
  if (!D.shadow) return D;
  if (D.type == VGI_LIGHT_DIRECTIONAL) {
    // the step toward the light is the same from every voxel: -vec / |vec_a| in 16.16
    const float l[3] = {-L.vec[0], -L.vec[1], -L.vec[2]};
    const float a0 = std::fabs(l[0]), a1 = std::fabs(l[1]), a2 = std::fabs(l[2]);
    const float la = (a0 >= a1 && a0 >= a2) ? a0 : (a1 >= a2 ? a1 : a2);
    if (!(la > 0.0f)) {
      D.shadow = 0;  // no direction to walk: visible from everywhere
      return D;
    }
    for (int a = 0; a < 3; a++) D.step[a] = (int32_t)std::nearbyint((l[a] / la) * 65536.0f);
  } else {
    for (int a = 0; a < 3; a++) D.pv[a] = (L.vec[a] - origin[a]) / h;
  }
  return D;
}

}  // namespace vgi
