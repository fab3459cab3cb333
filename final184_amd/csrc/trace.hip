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
// 8x8 block) of one scene of the batch; tiles are dealt to the XCDs in bands of one tile
// row (neighbouring tiles share that XCD's L2, and every XCD works across the whole
// screen), and every lane of a wavefront marches the same cone index at the same time,
// so directions stay coherent across the wavefront.
//
// Gather: a quadrilinear sample is 16 weighted texels, sum_i w_i c_i with the level blend
// folded into the weights, accumulated in FP32 (FP16 texels feed v_fma_mix directly; no
// unpacking). For cones whose schedule t_i does not depend on the pixel (all diffuse
// cones, and the specular cone without a roughness buffer) the host computes the schedule
// once per aperture (distance, level, level weights per sample) and the march reads it
// with scalar loads; the level is then wave-uniform:
// the level's buffer descriptor and its row/slice pitches live in SGPRs and the eight
// texels of a level come from one VGPR byte offset (plus SGPR pitch and immediate +1
// texel), four 16-byte loads per level for FP16 (texel x and x+1 share a load), eight for
// FP32. In the FP32 form a gather whose 2x2x2 block is the same for every active lane
// is read with scalar loads instead. The per-pixel-roughness specular cone uses per-lane
// global loads.
//
// Morton order (for G-buffers that are not spatially coherent, e.g. points sampled on
// triangles): k_morton_keys keys every pixel by (scene << 32 | 30-bit Morton code of its
// position quantized to 1024^3 over the grid), one radix sort orders the batch, and a
// workgroup then takes 256 consecutive sorted pixels of one scene, so the cones of a
// wavefront start close together and share cache lines. Every pixel's result is the
// same in either order.
#include <rocprim/device/device_radix_sort.hpp>

#include "vgi_internal.hpp"

namespace vgi {

typedef _Float16 half2_t __attribute__((ext_vector_type(2)));

// Accumulate w * texel into a (FP32) from the raw 32-bit words of one texel.
template <int PREC> struct Acc;
template <> struct Acc<VGI_PREC_FP32> {
  __device__ static void add(float4& a, float w, uint32_t d0, uint32_t d1, uint32_t d2, uint32_t d3) {
    a.x = fmaf(__builtin_bit_cast(float, d0), w, a.x);
    a.y = fmaf(__builtin_bit_cast(float, d1), w, a.y);
    a.z = fmaf(__builtin_bit_cast(float, d2), w, a.z);
    a.w = fmaf(__builtin_bit_cast(float, d3), w, a.w);
  }
};
template <> struct Acc<VGI_PREC_FP16> {
  __device__ static void add(float4& a, float w, uint32_t d0, uint32_t d1) {
    const half2_t rg = __builtin_bit_cast(half2_t, d0);
    const half2_t ba = __builtin_bit_cast(half2_t, d1);
    a.x = fmaf((float)rg.x, w, a.x);
    a.y = fmaf((float)rg.y, w, a.y);
    a.z = fmaf((float)ba.x, w, a.z);
    a.w = fmaf((float)ba.y, w, a.w);
  }
};
template <> struct Acc<VGI_PREC_BF16> {  // RGB bf16, opacity fp16
  __device__ static void add(float4& a, float w, uint32_t d0, uint32_t d1) {
    const half2_t ba = __builtin_bit_cast(half2_t, d1);
    a.x = fmaf(__builtin_bit_cast(float, d0 << 16), w, a.x);
    a.y = fmaf(__builtin_bit_cast(float, d0 & 0xFFFF0000u), w, a.y);
    a.z = fmaf(__builtin_bit_cast(float, d1 << 16), w, a.z);
    a.w = fmaf((float)ba.y, w, a.w);
  }
};

// FP16 x-pairs: one dword holds a channel of texels x and x + 1, and one packed dot
// product (FP16 products, FP32 sum) interpolates it along x with the weight pair
// (w_x0 * w_row, w_x1 * w_row): four per row of the 2x2x2 footprint, where the other forms
// need eight multiply-adds.
struct AccPair {
  __device__ static void add(float4& a, half2_t w, uint32_t rr, uint32_t gg, uint32_t bb, uint32_t aa) {
    a.x = __builtin_amdgcn_fdot2(__builtin_bit_cast(half2_t, rr), w, a.x, false);
    a.y = __builtin_amdgcn_fdot2(__builtin_bit_cast(half2_t, gg), w, a.y, false);
    a.z = __builtin_amdgcn_fdot2(__builtin_bit_cast(half2_t, bb), w, a.z, false);
    a.w = __builtin_amdgcn_fdot2(__builtin_bit_cast(half2_t, aa), w, a.w, false);
  }
  // The first row of a sample's first gather, as the three-operand dot product with a zero
  // addend: the compiler would select the two-operand form, whose destination is the
  // accumulator, and first move 0 into each accumulator. SGPR: the texel words are
  // wave-uniform scalars.
  template <bool SGPR>
  __device__ static void first(float4& a, half2_t w, uint32_t rr, uint32_t gg, uint32_t bb, uint32_t aa) {
    const uint32_t wv = __builtin_bit_cast(uint32_t, w);
    const uint32_t t[4] = {rr, gg, bb, aa};
    float* o[4] = {&a.x, &a.y, &a.z, &a.w};
#pragma unroll
    for (int c = 0; c < 4; c++) {
      if constexpr (SGPR) __asm__("v_dot2_f32_f16 %0, %1, %2, 0" : "=v"(*o[c]) : "s"(t[c]), "v"(wv));
      else __asm__("v_dot2_f32_f16 %0, %1, %2, 0" : "=v"(*o[c]) : "v"(t[c]), "v"(wv));
    }
  }
};

template <int PREC>
constexpr uint32_t kTexelBytes = (PREC == VGI_PREC_FP32 || PREC == VGI_PREC_FP16X2) ? 16u : 8u;

// Whether a zero-opacity texel of level l0 + 1 proves its whole subtree empty, so that the
// fine gather at level l0 may be skipped. Occupied level-0 voxels have opacity 1 and a
// texel is the rounded mean of its 8 children, all non-negative: the sum is at least its
// largest term and the mean of a nonzero child is >= 8^-k at level k, which round-to-
// nearest maps to a value >= 8^-k as long as 8^-k is representable. FP32 represents it at
// every level; FP16 opacity (the FP16 and BF16 forms) down to 2^-24 = 8^-8, so from level
// 9 on (512^3 grids with 10 levels, 1024^3) a parent can round to zero over a nonzero
// child, and there the fine gather is always taken.
constexpr uint32_t kHalfExactSkipTop = 8;
template <int PREC>
__device__ __forceinline__ bool skip_fine_ok(uint32_t l0) {
  return PREC == VGI_PREC_FP32 || l0 + 1 <= kHalfExactSkipTop;
}

// Trilinear footprint of level k at level-0 voxel coordinates u (texel centres at i+0.5):
// texel (ix, iy, iz) of the level (unpadded) and the fractions toward the next texel on
// each axis.
struct Cell {
  int ix, iy, iz;
  float fx, fy, fz;
};

// floor(x) as an integer in one instruction (the compiler emits floor, then convert); x is
// finite and within the padded grid here (positions outside [0, n]^3 end the march first)
__device__ __forceinline__ int cvt_floor(float x) {
  int r;
  __asm__("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(x));
  return r;
}

// a * b + c with 24-bit signed factors, one full-rate instruction (a 32-bit multiply issues
// at a quarter of the rate); B_SGPR: b is wave-uniform and held in an SGPR
template <bool B_SGPR>
__device__ __forceinline__ uint32_t mad_i24(int a, uint32_t b, uint32_t c) {
  uint32_t d;
  if constexpr (B_SGPR) __asm__("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(d) : "v"(a), "s"(b), "v"(c));
  else __asm__("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c));
  return d;
}

template <bool UNIFORM_LEVEL>
__device__ __forceinline__ Cell cell_of(uint32_t k, float ux, float uy, float uz) {
  // 2^-k: for a wave-uniform level built from its bit pattern (k <= 126) by scalar
  // instructions, per lane by one ldexp
  const float sc = UNIFORM_LEVEL ? __builtin_bit_cast(float, (127u - k) << 23) : __builtin_ldexpf(1.0f, -(int)k);
  const float gx = ux * sc - 0.5f, gy = uy * sc - 0.5f, gz = uz * sc - 0.5f;
  // integer part by a flooring convert, fraction by v_fract (x - floor(x))
  return Cell{cvt_floor(gx), cvt_floor(gy), cvt_floor(gz), __builtin_amdgcn_fractf(gx),
              __builtin_amdgcn_fractf(gy), __builtin_amdgcn_fractf(gz)};
}

// Index of texel (ix + 1, iy + 1, iz + 1) (the +1: padded border) of a padded level of row
// pitch p1, per lane: iz >= -1 and every factor is < 2^11, so 24-bit multiplies are exact;
// row = (iz + 1) p1 + iy + 1, written as iz p1 + (p1 + iy + 1) for one multiply-add.
__device__ __forceinline__ uint32_t padded_texel(const Cell& c, uint32_t p1) {
  const uint32_t row = mad_i24<false>(c.iz, p1, p1 + (uint32_t)(c.iy + 1));
  return __umul24(p1, row) + (uint32_t)(c.ix + 1);  // < 1026^3 < 2^31
}

// Byte offset of the same texel with wave-uniform pitches (TB bytes a texel, rowb a row,
// slab a slice): ix TB + iy rowb + iz slab + (TB + rowb + slab), one shift-add and two
// 24-bit multiply-adds. ix, iy, iz >= -1, and on a level under 4 GiB slab < 2^23 (slab =
// TB p1^2 >= 2^23 needs p1 >= 724 for 16-byte texels and p1 >= 1024 for 8-byte ones, and
// then TB p1^3 >= 4 GiB), so every factor fits 24 signed bits; the terms add modulo 2^32
// to the offset, which is below 2^32.
template <uint32_t TB>
__device__ __forceinline__ uint32_t texel_offset(const Cell& c, uint32_t rowb, uint32_t slab, uint32_t base) {
  static_assert(TB == 8 || TB == 16, "texel size");
  uint32_t o;  // ix TB + base as one shift-add (the compiler splits it around the scalar sum)
  __asm__("v_lshl_add_u32 %0, %1, %2, %3" : "=v"(o) : "v"(c.ix), "i"(TB == 16 ? 4 : 3), "s"(base));
  return mad_i24<true>(c.iz, slab, mad_i24<true>(c.iy, rowb, o));
}

// The 8 weights of a footprint, scaled by the level weight wl: xy weights, z weights.
struct Foot {
  float w00, w10, w01, w11;
  float wz0, wz1;
};
__device__ __forceinline__ Foot foot_weights(const Cell& c, float wl) {
  const float ex = 1.0f - c.fx, ey = 1.0f - c.fy;
  Foot f;
  f.w00 = ex * ey;
  f.w10 = c.fx * ey;
  f.w01 = ex * c.fy;
  f.w11 = c.fx * c.fy;
  f.wz0 = (1.0f - c.fz) * wl;
  f.wz1 = c.fz * wl;
  return f;
}

// FP16 weight pairs of the four rows (y0 z0, y1 z0, y0 z1, y1 z1) of a footprint, in
// packed-half arithmetic: the fractions rounded to FP16, the x pair (w (1 - fx), w fx) =
// (-fx, fx) (w, w) + (w, 0) from the level weight w given as the packed halves (w, 0)
// (one multiply-add that reads that scalar register twice: a sign modifier and operand
// selection), the y and z pairs (1 - f, f), the four row products wy wz, and each row's
// pair as x pair times row product. Eleven instructions for the 16 weights.
struct PairW {
  half2_t r[4];
};
// (a, b) rounded to FP16 (to nearest even) in one instruction; written out, since the
// compiler converts the halves one at a time when they are used apart
__device__ __forceinline__ half2_t cvt_pk_half(float a, float b) {
  half2_t r;
  __asm__("v_cvt_pk_f16_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}
__device__ __forceinline__ void pair_weights(const Cell& c, uint32_t hw, PairW& pw) {
  const half2_t np1 = {(_Float16)-1.0f, (_Float16)1.0f}, one0 = {(_Float16)1.0f, (_Float16)0.0f};
  const half2_t fxy = cvt_pk_half(c.fx, c.fy);
  const half2_t fz = cvt_pk_half(c.fz, c.fz);
  const half2_t w0 = __builtin_bit_cast(half2_t, hw);
  const half2_t xw = (half2_t){-fxy.x, fxy.x} * (half2_t){w0.x, w0.x} + w0;
  const half2_t yw = (half2_t){fxy.y, fxy.y} * np1 + one0;
  const half2_t zw = fz * np1 + one0;
  const half2_t r01 = yw * (half2_t){zw.x, zw.x};  // rows y0 z0, y1 z0
  const half2_t r23 = yw * (half2_t){zw.y, zw.y};  // rows y0 z1, y1 z1
  pw.r[0] = xw * (half2_t){r01.x, r01.x};
  pw.r[1] = xw * (half2_t){r01.y, r01.y};
  pw.r[2] = xw * (half2_t){r23.x, r23.x};
  pw.r[3] = xw * (half2_t){r23.y, r23.y};
}

// (w, 0) as packed halves, per lane (the schedule holds them for fixed cones)
__device__ __forceinline__ uint32_t half_weight(float w) {
  const half2_t w0 = {(_Float16)w, (_Float16)0.0f};
  return __builtin_bit_cast(uint32_t, w0);
}
constexpr uint32_t kHalfOne = 0x00003C00u;  // w = 1: (1, 0)

// ---- wave-uniform level: buffer loads with SGPR descriptor and pitches ----
// When every active lane's footprint is the same 2x2x2 texel block (usual at coarse
// levels, where a texel spans many pixels' cone positions), the block is read once with
// scalar loads and broadcast, so the texture-address path does no work for it. Vector
// loads of 16 bytes a lane (eight a gather in FP32, four in the 8-byte and x-pair forms)
// keep that path near its limit even where VALU issue bounds the kernel: the vote took
// the FP16 x-pair trace from 0.86 to 0.71 ms and plain FP16 from 0.92 to 0.87 ms. The
// 8-byte and x-pair forms vote only for a sample's first (coarser) gather, whose
// footprint is uniform more often; voting for the second as well was slower. The
// weighted sum is the same either way.
typedef uint32_t u32x8_t __attribute__((ext_vector_type(8)));
#define VGI_CONSTANT __attribute__((address_space(4)))

// INIT: acc is zero on entry (the sample's first gather). The level weight is wl (FP32
// forms) or hw (the FP16 x-pair form).
template <int PREC, bool INIT>
__device__ __forceinline__ void gather_uniform(const TraceParams& P, uint32_t k, float ux, float uy,
                                               float uz, float wl, uint32_t hw, float4& acc, uint32_t& abits) {
  constexpr uint32_t TB = kTexelBytes<PREC>;
  const LevelDesc L = P.lv[k];  // one scalar load: the level is wave-uniform
  const uint32_t rowb = L.p1 * TB, slab = L.p1 * rowb;
  const Cell c = cell_of<true>(k, ux, uy, uz);
  const uint32_t off = texel_offset<TB>(c, rowb, slab, TB + rowb + slab);  // levels here are < 2^32 bytes
  const int soff[4] = {0, (int)rowb, (int)slab, (int)(rowb + slab)};
  if constexpr (PREC == VGI_PREC_FP16X2) {
    PairW pw;
    pair_weights(c, hw, pw);
    // When every active lane's footprint is the same 2x2x2 block (usual at coarse levels),
    // its four rows are read once with scalar loads and broadcast, sparing the
    // texture-address path four 16-byte vector loads. Only for a sample's first gather:
    // the second one accumulates in place into the first's sums, which the two-path
    // form would copy back after it (four moves, beside the vote's two instructions).
    const uint32_t off0 = __builtin_amdgcn_readfirstlane(off);
    if (INIT && __all(off == off0)) {
      const char* base = (const char*)L.ptr + off0;
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const vgi_u32x4 e = *(const VGI_CONSTANT vgi_u32x4*)(base + soff[q]);
        if (q == 0 && INIT) AccPair::first<true>(acc, pw.r[q], e[0], e[1], e[2], e[3]);
        else AccPair::add(acc, pw.r[q], e[0], e[1], e[2], e[3]);
        abits |= e[3];
      }
      return;
    }
    const __amdgpu_buffer_rsrc_t r =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(L.ptr), 0, (int)L.bytes32, 0x00020000);
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const auto e = __builtin_amdgcn_raw_buffer_load_b128(r, off, soff[q], 0);
      const uint32_t rr = e[0], gg = e[1], bb = e[2], aa = e[3];
      if (q == 0 && INIT) AccPair::first<false>(acc, pw.r[q], rr, gg, bb, aa);
      else AccPair::add(acc, pw.r[q], rr, gg, bb, aa);
      abits |= aa;
    }
  } else {
    const Foot f = foot_weights(c, wl);
    const float wxy[4][2] = {{f.w00, f.w10}, {f.w01, f.w11}, {f.w00, f.w10}, {f.w01, f.w11}};
    const uint32_t off0 = __builtin_amdgcn_readfirstlane(off);
    if constexpr (PREC == VGI_PREC_FP32) {
      if (__all(off == off0)) {
        const char* base = (const char*)L.ptr + off0;
  #pragma unroll
        for (int q = 0; q < 4; q++) {
          const float wz = q < 2 ? f.wz0 : f.wz1;
          const u32x8_t v = *(const VGI_CONSTANT u32x8_t*)(base + soff[q]);
          const uint32_t a0 = v[0], a1 = v[1], a2 = v[2], a3 = v[3];
          const uint32_t b0 = v[4], b1 = v[5], b2 = v[6], b3 = v[7];
          Acc<VGI_PREC_FP32>::add(acc, wxy[q][0] * wz, a0, a1, a2, a3);
          Acc<VGI_PREC_FP32>::add(acc, wxy[q][1] * wz, b0, b1, b2, b3);
          abits |= a3 | b3;
        }
        return;
      }
    } else if (INIT && __all(off == off0)) {
      const char* base = (const char*)L.ptr + off0;
  #pragma unroll
      for (int q = 0; q < 4; q++) {
        const float wz = q < 2 ? f.wz0 : f.wz1;
        const vgi_u32x4 e = *(const VGI_CONSTANT vgi_u32x4*)(base + soff[q]);
        const uint32_t a0 = e[0], a1 = e[1], a2 = e[2], a3 = e[3];
        Acc<PREC>::add(acc, wxy[q][0] * wz, a0, a1);
        Acc<PREC>::add(acc, wxy[q][1] * wz, a2, a3);
        abits |= (a1 | a3) & 0xFFFF0000u;  // opacity is the high half
      }
      return;
    }
    const __amdgpu_buffer_rsrc_t r =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(L.ptr), 0, (int)L.bytes32, 0x00020000);  // unsigned field
  #pragma unroll
    for (int q = 0; q < 4; q++) {
      const float wz = q < 2 ? f.wz0 : f.wz1;
      if constexpr (PREC == VGI_PREC_FP32) {
        const auto a = __builtin_amdgcn_raw_buffer_load_b128(r, off, soff[q], 0);
        const auto b = __builtin_amdgcn_raw_buffer_load_b128(r, off + 16, soff[q], 0);
        const uint32_t a0 = a[0], a1 = a[1], a2 = a[2], a3 = a[3];
        const uint32_t b0 = b[0], b1 = b[1], b2 = b[2], b3 = b[3];
        Acc<VGI_PREC_FP32>::add(acc, wxy[q][0] * wz, a0, a1, a2, a3);
        Acc<VGI_PREC_FP32>::add(acc, wxy[q][1] * wz, b0, b1, b2, b3);
        abits |= a3 | b3;
      } else {
        const auto a = __builtin_amdgcn_raw_buffer_load_b128(r, off, soff[q], 0);
        const uint32_t a0 = a[0], a1 = a[1], a2 = a[2], a3 = a[3];
        Acc<PREC>::add(acc, wxy[q][0] * wz, a0, a1);
        Acc<PREC>::add(acc, wxy[q][1] * wz, a2, a3);
        abits |= (a1 | a3) & 0xFFFF0000u;  // opacity is the high half
      }
    }
  }
}

// ---- per-lane level: 64-bit global loads ----
template <int PREC, bool INIT>
__device__ __forceinline__ void gather_lane(const TraceParams& P, uint32_t k, float ux, float uy,
                                            float uz, float wl, uint32_t hw, float4& acc, uint32_t& abits) {
  constexpr uint32_t TB = kTexelBytes<PREC>;
  const uint32_t p1 = P.nk[k] + 2;
  const size_t rowb = (size_t)p1 * TB, slab = (size_t)p1 * p1 * TB;
  const Cell c = cell_of<false>(k, ux, uy, uz);
  const char* base = (const char*)P.level[k] + (size_t)padded_texel(c, p1) * TB;
  const size_t offs[4] = {0, rowb, slab, rowb + slab};
  if constexpr (PREC == VGI_PREC_FP16X2) {
    PairW pw;
    pair_weights(c, hw, pw);
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const uint4 e = gld4(base + offs[q]);
      if (q == 0 && INIT) AccPair::first<false>(acc, pw.r[q], e.x, e.y, e.z, e.w);
      else AccPair::add(acc, pw.r[q], e.x, e.y, e.z, e.w);
      abits |= e.w;
    }
  } else {
    const Foot f = foot_weights(c, wl);
    const float wxy[4][2] = {{f.w00, f.w10}, {f.w01, f.w11}, {f.w00, f.w10}, {f.w01, f.w11}};
  #pragma unroll
    for (int q = 0; q < 4; q++) {
      const float wz = q < 2 ? f.wz0 : f.wz1;
      if constexpr (PREC == VGI_PREC_FP32) {
        const uint4 a = gld4(base + offs[q]);
        const uint4 b = gld4(base + offs[q] + 16);
        Acc<VGI_PREC_FP32>::add(acc, wxy[q][0] * wz, a.x, a.y, a.z, a.w);
        Acc<VGI_PREC_FP32>::add(acc, wxy[q][1] * wz, b.x, b.y, b.z, b.w);
        abits |= a.w | b.w;
      } else {
        const uint4 a = gld4(base + offs[q]);  // texels x and x+1 (8-byte aligned)
        Acc<PREC>::add(acc, wxy[q][0] * wz, a.x, a.y);
        Acc<PREC>::add(acc, wxy[q][1] * wz, a.z, a.w);
        abits |= (a.y | a.w) & 0xFFFF0000u;
      }
    }
  }
}

// Compositing of one sample (front to back). The sample's opacity is a convex combination
// of values in [0, 1]; its rounded weights can sum to 1 + ulp, and the clamp keeps the
// accumulated a within [0, 1].
__device__ __forceinline__ void composite(float4 s, float& cr, float& cg, float& cb, float& a) {
  s.w = __builtin_amdgcn_fmed3f(s.w, 0.0f, 1.0f);  // min(s.w, 1) for s.w >= 0, one instruction
  const float w = 1.0f - a;
  cr = fmaf(w, s.x, cr);
  cg = fmaf(w, s.y, cg);
  cb = fmaf(w, s.z, cb);
  a = fmaf(w, s.w, a);
}

// inside [0, n]^3, as one unsigned maximum3 and one compare of the bit patterns: for
// x >= +0 the patterns order as the values do, a negative x or a NaN has a pattern above
// n's (sign bit, or an all-ones exponent), so it ends the march. -0.0, which the float
// test would accept, never occurs: the cone origin is canonicalized (+0.0 added), and a
// sample fma(ud, t, uo) with uo != -0.0 is never -0.0.
__device__ __forceinline__ bool inside_grid(float ux, float uy, float uz, float nf) {
  const uint32_t m = __builtin_elementwise_max(
      __builtin_elementwise_max(__builtin_bit_cast(uint32_t, ux), __builtin_bit_cast(uint32_t, uy)),
      __builtin_bit_cast(uint32_t, uz));
  return m <= __builtin_bit_cast(uint32_t, nf);
}

// cone origin in level-0 voxel units, with -0.0 made +0.0 (inside_grid relies on it)
__device__ __forceinline__ float cone_origin(float o, float origin, float inv_h) {
  return (o - origin) * inv_h + 0.0f;
}

// One fixed-aperture cone: its samples (distance, levels, level weights) come from the
// schedule the host computed once for the aperture, read with scalar loads, so the whole
// wavefront takes each sample at the same distance and level with no per-lane schedule
// arithmetic. Sample: quadrilinear gather (levels l0 + 1 and l0, or the top level alone),
// then front-to-back compositing; the march ends at the schedule's end (max distance),
// the grid boundary or opacity >= cutoff.
//
// The coarse footprint (level l0 + 1) holds the parents of every fine (level l0) footprint
// texel. When the coarse footprint's opacity is zero and zero opacity implies an all-zero
// subtree, the fine gather is exactly zero and is skipped (skip_fine_ok).
template <int PREC>
__device__ __forceinline__ float4 cone_march_fixed(const TraceParams& P, const SchedEntry* __restrict__ sched,
                                                   uint32_t nsched, float ox, float oy, float oz, float dx,
                                                   float dy, float dz, uint32_t& count) {
  float cr = 0.f, cg = 0.f, cb = 0.f, a = 0.f;
  const float nf = (float)P.n;
  // cone origin and direction in level-0 voxel units: u(t) = uo + ud t
  const float uox = cone_origin(ox, P.origin[0], P.inv_h), udx = dx * P.inv_h;
  const float uoy = cone_origin(oy, P.origin[1], P.inv_h), udy = dy * P.inv_h;
  const float uoz = cone_origin(oz, P.origin[2], P.inv_h), udz = dz * P.inv_h;
  const float cutoff = P.cutoff;
  // The sample index is the same for every lane still marching (the entry is read with
  // scalar loads); a lane whose march ends leaves the loop, and the loop ends when none is
  // left.
  const uint32_t n = __builtin_amdgcn_readfirstlane(nsched);
  for (uint32_t it = 0; it < n; it++) {
    const SchedEntry e = sld_rec(sched + it);  // scalar load: the index is wave-uniform
    const float ux = fmaf(udx, e.t, uox);
    const float uy = fmaf(udy, e.t, uoy);
    const float uz = fmaf(udz, e.t, uoz);
    if (!inside_grid(ux, uy, uz, nf)) break;
    float4 s = make_float4(0.f, 0.f, 0.f, 0.f);
    uint32_t abits = 0;
    if (e.code & kSchedTop) {
      gather_uniform<PREC, true>(P, e.code & ~kSchedTop, ux, uy, uz, 1.0f, kHalfOne, s, abits);
      count += 1;
    } else {
      gather_uniform<PREC, true>(P, e.code + 1, ux, uy, uz, e.fr, e.hfr, s, abits);
      if (abits || !skip_fine_ok<PREC>(e.code))
        gather_uniform<PREC, false>(P, e.code, ux, uy, uz, e.gr, e.hgr, s, abits);
      count += 2;
    }
    composite(s, cr, cg, cb, a);
    if (a >= cutoff) break;
  }
  return make_float4(cr, cg, cb, a);
}

// One cone with a per-pixel aperture (the roughness-driven specular cone), or any cone when
// a level is too large for 32-bit buffer offsets (BUF false: every gather then uses 64-bit
// addresses): sample diameter d(t) = max(h, 2 t tan), level log2(d / h), step
// t += step_factor * d, computed per lane.
template <int PREC, bool BUF>
__device__ __forceinline__ float4 cone_march_lane(const TraceParams& P, float ox, float oy, float oz, float dx,
                                                  float dy, float dz, float tan_half, uint32_t& count) {
  float cr = 0.f, cg = 0.f, cb = 0.f, a = 0.f;
  float t = P.start_off * P.h;
  const float nf = (float)P.n;
  const float top = (float)(P.levels - 1);
  const float uox = cone_origin(ox, P.origin[0], P.inv_h), udx = dx * P.inv_h;
  const float uoy = cone_origin(oy, P.origin[1], P.inv_h), udy = dy * P.inv_h;
  const float uoz = cone_origin(oz, P.origin[2], P.inv_h), udz = dz * P.inv_h;
  // loop invariants, read once (otherwise reloaded from the descriptor every sample)
  const float h = P.h, inv_h = P.inv_h, step = P.step, cutoff = P.cutoff, max_dist = P.max_dist;
  // the step cap bounds the loop even when t stops growing in float (zero aperture,
  // zero direction, huge max distance); no workload comes near it
  for (int it = 0; it < VGI_MAX_MARCH_STEPS && t <= max_dist; it++) {
    const float ux = fmaf(udx, t, uox);
    const float uy = fmaf(udy, t, uoy);
    const float uz = fmaf(udz, t, uoz);
    if (!inside_grid(ux, uy, uz, nf)) break;
    const float diam = fmaxf(h, 2.0f * t * tan_half);
    const float lvl = fmaxf(0.0f, __log2f(diam * inv_h));
    float4 s = make_float4(0.f, 0.f, 0.f, 0.f);
    uint32_t abits = 0;
    if (lvl >= top) {
      if constexpr (BUF) gather_uniform<PREC, true>(P, P.levels - 1, ux, uy, uz, 1.0f, kHalfOne, s, abits);
      else gather_lane<PREC, true>(P, P.levels - 1, ux, uy, uz, 1.0f, kHalfOne, s, abits);
      count += 1;
    } else {
      const float fl = floorf(lvl);
      const float fr = lvl - fl, gr = 1.0f - fr;
      const uint32_t l0 = (uint32_t)fl;
      const uint32_t hfr = half_weight(fr), hgr = half_weight(gr);  // used by the FP16 x-pair form
      // per-lane schedules usually still agree on the level across the wavefront: then the
      // level's descriptor and pitches are scalars
      const uint32_t lu = __builtin_amdgcn_readfirstlane(l0);
      if (BUF && __all(l0 == lu)) {
        gather_uniform<PREC, true>(P, lu + 1, ux, uy, uz, fr, hfr, s, abits);
        if (abits || !skip_fine_ok<PREC>(lu)) gather_uniform<PREC, false>(P, lu, ux, uy, uz, gr, hgr, s, abits);
      } else {
        gather_lane<PREC, true>(P, l0 + 1, ux, uy, uz, fr, hfr, s, abits);
        if (abits || !skip_fine_ok<PREC>(l0)) gather_lane<PREC, false>(P, l0, ux, uy, uz, gr, hgr, s, abits);
      }
      count += 2;
    }
    composite(s, cr, cg, cb, a);
    if (a >= cutoff) break;
    t = t + step * diam;
  }
  return make_float4(cr, cg, cb, a);
}

template <int PREC, bool BUF, bool SORTED>
__global__ __launch_bounds__(256) void k_trace(TraceBatch Bt, const TraceParams* __restrict__ scenes,
                                               const uint2* __restrict__ blocks,
                                               const unsigned long long* __restrict__ skeys,
                                               const uint32_t* __restrict__ spix) {
  // XCD-aware order: block b runs on XCD b % 8. The batch's work list (tiles of one scene
  // are consecutive, row-major; sorted runs are consecutive in Morton order) is dealt to
  // the XCDs in bands of Bt.band consecutive items, XCD x taking bands x, x + 8, ...:
  // neighbouring items share that XCD's L2, and every XCD draws from the whole list, so
  // they finish together (one contiguous eighth each left the room's frame with XCD
  // workloads 20% apart). The items after the last whole round of bands are split into
  // eight contiguous parts.
  const uint32_t b = blockIdx.x;
  const uint32_t xcd = b & 7u;
  const uint32_t per = Bt.band * 8u, rounds = Bt.band ? Bt.nblocks / per : 0u;
  uint32_t idx;
  if (b < rounds * per) {
    const uint32_t slot = b >> 3;
    idx = ((slot / Bt.band) * 8u + xcd) * Bt.band + slot % Bt.band;
  } else {
    const uint32_t r = b - rounds * per, n = Bt.nblocks - rounds * per;  // r % 8 == xcd
    const uint32_t full = n >> 3, rem = n & 7u;
    idx = rounds * per + xcd * full + (xcd < rem ? xcd : rem) + (r >> 3);
  }
  const uint32_t lane = threadIdx.x & 63u;
  uint32_t scene;
  size_t p;
  bool active;
  if constexpr (SORTED) {
    const size_t e = (size_t)idx * kSortBlock + threadIdx.x;
    scene = (uint32_t)(skeys[(size_t)idx * kSortBlock] >> 32);  // a run never spans two scenes
    const uint32_t q = gld(spix + e);
    active = q != kNoPixel;
    p = q;
  } else {
    const uint2 st = Bt.nscenes == 1 ? make_uint2(0u, idx) : blocks[idx];
    scene = st.x;
    const TraceParams& T = scenes[scene];
    const uint32_t tile = st.y;
    const uint32_t tx = tile % T.tiles_x, ty = tile / T.tiles_x;
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t wpr = Bt.tile_w >> 3;  // 8x8 wave blocks per tile row
    const uint32_t px = tx * Bt.tile_w + (wave % wpr) * 8 + (lane & 7u);
    const uint32_t py = ty * Bt.tile_h + (wave / wpr) * 8 + (lane >> 3);
    active = px < T.width && py < T.height;
    p = (size_t)py * T.width + px;
  }
  const TraceParams& P = scenes[scene];
  uint32_t count = 0;
  if (active) {
    const float Px = gld(P.pos + p * 3), Py = gld(P.pos + p * 3 + 1), Pz = gld(P.pos + p * 3 + 2);
    const float Nx = gld(P.nrm + p * 3), Ny = gld(P.nrm + p * 3 + 1), Nz = gld(P.nrm + p * 3 + 2);
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
        const float4 r = BUF ? cone_march_fixed<PREC>(P, P.sched[0], P.sched_len[0], ox, oy, oz, dx, dy, dz, count)
                             : cone_march_lane<PREC, BUF>(P, ox, oy, oz, dx, dy, dz, P.tan_d, count);
        const float wc = P.w[c];
        ar = fmaf(wc, r.x, ar);
        ag = fmaf(wc, r.y, ag);
        abl = fmaf(wc, r.z, abl);
        aa = fmaf(wc, r.w, aa);
      }
      od = make_float4(ar, ag, abl, 1.0f - aa);
      if (P.out_s && P.specular) {
        const float vx = P.eye[0] - Px, vy = P.eye[1] - Py, vz = P.eye[2] - Pz;
        const float vl2 = (vx * vx + vy * vy) + vz * vz;
        const float vi = vl2 > 0.f ? 1.0f / __builtin_sqrtf(vl2) : 0.f;
        const float ix = -vx * vi, iy = -vy * vi, iz = -vz * vi;
        const float dn = (nx * ix + ny * iy) + nz * iz;
        const float rx = ix - 2.0f * dn * nx, ry = iy - 2.0f * dn * ny, rz = iz - 2.0f * dn * nz;
        if (P.rough) {
          const float th = fminf(fmaxf(gld(P.rough + p), 0.0087f), 1.0f);
          os = cone_march_lane<PREC, BUF>(P, ox, oy, oz, rx, ry, rz, th, count);
        } else {
          os = BUF ? cone_march_fixed<PREC>(P, P.sched[1], P.sched_len[1], ox, oy, oz, rx, ry, rz, count)
                   : cone_march_lane<PREC, BUF>(P, ox, oy, oz, rx, ry, rz, P.tan_s, count);
        }
      }
    }
    gst4(P.out_d + p, make_uint4(__builtin_bit_cast(uint32_t, od.x), __builtin_bit_cast(uint32_t, od.y),
                                 __builtin_bit_cast(uint32_t, od.z), __builtin_bit_cast(uint32_t, od.w)));
    if (P.out_s)
      gst4(P.out_s + p, make_uint4(__builtin_bit_cast(uint32_t, os.x), __builtin_bit_cast(uint32_t, os.y),
                                   __builtin_bit_cast(uint32_t, os.z), __builtin_bit_cast(uint32_t, os.w)));
  }
  if (Bt.counter) {  // one global atomic per workgroup (all add to one address)
    __shared__ unsigned long long s_count[4];
    unsigned long long c = count;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
    if (lane == 0) s_count[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {  // workgroups of 64 (8x8 tiles) or 256 threads
      unsigned long long t = 0;
      for (uint32_t i = 0; i < (blockDim.x >> 6); i++) t += s_count[i];
      atomicAdd(Bt.counter, t);
    }
  }
}

template <int PREC, bool SORTED>
static void launch_trace_prec(const TraceBatch& B, bool buffer_path, dim3 grid, dim3 block, hipStream_t s) {
  if (buffer_path)
    hipLaunchKernelGGL((k_trace<PREC, true, SORTED>), grid, block, 0, s, B, B.scenes, B.blocks, B.skeys, B.spix);
  else
    hipLaunchKernelGGL((k_trace<PREC, false, SORTED>), grid, block, 0, s, B, B.scenes, B.blocks, B.skeys, B.spix);
}

hipError_t launch_trace(const TraceBatch& B, bool buffer_path, hipStream_t s) {
  if (B.nblocks == 0) return hipSuccess;
  const bool sorted = B.skeys != nullptr;
  const dim3 block(sorted ? kSortBlock : B.tile_w * B.tile_h), grid(B.nblocks);
  switch (B.prec) {
    case VGI_PREC_FP32:
      if (sorted) launch_trace_prec<VGI_PREC_FP32, true>(B, buffer_path, grid, block, s);
      else launch_trace_prec<VGI_PREC_FP32, false>(B, buffer_path, grid, block, s);
      break;
    case VGI_PREC_FP16:
      if (sorted) launch_trace_prec<VGI_PREC_FP16, true>(B, buffer_path, grid, block, s);
      else launch_trace_prec<VGI_PREC_FP16, false>(B, buffer_path, grid, block, s);
      break;
    case VGI_PREC_BF16:
      if (sorted) launch_trace_prec<VGI_PREC_BF16, true>(B, buffer_path, grid, block, s);
      else launch_trace_prec<VGI_PREC_BF16, false>(B, buffer_path, grid, block, s);
      break;
    case VGI_PREC_FP16X2:
      if (sorted) launch_trace_prec<VGI_PREC_FP16X2, true>(B, buffer_path, grid, block, s);
      else launch_trace_prec<VGI_PREC_FP16X2, false>(B, buffer_path, grid, block, s);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Morton order
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint32_t spread3(uint32_t v) {  // 10 bits -> every third bit
  v &= 0x3FFu;
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

// One thread per entry of the batch's padded pixel list; a workgroup's 256 entries belong
// to one scene (segments are multiples of 256).
__global__ __launch_bounds__(256) void k_morton_keys(const TraceParams* __restrict__ scenes,
                                                     const uint32_t* __restrict__ seg, uint32_t nscenes,
                                                     unsigned long long* __restrict__ keys,
                                                     uint32_t* __restrict__ pix) {
  const uint32_t e0 = blockIdx.x * kSortBlock;
  uint32_t lo = 0, hi = nscenes - 1;  // last scene whose segment starts at or before e0
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (seg[mid] <= e0) lo = mid;
    else hi = mid - 1;
  }
  const uint32_t sc = lo;
  const TraceParams& P = scenes[sc];
  const uint32_t e = e0 + threadIdx.x;
  const uint32_t i = e - seg[sc];
  const uint32_t npx = P.width * P.height;
  unsigned long long key = ((unsigned long long)sc << 32) | 0xFFFFFFFFull;  // padding sorts last
  uint32_t q = kNoPixel;
  if (i < npx) {
    const float s = 1024.0f / (float)P.n * P.inv_h;
    uint32_t c[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
      const float u = (gld(P.pos + (size_t)i * 3 + a) - P.origin[a]) * s;
      c[a] = (uint32_t)fminf(fmaxf(u, 0.0f), 1023.0f);  // NaN -> 0
    }
    key = ((unsigned long long)sc << 32) | (spread3(c[0]) | (spread3(c[1]) << 1) | (spread3(c[2]) << 2));
    q = i;
  }
  gst(keys + e, key);
  gst(pix + e, q);
}

hipError_t launch_morton_keys(const TraceBatch& B, unsigned long long* keys, uint32_t* pix, uint32_t entries,
                              hipStream_t s) {
  if (entries == 0) return hipSuccess;
  hipLaunchKernelGGL(k_morton_keys, dim3(entries / kSortBlock), dim3(kSortBlock), 0, s, B.scenes, B.seg,
                     B.nscenes, keys, pix);
  return hipGetLastError();
}

hipError_t sort_pixels(void* tmp, size_t* tmp_bytes, const unsigned long long* keys_in,
                       unsigned long long* keys_out, const uint32_t* pix_in, uint32_t* pix_out, uint32_t n,
                       int end_bit, hipStream_t s) {
  return rocprim::radix_sort_pairs(tmp, *tmp_bytes, keys_in, keys_out, pix_in, pix_out, (size_t)n, 0u,
                                   (unsigned)end_bit, s);
}

}  // namespace vgi
