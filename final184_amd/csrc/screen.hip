// screen.hip — the screen-space passes around the cone march: G-buffer reconstruction
// from the depth and packed-normal targets, the temporal blend with last frame's
// result, and the one-axis cross-bilateral blur.
//
// All three are HBM-bound streaming passes (a few tens of bytes per pixel against a
// handful of arithmetic operations). Decode and temporal run one thread per pixel in
// 64 x 4 blocks, so a wavefront covers 64 consecutive pixels of one row and every load
// and store is a contiguous run. The blur stages its tile plus the 15-pixel reach on
// both sides in LDS (colour rgb and depth packed in one float4 per texel), then reads
// its 13 taps from LDS; workgroups walk the tiles along the blur axis within each XCD
// so that the reach, re-read by the neighbouring tile, is an L2 hit.
//
// Arithmetic is FP32 in the order the definitions write it, with fused multiply-adds
// only where written (the blur's weighted sums and its depth term), IEEE division and
// square root, and a polynomial exp2, so the results are reproducible bit for bit on
// any IEEE machine. The blur is VALU-bound (about 300 operations per pixel, mostly the
// 12 exp2 evaluations), the other two passes HBM-bound.
#include "vgi_internal.hpp"

#pragma clang fp contract(off)

namespace vgi {

namespace {

constexpr uint32_t kRowBlockX = 64;  // decode / temporal block: 64 x 4 pixels
constexpr uint32_t kRowBlockY = 4;

template <bool HALF>
__device__ __forceinline__ float4 load_texel(const void* img, size_t i) {
  if constexpr (HALF) {
    const uint2 u = gld2(static_cast<const uint16_t*>(img) + 4 * i);
    return make_float4(f16_bits_to_f32((uint16_t)(u.x & 0xFFFFu)), f16_bits_to_f32((uint16_t)(u.x >> 16)),
                       f16_bits_to_f32((uint16_t)(u.y & 0xFFFFu)), f16_bits_to_f32((uint16_t)(u.y >> 16)));
  } else {
    const uint4 u = gld4(static_cast<const float*>(img) + 4 * i);
    return make_float4(__uint_as_float(u.x), __uint_as_float(u.y), __uint_as_float(u.z), __uint_as_float(u.w));
  }
}

template <bool HALF>
__device__ __forceinline__ void store_texel(void* img, size_t i, float4 v) {
  if constexpr (HALF) {
    gst2(static_cast<uint16_t*>(img) + 4 * i,
         make_uint2((uint32_t)f32_to_f16_bits(v.x) | ((uint32_t)f32_to_f16_bits(v.y) << 16),
                    (uint32_t)f32_to_f16_bits(v.z) | ((uint32_t)f32_to_f16_bits(v.w) << 16)));
  } else {
    gst4(static_cast<float*>(img) + 4 * i,
         make_uint4(__float_as_uint(v.x), __float_as_uint(v.y), __float_as_uint(v.z), __float_as_uint(v.w)));
  }
}

// r = M * (a, b, c, d) for a column-major 4x4 matrix, sums left to right.
__device__ __forceinline__ float mat_row(const float* m, int r, float a, float b, float c, float d) {
  return ((m[r] * a + m[4 + r] * b) + m[8 + r] * c) + m[12 + r] * d;
}
// r = mat3(M) * (a, b, c): the upper-left 3x3 block.
__device__ __forceinline__ float mat3_row(const float* m, int r, float a, float b, float c) {
  return (m[r] * a + m[4 + r] * b) + m[8 + r] * c;
}

// ---------------------------------------------------------------------------
// G-buffer reconstruction (indirect.frag:44-58, 189-196)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_gbuffer_decode(vgi_gbuffer_ref in, float* __restrict__ pos,
                                                        float* __restrict__ nrm, float* __restrict__ view_depth) {
  const uint32_t x = blockIdx.x * kRowBlockX + (threadIdx.x & (kRowBlockX - 1));
  const uint32_t y = blockIdx.y * kRowBlockY + threadIdx.x / kRowBlockX;
  if (x >= in.width || y >= in.height) return;
  const size_t i = (size_t)y * in.width + x;
  const float* P = in.inv_proj;
  const float* M = in.inv_model_view;

  // texelFetch at ivec2(uv * size) of the pixel-centre uv is the pixel itself
  const float u = ((float)x + 0.5f) / (float)in.width;
  const float v = ((float)y + 0.5f) / (float)in.height;
  const float p0 = u * 2.0f - 1.0f, p1 = v * 2.0f - 1.0f, p2 = gld(in.depth + i);
  const float cw = mat_row(P, 3, p0, p1, p2, 1.0f);
  const float cx = mat_row(P, 0, p0, p1, p2, 1.0f) / cw;
  const float cy = mat_row(P, 1, p0, p1, p2, 1.0f) / cw;
  const float cz = mat_row(P, 2, p0, p1, p2, 1.0f) / cw;
  gst(pos + 3 * i + 0, mat_row(M, 0, cx, cy, cz, 1.0f));
  gst(pos + 3 * i + 1, mat_row(M, 1, cx, cy, cz, 1.0f));
  gst(pos + 3 * i + 2, mat_row(M, 2, cx, cy, cz, 1.0f));
  if (view_depth) gst(view_depth + i, -cz);

  // RGBA16 UNORM normal: value / 65535, then fma(t, 2, -1) and normalize
  const uint2 nb = gld2(in.normals + 4 * i);
  float n0 = __builtin_fmaf((float)(nb.x & 0xFFFFu) / 65535.0f, 2.0f, -1.0f);
  float n1 = __builtin_fmaf((float)(nb.x >> 16) / 65535.0f, 2.0f, -1.0f);
  float n2 = __builtin_fmaf((float)(nb.y & 0xFFFFu) / 65535.0f, 2.0f, -1.0f);
  const float len2 = (n0 * n0 + n1 * n1) + n2 * n2;
  if (len2 > 0.0f) {
    const float len = __builtin_sqrtf(len2);
    n0 = n0 / len;
    n1 = n1 / len;
    n2 = n2 / len;
  } else {
    n0 = n1 = n2 = 0.0f;
  }
  gst(nrm + 3 * i + 0, mat3_row(M, 0, n0, n1, n2));
  gst(nrm + 3 * i + 1, mat3_row(M, 1, n0, n1, n2));
  gst(nrm + 3 * i + 2, mat3_row(M, 2, n0, n1, n2));
}

// ---------------------------------------------------------------------------
// Temporal blend (indirect.frag:224-242)
// ---------------------------------------------------------------------------
// Bilinear sample of a texture with repeat addressing at uv in [0, 1]^2, weights in
// FP32: x = u * W - 0.5 lies in [-0.5, W - 0.5], so floor(x) is -1 .. W-1 and the two
// texel columns wrap with one conditional step each (likewise rows).
template <bool HALF>
__device__ __forceinline__ float4 bilinear_repeat(const void* img, uint32_t W, uint32_t H, float u, float v) {
  const float x = u * (float)W - 0.5f, y = v * (float)H - 0.5f;
  const float fx0 = floorf(x), fy0 = floorf(y);
  const float fx = x - fx0, fy = y - fy0;
  int x0 = (int)fx0, y0 = (int)fy0, x1 = x0 + 1, y1 = y0 + 1;
  if (x0 < 0) x0 += (int)W;
  if (x1 >= (int)W) x1 -= (int)W;
  if (y0 < 0) y0 += (int)H;
  if (y1 >= (int)H) y1 -= (int)H;
  const float4 c00 = load_texel<HALF>(img, (size_t)y0 * W + x0);
  const float4 c10 = load_texel<HALF>(img, (size_t)y0 * W + x1);
  const float4 c01 = load_texel<HALF>(img, (size_t)y1 * W + x0);
  const float4 c11 = load_texel<HALF>(img, (size_t)y1 * W + x1);
  const float gx = 1.0f - fx, gy = 1.0f - fy;
  float4 r;
  r.x = (c00.x * gx + c10.x * fx) * gy + (c01.x * gx + c11.x * fx) * fy;
  r.y = (c00.y * gx + c10.y * fx) * gy + (c01.y * gx + c11.y * fx) * fy;
  r.z = (c00.z * gx + c10.z * fx) * gy + (c01.z * gx + c11.z * fx) * fy;
  r.w = (c00.w * gx + c10.w * fx) * gy + (c01.w * gx + c11.w * fx) * fy;
  return r;
}

__device__ __forceinline__ float clamp16(float v) { return fminf(fmaxf(v, 0.0f), 16.0f); }

template <bool HIST_HALF, bool OUT_HALF>
__global__ __launch_bounds__(256) void k_temporal(vgi_temporal_params p) {
  const uint32_t x = blockIdx.x * kRowBlockX + (threadIdx.x & (kRowBlockX - 1));
  const uint32_t y = blockIdx.y * kRowBlockY + threadIdx.x / kRowBlockX;
  if (x >= p.width || y >= p.height) return;
  const size_t i = (size_t)y * p.width + x;
  const float4 ind = load_texel<false>(p.indirect, i);
  const float w0 = gld(p.position + 3 * i), w1 = gld(p.position + 3 * i + 1), w2 = gld(p.position + 3 * i + 2);
  const float csz = -gld(p.view_depth + i);  // view-space z of the pixel

  const float* V = p.prev_model_view;
  const float* Q = p.prev_projection;
  const float c0 = mat_row(V, 0, w0, w1, w2, 1.0f), c1 = mat_row(V, 1, w0, w1, w2, 1.0f);
  const float c2 = mat_row(V, 2, w0, w1, w2, 1.0f), c3 = mat_row(V, 3, w0, w1, w2, 1.0f);
  const float qw = mat_row(Q, 3, c0, c1, c2, c3);
  const float ru = (mat_row(Q, 0, c0, c1, c2, c3) / qw) * 0.5f + 0.5f;
  const float rv = (mat_row(Q, 1, c0, c1, c2, c3) / qw) * 0.5f + 0.5f;

  float r = ind.x, g = ind.y, b = ind.z;
  if (ru >= 0.0f && ru <= 1.0f && rv >= 0.0f && rv <= 1.0f) {  // false for NaN
    const float4 h = bilinear_repeat<HIST_HALF>(p.history, p.width, p.height, ru, rv);
    float t = 1.0f - fabsf(h.w + csz);
    t = fminf(fmaxf(t, 0.0f), 1.0f);
    const float blend = 0.95f * ((t * t) * (3.0f - 2.0f * t));
    const float keep = 1.0f - blend;
    r = clamp16(r * keep + h.x * blend);
    g = clamp16(g * keep + h.y * blend);
    b = clamp16(b * keep + h.z * blend);
  }
  store_texel<OUT_HALF>(p.out, i, make_float4(r, g, b, -csz));
}

// ---------------------------------------------------------------------------
// Cross-bilateral blur (bilateralBlur.inc:15-72)
// ---------------------------------------------------------------------------
// 2^x for x <= 0: 2^k * p(f), k = rint(x), f = x - k (exact, |f| <= 1/2), p the
// degree-7 Taylor polynomial of 2^f (coefficients ln2^j / j! rounded to FP32; the
// truncation error is below 6e-9 relative) evaluated by Horner's rule with FMAs;
// 0 for x <= -125 and NaN, so every non-zero result is a normal number.
__device__ __forceinline__ float exp2_poly(float x) {
  // Branch-free: the polynomial runs on max(x, -126) (NaN -> -126) and a select returns
  // 0 below -125, so the taps of a pixel keep their loads in flight together.
  const float xc = fmaxf(x, -126.0f);
  const float k = rintf(xc);
  const float f = xc - k;
  float q = 1.525273365e-05f;
  q = __builtin_fmaf(q, f, 1.540352969e-04f);
  q = __builtin_fmaf(q, f, 1.333355787e-03f);
  q = __builtin_fmaf(q, f, 9.618128650e-03f);
  q = __builtin_fmaf(q, f, 5.550410971e-02f);
  q = __builtin_fmaf(q, f, 2.402265072e-01f);
  q = __builtin_fmaf(q, f, 6.931471825e-01f);
  q = __builtin_fmaf(q, f, 1.0f);
  const float r = ldexpf(q, (int)k);
  return x > -125.0f ? r : 0.0f;
}

constexpr int kBlurReach = 15;  // farthest tap from the centre, pixels
constexpr int kBlurTaps = 12;
// Tap offsets (pixels) and their distance weights -r^2 / 32 (r = 1..4 for the taps at
// 2r, r = 5, 7 for the taps at 2r + 1), in the order the sums take them.
__constant__ int kTapOffset[kBlurTaps] = {2, -2, 4, -4, 6, -6, 8, -8, 11, -11, 15, -15};

template <int AXIS>
struct BlurTile;
template <>
struct BlurTile<0> {  // along x: 256 x 2 pixel tiles, 286 x 2 staged
  static constexpr int TX = 256, TY = 2, LX = TX + 2 * kBlurReach, LY = TY, STEP = 1;
};
template <>
struct BlurTile<1> {  // along y: 32 x 32 pixel tiles, 32 x 62 staged (31 KB: 5 workgroups per CU)
  static constexpr int TX = 32, TY = 32, LX = TX, LY = TY + 2 * kBlurReach, STEP = LX;
};

template <int R>
__device__ __forceinline__ float blur_dist() {
  return (-(float)(R * R)) * 0.03125f;
}

// exp2(-r^2/32 - dz^2) with dz = (d0 - d) * 512; the square is fused into the subtraction.
__device__ __forceinline__ float blur_weight(float dist, float d, float d0) {
  const float dz = (d0 - d) * 512.0f;
  return exp2_poly(__builtin_fmaf(-dz, dz, dist));
}

template <int AXIS, bool IN_HALF, bool OUT_HALF>
__global__ __launch_bounds__(256) void k_blur(vgi_blur_params p, uint32_t ntx, uint32_t nty) {
  using T = BlurTile<AXIS>;
  __shared__ float4 s_tex[T::LX * T::LY];

  // XCD-aware tile order: the hardware deals workgroups to the 8 XCDs round robin, so
  // workgroup b runs on XCD b % 8; each XCD takes a contiguous run of the tile sequence,
  // which walks along the blur axis first.
  const uint32_t nblk = gridDim.x, b = blockIdx.x, xcd = b & 7u;
  const uint32_t t = xcd * (nblk >> 3) + min(xcd, nblk & 7u) + (b >> 3);
  const uint32_t tx = AXIS == 0 ? t % ntx : t / nty;
  const uint32_t ty = AXIS == 0 ? t / ntx : t % nty;
  const int W = (int)p.width, H = (int)p.height;
  const int x0 = (int)tx * T::TX - (AXIS == 0 ? kBlurReach : 0);
  const int y0 = (int)ty * T::TY - (AXIS == 1 ? kBlurReach : 0);

  for (int e = threadIdx.x; e < T::LX * T::LY; e += 256) {
    const int lx = e % T::LX, ly = e / T::LX;
    const int gx = min(max(x0 + lx, 0), W - 1), gy = min(max(y0 + ly, 0), H - 1);  // clamp to edge
    const size_t gi = (size_t)gy * (size_t)W + (size_t)gx;
    float4 c = load_texel<IN_HALF>(p.color, gi);
    c.w = gld(p.depth + gi);
    s_tex[e] = c;
  }
  __syncthreads();

  for (int o = threadIdx.x; o < T::TX * T::TY; o += 256) {
    const int ox = o % T::TX, oy = o / T::TX;
    const int gx = (int)tx * T::TX + ox, gy = (int)ty * T::TY + oy;
    if (gx >= W || gy >= H) continue;
    const int lc = AXIS == 0 ? oy * T::LX + ox + kBlurReach : (oy + kBlurReach) * T::LX + ox;
    const float4 c = s_tex[lc];
    const float d0 = c.w;
    float tr = c.x, tg = c.y, tb = c.z, tw = 1.0f;
#define VGI_TAP(J, R)                                                  \
    {                                                                  \
      const float4 s = s_tex[lc + kTapOffset[J] * T::STEP];            \
      const float w = blur_weight(blur_dist<R>(), s.w, d0);            \
      tr = __builtin_fmaf(s.x, w, tr);                                 \
      tg = __builtin_fmaf(s.y, w, tg);                                 \
      tb = __builtin_fmaf(s.z, w, tb);                                 \
      tw = tw + w;                                                     \
    }
    VGI_TAP(0, 1) VGI_TAP(1, 1) VGI_TAP(2, 2) VGI_TAP(3, 2) VGI_TAP(4, 3) VGI_TAP(5, 3)
    VGI_TAP(6, 4) VGI_TAP(7, 4) VGI_TAP(8, 5) VGI_TAP(9, 5) VGI_TAP(10, 7) VGI_TAP(11, 7)
#undef VGI_TAP
    store_texel<OUT_HALF>(p.out, (size_t)gy * (size_t)W + (size_t)gx, make_float4(tr / tw, tg / tw, tb / tw, 0.0f));
  }
}

template <int AXIS>
hipError_t launch_blur_axis(const vgi_blur_params& p, hipStream_t s) {
  using T = BlurTile<AXIS>;
  const uint32_t ntx = (p.width + T::TX - 1) / T::TX, nty = (p.height + T::TY - 1) / T::TY;
  const dim3 grid(ntx * nty), block(256);
  const bool ih = p.color_format == VGI_IMAGE_RGBA16F, oh = p.out_format == VGI_IMAGE_RGBA16F;
  if (ih && oh) hipLaunchKernelGGL((k_blur<AXIS, true, true>), grid, block, 0, s, p, ntx, nty);
  else if (ih) hipLaunchKernelGGL((k_blur<AXIS, true, false>), grid, block, 0, s, p, ntx, nty);
  else if (oh) hipLaunchKernelGGL((k_blur<AXIS, false, true>), grid, block, 0, s, p, ntx, nty);
  else hipLaunchKernelGGL((k_blur<AXIS, false, false>), grid, block, 0, s, p, ntx, nty);
  return hipGetLastError();
}

dim3 row_grid(uint32_t w, uint32_t h) {
  return dim3((w + kRowBlockX - 1) / kRowBlockX, (h + kRowBlockY - 1) / kRowBlockY);
}

}  // namespace

hipError_t launch_gbuffer_decode(const vgi_gbuffer_ref& in, float* pos, float* nrm, float* view_depth,
                                 hipStream_t s) {
  if (!in.width || !in.height) return hipSuccess;
  hipLaunchKernelGGL(k_gbuffer_decode, row_grid(in.width, in.height), dim3(256), 0, s, in, pos, nrm, view_depth);
  return hipGetLastError();
}

hipError_t launch_temporal(const vgi_temporal_params& p, hipStream_t s) {
  if (!p.width || !p.height) return hipSuccess;
  const dim3 grid = row_grid(p.width, p.height);
  const bool hh = p.history_format == VGI_IMAGE_RGBA16F, oh = p.out_format == VGI_IMAGE_RGBA16F;
  if (hh && oh) hipLaunchKernelGGL((k_temporal<true, true>), grid, dim3(256), 0, s, p);
  else if (hh) hipLaunchKernelGGL((k_temporal<true, false>), grid, dim3(256), 0, s, p);
  else if (oh) hipLaunchKernelGGL((k_temporal<false, true>), grid, dim3(256), 0, s, p);
  else hipLaunchKernelGGL((k_temporal<false, false>), grid, dim3(256), 0, s, p);
  return hipGetLastError();
}

hipError_t launch_blur(const vgi_blur_params& p, hipStream_t s) {
  if (!p.width || !p.height) return hipSuccess;
  return p.axis == 0 ? launch_blur_axis<0>(p, s) : launch_blur_axis<1>(p, s);
}

}  // namespace vgi
