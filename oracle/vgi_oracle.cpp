// vgi_oracle.cpp — CPU oracle for the voxel global-illumination path.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker the HIP path is compared
// against; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load it. The product (final184_amd/libvgi.so) never links or calls it.
//
// It restates, in plain scalar C++ (compiled with -ffp-contract=off so that float
// results do not depend on FMA contraction), the algorithm of each stage:
//
//  (1) orc_voxelize — conservative voxelization into an integer accumulator.
//      The reference voxelizes with a raster pipeline: VoxelGS picks the dominant axis
//      of the view-space face normal (Pipelang/Internal/main.lua:100-115) and VoxelPS
//      imageStores one RG16UI texel per fragment, last writer wins (main.lua:242-275).
//      The order-independent replacement keeps, per voxel, integer sums of the RGB8
//      albedo and a count (the sum/count form of an RGBA8 running average) plus sums of
//      the quantized unit normal (the reference stores a 5:6:5 normal per voxel,
//      main.lua:271). Coverage is the exact closed-box triangle/box overlap test of
//      Akenine-Moller 2001 in the form of Schwarz & Seidel 2010 (plane test + three
//      axis-plane projections, each with three edge functions), evaluated on vertices
//      snapped to 1/256 voxel with 128-bit integer products: brute force over the
//      triangle's voxel bounding box, no pruning. A textured mesh's albedo is sampled per
//      covered voxel and a voxel whose sampled alpha is below 0.05 gets nothing from the
//      triangle (BasicMaterial's texture path and discard, main.lua:188-205; restated in
//      texture_albedo below).
//  (2) orc_inject_mips_* — radiance at each occupied voxel = linear albedo x direct
//      light at the voxel centre (Lambert with the voxel's mean normal; directional
//      lights and point lights with 1/max(d^2, r_min^2) falloff); the reference
//      lights voxels at lookup with lambert(-sun, voxel normal) x pow(colour, 2.2)
//      (Shader/Lighting/indirect.frag:157-169). A light whose type carries the shadow
//      flag (0x100) is cut to zero at voxels from which a walk toward it meets an occupied
//      voxel first: the voxel-space counterpart of the reference's shadow-map test of
//      every lit hit (indirect.frag:160-165), restated in light_visible below. Then the 2x2x2 box mip chain
//      (no reference code; Crassin et al. 2011), children summed x fastest then y then
//      z, times 1/8, each level rounded to its storage precision before the next.
//  (3) orc_trace_* — front-to-back cone march with quadrilinear (trilinear x level)
//      gathers, replacing the reference's per-pixel ray march
//      (Shader/Lighting/indirect.frag:104-184, called 4x at :200-222).
//  (4) orc_ref_* — line-for-line restatements of the reference shader helpers used
//      by tests that document reference behaviour: VoxelGS's dominant-axis choice
//      (main.lua:100-115), VoxelPS's RGB565 colour/normal packing (main.lua:270-271),
//      indirect.frag's unpackColor565/unpackNormal565 (:88-102), make_coord_space
//      (:71-86), hemisphereSample_cos (Shader/math.inc:76-81) and getIndirect's ray
//      march (indirect.frag:104-184) over a nearest-voxel RG16UI volume.
//  (5) orc_gbuffer_decode_* — the reference's G-buffer reconstruction
//      (Shader/Lighting/indirect.frag:44-58 getDepth/getCSpos/getNormal, :189-196): D32
//      depth at the pixel, view position = InvProj * (uv*2-1, depth, 1) divided by w,
//      world position and normal through InvModelView, normal = normalize(unorm*2-1).
//  (6) orc_temporal — the reference's temporal blend (indirect.frag:224-242): reproject
//      the world position with last frame's matrices, bilinear (repeat-addressed)
//      history sample, weight 0.95 * smoothstep(0, 1, 1 - |prevDepth + cspos.z|),
//      clamp [0, 16], out = (colour, -cspos.z).
//  (7) orc_bilateral_blur — bilateralBlur.inc:15-72 with blurX.frag / blurY.frag's
//      directions: taps at +-2i (i = 1..4) and +-(2i+1) (i = 5, 7) pixels (invres is
//      2 / resolution, so every tap lands on a texel centre), clamp-to-edge,
//      w = exp2(-r^2/32 - ((d0 - d) * 512)^2), out = (sum w*c / sum w, 0); the square
//      and the weighted sums are fused multiply-adds (what a shader compiler emits).
//
// Data conventions shared with the product are the spec, restated here:
//  * grid cube [origin, origin+extent]^3, n^3 voxels, voxel (i,j,k) linear index
//    i + n*(j + n*k); level k has n>>k voxels per axis.
//  * snap: s = float(n*256)/extent; v = (p - origin)*s in float; reject unless
//    |v| <= 262144; q = rint(v) (round half to even).
//  * accumulator voxel = 8 x uint32: R, G, B sums, count, sum(nx+1024), sum(ny+1024),
//    sum(nz+1024), 0; nq = rint(N*1024/|N|) in double from the exact integer normal N.
//  * texel = premultiplied RGBA; occupied level-0 voxels have alpha 1.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef __int128 i128;

#define OAPI extern "C"

extern "C" {

struct orc_vox_stats {
  uint64_t triangles_in;
  uint64_t rejected;
  uint64_t degenerate;
  uint64_t voxel_hits;  // (triangle, voxel) overlaps
};

struct orc_light {
  uint32_t type;  // 0 directional, 1 point
  float vec[3];
  float color[3];
  float min_distance;
};

struct orc_cones {
  uint32_t num_diffuse;
  uint32_t specular;
  float diffuse_aperture_deg;
  float diffuse_ring_deg;
  float specular_aperture_deg;
  float step_factor;
  float max_distance;
  float alpha_cutoff;
  float normal_offset;
  float start_offset;
  float eye[3];
};

}  // extern "C" (types)

OAPI int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

OAPI void orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

// ---------------------------------------------------------------------------
// (1) voxelization
// ---------------------------------------------------------------------------

static inline int64_t floor_div256(int64_t a) {
  // floor(a / 256) for signed a
  return (a >= 0) ? (a / 256) : -((-a + 255) / 256);
}

// Exact closed-box overlap of triangle (q[0..2], integer coordinates) with the axis-
// aligned box [b, b + delta]^3. Straight from the separating-axis statement.
static bool tri_box_overlap(const int64_t q[3][3], const int64_t b[3], int64_t delta) {
  int64_t e0[3], e1[3];
  for (int a = 0; a < 3; a++) {
    e0[a] = q[1][a] - q[0][a];
    e1[a] = q[2][a] - q[0][a];
  }
  i128 N[3] = {(i128)e0[1] * e1[2] - (i128)e0[2] * e1[1],
               (i128)e0[2] * e1[0] - (i128)e0[0] * e1[2],
               (i128)e0[0] * e1[1] - (i128)e0[1] * e1[0]};
  // bounding boxes
  for (int a = 0; a < 3; a++) {
    int64_t mn = std::min(q[0][a], std::min(q[1][a], q[2][a]));
    int64_t mx = std::max(q[0][a], std::max(q[1][a], q[2][a]));
    if (mx < b[a] || mn > b[a] + delta) return false;
  }
  // plane: the box spans both sides of (or touches) N.(p - q0) = 0
  i128 s_lo = 0, s_hi = 0;
  for (int a = 0; a < 3; a++) {
    i128 t0 = N[a] * (i128)(b[a] - q[0][a]);
    i128 t1 = N[a] * (i128)(b[a] + delta - q[0][a]);
    s_lo += std::min(t0, t1);
    s_hi += std::max(t0, t1);
  }
  if (s_lo > 0 || s_hi < 0) return false;
  // projections onto the three axis planes; "up" axis u, plane axes (u+1)%3, (u+2)%3
  for (int u = 0; u < 3; u++) {
    int A = (u + 1) % 3, B = (u + 2) % 3;
    i128 sgn = (N[u] < 0) ? -1 : 1;
    for (int i = 0; i < 3; i++) {
      const int64_t* p0 = q[i];
      const int64_t* p1 = q[(i + 1) % 3];
      i128 na = -(i128)(p1[B] - p0[B]) * sgn;  // left normal of the edge in (A,B)
      i128 nb = (i128)(p1[A] - p0[A]) * sgn;
      // maximum of the edge function over the projected box
      i128 ea = std::max(na * (i128)(b[A] - p0[A]), na * (i128)(b[A] + delta - p0[A]));
      i128 eb = std::max(nb * (i128)(b[B] - p0[B]), nb * (i128)(b[B] + delta - p0[B]));
      if (ea + eb < 0) return false;
    }
  }
  return true;
}

OAPI int orc_tri_box_overlap(const int64_t* q9, const int64_t* b3, int64_t delta) {
  int64_t q[3][3];
  for (int i = 0; i < 3; i++)
    for (int a = 0; a < 3; a++) q[i][a] = q9[i * 3 + a];
  return tri_box_overlap(q, b3, delta) ? 1 : 0;
}

// Snap one world-space point; returns false when rejected.
static bool snap_point(const float p[3], const float origin[3], float s, int64_t out[3]) {
  for (int a = 0; a < 3; a++) {
    float v = (p[a] - origin[a]) * s;
    if (!(std::fabs(v) <= 262144.0f)) return false;
    out[a] = (int64_t)std::rint(v);
  }
  return true;
}

// World transform of a vertex with a row-major 3x4 matrix, fixed evaluation order.
static void transform_point(const float m[12], const float x[3], float out[3]) {
  for (int r = 0; r < 3; r++) {
    float acc = m[r * 4 + 0] * x[0];
    acc = acc + m[r * 4 + 1] * x[1];
    acc = acc + m[r * 4 + 2] * x[2];
    acc = acc + m[r * 4 + 3];
    out[r] = acc;
  }
}

// Snapped vertices + quantized normal of one triangle. Returns 0 ok, 1 rejected,
// 2 degenerate.
OAPI int orc_triangle_setup(const float* positions, uint32_t num_vertices, const uint32_t* indices,
                       uint32_t tri, const float* model, const float* origin, float extent,
                       uint32_t n, int64_t* q_out9, int32_t* nq_out3) {
  float s = (float)(n * 256u) / extent;
  uint32_t vid[3];
  for (int i = 0; i < 3; i++) {
    vid[i] = indices ? indices[tri * 3 + i] : tri * 3 + i;
    if (vid[i] >= num_vertices) return 1;
  }
  int64_t q[3][3];
  for (int i = 0; i < 3; i++) {
    float w[3];
    transform_point(model, positions + (size_t)vid[i] * 3, w);
    if (!snap_point(w, origin, s, q[i])) return 1;
  }
  int64_t e0[3], e1[3];
  for (int a = 0; a < 3; a++) {
    e0[a] = q[1][a] - q[0][a];
    e1[a] = q[2][a] - q[0][a];
  }
  int64_t N[3] = {e0[1] * e1[2] - e0[2] * e1[1], e0[2] * e1[0] - e0[0] * e1[2],
                  e0[0] * e1[1] - e0[1] * e1[0]};
  for (int i = 0; i < 3; i++)
    for (int a = 0; a < 3; a++) q_out9[i * 3 + a] = q[i][a];
  if (N[0] == 0 && N[1] == 0 && N[2] == 0) return 2;
  double len = std::sqrt(((double)N[0] * (double)N[0] + (double)N[1] * (double)N[1]) +
                         (double)N[2] * (double)N[2]);
  for (int a = 0; a < 3; a++) nq_out3[a] = (int32_t)std::rint((double)N[a] * 1024.0 / len);
  return 0;
}

// A textured material (the reference's BasicMaterial with UseTextures,
// Pipelang/Internal/main.lua:188-205: BaseColor = texture(uv) * BaseColorFactor, discard
// when BaseColor.a < 0.05). uv: 2 floats per vertex; texels: RGBA8, row-major w x h.
struct orc_texture {
  const float* uv;
  const uint8_t* texels;
  uint32_t w, h;
  float factor[4];
};

// Albedo of one covered voxel from a textured triangle. The voxel centre is projected
// along the triangle's dominant axis u (VoxelGS's rule, as its fragment is rasterized in
// that projection) onto the plane of the other two axes (A = u+1, B = u+2 mod 3), where
// the barycentric numerator of vertex k is sign(N_u) * cross(q_j - q_i, c - q_i) for the
// edge (i, j) opposite k, in the exact snapped units (c = 256 * voxel + 128). Negative
// numerators (centre outside the triangle: conservative coverage) are clamped to 0, so
// the weights are those of the nearest point's region. Then, in single precision:
// uv = ((w0 uv0 + w1 uv1) + w2 uv2) / (w0 + w1 + w2); a non-finite uv is taken as (0, 0);
// wrap addressing (u - floor(u)); bilinear filtering at level 0 between texel centres
// (x = u W - 1/2), texel values c * (1/255), lerp(a, b, t) = a + t (b - a), first along x
// then y; times the factor; discarded when alpha < 0.05; RGB8 = (uint)(clamp(c, 0, 1) *
// 255 + 0.5). Returns false for a discarded fragment.
static bool texture_albedo(const orc_texture& T, const uint32_t vid[3], const int64_t q[3][3],
                           const int64_t N[3], int64_t vx, int64_t vy, int64_t vz, uint32_t rgb[3]) {
  const int64_t ax = std::llabs((long long)N[0]), ay = std::llabs((long long)N[1]), az = std::llabs((long long)N[2]);
  const int u = ax > ay ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
  const int A = (u + 1) % 3, B = (u + 2) % 3;
  const int64_t v[3] = {vx, vy, vz};
  const int64_t cA = v[A] * 256 + 128, cB = v[B] * 256 + 128;
  const int64_t sg = N[u] < 0 ? -1 : 1;
  int64_t e[3];
  for (int k = 0; k < 3; k++) {
    const int i = (k + 1) % 3, j = (k + 2) % 3;  // the edge opposite vertex k
    const int64_t dA = q[j][A] - q[i][A], dB = q[j][B] - q[i][B];
    const int64_t pA = cA - q[i][A], pB = cB - q[i][B];
    e[k] = std::max<int64_t>(0, sg * (dA * pB - dB * pA));
  }
  const float w0 = (float)e[0], w1 = (float)e[1], w2 = (float)e[2], ws = (float)(e[0] + e[1] + e[2]);
  const float* uv0 = T.uv + (size_t)vid[0] * 2;
  const float* uv1 = T.uv + (size_t)vid[1] * 2;
  const float* uv2 = T.uv + (size_t)vid[2] * 2;
  float tu = ((w0 * uv0[0] + w1 * uv1[0]) + w2 * uv2[0]) / ws;
  float tv = ((w0 * uv0[1] + w1 * uv1[1]) + w2 * uv2[1]) / ws;
  if (!std::isfinite(tu) || !std::isfinite(tv)) tu = tv = 0.0f;
  const float fu = tu - std::floor(tu), fv = tv - std::floor(tv);
  const float x = fu * (float)T.w - 0.5f, y = fv * (float)T.h - 0.5f;
  const float x0 = std::floor(x), y0 = std::floor(y);
  const float gx = x - x0, gy = y - y0;
  int64_t ix0 = (int64_t)x0, iy0 = (int64_t)y0, ix1 = ix0 + 1, iy1 = iy0 + 1;
  if (ix0 < 0) ix0 += T.w;
  if (iy0 < 0) iy0 += T.h;
  if (ix1 >= (int64_t)T.w) ix1 -= T.w;
  if (iy1 >= (int64_t)T.h) iy1 -= T.h;
  const uint8_t* t00 = T.texels + ((size_t)iy0 * T.w + (size_t)ix0) * 4;
  const uint8_t* t10 = T.texels + ((size_t)iy0 * T.w + (size_t)ix1) * 4;
  const uint8_t* t01 = T.texels + ((size_t)iy1 * T.w + (size_t)ix0) * 4;
  const uint8_t* t11 = T.texels + ((size_t)iy1 * T.w + (size_t)ix1) * 4;
  const float k255 = 1.0f / 255.0f;
  float col[4];
  for (int ch = 0; ch < 4; ch++) {
    const float a = (float)t00[ch] * k255, b = (float)t10[ch] * k255;
    const float c = (float)t01[ch] * k255, d = (float)t11[ch] * k255;
    const float top = a + gx * (b - a), bot = c + gx * (d - c);
    col[ch] = (top + gy * (bot - top)) * T.factor[ch];
  }
  if (col[3] < 0.05f) return false;
  for (int ch = 0; ch < 3; ch++)
    rgb[ch] = (uint32_t)(std::fmin(std::fmax(col[ch], 0.0f), 1.0f) * 255.0f + 0.5f);
  return true;
}

// Adds one mesh's triangles into acc (n^3 * 8 uint32, not cleared). indices may be NULL
// (non-indexed). albedo: 3 bytes per triangle (or NULL -> base_rgb), unless tex is given.
static void voxelize_into(const float* positions, uint32_t num_vertices, const uint32_t* indices,
                          uint32_t num_triangles, const uint8_t* albedo, const uint8_t* base_rgb,
                          const orc_texture* tex, const float* model, const float* origin, float extent,
                          uint32_t n, uint32_t* acc, orc_vox_stats* stats) {
  uint64_t rejected = 0, degenerate = 0, hits = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : rejected, degenerate, hits)
  for (int64_t t = 0; t < (int64_t)num_triangles; t++) {
    int64_t q9[9];
    int32_t nq[3];
    int rc = orc_triangle_setup(positions, num_vertices, indices, (uint32_t)t, model, origin,
                                extent, n, q9, nq);
    if (rc == 1) { rejected++; continue; }
    if (rc == 2) { degenerate++; continue; }
    int64_t q[3][3];
    for (int i = 0; i < 3; i++)
      for (int a = 0; a < 3; a++) q[i][a] = q9[i * 3 + a];
    int64_t N[3];
    {
      int64_t e0[3], e1[3];
      for (int a = 0; a < 3; a++) {
        e0[a] = q[1][a] - q[0][a];
        e1[a] = q[2][a] - q[0][a];
      }
      N[0] = e0[1] * e1[2] - e0[2] * e1[1];
      N[1] = e0[2] * e1[0] - e0[0] * e1[2];
      N[2] = e0[0] * e1[1] - e0[1] * e1[0];
    }
    uint32_t vid[3];
    for (int i = 0; i < 3; i++) vid[i] = indices ? indices[t * 3 + i] : (uint32_t)t * 3 + i;
    const uint8_t* c = albedo ? albedo + (size_t)t * 3 : base_rgb;
    int64_t lo[3], hi[3];
    bool empty = false;
    for (int a = 0; a < 3; a++) {
      int64_t mn = std::min(q[0][a], std::min(q[1][a], q[2][a]));
      int64_t mx = std::max(q[0][a], std::max(q[1][a], q[2][a]));
      // voxel k spans [256k, 256k+256]; closed boxes touching [mn, mx]
      lo[a] = std::max<int64_t>(floor_div256(mn - 1), 0);
      hi[a] = std::min<int64_t>(floor_div256(mx), (int64_t)n - 1);
      if (lo[a] > hi[a]) empty = true;
    }
    if (empty) continue;
    for (int64_t z = lo[2]; z <= hi[2]; z++)
      for (int64_t y = lo[1]; y <= hi[1]; y++)
        for (int64_t x = lo[0]; x <= hi[0]; x++) {
          int64_t b[3] = {x * 256, y * 256, z * 256};
          if (!tri_box_overlap(q, b, 256)) continue;
          uint32_t rgb[3] = {c[0], c[1], c[2]};
          if (tex && !texture_albedo(*tex, vid, q, N, x, y, z, rgb)) continue;  // discarded
          hits++;
          uint32_t* v = acc + ((size_t)x + (size_t)n * ((size_t)y + (size_t)n * (size_t)z)) * 8;
          __atomic_fetch_add(&v[0], rgb[0], __ATOMIC_RELAXED);
          __atomic_fetch_add(&v[1], rgb[1], __ATOMIC_RELAXED);
          __atomic_fetch_add(&v[2], rgb[2], __ATOMIC_RELAXED);
          __atomic_fetch_add(&v[3], 1u, __ATOMIC_RELAXED);
          __atomic_fetch_add(&v[4], (uint32_t)(nq[0] + 1024), __ATOMIC_RELAXED);
          __atomic_fetch_add(&v[5], (uint32_t)(nq[1] + 1024), __ATOMIC_RELAXED);
          __atomic_fetch_add(&v[6], (uint32_t)(nq[2] + 1024), __ATOMIC_RELAXED);
        }
  }
  if (stats) {
    stats->triangles_in += num_triangles;
    stats->rejected += rejected;
    stats->degenerate += degenerate;
    stats->voxel_hits += hits;
  }
}

// acc: n^3 * 8 uint32, zeroed by this call. indices may be NULL (non-indexed).
// albedo: 3 bytes per triangle (or NULL -> base_rgb). model: row-major 3x4.
OAPI int orc_voxelize(const float* positions, uint32_t num_vertices, const uint32_t* indices,
                 uint32_t num_triangles, const uint8_t* albedo, const uint8_t* base_rgb,
                 const float* model, const float* origin, float extent, uint32_t n,
                 uint32_t* acc, orc_vox_stats* stats) {
  if (n == 0 || n > 1024 || (n & (n - 1))) return -1;
  std::memset(acc, 0, (size_t)n * n * n * 8 * sizeof(uint32_t));
  if (stats) *stats = orc_vox_stats{0, 0, 0, 0};
  voxelize_into(positions, num_vertices, indices, num_triangles, albedo, base_rgb, nullptr, model, origin,
                extent, n, acc, stats);
  return 0;
}

// The same for a textured mesh, added into acc (not cleared, so meshes accumulate).
// uv: 2 floats per vertex; texels: RGBA8 w x h; factor: RGBA.
OAPI int orc_voxelize_textured(const float* positions, uint32_t num_vertices, const uint32_t* indices,
                               uint32_t num_triangles, const float* uv, const uint8_t* texels, uint32_t w,
                               uint32_t h, const float* factor, const float* model, const float* origin,
                               float extent, uint32_t n, uint32_t* acc, orc_vox_stats* stats) {
  if (n == 0 || n > 1024 || (n & (n - 1)) || !uv || !texels || w == 0 || h == 0) return -1;
  orc_texture T{uv, texels, w, h, {factor[0], factor[1], factor[2], factor[3]}};
  const uint8_t white[3] = {255, 255, 255};
  voxelize_into(positions, num_vertices, indices, num_triangles, nullptr, white, &T, model, origin, extent, n,
                acc, stats);
  return 0;
}

// ---------------------------------------------------------------------------
// storage rounding (float -> fp16 / bf16, round to nearest even)
// ---------------------------------------------------------------------------

static inline uint32_t f2u(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
static inline float u2f(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

OAPI uint16_t orc_float_to_half(float f) {
  uint32_t x = f2u(f);
  uint32_t sign = (x >> 16) & 0x8000u;
  uint32_t ax = x & 0x7fffffffu;
  if (ax >= 0x7f800000u) return (uint16_t)(sign | (ax > 0x7f800000u ? 0x7e00u : 0x7c00u));
  if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);  // rounds to >= 65520 -> inf
  if (ax < 0x38800000u) {                                      // below 2^-14: subnormal
    // value = ax as float; half subnormal unit = 2^-24
    float a = u2f(ax);
    float scaled = a * 16777216.0f;  // exact (power of two)
    uint32_t m = (uint32_t)std::nearbyint(scaled);  // RNE, 0..1024
    return (uint16_t)(sign | m);
  }
  uint32_t exp = (ax >> 23) - 127 + 15;
  uint32_t mant = ax & 0x7fffffu;
  uint32_t h = (exp << 10) | (mant >> 13);
  uint32_t rem = mant & 0x1fffu;
  if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;
  return (uint16_t)(sign | h);
}

OAPI float orc_half_to_float(uint16_t h) {
  uint32_t sign = ((uint32_t)h & 0x8000u) << 16;
  uint32_t exp = (h >> 10) & 0x1fu;
  uint32_t mant = h & 0x3ffu;
  if (exp == 0) {
    float v = (float)mant * 5.9604644775390625e-08f;  // 2^-24
    return sign ? -v : v;
  }
  if (exp == 31) return u2f(sign | 0x7f800000u | (mant << 13));
  return u2f(sign | ((exp - 15 + 127) << 23) | (mant << 13));
}

OAPI uint16_t orc_float_to_bf16(float f) {
  uint32_t x = f2u(f);
  if ((x & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((x >> 16) | 0x40u);
  uint32_t lsb = (x >> 16) & 1u;
  x += 0x7fffu + lsb;
  return (uint16_t)(x >> 16);
}

OAPI float orc_bf16_to_float(uint16_t b) { return u2f((uint32_t)b << 16); }

template <typename T> static inline T store_round(T v, int prec, bool alpha);
template <> inline float store_round<float>(float v, int prec, bool alpha) {
  if (prec == 0) return v;
  if (prec == 1 || alpha) return orc_half_to_float(orc_float_to_half(v));
  return orc_bf16_to_float(orc_float_to_bf16(v));
}
template <> inline double store_round<double>(double v, int, bool) { return v; }

// ---------------------------------------------------------------------------
// (2) radiance injection + mip chain
// ---------------------------------------------------------------------------

template <typename T> static T gamma_lut_value(unsigned i) {
  return (T)std::pow((double)i / 255.0, 2.2);
}

// Shadow flag in orc_light.type; the low byte is the light's kind.
static const uint32_t kLightShadow = 0x100u;

// Visibility of light L from occupied voxel (i, j, k) of the level-0 grid (acc: dense
// accumulator, occupied = nonzero count). Integers and single-precision IEEE operations
// only, so the decision is the same at every arithmetic width of the radiance:
//  * start: the voxel centre moved one voxel along the dominant axis of the voxel's
//    integer normal sum ns (ties to the lowest axis; no move when ns is zero), in voxel
//    units p0 = (i, j, k) + o + 0.5;
//  * direction toward the light: -vec for a directional light; Pv - p0 for a point light,
//    Pv = (vec - origin) / h in float (h = extent / n);
//  * dominant axis a of the direction l (ties to the lowest axis); step s_x = l_x / |l_a|
//    (so s_a = +-1), in 16.16 fixed point rint(s_x * 65536);
//  * samples q_m = p0 + m s for m = 0, 1, ... in fixed point, each naming the voxel
//    floor(q_m); the walk ends lit when a sample leaves the grid and, for a point light,
//    when m reaches ceil(|l_a|) (the sample has reached the light along a); a sample
//    naming the voxel itself is passed over; the first occupied one ends it unlit.
// Every step moves exactly one voxel along a, so a walk takes at most n + 1 samples.
static bool light_visible(const uint32_t* acc, uint32_t n, const float origin[3], float h,
                          const orc_light& L, uint32_t i, uint32_t j, uint32_t k, const int32_t ns[3]) {
  const int32_t v[3] = {(int32_t)i, (int32_t)j, (int32_t)k};
  int32_t o[3] = {0, 0, 0};
  {
    const int64_t m0 = std::llabs((long long)ns[0]), m1 = std::llabs((long long)ns[1]),
                  m2 = std::llabs((long long)ns[2]);
    const int na = (m0 >= m1 && m0 >= m2) ? 0 : (m1 >= m2 ? 1 : 2);
    o[na] = ns[na] > 0 ? 1 : (ns[na] < 0 ? -1 : 0);
  }
  float l[3];
  for (int x = 0; x < 3; x++) {
    if ((L.type & 0xffu) == 0) {
      l[x] = -L.vec[x];
    } else {
      const float pv = (L.vec[x] - origin[x]) / h;
      l[x] = pv - ((float)(v[x] + o[x]) + 0.5f);
    }
  }
  const float m0 = std::fabs(l[0]), m1 = std::fabs(l[1]), m2 = std::fabs(l[2]);
  const int a = (m0 >= m1 && m0 >= m2) ? 0 : (m1 >= m2 ? 1 : 2);
  const float la = std::fabs(l[a]);
  if (!(la > 0.0f)) return true;
  int32_t s[3];
  for (int x = 0; x < 3; x++) s[x] = (int32_t)std::nearbyint((l[x] / la) * 65536.0f);
  int64_t steps = (int64_t)n + 2;  // leaves the grid before this
  if ((L.type & 0xffu) != 0) {
    const float c = std::ceil(la);
    if (c < (float)steps) steps = (int64_t)c;
  }
  int32_t q[3];
  for (int x = 0; x < 3; x++) q[x] = (v[x] + o[x]) * 65536 + 32768;
  for (int64_t m = 0; m < steps; m++) {
    int32_t c[3];
    bool out = false;
    for (int x = 0; x < 3; x++) {
      c[x] = q[x] >> 16;  // floor (arithmetic shift)
      out = out || c[x] < 0 || c[x] >= (int32_t)n;
    }
    if (out) return true;
    if (!(c[0] == v[0] && c[1] == v[1] && c[2] == v[2])) {
      const size_t lin = (size_t)c[0] + (size_t)n * ((size_t)c[1] + (size_t)n * (size_t)c[2]);
      if (acc[lin * 8 + 3] != 0) return false;
    }
    for (int x = 0; x < 3; x++) q[x] += s[x];
  }
  return true;
}

template <typename T>
static void inject_voxel(const uint32_t* v, uint32_t i, uint32_t j, uint32_t k, const T origin[3],
                         T h, const orc_light* lights, uint32_t nl, const T* lut, T out[4],
                         const uint32_t* acc, uint32_t n, const float originf[3], float hf) {
  uint32_t c = v[3];
  if (c == 0) { out[0] = out[1] = out[2] = out[3] = (T)0; return; }
  T lin[3];
  for (int ch = 0; ch < 3; ch++) lin[ch] = lut[(v[ch] + c / 2) / c];
  T inv = (T)1 / ((T)c * (T)1024);
  T nrm[3];
  for (int a = 0; a < 3; a++) nrm[a] = (T)((int32_t)v[4 + a] - (int32_t)(1024u * c)) * inv;
  T pos[3];
  uint32_t idx[3] = {i, j, k};
  for (int a = 0; a < 3; a++) pos[a] = origin[a] + ((T)idx[a] + (T)0.5) * h;
  int32_t ns[3];
  for (int a = 0; a < 3; a++) ns[a] = (int32_t)v[4 + a] - (int32_t)(1024u * c);
  T E[3] = {0, 0, 0};
  for (uint32_t l = 0; l < nl; l++) {
    const orc_light& L = lights[l];
    const bool shadow = (L.type & kLightShadow) != 0;
    if ((L.type & 0xffu) == 0) {
      T cosv = ((nrm[0] * -(T)L.vec[0] + nrm[1] * -(T)L.vec[1]) + nrm[2] * -(T)L.vec[2]);
      if (cosv > (T)0 && (!shadow || light_visible(acc, n, originf, hf, L, i, j, k, ns)))
        for (int ch = 0; ch < 3; ch++) E[ch] = E[ch] + (T)L.color[ch] * cosv;
    } else {
      T d[3];
      for (int a = 0; a < 3; a++) d[a] = (T)L.vec[a] - pos[a];
      T d2 = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2];
      if (d2 > (T)0) {
        T cosv = ((nrm[0] * d[0] + nrm[1] * d[1]) + nrm[2] * d[2]) / std::sqrt(d2);
        if (cosv > (T)0 && (!shadow || light_visible(acc, n, originf, hf, L, i, j, k, ns))) {
          T md = (T)L.min_distance;
          T att = std::max(d2, md * md);
          T f = cosv / att;
          for (int ch = 0; ch < 3; ch++) E[ch] = E[ch] + (T)L.color[ch] * f;
        }
      }
    }
  }
  for (int ch = 0; ch < 3; ch++) out[ch] = lin[ch] * E[ch];
  out[3] = (T)1;
}

// out: concatenated levels, level k = (n>>k)^3 texels x 4, x fastest.
template <typename T>
static int inject_mips(const uint32_t* acc, const float* originf, float extentf, uint32_t n,
                       uint32_t levels, const orc_light* lights, uint32_t nl, int prec, T* out) {
  if (levels == 0 || (n >> (levels - 1)) == 0) return -1;
  T lut[256];
  for (unsigned i = 0; i < 256; i++) lut[i] = gamma_lut_value<T>(i);
  T origin[3] = {(T)originf[0], (T)originf[1], (T)originf[2]};
  T h = (T)extentf / (T)n;
  const float hf = extentf / (float)n;  // the visibility walk's voxel size, single precision
  // level 0
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < (int64_t)n; k++)
    for (uint32_t j = 0; j < n; j++)
      for (uint32_t i = 0; i < n; i++) {
        size_t lin = (size_t)i + (size_t)n * ((size_t)j + (size_t)n * (size_t)k);
        T t[4];
        inject_voxel<T>(acc + lin * 8, i, j, (uint32_t)k, origin, h, lights, nl, lut, t, acc, n, originf, hf);
        for (int c = 0; c < 4; c++) out[lin * 4 + c] = store_round<T>(t[c], prec, c == 3);
      }
  // levels 1..L-1
  T* src = out;
  for (uint32_t l = 1; l < levels; l++) {
    uint32_t ns = n >> (l - 1), nd = n >> l;
    T* dst = src + (size_t)ns * ns * ns * 4;
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < (int64_t)nd; k++)
      for (uint32_t j = 0; j < nd; j++)
        for (uint32_t i = 0; i < nd; i++) {
          T s[4] = {0, 0, 0, 0};
          for (uint32_t dz = 0; dz < 2; dz++)
            for (uint32_t dy = 0; dy < 2; dy++)
              for (uint32_t dx = 0; dx < 2; dx++) {
                size_t cl = (size_t)(2 * i + dx) +
                            (size_t)ns * ((size_t)(2 * j + dy) + (size_t)ns * (size_t)(2 * k + dz));
                for (int c = 0; c < 4; c++) s[c] = s[c] + src[cl * 4 + c];
              }
          size_t dl = (size_t)i + (size_t)nd * ((size_t)j + (size_t)nd * (size_t)k);
          for (int c = 0; c < 4; c++) dst[dl * 4 + c] = store_round<T>(s[c] * (T)0.125, prec, c == 3);
        }
    src = dst;
  }
  return 0;
}

// Visibility walks (light_visible) of up to 32 lights from every occupied voxel: bit l of
// out[voxel] is set when light l is visible from it (whatever its shadow flag); 0 for
// unoccupied voxels. For tests of the walk itself.
OAPI int orc_light_visibility(const uint32_t* acc, const float* origin, float extent, uint32_t n,
                              const orc_light* lights, uint32_t nl, uint32_t* out) {
  if (nl > 32 || n == 0) return -1;
  const float hf = extent / (float)n;
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < (int64_t)n; k++)
    for (uint32_t j = 0; j < n; j++)
      for (uint32_t i = 0; i < n; i++) {
        const size_t lin = (size_t)i + (size_t)n * ((size_t)j + (size_t)n * (size_t)k);
        const uint32_t* v = acc + lin * 8;
        uint32_t bits = 0;
        if (v[3] != 0) {
          int32_t ns[3];
          for (int a = 0; a < 3; a++) ns[a] = (int32_t)v[4 + a] - (int32_t)(1024u * v[3]);
          for (uint32_t l = 0; l < nl; l++)
            if (light_visible(acc, n, origin, hf, lights[l], i, j, (uint32_t)k, ns)) bits |= 1u << l;
        }
        out[lin] = bits;
      }
  return 0;
}

OAPI int orc_inject_mips_f32(const uint32_t* acc, const float* origin, float extent, uint32_t n,
                        uint32_t levels, const orc_light* lights, uint32_t nl, int prec,
                        float* out) {
  return inject_mips<float>(acc, origin, extent, n, levels, lights, nl, prec, out);
}

OAPI int orc_inject_mips_f64(const uint32_t* acc, const float* origin, float extent, uint32_t n,
                        uint32_t levels, const orc_light* lights, uint32_t nl, double* out) {
  return inject_mips<double>(acc, origin, extent, n, levels, lights, nl, 0, out);
}

// ---------------------------------------------------------------------------
// (3) cone trace
// ---------------------------------------------------------------------------

template <typename T> struct Vol {
  const T* level[16];
  uint32_t nk[16];
  uint32_t levels;
  T origin[3];
  T h, inv_h;
  uint32_t n;
};

template <typename T>
static void trilinear(const Vol<T>& V, uint32_t k, const T u[3], T out[4]) {
  T sc = (T)1 / (T)(1u << k);
  int64_t i0[3];
  T f[3];
  for (int a = 0; a < 3; a++) {
    T g = u[a] * sc - (T)0.5;
    T fl = std::floor(g);
    i0[a] = (int64_t)fl;
    f[a] = g - fl;
  }
  int64_t nk = V.nk[k];
  const T* L = V.level[k];
  T c[2][2][2][4];
  for (int dz = 0; dz < 2; dz++)
    for (int dy = 0; dy < 2; dy++)
      for (int dx = 0; dx < 2; dx++) {
        int64_t x = i0[0] + dx, y = i0[1] + dy, z = i0[2] + dz;
        bool in = x >= 0 && y >= 0 && z >= 0 && x < nk && y < nk && z < nk;
        for (int ch = 0; ch < 4; ch++)
          c[dz][dy][dx][ch] = in ? L[((size_t)x + (size_t)nk * ((size_t)y + (size_t)nk * (size_t)z)) * 4 + ch]
                                 : (T)0;
      }
  for (int ch = 0; ch < 4; ch++) {
    T cy[2][2];
    for (int dz = 0; dz < 2; dz++)
      for (int dy = 0; dy < 2; dy++)
        cy[dz][dy] = c[dz][dy][0][ch] + (c[dz][dy][1][ch] - c[dz][dy][0][ch]) * f[0];
    T cz[2];
    for (int dz = 0; dz < 2; dz++) cz[dz] = cy[dz][0] + (cy[dz][1] - cy[dz][0]) * f[1];
    out[ch] = cz[0] + (cz[1] - cz[0]) * f[2];
  }
}

template <typename T>
static void cone_march(const Vol<T>& V, const T o[3], const T d[3], T tan_half, const orc_cones& P,
                       T out[4], uint64_t* tri_count) {
  T C[3] = {0, 0, 0}, a = 0;
  T t = (T)P.start_offset * V.h;
  T maxd = (T)P.max_distance;
  uint64_t cnt = 0;
  for (int it = 0; it < 4096 && t <= maxd; it++) {  // step cap, as in the product
    T u[3];
    bool inside = true;
    for (int ax = 0; ax < 3; ax++) {
      T x = o[ax] + d[ax] * t;
      u[ax] = (x - V.origin[ax]) * V.inv_h;
      if (!(u[ax] >= (T)0 && u[ax] <= (T)V.n)) inside = false;
    }
    if (!inside) break;
    T diam = std::max(V.h, (T)2 * t * tan_half);
    T lvl = std::max((T)0, std::log2(diam * V.inv_h));
    T s[4];
    if (lvl >= (T)(V.levels - 1)) {
      trilinear<T>(V, V.levels - 1, u, s);
      cnt += 1;
    } else {
      T fl = std::floor(lvl);
      uint32_t l0 = (uint32_t)fl;
      T fr = lvl - fl;
      T s0[4], s1[4];
      trilinear<T>(V, l0, u, s0);
      trilinear<T>(V, l0 + 1, u, s1);
      cnt += 2;
      for (int ch = 0; ch < 4; ch++) s[ch] = s0[ch] + (s1[ch] - s0[ch]) * fr;
    }
    // a sample's opacity is a convex combination of values in [0, 1]; rounding of its
    // weights can push it an ulp above 1, and the clamp keeps 0 <= a <= 1 (as the product)
    s[3] = std::min(s[3], (T)1);
    T w = (T)1 - a;
    for (int ch = 0; ch < 3; ch++) C[ch] = C[ch] + w * s[ch];
    a = a + w * s[3];
    if (a >= (T)P.alpha_cutoff) break;
    t = t + (T)P.step_factor * diam;
  }
  out[0] = C[0]; out[1] = C[1]; out[2] = C[2]; out[3] = a;
  if (tri_count) *tri_count += cnt;
}

// Diffuse cone table in the tangent frame (x, y, z = normal) and weights, computed
// in double and rounded to T.
template <typename T>
static void cone_table(const orc_cones& P, T dirs[][3], T* w) {
  uint32_t K = P.num_diffuse;
  if (K == 0) return;
  const double PI = 3.14159265358979323846;
  double ring = (double)P.diffuse_ring_deg * PI / 180.0;
  double wsum = 1.0 + (double)(K - 1) * std::cos(ring);
  for (uint32_t c = 0; c < K; c++) {
    if (c == 0) {
      dirs[c][0] = 0; dirs[c][1] = 0; dirs[c][2] = 1;
      w[c] = (T)(1.0 / wsum);
    } else {
      double phi = 2.0 * PI * (double)(c - 1) / (double)(K - 1);
      dirs[c][0] = (T)(std::sin(ring) * std::cos(phi));
      dirs[c][1] = (T)(std::sin(ring) * std::sin(phi));
      dirs[c][2] = (T)std::cos(ring);
      w[c] = (T)(std::cos(ring) / wsum);
    }
  }
}

// Orthonormal basis around unit n: Duff et al. 2017, "Building an Orthonormal Basis,
// Revisited".
template <typename T> static void onb(const T n[3], T t[3], T b[3]) {
  T sign = std::copysign((T)1, n[2]);
  T a = (T)-1 / (sign + n[2]);
  T bb = n[0] * n[1] * a;
  t[0] = (T)1 + sign * n[0] * n[0] * a;
  t[1] = sign * bb;
  t[2] = -sign * n[0];
  b[0] = bb;
  b[1] = sign + n[1] * n[1] * a;
  b[2] = -n[1];
}

// levels_data: concatenated levels as produced by orc_inject_mips_*.
// pixels: indices of the pixels to trace (or NULL for all num_pixels).
template <typename T>
static int trace(const T* levels_data, const float* originf, float extentf, uint32_t n,
                 uint32_t levels, const float* pos, const float* nrm, const float* rough,
                 const int64_t* pixels, uint64_t num_pixels, const orc_cones* Pp, T* out_diffuse,
                 T* out_specular, uint64_t* samples_out) {
  const orc_cones& P = *Pp;
  if (P.num_diffuse > 32) return -1;
  Vol<T> V;
  V.levels = levels;
  V.n = n;
  size_t off = 0;
  for (uint32_t k = 0; k < levels; k++) {
    V.nk[k] = n >> k;
    V.level[k] = levels_data + off;
    off += (size_t)V.nk[k] * V.nk[k] * V.nk[k] * 4;
  }
  for (int a = 0; a < 3; a++) V.origin[a] = (T)originf[a];
  V.h = (T)extentf / (T)n;
  V.inv_h = (T)n / (T)extentf;
  T dirs[32][3], w[32];
  cone_table<T>(P, dirs, w);
  const double PI = 3.14159265358979323846;
  T tan_d = (T)std::tan((double)P.diffuse_aperture_deg * PI / 360.0);
  T tan_s = (T)std::tan((double)P.specular_aperture_deg * PI / 360.0);
  uint64_t total = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : total)
  for (int64_t q = 0; q < (int64_t)num_pixels; q++) {
    int64_t p = pixels ? pixels[q] : q;
    T* od = out_diffuse + q * 4;
    T* os = out_specular ? out_specular + q * 4 : nullptr;
    T Pw[3] = {(T)pos[p * 3], (T)pos[p * 3 + 1], (T)pos[p * 3 + 2]};
    T Nn[3] = {(T)nrm[p * 3], (T)nrm[p * 3 + 1], (T)nrm[p * 3 + 2]};
    T len2 = (Nn[0] * Nn[0] + Nn[1] * Nn[1]) + Nn[2] * Nn[2];
    if (!(len2 > (T)0)) {
      for (int c = 0; c < 4; c++) { od[c] = 0; if (os) os[c] = 0; }
      continue;
    }
    T inv = (T)1 / std::sqrt(len2);
    T nv[3] = {Nn[0] * inv, Nn[1] * inv, Nn[2] * inv};
    T o[3];
    for (int a = 0; a < 3; a++) o[a] = Pw[a] + nv[a] * ((T)P.normal_offset * V.h);
    T tv[3], bv[3];
    onb<T>(nv, tv, bv);
    T acc[4] = {0, 0, 0, 0};
    uint64_t cnt = 0;
    for (uint32_t c = 0; c < P.num_diffuse; c++) {
      T d[3];
      for (int a = 0; a < 3; a++) d[a] = (tv[a] * dirs[c][0] + bv[a] * dirs[c][1]) + nv[a] * dirs[c][2];
      T r[4];
      cone_march<T>(V, o, d, tan_d, P, r, &cnt);
      for (int ch = 0; ch < 4; ch++) acc[ch] = acc[ch] + w[c] * r[ch];
    }
    od[0] = acc[0]; od[1] = acc[1]; od[2] = acc[2]; od[3] = (T)1 - acc[3];
    if (os) {
      if (P.specular) {
        T v[3] = {(T)P.eye[0] - Pw[0], (T)P.eye[1] - Pw[1], (T)P.eye[2] - Pw[2]};
        T vl2 = (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2];
        T vi = vl2 > (T)0 ? (T)1 / std::sqrt(vl2) : (T)0;
        T I[3] = {-v[0] * vi, -v[1] * vi, -v[2] * vi};
        T dn = (nv[0] * I[0] + nv[1] * I[1]) + nv[2] * I[2];
        T r[3];
        for (int a = 0; a < 3; a++) r[a] = I[a] - (T)2 * dn * nv[a];
        T th = tan_s;
        if (rough) th = std::min(std::max((T)rough[p], (T)0.0087), (T)1.0);
        T res[4];
        cone_march<T>(V, o, r, th, P, res, &cnt);
        for (int ch = 0; ch < 4; ch++) os[ch] = res[ch];
      } else {
        for (int c = 0; c < 4; c++) os[c] = 0;
      }
    }
    total += cnt;
  }
  if (samples_out) *samples_out = total;
  return 0;
}

OAPI int orc_trace_f32(const float* levels_data, const float* origin, float extent, uint32_t n,
                  uint32_t levels, const float* pos, const float* nrm, const float* rough,
                  const int64_t* pixels, uint64_t num_pixels, const orc_cones* P,
                  float* out_diffuse, float* out_specular, uint64_t* samples) {
  return trace<float>(levels_data, origin, extent, n, levels, pos, nrm, rough, pixels, num_pixels,
                      P, out_diffuse, out_specular, samples);
}

OAPI int orc_trace_f64(const double* levels_data, const float* origin, float extent, uint32_t n,
                  uint32_t levels, const float* pos, const float* nrm, const float* rough,
                  const int64_t* pixels, uint64_t num_pixels, const orc_cones* P,
                  double* out_diffuse, double* out_specular, uint64_t* samples) {
  return trace<double>(levels_data, origin, extent, n, levels, pos, nrm, rough, pixels,
                       num_pixels, P, out_diffuse, out_specular, samples);
}

// Closed-form step schedule of one cone (for tests): the distances t_i at which a
// cone of half-angle tangent k samples, with no early exit, up to max_distance.
OAPI int orc_cone_schedule_f64(double h, double k, double step_factor, double start, double max_d,
                          double* out, int cap) {
  int m = 0;
  double t = start * h;
  while (t <= max_d && m < cap) {
    out[m++] = t;
    double diam = std::max(h, 2.0 * t * k);
    t = t + step_factor * diam;
  }
  return m;
}

// ---------------------------------------------------------------------------
// (4) reference-behaviour restatements (ref_mode)
// ---------------------------------------------------------------------------

// VoxelGS orientation code from a face normal (main.lua:100-115):
// 0 = Z (also on ties), 1 = X, 2 = Y.
OAPI int orc_ref_orientation(float nx, float ny, float nz) {
  float x = std::fabs(nx), y = std::fabs(ny), z = std::fabs(nz);
  if (x > y) return (x > z) ? 1 : 0;
  return (y > z) ? 2 : 0;
}

// VoxelPS packing (main.lua:270-271); components beyond the encodable range wrap the
// way an unsigned conversion of the float would after truncation toward zero.
static uint32_t glsl_uint(float f) {
  if (f >= 0.0f) return (uint32_t)f;
  return (uint32_t)(int32_t)f;  // negative: undefined in GLSL; two's complement here
}
OAPI uint32_t orc_ref_pack_color565(float r, float g, float b) {
  return (glsl_uint(r * 31.0f) << 11) | (glsl_uint(g * 63.0f) << 5) | glsl_uint(b * 31.0f);
}
OAPI uint32_t orc_ref_pack_normal565(float nx, float ny, float nz) {
  return (glsl_uint(nx * 16.0f + 15.0f) << 11) | (glsl_uint(ny * 32.0f + 31.0f) << 5) |
         glsl_uint(nz * 16.0f + 15.0f);
}
// indirect.frag:88-94
OAPI void orc_ref_unpack_color565(uint32_t x, float* out3) {
  out3[0] = (float)((x & 0xF800u) >> 11) / 31.0f;
  out3[1] = (float)((x & 0x7E0u) >> 5) / 63.0f;
  out3[2] = (float)(x & 0x1Fu) / 31.0f;
}
// indirect.frag:96-102 (reads x from the low bits: an x<->z swap against the packer)
OAPI void orc_ref_unpack_normal565(uint32_t x, float* out3) {
  float v[3] = {(float)(x & 0x1Fu) / 16.0f - 1.0f, (float)((x & 0x7E0u) >> 5) / 32.0f - 1.0f,
                (float)((x & 0xF800u) >> 11) / 16.0f - 1.0f};
  float l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  for (int a = 0; a < 3; a++) out3[a] = v[a] / l;
}

// hemisphereSample_cos (Shader/math.inc:76-81)
OAPI void orc_ref_hemisphere_cos(float u, float v, float* out3) {
  const float PI = 3.14159265358979f;
  float phi = v * 2.0f * PI;
  float cosTheta = std::sqrt(1.0f - u);
  float sinTheta = std::sqrt(1.0f - cosTheta * cosTheta);
  out3[0] = std::cos(phi) * sinTheta;
  out3[1] = std::sin(phi) * sinTheta;
  out3[2] = cosTheta;
}

static void cross3(const float a[3], const float b[3], float o[3]) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}
static void normalize3(float v[3]) {
  float l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  for (int a = 0; a < 3; a++) v[a] /= l;
}

// make_coord_space (indirect.frag:71-86); out9 = columns x, y, z.
OAPI void orc_ref_make_coord_space(const float* n, float* out9) {
  float z[3] = {n[0], n[1], n[2]};
  float h[3] = {z[0], z[1], z[2]};
  float ax = std::fabs(h[0]), ay = std::fabs(h[1]), az = std::fabs(h[2]);
  if (ax <= ay && ax <= az) h[0] = 1.0f;
  else if (ay <= ax && ay <= az) h[1] = 1.0f;
  else h[2] = 1.0f;
  normalize3(z);
  float y[3];
  cross3(h, z, y);
  normalize3(y);
  float x[3];
  cross3(z, y, x);
  normalize3(x);
  for (int a = 0; a < 3; a++) { out9[a] = x[a]; out9[3 + a] = y[a]; out9[6 + a] = z[a]; }
}

// getIndirect's march (indirect.frag:104-184) with the random numbers given and the
// shadow test disabled (shade = 1), over a 128^3 RG16UI volume (2 x uint16 per voxel,
// x fastest). w2voxel: column-major 4x4 VoxelProj*VoxelView. Returns 1 on hit and
// writes Lo (3), hit position (3) and the march step index.
OAPI int orc_ref_get_indirect(const uint16_t* voxels, const float* w2voxel, const float* wpos,
                         const float* wnorm, float r0, float r1, const float* sun_dir,
                         const float* sun_lum, float* Lo_out, float* hit_pos, int* steps) {
  const float step_size = 0.2f;
  float dir_l[3];
  orc_ref_hemisphere_cos(r0, r1, dir_l);
  float m[9];
  orc_ref_make_coord_space(wnorm, m);
  float dir[3];
  for (int a = 0; a < 3; a++) dir[a] = m[a] * dir_l[0] + m[3 + a] * dir_l[1] + m[6 + a] * dir_l[2];
  float dn = dir[0] * wnorm[0] + dir[1] * wnorm[1] + dir[2] * wnorm[2];
  if (dn < 0.0f) { for (int a = 0; a < 3; a++) dir[a] = -dir[a]; dn = -dn; }
  float NdotD = dn;
  auto to_voxel = [&](const float p[3], float out[3]) {
    for (int r = 0; r < 3; r++)
      out[r] = w2voxel[0 * 4 + r] * p[0] + w2voxel[1 * 4 + r] * p[1] + w2voxel[2 * 4 + r] * p[2] +
               w2voxel[3 * 4 + r];
    out[0] = out[0] * 0.5f + 0.5f;
    out[1] = out[1] * 0.5f + 0.5f;
    for (int a = 0; a < 3; a++) out[a] *= 128.0f;
  };
  float march[3];
  for (int a = 0; a < 3; a++) march[a] = wpos[a] + dir[a] * (1.0f + r1) * step_size / dn;
  float sv[3];
  to_voxel(wpos, sv);
  int prev[3] = {(int)sv[0], (int)sv[1], (int)sv[2]};
  Lo_out[0] = Lo_out[1] = Lo_out[2] = 0.0f;
  int i = 0;
  for (; i < 60; i++) {
    for (int a = 0; a < 3; a++) march[a] += dir[a] * step_size;
    float vp[3];
    to_voxel(march, vp);
    if (vp[0] < 0 || vp[1] < 0 || vp[2] < 0 || vp[0] > 127 || vp[1] > 127 || vp[2] > 127) break;
    int iv[3] = {(int)vp[0], (int)vp[1], (int)vp[2]};
    if (iv[0] != prev[0] || iv[1] != prev[1] || iv[2] != prev[2]) {
      const uint16_t* vd = voxels + ((size_t)iv[0] + 128u * ((size_t)iv[1] + 128u * (size_t)iv[2])) * 2;
      prev[0] = iv[0]; prev[1] = iv[1]; prev[2] = iv[2];
      if (vd[0] != 0) {
        float col[3], vn[3];
        orc_ref_unpack_color565(vd[0], col);
        orc_ref_unpack_normal565(vd[1], vn);
        for (int a = 0; a < 3; a++) col[a] = std::pow(col[a], 2.2f);
        float lam = std::fabs(-sun_dir[0] * vn[0] - sun_dir[1] * vn[1] - sun_dir[2] * vn[2]);
        float den = std::max(0.01f, NdotD);
        for (int a = 0; a < 3; a++) Lo_out[a] += sun_lum[a] * (lam * col[a] / den);
        for (int a = 0; a < 3; a++) hit_pos[a] = march[a];
        *steps = i;
        return 1;
      }
    }
  }
  *steps = i;
  if (i == 60) {
    float den = std::max(0.01f, NdotD);
    float t = std::min(std::max(NdotD / 0.01f, 0.0f), 1.0f);
    float ss = t * t * (3.0f - 2.0f * t);
    const float sky[3] = {0.7f, 0.8f, 1.0f};
    for (int a = 0; a < 3; a++) Lo_out[a] += sky[a] * 0.4f / den * ss;
  }
  return 0;
}


// ---------------------------------------------------------------------------
// (5) G-buffer reconstruction (indirect.frag:44-58, 189-196)
// ---------------------------------------------------------------------------
// Matrices are GLSL mat4, column-major: m[c * 4 + r]. Pixel (x, y) is sampled at its
// centre uv = ((x + 0.5) / W, (y + 0.5) / H). Sums run in the order written; the FP32
// form is the arithmetic the product performs (IEEE division and square root).
template <typename T>
static void gbuffer_decode(const float* depth, const uint16_t* normals, const float* inv_proj,
                           const float* inv_mv, uint32_t W, uint32_t H, float* out_pos, float* out_nrm,
                           float* out_view_depth) {
  T P[16], M[16];
  for (int i = 0; i < 16; i++) {
    P[i] = (T)inv_proj[i];
    M[i] = (T)inv_mv[i];
  }
#pragma omp parallel for schedule(static)
  for (int64_t y = 0; y < (int64_t)H; y++)
    for (uint32_t x = 0; x < W; x++) {
      const size_t i = (size_t)y * W + x;
      const T u = ((T)x + (T)0.5) / (T)W, v = ((T)y + (T)0.5) / (T)H;
      const T p[4] = {u * (T)2 - (T)1, v * (T)2 - (T)1, (T)depth[i], (T)1};
      T c[4];
      for (int r = 0; r < 4; r++) c[r] = ((P[r] * p[0] + P[4 + r] * p[1]) + P[8 + r] * p[2]) + P[12 + r] * p[3];
      const T cs[3] = {c[0] / c[3], c[1] / c[3], c[2] / c[3]};
      for (int r = 0; r < 3; r++)
        out_pos[i * 3 + r] = (float)(((M[r] * cs[0] + M[4 + r] * cs[1]) + M[8 + r] * cs[2]) + M[12 + r]);
      T n[3];
      for (int k = 0; k < 3; k++) n[k] = std::fma((T)normals[i * 4 + k] / (T)65535, (T)2, (T)-1);
      const T len2 = (n[0] * n[0] + n[1] * n[1]) + n[2] * n[2];
      if (len2 > (T)0) {
        const T len = std::sqrt(len2);
        for (int k = 0; k < 3; k++) n[k] = n[k] / len;
      } else {
        n[0] = n[1] = n[2] = (T)0;  // no surface: the trace skips zero normals
      }
      for (int r = 0; r < 3; r++) out_nrm[i * 3 + r] = (float)((M[r] * n[0] + M[4 + r] * n[1]) + M[8 + r] * n[2]);
      if (out_view_depth) out_view_depth[i] = (float)(-cs[2]);
    }
}

OAPI void orc_gbuffer_decode_f32(const float* depth, const uint16_t* normals, const float* inv_proj,
                                 const float* inv_mv, uint32_t W, uint32_t H, float* out_pos, float* out_nrm,
                                 float* out_view_depth) {
  gbuffer_decode<float>(depth, normals, inv_proj, inv_mv, W, H, out_pos, out_nrm, out_view_depth);
}

OAPI void orc_gbuffer_decode_f64(const float* depth, const uint16_t* normals, const float* inv_proj,
                                 const float* inv_mv, uint32_t W, uint32_t H, float* out_pos, float* out_nrm,
                                 float* out_view_depth) {
  gbuffer_decode<double>(depth, normals, inv_proj, inv_mv, W, H, out_pos, out_nrm, out_view_depth);
}

// ---------------------------------------------------------------------------
// (6) temporal blend (indirect.frag:224-242), FP32 in the written order
// ---------------------------------------------------------------------------
static inline float mrow(const float* m, int r, float a, float b, float c, float d) {
  return ((m[r] * a + m[4 + r] * b) + m[8 + r] * c) + m[12 + r] * d;
}

OAPI void orc_temporal(const float* indirect, const float* pos, const float* view_depth, const float* history,
                       const float* prev_proj, const float* prev_mv, uint32_t W, uint32_t H, float* out) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)W * H; i++) {
    const float* wp = pos + i * 3;
    const float csz = -view_depth[i];
    float c[4];
    for (int r = 0; r < 4; r++) c[r] = mrow(prev_mv, r, wp[0], wp[1], wp[2], 1.0f);
    float q[4];
    for (int r = 0; r < 4; r++) q[r] = mrow(prev_proj, r, c[0], c[1], c[2], c[3]);
    // prevProjPos /= prevProjPos.w; st * 0.5 + 0.5
    const float u = (q[0] / q[3]) * 0.5f + 0.5f, v = (q[1] / q[3]) * 0.5f + 0.5f;
    float col[3] = {indirect[i * 4], indirect[i * 4 + 1], indirect[i * 4 + 2]};
    if (u >= 0.0f && u <= 1.0f && v >= 0.0f && v <= 1.0f) {
      // GL_REPEAT bilinear: texel centres at (k + 0.5) / size, indices taken modulo size
      const float x = u * (float)W - 0.5f, y = v * (float)H - 0.5f;
      const float fx0 = std::floor(x), fy0 = std::floor(y);
      const float fx = x - fx0, fy = y - fy0;
      auto wrap = [](int64_t k, int64_t n) { return ((k % n) + n) % n; };
      const int64_t xa = wrap((int64_t)fx0, W), xb = wrap((int64_t)fx0 + 1, W);
      const int64_t ya = wrap((int64_t)fy0, H), yb = wrap((int64_t)fy0 + 1, H);
      const float* t00 = history + (ya * W + xa) * 4;
      const float* t10 = history + (ya * W + xb) * 4;
      const float* t01 = history + (yb * W + xa) * 4;
      const float* t11 = history + (yb * W + xb) * 4;
      float h[4];
      for (int k = 0; k < 4; k++)
        h[k] = (t00[k] * (1.0f - fx) + t10[k] * fx) * (1.0f - fy) + (t01[k] * (1.0f - fx) + t11[k] * fx) * fy;
      // smoothstep(0, 1, s) = t * t * (3 - 2t), t = clamp(s, 0, 1); mix(a, b, w) = a(1-w) + bw
      float t = 1.0f - std::fabs(h[3] + csz);
      t = std::fmin(std::fmax(t, 0.0f), 1.0f);
      const float blend = 0.95f * ((t * t) * (3.0f - 2.0f * t));
      for (int k = 0; k < 3; k++) col[k] = std::fmin(std::fmax(col[k] * (1.0f - blend) + h[k] * blend, 0.0f), 16.0f);
    }
    float* o = out + i * 4;
    o[0] = col[0];
    o[1] = col[1];
    o[2] = col[2];
    o[3] = -csz;
  }
}

// ---------------------------------------------------------------------------
// (7) cross-bilateral blur (bilateralBlur.inc:15-72)
// ---------------------------------------------------------------------------
// exp2 as the product defines it: 2^rint(x) * P7(x - rint(x)), P7 the Taylor polynomial
// of 2^f with FP32 coefficients ln2^j / j!, Horner with fused multiply-adds; 0 for
// x <= -125 or NaN.
OAPI float orc_exp2_poly(float x) {
  if (!(x > -125.0f)) return 0.0f;
  const float k = std::nearbyint(x);
  const float f = x - k;
  static const float c[8] = {1.0f, 6.931471825e-01f, 2.402265072e-01f, 5.550410971e-02f,
                             9.618128650e-03f, 1.333355787e-03f, 1.540352969e-04f, 1.525273365e-05f};
  float q = c[7];
  for (int j = 6; j >= 0; j--) q = std::fma(q, f, c[j]);
  return std::ldexp(q, (int)k);
}

OAPI void orc_bilateral_blur(const float* color, const float* depth, uint32_t W, uint32_t H, uint32_t axis,
                             float* out) {
  // (offset in pixels, r): SampleColor(DIR(i)) is i * 2 / resolution away in uv
  static const int off[12] = {2, -2, 4, -4, 6, -6, 8, -8, 11, -11, 15, -15};
  static const int rad[12] = {1, 1, 2, 2, 3, 3, 4, 4, 5, 5, 7, 7};
  const float sigma = 8.0f * 0.5f, falloff = 1.0f / (2.0f * sigma * sigma);  // KERNEL_RADIUS 8
#pragma omp parallel for schedule(static)
  for (int64_t y = 0; y < (int64_t)H; y++)
    for (int64_t x = 0; x < (int64_t)W; x++) {
      const int64_t i = y * W + x;
      const float d0 = depth[i];
      float tot[3] = {color[i * 4], color[i * 4 + 1], color[i * 4 + 2]}, tw = 1.0f;
      for (int j = 0; j < 12; j++) {
        const int64_t sx = axis == 0 ? std::min<int64_t>(std::max<int64_t>(x + off[j], 0), W - 1) : x;
        const int64_t sy = axis == 1 ? std::min<int64_t>(std::max<int64_t>(y + off[j], 0), H - 1) : y;
        const int64_t si = sy * W + sx;
        const float r = (float)rad[j];
        // -r*r*falloff - dz*dz and total += c * w, each as one fused multiply-add
        const float dz = (d0 - depth[si]) * 512.0f;
        const float w = orc_exp2_poly(std::fma(-dz, dz, -r * r * falloff));
        for (int k = 0; k < 3; k++) tot[k] = std::fma(color[si * 4 + k], w, tot[k]);
        tw = tw + w;
      }
      out[i * 4] = tot[0] / tw;
      out[i * 4 + 1] = tot[1] / tw;
      out[i * 4 + 2] = tot[2] / tw;
      out[i * 4 + 3] = 0.0f;
    }
}
