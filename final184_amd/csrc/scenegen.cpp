// scenegen.cpp — seeded synthetic inputs for the voxel GI path (host only).
//
// The reference renders Sponza (262,267 triangles; Models/Sponza.gltf, whose .bin is
// absent) through a G-buffer pass (Foreground/Renderer/MegaPipeline.cpp:416-467). Here the
// inputs are generated: random triangle soups, a procedural room, G-buffers sampled on
// triangles or ray cast from a pinhole camera, and light sets. Every generator is a pure
// function of its seed (splitmix64); results do not depend on the thread count.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
  double range(double a, double b) { return a + (b - a) * uniform(); }
  uint8_t byte() { return (uint8_t)(next() >> 56); }
};

// Per-stream seeds so that independent parts of a scene do not shift each other.
uint64_t sub_seed(uint64_t seed, uint64_t stream) {
  Rng r(seed ^ (0xA0761D6478BD642Full * (stream + 1)));
  return r.next();
}

struct V3 {
  double x, y, z;
};
V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
V3 normalize(V3 a) {
  double l = std::sqrt(dot(a, a));
  return l > 0 ? a * (1.0 / l) : a;
}

// Uniform point in the ball of radius r (rejection from the cube: no transcendentals).
V3 ball(Rng& g, double r) {
  for (;;) {
    V3 p{g.range(-1, 1), g.range(-1, 1), g.range(-1, 1)};
    if (dot(p, p) <= 1.0) return p * r;
  }
}

struct Soup {
  float* pos;       // 9 floats per triangle
  uint8_t* albedo;  // 3 per triangle
  float* rough;     // 1 per triangle (may be null)
  uint32_t cap;
  uint32_t n = 0;
  void tri(V3 a, V3 b, V3 c, const uint8_t rgb[3], float r) {
    if (n >= cap) return;
    float* p = pos + (size_t)n * 9;
    V3 v[3] = {a, b, c};
    for (int i = 0; i < 3; i++) {
      p[i * 3 + 0] = (float)v[i].x;
      p[i * 3 + 1] = (float)v[i].y;
      p[i * 3 + 2] = (float)v[i].z;
    }
    std::memcpy(albedo + (size_t)n * 3, rgb, 3);
    if (rough) rough[n] = r;
    n++;
  }
  // quad a-b-c-d (counter-clockwise seen from the side the normal points to)
  void quad(V3 a, V3 b, V3 c, V3 d, const uint8_t rgb[3], float r) {
    tri(a, b, c, rgb, r);
    tri(a, c, d, rgb, r);
  }
};

// ---- ray casting over a uniform grid of triangle lists ---------------------
struct Accel {
  const float* pos;
  uint32_t ntri;
  V3 lo, hi;
  int res[3];
  V3 cell;
  std::vector<uint32_t> start, items;

  void build(const float* p, uint32_t n, int target_per_axis) {
    pos = p;
    ntri = n;
    lo = {1e30, 1e30, 1e30};
    hi = {-1e30, -1e30, -1e30};
    for (size_t i = 0; i < (size_t)n * 3; i++) {
      lo.x = std::min(lo.x, (double)p[i * 3]);
      lo.y = std::min(lo.y, (double)p[i * 3 + 1]);
      lo.z = std::min(lo.z, (double)p[i * 3 + 2]);
      hi.x = std::max(hi.x, (double)p[i * 3]);
      hi.y = std::max(hi.y, (double)p[i * 3 + 1]);
      hi.z = std::max(hi.z, (double)p[i * 3 + 2]);
    }
    V3 pad{1e-4, 1e-4, 1e-4};
    lo = lo - pad;
    hi = hi + pad;
    V3 ext = hi - lo;
    double m = std::max(ext.x, std::max(ext.y, ext.z));
    double ax[3] = {ext.x, ext.y, ext.z};
    for (int a = 0; a < 3; a++) res[a] = std::max(1, (int)std::ceil(target_per_axis * ax[a] / m));
    cell = {ext.x / res[0], ext.y / res[1], ext.z / res[2]};
    size_t nc = (size_t)res[0] * res[1] * res[2];
    std::vector<uint32_t> cnt(nc + 1, 0);
    auto range = [&](uint32_t t, int lo_i[3], int hi_i[3]) {
      const float* q = pos + (size_t)t * 9;
      double l[3] = {q[0], q[1], q[2]}, h[3] = {q[0], q[1], q[2]};
      for (int v = 1; v < 3; v++)
        for (int a = 0; a < 3; a++) {
          l[a] = std::min(l[a], (double)q[v * 3 + a]);
          h[a] = std::max(h[a], (double)q[v * 3 + a]);
        }
      double o[3] = {lo.x, lo.y, lo.z}, c[3] = {cell.x, cell.y, cell.z};
      for (int a = 0; a < 3; a++) {
        lo_i[a] = std::max(0, std::min(res[a] - 1, (int)std::floor((l[a] - o[a]) / c[a])));
        hi_i[a] = std::max(0, std::min(res[a] - 1, (int)std::floor((h[a] - o[a]) / c[a])));
      }
    };
    for (uint32_t t = 0; t < n; t++) {
      int a[3], b[3];
      range(t, a, b);
      for (int z = a[2]; z <= b[2]; z++)
        for (int y = a[1]; y <= b[1]; y++)
          for (int x = a[0]; x <= b[0]; x++) cnt[(size_t)x + res[0] * ((size_t)y + res[1] * (size_t)z)]++;
    }
    start.assign(nc + 1, 0);
    for (size_t i = 0; i < nc; i++) start[i + 1] = start[i] + cnt[i];
    items.resize(start[nc]);
    std::vector<uint32_t> cur(start.begin(), start.end() - 1);
    for (uint32_t t = 0; t < n; t++) {
      int a[3], b[3];
      range(t, a, b);
      for (int z = a[2]; z <= b[2]; z++)
        for (int y = a[1]; y <= b[1]; y++)
          for (int x = a[0]; x <= b[0]; x++) items[cur[(size_t)x + res[0] * ((size_t)y + res[1] * (size_t)z)]++] = t;
    }
  }

  // Moller-Trumbore; returns t or -1
  double hit_tri(uint32_t t, V3 o, V3 d) const {
    const float* q = pos + (size_t)t * 9;
    V3 a{q[0], q[1], q[2]}, b{q[3], q[4], q[5]}, c{q[6], q[7], q[8]};
    V3 e1 = b - a, e2 = c - a;
    V3 p = cross(d, e2);
    double det = dot(e1, p);
    if (std::fabs(det) < 1e-18) return -1;
    double inv = 1.0 / det;
    V3 s = o - a;
    double u = dot(s, p) * inv;
    if (u < 0 || u > 1) return -1;
    V3 qv = cross(s, e1);
    double v = dot(d, qv) * inv;
    if (v < 0 || u + v > 1) return -1;
    double tt = dot(e2, qv) * inv;
    return tt > 1e-7 ? tt : -1;
  }

  // nearest hit along o + t d (d unit); returns triangle id or -1
  int64_t cast(V3 o, V3 d, double* t_out) const {
    double od[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
    double l[3] = {lo.x, lo.y, lo.z}, h[3] = {hi.x, hi.y, hi.z}, c[3] = {cell.x, cell.y, cell.z};
    double t0 = 0, t1 = 1e30;
    for (int a = 0; a < 3; a++) {
      if (std::fabs(dd[a]) < 1e-30) {
        if (od[a] < l[a] || od[a] > h[a]) return -1;
        continue;
      }
      double ta = (l[a] - od[a]) / dd[a], tb = (h[a] - od[a]) / dd[a];
      if (ta > tb) std::swap(ta, tb);
      t0 = std::max(t0, ta);
      t1 = std::min(t1, tb);
    }
    if (t0 > t1) return -1;
    int ci[3];
    double tmax[3], tdelta[3];
    int step[3];
    for (int a = 0; a < 3; a++) {
      double p = od[a] + dd[a] * t0;
      ci[a] = std::max(0, std::min(res[a] - 1, (int)std::floor((p - l[a]) / c[a])));
      if (dd[a] > 0) {
        step[a] = 1;
        tmax[a] = (l[a] + (ci[a] + 1) * c[a] - od[a]) / dd[a];
        tdelta[a] = c[a] / dd[a];
      } else if (dd[a] < 0) {
        step[a] = -1;
        tmax[a] = (l[a] + ci[a] * c[a] - od[a]) / dd[a];
        tdelta[a] = -c[a] / dd[a];
      } else {
        step[a] = 0;
        tmax[a] = 1e30;
        tdelta[a] = 1e30;
      }
    }
    for (;;) {
      size_t cidx = (size_t)ci[0] + res[0] * ((size_t)ci[1] + res[1] * (size_t)ci[2]);
      double best = 1e30;
      int64_t bid = -1;
      for (uint32_t k = start[cidx]; k < start[cidx + 1]; k++) {
        double tt = hit_tri(items[k], o, d);
        if (tt > 0 && (tt < best || (tt == best && (int64_t)items[k] < bid))) {
          best = tt;
          bid = items[k];
        }
      }
      int a = (tmax[0] < tmax[1]) ? ((tmax[0] < tmax[2]) ? 0 : 2) : ((tmax[1] < tmax[2]) ? 1 : 2);
      if (bid >= 0 && best <= tmax[a] + 1e-12) {
        *t_out = best;
        return bid;
      }
      ci[a] += step[a];
      if (ci[a] < 0 || ci[a] >= res[a]) return -1;
      tmax[a] += tdelta[a];
    }
  }
};

}  // namespace

extern "C" {

// Random triangle soup: centres uniform in [lo+r, hi-r]^3, vertices in the ball of
// radius r around the centre (so every edge is <= 2r and the soup stays in [lo,hi]^3);
// albedo uniform RGB8.
int sg_random_triangles(uint64_t seed, uint32_t ntri, float lo, float hi, float radius,
                        float* pos, uint8_t* albedo) {
  Rng g(sub_seed(seed, 1));
  for (uint32_t t = 0; t < ntri; t++) {
    V3 c{g.range(lo + radius, hi - radius), g.range(lo + radius, hi - radius),
         g.range(lo + radius, hi - radius)};
    for (int v = 0; v < 3; v++) {
      V3 p = c + ball(g, radius);
      pos[(size_t)t * 9 + v * 3 + 0] = (float)p.x;
      pos[(size_t)t * 9 + v * 3 + 1] = (float)p.y;
      pos[(size_t)t * 9 + v * 3 + 2] = (float)p.z;
    }
    albedo[(size_t)t * 3 + 0] = g.byte();
    albedo[(size_t)t * 3 + 1] = g.byte();
    albedo[(size_t)t * 3 + 2] = g.byte();
  }
  return 0;
}

// Procedural room: axis-aligned box [0,dx]x[0,dy]x[0,dz] (y up) with 6 inward-facing
// walls of 12 triangles, 64 capped 32-segment columns (radius 0.02) on an 8x8 layout,
// 20,000 floor and 20,000 ceiling tiles (2 triangles each), padded with small debris
// triangles (vertices within 0.008 of a centre in the lower 0.12 of the room) to
// ntri_total. Albedo is uniform RGB8 per object (per triangle for debris) and roughness
// uniform in [0.05, 0.5] per object. Returns the number of triangles written.
int64_t sg_room(uint64_t seed, uint32_t ntri_total, float dx, float dy, float dz, float* pos,
                uint8_t* albedo, float* rough) {
  Soup s{pos, albedo, rough, ntri_total};
  Rng g(sub_seed(seed, 2));
  auto color = [&](uint8_t c[3]) { c[0] = g.byte(); c[1] = g.byte(); c[2] = g.byte(); };
  auto rgh = [&]() { return (float)g.range(0.05, 0.5); };
  const double X = dx, Y = dy, Z = dz;
  // walls: 2x3 quads each, normals into the room
  struct Wall { V3 o, u, v; };  // o + a u + b v, normal = u x v (into the room)
  Wall walls[6] = {
      {{0, 0, 0}, {0, 0, Z}, {0, Y, 0}},  // x = 0, normal +x: (0,0,Z)x(0,Y,0) = (-ZY..)
      {{X, 0, 0}, {0, Y, 0}, {0, 0, Z}},  // x = X, normal -x
      {{0, 0, 0}, {X, 0, 0}, {0, 0, Z}},  // y = 0 (floor), normal +y
      {{0, Y, 0}, {0, 0, Z}, {X, 0, 0}},  // y = Y (ceiling), normal -y
      {{0, 0, 0}, {0, Y, 0}, {X, 0, 0}},  // z = 0, normal +z
      {{0, 0, Z}, {X, 0, 0}, {0, Y, 0}},  // z = Z, normal -z
  };
  for (auto& w : walls) {
    V3 nrm = cross(w.u, w.v);
    V3 centre{X / 2, Y / 2, Z / 2};
    V3 mid = w.o + w.u * 0.5 + w.v * 0.5;
    bool flip = dot(nrm, centre - mid) < 0;
    uint8_t c[3];
    color(c);
    float r = rgh();
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 3; j++) {
        V3 a = w.o + w.u * (i / 2.0) + w.v * (j / 3.0);
        V3 b = w.o + w.u * ((i + 1) / 2.0) + w.v * (j / 3.0);
        V3 cc = w.o + w.u * ((i + 1) / 2.0) + w.v * ((j + 1) / 3.0);
        V3 d = w.o + w.u * (i / 2.0) + w.v * ((j + 1) / 3.0);
        if (flip) s.quad(a, d, cc, b, c, r);
        else s.quad(a, b, cc, d, c, r);
      }
  }
  // columns
  const double PI = 3.14159265358979323846;
  const int SEG = 32;
  const double R = 0.02;
  for (int ci = 0; ci < 8; ci++)
    for (int cj = 0; cj < 8; cj++) {
      double cx = X * (0.1 + 0.8 * ci / 7.0), cz = Z * (0.1 + 0.8 * cj / 7.0);
      uint8_t c[3];
      color(c);
      float r = rgh();
      for (int k = 0; k < SEG; k++) {
        double a0 = 2 * PI * k / SEG, a1 = 2 * PI * (k + 1) / SEG;
        V3 p0{cx + R * std::cos(a0), 0, cz + R * std::sin(a0)};
        V3 p1{cx + R * std::cos(a1), 0, cz + R * std::sin(a1)};
        V3 q0{p0.x, Y, p0.z}, q1{p1.x, Y, p1.z};
        s.quad(p0, q0, q1, p1, c, r);  // outward
        V3 top{cx, Y, cz}, bot{cx, 0, cz};
        s.tri(top, q1, q0, c, r);
        s.tri(bot, p0, p1, c, r);
      }
    }
  // floor and ceiling tiles: 200 x 100 cells, tile = inner 80% of the cell
  const int TX = 200, TZ = 100;
  for (int side = 0; side < 2; side++) {
    double y = side == 0 ? 0.0015 : Y - 0.0015;
    for (int j = 0; j < TZ; j++)
      for (int i = 0; i < TX; i++) {
        double x0 = X * (i + 0.1) / TX, x1 = X * (i + 0.9) / TX;
        double z0 = Z * (j + 0.1) / TZ, z1 = Z * (j + 0.9) / TZ;
        uint8_t c[3];
        color(c);
        float r = rgh();
        V3 a{x0, y, z0}, b{x1, y, z0}, cc{x1, y, z1}, d{x0, y, z1};
        if (side == 0) s.quad(a, d, cc, b, c, r);  // normal +y
        else s.quad(a, b, cc, d, c, r);            // normal -y
      }
  }
  // debris
  Rng gd(sub_seed(seed, 3));
  while (s.n < ntri_total) {
    V3 c{gd.range(0.02 * X, 0.98 * X), gd.range(0.008, 0.12), gd.range(0.02 * Z, 0.98 * Z)};
    V3 v0 = c + ball(gd, 0.008), v1 = c + ball(gd, 0.008), v2 = c + ball(gd, 0.008);
    uint8_t col[3] = {gd.byte(), gd.byte(), gd.byte()};
    s.tri(v0, v1, v2, col, (float)gd.range(0.05, 0.5));
  }
  return s.n;
}

// G-buffer of npix points sampled on the triangles: triangle chosen with probability
// proportional to area, barycentrics (1-sqrt(u), sqrt(u)(1-v), sqrt(u) v), normal = unit
// face normal of the triangle as wound.
int sg_gbuffer_on_triangles(uint64_t seed, const float* pos, uint32_t ntri, uint64_t npix,
                            float* out_pos, float* out_nrm) {
  std::vector<double> cdf(ntri + 1, 0.0);
  for (uint32_t t = 0; t < ntri; t++) {
    const float* q = pos + (size_t)t * 9;
    V3 a{q[0], q[1], q[2]}, b{q[3], q[4], q[5]}, c{q[6], q[7], q[8]};
    V3 n = cross(b - a, c - a);
    cdf[t + 1] = cdf[t] + 0.5 * std::sqrt(dot(n, n));
  }
  if (cdf[ntri] <= 0) return -1;
  Rng g(sub_seed(seed, 4));
  for (uint64_t p = 0; p < npix; p++) {
    double x = g.uniform() * cdf[ntri];
    uint32_t t = (uint32_t)(std::upper_bound(cdf.begin() + 1, cdf.end(), x) - (cdf.begin() + 1));
    if (t >= ntri) t = ntri - 1;
    const float* q = pos + (size_t)t * 9;
    V3 a{q[0], q[1], q[2]}, b{q[3], q[4], q[5]}, c{q[6], q[7], q[8]};
    double su = std::sqrt(g.uniform()), v = g.uniform();
    V3 P = a * (1 - su) + b * (su * (1 - v)) + c * (su * v);
    V3 n = normalize(cross(b - a, c - a));
    out_pos[p * 3 + 0] = (float)P.x;
    out_pos[p * 3 + 1] = (float)P.y;
    out_pos[p * 3 + 2] = (float)P.z;
    out_nrm[p * 3 + 0] = (float)n.x;
    out_nrm[p * 3 + 1] = (float)n.y;
    out_nrm[p * 3 + 2] = (float)n.z;
  }
  return 0;
}

// Camera pose drawn in [lo, hi] (position), yaw uniform, pitch in [-0.25, 0.1] rad.
// out_cam: eye(3), forward(3), right(3), up(3).
int sg_camera_pose(uint64_t seed, const float* lo, const float* hi, float* out_cam) {
  Rng g(sub_seed(seed, 5));
  V3 e{g.range(lo[0], hi[0]), g.range(lo[1], hi[1]), g.range(lo[2], hi[2])};
  double yaw = g.range(0, 2 * 3.14159265358979323846), pitch = g.range(-0.25, 0.1);
  V3 f{std::cos(pitch) * std::sin(yaw), std::sin(pitch), std::cos(pitch) * std::cos(yaw)};
  V3 r = normalize(cross(f, V3{0, 1, 0}));
  V3 u = cross(r, f);
  V3 all[4] = {e, f, r, u};
  for (int i = 0; i < 4; i++) {
    out_cam[i * 3 + 0] = (float)all[i].x;
    out_cam[i * 3 + 1] = (float)all[i].y;
    out_cam[i * 3 + 2] = (float)all[i].z;
  }
  return 0;
}

// Ray-cast G-buffer from a pinhole camera (cam as written by sg_camera_pose), vertical
// field of view vfov_deg. Per pixel: hit position, unit face normal turned toward the
// camera, and the hit triangle's roughness (0 with a zero normal where nothing is hit).
int sg_gbuffer_camera(const float* pos, const float* tri_rough, uint32_t ntri, const float* cam,
                      uint32_t W, uint32_t H, float vfov_deg, float* out_pos, float* out_nrm,
                      float* out_rough) {
  Accel acc;
  acc.build(pos, ntri, 160);
  V3 e{cam[0], cam[1], cam[2]}, f{cam[3], cam[4], cam[5]}, r{cam[6], cam[7], cam[8]},
      u{cam[9], cam[10], cam[11]};
  double th = std::tan(vfov_deg * 3.14159265358979323846 / 360.0);
  double aspect = (double)W / (double)H;
#pragma omp parallel for schedule(dynamic, 4)
  for (int64_t y = 0; y < (int64_t)H; y++)
    for (uint32_t x = 0; x < W; x++) {
      size_t p = (size_t)y * W + x;
      double sx = (2.0 * (x + 0.5) / W - 1.0) * th * aspect;
      double sy = (1.0 - 2.0 * (y + 0.5) / H) * th;
      V3 d = normalize(f + r * sx + u * sy);
      double t = 0;
      int64_t id = acc.cast(e, d, &t);
      if (id < 0) {
        for (int a = 0; a < 3; a++) { out_pos[p * 3 + a] = 0; out_nrm[p * 3 + a] = 0; }
        if (out_rough) out_rough[p] = 0;
        continue;
      }
      const float* q = pos + (size_t)id * 9;
      V3 a{q[0], q[1], q[2]}, b{q[3], q[4], q[5]}, c{q[6], q[7], q[8]};
      V3 n = normalize(cross(b - a, c - a));
      if (dot(n, d) > 0) n = n * -1.0;
      V3 P = e + d * t;
      out_pos[p * 3 + 0] = (float)P.x;
      out_pos[p * 3 + 1] = (float)P.y;
      out_pos[p * 3 + 2] = (float)P.z;
      out_nrm[p * 3 + 0] = (float)n.x;
      out_nrm[p * 3 + 1] = (float)n.y;
      out_nrm[p * 3 + 2] = (float)n.z;
      if (out_rough) out_rough[p] = tri_rough ? tri_rough[id] : 0.0f;
    }
  return 0;
}

// count point lights, positions uniform in [lo, hi], white, intensity uniform in
// [imin, imax]. out: count x (x, y, z, intensity).
int sg_point_lights(uint64_t seed, uint32_t count, const float* lo, const float* hi, float imin,
                    float imax, float* out) {
  Rng g(sub_seed(seed, 6));
  for (uint32_t i = 0; i < count; i++) {
    out[i * 4 + 0] = (float)g.range(lo[0], hi[0]);
    out[i * 4 + 1] = (float)g.range(lo[1], hi[1]);
    out[i * 4 + 2] = (float)g.range(lo[2], hi[2]);
    out[i * 4 + 3] = (float)g.range(imin, imax);
  }
  return 0;
}

// Lognormal integer draws (median m, sigma s) clipped to [lo, hi]; Box-Muller.
int sg_lognormal_ints(uint64_t seed, uint32_t count, double median, double sigma, int64_t lo,
                      int64_t hi, int64_t* out) {
  Rng g(sub_seed(seed, 7));
  for (uint32_t i = 0; i < count; i++) {
    double u1 = g.uniform(), u2 = g.uniform();
    if (u1 < 1e-300) u1 = 1e-300;
    double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(2 * 3.14159265358979323846 * u2);
    double v = median * std::exp(sigma * z);
    int64_t k = (int64_t)std::llround(v);
    out[i] = std::max(lo, std::min(hi, k));
  }
  return 0;
}

}  // extern "C"
