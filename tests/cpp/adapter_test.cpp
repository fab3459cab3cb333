// Drives include/vgi_foreground.hpp the way Foreground/Renderer/MegaPipeline.cpp drives its
// passes (RenderList with per-primitive model matrices, then the indirect pass).
//   adapter_test nodevice : cache and error behaviour without a GPU
//   adapter_test gpu      : two primitives voxelized through the adapter == the oracle, then
//                           two RenderFrame calls whose every stage == the oracle chain
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "vgi_foreground.hpp"

extern "C" {
struct orc_vox_stats { uint64_t a, b, c, d; };
int orc_voxelize(const float*, uint32_t, const uint32_t*, uint32_t, const uint8_t*, const uint8_t*,
                 const float*, const float*, float, uint32_t, uint32_t*, orc_vox_stats*);
int orc_voxelize_textured(const float*, uint32_t, const uint32_t*, uint32_t, const float*, const uint8_t*, uint32_t,
                          uint32_t, const float*, const float*, const float*, float, uint32_t, uint32_t*,
                          orc_vox_stats*);
void orc_gbuffer_decode_f32(const float*, const uint16_t*, const float*, const float*, uint32_t, uint32_t,
                            float*, float*, float*);
void orc_temporal(const float*, const float*, const float*, const float*, const float*, const float*, uint32_t,
                  uint32_t, float*);
void orc_bilateral_blur(const float*, const float*, uint32_t, uint32_t, uint32_t, float*);
uint16_t orc_float_to_half(float);
float orc_half_to_float(uint16_t);
}

#define CHECK(c)                                                      \
  do {                                                                \
    if (!(c)) {                                                       \
      std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #c);         \
      return 1;                                                       \
    }                                                                 \
  } while (0)

using namespace Foreground;

static int nodevice() {
  vgi_grid g{{0, 0, 0}, 1.0f, 32, 5};
  CVoxelizeRenderer vr(nullptr, nullptr, g);
  vr.SetRenderPass(CRenderPass{});
  auto mesh = std::make_shared<CTriangleMesh>();
  static float fake[9] = {0};
  mesh->SetPositions(fake, 3);
  mesh->SetElementCount(3);
  auto prim = std::make_shared<CPrimitive>();
  prim->SetShape(mesh);
  prim->SetMaterial(std::make_shared<CBasicMaterial>());
  {
    auto dead = std::make_shared<CPrimitive>();
    dead->SetShape(mesh);
    vr.PreparePrimitiveResources(dead);
    vr.PreparePrimitiveResources(prim);
    CHECK(vr.CachedPrimitives() == 2);
  }
  vr.GarbageCollectResourceCache();  // the expired primitive goes, the live one stays
  CHECK(vr.CachedPrimitives() == 1);
  CRenderContext ctx;
  vr.RenderList(ctx, {tc::Matrix3x4::Identity()}, {prim.get()});
  CHECK(vr.LastStatus() == VGI_E_INVALID);
  CHECK(!vr.LastError().empty());
  vr.ClearResourceCache();
  CHECK(vr.CachedPrimitives() == 0);
  CIndirectPass ip(nullptr, nullptr);
  vgi_gbuffer gb{};
  ip.Render(ctx, gb, nullptr, nullptr);
  CHECK(ip.LastStatus() == VGI_E_INVALID);
  std::printf("ADAPTER NODEVICE OK\n");
  return 0;
}

static void* dev_copy(vgi_context* c, const void* src, size_t bytes) {
  void* p = nullptr;
  if (vgi_malloc(c, &p, bytes) != VGI_OK) return nullptr;
  vgi_memcpy_h2d(c, p, src, bytes);
  return p;
}

static std::vector<float> widen(const std::vector<uint16_t>& h) {
  std::vector<float> f(h.size());
  for (size_t i = 0; i < h.size(); i++) f[i] = orc_half_to_float(h[i]);
  return f;
}

static std::vector<uint16_t> narrow(const std::vector<float>& f) {
  std::vector<uint16_t> h(f.size());
  for (size_t i = 0; i < f.size(); i++) h[i] = orc_float_to_half(f[i]);
  return h;
}

template <class T>
static std::vector<T> readback(vgi_context* c, const T* d, size_t count) {
  std::vector<T> h(count);
  vgi_memcpy_d2h(c, h.data(), d, count * sizeof(T));
  return h;
}

// Two frames of CIndirectPass::RenderFrame on a camera in front of the voxelized
// primitives (the reference's projection form, Camera.cpp:85-101, and a view that only
// translates). Every stage must equal the oracle on that stage's inputs, and the second
// frame must take the first frame's blended image as its history.
static int frames(vgi_context* c, vgi_scene* s, const vgi_lights& L) {
  const uint32_t W = 48, H = 32;
  const size_t px = (size_t)W * H;
  // depth 0.87 .. 0.90 puts the surface at view z -1.35 .. -1.68, inside the grid; normals
  // face the camera with some noise
  std::mt19937 rng(11);
  std::uniform_real_distribution<float> N(-0.0005f, 0.0005f);
  std::uniform_int_distribution<int> Q(-6000, 6000);
  std::vector<float> depth(px);
  std::vector<uint16_t> nrm(px * 4);
  for (size_t i = 0; i < px; i++) {
    depth[i] = 0.87f + 0.03f * (float)(i % W) / W + N(rng);
    nrm[i * 4 + 0] = (uint16_t)(32768 + Q(rng));
    nrm[i * 4 + 1] = (uint16_t)(32768 + Q(rng));
    nrm[i * 4 + 2] = 65535;
    nrm[i * 4 + 3] = 65535;
  }
  // P = [[a,0,0,0],[0,-b,0,0],[0,0,c,d],[0,0,-1,0]]; its inverse, column-major
  const float yslope = std::tan(0.5f * 60.f * 3.14159265f / 180.f), aspect = (float)W / H;
  const float n = 0.1f, f = 10.f, a = 1.f / (yslope * aspect), b = 1.f / yslope;
  const float cc = (f + n) / (n - f), dd = -2.f * f * n / (f - n);
  CIndirectFrameConstants m{};
  m.InvProj[0] = 1.f / a;
  m.InvProj[5] = -1.f / b;
  m.InvProj[11] = 1.f / dd;  // row 3, column 2
  m.InvProj[14] = -1.f;      // row 2, column 3
  m.InvProj[15] = cc / dd;
  const float eye[3] = {0.f, 0.f, 1.5f};
  for (int i = 0; i < 16; i++) m.InvModelView[i] = (i % 5 == 0) ? 1.f : 0.f;
  for (int k = 0; k < 3; k++) m.InvModelView[12 + k] = eye[k];
  // last frame's camera: the same projection, the eye 0.02 to the left
  for (int i = 0; i < 16; i++) m.PrevProjection[i] = m.PrevModelView[i] = 0.f;
  m.PrevProjection[0] = a;
  m.PrevProjection[5] = -b;
  m.PrevProjection[10] = cc;
  m.PrevProjection[11] = -1.f;
  m.PrevProjection[14] = dd;
  for (int i = 0; i < 4; i++) m.PrevModelView[i * 5] = 1.f;
  m.PrevModelView[12] = -(eye[0] - 0.02f);
  m.PrevModelView[13] = -eye[1];
  m.PrevModelView[14] = -eye[2];

  CGBufferTargets t;
  t.Depth = (const float*)dev_copy(c, depth.data(), px * 4);
  t.Normals = (const uint16_t*)dev_copy(c, nrm.data(), px * 8);
  t.Width = W;
  t.Height = H;
  CIndirectPass ip(c, s);
  ip.SetLights(L);
  vgi_cone_params cones{6, 0, 60.f, 60.f, 10.f, 0.5f, 0.8f, 0.99f, 1.f, 1.f, {eye[0], eye[1], eye[2]}};
  ip.SetCones(cones);
  CRenderContext ctx;
  std::vector<float> history(px * 4, 0.f);  // the first frame's history is cleared
  for (int frame = 0; frame < 2; frame++) {
    const uint16_t* fin = ip.RenderFrame(ctx, t, m);
    CHECK(fin != nullptr && ip.LastStatus() == VGI_OK);
    CHECK(vgi_sync(c, nullptr) == VGI_OK);
    std::vector<float> rp(px * 3), rn(px * 3), rv(px);
    orc_gbuffer_decode_f32(depth.data(), nrm.data(), m.InvProj, m.InvModelView, W, H, rp.data(), rn.data(),
                           rv.data());
    CHECK(readback(c, ip.PositionImage(), px * 3) == rp);
    CHECK(readback(c, ip.NormalImage(), px * 3) == rn);
    CHECK(readback(c, ip.ViewDepthImage(), px) == rv);
    const std::vector<float> diffuse = readback(c, ip.DiffuseImage(), px * 4);
    std::vector<float> taa(px * 4);
    orc_temporal(diffuse.data(), rp.data(), rv.data(), history.data(), m.PrevProjection, m.PrevModelView, W,
                 H, taa.data());
    const std::vector<uint16_t> taa16 = narrow(taa);
    CHECK(readback(c, ip.TemporalImage(), px * 4) == taa16);
    std::vector<float> bx(px * 4), by(px * 4);
    orc_bilateral_blur(widen(taa16).data(), depth.data(), W, H, 1, bx.data());
    const std::vector<uint16_t> bx16 = narrow(bx);
    CHECK(readback(c, ip.BlurImage(), px * 4) == bx16);
    orc_bilateral_blur(widen(bx16).data(), depth.data(), W, H, 0, by.data());
    CHECK(readback(c, fin, px * 4) == narrow(by));
    size_t blended = 0;
    for (size_t i = 0; i < px; i++) blended += taa[i * 4] != diffuse[i * 4];
    std::printf("frame %d: %zu of %zu pixels blended with the history\n", frame, blended, px);
    CHECK(frame == 0 || blended > px / 2);
    history = widen(taa16);
  }
  vgi_free(c, (void*)t.Depth);
  vgi_free(c, (void*)t.Normals);
  return 0;
}

static int gpu() {
  vgi_context* c = nullptr;
  CHECK(vgi_create(&c, 0) == VGI_OK);
  vgi_scene* s = nullptr;
  CHECK(vgi_scene_create(c, &s) == VGI_OK);
  const uint32_t n = 32;
  vgi_grid g{{-1.f, -1.f, -1.f}, 2.0f, n, 5};
  std::mt19937 rng(7);
  std::uniform_real_distribution<float> U(-0.6f, 0.6f);
  // primitive A: 200 shared vertices, 300 u16-indexed triangles, constant albedo
  std::vector<float> va(600);
  for (auto& v : va) v = U(rng);
  std::vector<uint16_t> ia(900);
  for (auto& i : ia) i = (uint16_t)(rng() % 200);
  // primitive B: 150 non-indexed triangles with per-triangle albedo
  std::vector<float> vb(1350);
  for (size_t i = 0; i < vb.size(); i++) vb[i] = U(rng) * 0.5f + (i % 3 == 0 ? 0.2f : 0.f);
  std::vector<uint8_t> ab(450);
  for (auto& a : ab) a = (uint8_t)(rng() & 255);
  auto meshA = std::make_shared<CTriangleMesh>();
  meshA->SetPositions((const float*)dev_copy(c, va.data(), va.size() * 4), 200);
  meshA->SetIndexBuffer(dev_copy(c, ia.data(), ia.size() * 2), VGI_INDEX_U16);
  meshA->SetElementCount(900);
  auto matA = std::make_shared<CBasicMaterial>();
  matA->SetAlbedo(0.2f, 0.4f, 0.8f);
  auto meshB = std::make_shared<CTriangleMesh>();
  meshB->SetPositions((const float*)dev_copy(c, vb.data(), vb.size() * 4), 450);
  meshB->SetElementCount(450);
  auto matB = std::make_shared<CBasicMaterial>();
  matB->SetAlbedoPerTriangle((const uint8_t*)dev_copy(c, ab.data(), ab.size()));
  // primitive C: 100 non-indexed triangles with TEXCOORD_0 and a 24 x 20 RGBA8 base colour
  // texture (a third of its texels below the 0.05 alpha discard) times the albedo factor
  std::vector<float> vc(900), uvc(600);
  for (auto& v : vc) v = U(rng) * 0.8f;
  for (auto& v : uvc) v = U(rng) * 3.0f;
  const uint32_t TW = 24, TH = 20;
  std::vector<uint8_t> texc(TW * TH * 4);
  for (size_t i = 0; i < texc.size(); i++) texc[i] = (uint8_t)(rng() & 255);
  for (size_t i = 3; i < texc.size(); i += 4)
    if (rng() % 3 == 0) texc[i] = (uint8_t)(rng() % 20);
  auto meshC = std::make_shared<CTriangleMesh>();
  meshC->SetPositions((const float*)dev_copy(c, vc.data(), vc.size() * 4), 300);
  meshC->SetElementCount(300);
  meshC->SetTexCoords((const float*)dev_copy(c, uvc.data(), uvc.size() * 4));
  auto matC = std::make_shared<CBasicMaterial>();
  matC->SetAlbedo(0.9f, 0.8f, 1.0f, 1.0f);
  matC->SetAlbedoImage((const uint8_t*)dev_copy(c, texc.data(), texc.size()), TW, TH);
  auto pA = std::make_shared<CPrimitive>(), pB = std::make_shared<CPrimitive>();
  auto pC = std::make_shared<CPrimitive>();
  pA->SetShape(meshA);
  pA->SetMaterial(matA);
  pB->SetShape(meshB);
  pB->SetMaterial(matB);
  pC->SetShape(meshC);
  pC->SetMaterial(matC);
  tc::Matrix3x4 mA, mB;
  const float cs = std::cos(0.4f), sn = std::sin(0.4f);
  float ma[12] = {cs, -sn, 0, 0.1f, sn, cs, 0, 0, 0, 0, 1, -0.05f};
  std::memcpy(mA.m, ma, sizeof(ma));
  float mb[12] = {1.1f, 0, 0, 0, 0, 0.9f, 0, 0.1f, 0, 0, 1, 0};
  std::memcpy(mB.m, mb, sizeof(mb));

  CVoxelizeRenderer vr(c, s, g);
  CRenderContext ctx;
  vr.RenderList(ctx, {mA, mB, mB}, {pA.get(), pB.get(), pC.get()});
  CHECK(vr.LastStatus() == VGI_OK);
  std::vector<uint32_t> got((size_t)n * n * n * 8), ra(got.size()), rb(got.size()), rc(got.size(), 0u);
  CHECK(vgi_scene_read_voxels(c, s, got.data()) == VGI_OK);
  std::vector<uint32_t> ia32(ia.begin(), ia.end());
  const uint8_t baseA[3] = {matA->Albedo[0], matA->Albedo[1], matA->Albedo[2]};
  orc_vox_stats st;
  orc_voxelize(va.data(), 200, ia32.data(), 300, nullptr, baseA, ma, g.origin, g.extent, n, ra.data(), &st);
  orc_voxelize(vb.data(), 450, nullptr, 150, ab.data(), nullptr, mb, g.origin, g.extent, n, rb.data(), &st);
  orc_voxelize_textured(vc.data(), 300, nullptr, 100, uvc.data(), texc.data(), TW, TH, matC->Factor, mb, g.origin,
                        g.extent, n, rc.data(), &st);
  size_t bad = 0, occ = 0;
  for (size_t i = 0; i < got.size(); i++) {
    bad += got[i] != ra[i] + rb[i] + rc[i];
    if (i % 8 == 3 && got[i]) occ++;
  }
  std::printf("occupied voxels %zu, mismatched words %zu\n", occ, bad);
  CHECK(bad == 0 && occ > 100);

  // indirect pass on a tiny G-buffer
  const uint32_t W = 16, H = 16;
  std::vector<float> gp(W * H * 3), gn(W * H * 3);
  for (uint32_t i = 0; i < W * H; i++) {
    gp[i * 3] = U(rng) * 0.5f; gp[i * 3 + 1] = U(rng) * 0.5f; gp[i * 3 + 2] = U(rng) * 0.5f;
    gn[i * 3] = 0; gn[i * 3 + 1] = 0; gn[i * 3 + 2] = 1;
  }
  vgi_gbuffer gb{(const float*)dev_copy(c, gp.data(), gp.size() * 4),
                 (const float*)dev_copy(c, gn.data(), gn.size() * 4), nullptr, W, H};
  float* od = nullptr;
  CHECK(vgi_malloc(c, (void**)&od, W * H * 16) == VGI_OK);
  vgi_lights L{};
  L.count = 1;
  L.light[0] = vgi_light{VGI_LIGHT_DIRECTIONAL, {0.f, 0.f, -1.f}, {1.f, 1.f, 1.f}, 0.f};
  CIndirectPass ip(c, s);
  ip.SetLights(L);
  ip.Render(ctx, gb, od, nullptr);
  CHECK(ip.LastStatus() == VGI_OK);
  std::vector<float> out(W * H * 4);
  CHECK(vgi_memcpy_d2h(c, out.data(), od, out.size() * 4) == VGI_OK);
  float amin = 1e30f, amax = -1e30f;
  size_t nonfinite = 0;
  for (uint32_t i = 0; i < W * H; i++) {
    const float v = out[i * 4 + 3];
    if (!std::isfinite(v)) nonfinite++;
    else {
      amin = std::min(amin, v);
      amax = std::max(amax, v);
    }
  }
  std::printf("indirect alpha: min %.9g max %.9g non-finite %zu\n", amin, amax, nonfinite);
  // the diffuse alpha is a weighted mean of per-cone opacities in [0, 1] (weights rounded)
  CHECK(nonfinite == 0 && amin >= 0.f && amax <= 1.f + 1e-6f);
  CHECK(frames(c, s, L) == 0);
  vgi_scene_release(s);
  vgi_destroy(c);
  std::printf("ADAPTER GPU OK\n");
  return 0;
}

int main(int argc, char** argv) {
  if (argc > 1 && std::strcmp(argv[1], "gpu") == 0) return gpu();
  return nodevice();
}
