// vgi_foreground.hpp — header-only C++ adapter that keeps the reference's pass interface
// over the vgi C ABI, so a renderer written against tobyc11/Final184's Foreground layer
// can swap its voxel GI passes for the MI355X path.
//
//   Foreground::CVoxelizeRenderer   <- Foreground/Renderer/VoxelizeRenderer.h:14-27
//       SetRenderPass, RenderList(context, modelMats, primitives),
//       PreparePrimitiveResources, ClearResourceCache, GarbageCollectResourceCache
//   Foreground::CIndirectPass       <- the lighting_indirect CMaterial pass,
//       Foreground/Renderer/MegaPipeline.cpp:252-268 (Shader/Lighting/indirect.frag),
//       and with RenderFrame the chain around it: the G-buffer decode, the temporal
//       blend and the indirect_blurX / indirect_blurY passes (MegaPipeline.cpp:196-284)
//   Foreground::CTriangleMesh / CBasicMaterial / CPrimitive
//                                   <- Foreground/Shape/TriangleMesh.h:26-103,
//                                      Foreground/Material/BasicMaterial.h,
//                                      Foreground/SceneGraph/Primitive.h:12-25
//   tc::Matrix3x4                   <- Math/Matrix3x4.h (row-major 3x4, m00..m23)
//
// RHI is an empty submodule in the reference, so buffers here are plain device pointers
// and the render context is a HIP stream. Error behaviour follows the reference: a pass
// that cannot run returns without drawing (VoxelizeRenderer.cpp:49-50,87-88); the status
// of the last call is kept in LastStatus() / LastError() instead of being lost.
#pragma once

#include <map>
#include <memory>
#include <string>
#include <vector>

#include "vgi.h"

namespace tc {
// Row-major 3x4 affine matrix, the layout of the reference's tc::Matrix3x4
// (m00 m01 m02 m03 / m10 ... / m20 ... m23).
struct Matrix3x4 {
  float m[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
  static Matrix3x4 Identity() { return Matrix3x4{}; }
};
}  // namespace tc

namespace Foreground {

// Stand-in for RHI::IRenderContext: the stream every pass records onto.
struct CRenderContext {
  void* Stream = nullptr;  // hipStream_t; NULL = the vgi context's stream
};

// Stand-in for RHI::CRenderPass::Ref (the voxelize pass owns no render target here).
struct CRenderPass {};

class CTriangleMesh {
public:
  // Vertex positions (float32 x3, device memory) with the byte stride of the binding.
  void SetPositions(const float* device_positions, uint32_t vertex_count, uint32_t stride = 12) {
    Positions = device_positions;
    VertexCount = vertex_count;
    Stride = stride;
  }
  // Index buffer in device memory: VGI_INDEX_U16 or VGI_INDEX_U32 (the reference's
  // SetIndexBuffer(buffer, format, offset), TriangleMesh.h:57-62); NULL for non-indexed.
  void SetIndexBuffer(const void* device_indices, vgi_index_type format) {
    Indices = device_indices;
    IndexFormat = device_indices ? format : VGI_INDEX_NONE;
  }
  // Number of indices (or vertices when non-indexed) drawn, as DrawIndexed(ElementCount).
  void SetElementCount(uint32_t count) { ElementCount = count; }
  uint32_t GetElementCount() const { return ElementCount; }
  // TEXCOORD_0 (float32 x2 per vertex, device memory) with the byte stride of its binding
  // (the importer's TexCoord0 attribute, glTFSceneImporter.cpp:262-263).
  void SetTexCoords(const float* device_uv, uint32_t stride = 8) {
    TexCoords = device_uv;
    TexCoordStride = stride;
  }

  const float* Positions = nullptr;
  const float* TexCoords = nullptr;
  uint32_t TexCoordStride = 8;
  uint32_t VertexCount = 0;
  uint32_t Stride = 12;
  const void* Indices = nullptr;
  vgi_index_type IndexFormat = VGI_INDEX_NONE;
  uint32_t ElementCount = 0;
};

class CBasicMaterial {
public:
  // BaseColorFactor (main.lua:169-177); RGB in [0, 1], stored as RGB8 for untextured
  // draws and kept as floats to multiply a texture sample.
  void SetAlbedo(float r, float g, float b, float a = 1.0f) {
    auto q = [](float v) { return (uint8_t)(v <= 0.f ? 0 : v >= 1.f ? 255 : (int)(v * 255.f + 0.5f)); };
    Albedo[0] = q(r);
    Albedo[1] = q(g);
    Albedo[2] = q(b);
    Factor[0] = r;
    Factor[1] = g;
    Factor[2] = b;
    Factor[3] = a;
  }
  // Optional per-triangle RGB8 albedo in device memory (replaces the texture lookup).
  void SetAlbedoPerTriangle(const uint8_t* device_rgb8) { PerTriangle = device_rgb8; }
  // The base colour texture (SetAlbedoImage, BasicMaterial.h; UseTextures when set,
  // BasicMaterial.cpp:26): RGBA8 texels in device memory, row-major width x height. A
  // primitive whose shape has texcoords then voxelizes with texture(uv) x the albedo
  // factor and the alpha < 0.05 discard (main.lua:188-205).
  void SetAlbedoImage(const uint8_t* device_rgba8, uint32_t width, uint32_t height) {
    Image = device_rgba8;
    ImageWidth = width;
    ImageHeight = height;
  }

  uint8_t Albedo[3] = {255, 255, 255};
  float Factor[4] = {1.f, 1.f, 1.f, 1.f};
  const uint8_t* PerTriangle = nullptr;
  const uint8_t* Image = nullptr;
  uint32_t ImageWidth = 0, ImageHeight = 0;
};

class CPrimitive : public std::enable_shared_from_this<CPrimitive> {
public:
  std::shared_ptr<CBasicMaterial> GetMaterial() const { return Material; }
  void SetMaterial(std::shared_ptr<CBasicMaterial> material) { Material = std::move(material); }
  std::shared_ptr<CTriangleMesh> GetShape() const { return Shape; }
  void SetShape(std::shared_ptr<CTriangleMesh> shape) { Shape = std::move(shape); }

private:
  std::shared_ptr<CBasicMaterial> Material;
  std::shared_ptr<CTriangleMesh> Shape;
};

class CVoxelizeRenderer {
public:
  // The reference takes the owning CMegaPipeline; here the owner is a vgi context and
  // the scene that holds the voxel volume, plus the voxel grid (the voxelizer view).
  CVoxelizeRenderer(vgi_context* ctx, vgi_scene* scene, const vgi_grid& grid)
      : Ctx(ctx), Scene(scene), Grid(grid) {}

  void SetGrid(const vgi_grid& grid) { Grid = grid; }
  // The voxelizer view's matrices as the reference binds them (VoxelView, VoxelProjection:
  // GLSL column-major, MegaPipeline.cpp:241-246): the grid becomes the cube holding that
  // orthographic volume, n voxels per axis (the reference's image is 128^3). Returns false
  // (and keeps the previous grid) for a camera that is not orthographic and axis-aligned.
  bool SetVoxelCamera(const float voxelView[16], const float voxelProj[16], uint32_t n = 128,
                      uint32_t levels = 7) {
    vgi_grid g;
    Status = vgi_grid_from_voxel_camera(voxelView, voxelProj, n, levels, &g);
    if (Status != VGI_OK) {
      Error = vgi_last_error();
      return false;
    }
    Grid = g;
    return true;
  }
  const vgi_grid& GetGrid() const { return Grid; }

  // Kept for interface parity; voxelization needs no render pass here.
  void SetRenderPass(const CRenderPass& /*renderPass*/, uint32_t /*subpass*/ = 0) {}

  // One voxelize pass over the visible primitives with their model matrices
  // (VoxelizeRenderer.cpp:19-30): clears the volume and voxelizes every primitive in one
  // batched launch on the context's stream.
  void RenderList(CRenderContext& context, const std::vector<tc::Matrix3x4>& modelMats,
                  const std::vector<CPrimitive*>& primitives) {
    GarbageCollectResourceCache();
    Meshes.clear();
    const size_t count = std::min(modelMats.size(), primitives.size());
    for (size_t i = 0; i < count; i++) {
      CPrimitive* prim = primitives[i];
      if (!prim) continue;
      auto it = Cache.find(prim->weak_from_this());
      if (it == Cache.end()) {
        PreparePrimitiveResources(prim->shared_from_this());
        it = Cache.find(prim->weak_from_this());
        if (it == Cache.end()) continue;  // not drawable: skipped, as the reference does
      }
      vgi_mesh m = it->second;
      for (int k = 0; k < 12; k++) m.model[k] = modelMats[i].m[k];
      Meshes.push_back(m);
    }
    Status = vgi_voxelize(Ctx, Scene, Meshes.data(), (uint32_t)Meshes.size(), &Grid, context.Stream);
    if (Status != VGI_OK) Error = vgi_last_error();
  }

  // Builds the cached draw description of a primitive (VoxelizeRenderer.cpp:32-67).
  void PreparePrimitiveResources(std::shared_ptr<CPrimitive> primitive) {
    if (!primitive) return;
    auto shape = primitive->GetShape();
    auto mat = primitive->GetMaterial();
    if (!shape || !shape->Positions) return;
    vgi_mesh m{};
    m.positions = shape->Positions;
    m.position_stride = shape->Stride;
    m.num_vertices = shape->VertexCount;
    m.indices = shape->Indices;
    m.index_type = shape->IndexFormat;
    m.num_triangles = shape->ElementCount / 3;
    if (mat) {
      m.albedo_rgb8 = mat->PerTriangle;
      m.base_color[0] = mat->Albedo[0];
      m.base_color[1] = mat->Albedo[1];
      m.base_color[2] = mat->Albedo[2];
      if (mat->Image && shape->TexCoords) {
        m.texcoords = shape->TexCoords;
        m.texcoord_stride = shape->TexCoordStride;
        m.texture_rgba8 = mat->Image;
        m.texture_width = mat->ImageWidth;
        m.texture_height = mat->ImageHeight;
        for (int k = 0; k < 4; k++) m.base_color_factor[k] = mat->Factor[k];
      }
    } else {
      m.base_color[0] = m.base_color[1] = m.base_color[2] = 255;
    }
    Cache[primitive] = m;
  }

  void ClearResourceCache() { Cache.clear(); }

  // Drops entries of primitives that no longer exist (erase returns the next iterator).
  void GarbageCollectResourceCache() {
    for (auto it = Cache.begin(); it != Cache.end();) {
      if (it->first.expired()) it = Cache.erase(it);
      else ++it;
    }
  }

  size_t CachedPrimitives() const { return Cache.size(); }
  int LastStatus() const { return Status; }
  const std::string& LastError() const { return Error; }

private:
  vgi_context* Ctx;
  vgi_scene* Scene;
  vgi_grid Grid;
  std::map<std::weak_ptr<CPrimitive>, vgi_mesh, std::owner_less<std::weak_ptr<CPrimitive>>> Cache;
  std::vector<vgi_mesh> Meshes;
  int Status = VGI_OK;
  std::string Error;
};

// The G-buffer targets the indirect pass samples (MegaPipeline.cpp:254-255): GBufferDepth
// (D32) and GBuffer1 (RGBA16 UNORM view-space normals), device memory, width x height.
struct CGBufferTargets {
  const float* Depth = nullptr;
  const uint16_t* Normals = nullptr;
  const float* Roughness = nullptr;  // optional per-pixel specular roughness
  uint32_t Width = 0;
  uint32_t Height = 0;
};

// The matrices of one frame as the reference uploads them, GLSL column-major: InvProj
// (SceneView.cpp:30), InvModelView (MegaPipeline.cpp:242) and last frame's ProjMat and
// ViewMat (the prevProj block, MegaPipeline.cpp:70-71,321-322).
struct CIndirectFrameConstants {
  float InvProj[16];
  float InvModelView[16];
  float PrevProjection[16];
  float PrevModelView[16];
};

// The lighting_indirect pass: direct light injected into the volume, the mip chain, and
// the cone trace of every G-buffer pixel into the indirect buffer.
class CIndirectPass {
public:
  CIndirectPass(vgi_context* ctx, vgi_scene* scene) : Ctx(ctx), Scene(scene) {}
  CIndirectPass(const CIndirectPass&) = delete;
  CIndirectPass& operator=(const CIndirectPass&) = delete;
  ~CIndirectPass() { FreeImages(); }

  void SetLights(const vgi_lights& lights) { Lights = lights; }
  void SetCones(const vgi_cone_params& cones) { Cones = cones; }
  void SetPrecision(vgi_precision p) { Precision = p; }

  // out_diffuse / out_specular: float4 per pixel, device memory (specular may be NULL).
  void Render(CRenderContext& context, const vgi_gbuffer& gbuffer, float* out_diffuse,
              float* out_specular) {
    Status = vgi_build_mips(Ctx, Scene, &Lights, (uint32_t)Precision, context.Stream);
    if (Status == VGI_OK)
      Status = vgi_trace(Ctx, Scene, &gbuffer, &Cones, out_diffuse, out_specular, context.Stream);
    if (Status != VGI_OK) Error = vgi_last_error();
  }

  // The reference's whole indirect chain for one frame (MegaPipeline.cpp:196-284): the
  // targets are decoded to world position and normal (indirect.frag:44-58,189-196), the
  // cones traced, the result blended with the previous frame's (indirect.frag:224-242),
  // then blurred along y (indirect_blurX) and along x (indirect_blurY). The history is
  // this pass's previous blended output, which the reference copies into
  // indirectTemporalImage at the start of a frame; here two images alternate instead.
  // Images are RGBA16F like the reference's render targets, owned by the pass and
  // re-created (history cleared, as on the reference's first frame) when the target
  // size changes. Returns the final image (device, RGBA16F) or NULL on error.
  const uint16_t* RenderFrame(CRenderContext& context, const CGBufferTargets& targets,
                              const CIndirectFrameConstants& m) {
    Status = EnsureImages(targets.Width, targets.Height);
    if (Status == VGI_OK) {
      vgi_gbuffer_ref ref{targets.Depth, targets.Normals, {}, {}, targets.Width, targets.Height};
      for (int i = 0; i < 16; i++) {
        ref.inv_proj[i] = m.InvProj[i];
        ref.inv_model_view[i] = m.InvModelView[i];
      }
      Status = vgi_gbuffer_decode(Ctx, &ref, Position, Normal, ViewDepth, context.Stream);
    }
    if (Status == VGI_OK) {
      const vgi_gbuffer gb{Position, Normal, targets.Roughness, Width, Height};
      Render(context, gb, Diffuse, nullptr);
    }
    if (Status == VGI_OK) {
      vgi_temporal_params t{};
      t.indirect = Diffuse;
      t.position = Position;
      t.view_depth = ViewDepth;
      t.history = Indirect[Frame & 1];
      t.out = Indirect[(Frame + 1) & 1];
      t.history_format = VGI_IMAGE_RGBA16F;
      t.out_format = VGI_IMAGE_RGBA16F;
      for (int i = 0; i < 16; i++) {
        t.prev_projection[i] = m.PrevProjection[i];
        t.prev_model_view[i] = m.PrevModelView[i];
      }
      t.width = Width;
      t.height = Height;
      Status = vgi_temporal_resolve(Ctx, &t, context.Stream);
    }
    if (Status == VGI_OK) {
      const vgi_blur_params bx{Indirect[(Frame + 1) & 1], targets.Depth, Blur, VGI_IMAGE_RGBA16F,
                               VGI_IMAGE_RGBA16F, Width, Height, 1};
      Status = vgi_bilateral_blur(Ctx, &bx, context.Stream);
    }
    if (Status == VGI_OK) {
      const vgi_blur_params by{Blur, targets.Depth, Final, VGI_IMAGE_RGBA16F, VGI_IMAGE_RGBA16F,
                               Width, Height, 0};
      Status = vgi_bilateral_blur(Ctx, &by, context.Stream);
    }
    if (Status != VGI_OK) {
      Error = vgi_last_error();
      return nullptr;
    }
    Frame++;
    return Final;
  }

  // The reference's first-frame clear of the history (MegaPipeline.cpp:197-204).
  void ResetHistory() {
    if (Indirect[0]) Status = vgi_memset(Ctx, Indirect[Frame & 1], 0, (size_t)Width * Height * 8);
  }

  // Intermediate images of the last RenderFrame (device memory): world position and normal
  // (float x3), view depth, the traced irradiance (float4), the blended result (RGBA16F,
  // the next frame's history) and the image blurred along y (RGBA16F).
  const float* PositionImage() const { return Position; }
  const float* NormalImage() const { return Normal; }
  const float* ViewDepthImage() const { return ViewDepth; }
  const float* DiffuseImage() const { return Diffuse; }
  const uint16_t* TemporalImage() const { return Indirect[Frame & 1]; }
  const uint16_t* BlurImage() const { return Blur; }

  int LastStatus() const { return Status; }
  const std::string& LastError() const { return Error; }

private:
  int EnsureImages(uint32_t w, uint32_t h) {
    if (!Ctx) return VGI_E_INVALID;
    if (w == Width && h == Height && Final) return VGI_OK;
    FreeImages();
    const size_t px = (size_t)w * h;
    const size_t bytes[8] = {px * 12, px * 12, px * 4, px * 16, px * 8, px * 8, px * 8, px * 8};
    void** dst[8] = {(void**)&Position, (void**)&Normal, (void**)&ViewDepth, (void**)&Diffuse,
                     (void**)&Indirect[0], (void**)&Indirect[1], (void**)&Blur, (void**)&Final};
    for (int i = 0; i < 8; i++) {
      const int rc = vgi_malloc(Ctx, dst[i], bytes[i] ? bytes[i] : 1);
      if (rc != VGI_OK) {
        FreeImages();
        return rc;
      }
    }
    Width = w;
    Height = h;
    Frame = 0;
    const int rc = vgi_memset(Ctx, Indirect[0], 0, px * 8);
    return rc != VGI_OK ? rc : vgi_memset(Ctx, Indirect[1], 0, px * 8);
  }

  void FreeImages() {
    void* all[8] = {Position, Normal, ViewDepth, Diffuse, Indirect[0], Indirect[1], Blur, Final};
    for (void* p : all)
      if (p && Ctx) vgi_free(Ctx, p);
    Position = Normal = ViewDepth = Diffuse = nullptr;
    Indirect[0] = Indirect[1] = Blur = Final = nullptr;
    Width = Height = 0;
  }

  vgi_context* Ctx;
  vgi_scene* Scene;
  vgi_lights Lights{};
  vgi_cone_params Cones{6, 1, 60.f, 60.f, 10.f, 0.5f, 1.f, 0.99f, 1.f, 1.f, {0.f, 0.f, 0.f}};
  vgi_precision Precision = VGI_PREC_FP16;
  int Status = VGI_OK;
  std::string Error;
  uint32_t Width = 0, Height = 0;
  uint64_t Frame = 0;
  float* Position = nullptr;
  float* Normal = nullptr;
  float* ViewDepth = nullptr;
  float* Diffuse = nullptr;
  uint16_t* Indirect[2] = {nullptr, nullptr};
  uint16_t* Blur = nullptr;
  uint16_t* Final = nullptr;
};

}  // namespace Foreground
