// gltf.hpp — glTF 2.0 ingestion for the voxelizer (host C++; no GPU code).
//
// Reads the subset of glTF the voxel path needs: the default scene's node tree (matrix
// or translation/rotation/scale, composed parent * local), each mesh primitive's
// POSITION, TEXCOORD_0 and index accessors, its topology (triangle lists; strips and fans
// assembled into lists the way Vulkan assembles them), its material's base colour and
// base colour texture, and the images as stored (decoding them is the caller's).
// .gltf (JSON with external or base64 data-URI buffers) and .glb (binary container)
// are both read. Replaces the mesh and node part of CglTFSceneImporter
// (Foreground/SceneGraph/glTFSceneImporter.cpp:210-326 GetMesh, :333-372 Visit,
// :387-425 ImportFile), which parses with tinygltf.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

namespace vgi {

struct GltfGeometry {          // one glTF mesh primitive's triangles (shared by instances)
  std::vector<float> positions;    // xyz per vertex, object space
  std::vector<float> texcoords;    // uv per vertex (TEXCOORD_0), empty when absent
  std::vector<uint32_t> indices;   // 3 per triangle
  int32_t mesh = -1, primitive = -1;
};

struct GltfInstance {          // one primitive placed by one node
  uint32_t geometry = 0;       // index into GltfScene::geometry
  int32_t node = -1, material = -1;
  float model[12];             // row-major 3x4 world transform
  uint8_t base_color[4];
  int32_t image = -1;          // image of the material's baseColorTexture, -1: none
  float factor[4];             // RGBA multiplier of that texture (the albedo mode's base colour)
};

struct GltfImage {             // an image as the file stores it
  std::string mime_type;       // "" when the file does not say
  std::string uri;             // relative file name, "" when embedded
  std::vector<uint8_t> data;   // encoded bytes (embedded ones at parse time, files on request)
  bool loaded = false;
};

struct GltfScene {
  std::vector<GltfGeometry> geometry;
  std::vector<GltfInstance> instances;
  std::vector<GltfImage> images;
  std::string base_dir;             // where relative image files are read from
  uint32_t skipped_primitives = 0;  // points / lines (the voxelizer draws triangles only)
};

// Reads an external image's bytes (relative to the scene's base_dir) into img.data.
int gltf_load_image(const GltfScene& scene, GltfImage& img, std::string& err);

// albedo_mode: 0 = as the reference's importer (its material lookup of "baseColor" never
// matches glTF's "baseColorFactor", so every material is white), 1 = baseColorFactor.
// base_dir: directory that relative buffer URIs are resolved against ("" = none allowed).
// Returns 0 or a negative VGI_E_* code with a message in err.
int gltf_parse(const uint8_t* data, size_t bytes, const std::string& base_dir, uint32_t albedo_mode,
               GltfScene& out, std::string& err);

}  // namespace vgi
