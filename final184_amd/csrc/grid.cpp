// grid.cpp — the reference's voxelizer camera as a vgi_grid (host only).
//
// The reference voxelizes through an orthographic camera: App/MainBehaviour.cpp:54-64 (a
// node rotated by pitch -90 degrees at y = 16, CCamera with MagX = MagY = 16, near 1, far
// 20), whose matrices CSceneView::PrepareToRender uploads as GLSL column-major
// (ViewMat = inverse(world)^T, ProjMat = P^T; Foreground/SceneGraph/SceneView.cpp:26-30)
// with P from CCamera::CalcOrtho (Foreground/SceneGraph/Camera.cpp:104-111): clip x = x /
// MagX, y = y / MagY, z = (z + n) / (n - f). Its shaders address the voxel image at
// ((clip.xy * 0.5 + 0.5), clip.z) * 128 (Shader/Lighting/indirect.frag:128-133), so the
// volume is the world-space image of the clip box [-1, 1]^2 x [0, 1] under
// inverse(VoxelProj * VoxelView).
//
// vgi grids are world-axis-aligned cubes of isotropic voxels. For a camera whose view
// rotation maps axes to axes (any multiple of 90 degrees about the world axes) that
// image is a world-axis-aligned box; the grid is the cube with the box's centre whose
// edge is the box's longest side. The reference's box is 32 x 19 x 32 units with 128
// voxels per axis (0.25 x 0.148 x 0.25 units per voxel); the grid holds it in a 32-unit
// cube, 0.25 units per voxel along every axis.
#include <cmath>
#include <cstdint>
#include <cstdio>

#include "../../include/vgi.h"

namespace vgi {
int set_error(int code, const char* fmt, ...);
}

namespace {

// GLSL column-major 4x4: element (row r, column c) is m[c * 4 + r].
inline double at(const float* m, int r, int c) { return (double)m[c * 4 + r]; }

}  // namespace

extern "C" int vgi_grid_from_voxel_camera(const float voxel_view[16], const float voxel_proj[16], uint32_t n,
                                          uint32_t levels, vgi_grid* out) {
  if (!voxel_view || !voxel_proj || !out) return vgi::set_error(VGI_E_INVALID, "NULL argument");
  for (int i = 0; i < 16; i++)
    if (!std::isfinite(voxel_view[i]) || !std::isfinite(voxel_proj[i]))
      return vgi::set_error(VGI_E_INVALID, "voxel camera matrices must be finite");
  // M = VoxelProj * VoxelView (world -> clip), in double
  double M[4][4];
  for (int r = 0; r < 4; r++)
    for (int c = 0; c < 4; c++) {
      double s = 0.0;
      for (int k = 0; k < 4; k++) s += at(voxel_proj, r, k) * at(voxel_view, k, c);
      M[r][c] = s;
    }
  // orthographic: the last row is (0, 0, 0, 1)
  double scale = 0.0;
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) scale = std::fmax(scale, std::fabs(M[r][c]));
  if (!(scale > 0.0)) return vgi::set_error(VGI_E_INVALID, "voxel camera maps the world to a point");
  const double tol = 1e-6 * scale;
  if (std::fabs(M[3][0]) > tol || std::fabs(M[3][1]) > tol || std::fabs(M[3][2]) > tol ||
      std::fabs(M[3][3] - 1.0) > 1e-6)
    return vgi::set_error(VGI_E_INVALID, "voxel camera projection is not orthographic");
  // axis to axis: every row and column of the linear part has exactly one nonzero entry
  int axis_of_row[3];
  for (int r = 0; r < 3; r++) {
    int nz = 0;
    for (int c = 0; c < 3; c++)
      if (std::fabs(M[r][c]) > tol) {
        axis_of_row[r] = c;
        nz++;
      }
    if (nz != 1)
      return vgi::set_error(VGI_E_INVALID,
                            "voxel camera view is not aligned with the world axes (row %d has %d nonzero entries)", r,
                            nz);
  }
  if (axis_of_row[0] == axis_of_row[1] || axis_of_row[0] == axis_of_row[2] || axis_of_row[1] == axis_of_row[2])
    return vgi::set_error(VGI_E_INVALID, "voxel camera maps two clip axes to one world axis");
  // world interval of each clip axis: clip_r = M[r][a] * world_a + M[r][3]
  const double clip_lo[3] = {-1.0, -1.0, 0.0}, clip_hi[3] = {1.0, 1.0, 1.0};
  double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  for (int r = 0; r < 3; r++) {
    const int a = axis_of_row[r];
    const double w0 = (clip_lo[r] - M[r][3]) / M[r][a], w1 = (clip_hi[r] - M[r][3]) / M[r][a];
    lo[a] = std::fmin(w0, w1);
    hi[a] = std::fmax(w0, w1);
  }
  double edge = 0.0;
  for (int a = 0; a < 3; a++) edge = std::fmax(edge, hi[a] - lo[a]);
  vgi_grid g;
  for (int a = 0; a < 3; a++) g.origin[a] = (float)(0.5 * (lo[a] + hi[a]) - 0.5 * edge);
  g.extent = (float)edge;
  g.n = n;
  g.levels = levels;
  if (!(g.extent > 0.0f) || !std::isfinite(g.extent)) return vgi::set_error(VGI_E_INVALID, "voxel camera box is empty");
  *out = g;
  return VGI_OK;
}
