"""Seeded synthetic workloads (ctypes over libvgi_gen.so).

The configurations follow the workloads the path is measured on:

  random_soup  10,000 random triangles in [0,1]^3 (edges <= 0.05, albedo RGB8), one
               directional light (0.3,-1,0.2)/|.|, 64^3 grid with 6 levels, a 256x256
               G-buffer of points sampled on the triangles, 6 diffuse 60-degree cones +
               1 specular 10-degree cone, step 0.5 voxel, max distance 1.0, cutoff 0.99.
  room         procedural room 1.0 x 0.45 x 0.6 (walls, 64 columns, 40,000 tiles, debris)
               padded to N triangles, 4 point lights (intensity 1-5), a pinhole-camera
               G-buffer with per-pixel roughness (0.05-0.5), 9 (or 16) diffuse cones +
               1 roughness-driven specular cone. Default: 262,144 triangles, 256^3 grid,
               8 levels, 1920x1080.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass, field

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
GEN_PATH = os.path.join(_PKG, "libvgi_gen.so")


def _load():
    if not os.path.exists(GEN_PATH):
        raise ImportError(f"{GEN_PATH} is missing: build it with `python -m final184_amd.build`")
    g = C.CDLL(GEN_PATH)
    P = C.c_void_p
    g.sg_random_triangles.argtypes = [C.c_uint64, C.c_uint32, C.c_float, C.c_float, C.c_float, P, P]
    g.sg_random_triangles.restype = C.c_int
    g.sg_room.argtypes = [C.c_uint64, C.c_uint32, C.c_float, C.c_float, C.c_float, P, P, P]
    g.sg_room.restype = C.c_int64
    g.sg_gbuffer_on_triangles.argtypes = [C.c_uint64, P, C.c_uint32, C.c_uint64, P, P]
    g.sg_gbuffer_on_triangles.restype = C.c_int
    g.sg_camera_pose.argtypes = [C.c_uint64, P, P, P]
    g.sg_camera_pose.restype = C.c_int
    g.sg_gbuffer_camera.argtypes = [P, P, C.c_uint32, P, C.c_uint32, C.c_uint32, C.c_float, P, P, P]
    g.sg_gbuffer_camera.restype = C.c_int
    g.sg_point_lights.argtypes = [C.c_uint64, C.c_uint32, P, P, C.c_float, C.c_float, P]
    g.sg_point_lights.restype = C.c_int
    g.sg_lognormal_ints.argtypes = [C.c_uint64, C.c_uint32, C.c_double, C.c_double, C.c_int64,
                                    C.c_int64, P]
    g.sg_lognormal_ints.restype = C.c_int
    return g


_gen = _load()


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def random_triangles(seed: int, ntri: int, lo=0.0, hi=1.0, radius=0.025):
    pos = np.empty((ntri, 9), np.float32)
    alb = np.empty((ntri, 3), np.uint8)
    rc = _gen.sg_random_triangles(seed, ntri, lo, hi, radius, pos.ctypes.data, alb.ctypes.data)
    assert rc == 0
    return pos, alb


def room_triangles(seed: int, ntri: int, dims=(1.0, 0.45, 0.6)):
    pos = np.empty((ntri, 9), np.float32)
    alb = np.empty((ntri, 3), np.uint8)
    rough = np.empty((ntri,), np.float32)
    n = _gen.sg_room(seed, ntri, dims[0], dims[1], dims[2], pos.ctypes.data, alb.ctypes.data,
                     rough.ctypes.data)
    assert n == ntri, (n, ntri)
    return pos, alb, rough


def gbuffer_on_triangles(seed: int, pos: np.ndarray, npix: int):
    pos = _f32(pos)
    gp = np.empty((npix, 3), np.float32)
    gn = np.empty((npix, 3), np.float32)
    rc = _gen.sg_gbuffer_on_triangles(seed, pos.ctypes.data, pos.shape[0], npix, gp.ctypes.data,
                                      gn.ctypes.data)
    assert rc == 0
    return gp, gn


def camera_pose(seed: int, lo, hi):
    cam = np.empty(12, np.float32)
    lo_a, hi_a = _f32(lo), _f32(hi)  # keep the arrays alive across the call
    _gen.sg_camera_pose(seed, lo_a.ctypes.data, hi_a.ctypes.data, cam.ctypes.data)
    return cam


def gbuffer_camera(pos, tri_rough, cam, width, height, vfov_deg=60.0):
    pos = _f32(pos)
    tri_rough = _f32(tri_rough)
    cam = _f32(cam)
    gp = np.empty((width * height, 3), np.float32)
    gn = np.empty((width * height, 3), np.float32)
    gr = np.empty((width * height,), np.float32)
    rc = _gen.sg_gbuffer_camera(pos.ctypes.data, tri_rough.ctypes.data, pos.shape[0],
                                cam.ctypes.data, width, height, vfov_deg, gp.ctypes.data,
                                gn.ctypes.data, gr.ctypes.data)
    assert rc == 0
    return gp, gn, gr


def camera_view_matrix(cam) -> np.ndarray:
    """World -> view (float64 4x4, row-major) of a camera_pose (eye, forward, right, up):
    view x = right, y = up, z = -forward (the camera looks down -z)."""
    c = np.asarray(cam, np.float64)
    e, f, r, u = c[0:3], c[3:6], c[6:9], c[9:12]
    m = np.eye(4)
    m[0, :3], m[1, :3], m[2, :3] = r, u, -f
    m[:3, 3] = -(m[:3, :3] @ e)
    return m


def perspective_matrix(vfov_deg: float, aspect: float, near: float, far: float) -> np.ndarray:
    """The reference's CCamera::CalcPerspective (Foreground/SceneGraph/Camera.cpp:85-101),
    float64 4x4 row-major: OpenGL-style depth terms, y negated for Vulkan's NDC (row 0 of
    the screen is +y in view space)."""
    ys = math.tan(math.radians(vfov_deg) / 2.0)
    xs = ys * aspect
    a, b = 1.0 / xs, 1.0 / ys
    c = (far + near) / (near - far)
    d = -2.0 * far * near / (far - near)
    return np.array([[a, 0, 0, 0], [0, -b, 0, 0], [0, 0, c, d], [0, 0, -1, 0]], np.float64)


def glsl_mat4(m: np.ndarray) -> np.ndarray:
    """16 float32 in GLSL's column-major order (m[column * 4 + row])."""
    return np.ascontiguousarray(np.asarray(m, np.float64).T.reshape(16), np.float32)


def reference_gbuffer(gb_pos, gb_nrm, cam, width: int, height: int, vfov_deg: float = 60.0,
                      near: float = 0.05, far: float = 10.0):
    """A world-space G-buffer (camera_pose + gbuffer_camera) in the reference's render-target
    form: D32 depth = z_ndc of the reference's projection, RGBA16 UNORM view-space normals
    (n * 0.5 + 0.5, rounded), and the column-major InvProj / InvModelView the reference
    uploads (SceneView.cpp:28-30, MegaPipeline.cpp:242). Pixels without a surface get depth 1
    and normal 0. Returns depth (H*W float32), normals (H*W x 4 uint16), inv_proj,
    inv_model_view (16 float32 each), view and proj (float64 4x4)."""
    V = camera_view_matrix(cam)
    P = perspective_matrix(vfov_deg, width / height, near, far)
    pos = np.asarray(gb_pos, np.float64).reshape(-1, 3)
    nrm = np.asarray(gb_nrm, np.float64).reshape(-1, 3)
    hit = np.any(nrm != 0, axis=1)
    vp = pos @ V[:3, :3].T + V[:3, 3]
    clip_z = vp[:, 2] * P[2, 2] + P[2, 3]
    clip_w = -vp[:, 2]
    depth = np.where(hit, clip_z / np.where(hit, clip_w, 1.0), 1.0).astype(np.float32)
    vn = nrm @ V[:3, :3].T
    q = np.rint((vn * 0.5 + 0.5) * 65535.0)
    normals = np.zeros((pos.shape[0], 4), np.uint16)
    normals[:, :3] = np.where(hit[:, None], np.clip(q, 0, 65535), 0).astype(np.uint16)
    normals[:, 3] = 65535
    return depth, normals, glsl_mat4(np.linalg.inv(P)), glsl_mat4(np.linalg.inv(V)), V, P


def point_lights(seed: int, count: int, lo, hi, imin=1.0, imax=5.0):
    out = np.empty((count, 4), np.float32)
    lo_a, hi_a = _f32(lo), _f32(hi)  # keep the arrays alive across the call
    _gen.sg_point_lights(seed, count, lo_a.ctypes.data, hi_a.ctypes.data, imin, imax,
                         out.ctypes.data)
    return out


def lognormal_ints(seed: int, count: int, median: float, sigma: float, lo: int, hi: int):
    out = np.empty(count, np.int64)
    _gen.sg_lognormal_ints(seed, count, median, sigma, lo, hi, out.ctypes.data)
    return out


@dataclass
class Workload:
    """Everything one frame of the path needs, as host arrays."""
    name: str
    positions: np.ndarray            # (T, 9) float32, non-indexed triangles
    albedo: np.ndarray               # (T, 3) uint8
    grid: dict                       # origin, extent, n, levels
    lights: list                     # dicts for vgi.make_lights
    gb_pos: np.ndarray               # (P, 3) float32
    gb_nrm: np.ndarray               # (P, 3) float32
    width: int
    height: int
    cones: dict                      # kwargs for vgi.make_cones
    gb_rough: np.ndarray | None = None
    meta: dict = field(default_factory=dict)

    @property
    def num_triangles(self) -> int:
        return self.positions.shape[0]

    @property
    def num_pixels(self) -> int:
        return self.width * self.height


SEED_SOUP = 0xF1840001
SEED_ROOM = 0xF1840002


def random_soup(seed: int = SEED_SOUP, ntri: int = 10_000, n: int = 64, levels: int = 6,
                width: int = 256, height: int = 256) -> Workload:
    pos, alb = random_triangles(seed, ntri)
    gp, gn = gbuffer_on_triangles(seed, pos, width * height)
    d = np.array([0.3, -1.0, 0.2], np.float64)
    d /= np.linalg.norm(d)
    return Workload(
        name=f"random_soup_{ntri}tri_{n}^3_{width}x{height}",
        positions=pos, albedo=alb,
        grid=dict(origin=(0.0, 0.0, 0.0), extent=1.0, n=n, levels=levels),
        lights=[dict(type="directional", vec=tuple(float(v) for v in d), color=(1.0, 1.0, 1.0))],
        gb_pos=gp, gb_nrm=gn, width=width, height=height,
        cones=dict(num_diffuse=6, specular=True, diffuse_aperture_deg=60.0, diffuse_ring_deg=60.0,
                   specular_aperture_deg=10.0, step_factor=0.5, max_distance=1.0,
                   alpha_cutoff=0.99, normal_offset=1.0, start_offset=1.0, eye=(0.5, 0.5, 2.0)),
        meta=dict(seed=seed))


ROOM_DIMS = (1.0, 0.45, 0.6)


def room(seed: int = SEED_ROOM, ntri: int = 262_144, n: int = 256, levels: int = 8,
         width: int = 1920, height: int = 1080, num_diffuse: int = 9,
         max_distance: float = 1.0) -> Workload:
    pos, alb, rough = room_triangles(seed, ntri, ROOM_DIMS)
    cam = camera_pose(seed, (0.25, 0.15, 0.15), (0.75, 0.30, 0.45))
    gp, gn, gr = gbuffer_camera(pos, rough, cam, width, height, 60.0)
    L = point_lights(seed, 4, (0.05, 0.10, 0.05), (0.95, 0.40, 0.55), 1.0, 5.0)
    lights = [dict(type="point", vec=tuple(float(v) for v in L[i, :3]),
                   color=(float(L[i, 3]),) * 3, min_distance=0.05) for i in range(4)]
    return Workload(
        name=f"room_{ntri}tri_{n}^3_{width}x{height}_{num_diffuse}+1cones",
        positions=pos, albedo=alb,
        grid=dict(origin=(-0.02, -0.02, -0.02), extent=1.04, n=n, levels=levels),
        lights=lights, gb_pos=gp, gb_nrm=gn, gb_rough=gr, width=width, height=height,
        cones=dict(num_diffuse=num_diffuse, specular=True, diffuse_aperture_deg=60.0,
                   diffuse_ring_deg=60.0, specular_aperture_deg=10.0, step_factor=0.5,
                   max_distance=max_distance, alpha_cutoff=0.99, normal_offset=1.0,
                   start_offset=1.0, eye=tuple(float(v) for v in cam[:3])),
        meta=dict(seed=seed, camera=cam.tolist()))


SEED_BATCH = 0xF1840004


def batch_scenes(seed: int = SEED_BATCH, count: int = 64, median: float = 20_000.0,
                 sigma: float = 1.0, tri_range=(1_000, 200_000), grids=(32, 64, 128, 256),
                 sides=(256, 512, 1024), num_diffuse: int = 6) -> list[Workload]:
    """Independent scenes of uneven size for one batch: triangle counts lognormal (median,
    sigma) clipped to tri_range, grid edge and square G-buffer side drawn uniformly from
    the given sets (mip chain down to 2^3), seeded random triangle soups lit by one
    directional light, G-buffer points sampled on the triangles."""
    rng = np.random.default_rng(seed)
    counts = lognormal_ints(seed, count, median, sigma, tri_range[0], tri_range[1])
    out = []
    for i in range(count):
        n = int(rng.choice(grids))
        side = int(rng.choice(sides))
        s = seed + 0x9E37 * (i + 1)
        pos, alb = random_triangles(s, int(counts[i]))
        gp, gn = gbuffer_on_triangles(s, pos, side * side)
        d = rng.normal(size=3)
        d[1] = -abs(d[1]) - 0.5
        d /= np.linalg.norm(d)
        out.append(Workload(
            name=f"batch{i}_{int(counts[i])}tri_{n}^3_{side}^2",
            positions=pos, albedo=alb,
            grid=dict(origin=(0.0, 0.0, 0.0), extent=1.0, n=n, levels=int(math.log2(n))),
            lights=[dict(type="directional", vec=tuple(float(v) for v in d), color=(1.0, 1.0, 1.0))],
            gb_pos=gp, gb_nrm=gn, width=side, height=side,
            cones=dict(num_diffuse=num_diffuse, specular=True, diffuse_aperture_deg=60.0,
                       diffuse_ring_deg=60.0, specular_aperture_deg=10.0, step_factor=0.5,
                       max_distance=1.0, alpha_cutoff=0.99, normal_offset=1.0, start_offset=1.0,
                       eye=(0.5, 0.5, 2.0)),
            meta=dict(seed=s)))
    return out


def cone_samples_estimate(h: float, tan_half: float, step: float, start: float, max_d: float) -> int:
    """Number of samples of one cone with no early exit (closed-form schedule)."""
    t, m = start * h, 0
    while t <= max_d:
        m += 1
        t += step * max(h, 2.0 * t * tan_half)
    return m


def deg_tan_half(deg: float) -> float:
    return math.tan(math.radians(deg) / 2.0)
