"""ctypes binding of libvgi.so (the C ABI in include/vgi.h).

This is plumbing for tests, benchmarks and Python callers: device memory comes from
the library's own vgi_malloc (no torch), host arrays are numpy. Every call checks its
status and raises VgiError with the library's message. There is no CPU fallback: if
libvgi.so is missing or cannot be loaded, importing this module raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libvgi.so")

VGI_OK = 0
ERRORS = {
    -1: "VGI_E_INVALID", -2: "VGI_E_NOMEM", -3: "VGI_E_HIP", -4: "VGI_E_BUDGET",
    -5: "VGI_E_NOT_READY", -6: "VGI_E_RANGE", -7: "VGI_E_NODEVICE", -8: "VGI_E_STATE",
}
PREC_FP32, PREC_FP16, PREC_BF16, PREC_FP16X2 = 0, 1, 2, 3
PREC_NAMES = {"fp32": PREC_FP32, "fp16": PREC_FP16, "bf16": PREC_BF16, "fp16x2": PREC_FP16X2}
INDEX_NONE, INDEX_U16, INDEX_U32 = 0, 1, 2
LIGHT_DIRECTIONAL, LIGHT_POINT = 0, 1
LIGHT_SHADOW = 0x100  # vgi.h VGI_LIGHT_SHADOW: OR'd into the type of a light that casts shadows
MAX_LEVELS, MAX_LIGHTS, MAX_DIFFUSE_CONES = 11, 128, 32


class VgiError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Grid(C.Structure):
    _fields_ = [("origin", C.c_float * 3), ("extent", C.c_float), ("n", C.c_uint32),
                ("levels", C.c_uint32)]


class Mesh(C.Structure):
    _fields_ = [("positions", C.c_void_p), ("position_stride", C.c_uint32),
                ("num_vertices", C.c_uint32), ("indices", C.c_void_p), ("index_type", C.c_uint32),
                ("num_triangles", C.c_uint32), ("albedo_rgb8", C.c_void_p),
                ("base_color", C.c_uint8 * 4), ("model", C.c_float * 12),
                ("texcoords", C.c_void_p), ("texcoord_stride", C.c_uint32), ("texture_width", C.c_uint32),
                ("texture_height", C.c_uint32), ("texture_reserved", C.c_uint32),
                ("texture_rgba8", C.c_void_p), ("base_color_factor", C.c_float * 4)]


assert C.sizeof(Mesh) == 144, "vgi_mesh layout"


class Light(C.Structure):
    _fields_ = [("type", C.c_uint32), ("vec", C.c_float * 3), ("color", C.c_float * 3),
                ("min_distance", C.c_float)]


class Lights(C.Structure):
    _fields_ = [("count", C.c_uint32), ("light", Light * MAX_LIGHTS)]


class ConeParams(C.Structure):
    _fields_ = [("num_diffuse", C.c_uint32), ("specular", C.c_uint32),
                ("diffuse_aperture_deg", C.c_float), ("diffuse_ring_deg", C.c_float),
                ("specular_aperture_deg", C.c_float), ("step_factor", C.c_float),
                ("max_distance", C.c_float), ("alpha_cutoff", C.c_float),
                ("normal_offset", C.c_float), ("start_offset", C.c_float), ("eye", C.c_float * 3)]


class GBuffer(C.Structure):
    _fields_ = [("position", C.c_void_p), ("normal", C.c_void_p), ("roughness", C.c_void_p),
                ("width", C.c_uint32), ("height", C.c_uint32)]


ORDER_TILES, ORDER_MORTON = 0, 1
ADDRESS_AUTO, ADDRESS_WIDE = 0, 1  # cone-march texel addressing (vgi.h VGI_ADDRESS_*)
IMAGE_RGBA32F, IMAGE_RGBA16F = 0, 1
PROBE_FMA, PROBE_ATOMIC_RETURN, PROBE_ATOMIC_NORETURN, PROBE_INT = 0, 1, 2, 3  # vgi.h VGI_PROBE_*


class GBufferRef(C.Structure):
    """The reference's G-buffer: D32 depth, RGBA16 UNORM view-space normals and the
    column-major inverse projection / inverse model-view."""
    _fields_ = [("depth", C.c_void_p), ("normals", C.c_void_p), ("inv_proj", C.c_float * 16),
                ("inv_model_view", C.c_float * 16), ("width", C.c_uint32), ("height", C.c_uint32)]


class TemporalParams(C.Structure):
    _fields_ = [("indirect", C.c_void_p), ("position", C.c_void_p), ("view_depth", C.c_void_p),
                ("history", C.c_void_p), ("out", C.c_void_p), ("history_format", C.c_uint32),
                ("out_format", C.c_uint32), ("prev_projection", C.c_float * 16),
                ("prev_model_view", C.c_float * 16), ("width", C.c_uint32), ("height", C.c_uint32)]


GLTF_ALBEDO_REFERENCE, GLTF_ALBEDO_FACTOR = 0, 1


class GltfInfo(C.Structure):
    _fields_ = [("num_primitives", C.c_uint32), ("num_geometries", C.c_uint32),
                ("skipped_primitives", C.c_uint32), ("num_images", C.c_uint32), ("num_triangles", C.c_uint64)]


class GltfPrimitive(C.Structure):
    _fields_ = [("positions", C.c_void_p), ("indices", C.c_void_p), ("num_vertices", C.c_uint32),
                ("num_triangles", C.c_uint32), ("model", C.c_float * 12), ("base_color", C.c_uint8 * 4),
                ("node", C.c_int32), ("mesh", C.c_int32), ("primitive", C.c_int32), ("material", C.c_int32),
                ("texcoords", C.c_void_p), ("image", C.c_int32), ("reserved", C.c_uint32),
                ("base_color_factor", C.c_float * 4)]


class GltfImage(C.Structure):
    _fields_ = [("data", C.c_void_p), ("bytes", C.c_size_t), ("mime_type", C.c_char_p), ("uri", C.c_char_p)]


class BlurParams(C.Structure):
    _fields_ = [("color", C.c_void_p), ("depth", C.c_void_p), ("out", C.c_void_p),
                ("color_format", C.c_uint32), ("out_format", C.c_uint32), ("width", C.c_uint32),
                ("height", C.c_uint32), ("axis", C.c_uint32)]


class Tuning(C.Structure):
    _fields_ = [("bin_size", C.c_uint32), ("chunk_tris", C.c_uint32), ("tile_w", C.c_uint32),
                ("tile_h", C.c_uint32), ("pixel_order", C.c_uint32), ("address_mode", C.c_uint32)]


class SceneStats(C.Structure):
    _fields_ = [("triangles_in", C.c_uint64), ("triangles_rejected", C.c_uint64),
                ("triangles_degenerate", C.c_uint64), ("bin_pairs", C.c_uint64),
                ("work_items", C.c_uint64), ("device_bytes", C.c_uint64), ("n", C.c_uint32),
                ("levels", C.c_uint32), ("precision", C.c_uint32), ("has_voxels", C.c_uint32),
                ("has_mips", C.c_uint32), ("bin_size", C.c_uint32), ("voxelize_fallback", C.c_uint32),
                ("pad", C.c_uint32)]


class Timings(C.Structure):
    _fields_ = [("voxelize_setup_ms", C.c_float), ("voxelize_sort_ms", C.c_float),
                ("voxelize_raster_ms", C.c_float), ("voxelize_total_ms", C.c_float),
                ("mips_ms", C.c_float), ("trace_ms", C.c_float),
                ("trace_trilinear_samples", C.c_uint64)]


class VoxelizeItem(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("meshes", C.c_void_p), ("num_meshes", C.c_uint32),
                ("grid", Grid)]


class MipsItem(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("lights", C.c_void_p)]


class TraceItem(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("gbuffer", GBuffer), ("cones", ConeParams),
                ("out_diffuse", C.c_void_p), ("out_specular", C.c_void_p)]


class AutotuneReport(C.Structure):
    _fields_ = [("best", Tuning), ("bin_sizes", C.c_uint32 * 3), ("tiles", (C.c_uint32 * 3) * 4),
                ("build_ms", C.c_float * 3), ("trace_ms", C.c_float * 4),
                ("frame_ms", (C.c_float * 4) * 3)]

    @staticmethod
    def schedule_name(w, h, order) -> str:
        return "morton" if order == ORDER_MORTON else f"{w}x{h}"

    def as_dict(self) -> dict:
        b = self.best
        return dict(best=dict(bin_size=b.bin_size, schedule=self.schedule_name(b.tile_w, b.tile_h, b.pixel_order)),
                    build_ms={str(self.bin_sizes[i]): round(self.build_ms[i], 4) for i in range(3)},
                    trace_ms={self.schedule_name(*self.tiles[j]): round(self.trace_ms[j], 4) for j in range(4)})


class ServerConfig(C.Structure):
    _fields_ = [("memory_budget", C.c_uint64), ("precision", C.c_uint32), ("slots", C.c_uint32),
                ("max_batch_requests", C.c_uint32), ("max_batch_pixels", C.c_uint64)]


class Frame(C.Structure):
    _fields_ = [("gbuffer", GBuffer), ("cones", ConeParams), ("out_diffuse", C.c_void_p),
                ("out_specular", C.c_void_p)]


class Request(C.Structure):
    _fields_ = [("scene_key", C.c_uint64), ("meshes", C.c_void_p), ("num_meshes", C.c_uint32),
                ("revoxelize", C.c_uint32), ("grid", Grid), ("lights", Lights),
                ("frames", C.c_void_p), ("num_frames", C.c_uint32), ("pad", C.c_uint32),
                ("arrival_s", C.c_double), ("tag", C.c_uint64)]


class Completion(C.Structure):
    _fields_ = [("tag", C.c_uint64), ("status", C.c_int32), ("voxelized", C.c_uint32),
                ("arrival_s", C.c_double), ("start_s", C.c_double), ("done_s", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [(f, C.c_uint64) for f in ("submitted", "completed", "failed", "builds", "evictions",
                                          "batches", "frames", "pixels", "resident_scenes",
                                          "resident_bytes", "peak_resident_bytes")] + \
               [("max_concurrent_batches", C.c_uint32), ("max_batch_requests_seen", C.c_uint32)]

    def as_dict(self) -> dict:
        return {f: int(getattr(self, f)) for f, _ in self._fields_}


# (name, restype, argtypes) for every exported entry point of include/vgi.h
_P = C.c_void_p
SIGNATURES = [
    ("vgi_version", C.c_char_p, []),
    ("vgi_abi_version", C.c_uint32, []),
    ("vgi_last_error", C.c_char_p, []),
    ("vgi_create", C.c_int, [C.POINTER(_P), C.c_int]),
    ("vgi_destroy", C.c_int, [_P]),
    ("vgi_set_tuning", C.c_int, [_P, C.POINTER(Tuning)]),
    ("vgi_get_tuning", C.c_int, [_P, C.POINTER(Tuning)]),
    ("vgi_set_profiling", C.c_int, [_P, C.c_int]),
    ("vgi_get_timings", C.c_int, [_P, C.POINTER(Timings)]),
    ("vgi_sync", C.c_int, [_P, _P]),
    ("vgi_scene_create", C.c_int, [_P, C.POINTER(_P)]),
    ("vgi_scene_release", C.c_int, [_P]),
    ("vgi_scene_get_stats", C.c_int, [_P, C.POINTER(SceneStats)]),
    ("vgi_scene_bytes_estimate", C.c_uint64, [C.POINTER(Grid), C.c_uint32, C.c_uint64]),
    ("vgi_voxelize", C.c_int, [_P, _P, C.POINTER(Mesh), C.c_uint32, C.POINTER(Grid), _P]),
    ("vgi_build_mips", C.c_int, [_P, _P, C.POINTER(Lights), C.c_uint32, _P]),
    ("vgi_trace", C.c_int, [_P, _P, C.POINTER(GBuffer), C.POINTER(ConeParams), _P, _P, _P]),
    ("vgi_voxelize_batch", C.c_int, [_P, C.POINTER(VoxelizeItem), C.c_uint32, _P]),
    ("vgi_build_mips_batch", C.c_int, [_P, C.POINTER(MipsItem), C.c_uint32, C.c_uint32, _P]),
    ("vgi_trace_batch", C.c_int, [_P, C.POINTER(TraceItem), C.c_uint32, _P]),
    ("vgi_gbuffer_decode", C.c_int, [_P, C.POINTER(GBufferRef), _P, _P, _P, _P]),
    ("vgi_temporal_resolve", C.c_int, [_P, C.POINTER(TemporalParams), _P]),
    ("vgi_bilateral_blur", C.c_int, [_P, C.POINTER(BlurParams), _P]),
    ("vgi_gltf_load_file", C.c_int, [C.c_char_p, C.c_uint32, C.POINTER(_P)]),
    ("vgi_gltf_load_memory", C.c_int, [_P, C.c_size_t, C.c_char_p, C.c_uint32, C.POINTER(_P)]),
    ("vgi_gltf_get_info", C.c_int, [_P, C.POINTER(GltfInfo)]),
    ("vgi_gltf_get_primitive", C.c_int, [_P, C.c_uint32, C.POINTER(GltfPrimitive)]),
    ("vgi_gltf_get_image", C.c_int, [_P, C.c_uint32, C.POINTER(GltfImage)]),
    ("vgi_gltf_set_image_rgba8", C.c_int, [_P, C.c_uint32, _P, C.c_uint32, C.c_uint32]),
    ("vgi_gltf_upload", C.c_int, [_P, _P, C.POINTER(Mesh), C.c_uint32, C.POINTER(C.c_uint32)]),
    ("vgi_gltf_release", C.c_int, [_P]),
    ("vgi_autotune_batch", C.c_int, [_P, C.POINTER(VoxelizeItem), C.POINTER(MipsItem),
                                     C.POINTER(TraceItem), C.c_uint32, C.c_uint32, C.c_uint32,
                                     C.POINTER(AutotuneReport), _P]),
    ("vgi_now", C.c_double, []),
    ("vgi_server_create", C.c_int, [_P, C.POINTER(ServerConfig), C.POINTER(_P)]),
    ("vgi_server_submit", C.c_int, [_P, C.POINTER(Request)]),
    ("vgi_server_poll", C.c_int, [_P, C.POINTER(Completion), C.c_uint32, C.POINTER(C.c_uint32),
                                  C.c_uint32]),
    ("vgi_server_drain", C.c_int, [_P]),
    ("vgi_server_get_stats", C.c_int, [_P, C.POINTER(ServerStats)]),
    ("vgi_server_destroy", C.c_int, [_P]),
    ("vgi_scene_read_voxels", C.c_int, [_P, _P, _P]),
    ("vgi_scene_read_level", C.c_int, [_P, _P, C.c_uint32, _P]),
    ("vgi_malloc", C.c_int, [_P, C.POINTER(_P), C.c_size_t]),
    ("vgi_free", C.c_int, [_P, _P]),
    ("vgi_memcpy_h2d", C.c_int, [_P, _P, _P, C.c_size_t]),
    ("vgi_memcpy_d2h", C.c_int, [_P, _P, _P, C.c_size_t]),
    ("vgi_memset", C.c_int, [_P, _P, C.c_int, C.c_size_t]),
    ("vgi_memcpy_d2d", C.c_int, [_P, _P, _P, C.c_size_t, _P]),
    ("vgi_stream_create", C.c_int, [_P, C.POINTER(_P)]),
    ("vgi_stream_destroy", C.c_int, [_P, _P]),
    ("vgi_stream_wait", C.c_int, [_P, _P, _P]),
    ("vgi_stream_query", C.c_int, [_P, _P]),
    ("vgi_voxelize_noalloc", C.c_int, [_P, _P, _P, C.c_uint32, _P, _P]),
    ("vgi_grid_from_voxel_camera", C.c_int, [_P, _P, C.c_uint32, C.c_uint32, _P]),
    ("vgi_probe", C.c_int, [_P, C.c_uint32, _P, C.c_size_t, C.c_uint32, C.c_uint32, _P]),
    ("vgi_workspace_bytes", C.c_int, [_P, C.POINTER(C.c_uint64)]),
    ("vgi_timer_start", C.c_int, [_P, _P]),
    ("vgi_timer_stop", C.c_int, [_P, _P, C.POINTER(C.c_float)]),
]


def load_library(path: str = LIB_PATH) -> C.CDLL:
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing: build it with `python -m final184_amd.build` "
                          "(there is no CPU fallback)")
    lib = C.CDLL(path)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = load_library()


def _check(rc: int) -> None:
    if rc != VGI_OK:
        raise VgiError(rc, lib.vgi_last_error().decode())


def make_grid(origin, extent: float, n: int, levels: int) -> Grid:
    g = Grid()
    g.origin[:] = [float(v) for v in origin]
    g.extent = float(extent)
    g.n = int(n)
    g.levels = int(levels)
    return g


def grid_from_voxel_camera(voxel_view, voxel_proj, n: int, levels: int) -> Grid:
    """The grid holding the reference's voxelizer camera volume (GLSL column-major mat4s)."""
    v = np.ascontiguousarray(voxel_view, np.float32).reshape(16)
    p = np.ascontiguousarray(voxel_proj, np.float32).reshape(16)
    g = Grid()
    _check(lib.vgi_grid_from_voxel_camera(v.ctypes.data, p.ctypes.data, n, levels, C.byref(g)))
    return g


def make_lights(lights) -> Lights:
    """lights: iterable of dicts {type: 'directional'|'point', vec, color, min_distance, shadow}."""
    L = Lights()
    items = list(lights)
    if len(items) > MAX_LIGHTS:
        raise ValueError("too many lights")
    L.count = len(items)
    for i, d in enumerate(items):
        L.light[i].type = ((LIGHT_DIRECTIONAL if d["type"] == "directional" else LIGHT_POINT)
                           | (LIGHT_SHADOW if d.get("shadow") else 0))
        L.light[i].vec[:] = [float(v) for v in d["vec"]]
        L.light[i].color[:] = [float(v) for v in d["color"]]
        L.light[i].min_distance = float(d.get("min_distance", 0.0))
    return L


def make_cones(num_diffuse=6, specular=True, diffuse_aperture_deg=60.0, diffuse_ring_deg=60.0,
               specular_aperture_deg=10.0, step_factor=0.5, max_distance=1.0, alpha_cutoff=0.99,
               normal_offset=1.0, start_offset=1.0, eye=(0.5, 0.5, 2.0)) -> ConeParams:
    c = ConeParams()
    c.num_diffuse = num_diffuse
    c.specular = 1 if specular else 0
    c.diffuse_aperture_deg = diffuse_aperture_deg
    c.diffuse_ring_deg = diffuse_ring_deg
    c.specular_aperture_deg = specular_aperture_deg
    c.step_factor = step_factor
    c.max_distance = max_distance
    c.alpha_cutoff = alpha_cutoff
    c.normal_offset = normal_offset
    c.start_offset = start_offset
    c.eye[:] = [float(v) for v in eye]
    return c


IDENTITY_3X4 = (1.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0)


class DeviceArray:
    """A device allocation owned by a Context, with numpy-like shape/dtype metadata."""

    def __init__(self, ctx: "Context", nbytes: int, shape=None, dtype=np.uint8):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        self.shape = tuple(shape) if shape is not None else (self.nbytes,)
        self.dtype = np.dtype(dtype)
        p = C.c_void_p()
        _check(lib.vgi_malloc(ctx.handle, C.byref(p), self.nbytes))
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, ctx: "Context", arr: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(arr)
        d = cls(ctx, a.nbytes, a.shape, a.dtype)
        if a.nbytes:
            _check(lib.vgi_memcpy_h2d(ctx.handle, d.ptr, a.ctypes.data, a.nbytes))
        return d

    @classmethod
    def empty(cls, ctx: "Context", shape, dtype) -> "DeviceArray":
        dt = np.dtype(dtype)
        n = int(np.prod(shape)) * dt.itemsize
        return cls(ctx, n, shape, dt)

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        if out.nbytes:
            _check(lib.vgi_memcpy_d2h(self.ctx.handle, out.ctypes.data, self.ptr, out.nbytes))
        return out

    def zero(self) -> None:
        _check(lib.vgi_memset(self.ctx.handle, self.ptr, 0, self.nbytes))

    def free(self) -> None:
        if self.ptr:
            _check(lib.vgi_free(self.ctx.handle, self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            if self.ptr and self.ctx.handle:
                lib.vgi_free(self.ctx.handle, self.ptr)
        except Exception:
            pass


class MeshData:
    """Device-resident mesh (positions + optional indices/albedo, or texcoords + an RGBA8
    texture) and its vgi_mesh view."""

    def __init__(self, ctx: "Context", positions: np.ndarray, indices: np.ndarray | None = None,
                 albedo: np.ndarray | None = None, base_color=(255, 255, 255),
                 model=IDENTITY_3X4, index_type: int | None = None, position_stride: int = 0,
                 texcoords: np.ndarray | None = None, texture: np.ndarray | None = None,
                 base_color_factor=(1.0, 1.0, 1.0, 1.0)):
        pos = np.ascontiguousarray(positions, dtype=np.float32)
        self.pos = DeviceArray.from_numpy(ctx, pos)
        self.num_vertices = pos.size * 4 // (position_stride or 12)
        if indices is not None:
            idx = np.ascontiguousarray(indices)
            if index_type is None:
                index_type = INDEX_U16 if idx.dtype == np.uint16 else INDEX_U32
            idx = idx.astype(np.uint16 if index_type == INDEX_U16 else np.uint32, copy=False)
            self.idx = DeviceArray.from_numpy(ctx, idx)
            self.num_triangles = idx.size // 3
        else:
            self.idx = None
            index_type = INDEX_NONE
            self.num_triangles = self.num_vertices // 3
        self.albedo = DeviceArray.from_numpy(ctx, np.ascontiguousarray(albedo, np.uint8)) \
            if albedo is not None else None
        m = Mesh()
        m.positions = self.pos.ptr
        m.position_stride = position_stride
        m.num_vertices = self.num_vertices
        m.indices = self.idx.ptr if self.idx is not None else None
        m.index_type = index_type
        m.num_triangles = self.num_triangles
        m.albedo_rgb8 = self.albedo.ptr if self.albedo is not None else None
        m.base_color[:] = [int(base_color[0]), int(base_color[1]), int(base_color[2]), 0]
        m.model[:] = [float(v) for v in model]
        self.uv = self.tex = None
        if texcoords is not None and texture is not None:
            tex = np.ascontiguousarray(texture, np.uint8)
            if tex.ndim != 3 or tex.shape[2] != 4:
                raise ValueError("texture must be (height, width, 4) RGBA8")
            self.uv = DeviceArray.from_numpy(ctx, np.ascontiguousarray(texcoords, np.float32))
            self.tex = DeviceArray.from_numpy(ctx, tex)
            m.texcoords = self.uv.ptr
            m.texcoord_stride = 0
            m.texture_height, m.texture_width = int(tex.shape[0]), int(tex.shape[1])
            m.texture_rgba8 = self.tex.ptr
        m.base_color_factor[:] = [float(v) for v in base_color_factor]
        self.desc = m


class GBufferData:
    def __init__(self, ctx: "Context", position: np.ndarray, normal: np.ndarray, width: int,
                 height: int, roughness: np.ndarray | None = None):
        self.pos = DeviceArray.from_numpy(ctx, np.ascontiguousarray(position, np.float32))
        self.nrm = DeviceArray.from_numpy(ctx, np.ascontiguousarray(normal, np.float32))
        self.rough = DeviceArray.from_numpy(ctx, np.ascontiguousarray(roughness, np.float32)) \
            if roughness is not None else None
        g = GBuffer()
        g.position = self.pos.ptr
        g.normal = self.nrm.ptr
        g.roughness = self.rough.ptr if self.rough is not None else None
        g.width = width
        g.height = height
        self.desc = g
        self.num_pixels = width * height


class Scene:
    def __init__(self, ctx: "Context"):
        self.ctx = ctx
        h = C.c_void_p()
        _check(lib.vgi_scene_create(ctx.handle, C.byref(h)))
        self.handle = h.value
        self.grid = None

    def voxelize(self, meshes, grid: Grid, stream=None) -> None:
        arr = (Mesh * max(1, len(meshes)))()
        for i, m in enumerate(meshes):
            arr[i] = m.desc if isinstance(m, MeshData) else m
        _check(lib.vgi_voxelize(self.ctx.handle, self.handle, arr, len(meshes), C.byref(grid), stream))
        self.grid = grid

    def voxelize_noalloc(self, meshes, grid: Grid, stream=None) -> None:
        """vgi_voxelize_noalloc: the same pass with every per-call buffer in the library's
        static per-GPU scratch (the scene must have been voxelized once at this grid size)."""
        arr = (Mesh * max(1, len(meshes)))()
        for i, m in enumerate(meshes):
            arr[i] = m.desc if isinstance(m, MeshData) else m
        _check(lib.vgi_voxelize_noalloc(self.ctx.handle, self.handle, arr, len(meshes), C.byref(grid), stream))
        self.grid = grid

    def build_mips(self, lights: Lights, precision: int = PREC_FP32, stream=None) -> None:
        _check(lib.vgi_build_mips(self.ctx.handle, self.handle, C.byref(lights), precision, stream))

    def trace(self, gbuffer: GBufferData, cones: ConeParams, out_diffuse: DeviceArray,
              out_specular: DeviceArray | None = None, stream=None) -> None:
        _check(lib.vgi_trace(self.ctx.handle, self.handle, C.byref(gbuffer.desc), C.byref(cones),
                             out_diffuse.ptr, out_specular.ptr if out_specular is not None else None,
                             stream))

    def read_voxels(self) -> np.ndarray:
        n = self.grid.n
        out = np.empty((n, n, n, 8), np.uint32)  # [z][y][x][word]
        _check(lib.vgi_scene_read_voxels(self.ctx.handle, self.handle, out.ctypes.data))
        return out

    def read_level(self, level: int) -> np.ndarray:
        nk = self.grid.n >> level
        out = np.empty((nk, nk, nk, 4), np.float32)
        _check(lib.vgi_scene_read_level(self.ctx.handle, self.handle, level, out.ctypes.data))
        return out

    def stats(self) -> SceneStats:
        s = SceneStats()
        _check(lib.vgi_scene_get_stats(self.handle, C.byref(s)))
        return s

    def release(self) -> None:
        if self.handle:
            _check(lib.vgi_scene_release(self.handle))
            self.handle = None

    def __del__(self):
        try:
            if self.handle:
                lib.vgi_scene_release(self.handle)
        except Exception:
            pass


def _vox_items(jobs):
    items = (VoxelizeItem * max(1, len(jobs)))()
    keep = []
    for i, (scene, meshes, grid) in enumerate(jobs):
        arr = (Mesh * max(1, len(meshes)))()
        for k, m in enumerate(meshes):
            arr[k] = m.desc if isinstance(m, MeshData) else m
        keep.append(arr)
        items[i].scene = scene.handle
        items[i].meshes = C.cast(arr, C.c_void_p)
        items[i].num_meshes = len(meshes)
        items[i].grid = grid
    return items, keep


def _mips_items(jobs):
    items = (MipsItem * max(1, len(jobs)))()
    keep = []
    for i, (scene, lights) in enumerate(jobs):
        items[i].scene = scene.handle
        if lights is not None:
            keep.append(lights)
            items[i].lights = C.addressof(lights)
    return items, keep


def _trace_items(jobs):
    items = (TraceItem * max(1, len(jobs)))()
    for i, (scene, gb, cones, out_d, out_s) in enumerate(jobs):
        items[i].scene = scene.handle
        items[i].gbuffer = gb.desc
        items[i].cones = cones
        items[i].out_diffuse = out_d.ptr
        items[i].out_specular = out_s.ptr if out_s is not None else None
    return items


class Gltf:
    """A glTF 2.0 file read by the library (host side; no device needed until upload)."""

    def __init__(self, path: str | None = None, data: bytes | None = None, base_dir: str | None = None,
                 albedo_mode: int = GLTF_ALBEDO_REFERENCE):
        h = C.c_void_p()
        if path is not None:
            _check(lib.vgi_gltf_load_file(path.encode(), albedo_mode, C.byref(h)))
        else:
            buf = C.create_string_buffer(bytes(data), len(data))
            _check(lib.vgi_gltf_load_memory(buf, len(data), base_dir.encode() if base_dir else None,
                                            albedo_mode, C.byref(h)))
        self.handle = h.value

    def info(self) -> GltfInfo:
        i = GltfInfo()
        _check(lib.vgi_gltf_get_info(self.handle, C.byref(i)))
        return i

    def primitive(self, index: int) -> dict:
        """Copies of one placed primitive's arrays and attributes."""
        p = GltfPrimitive()
        _check(lib.vgi_gltf_get_primitive(self.handle, index, C.byref(p)))
        pos = np.ctypeslib.as_array(C.cast(p.positions, C.POINTER(C.c_float)), (p.num_vertices * 3,)).copy() \
            if p.num_vertices else np.zeros(0, np.float32)
        idx = np.ctypeslib.as_array(C.cast(p.indices, C.POINTER(C.c_uint32)), (p.num_triangles * 3,)).copy() \
            if p.num_triangles else np.zeros(0, np.uint32)
        uv = np.ctypeslib.as_array(C.cast(p.texcoords, C.POINTER(C.c_float)), (p.num_vertices * 2,)).copy() \
            if p.texcoords and p.num_vertices else None
        return dict(positions=pos.reshape(-1, 3), indices=idx.reshape(-1, 3), model=np.array(p.model[:], np.float32),
                    base_color=tuple(p.base_color[:]), node=p.node, mesh=p.mesh, primitive=p.primitive,
                    material=p.material, texcoords=None if uv is None else uv.reshape(-1, 2), image=p.image,
                    base_color_factor=tuple(p.base_color_factor[:]))

    def image(self, index: int) -> dict:
        """The image's encoded bytes as stored in the file, with its mime type and uri."""
        im = GltfImage()
        _check(lib.vgi_gltf_get_image(self.handle, index, C.byref(im)))
        data = C.string_at(im.data, im.bytes) if im.bytes else b""
        return dict(data=data, mime_type=(im.mime_type or b"").decode(), uri=(im.uri or b"").decode())

    def set_image_rgba8(self, index: int, rgba: np.ndarray) -> None:
        """Decoded texels (height, width, 4) uint8 for image `index`."""
        a = np.ascontiguousarray(rgba, np.uint8)
        if a.ndim != 3 or a.shape[2] != 4:
            raise ValueError("texels must be (height, width, 4) RGBA8")
        _check(lib.vgi_gltf_set_image_rgba8(self.handle, index, a.ctypes.data, a.shape[1], a.shape[0]))

    def primitives(self) -> list:
        return [self.primitive(i) for i in range(self.info().num_primitives)]

    def upload(self, ctx: "Context") -> list:
        """Device copies of the geometry and one vgi_mesh per placed primitive."""
        n = self.info().num_primitives
        arr = (Mesh * max(1, n))()
        cnt = C.c_uint32()
        _check(lib.vgi_gltf_upload(ctx.handle, self.handle, arr, n, C.byref(cnt)))
        return [arr[i] for i in range(cnt.value)]

    def close(self) -> None:
        if self.handle:
            _check(lib.vgi_gltf_release(self.handle))
            self.handle = None

    def __del__(self):
        try:
            if self.handle:
                lib.vgi_gltf_release(self.handle)
        except Exception:
            pass


class Context:
    def __init__(self, device: int = 0):
        h = C.c_void_p()
        _check(lib.vgi_create(C.byref(h), device))
        self.handle = h.value

    def close(self) -> None:
        if self.handle:
            _check(lib.vgi_destroy(self.handle))
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def scene(self) -> Scene:
        return Scene(self)

    def sync(self, stream=None) -> None:
        _check(lib.vgi_sync(self.handle, stream))

    def set_tuning(self, bin_size=8, chunk_tris=64, tile_w=16, tile_h=16, pixel_order=ORDER_TILES,
                   address_mode=ADDRESS_AUTO) -> None:
        t = Tuning(bin_size, chunk_tris, tile_w, tile_h, pixel_order, address_mode)
        _check(lib.vgi_set_tuning(self.handle, C.byref(t)))

    def put_tuning(self, t: Tuning) -> None:
        _check(lib.vgi_set_tuning(self.handle, C.byref(t)))

    def get_tuning(self) -> Tuning:
        t = Tuning()
        _check(lib.vgi_get_tuning(self.handle, C.byref(t)))
        return t

    def set_profiling(self, on: bool) -> None:
        _check(lib.vgi_set_profiling(self.handle, 1 if on else 0))

    def timings(self) -> Timings:
        t = Timings()
        _check(lib.vgi_get_timings(self.handle, C.byref(t)))
        return t

    def copy(self, dst: DeviceArray, src: DeviceArray, nbytes: int | None = None, stream=None) -> None:
        n = min(dst.nbytes, src.nbytes) if nbytes is None else nbytes
        _check(lib.vgi_memcpy_d2d(self.handle, dst.ptr, src.ptr, n, stream))

    def stream_create(self):
        s = C.c_void_p()
        _check(lib.vgi_stream_create(self.handle, C.byref(s)))
        return s.value

    def stream_destroy(self, s) -> None:
        _check(lib.vgi_stream_destroy(self.handle, s))

    def stream_wait(self, waiting, signalling) -> None:
        """Later work on `waiting` waits for the work enqueued on `signalling` so far."""
        _check(lib.vgi_stream_wait(self.handle, waiting, signalling))

    def stream_busy(self, stream=None) -> bool:
        """True while work enqueued on the stream has not completed (does not block)."""
        rc = lib.vgi_stream_query(self.handle, stream)
        if rc == -5:
            return True
        _check(rc)
        return False

    def probe(self, kind: int, buf: "DeviceArray", blocks: int, iters: int, stream=None) -> None:
        """One calibration kernel (PROBE_FMA, PROBE_ATOMIC_RETURN, PROBE_ATOMIC_NORETURN)."""
        _check(lib.vgi_probe(self.handle, kind, buf.ptr, buf.nbytes, blocks, iters, stream))

    def workspace_bytes(self) -> int:
        b = C.c_uint64()
        _check(lib.vgi_workspace_bytes(self.handle, C.byref(b)))
        return int(b.value)

    def voxelize_batch(self, jobs, stream=None) -> None:
        """jobs: sequence of (Scene, [MeshData | Mesh], Grid); one batched voxelize pass."""
        items, keep = _vox_items(jobs)
        _check(lib.vgi_voxelize_batch(self.handle, items, len(jobs), stream))
        for scene, _, grid in jobs:
            scene.grid = grid

    def build_mips_batch(self, jobs, precision: int = PREC_FP32, stream=None) -> None:
        """jobs: sequence of (Scene, Lights | None)."""
        items, keep = _mips_items(jobs)
        _check(lib.vgi_build_mips_batch(self.handle, items, len(jobs), precision, stream))

    def trace_batch(self, jobs, stream=None) -> None:
        """jobs: sequence of (Scene, GBufferData, ConeParams, out_diffuse, out_specular | None)."""
        _check(lib.vgi_trace_batch(self.handle, _trace_items(jobs), len(jobs), stream))

    def autotune_batch(self, vox_jobs, mips_jobs, trace_jobs, precision: int, reps: int = 3,
                       stream=None) -> AutotuneReport:
        """Times the batch under every bin size and screen tile; keeps the best in the
        context's tuning. The three job lists name the same scenes in the same order."""
        vi, k1 = _vox_items(vox_jobs)
        mi, k2 = _mips_items(mips_jobs)
        ti = _trace_items(trace_jobs)
        r = AutotuneReport()
        _check(lib.vgi_autotune_batch(self.handle, vi, mi, ti, len(vox_jobs), precision, reps,
                                      C.byref(r), stream))
        for scene, _, grid in vox_jobs:
            scene.grid = grid
        return r

    # ---- screen-space passes (device buffers in, device buffers out) ----
    def gbuffer_decode(self, depth: DeviceArray, normals: DeviceArray, inv_proj, inv_model_view,
                       width: int, height: int, out_position: DeviceArray, out_normal: DeviceArray,
                       out_view_depth: DeviceArray | None = None, stream=None) -> None:
        """inv_proj / inv_model_view: 4x4 arrays as GLSL stores them (column-major, 16 floats)."""
        g = GBufferRef()
        g.depth, g.normals = depth.ptr, normals.ptr
        g.inv_proj[:] = [float(v) for v in np.asarray(inv_proj, np.float32).reshape(16)]
        g.inv_model_view[:] = [float(v) for v in np.asarray(inv_model_view, np.float32).reshape(16)]
        g.width, g.height = width, height
        _check(lib.vgi_gbuffer_decode(self.handle, C.byref(g), out_position.ptr, out_normal.ptr,
                                      out_view_depth.ptr if out_view_depth is not None else None,
                                      stream))

    def temporal_resolve(self, indirect: DeviceArray, position: DeviceArray, view_depth: DeviceArray,
                         history: DeviceArray, out: DeviceArray, prev_projection, prev_model_view,
                         width: int, height: int, history_format: int = IMAGE_RGBA32F,
                         out_format: int = IMAGE_RGBA32F, stream=None) -> None:
        t = TemporalParams()
        t.indirect, t.position, t.view_depth = indirect.ptr, position.ptr, view_depth.ptr
        t.history, t.out = history.ptr, out.ptr
        t.history_format, t.out_format = history_format, out_format
        t.prev_projection[:] = [float(v) for v in np.asarray(prev_projection, np.float32).reshape(16)]
        t.prev_model_view[:] = [float(v) for v in np.asarray(prev_model_view, np.float32).reshape(16)]
        t.width, t.height = width, height
        _check(lib.vgi_temporal_resolve(self.handle, C.byref(t), stream))

    def bilateral_blur(self, color: DeviceArray, depth: DeviceArray, out: DeviceArray, width: int,
                       height: int, axis: int, color_format: int = IMAGE_RGBA32F,
                       out_format: int = IMAGE_RGBA32F, stream=None) -> None:
        b = BlurParams(color.ptr, depth.ptr, out.ptr, color_format, out_format, width, height, axis)
        _check(lib.vgi_bilateral_blur(self.handle, C.byref(b), stream))

    def timer_start(self, stream=None) -> None:
        _check(lib.vgi_timer_start(self.handle, stream))

    def timer_stop(self, stream=None) -> float:
        ms = C.c_float()
        _check(lib.vgi_timer_stop(self.handle, stream, C.byref(ms)))
        return ms.value


class Server:
    """Continuous-batching request loop over resident scenes (vgi_server_* in include/vgi.h)."""

    def __init__(self, ctx: "Context", memory_budget: int, precision: int = PREC_FP16, slots: int = 4,
                 max_batch_requests: int = 16, max_batch_pixels: int = 0):
        self.ctx = ctx
        cfg = ServerConfig(memory_budget, precision, slots, max_batch_requests, max_batch_pixels)
        h = C.c_void_p()
        _check(lib.vgi_server_create(ctx.handle, C.byref(cfg), C.byref(h)))
        self.handle = h.value
        self._buf = (Completion * 256)()

    def submit(self, scene_key: int, meshes, grid: Grid, lights: Lights, frames, tag: int,
               revoxelize: bool = False, arrival_s: float = 0.0) -> None:
        """frames: sequence of (GBufferData | GBuffer, ConeParams, out_diffuse, out_specular)."""
        marr = (Mesh * max(1, len(meshes)))()
        for k, m in enumerate(meshes):
            marr[k] = m.desc if isinstance(m, MeshData) else m
        farr = (Frame * max(1, len(frames)))()
        for k, (gb, cones, od, os_) in enumerate(frames):
            farr[k].gbuffer = gb.desc if isinstance(gb, GBufferData) else gb
            farr[k].cones = cones
            farr[k].out_diffuse = od.ptr if isinstance(od, DeviceArray) else od
            farr[k].out_specular = (os_.ptr if isinstance(os_, DeviceArray) else os_) if os_ is not None \
                else None
        r = Request()
        r.scene_key = scene_key
        r.meshes = C.cast(marr, C.c_void_p)
        r.num_meshes = len(meshes)
        r.revoxelize = 1 if revoxelize else 0
        r.grid = grid
        r.lights = lights
        r.frames = C.cast(farr, C.c_void_p)
        r.num_frames = len(frames)
        r.arrival_s = arrival_s
        r.tag = tag
        _check(lib.vgi_server_submit(self.handle, C.byref(r)))  # the server copies both tables

    def poll(self, wait_us: int = 0) -> list:
        n = C.c_uint32()
        _check(lib.vgi_server_poll(self.handle, self._buf, len(self._buf), C.byref(n), wait_us))
        return [Completion.from_buffer_copy(self._buf[i]) for i in range(n.value)]

    def drain(self) -> None:
        _check(lib.vgi_server_drain(self.handle))

    def stats(self) -> ServerStats:
        s = ServerStats()
        _check(lib.vgi_server_get_stats(self.handle, C.byref(s)))
        return s

    def close(self) -> None:
        if self.handle:
            _check(lib.vgi_server_destroy(self.handle))
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def now() -> float:
    return float(lib.vgi_now())


def scene_bytes_estimate(grid: Grid, precision: int, num_triangles: int) -> int:
    return int(lib.vgi_scene_bytes_estimate(C.byref(grid), precision, num_triangles))
