"""ctypes wrapper of the CPU oracle (oracle/_build/libvgi_oracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg as the checker. Nothing in final184_amd/ imports this module.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_DIR, "_build", "libvgi_oracle.so")


class OrcVoxStats(C.Structure):
    _fields_ = [("triangles_in", C.c_uint64), ("rejected", C.c_uint64),
                ("degenerate", C.c_uint64), ("voxel_hits", C.c_uint64)]


class OrcLight(C.Structure):
    _fields_ = [("type", C.c_uint32), ("vec", C.c_float * 3), ("color", C.c_float * 3),
                ("min_distance", C.c_float)]


class OrcCones(C.Structure):
    _fields_ = [("num_diffuse", C.c_uint32), ("specular", C.c_uint32),
                ("diffuse_aperture_deg", C.c_float), ("diffuse_ring_deg", C.c_float),
                ("specular_aperture_deg", C.c_float), ("step_factor", C.c_float),
                ("max_distance", C.c_float), ("alpha_cutoff", C.c_float),
                ("normal_offset", C.c_float), ("start_offset", C.c_float), ("eye", C.c_float * 3)]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} missing: run __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    P = C.c_void_p
    sig = {
        "orc_num_threads": (C.c_int, []),
        "orc_set_threads": (None, [C.c_int]),
        "orc_tri_box_overlap": (C.c_int, [P, P, C.c_int64]),
        "orc_triangle_setup": (C.c_int, [P, C.c_uint32, P, C.c_uint32, P, P, C.c_float, C.c_uint32, P, P]),
        "orc_voxelize": (C.c_int, [P, C.c_uint32, P, C.c_uint32, P, P, P, P, C.c_float, C.c_uint32, P, P]),
        "orc_voxelize_textured": (C.c_int, [P, C.c_uint32, P, C.c_uint32, P, P, C.c_uint32, C.c_uint32, P, P, P,
                                            C.c_float, C.c_uint32, P, P]),
        "orc_float_to_half": (C.c_uint16, [C.c_float]),
        "orc_half_to_float": (C.c_float, [C.c_uint16]),
        "orc_float_to_bf16": (C.c_uint16, [C.c_float]),
        "orc_bf16_to_float": (C.c_float, [C.c_uint16]),
        "orc_inject_mips_f32": (C.c_int, [P, P, C.c_float, C.c_uint32, C.c_uint32, P, C.c_uint32, C.c_int, P]),
        "orc_inject_mips_f64": (C.c_int, [P, P, C.c_float, C.c_uint32, C.c_uint32, P, C.c_uint32, P]),
        "orc_light_visibility": (C.c_int, [P, P, C.c_float, C.c_uint32, P, C.c_uint32, P]),
        "orc_trace_f32": (C.c_int, [P, P, C.c_float, C.c_uint32, C.c_uint32, P, P, P, P, C.c_uint64, P, P, P, P]),
        "orc_trace_f64": (C.c_int, [P, P, C.c_float, C.c_uint32, C.c_uint32, P, P, P, P, C.c_uint64, P, P, P, P]),
        "orc_cone_schedule_f64": (C.c_int, [C.c_double] * 5 + [P, C.c_int]),
        "orc_ref_orientation": (C.c_int, [C.c_float] * 3),
        "orc_ref_pack_color565": (C.c_uint32, [C.c_float] * 3),
        "orc_ref_pack_normal565": (C.c_uint32, [C.c_float] * 3),
        "orc_ref_unpack_color565": (None, [C.c_uint32, P]),
        "orc_ref_unpack_normal565": (None, [C.c_uint32, P]),
        "orc_ref_hemisphere_cos": (None, [C.c_float, C.c_float, P]),
        "orc_ref_make_coord_space": (None, [P, P]),
        "orc_ref_get_indirect": (C.c_int, [P, P, P, P, C.c_float, C.c_float, P, P, P, P, P]),
        "orc_gbuffer_decode_f32": (None, [P, P, P, P, C.c_uint32, C.c_uint32, P, P, P]),
        "orc_gbuffer_decode_f64": (None, [P, P, P, P, C.c_uint32, C.c_uint32, P, P, P]),
        "orc_temporal": (None, [P, P, P, P, P, P, C.c_uint32, C.c_uint32, P]),
        "orc_exp2_poly": (C.c_float, [C.c_float]),
        "orc_bilateral_blur": (None, [P, P, C.c_uint32, C.c_uint32, C.c_uint32, P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return L


lib = _load()

IDENTITY_3X4 = np.array([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0], np.float32)


def _p(a):
    return a.ctypes.data if a is not None else None


def set_threads(n: int) -> None:
    lib.orc_set_threads(int(n))


def num_threads() -> int:
    return int(lib.orc_num_threads())


def voxelize(positions, grid, albedo=None, base_rgb=(255, 255, 255), indices=None,
             model=None, num_vertices=None):
    """Accumulator (n, n, n, 8) uint32 indexed [z][y][x][word], and stats."""
    pos = np.ascontiguousarray(positions, np.float32).reshape(-1)
    nv = pos.size // 3 if num_vertices is None else int(num_vertices)
    idx = None if indices is None else np.ascontiguousarray(indices, np.uint32).reshape(-1)
    ntri = (idx.size // 3) if idx is not None else nv // 3
    alb = None if albedo is None else np.ascontiguousarray(albedo, np.uint8).reshape(-1)
    base = np.array(list(base_rgb)[:3], np.uint8)
    m = IDENTITY_3X4 if model is None else np.ascontiguousarray(model, np.float32).reshape(12)
    origin = np.array(grid["origin"], np.float32)
    n = int(grid["n"])
    acc = np.empty((n, n, n, 8), np.uint32)
    st = OrcVoxStats()
    rc = lib.orc_voxelize(_p(pos), nv, _p(idx), ntri, _p(alb), _p(base), _p(m), _p(origin),
                          float(grid["extent"]), n, _p(acc), C.byref(st))
    assert rc == 0
    return acc, st


def voxelize_textured(positions, grid, uv, texels, factor=(1.0, 1.0, 1.0, 1.0), indices=None, model=None,
                      num_vertices=None, acc=None, stats=None):
    """Adds a textured mesh into acc ((n, n, n, 8) uint32 [z][y][x][word], zeros when None).
    uv: (num_vertices, 2) float32; texels: (h, w, 4) uint8 RGBA, row 0 first."""
    pos = np.ascontiguousarray(positions, np.float32).reshape(-1)
    nv = pos.size // 3 if num_vertices is None else int(num_vertices)
    idx = None if indices is None else np.ascontiguousarray(indices, np.uint32).reshape(-1)
    ntri = (idx.size // 3) if idx is not None else nv // 3
    uvs = np.ascontiguousarray(uv, np.float32).reshape(-1)
    assert uvs.size >= 2 * nv
    tx = np.ascontiguousarray(texels, np.uint8)
    assert tx.ndim == 3 and tx.shape[2] == 4
    fac = np.array(list(factor), np.float32)
    m = IDENTITY_3X4 if model is None else np.ascontiguousarray(model, np.float32).reshape(12)
    origin = np.array(grid["origin"], np.float32)
    n = int(grid["n"])
    if acc is None:
        acc = np.zeros((n, n, n, 8), np.uint32)
    assert acc.dtype == np.uint32 and acc.flags.c_contiguous and acc.shape == (n, n, n, 8)
    st = stats if stats is not None else OrcVoxStats()
    rc = lib.orc_voxelize_textured(_p(pos), nv, _p(idx), ntri, _p(uvs), _p(tx), tx.shape[1], tx.shape[0],
                                   _p(fac), _p(m), _p(origin), float(grid["extent"]), n, _p(acc), C.byref(st))
    assert rc == 0
    return acc, st


def triangle_setup(positions, grid, tri, indices=None, model=None):
    pos = np.ascontiguousarray(positions, np.float32).reshape(-1)
    idx = None if indices is None else np.ascontiguousarray(indices, np.uint32).reshape(-1)
    m = IDENTITY_3X4 if model is None else np.ascontiguousarray(model, np.float32).reshape(12)
    origin = np.array(grid["origin"], np.float32)
    q = np.empty(9, np.int64)
    nq = np.empty(3, np.int32)
    rc = lib.orc_triangle_setup(_p(pos), pos.size // 3, _p(idx), tri, _p(m), _p(origin),
                                float(grid["extent"]), int(grid["n"]), _p(q), _p(nq))
    return rc, q.reshape(3, 3), nq


def tri_box_overlap(q, b, delta) -> bool:
    q = np.ascontiguousarray(q, np.int64).reshape(9)
    b = np.ascontiguousarray(b, np.int64).reshape(3)
    return bool(lib.orc_tri_box_overlap(_p(q), _p(b), int(delta)))


LIGHT_SHADOW = 0x100  # orc_light.type flag: the light is cut where the visibility walk meets a voxel


def _lights(lights):
    arr = (OrcLight * max(1, len(lights)))()
    for i, d in enumerate(lights):
        arr[i].type = (0 if d["type"] == "directional" else 1) | (LIGHT_SHADOW if d.get("shadow") else 0)
        arr[i].vec[:] = [float(v) for v in d["vec"]]
        arr[i].color[:] = [float(v) for v in d["color"]]
        arr[i].min_distance = float(d.get("min_distance", 0.0))
    return arr


PREC = {"fp32": 0, "fp16": 1, "bf16": 2}


def inject_mips(acc, grid, lights, prec="fp32", f64=False):
    """List of levels, level k as (n_k, n_k, n_k, 4) [z][y][x][rgba]."""
    n, L = int(grid["n"]), int(grid["levels"])
    total = sum((n >> k) ** 3 for k in range(L)) * 4
    acc = np.ascontiguousarray(acc, np.uint32)
    origin = np.array(grid["origin"], np.float32)
    la = _lights(lights)
    if f64:
        out = np.empty(total, np.float64)
        rc = lib.orc_inject_mips_f64(_p(acc), _p(origin), float(grid["extent"]), n, L, la,
                                     len(lights), _p(out))
    else:
        out = np.empty(total, np.float32)
        rc = lib.orc_inject_mips_f32(_p(acc), _p(origin), float(grid["extent"]), n, L, la,
                                     len(lights), PREC[prec] if isinstance(prec, str) else prec, _p(out))
    assert rc == 0
    levels, off = [], 0
    for k in range(L):
        nk = n >> k
        levels.append(out[off:off + nk ** 3 * 4].reshape(nk, nk, nk, 4))
        off += nk ** 3 * 4
    return levels


def light_visibility(acc, grid, lights):
    """(n, n, n) [z][y][x] uint32: bit l set where light l is visible from an occupied voxel."""
    n = int(grid["n"])
    acc = np.ascontiguousarray(acc, np.uint32)
    origin = np.array(grid["origin"], np.float32)
    out = np.empty(n ** 3, np.uint32)
    rc = lib.orc_light_visibility(_p(acc), _p(origin), float(grid["extent"]), n, _lights(lights),
                                  len(lights), _p(out))
    assert rc == 0
    return out.reshape(n, n, n)


def _cones(c):
    o = OrcCones()
    for k, _ in OrcCones._fields_:
        if k == "eye":
            o.eye[:] = [float(v) for v in c["eye"]]
        elif k == "specular":
            o.specular = 1 if c["specular"] else 0
        else:
            setattr(o, k, c[k])
    return o


def trace(levels, grid, gb_pos, gb_nrm, cones, gb_rough=None, pixels=None, f64=False,
          want_specular=True):
    """(diffuse (P,4), specular (P,4) or None, trilinear sample count)."""
    dt = np.float64 if f64 else np.float32
    flat = np.concatenate([np.ascontiguousarray(l, dt).reshape(-1) for l in levels])
    gp = np.ascontiguousarray(gb_pos, np.float32)
    gn = np.ascontiguousarray(gb_nrm, np.float32)
    gr = None if gb_rough is None else np.ascontiguousarray(gb_rough, np.float32)
    px = None if pixels is None else np.ascontiguousarray(pixels, np.int64)
    npx = gp.shape[0] if px is None else px.size
    od = np.empty((npx, 4), dt)
    os_ = np.empty((npx, 4), dt) if want_specular else None
    cnt = C.c_uint64()
    origin = np.array(grid["origin"], np.float32)
    fn = lib.orc_trace_f64 if f64 else lib.orc_trace_f32
    rc = fn(_p(flat), _p(origin), float(grid["extent"]), int(grid["n"]), int(grid["levels"]),
            _p(gp), _p(gn), _p(gr), _p(px), npx, C.byref(_cones(cones)), _p(od), _p(os_),
            C.byref(cnt))
    assert rc == 0
    return od, os_, int(cnt.value)


def cone_schedule(h, tan_half, step, start, max_d, cap=100000):
    out = np.empty(cap, np.float64)
    m = lib.orc_cone_schedule_f64(h, tan_half, step, start, max_d, _p(out), cap)
    return out[:m].copy()


def float_to_half_bits(x: float) -> int:
    return int(lib.orc_float_to_half(float(x)))


def half_bits_to_float(b: int) -> float:
    return float(lib.orc_half_to_float(int(b)))


def float_to_bf16_bits(x: float) -> int:
    return int(lib.orc_float_to_bf16(float(x)))


def bf16_bits_to_float(b: int) -> float:
    return float(lib.orc_bf16_to_float(int(b)))


def ref_orientation(n) -> int:
    return int(lib.orc_ref_orientation(float(n[0]), float(n[1]), float(n[2])))


def ref_pack_color565(r, g, b) -> int:
    return int(lib.orc_ref_pack_color565(r, g, b))


def ref_pack_normal565(x, y, z) -> int:
    return int(lib.orc_ref_pack_normal565(x, y, z))


def ref_unpack_color565(v):
    o = np.empty(3, np.float32)
    lib.orc_ref_unpack_color565(int(v), _p(o))
    return o


def ref_unpack_normal565(v):
    o = np.empty(3, np.float32)
    lib.orc_ref_unpack_normal565(int(v), _p(o))
    return o


def ref_hemisphere_cos(u, v):
    o = np.empty(3, np.float32)
    lib.orc_ref_hemisphere_cos(u, v, _p(o))
    return o


def ref_make_coord_space(n):
    nn = np.ascontiguousarray(n, np.float32)
    o = np.empty(9, np.float32)
    lib.orc_ref_make_coord_space(_p(nn), _p(o))
    return o.reshape(3, 3)  # rows = x, y, z axes


def ref_get_indirect(voxels_rg16, w2voxel_colmajor, wpos, wnorm, r0, r1, sun_dir, sun_lum):
    vox = np.ascontiguousarray(voxels_rg16, np.uint16)
    m = np.ascontiguousarray(w2voxel_colmajor, np.float32).reshape(16)
    wp = np.ascontiguousarray(wpos, np.float32)
    wn = np.ascontiguousarray(wnorm, np.float32)
    sd = np.ascontiguousarray(sun_dir, np.float32)
    sl = np.ascontiguousarray(sun_lum, np.float32)
    Lo = np.empty(3, np.float32)
    hp = np.zeros(3, np.float32)
    steps = C.c_int()
    hit = lib.orc_ref_get_indirect(_p(vox), _p(m), _p(wp), _p(wn), r0, r1, _p(sd), _p(sl), _p(Lo),
                                   _p(hp), C.byref(steps))
    return bool(hit), Lo, hp, int(steps.value)


def gbuffer_decode(depth, normals, inv_proj, inv_mv, width, height, f64=False):
    """World position, normal (float32 x 3 per pixel) and view depth (-z) from the reference's
    G-buffer: D32 depth, RGBA16 UNORM view-space normals, column-major InvProj and
    InvModelView (indirect.frag:44-58, 189-196)."""
    npx = width * height
    depth = np.ascontiguousarray(depth, np.float32).reshape(npx)
    normals = np.ascontiguousarray(normals, np.uint16).reshape(npx, 4)
    ip = np.ascontiguousarray(inv_proj, np.float32).reshape(16)
    im = np.ascontiguousarray(inv_mv, np.float32).reshape(16)
    pos = np.empty((npx, 3), np.float32)
    nrm = np.empty((npx, 3), np.float32)
    vd = np.empty(npx, np.float32)
    fn = lib.orc_gbuffer_decode_f64 if f64 else lib.orc_gbuffer_decode_f32
    fn(_p(depth), _p(normals), _p(ip), _p(im), width, height, _p(pos), _p(nrm), _p(vd))
    return pos, nrm, vd


def temporal(indirect, position, view_depth, history, prev_proj, prev_mv, width, height):
    """Temporal blend of one frame (indirect.frag:224-242): float32 (H*W, 4) out, rgb =
    blended colour, a = view depth. history is float32 (H*W, 4) (an FP16 history is
    widened by the caller, exactly)."""
    npx = width * height
    ind = np.ascontiguousarray(indirect, np.float32).reshape(npx, 4)
    pos = np.ascontiguousarray(position, np.float32).reshape(npx, 3)
    vd = np.ascontiguousarray(view_depth, np.float32).reshape(npx)
    hist = np.ascontiguousarray(history, np.float32).reshape(npx, 4)
    pp = np.ascontiguousarray(prev_proj, np.float32).reshape(16)
    pm = np.ascontiguousarray(prev_mv, np.float32).reshape(16)
    out = np.empty((npx, 4), np.float32)
    lib.orc_temporal(_p(ind), _p(pos), _p(vd), _p(hist), _p(pp), _p(pm), width, height, _p(out))
    return out


def exp2_poly(x):
    return float(lib.orc_exp2_poly(float(x)))


def bilateral_blur(color, depth, width, height, axis):
    """One cross-bilateral blur pass (bilateralBlur.inc:15-72); axis 0 along x (blurY.frag),
    1 along y (blurX.frag). float32 (H*W, 4) in (rgb used) and out (a = 0)."""
    npx = width * height
    col = np.ascontiguousarray(color, np.float32).reshape(npx, 4)
    dep = np.ascontiguousarray(depth, np.float32).reshape(npx)
    out = np.empty((npx, 4), np.float32)
    lib.orc_bilateral_blur(_p(col), _p(dep), width, height, axis, _p(out))
    return out
