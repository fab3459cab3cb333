"""C ABI checks that need no GPU: the library loads, exports every entry point that
include/vgi.h declares, the ctypes mirrors match the C struct layouts (checked by a
small C program compiled against the header), and entry points report errors through
status codes and vgi_last_error without a device."""
import ctypes as C
import os
import re
import subprocess

import pytest

from final184_amd import vgi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "vgi.h")
HAS_GPU = os.path.exists("/dev/kfd")


def declared_functions():
    text = open(HEADER).read()
    return re.findall(r"^\s*(?:const\s+char\s*\*|int|uint32_t|uint64_t|double)\s+(vgi_\w+)\s*\(", text, re.M)


def test_every_declared_symbol_is_exported_and_bound():
    names = declared_functions()
    assert len(names) == len(vgi.SIGNATURES)
    bound = {n for n, _, _ in vgi.SIGNATURES}
    for n in names:
        assert hasattr(vgi.lib, n), f"{n} not exported by libvgi.so"
        assert n in bound, f"{n} has no ctypes signature"
    assert set(bound) <= set(names)


STRUCTS = {
    "vgi_grid": (vgi.Grid, ["origin", "extent", "n", "levels"]),
    "vgi_mesh": (vgi.Mesh, ["positions", "position_stride", "num_vertices", "indices", "index_type",
                            "num_triangles", "albedo_rgb8", "base_color", "model"]),
    "vgi_light": (vgi.Light, ["type", "vec", "color", "min_distance"]),
    "vgi_lights": (vgi.Lights, ["count", "light"]),
    "vgi_cone_params": (vgi.ConeParams, [f for f, _ in vgi.ConeParams._fields_]),
    "vgi_gbuffer": (vgi.GBuffer, [f for f, _ in vgi.GBuffer._fields_]),
    "vgi_tuning": (vgi.Tuning, [f for f, _ in vgi.Tuning._fields_]),
    "vgi_scene_stats": (vgi.SceneStats, [f for f, _ in vgi.SceneStats._fields_]),
    "vgi_timings": (vgi.Timings, [f for f, _ in vgi.Timings._fields_]),
    "vgi_voxelize_item": (vgi.VoxelizeItem, [f for f, _ in vgi.VoxelizeItem._fields_]),
    "vgi_mips_item": (vgi.MipsItem, [f for f, _ in vgi.MipsItem._fields_]),
    "vgi_trace_item": (vgi.TraceItem, [f for f, _ in vgi.TraceItem._fields_]),
    "vgi_autotune_report": (vgi.AutotuneReport, [f for f, _ in vgi.AutotuneReport._fields_]),
    "vgi_server_config": (vgi.ServerConfig, [f for f, _ in vgi.ServerConfig._fields_]),
    "vgi_frame": (vgi.Frame, [f for f, _ in vgi.Frame._fields_]),
    "vgi_request": (vgi.Request, [f for f, _ in vgi.Request._fields_]),
    "vgi_completion": (vgi.Completion, [f for f, _ in vgi.Completion._fields_]),
    "vgi_server_stats": (vgi.ServerStats, [f for f, _ in vgi.ServerStats._fields_]),
    "vgi_gbuffer_ref": (vgi.GBufferRef, [f for f, _ in vgi.GBufferRef._fields_]),
    "vgi_temporal_params": (vgi.TemporalParams, [f for f, _ in vgi.TemporalParams._fields_]),
    "vgi_blur_params": (vgi.BlurParams, [f for f, _ in vgi.BlurParams._fields_]),
    "vgi_gltf_info": (vgi.GltfInfo, [f for f, _ in vgi.GltfInfo._fields_]),
    "vgi_gltf_primitive": (vgi.GltfPrimitive, [f for f, _ in vgi.GltfPrimitive._fields_]),
}


def test_struct_layouts_match_c(tmp_path):
    lines = ["#include <stdio.h>", "#include <stddef.h>", f'#include "{HEADER}"', "int main(void){"]
    for cname, (_, fields) in STRUCTS.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f in fields:
            lines.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", str(src), "-o", str(exe)], check=True)
    out = dict(l.rsplit(" ", 1) for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                        check=True).stdout.split("\n") if l)
    for cname, (cls, fields) in STRUCTS.items():
        assert int(out[cname]) == C.sizeof(cls), cname
        for f in fields:
            assert int(out[f"{cname}.{f}"]) == getattr(cls, f).offset, f"{cname}.{f}"


def test_version_and_abi():
    assert vgi.lib.vgi_abi_version() == 3
    assert b"gfx950" in vgi.lib.vgi_version()


@pytest.mark.skipif(HAS_GPU, reason="checks the no-device path")
def test_create_without_device_reports_nodevice():
    h = C.c_void_p()
    rc = vgi.lib.vgi_create(C.byref(h), 0)
    assert rc == -7  # VGI_E_NODEVICE
    assert h.value is None
    assert len(vgi.lib.vgi_last_error()) > 0
    with pytest.raises(vgi.VgiError):
        vgi.Context(0)


def test_null_arguments_are_rejected_with_a_message():
    assert vgi.lib.vgi_create(None, 0) == -1
    assert b"NULL" in vgi.lib.vgi_last_error()
    assert vgi.lib.vgi_destroy(None) == 0
    assert vgi.lib.vgi_scene_release(None) == 0
    assert vgi.lib.vgi_voxelize(None, None, None, 0, None, None) == -1
    assert vgi.lib.vgi_build_mips(None, None, None, 0, None) == -1
    assert vgi.lib.vgi_trace(None, None, None, None, None, None, None) == -1
    assert vgi.lib.vgi_scene_get_stats(None, None) == -1
    assert vgi.lib.vgi_voxelize_batch(None, None, 0, None) == -1
    assert vgi.lib.vgi_build_mips_batch(None, None, 0, 0, None) == -1
    assert vgi.lib.vgi_trace_batch(None, None, 0, None) == -1
    assert vgi.lib.vgi_autotune_batch(None, None, None, None, 0, 0, 1, None, None) == -1
    assert vgi.lib.vgi_server_create(None, None, None) == -1
    assert vgi.lib.vgi_server_submit(None, None) == -1
    assert vgi.lib.vgi_gbuffer_decode(None, None, None, None, None, None) == -1
    assert vgi.lib.vgi_temporal_resolve(None, None, None) == -1
    assert vgi.lib.vgi_bilateral_blur(None, None, None) == -1
    assert vgi.lib.vgi_gltf_load_file(None, 0, None) == -1
    assert vgi.lib.vgi_gltf_load_memory(None, 0, None, 0, None) == -1
    assert vgi.lib.vgi_gltf_get_info(None, None) == -1
    assert vgi.lib.vgi_gltf_upload(None, None, None, 0, None) == -1
    assert vgi.lib.vgi_gltf_release(None) == -1
    assert vgi.lib.vgi_server_destroy(None) == 0
    assert vgi.lib.vgi_now() > 0


def test_bytes_estimate_scales_with_grid_and_precision():
    g256 = vgi.make_grid((0, 0, 0), 1.0, 256, 8)
    g128 = vgi.make_grid((0, 0, 0), 1.0, 128, 7)
    b32 = vgi.scene_bytes_estimate(g256, vgi.PREC_FP32, 262144)
    b16 = vgi.scene_bytes_estimate(g256, vgi.PREC_FP16, 262144)
    assert b32 > b16 > vgi.scene_bytes_estimate(g128, vgi.PREC_FP16, 262144)
    # accumulator (32 B/voxel) + padded fp16 chain (8 B/texel) at 256^3
    assert b16 >= 256 ** 3 * 32 + 258 ** 3 * 8
    assert vgi.scene_bytes_estimate(vgi.make_grid((0, 0, 0), 1.0, 100, 3), 0, 10) == 0
