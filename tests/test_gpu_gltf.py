"""A glTF scene voxelized on the GPU equals the oracle's voxelization of its primitives.

The reader's primitives (object-space vertices, u32 indices, node model matrix, base
colour) are uploaded with vgi_gltf_upload and voxelized in one vgi_voxelize call; the
oracle voxelizes each placed primitive with the same model matrix and colour, and the
integer accumulators must sum to the GPU's, word for word. Both albedo modes.
"""
import numpy as np
import pytest

from oracle import oracle as orc
from tests import gltf_writer as gw

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", [0, 1])
def test_gltf_scene_voxelizes_like_the_oracle(ctx, tmp_path, mode):
    from final184_amd import vgi
    print(f"case gltf voxelize mode={mode}", flush=True)
    w, _ = gw.kitchen_sink()
    path = str(tmp_path / "scene.glb")
    w.write_glb(path)
    g = vgi.Gltf(path, albedo_mode=mode)
    meshes = g.upload(ctx)
    prims = g.primitives()
    assert len(meshes) == len(prims) == g.info().num_primitives
    grid = dict(origin=(-1.5, -1.5, -1.5), extent=3.0, n=64, levels=3)
    scene = ctx.scene()
    scene.voxelize(meshes, vgi.make_grid(**grid))
    got = scene.read_voxels()
    ref = np.zeros_like(got, dtype=np.uint64)
    for p in prims:
        acc, _ = orc.voxelize(p["positions"], grid, base_rgb=p["base_color"][:3], indices=p["indices"],
                              model=p["model"])
        ref += acc
    assert np.array_equal(got.astype(np.uint64), ref)
    occupied = int((got[..., 3] > 0).sum())
    print(f"occupied voxels {occupied}", flush=True)
    assert occupied > 200
    # a second upload with the same context reuses the device copies
    again = g.upload(ctx)
    assert [m.positions for m in again] == [m.positions for m in meshes]
    g.close()


@pytest.mark.parametrize("mode", [0, 1])
def test_textured_gltf_scene_voxelizes_like_the_oracle(ctx, tmp_path, mode):
    """Textured primitives: the caller decodes each image (here the test's own PNG decoder)
    and hands the texels to the handle; the upload then makes those primitives textured
    meshes. Each textured primitive must voxelize as the oracle's textured path with its
    TEXCOORD_0, its image (the texture's index in the reference mode, its source in the
    factor mode) and its factor; the untextured one with its base colour. An image left
    undecoded leaves its primitives untextured."""
    from final184_amd import vgi
    print(f"case textured gltf voxelize mode={mode}", flush=True)
    w, texels, _ = gw.textured_scene(file_dir=str(tmp_path))
    path = str(tmp_path / "tex.gltf")
    w.write_gltf(path)
    g = vgi.Gltf(path, albedo_mode=mode)
    prims = g.primitives()
    decoded = {}
    for k in range(g.info().num_images):
        if k == 2 and mode == 1:
            continue  # left undecoded: its primitive keeps the base colour
        decoded[k] = gw.png_decode_rgba(g.image(k)["data"])
        g.set_image_rgba8(k, decoded[k])
    meshes = g.upload(ctx)
    grid = dict(origin=(-1.0, -1.0, -1.0), extent=2.0, n=64, levels=3)
    scene = ctx.scene()
    try:
        scene.voxelize(meshes, vgi.make_grid(**grid))
        got = scene.read_voxels()
    finally:
        scene.release()
    ref = np.zeros(got.shape, np.uint32)
    textured = 0
    for p in prims:
        if p["texcoords"] is not None and p["image"] in decoded:
            textured += 1
            orc.voxelize_textured(p["positions"], grid, p["texcoords"], decoded[p["image"]],
                                  factor=p["base_color_factor"], indices=p["indices"], model=p["model"], acc=ref)
        else:
            acc, _ = orc.voxelize(p["positions"], grid, base_rgb=p["base_color"][:3], indices=p["indices"],
                                  model=p["model"])
            ref += acc
    assert textured == (3 if mode == 0 else 2)
    assert np.array_equal(got, ref), f"{(got != ref).any(axis=-1).sum()} voxel records differ"
    assert (got[..., 3] > 0).sum() > 200
    g.close()
