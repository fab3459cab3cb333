"""Pins for the oracle's restatements of the reference shaders (reference behaviour).

Covers VoxelGS's dominant-axis choice (Pipelang/Internal/main.lua:100-115), VoxelPS's
RGB565 colour / 5:6:5 normal packing (main.lua:270-271) against indirect.frag's
unpackers (Shader/Lighting/indirect.frag:88-102), make_coord_space (:71-86),
hemisphereSample_cos (Shader/math.inc:76-81) and getIndirect's voxel ray march
(indirect.frag:104-184) with its fixed random numbers given.
"""
import math

import numpy as np

from oracle import oracle as orc


def test_orientation_rule_and_ties():
    assert orc.ref_orientation((3, 1, 1)) == 1      # X dominant
    assert orc.ref_orientation((1, 3, 1)) == 2      # Y dominant
    assert orc.ref_orientation((1, 1, 3)) == 0      # Z dominant
    assert orc.ref_orientation((2, 2, 1)) == 2      # x == y > z: Y (y >= x branch)
    assert orc.ref_orientation((2, 1, 2)) == 0      # x == z: Z
    assert orc.ref_orientation((1, 2, 2)) == 0      # y == z: Z
    assert orc.ref_orientation((1, 1, 1)) == 0      # all equal: Z
    assert orc.ref_orientation((-5, 1, 2)) == 1     # absolute values


def test_color565_round_trip_truncates():
    rng = np.random.default_rng(0)
    for r, g, b in rng.uniform(0, 1, (500, 3)):
        c = orc.ref_unpack_color565(orc.ref_pack_color565(r, g, b))
        # uint(x * 31) truncates: decoded <= true, within one code
        assert c[0] <= r + 1e-6 and r - c[0] < 1 / 31 + 1e-6
        assert c[1] <= g + 1e-6 and g - c[1] < 1 / 63 + 1e-6
        assert c[2] <= b + 1e-6 and b - c[2] < 1 / 31 + 1e-6
    assert orc.ref_pack_color565(1, 1, 1) == 0xFFFF
    assert orc.ref_pack_color565(0, 0, 0) == 0


def test_normal565_axis_swap_quirk():
    # the packer puts x in bits 11-15, the unpacker reads x from bits 0-4: x and z swap,
    # and decoding v/16 - 1 after encoding n*16 + 15 biases every component by -1/16
    n = orc.ref_unpack_normal565(orc.ref_pack_normal565(0.0, 0.0, 1.0))
    raw = np.array([31 / 16 - 1, 31 / 32 - 1, 15 / 16 - 1])
    np.testing.assert_allclose(n, raw / np.linalg.norm(raw), rtol=1e-6)
    assert n[0] > 0.99  # a +z normal comes back as (almost) +x
    m = orc.ref_unpack_normal565(orc.ref_pack_normal565(1.0, 0.0, 0.0))
    assert m[2] > 0.99


def test_hemisphere_cos_distribution():
    rng = np.random.default_rng(1)
    us, vs = rng.uniform(0, 1, 20000), rng.uniform(0, 1, 20000)
    d = np.array([orc.ref_hemisphere_cos(u, v) for u, v in zip(us, vs)])
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-5)
    assert np.all(d[:, 2] >= 0)
    # cosine-weighted hemisphere: E[cos] = 2/3, E[cos^2] = 1/2
    assert abs(d[:, 2].mean() - 2 / 3) < 0.01
    assert abs((d[:, 2] ** 2).mean() - 0.5) < 0.01


def test_make_coord_space_is_orthonormal_with_z_normal():
    rng = np.random.default_rng(2)
    for n in rng.normal(size=(200, 3)):
        m = orc.ref_make_coord_space(n)
        np.testing.assert_allclose(m @ m.T, np.eye(3), atol=1e-5)
        np.testing.assert_allclose(m[2], n / np.linalg.norm(n), atol=1e-5)


def _volume_with_wall(z_wall):
    vox = np.zeros((128, 128, 128, 2), np.uint16)  # [z][y][x][rg]
    vox[z_wall, :, :, 0] = 0xFFFF
    vox[z_wall, :, :, 1] = orc.ref_pack_normal565(0.0, 0.0, 1.0)
    return vox


# voxel coordinates == world coordinates: x' = x/64 - 1, y' = y/64 - 1, z' = z/128
W2V = np.array([[1 / 64, 0, 0, 0], [0, 1 / 64, 0, 0], [0, 0, 1 / 128, 0], [-1, -1, 0, 1]],
               np.float32)  # column-major: columns are rows of this array


def test_get_indirect_hits_wall_at_predicted_step():
    vox = _volume_with_wall(30)
    wpos, wnorm = np.array([60.3, 70.2, 20.0]), np.array([0.0, 0.0, 1.0])
    sun_dir, lum = np.array([0.0, 0.0, -1.0]), np.array([2.0, 2.0, 2.0])
    hit, Lo, hp, steps = orc.ref_get_indirect(vox, W2V, wpos, wnorm, 0.0, 0.5, sun_dir, lum)
    # u = 0 samples the normal; start offset (1 + 0.5) * 0.2, then 0.2 per step:
    # z_i = 20.3 + 0.2 (i + 1) first reaches voxel 30 at i = 48
    assert hit and steps == 48
    assert abs(hp[2] - 30.1) < 1e-3
    vn = orc.ref_unpack_normal565(orc.ref_pack_normal565(0.0, 0.0, 1.0))
    expect = lum * abs(float(np.dot(-sun_dir, vn)))  # colour 0xFFFF -> 1, N.D = 1
    np.testing.assert_allclose(Lo, expect, rtol=1e-5)


def test_get_indirect_sky_only_after_all_steps():
    vox = _volume_with_wall(0)[:]  # wall at z = 0, behind the ray
    wnorm = np.array([0.0, 0.0, 1.0])
    sun_dir, lum = np.array([0.0, 0.0, -1.0]), np.array([1.0, 1.0, 1.0])
    # inside the volume for all 60 steps: sky term (0.7, 0.8, 1.0) * 0.4
    hit, Lo, _, steps = orc.ref_get_indirect(vox, W2V, np.array([60.0, 60.0, 20.0]), wnorm, 0.0, 0.0,
                                             sun_dir, lum)
    assert not hit and steps == 60
    np.testing.assert_allclose(Lo, [0.28, 0.32, 0.4], rtol=1e-6)
    # leaving the volume before 60 steps: no sky at all (indirect.frag:145-149 vs :179)
    hit, Lo, _, steps = orc.ref_get_indirect(vox, W2V, np.array([60.0, 60.0, 120.0]), wnorm, 0.0, 0.0,
                                             sun_dir, lum)
    assert not hit and steps < 60 and not Lo.any()
