"""The oracle's textured albedo (BasicMaterial's texture path and alpha discard) against
closed forms.

The reference's material multiplies a bilinear, wrap-addressed texture sample by
BaseColorFactor and discards fragments whose alpha is below 0.05
(Pipelang/Internal/main.lua:188-205; the sampler: Foreground/Renderer/MegaPipeline.cpp:39-43).
Here a voxel covered by a textured triangle takes the sample at its centre projected along
the triangle's dominant axis. On a triangle in the plane z = 8.5 of a 16^3 grid with one
world unit per voxel and uv = (x, y) / 16, a voxel whose centre is inside the triangle
samples uv = ((i + 1/2) / 16, (j + 1/2) / 16): the centre of texel (i, j) of a 16 x 16
texture, so it takes that texel exactly; in a texture twice as wide it falls halfway
between texels 2i and 2i + 1; shifted by whole periods it wraps to the same texels.
"""
import numpy as np
import pytest

from oracle import oracle as orc

N = 16
GRID = dict(origin=(0.0, 0.0, 0.0), extent=float(N), n=N, levels=1)
Z = 8.5
# one triangle over most of the layer z = 8 (vertices on voxel corners, centre rows inside)
TRI = np.array([[0.0, 0.0, Z], [16.0, 0.0, Z], [0.0, 16.0, Z]], np.float32)


def _uv(scale=1.0 / 16.0, shift=(0.0, 0.0)):
    return (TRI[:, :2] * scale + np.array(shift, np.float32)).astype(np.float32)


def _inside():
    """Voxels (i, j) of layer 8 whose centre is strictly inside the triangle."""
    return [(i, j) for j in range(N) for i in range(N) if (i + 0.5) + (j + 0.5) < 16.0]


def _texture(w, h, seed):
    rng = np.random.default_rng(seed)
    t = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    t[..., 3] = 255
    return t


def test_texel_centres_give_the_texel():
    tex = _texture(16, 16, 1)
    acc, st = orc.voxelize_textured(TRI, GRID, _uv(), tex)
    assert st.rejected == 0 and st.degenerate == 0
    for i, j in _inside():
        rec = acc[8, j, i]
        assert rec[3] == 1, (i, j)
        assert tuple(rec[:3]) == tuple(tex[j, i, :3]), (i, j)
    # only layer 8 is touched: the plane z = 8.5 crosses no other layer's closed box
    assert acc[..., 3].sum() == acc[8, :, :, 3].sum()


def test_halfway_between_texels_is_their_mean():
    tex = np.zeros((16, 32, 4), np.uint8)
    tex[..., 3] = 255
    rng = np.random.default_rng(2)
    even = rng.integers(0, 120, (16, 16, 3)) * 2  # pairs with even sums
    tex[:, 0::2, :3] = even
    tex[:, 1::2, :3] = even + 2 * rng.integers(0, 60, (16, 16, 3))
    # 32 texels across: the centre of voxel i samples x = 32 (i + 1/2) / 16 - 1/2 = 2i + 1/2
    acc, _ = orc.voxelize_textured(TRI, GRID, _uv(), tex)
    for i, j in _inside():
        want = (tex[j, 2 * i, :3].astype(int) + tex[j, 2 * i + 1, :3].astype(int)) // 2
        assert tuple(acc[8, j, i, :3]) == tuple(want), (i, j)


def test_wrap_addressing_repeats_the_texture():
    tex = _texture(16, 16, 3)
    a, _ = orc.voxelize_textured(TRI, GRID, _uv(), tex)
    b, _ = orc.voxelize_textured(TRI, GRID, _uv(shift=(3.0, -2.0)), tex)
    assert np.array_equal(a, b)


def test_alpha_below_five_percent_is_discarded():
    tex = _texture(16, 16, 4)
    alpha = np.random.default_rng(5).integers(0, 26, (16, 16)).astype(np.uint8)  # around 0.05 * 255
    tex[..., 3] = alpha
    acc, st = orc.voxelize_textured(TRI, GRID, _uv(), tex)
    for i, j in _inside():
        kept = alpha[j, i] >= 13  # 12/255 < 0.05 <= 13/255
        assert acc[8, j, i, 3] == (1 if kept else 0), (i, j, alpha[j, i])
        if not kept:
            assert (acc[8, j, i] == 0).all()  # no colour, normal or count: not occupied


def test_factor_multiplies_the_sample():
    tex = _texture(16, 16, 6)
    tex[..., :3] &= 0xFE  # even values: halves are exact
    acc, _ = orc.voxelize_textured(TRI, GRID, _uv(), tex, factor=(0.5, 1.0, 0.0, 1.0))
    for i, j in _inside():
        r, g, b = (int(v) for v in acc[8, j, i, :3])
        assert (r, g, b) == (tex[j, i, 0] // 2, tex[j, i, 1], 0), (i, j)
    # a factor alpha below 0.05 discards everything
    acc, _ = orc.voxelize_textured(TRI, GRID, _uv(), tex, factor=(1.0, 1.0, 1.0, 0.04))
    assert acc[..., 3].sum() == 0


@pytest.mark.parametrize("rgb", [(255, 255, 255), (200, 31, 7)])
def test_white_texture_equals_the_untextured_base_colour(rgb):
    """A white opaque texture with factor rgb / 255 reproduces the untextured voxelization
    with that base colour exactly: same voxels, counts, normals and colour sums."""
    rng = np.random.default_rng(7)
    grid = dict(origin=(0.0, 0.0, 0.0), extent=1.0, n=32, levels=1)
    ntri = 300
    c = rng.uniform(0.1, 0.9, (ntri, 1, 3))
    pos = (c + rng.uniform(-0.06, 0.06, (ntri, 3, 3))).astype(np.float32).reshape(-1, 3)
    uv = rng.uniform(-2.0, 2.0, (pos.shape[0], 2)).astype(np.float32)
    white = np.full((4, 8, 4), 255, np.uint8)
    factor = tuple(np.float32(v) / np.float32(255.0) for v in rgb) + (1.0,)
    a, _ = orc.voxelize(pos, grid, base_rgb=rgb)
    b, _ = orc.voxelize_textured(pos, grid, uv, white, factor=factor)
    assert np.array_equal(a, b)


def test_meshes_accumulate():
    tex = _texture(16, 16, 8)
    a, _ = orc.voxelize_textured(TRI, GRID, _uv(), tex)
    acc, _ = orc.voxelize_textured(TRI, GRID, _uv(), tex)
    acc, _ = orc.voxelize_textured(TRI, GRID, _uv(), tex, acc=acc)
    assert np.array_equal(acc, 2 * a)


def test_sample_point_is_the_centre_projected_along_the_dominant_axis():
    """A slanted triangle, z = 8.5 + x / 4 (dominant axis z), with uv = (x, y) / 16: every
    covered voxel of column (i, j) -- several layers, as the plane crosses them -- samples
    at the column's (x, y) centre, so it takes texel (i, j). (Projecting along the normal
    instead would move the sample by up to a quarter voxel per layer.)"""
    tri = np.array([[0.0, 0.0, Z], [12.0, 0.0, Z + 3.0], [0.0, 12.0, Z]], np.float32)
    uv = (tri[:, :2] / 16.0).astype(np.float32)
    tex = _texture(16, 16, 9)
    acc, _ = orc.voxelize_textured(tri, GRID, uv, tex)
    seen = 0
    for j in range(N):
        for i in range(N):
            x, y = i + 0.5, j + 0.5
            if not (x >= 0.5 and y >= 0.5 and x + y < 12.0):  # centre inside the projection
                continue
            ks = np.nonzero(acc[:, j, i, 3])[0]
            assert len(ks) >= 1
            for k in ks:
                assert acc[k, j, i, 3] == 1 and tuple(acc[k, j, i, :3]) == tuple(tex[j, i, :3]), (i, j, k)
                seen += 1
    assert seen > 70


def test_centres_outside_clamp_to_the_boundary():
    """Conservative coverage takes voxels whose centre lies outside the triangle; their
    negative barycentric numerator is clamped to zero. Past the hypotenuse x + y = 16 of the
    triangle (0,0), (16,0), (0,16) with uv = (x, y) / 16, voxel (i, j) with i + j = 16 keeps
    the weights x_c, y_c of vertices 1 and 2: u = x_c / (x_c + y_c) = (i + 1/2) / 17, so on a
    texture whose red channel rises by 16 per column it reads red = 16 (16 u - 1/2)."""
    tex = np.zeros((16, 16, 4), np.uint8)
    tex[..., 0] = (16 * np.arange(16)).astype(np.uint8)[None, :]
    tex[..., 3] = 255
    acc, _ = orc.voxelize_textured(TRI, GRID, _uv(), tex)
    for i in range(2, 14):
        j = 16 - i
        assert acc[8, j, i, 3] == 1, (i, j)
        want = 16.0 * (16.0 * (i + 0.5) / 17.0 - 0.5)
        assert abs(int(acc[8, j, i, 0]) - want) <= 1.0, (i, j, int(acc[8, j, i, 0]), want)
