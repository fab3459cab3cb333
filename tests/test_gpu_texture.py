"""GPU textured voxelization: accumulators bit-exact against the oracle's textured path.

A textured mesh's albedo is BasicMaterial's texture(uv) * BaseColorFactor with the alpha
discard (Pipelang/Internal/main.lua:188-205), sampled per covered voxel; the sample point,
filtering and rounding are pinned on the CPU (tests/test_oracle_texture.py). Here the
binned kernel (one sample per column of the dominant projection, from the coverage test's
edge functions) and the per-triangle fallback (one sample per voxel, from the snapped
vertices) must give the oracle's integers exactly: discarded voxels absent, wrap
addressing, odd texture sizes, textured and untextured meshes in one call, 8^3 and 16^3
bins, the overflow fallback, batches and the allocation-free entry point.
"""
import numpy as np
import pytest

from final184_amd import scenegen as sg
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _vgi():
    from final184_amd import vgi
    return vgi


def _diff_msg(a, b):
    bad = np.argwhere(a != b)
    return f"{len(bad)} words differ; first {bad[:5].tolist()} gpu={a[tuple(bad[0])]} ref={b[tuple(bad[0])]}"


def _texture(h, w, seed, discard_frac=0.3):
    rng = np.random.default_rng(seed)
    t = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    t[..., 3] = 255
    low = rng.random((h, w)) < discard_frac
    t[..., 3][low] = rng.integers(0, 26, int(low.sum()))  # around the 0.05 threshold
    return t


def _uv(nverts, seed):
    return np.random.default_rng(seed).uniform(-1.5, 2.5, (nverts, 2)).astype(np.float32)


FACTOR = (0.9, 0.75, 1.0, 1.0)


@pytest.fixture(scope="module")
def soup():
    w = sg.random_soup(ntri=20_000, n=64)
    pos = w.positions.reshape(-1, 3)
    uv = _uv(len(pos), 11)
    tex = _texture(37, 53, 12)
    ref, st = orc.voxelize_textured(pos, w.grid, uv, tex, factor=FACTOR)
    return dict(w=w, pos=pos, uv=uv, tex=tex, ref=ref, stats=st)


@pytest.mark.parametrize("bin_size", [8, 16])
def test_textured_soup_bit_exact(ctx, soup, bin_size):
    vgi = _vgi()
    w = soup["w"]
    print(f"case textured soup 20000 tris 64^3 bin {bin_size}", flush=True)
    mesh = vgi.MeshData(ctx, soup["pos"], texcoords=soup["uv"], texture=soup["tex"], base_color_factor=FACTOR)
    sc = ctx.scene()
    try:
        ctx.set_tuning(bin_size=bin_size)
        sc.voxelize([mesh], vgi.make_grid(**w.grid))
        acc = sc.read_voxels()
    finally:
        ctx.set_tuning()
        sc.release()
    ref = soup["ref"]
    assert np.array_equal(acc, ref), _diff_msg(acc, ref)
    # the discard took voxels away: the untextured coverage has more
    plain, _ = orc.voxelize(w.positions, w.grid, albedo=w.albedo)
    assert (plain[..., 3] > 0).sum() > (ref[..., 3] > 0).sum() > 1000


def test_textured_and_plain_meshes_in_one_call(ctx, soup):
    """A textured mesh beside an untextured one in the same scene: the sums add."""
    vgi = _vgi()
    w = soup["w"]
    print("case textured + plain meshes", flush=True)
    other = sg.random_soup(seed=99, ntri=5000, n=64)
    ref = soup["ref"].copy()
    ref_plain, _ = orc.voxelize(other.positions, w.grid, albedo=other.albedo)
    ref += ref_plain
    ref[..., 7] = 0
    m1 = vgi.MeshData(ctx, soup["pos"], texcoords=soup["uv"], texture=soup["tex"], base_color_factor=FACTOR)
    m2 = vgi.MeshData(ctx, other.positions.reshape(-1, 3), albedo=other.albedo)
    sc = ctx.scene()
    try:
        sc.voxelize([m2, m1], vgi.make_grid(**w.grid))
        acc = sc.read_voxels()
    finally:
        sc.release()
    assert np.array_equal(acc, ref), _diff_msg(acc, ref)


def test_textured_walls_fallback_and_batch(ctx):
    """Full-width textured planes at 256^3: large triangles (one wavefront each in the bin
    pass) whose pairs overflow a fresh stream's plan, so the first call takes the
    per-triangle path and the second the binned path; both equal the oracle. Then a batch
    of the planes and a textured soup in one call."""
    vgi = _vgi()
    print("case textured planes 256^3: fallback, binned, batch", flush=True)
    n = 256
    grid_d = dict(origin=(0.0, 0.0, 0.0), extent=1.0, n=n, levels=8)
    rng = np.random.default_rng(21)
    zs = (np.sort(rng.choice(n, 80, replace=False)) + rng.uniform(0.2, 0.8, 80)) / n
    tris, uvs = [], []
    for z in zs:
        a, b, c, d = (0.001, 0.001, z), (0.999, 0.001, z), (0.999, 0.999, z), (0.001, 0.999, z)
        tris += [a, b, c, a, c, d]
        s = float(rng.uniform(1.0, 3.0))  # texture repeats s times across the plane
        uvs += [(a[0] * s, a[1] * s), (b[0] * s, b[1] * s), (c[0] * s, c[1] * s)]
        uvs += [(a[0] * s, a[1] * s), (c[0] * s, c[1] * s), (d[0] * s, d[1] * s)]
    pos = np.array(tris, np.float32)
    uv = np.array(uvs, np.float32)
    tex = _texture(64, 48, 22)
    ref, _ = orc.voxelize_textured(pos, grid_d, uv, tex)
    mesh = vgi.MeshData(ctx, pos, texcoords=uv, texture=tex)
    grid = vgi.make_grid(**grid_d)
    st_a = ctx.stream_create()
    sc = ctx.scene()
    try:
        for call in range(2):
            sc.voxelize([mesh], grid, stream=st_a)
            ctx.sync(st_a)
            st = sc.stats()
            print(f"call {call}: pairs {st.bin_pairs} fallback {st.voxelize_fallback}", flush=True)
            assert st.voxelize_fallback == (1 if call == 0 else 0)
            acc = sc.read_voxels()
            assert np.array_equal(acc, ref), _diff_msg(acc, ref)
    finally:
        sc.release()
        ctx.stream_destroy(st_a)
    w = sg.random_soup(seed=5, ntri=3000, n=64)
    wuv = _uv(3 * w.num_triangles, 23)
    wtex = _texture(16, 16, 24)
    ref_w, _ = orc.voxelize_textured(w.positions, w.grid, wuv, wtex)
    mesh_w = vgi.MeshData(ctx, w.positions.reshape(-1, 3), texcoords=wuv, texture=wtex)
    s1, s2 = ctx.scene(), ctx.scene()
    try:
        ctx.voxelize_batch([(s1, [mesh], grid), (s2, [mesh_w], vgi.make_grid(**w.grid))])
        ctx.sync()
        assert np.array_equal(s1.read_voxels(), ref)
        assert np.array_equal(s2.read_voxels(), ref_w)
    finally:
        s1.release()
        s2.release()


def test_textured_noalloc_equals_voxelize(ctx, soup):
    vgi = _vgi()
    w = soup["w"]
    print("case textured soup through the allocation-free entry point", flush=True)
    mesh = vgi.MeshData(ctx, soup["pos"], texcoords=soup["uv"], texture=soup["tex"], base_color_factor=FACTOR)
    grid = vgi.make_grid(**w.grid)
    sc = ctx.scene()
    try:
        sc.voxelize([mesh], grid)  # the scene's volume
        sc.voxelize_noalloc([mesh], grid)
        ctx.sync()
        acc = sc.read_voxels()
    finally:
        sc.release()
    assert np.array_equal(acc, soup["ref"]), _diff_msg(acc, soup["ref"])


def test_textured_mesh_validation(ctx):
    vgi = _vgi()
    tri = np.array([[0.1, 0.1, 0.5], [0.9, 0.1, 0.5], [0.1, 0.9, 0.5]], np.float32)
    uv = np.zeros((3, 2), np.float32)
    mesh = vgi.MeshData(ctx, tri, texcoords=uv, texture=np.full((4, 4, 4), 255, np.uint8))
    sc = ctx.scene()
    grid = vgi.make_grid(origin=(0.0, 0.0, 0.0), extent=1.0, n=32, levels=1)
    try:
        mesh.desc.texture_width = 0
        with pytest.raises(vgi.VgiError):
            sc.voxelize([mesh], grid)
        mesh.desc.texture_width = 4
        mesh.desc.base_color_factor[3] = float("nan")
        with pytest.raises(vgi.VgiError):
            sc.voxelize([mesh], grid)
        mesh.desc.base_color_factor[3] = 1.0
        sc.voxelize([mesh], grid)
        assert sc.read_voxels()[..., 3].sum() > 0
    finally:
        sc.release()
