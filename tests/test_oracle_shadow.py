"""The oracle's light-visibility walk against closed forms.

A light flagged to cast shadows is cut at an occupied voxel when the walk from the voxel
toward the light meets another occupied voxel (the voxel-space counterpart of the
reference's shadow-map test of each lit hit, Shader/Lighting/indirect.frag:160-165). The
walk's definition: start at the voxel centre moved one voxel along the dominant axis of
the voxel's summed normal; step l / |l_a| toward the light (a = dominant axis of l) in
16.16 fixed point; lit when a sample leaves the grid or, for a point light, once the walk
reaches the light's distance along a; unlit at the first occupied sample other than the
voxel itself.

The volumes here are written directly as accumulators (count 1, chosen normal sums), so
each expected shadow is a closed form in integers: a slab's footprint under an overhead
light, the same footprint moved by the walk's rational slope under an oblique light, the
exact sample count of a point light's walk, the normal offset that keeps a surface from
shadowing itself at grazing angles, the tie rule of that offset, and a closed box that
no directional light enters.
"""
import numpy as np
import pytest

from oracle import oracle as orc

N = 32
GRID = dict(origin=(0.0, 0.0, 0.0), extent=float(N), n=N, levels=1)  # one world unit per voxel


def _acc():
    return np.zeros((N, N, N, 8), np.uint32)


def _put(acc, x, y, z, normal, rgb=(255, 255, 255)):
    """Occupied voxel(s) (count 1) with the given integer normal (units of 1/1024)."""
    acc[z, y, x, 0:3] = rgb
    acc[z, y, x, 3] = 1
    acc[z, y, x, 4:7] = [c + 1024 for c in normal]


def _floor(acc, y=2, normal=(0, 1024, 0)):
    _put(acc, slice(None), y, slice(None), normal)


def _slab(acc, y, x0, x1, z0, z1, normal=(0, 1024, 0)):
    _put(acc, slice(x0, x1), y, slice(z0, z1), normal)


def _vis(acc, lights):
    return orc.light_visibility(acc, GRID, lights)


def _sun(minus_vec, shadow=True):
    """Directional light whose walk direction (toward the light) is minus_vec."""
    return dict(type="directional", vec=tuple(-float(v) for v in minus_vec), color=(1.0, 1.0, 1.0),
                shadow=shadow)


def test_overhead_light_shadow_is_the_slab_footprint():
    acc = _acc()
    _floor(acc)
    _slab(acc, 20, 8, 16, 10, 20)
    vis = _vis(acc, [_sun((0.0, 1.0, 0.0))])
    floor_lit = (vis[:, 2, :] & 1).astype(bool)  # [z][x]
    xs, zs = np.meshgrid(np.arange(N), np.arange(N))
    want = ~((xs >= 8) & (xs < 16) & (zs >= 10) & (zs < 20))
    assert np.array_equal(floor_lit, want)
    # the slab's own top faces the open sky
    assert (vis[10:20, 20, 8:16] & 1).all()
    # unoccupied voxels carry no bits
    assert (vis[:, 5, :] == 0).all()


def test_oblique_light_moves_the_footprint_by_the_walk_slope():
    """Toward the light l = (0.5, 1, 0.25): the dominant axis is y, the step (1/2, 1, 1/4) is
    exact in 16.16. From floor voxel (x, 2, z) the walk starts at (x+.5, 3.5, z+.5) and is at
    layer y = 20 after 17 steps, in voxel (x + floor(.5 + 8.5), 20, z + floor(.5 + 4.25)) =
    (x + 9, 20, z + 4): shadowed iff the slab holds that voxel."""
    acc = _acc()
    _floor(acc)
    _slab(acc, 20, 8, 16, 10, 20)
    vis = _vis(acc, [_sun((0.5, 1.0, 0.25))])
    floor_lit = (vis[:, 2, :] & 1).astype(bool)
    xs, zs = np.meshgrid(np.arange(N), np.arange(N))
    want = ~((xs + 9 >= 8) & (xs + 9 < 16) & (zs + 4 >= 10) & (zs + 4 < 20))
    assert np.array_equal(floor_lit, want)
    assert (~floor_lit).sum() == 7 * 10


def test_dominant_axis_tie_goes_to_the_lowest_axis():
    """l = (1, 1, 0): x and y tie, x is taken, so the step is (1, 1, 0) and the walk from
    (x+.5, 3.5, z+.5) meets layer 20 after 17 steps at x + 17."""
    acc = _acc()
    _floor(acc)
    _slab(acc, 20, 20, 28, 10, 20)
    vis = _vis(acc, [_sun((1.0, 1.0, 0.0))])
    floor_lit = (vis[:, 2, :] & 1).astype(bool)
    xs, zs = np.meshgrid(np.arange(N), np.arange(N))
    want = ~((xs + 17 >= 20) & (xs + 17 < 28) & (zs >= 10) & (zs < 20))
    assert np.array_equal(floor_lit, want)


def test_point_light_walk_ends_at_the_light():
    """Light at the centre of voxel (16, 10, 16). From floor voxel (16, 2, 16) the walk starts
    at y = 3.5 and takes ceil(7) = 7 samples, voxels y = 3..9: a voxel at y = 9 blocks it, the
    voxel holding the light (y = 10) and everything beyond the light do not."""
    light = dict(type="point", vec=(16.5, 10.5, 16.5), color=(1.0, 1.0, 1.0), min_distance=0.1, shadow=True)
    for occ_y, lit in ((9, False), (10, True), (25, True)):
        acc = _acc()
        _floor(acc)
        _put(acc, 16, occ_y, 16, (0, 1024, 0))
        vis = _vis(acc, [light])
        assert bool(vis[16, 2, 16] & 1) == lit, occ_y
    # a full ceiling above the light shadows nothing on the floor
    acc = _acc()
    _floor(acc)
    _slab(acc, 25, 0, N, 0, N)
    assert (_vis(acc, [light])[:, 2, :] & 1).all()


def test_point_light_perspective_footprint():
    """Light at (16.5, 20.5, 16.5), a 3x3 occluder at y = 11, x and z in [15, 18). Floor voxel
    (x, 2, z) starts at (x+.5, 3.5, z+.5) with l = (16 - x, 17, 16 - z): for |16 - x| and
    |16 - z| <= 17 the dominant axis is y (exactly 17), and the walk reaches layer 11 after 8
    steps of (dx / 17, 1, dz / 17), each rounded once to 16.16, at x-coordinate
    (x + .5) * 65536 + 8 * rint(65536 * fl(dx / 17)). The shadow is the set of floor voxels
    whose walk is then inside the occluder: a magnified copy (17 / 9) of it."""
    light = dict(type="point", vec=(16.5, 20.5, 16.5), color=(1.0, 1.0, 1.0), min_distance=0.1, shadow=True)
    acc = _acc()
    _floor(acc)
    _slab(acc, 11, 15, 18, 15, 18)
    vis = _vis(acc, [light])
    f17 = np.float32(17.0)

    def step(d):  # the walk's 16.16 step: single-precision quotient, rounded to nearest even
        return int(np.rint(np.float32(np.float32(d) / f17) * np.float32(65536.0)))

    shadowed = set()
    for z in range(N):
        for x in range(N):
            dx, dz = 16 - x, 16 - z
            if max(abs(dx), abs(dz)) > 17:
                continue
            cx = (x * 65536 + 32768 + 8 * step(dx)) >> 16
            cz = (z * 65536 + 32768 + 8 * step(dz)) >> 16
            if 15 <= cx < 18 and 15 <= cz < 18:
                shadowed.add((x, z))
    got = {(x, z) for z in range(N) for x in range(N) if not vis[z, 2, x] & 1}
    assert got == shadowed
    assert (16, 16) in got and len(got) > 9  # magnified beyond the occluder's own 3x3


def test_normal_offset_keeps_a_surface_from_shadowing_itself():
    """A grazing light along the floor, l = (1, 1/8, 0): the walk starts one voxel above
    the floor and climbs, so no floor voxel blocks another. Flipping the floor's normal
    starts the walk below the floor at y = 1.5; it meets the floor layer after 4 steps."""
    sun = _sun((1.0, 0.125, 0.0))
    acc = _acc()
    _floor(acc)
    assert (_vis(acc, [sun])[:, 2, :] & 1).all()
    acc = _acc()
    _floor(acc, normal=(0, -1024, 0))
    lit = (_vis(acc, [sun])[:, 2, :] & 1).astype(bool)
    # walks from x >= 28 leave the grid (x + 4 >= 32) before reaching the floor layer
    xs = np.arange(N)[None, :].repeat(N, 0)
    assert np.array_equal(lit, xs + 4 >= N)


def test_normal_offset_tie_goes_to_the_lowest_axis():
    """Normal sums (724, 724, 0): x and y tie, the walk starts at (11.5, 10.5, 10.5) and goes
    straight up along x = 11."""
    sun = _sun((0.0, 1.0, 0.0))
    for occ_x, lit in ((11, False), (10, True)):
        acc = _acc()
        _put(acc, 10, 10, 10, (724, 724, 0))
        _put(acc, occ_x, 15, 10, (0, 1024, 0))
        assert bool(_vis(acc, [sun])[10, 10, 10] & 1) == lit, occ_x


def test_closed_box_admits_no_directional_light():
    """A closed shell with inward normals: from every inner face voxel that faces the light
    (normal . l > 0, the only case in which the light reaches the radiance) the walk heads
    into the box and crosses a wall layer before it can leave the grid (a sample moves at
    most one voxel per axis per step, so no layer is skipped), for any direction."""
    acc = _acc()
    lo, hi = 4, 27
    _put(acc, slice(lo, hi + 1), lo, slice(lo, hi + 1), (0, 1024, 0))
    _put(acc, slice(lo, hi + 1), hi, slice(lo, hi + 1), (0, -1024, 0))
    _put(acc, lo, slice(lo + 1, hi), slice(lo, hi + 1), (1024, 0, 0))
    _put(acc, hi, slice(lo + 1, hi), slice(lo, hi + 1), (-1024, 0, 0))
    _put(acc, slice(lo + 1, hi), slice(lo + 1, hi), lo, (0, 0, 1024))
    _put(acc, slice(lo + 1, hi), slice(lo + 1, hi), hi, (0, 0, -1024))
    rng = np.random.default_rng(3)
    dirs = np.concatenate([rng.normal(size=(6, 3)), [[0.0, 1.0, 0.0]]])
    vis = _vis(acc, [_sun(tuple(float(v) for v in d)) for d in dirs])
    occ = acc[..., 3] > 0
    nrm = acc[..., 4:7].astype(np.int64) - 1024
    checked = 0
    for b, d in enumerate(dirs):
        facing = occ & ((nrm @ d) > 0)
        checked += int(facing.sum())
        assert ((vis[facing] >> b) & 1 == 0).all(), b
    assert checked > 3 * occ.sum()


@pytest.mark.parametrize("prec", ["fp32", "fp16"])
def test_shadowed_radiance_is_zero_exactly_where_the_walk_is_blocked(prec):
    acc = _acc()
    _floor(acc)
    _slab(acc, 20, 8, 16, 10, 20)
    base = dict(type="directional", vec=(-0.5, -1.0, -0.25), color=(0.8, 0.9, 1.0))
    plain = orc.inject_mips(acc, GRID, [base], prec)[0]
    shad = orc.inject_mips(acc, GRID, [dict(base, shadow=True)], prec)[0]
    vis = _vis(acc, [dict(base, shadow=True)])
    blocked = (acc[..., 3] > 0) & ((vis & 1) == 0)
    assert blocked.sum() == 70
    assert (shad[blocked][:, :3] == 0).all() and (shad[blocked][:, 3] == 1).all()
    assert np.array_equal(shad[~blocked], plain[~blocked])
    assert (plain[blocked][:, :3] > 0).all()
