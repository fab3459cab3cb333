"""Synthetic workload generators: determinism, counts and geometric bounds."""
import numpy as np

from final184_amd import scenegen as sg


def test_random_soup_deterministic_and_bounded():
    a, ca = sg.random_triangles(7, 5000)
    b, cb = sg.random_triangles(7, 5000)
    c, _ = sg.random_triangles(8, 5000)
    assert np.array_equal(a, b) and np.array_equal(ca, cb) and not np.array_equal(a, c)
    t = a.reshape(-1, 3, 3).astype(np.float64)
    assert t.min() >= 0.0 and t.max() <= 1.0
    edges = np.linalg.norm(t - np.roll(t, 1, axis=1), axis=2)
    assert edges.max() <= 0.05 + 1e-6


def test_room_triangle_budget_and_layout():
    pos, alb, rough = sg.room_triangles(3, 120_000)
    assert pos.shape == (120_000, 9)
    t = pos.reshape(-1, 3, 3)
    assert t[..., 0].min() >= -1e-6 and t[..., 0].max() <= 1.0 + 1e-6
    assert t[..., 1].min() >= -1e-6 and t[..., 1].max() <= 0.45 + 1e-6
    assert t[..., 2].min() >= -1e-6 and t[..., 2].max() <= 0.6 + 1e-6
    assert rough.min() >= 0.05 and rough.max() <= 0.5
    # fixed part: 72 wall + 8,192 column + 80,000 tile triangles, the rest is debris
    wall = t[:72]
    n = np.cross(wall[:, 1] - wall[:, 0], wall[:, 2] - wall[:, 0])
    centre = np.array([0.5, 0.225, 0.3])
    inward = np.einsum("ij,ij->i", n, centre - wall.mean(axis=1))
    assert np.all(inward > 0)  # walls face the inside of the room


def test_gbuffer_on_triangles_points_lie_on_triangles():
    pos, _ = sg.random_triangles(1, 200, radius=0.05)
    gp, gn = sg.gbuffer_on_triangles(1, pos, 4000)
    np.testing.assert_allclose(np.linalg.norm(gn, axis=1), 1.0, atol=1e-5)
    tri = pos.reshape(-1, 3, 3).astype(np.float64)
    n = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    # each sample is on the plane of the triangle whose normal it carries
    for p, m in zip(gp[:200].astype(np.float64), gn[:200].astype(np.float64)):
        k = np.argmin(np.linalg.norm(n - m, axis=1))
        assert abs(np.dot(p - tri[k, 0], n[k])) < 1e-5


def test_camera_gbuffer_faces_camera():
    w = sg.room(ntri=30_000, n=64, levels=6, width=64, height=36)
    eye = np.array(w.cones["eye"])
    hit = np.linalg.norm(w.gb_nrm, axis=1) > 0
    assert hit.mean() > 0.99
    view = w.gb_pos[hit] - eye
    assert np.all(np.einsum("ij,ij->i", view, w.gb_nrm[hit]) <= 1e-6)
    assert np.all((w.gb_rough[hit] >= 0.05) & (w.gb_rough[hit] <= 0.5))
    w2 = sg.room(ntri=30_000, n=64, levels=6, width=64, height=36)
    assert np.array_equal(w.gb_pos, w2.gb_pos)


def test_lognormal_draws_are_clipped():
    x = sg.lognormal_ints(3, 10_000, 20_000, 1.0, 1_000, 200_000)
    assert x.min() >= 1_000 and x.max() <= 200_000
    assert 15_000 < np.median(x) < 25_000


def test_batch_scenes_shapes_and_determinism():
    a = sg.batch_scenes(count=12, median=2000.0, tri_range=(100, 5000))
    b = sg.batch_scenes(count=12, median=2000.0, tri_range=(100, 5000))
    for x, y in zip(a, b):
        assert x.name == y.name and np.array_equal(x.positions, y.positions)
        assert 100 <= x.num_triangles <= 5000
        assert x.grid["n"] in (32, 64, 128, 256) and 2 ** x.grid["levels"] == x.grid["n"]
        assert x.width == x.height and x.width in (256, 512, 1024)
        assert x.gb_pos.shape == (x.num_pixels, 3)
        d = np.array(x.lights[0]["vec"])
        assert abs(np.linalg.norm(d) - 1.0) < 1e-6 and d[1] < 0
    assert len({x.num_triangles for x in a}) > 6  # uneven sizes
