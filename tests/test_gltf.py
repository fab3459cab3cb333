"""glTF ingestion (host side, no GPU): the library's reader against closed forms and
against the reference's own parser.

Closed forms: a scene written by tests/gltf_writer.py with known vertices, index types,
topologies and node transforms must come out as the triangles numpy places in world
space, identically from .gltf + .bin, .gltf with data URIs and .glb. The reference's
parser (tinygltf, compiled from the reference sources into oracle/_ref by
final184_amd/build.py and driven as the reference's importer drives it) must yield the
same primitives, vertex for vertex. Malformed files must fail with an error code, never
crash (a seeded mutation sweep runs in-process).
"""
import ctypes as C
import os

import numpy as np
import pytest

from final184_amd import vgi
from tests import gltf_writer as gw

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libgltf_ref.so")


@pytest.fixture(scope="module")
def files(tmp_path_factory):
    d = tmp_path_factory.mktemp("gltf")
    w, expect = gw.kitchen_sink()
    paths = {"gltf": str(d / "scene.gltf"), "embedded": str(d / "embedded.gltf"), "glb": str(d / "scene.glb")}
    w.write_gltf(paths["gltf"])
    w.write_gltf(paths["embedded"], embed=True)
    w.write_glb(paths["glb"])
    return paths, expect, w


def _world(p):
    m = p["model"].astype(np.float64).reshape(3, 4)
    pos = p["positions"].astype(np.float64) @ m[:, :3].T + m[:, 3]
    return pos[p["indices"]]


@pytest.mark.parametrize("kind", ["gltf", "embedded", "glb"])
def test_reader_matches_closed_form(files, kind):
    paths, expect, _ = files
    g = vgi.Gltf(paths[kind], albedo_mode=vgi.GLTF_ALBEDO_FACTOR)
    info = g.info()
    prims = g.primitives()
    assert info.num_primitives == len(expect) == len(prims)
    assert info.num_geometries == 4          # mesh 0's two primitives are stored once
    assert info.skipped_primitives == 1      # the line primitive
    assert info.num_triangles == sum(len(e["world"]) for e in expect)
    for p, e in zip(prims, expect):
        got = _world(p)
        assert got.shape == e["world"].shape
        assert np.abs(got - e["world"]).max() < 1e-6, p["mesh"]
        assert p["material"] == e["material"]
        assert p["base_color"] == tuple(int(round(c * 255)) for c in e["rgba_factor"])
        assert np.allclose(p["model"].reshape(3, 4), e["model"][:3], atol=1e-7)


def test_containers_agree_bit_for_bit(files):
    paths, _, _ = files
    out = [vgi.Gltf(paths[k]).primitives() for k in ("gltf", "embedded", "glb")]
    for a, b in zip(out[0], out[1]):
        for k in ("positions", "indices", "model"):
            assert np.array_equal(a[k], b[k])
    for a, b in zip(out[0], out[2]):
        for k in ("positions", "indices", "model"):
            assert np.array_equal(a[k], b[k])


def test_reference_albedo_mode(files):
    """As the reference's importer: every material white, primitives without a material
    not drawn by the voxelizer."""
    paths, expect, _ = files
    prims = vgi.Gltf(paths["glb"]).primitives()
    with_mat = [e for e in expect if e["material"] >= 0]
    assert len(prims) == len(with_mat)
    for p, e in zip(prims, with_mat):
        assert p["base_color"] == (255, 255, 255, 255)
        assert np.abs(_world(p) - e["world"]).max() < 1e-6


def test_percent_encoded_buffer_uri(files, tmp_path):
    """URIs are RFC 3986 references: "my%20data.bin" names the file "my data.bin". (The
    tinygltf version vendored in the reference looks the encoded name up literally.)"""
    _, expect, w = files
    path = str(tmp_path / "spaced.gltf")
    w.write_gltf(path, bin_name="my data 100%.bin")
    assert os.path.exists(str(tmp_path / "my data 100%.bin"))
    prims = vgi.Gltf(path, albedo_mode=vgi.GLTF_ALBEDO_FACTOR).primitives()
    assert len(prims) == len(expect)
    assert np.abs(_world(prims[0]) - expect[0]["world"]).max() < 1e-6


def test_load_from_memory(files):
    paths, expect, w = files
    g = vgi.Gltf(data=w.glb_bytes())
    assert g.info().num_primitives == len([e for e in expect if e["material"] >= 0])
    # an external buffer needs a base directory
    with open(paths["gltf"], "rb") as f:
        js = f.read()
    with pytest.raises(vgi.VgiError):
        vgi.Gltf(data=js)
    g2 = vgi.Gltf(data=js, base_dir=os.path.dirname(paths["gltf"]))
    assert g2.info().num_primitives == g.info().num_primitives


# ---------------------------------------------------------------------------
# the reference's parser
# ---------------------------------------------------------------------------

def _ref():
    if not os.path.exists(REF_LIB):
        pytest.skip("oracle/_ref/libgltf_ref.so is not built (needs the reference sources)")
    L = C.CDLL(REF_LIB)
    L.gref_load.restype = C.c_void_p
    L.gref_load.argtypes = [C.c_char_p, C.c_int, C.c_char_p, C.c_int]
    L.gref_count.argtypes = [C.c_void_p]
    L.gref_get.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_void_p), C.POINTER(C.c_int), C.POINTER(C.c_void_p),
                           C.POINTER(C.c_int), C.c_void_p, C.c_void_p, C.POINTER(C.c_int)]
    L.gref_free.argtypes = [C.c_void_p]
    return L


def _ref_prims(L, path, mode):
    err = C.create_string_buffer(512)
    h = L.gref_load(path.encode(), mode, err, 512)
    assert h, err.value
    out = []
    for i in range(L.gref_count(h)):
        pp, ip = C.c_void_p(), C.c_void_p()
        nv, nt, mat = C.c_int(), C.c_int(), C.c_int()
        model = np.zeros(12, np.float32)
        rgba = np.zeros(4, np.uint8)
        L.gref_get(h, i, C.byref(pp), C.byref(nv), C.byref(ip), C.byref(nt), model.ctypes.data, rgba.ctypes.data,
                   C.byref(mat))
        pos = np.ctypeslib.as_array(C.cast(pp, C.POINTER(C.c_float)), (nv.value * 3,)).copy().reshape(-1, 3)
        idx = np.ctypeslib.as_array(C.cast(ip, C.POINTER(C.c_uint32)), (nt.value * 3,)).copy().reshape(-1, 3) \
            if nt.value else np.zeros((0, 3), np.uint32)
        out.append(dict(positions=pos, indices=idx, model=model, base_color=tuple(int(c) for c in rgba),
                        material=mat.value))
    L.gref_free(h)
    return out


@pytest.mark.parametrize("kind", ["gltf", "embedded", "glb"])
@pytest.mark.parametrize("mode", [0, 1])
def test_reader_matches_reference_parser(files, kind, mode):
    L = _ref()
    paths, _, _ = files
    ref = _ref_prims(L, paths[kind], mode)
    got = vgi.Gltf(paths[kind], albedo_mode=mode).primitives()
    assert len(got) == len(ref)
    for a, b in zip(got, ref):
        assert np.array_equal(a["positions"], b["positions"])
        assert np.array_equal(a["indices"], b["indices"])
        assert a["base_color"] == b["base_color"] and a["material"] == b["material"]
        assert np.allclose(a["model"], b["model"], rtol=0, atol=2e-7)


def test_reference_parser_on_generated_scenes(tmp_path):
    """Random scenes: node trees with matrices and TRS, all index types and topologies."""
    L = _ref()
    rng = np.random.default_rng(17)
    for case in range(12):
        w = gw.Writer()
        mats = [w.material(rng.uniform(0, 1, 4)) for _ in range(3)]
        meshes = []
        for _ in range(3):
            prims = []
            for _ in range(int(rng.integers(1, 4))):
                nv = int(rng.integers(3, 30))
                pos = rng.uniform(-1, 1, (nv, 3)).astype(np.float32)
                mode = int(rng.choice([4, 5, 6]))
                pr = {"attributes": {"POSITION": w.positions(pos, interleave_normals=bool(rng.integers(2)))},
                      "mode": mode}
                if rng.integers(2):
                    dt = [np.uint8, np.uint16, np.uint32][int(rng.integers(3))]
                    pr["indices"] = w.indices(rng.integers(0, nv, int(rng.integers(3, 40))), dt)
                if rng.integers(4):
                    pr["material"] = mats[int(rng.integers(3))]
                prims.append(pr)
            meshes.append(w.mesh(prims))
        nodes = []
        for k in range(6):
            kw = {"mesh": meshes[int(rng.integers(3))]} if rng.integers(3) else {}
            if rng.integers(2):
                kw["matrix"] = [float(v) for v in gw.trs_matrix(rng.uniform(-1, 1, 3), gw._quat(rng.normal(size=3), float(rng.uniform(0, 360))),
                                                                  rng.uniform(0.5, 2, 3)).T.reshape(-1)]
            else:
                kw["translation"] = [float(v) for v in rng.uniform(-1, 1, 3)]
                kw["rotation"] = gw._quat(rng.normal(size=3), float(rng.uniform(0, 360)))
                kw["scale"] = [float(v) for v in rng.uniform(0.5, 2, 3)]
            if k >= 2 and rng.integers(2):
                kw["children"] = [int(rng.integers(0, 2))]  # earlier leaves become children
            nodes.append(w.node(**kw))
        w.doc["scenes"][0]["nodes"] = [n for n in nodes if n >= 2]
        path = str(tmp_path / f"case{case}.glb")
        w.write_glb(path)
        for mode in (0, 1):
            ref = _ref_prims(L, path, mode)
            got = vgi.Gltf(path, albedo_mode=mode).primitives()
            assert len(got) == len(ref), case
            for a, b in zip(got, ref):
                assert np.array_equal(a["positions"], b["positions"]) and np.array_equal(a["indices"], b["indices"])
                assert a["base_color"] == b["base_color"]
                assert np.allclose(a["model"], b["model"], rtol=0, atol=1e-6), case


# ---------------------------------------------------------------------------
# malformed input
# ---------------------------------------------------------------------------

def _doc_with(w, **patch):
    d = w._finish()
    d["buffers"][0]["uri"] = "data:application/octet-stream;base64," + __import__("base64").b64encode(w.blob).decode()
    for k, v in patch.items():
        d[k] = v
    return d


def test_malformed_files_fail_cleanly(files, tmp_path):
    import json
    _, _, w = files
    good = _doc_with(w)
    cases = []
    d = json.loads(json.dumps(good)); d["asset"]["version"] = "1.0"; cases.append(("version", d))
    d = json.loads(json.dumps(good)); d["extensionsRequired"] = ["KHR_draco_mesh_compression"]; cases.append(("ext", d))
    d = json.loads(json.dumps(good)); d["accessors"][0]["count"] += 1000; cases.append(("past view", d))
    d = json.loads(json.dumps(good)); d["accessors"][0]["sparse"] = {"count": 1}; cases.append(("sparse", d))
    d = json.loads(json.dumps(good)); d["meshes"][0]["primitives"][0]["indices"] = 99; cases.append(("bad accessor", d))
    d = json.loads(json.dumps(good)); d["nodes"][0]["children"] = [0]; cases.append(("cycle", d))
    d = json.loads(json.dumps(good)); d["buffers"][0]["byteLength"] += 64; cases.append(("short buffer", d))
    d = json.loads(json.dumps(good)); d["buffers"][0]["uri"] = "missing.bin"; cases.append(("missing file", d))
    d = json.loads(json.dumps(good)); d["buffers"][0]["uri"] = "http://example.com/x.bin"; cases.append(("remote", d))
    # an index beyond the vertex count
    d = json.loads(json.dumps(good))
    blob = bytearray(w.blob)
    acc = d["accessors"][d["meshes"][0]["primitives"][0]["indices"]]
    view = d["bufferViews"][acc["bufferView"]]
    blob[view["byteOffset"]:view["byteOffset"] + 2] = (40000).to_bytes(2, "little")
    d["buffers"][0]["uri"] = "data:application/octet-stream;base64," + __import__("base64").b64encode(bytes(blob)).decode()
    cases.append(("index range", d))
    for name, doc in cases:
        with pytest.raises(vgi.VgiError) as e:
            vgi.Gltf(data=json.dumps(doc).encode(), base_dir=str(tmp_path))
        assert e.value.code in (-1, -6), name
    for text in (b"", b"{", b'{"asset": {"version": "2.0"}, "x": [1, 2,]}', b"[1]", b'{"a": "\\q"}',
                 b'{"asset": {"version": "2.0"}} trailing'):
        with pytest.raises(vgi.VgiError):
            vgi.Gltf(data=text if text else b" ")
    # a GLB whose length field overstates the file
    glb = bytearray(w.glb_bytes())
    glb[8:12] = (len(glb) + 100).to_bytes(4, "little")
    with pytest.raises(vgi.VgiError):
        vgi.Gltf(data=bytes(glb))


def test_mutated_glb_never_crashes(files):
    """Seeded byte mutations of a valid .glb: every load returns (success or an error)."""
    _, _, w = files
    base = w.glb_bytes()
    rng = np.random.default_rng(23)
    ok = bad = 0
    for _ in range(400):
        b = bytearray(base)
        for _ in range(int(rng.integers(1, 6))):
            pos = int(rng.integers(0, len(b)))
            b[pos] = int(rng.integers(0, 256))
        if rng.integers(4) == 0:
            b = b[:int(rng.integers(0, len(b)))]
        try:
            g = vgi.Gltf(data=bytes(b) if b else b" ")
            g.primitives()
            ok += 1
        except vgi.VgiError:
            bad += 1
    assert ok + bad == 400 and bad > 0


# ---------------------------------------------------------------------------
# texture coordinates, base colour textures and images
# ---------------------------------------------------------------------------

def _ref_textures(L, path, mode):
    L.gref_get_texture.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_void_p), C.POINTER(C.c_int),
                                   C.POINTER(C.c_int), C.c_void_p]
    L.gref_image.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                             C.POINTER(C.c_int), C.POINTER(C.c_void_p)]
    err = C.create_string_buffer(512)
    h = L.gref_load(path.encode(), mode, err, 512)
    assert h, err.value
    prims, images = [], []
    for i in range(L.gref_count(h)):
        uvp, nuv, img = C.c_void_p(), C.c_int(), C.c_int()
        fac = np.zeros(4, np.float32)
        L.gref_get_texture(h, i, C.byref(uvp), C.byref(nuv), C.byref(img), fac.ctypes.data)
        uv = np.ctypeslib.as_array(C.cast(uvp, C.POINTER(C.c_float)), (nuv.value * 2,)).copy().reshape(-1, 2) \
            if nuv.value else None
        prims.append(dict(texcoords=uv, image=img.value, base_color_factor=tuple(fac)))
    k = 0
    while True:
        w_, h_, comp, bits, px = C.c_int(), C.c_int(), C.c_int(), C.c_int(), C.c_void_p()
        if L.gref_image(h, k, C.byref(w_), C.byref(h_), C.byref(comp), C.byref(bits), C.byref(px)) != 0:
            break
        assert comp.value == 4 and bits.value == 8
        images.append(np.ctypeslib.as_array(C.cast(px, C.POINTER(C.c_uint8)),
                                            (h_.value * w_.value * 4,)).copy().reshape(h_.value, w_.value, 4))
        k += 1
    L.gref_free(h)
    return prims, images


@pytest.fixture(scope="module")
def textured(tmp_path_factory):
    d = tmp_path_factory.mktemp("gltf_tex")
    w, texels, expect = gw.textured_scene(file_dir=str(d))
    paths = {"gltf": str(d / "tex.gltf"), "glb": str(d / "tex.glb")}
    w.write_gltf(paths["gltf"])
    w.write_glb(paths["glb"])
    return paths, texels, expect, str(d)


@pytest.mark.parametrize("kind", ["gltf", "glb"])
@pytest.mark.parametrize("mode", [0, 1])
def test_texcoords_and_textures_closed_form(textured, kind, mode):
    """TEXCOORD_0 as written (float, and normalized u16 as c / 65535); the base colour
    texture's image: the texture's index in the reference mode (the reference's importer
    indexes images with it), the texture's source in the factor mode; the factor white in
    the reference mode; images as stored (bufferView, data URI, a file beside the scene),
    decoding to the texels written."""
    paths, texels, expect, _ = textured
    g = vgi.Gltf(paths[kind], albedo_mode=mode)
    assert g.info().num_images == 3
    prims = g.primitives()
    assert len(prims) == len(expect)
    source = {0: 1, 1: 0, 2: 2}  # texture -> image, as written
    for p, e in zip(prims, expect):
        if e["uv"] is None:
            assert p["texcoords"] is None and p["image"] == -1
            continue
        assert np.array_equal(p["texcoords"], e["uv"])
        assert p["image"] == (e["texture"] if mode == 0 else source[e["texture"]])
        if mode == 0:
            assert p["base_color_factor"] == (1.0, 1.0, 1.0, 1.0)
    im = [g.image(k) for k in range(3)]
    assert im[0]["mime_type"] == "image/png" and im[0]["uri"] == ""
    assert im[1]["mime_type"] == "image/png" and im[1]["uri"] == ""
    assert im[2]["uri"] == "tex_two.png"
    for k in range(3):
        assert np.array_equal(gw.png_decode_rgba(im[k]["data"]), texels[k])


@pytest.mark.parametrize("kind", ["gltf", "glb"])
@pytest.mark.parametrize("mode", [0, 1])
def test_textures_match_reference_parser(textured, kind, mode):
    L = _ref()
    paths, texels, _, _ = textured
    ref, ref_images = _ref_textures(L, paths[kind], mode)
    g = vgi.Gltf(paths[kind], albedo_mode=mode)
    got = g.primitives()
    assert len(got) == len(ref)
    for a, b in zip(got, ref):
        if b["texcoords"] is None:
            assert a["texcoords"] is None
        else:
            assert np.array_equal(a["texcoords"], b["texcoords"])
        assert a["image"] == b["image"]
        assert np.allclose(a["base_color_factor"], b["base_color_factor"], rtol=0, atol=0)
    # the bytes the library hands over decode to the pixels tinygltf decoded
    assert len(ref_images) == 3
    for k in range(3):
        assert np.array_equal(ref_images[k], texels[k])
        assert np.array_equal(gw.png_decode_rgba(g.image(k)["data"]), ref_images[k])


def test_percent_encoded_image_uri(tmp_path):
    """An image file whose name has a space is named "tex%20two.png" in the uri, which the
    reader decodes (glTF URIs are RFC 3986 references)."""
    w, texels, _ = gw.textured_scene(file_dir=str(tmp_path), file_name="tex two.png")
    path = str(tmp_path / "s.gltf")
    w.write_gltf(path)
    g = vgi.Gltf(path)
    assert g.image(2)["uri"] == "tex two.png"
    assert np.array_equal(gw.png_decode_rgba(g.image(2)["data"]), texels[2])


def test_external_image_is_read_on_request(textured, tmp_path):
    paths, texels, _, d = textured
    with open(paths["gltf"], "rb") as f:
        js = f.read()
    g = vgi.Gltf(data=js, base_dir=d)
    assert np.array_equal(gw.png_decode_rgba(g.image(2)["data"]), texels[2])
    # the same document elsewhere: the file is not there
    other = tmp_path / "elsewhere"
    other.mkdir()
    import shutil
    shutil.copy(paths["gltf"].replace(".gltf", ".bin"), str(other / "tex.bin"))
    g2 = vgi.Gltf(data=js, base_dir=str(other))
    assert g2.image(0)["data"]  # embedded images need no file
    with pytest.raises(vgi.VgiError):
        g2.image(2)
    with pytest.raises(vgi.VgiError):
        g2.set_image_rgba8(3, np.zeros((2, 2, 4), np.uint8))
