"""Writes small glTF 2.0 files with known contents for the glTF reader's tests.

A scene is built in memory (buffers, views, accessors, meshes, nodes, materials) and
emitted as .gltf with an external .bin, as .gltf with base64 data URIs, or as .glb. The
"kitchen sink" scene exercises what the reader supports: interleaved positions
(byteStride 24), an accessor byteOffset inside its view, u8 / u16 / u32 / no indices,
triangle lists, strips and fans, a line primitive (skipped), a primitive without a
material, a mesh placed by two nodes, node matrices and TRS, and a child under a parent.
"""
from __future__ import annotations

import base64
import json
import math
import struct
import zlib

import numpy as np


def png_rgba(texels: np.ndarray) -> bytes:
    """A PNG of RGBA8 texels (h, w, 4): filter 0 on every row, one zlib stream."""
    t = np.ascontiguousarray(texels, np.uint8)
    h, w, _ = t.shape
    raw = b"".join(b"\x00" + t[y].tobytes() for y in range(h))

    def chunk(tag, data):
        return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)

    return (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 6, 0, 0, 0)) +
            chunk(b"IDAT", zlib.compress(raw)) + chunk(b"IEND", b""))


def png_decode_rgba(data: bytes) -> np.ndarray:
    """Decodes the PNGs png_rgba writes (8-bit RGBA, filter 0 rows) back to (h, w, 4)."""
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    off, idat = 8, b""
    w = h = 0
    while off < len(data):
        n, tag = struct.unpack(">I4s", data[off:off + 8])
        body = data[off + 8:off + 8 + n]
        if tag == b"IHDR":
            w, h, depth, ctype = struct.unpack(">IIBB", body[:10])
            assert depth == 8 and ctype == 6
        elif tag == b"IDAT":
            idat += body
        off += 12 + n
    raw = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(h, 1 + 4 * w)
    assert (raw[:, 0] == 0).all()
    return raw[:, 1:].reshape(h, w, 4).copy()


class Writer:
    def __init__(self):
        self.blob = bytearray()
        self.doc = {"asset": {"version": "2.0", "generator": "tests"}, "scene": 0, "scenes": [{"nodes": []}],
                    "nodes": [], "meshes": [], "materials": [], "accessors": [], "bufferViews": []}

    def _view(self, data: bytes, stride: int | None = None, target: int | None = None) -> int:
        while len(self.blob) % 4:
            self.blob.append(0)
        v = {"buffer": 0, "byteOffset": len(self.blob), "byteLength": len(data)}
        if stride:
            v["byteStride"] = stride
        if target:
            v["target"] = target
        self.blob += data
        self.doc["bufferViews"].append(v)
        return len(self.doc["bufferViews"]) - 1

    def _accessor(self, view: int, comp: int, count: int, type_: str, offset: int = 0, minmax=None) -> int:
        a = {"bufferView": view, "componentType": comp, "count": count, "type": type_}
        if offset:
            a["byteOffset"] = offset
        if minmax is not None:
            a["min"], a["max"] = [float(x) for x in minmax[0]], [float(x) for x in minmax[1]]
        self.doc["accessors"].append(a)
        return len(self.doc["accessors"]) - 1

    def positions(self, pos: np.ndarray, interleave_normals: bool = False, lead_bytes: int = 0) -> int:
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        mm = (pos.min(0), pos.max(0))
        if interleave_normals:
            inter = np.zeros((len(pos), 6), np.float32)
            inter[:, :3] = pos
            inter[:, 5] = 1.0
            v = self._view(inter.tobytes(), stride=24, target=34962)
            return self._accessor(v, 5126, len(pos), "VEC3", 0, mm)
        v = self._view(b"\0" * lead_bytes + pos.tobytes(), target=34962)
        return self._accessor(v, 5126, len(pos), "VEC3", lead_bytes, mm)

    def indices(self, idx, dtype) -> int:
        a = np.ascontiguousarray(idx, dtype).reshape(-1)
        comp = {np.uint8: 5121, np.uint16: 5123, np.uint32: 5125}[dtype]
        v = self._view(a.tobytes(), target=34963)
        return self._accessor(v, comp, a.size, "SCALAR")

    def texcoords(self, uv: np.ndarray, normalized_u16: bool = False) -> int:
        """TEXCOORD_0: float VEC2, or normalized unsigned short (rounded from uv in [0, 1])."""
        uv = np.ascontiguousarray(uv, np.float32).reshape(-1, 2)
        if normalized_u16:
            q = np.round(np.clip(uv, 0, 1) * 65535).astype(np.uint16)
            v = self._view(q.tobytes(), target=34962)
            a = self._accessor(v, 5123, len(uv), "VEC2")
            self.doc["accessors"][a]["normalized"] = True
            return a
        v = self._view(uv.tobytes(), target=34962)
        return self._accessor(v, 5126, len(uv), "VEC2")

    def image(self, texels: np.ndarray, how: str = "bufferView", file_dir: str | None = None,
              name: str | None = None) -> int:
        """An RGBA8 PNG image: in a bufferView, as a data URI, or as a file beside the scene."""
        data = png_rgba(texels)
        im = {}
        if how == "bufferView":
            im = {"bufferView": self._view(data), "mimeType": "image/png"}
        elif how == "datauri":
            im = {"uri": "data:image/png;base64," + base64.b64encode(data).decode()}
        else:
            fname = name or f"image{len(self.doc.setdefault('images', []))}.png"
            with open(file_dir + "/" + fname, "wb") as f:
                f.write(data)
            im = {"uri": fname.replace(" ", "%20")}
        self.doc.setdefault("images", []).append(im)
        return len(self.doc["images"]) - 1

    def texture(self, source: int) -> int:
        self.doc.setdefault("samplers", [{"magFilter": 9729, "minFilter": 9987, "wrapS": 10497, "wrapT": 10497}])
        self.doc.setdefault("textures", []).append({"sampler": 0, "source": source})
        return len(self.doc["textures"]) - 1

    def textured_material(self, rgba, texture: int) -> int:
        self.doc["materials"].append({"pbrMetallicRoughness": {"baseColorFactor": [float(c) for c in rgba],
                                                               "baseColorTexture": {"index": texture},
                                                               "metallicFactor": 0.0}})
        return len(self.doc["materials"]) - 1

    def material(self, rgba) -> int:
        self.doc["materials"].append({"pbrMetallicRoughness": {"baseColorFactor": [float(c) for c in rgba],
                                                               "metallicFactor": 0.0}})
        return len(self.doc["materials"]) - 1

    def mesh(self, prims) -> int:
        self.doc["meshes"].append({"primitives": prims})
        return len(self.doc["meshes"]) - 1

    def node(self, **kw) -> int:
        self.doc["nodes"].append(kw)
        return len(self.doc["nodes"]) - 1

    def _finish(self) -> dict:
        d = json.loads(json.dumps(self.doc))
        d["buffers"] = [{"byteLength": len(self.blob)}]
        return d

    def write_gltf(self, path: str, embed: bool = False, bin_name: str | None = None) -> None:
        """bin_name may contain characters that the uri percent-encodes (e.g. a space)."""
        d = self._finish()
        if embed:
            d["buffers"][0]["uri"] = "data:application/octet-stream;base64," + base64.b64encode(self.blob).decode()
        else:
            binname = bin_name or path.rsplit("/", 1)[-1].replace(".gltf", ".bin")
            d["buffers"][0]["uri"] = binname.replace("%", "%25").replace(" ", "%20")
            with open(path.rsplit("/", 1)[0] + "/" + binname, "wb") as f:
                f.write(self.blob)
        with open(path, "w") as f:
            json.dump(d, f, indent=1)

    def glb_bytes(self) -> bytes:
        js = json.dumps(self._finish()).encode()
        js += b" " * ((4 - len(js) % 4) % 4)
        bin_ = bytes(self.blob) + b"\0" * ((4 - len(self.blob) % 4) % 4)
        total = 12 + 8 + len(js) + 8 + len(bin_)
        return (struct.pack("<III", 0x46546C67, 2, total) + struct.pack("<II", len(js), 0x4E4F534A) + js +
                struct.pack("<II", len(bin_), 0x004E4942) + bin_)

    def write_glb(self, path: str) -> None:
        with open(path, "wb") as f:
            f.write(self.glb_bytes())


def _quat(axis, deg):
    a = np.asarray(axis, np.float64)
    a = a / np.linalg.norm(a)
    h = math.radians(deg) / 2
    return [float(a[0] * math.sin(h)), float(a[1] * math.sin(h)), float(a[2] * math.sin(h)), float(math.cos(h))]


def trs_matrix(t=(0, 0, 0), q=(0, 0, 0, 1), s=(1, 1, 1)) -> np.ndarray:
    """4x4 T*R*S from the quaternion's rotation matrix (x, y, z, w), float64."""
    x, y, z, w = q
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    M = np.eye(4)
    M[:3, :3] = R @ np.diag(s)
    M[:3, 3] = t
    return M


def kitchen_sink(seed: int = 5):
    """The scene plus what a reader must produce: a list of placed primitives in visit
    order, each {"world": (T, 3, 3) float64 triangles, "rgba_factor", "material"}."""
    rng = np.random.default_rng(seed)
    w = Writer()
    m_red = w.material((0.8, 0.4, 0.2, 1.0))
    m_blue = w.material((0.1, 0.2, 0.9, 0.5))
    # mesh 0: interleaved list with u16 indices (material red) + a non-indexed strip (blue)
    p0 = rng.uniform(-0.5, 0.5, (40, 3)).astype(np.float32)
    i0 = rng.integers(0, 40, 60).astype(np.uint16)
    s0 = rng.uniform(-0.5, 0.5, (9, 3)).astype(np.float32)
    mesh0 = w.mesh([
        {"attributes": {"POSITION": w.positions(p0, interleave_normals=True)}, "indices": w.indices(i0, np.uint16),
         "material": m_red},
        {"attributes": {"POSITION": w.positions(s0)}, "mode": 5, "material": m_blue},
    ])
    # mesh 1: fan with u8 indices (no material) + a line primitive (skipped)
    p1 = rng.uniform(-0.3, 0.3, (12, 3)).astype(np.float32)
    f1 = np.array([0, 3, 5, 7, 9, 11, 2], np.uint8)
    mesh1 = w.mesh([
        {"attributes": {"POSITION": w.positions(p1)}, "indices": w.indices(f1, np.uint8), "mode": 6},
        {"attributes": {"POSITION": w.positions(p1)}, "mode": 1, "material": m_red},
    ])
    # mesh 2: u32 indices, POSITION at a byte offset inside its view (material blue)
    p2 = rng.uniform(-0.2, 0.2, (20, 3)).astype(np.float32)
    i2 = rng.integers(0, 20, 30).astype(np.uint32)
    mesh2 = w.mesh([{"attributes": {"POSITION": w.positions(p2, lead_bytes=16)},
                     "indices": w.indices(i2, np.uint32), "material": m_blue}])
    # nodes: 0 (matrix: 90 deg about z + translation) with children 1 (TRS, mesh 0) and
    # 2 (translation, mesh 1); 3 (scale, mesh 0 again) with child 4 (mesh 2)
    Mz = trs_matrix(t=(0.1, -0.2, 0.3), q=_quat((0, 0, 1), 90))
    q1 = _quat((1, 0, 0), 90)
    n1 = w.node(mesh=mesh0, translation=[0.05, 0.0, 0.0], rotation=q1, scale=[2.0, 2.0, 2.0])
    n2 = w.node(mesh=mesh1, translation=[0.0, 0.25, 0.0])
    n0 = w.node(matrix=[float(v) for v in Mz.T.reshape(-1)], children=[n1, n2])
    q4 = _quat((1, 1, 0), 30)
    n4 = w.node(mesh=mesh2, rotation=q4, translation=[0.0, 0.0, -0.1])
    n3 = w.node(mesh=mesh0, scale=[0.5, 0.5, 0.5], children=[n4])
    w.doc["scenes"][0]["nodes"] = [n0, n3]
    W1 = Mz @ trs_matrix(t=(0.05, 0, 0), q=q1, s=(2, 2, 2))
    W2 = Mz @ trs_matrix(t=(0, 0.25, 0))
    W3 = trs_matrix(s=(0.5, 0.5, 0.5))
    W4 = W3 @ trs_matrix(t=(0, 0, -0.1), q=q4)

    def world(M, pos, tri):
        p = np.asarray(pos, np.float64) @ M[:3, :3].T + M[:3, 3]
        return p[np.asarray(tri).reshape(-1, 3)]

    list0 = i0.reshape(-1, 3)
    strip0 = np.array([(i, i + 1 + (i & 1), i + 2 - (i & 1)) for i in range(len(s0) - 2)])
    fan1 = np.array([(f1[i + 1], f1[i + 2], f1[0]) for i in range(len(f1) - 2)])
    list2 = i2.reshape(-1, 3)
    red, blue, white = (0.8, 0.4, 0.2, 1.0), (0.1, 0.2, 0.9, 0.5), (1.0, 1.0, 1.0, 1.0)
    expect = [
        dict(world=world(W1, p0, list0), rgba_factor=red, material=m_red, model=W1),
        dict(world=world(W1, s0, strip0), rgba_factor=blue, material=m_blue, model=W1),
        dict(world=world(W2, p1, fan1), rgba_factor=white, material=-1, model=W2),
        dict(world=world(W3, p0, list0), rgba_factor=red, material=m_red, model=W3),
        dict(world=world(W3, s0, strip0), rgba_factor=blue, material=m_blue, model=W3),
        dict(world=world(W4, p2, list2), rgba_factor=blue, material=m_blue, model=W4),
    ]
    return w, expect


def textured_scene(seed: int = 9, file_dir: str | None = None, file_name: str = "tex_two.png"):
    """Three textured primitives and one plain one. Images: 0 in a bufferView, 1 as a data
    URI, 2 as a file beside the scene when file_dir is given (else a bufferView), named
    file_name (percent-encoded in the uri). Texture 0
    names image 1 and texture 1 names image 0, so a reader that takes the texture's index
    as the image index (the reference's importer) and one that follows "source" differ;
    texture 2 names image 2. Returns the writer, the texels of each image and, per placed
    primitive, what a reader must give: uv per vertex, texture index and material."""
    rng = np.random.default_rng(seed)
    w = Writer()
    texels = [rng.integers(0, 256, (h, wd, 4), dtype=np.uint8) for h, wd in ((8, 16), (12, 12), (5, 7))]
    for t in texels:
        t[..., 3] = np.where(rng.random(t.shape[:2]) < 0.3, rng.integers(0, 20, t.shape[:2]), 255)
    i0 = w.image(texels[0], "bufferView")
    i1 = w.image(texels[1], "datauri")
    i2 = w.image(texels[2], "file" if file_dir else "bufferView", file_dir=file_dir, name=file_name)
    t0, t1, t2 = w.texture(i1), w.texture(i0), w.texture(i2)
    mats = [w.textured_material((0.9, 0.8, 1.0, 1.0), t0), w.textured_material((1.0, 0.5, 0.25, 0.8), t1),
            w.textured_material((1.0, 1.0, 1.0, 1.0), t2), w.material((0.3, 0.6, 0.9, 1.0))]
    prims, expect = [], []
    for k, mat in enumerate(mats):
        nv = 30
        pos = rng.uniform(-0.6, 0.6, (nv, 3)).astype(np.float32)
        uv = rng.uniform(-1.0, 2.0, (nv, 2)).astype(np.float32)
        pr = {"attributes": {"POSITION": w.positions(pos)}, "indices": w.indices(rng.integers(0, nv, 60), np.uint16),
              "material": mat}
        u16 = k == 2
        if u16:
            uv = np.round(np.clip(uv, 0, 1) * 65535).astype(np.uint16).astype(np.float32) / np.float32(65535.0)
        if k < 3:
            pr["attributes"]["TEXCOORD_0"] = w.texcoords(uv, normalized_u16=u16)
        prims.append(pr)
        expect.append(dict(uv=uv if k < 3 else None, texture=[t0, t1, t2, -1][k], material=mat))
    mesh = w.mesh(prims)
    w.doc["scenes"][0]["nodes"] = [w.node(mesh=mesh, translation=[0.05, -0.02, 0.01])]
    return w, texels, expect
