"""glTF 2.0 input (aph_gltf_load), the translation of Plasma's glTF loader
(Applications/Plasma/glTFLoader.cpp:26-386).  CPU only: files are written here from
known arrays, and the loaded meshes, node transforms, materials and cameras are
checked against those arrays and closed forms - no reference glTF files exist in the
reference tree, and tinygltf (its parser) is not here."""
import base64
import json
import math
import os
import struct

import numpy as np
import pytest

from aurora_amd import capi

QUAD_P = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], np.float32)
QUAD_N = np.array([[0, 0, 1]] * 4, np.float32)
QUAD_T = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32)
QUAD_I = np.array([0, 1, 2, 0, 2, 3], np.uint32)


class Doc:
    """A glTF document under construction: one binary buffer, views and accessors."""

    def __init__(self):
        self.bin = bytearray()
        self.j = {"asset": {"version": "2.0"}, "buffers": [], "bufferViews": [], "accessors": [],
                  "meshes": [], "nodes": [], "scenes": [{"nodes": []}], "materials": []}

    def view(self, data, stride=None):
        while len(self.bin) % 4:
            self.bin.append(0)
        off = len(self.bin)
        self.bin += data
        v = {"buffer": 0, "byteOffset": off, "byteLength": len(data)}
        if stride:
            v["byteStride"] = stride
        self.j["bufferViews"].append(v)
        return len(self.j["bufferViews"]) - 1

    def accessor(self, arr, ctype=5126, typ=None, view=None, offset=0, count=None):
        arr = np.asarray(arr)
        if view is None:
            view = self.view(arr.tobytes())
        typ = typ or {1: "SCALAR", 2: "VEC2", 3: "VEC3", 4: "VEC4"}[arr.shape[1] if arr.ndim > 1 else 1]
        a = {"bufferView": view, "componentType": ctype, "count": count if count is not None else len(arr),
             "type": typ}
        if offset:
            a["byteOffset"] = offset
        self.j["accessors"].append(a)
        return len(self.j["accessors"]) - 1

    def mesh(self, prims):
        self.j["meshes"].append({"primitives": prims})
        return len(self.j["meshes"]) - 1

    def quad(self, normals=True, uvs=True, index_type=5123, material=None):
        attrs = {"POSITION": self.accessor(QUAD_P)}
        if normals:
            attrs["NORMAL"] = self.accessor(QUAD_N)
        if uvs:
            attrs["TEXCOORD_0"] = self.accessor(QUAD_T)
        p = {"attributes": attrs}
        if index_type is not None:
            dt = {5121: np.uint8, 5123: np.uint16, 5125: np.uint32, 5126: np.float32}[index_type]
            p["indices"] = self.accessor(QUAD_I.astype(dt), ctype=index_type, typ="SCALAR")
        if material is not None:
            p["material"] = material
        return p

    def node(self, root=True, **kw):
        self.j["nodes"].append(kw)
        n = len(self.j["nodes"]) - 1
        if root:
            self.j["scenes"][0]["nodes"].append(n)
        return n

    def write_gltf(self, path, external=None):
        j = dict(self.j)
        if external:
            with open(os.path.join(os.path.dirname(path), external), "wb") as f:
                f.write(self.bin)
            uri = external.replace(" ", "%20")
        else:
            uri = "data:application/octet-stream;base64," + base64.b64encode(bytes(self.bin)).decode()
        j["buffers"] = [{"uri": uri, "byteLength": len(self.bin)}]
        with open(path, "w") as f:
            json.dump(j, f)
        return path

    def write_glb(self, path):
        j = dict(self.j)
        j["buffers"] = [{"byteLength": len(self.bin)}]
        js = json.dumps(j).encode()
        js += b" " * (-len(js) % 4)
        bn = bytes(self.bin) + b"\0" * (-len(self.bin) % 4)
        total = 12 + 8 + len(js) + 8 + len(bn)
        with open(path, "wb") as f:
            f.write(struct.pack("<4sII", b"glTF", 2, total))
            f.write(struct.pack("<I4s", len(js), b"JSON") + js)
            f.write(struct.pack("<I4s", len(bn), b"BIN\0") + bn)
        return path


def quat_matrix(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def trs(t=(0, 0, 0), q=(0, 0, 0, 1), s=(1, 1, 1)):
    m = np.eye(4)
    m[:3, :3] = quat_matrix(q) * np.asarray(s, dtype=np.float64)
    m[:3, 3] = t
    return m


def test_quad_with_data_uri(tmp_path):
    d = Doc()
    m = d.mesh([d.quad()])
    d.node(mesh=m)
    g = capi.load_gltf(d.write_gltf(str(tmp_path / "q.gltf")))
    assert len(g["meshes"]) == 1 and len(g["instances"]) == 1
    me = g["meshes"][0]
    np.testing.assert_array_equal(me["positions"], QUAD_P)
    np.testing.assert_array_equal(me["normals"], QUAD_N)
    np.testing.assert_array_equal(me["uvs"], QUAD_T)  # as stored: no V flip (unlike OBJ)
    np.testing.assert_array_equal(me["indices"].reshape(-1), QUAD_I)
    inst = g["instances"][0]
    np.testing.assert_array_equal(inst["o2w"], np.eye(4)[:3].reshape(-1).astype(np.float32))
    assert inst["material"] == -1 and inst["node"] == 0 and inst["materialx"] is None


def test_node_hierarchy_composes_parent_times_trs(tmp_path):
    d = Doc()
    m = d.mesh([d.quad()])
    s = math.sqrt(0.5)
    child = d.node(root=False, mesh=m, translation=[0.5, 0, 0], rotation=[0, s, 0, s], scale=[2, 3, 4])
    d.node(children=[child], translation=[1, 2, 3], rotation=[s, 0, 0, s])
    mat = [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 7, 8, 9, 1]  # column-major: translation (7, 8, 9)
    d.node(mesh=m, matrix=mat)
    g = capi.load_gltf(d.write_gltf(str(tmp_path / "h.gltf")))
    assert len(g["meshes"]) == 1 and len(g["instances"]) == 2  # one mesh, two instances
    parent = trs((1, 2, 3), (s, 0, 0, s))
    local = trs((0.5, 0, 0), (0, s, 0, s), (2, 3, 4))
    want = (parent @ local)[:3].reshape(-1)
    by_node = {i["node"]: i for i in g["instances"]}
    np.testing.assert_allclose(by_node[child]["o2w"], want, atol=2e-6)
    want2 = np.eye(4)
    want2[:3, 3] = (7, 8, 9)
    np.testing.assert_array_equal(by_node[2]["o2w"], want2[:3].reshape(-1).astype(np.float32))


def test_quarter_turn_quaternion_is_exact(tmp_path):
    # q = (0, 0, 1, 0) is a half turn about z: the matrix entries are exact in float
    d = Doc()
    m = d.mesh([d.quad()])
    d.node(mesh=m, rotation=[0, 0, 1, 0], translation=[1, 0, 0])
    g = capi.load_gltf(d.write_gltf(str(tmp_path / "r.gltf")))
    np.testing.assert_array_equal(g["instances"][0]["o2w"],
                                  np.array([-1, 0, 0, 1, 0, -1, 0, 0, 0, 0, 1, 0], np.float32))


@pytest.mark.parametrize("index_type", [5121, 5123, 5125])
def test_index_types(tmp_path, index_type):
    d = Doc()
    d.node(mesh=d.mesh([d.quad(index_type=index_type)]))
    g = capi.load_gltf(d.write_gltf(str(tmp_path / "i.gltf")))
    np.testing.assert_array_equal(g["meshes"][0]["indices"].reshape(-1), QUAD_I)


@pytest.mark.parametrize("index_type", [None, 5126])
def test_missing_or_unsupported_indices_become_sequential(tmp_path, index_type):
    # glTFLoader.cpp:213-222: no (usable) index accessor -> 0, 1, 2, ...
    d = Doc()
    d.node(mesh=d.mesh([d.quad(index_type=index_type)]))
    g = capi.load_gltf(d.write_gltf(str(tmp_path / "s.gltf")))
    np.testing.assert_array_equal(g["meshes"][0]["indices"].reshape(-1), [0, 1, 2])  # 4 vertices: one triangle


def test_interleaved_strided_view(tmp_path):
    d = Doc()
    inter = np.concatenate([QUAD_P, QUAD_N], axis=1).astype(np.float32)  # 24-byte records
    v = d.view(inter.tobytes(), stride=24)
    pa = d.accessor(QUAD_P, view=v, count=4)
    na = d.accessor(QUAD_N, view=v, offset=12, count=4)
    ia = d.accessor(QUAD_I.astype(np.uint16), ctype=5123, typ="SCALAR")
    d.node(mesh=d.mesh([{"attributes": {"POSITION": pa, "NORMAL": na}, "indices": ia}]))
    g = capi.load_gltf(d.write_gltf(str(tmp_path / "st.gltf")))
    np.testing.assert_array_equal(g["meshes"][0]["positions"], QUAD_P)
    np.testing.assert_array_equal(g["meshes"][0]["normals"], QUAD_N)
    assert g["meshes"][0]["uvs"] is None


def test_missing_normals_equal_the_obj_loaders(tmp_path):
    # both loaders call Foundation's calculateNormals: a bent strip gives the same
    # angle-weighted normals bit for bit
    p = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0.3], [0, 1, 0.3], [2, 0, 0.5], [2, 1, 0.9]], np.float32)
    idx = np.array([0, 1, 2, 0, 2, 3, 1, 4, 5, 1, 5, 2], np.uint32)
    d = Doc()
    ia = d.accessor(idx, ctype=5125, typ="SCALAR")
    d.node(mesh=d.mesh([{"attributes": {"POSITION": d.accessor(p)}, "indices": ia}]))
    g = capi.load_gltf(d.write_gltf(str(tmp_path / "n.gltf")))
    with open(tmp_path / "n.obj", "w") as f:
        for v in p:
            f.write("v %.9g %.9g %.9g\n" % tuple(v))
        for t in idx.reshape(-1, 3):
            f.write("f %d %d %d\n" % tuple(t + 1))
    o = capi.load_obj(str(tmp_path / "n.obj"))
    np.testing.assert_array_equal(g["meshes"][0]["normals"], o["shapes"][0]["normals"])
    # and they are unit length, pointing to +z on the flat-ish strip
    assert np.allclose(np.linalg.norm(g["meshes"][0]["normals"], axis=1), 1, atol=1e-6)


def test_glb_and_external_buffer_equal_the_data_uri(tmp_path):
    d = Doc()
    d.j["materials"].append({"pbrMetallicRoughness": {"baseColorFactor": [0.2, 0.4, 0.6, 1],
                                                      "metallicFactor": 0.25, "roughnessFactor": 0.5}})
    d.node(mesh=d.mesh([d.quad(material=0)]), translation=[0, 1, 0])
    a = capi.load_gltf(d.write_gltf(str(tmp_path / "a.gltf")))
    b = capi.load_gltf(d.write_glb(str(tmp_path / "b.glb")))
    c = capi.load_gltf(d.write_gltf(str(tmp_path / "c.gltf"), external="my buffer.bin"))
    for x in (b, c):
        for k in ("positions", "normals", "uvs", "indices"):
            np.testing.assert_array_equal(x["meshes"][0][k], a["meshes"][0][k])
        np.testing.assert_array_equal(x["instances"][0]["o2w"], a["instances"][0]["o2w"])
        assert bytes(x["materials"][0]) == bytes(a["materials"][0])


def test_materials_translate_pbr_metallic_roughness(tmp_path):
    d = Doc()
    d.j["materials"] = [{"pbrMetallicRoughness": {"baseColorFactor": [0.2, 0.4, 0.6, 0.5],
                                                  "metallicFactor": 0.25, "roughnessFactor": 0.75}},
                        {"name": "glTF defaults"}]
    d.node(mesh=d.mesh([d.quad(material=0), d.quad(material=1), d.quad()]))
    g = capi.load_gltf(d.write_gltf(str(tmp_path / "m.gltf")))
    ref = capi.default_material()
    m0, m1 = g["materials"]
    assert tuple(m0.base_color) == tuple(np.float32([0.2, 0.4, 0.6]))
    assert m0.metalness == np.float32(0.25) and m0.specular_roughness == np.float32(0.75)
    assert tuple(m1.base_color) == (1.0, 1.0, 1.0) and m1.metalness == 1.0 and m1.specular_roughness == 1.0
    # every other field keeps Aurora's default
    for f, _ in capi.Material._fields_:
        if f not in ("base_color", "metalness", "specular_roughness"):
            assert np.array_equal(np.array(getattr(m0, f)), np.array(getattr(ref, f))), f
    assert [i["material"] for i in g["instances"]] == [0, 1, -1]
    assert len(g["meshes"]) == 3


def test_skipped_primitives(tmp_path):
    # a LINES primitive and a primitive whose POSITION is not FLOAT are skipped
    d = Doc()
    lines = d.quad()
    lines["mode"] = 1
    ushort = {"attributes": {"POSITION": d.accessor(QUAD_P.astype(np.uint16), ctype=5123)}}
    d.node(mesh=d.mesh([lines, ushort, d.quad()]))
    g = capi.load_gltf(d.write_gltf(str(tmp_path / "k.gltf")))
    assert len(g["meshes"]) == 1 and len(g["instances"]) == 1


def test_cameras(tmp_path):
    d = Doc()
    d.node(mesh=d.mesh([d.quad()]))
    d.j["cameras"] = [{"type": "perspective", "perspective": {"yfov": 0.8, "aspectRatio": 1.5, "znear": 0.1,
                                                                "zfar": 100}},
                      {"type": "orthographic", "orthographic": {"xmag": 2, "ymag": 1, "znear": 0, "zfar": 10}}]
    d.node(camera=0, translation=[1, 2, 10])
    d.node(camera=1, rotation=[0, 1, 0, 0])
    g = capi.load_gltf(d.write_gltf(str(tmp_path / "c.gltf")))
    c0, c1 = g["cameras"]
    assert c0["type"] == 0 and c1["type"] == 1
    assert c0["yfov"] == np.float32(0.8) and c0["aspect"] == np.float32(1.5) and c0["zfar"] == 100
    assert c1["xmag"] == 2 and c1["ymag"] == 1
    want = np.eye(4)
    want[:3, 3] = (-1, -2, -10)  # the inverse of a translation
    np.testing.assert_array_equal(c0["view"], want.astype(np.float32))
    np.testing.assert_array_equal(c1["view"], np.diag([-1, 1, -1, 1]).astype(np.float32))


def test_materialx_extension_is_reported(tmp_path):
    d = Doc()
    d.node(mesh=d.mesh([d.quad()]), extensions={"AURORA_materials_materialX": {"materialXFile": "m/gold.mtlx"}})
    g = capi.load_gltf(d.write_gltf(str(tmp_path / "x.gltf")))
    assert g["instances"][0]["materialx"] == str(tmp_path / "m" / "gold.mtlx")


def test_malformed_files_are_rejected(tmp_path):
    d = Doc()
    d.node(mesh=d.mesh([d.quad()]))
    good = d.write_gltf(str(tmp_path / "g.gltf"))
    capi.load_gltf(good)
    text = open(good).read()
    cases = {
        "truncated.gltf": text[:-5],
        "badjson.gltf": text.replace('"asset"', "asset"),
        "node.gltf": text.replace('"scenes": [{"nodes": [0]}]', '"scenes": [{"nodes": [7]}]'),
    }
    j = json.loads(text)
    j["accessors"][0]["count"] = 400  # reads past its buffer view
    cases["bounds.gltf"] = json.dumps(j)
    j = json.loads(text)
    j["buffers"][0]["uri"] = "missing.bin"
    cases["missing.gltf"] = json.dumps(j)
    for name, body in cases.items():
        p = tmp_path / name
        p.write_text(body)
        with pytest.raises(RuntimeError):
            capi.load_gltf(str(p))
    with pytest.raises(RuntimeError):
        capi.load_gltf(str(tmp_path / "nonexistent.gltf"))
    glb = d.write_glb(str(tmp_path / "g.glb"))
    raw = bytearray(open(glb, "rb").read())
    raw[4] = 3  # version 3
    (tmp_path / "v3.glb").write_bytes(bytes(raw))
    with pytest.raises(RuntimeError):
        capi.load_gltf(str(tmp_path / "v3.glb"))


def test_hostile_sizes_and_node_graphs_are_rejected(tmp_path):
    """Counts, offsets and lengths that are negative, fractional, beyond 2^53 or large
    enough to wrap the bounds arithmetic are errors, as are node indices beyond int; a node
    graph that shares children (not a tree) fails at the visit cap instead of walking
    2^64 paths."""
    d = Doc()
    d.node(mesh=d.mesh([d.quad()]))
    text = open(d.write_gltf(str(tmp_path / "g.gltf"))).read()
    edits = [
        ("accessors", 0, "count", 2 ** 62),  # (count - 1) * stride wraps a 64-bit size
        ("accessors", 0, "count", 1.5),
        ("accessors", 0, "count", -4),
        ("accessors", 0, "byteOffset", -16),
        ("accessors", 0, "byteOffset", 1e300),
        ("bufferViews", 0, "byteStride", 2 ** 61),
        ("bufferViews", 0, "byteLength", 2 ** 64),
        ("accessors", 0, "bufferView", 1e20),
        ("buffers", 0, "byteLength", -1),
    ]
    for k, (table, i, key, value) in enumerate(edits):
        j = json.loads(text)
        j[table][i][key] = value
        p = tmp_path / f"e{k}.gltf"
        p.write_text(json.dumps(j))
        with pytest.raises(RuntimeError):
            capi.load_gltf(str(p))
    j = json.loads(text)
    j["scenes"][0]["nodes"] = [1e20]
    (tmp_path / "root.gltf").write_text(json.dumps(j))
    with pytest.raises(RuntimeError):
        capi.load_gltf(str(tmp_path / "root.gltf"))
    # 60 levels, each node listing the next twice: 2^60 paths through 61 nodes
    j = json.loads(text)
    j["nodes"] = [{"children": [k + 1, k + 1]} for k in range(60)] + [{"mesh": 0}]
    j["scenes"][0]["nodes"] = [0]
    (tmp_path / "dag.gltf").write_text(json.dumps(j))
    with pytest.raises(RuntimeError):
        capi.load_gltf(str(tmp_path / "dag.gltf"))
    # the same chain as a tree loads, with its one instance
    j["nodes"] = [{"children": [k + 1]} for k in range(60)] + [{"mesh": 0}]
    (tmp_path / "chain.gltf").write_text(json.dumps(j))
    assert len(capi.load_gltf(str(tmp_path / "chain.gltf"))["instances"]) == 1


def test_json_escapes_and_numbers(tmp_path):
    # string escapes (\\u, surrogate pairs) and exponent numbers parse as JSON defines them
    d = Doc()
    d.node(mesh=d.mesh([d.quad()]), name="café \U0001F600 \"q\"", scale=[1e0, 2.5E-1, -4])
    g = capi.load_gltf(d.write_gltf(str(tmp_path / "e.gltf")))
    np.testing.assert_array_equal(g["instances"][0]["o2w"],
                                  np.array([1, 0, 0, 0, 0, 0.25, 0, 0, 0, 0, -4, 0], np.float32))


def hamilton(a, b):
    ax, ay, az, aw = a
    bx, by, bz, bw = b
    return np.array([aw * bx + ax * bw + ay * bz - az * by, aw * by - ax * bz + ay * bw + az * bx,
                     aw * bz + ax * by - ay * bx + az * bw, aw * bw - ax * bx - ay * by - az * bz])


def test_rotation_matches_the_quaternion_product(tmp_path):
    # independent of the matrix formula: the node's transform applied to a point equals
    # q (p, 0) q* for a unit quaternion
    rng = np.random.default_rng(3)
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    d = Doc()
    d.node(mesh=d.mesh([d.quad()]), rotation=[float(x) for x in q])
    g = capi.load_gltf(d.write_gltf(str(tmp_path / "q.gltf")))
    m = g["instances"][0]["o2w"].reshape(3, 4).astype(np.float64)
    for p in rng.normal(size=(5, 3)):
        qc = np.array([-q[0], -q[1], -q[2], q[3]])
        want = hamilton(hamilton(q, np.append(p, 0.0)), qc)[:3]
        np.testing.assert_allclose(m[:, :3] @ p, want, atol=1e-5)
