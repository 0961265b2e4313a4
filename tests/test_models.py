"""Model file readers (glTF 2.0, Panda3D egg) on files written here with known
contents, and on the reference's own model files when they are present."""
from __future__ import annotations

import base64
import json
import struct
from pathlib import Path

import numpy as np
import pytest

from pybatchrender.renderer import mesh as meshmod
from pybatchrender.renderer.model_io import load_egg, load_gltf

REF_MODELS = Path("/root/reference/pybatchrender/models")


def _quad_gltf(tmp: Path, *, glb=False, mode=4, normals=True, double_sided=False,
               node=None, data_uri=False, stride=0) -> Path:
    P = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], dtype=np.float32)
    N = np.tile(np.array([[0, 0, 1]], dtype=np.float32), (4, 1))
    if mode == 4:
        idx = np.array([0, 1, 2, 0, 2, 3], dtype=np.uint16)
    elif mode == 5:
        idx = np.array([0, 1, 3, 2], dtype=np.uint16)
    else:
        idx = np.array([0, 1, 2, 3], dtype=np.uint16)
    if stride:
        pad = np.zeros((4, stride // 4 - 3), dtype=np.float32)
        pbytes = np.concatenate([P, pad], 1).tobytes()
    else:
        pbytes = P.tobytes()
    blob = idx.tobytes() + b"\0" * (-len(idx.tobytes()) % 4)
    off_p = len(blob)
    blob += pbytes
    off_n = len(blob)
    blob += N.tobytes()
    views = [{"buffer": 0, "byteOffset": 0, "byteLength": idx.nbytes},
             {"buffer": 0, "byteOffset": off_p, "byteLength": len(pbytes),
              **({"byteStride": stride} if stride else {})},
             {"buffer": 0, "byteOffset": off_n, "byteLength": N.nbytes}]
    acc = [{"bufferView": 0, "componentType": 5123, "count": len(idx), "type": "SCALAR"},
           {"bufferView": 1, "componentType": 5126, "count": 4, "type": "VEC3"},
           {"bufferView": 2, "componentType": 5126, "count": 4, "type": "VEC3"}]
    attrs = {"POSITION": 1}
    if normals:
        attrs["NORMAL"] = 2
    doc = {"asset": {"version": "2.0"}, "accessors": acc, "bufferViews": views,
           "materials": [{"doubleSided": double_sided}],
           "meshes": [{"primitives": [{"attributes": attrs, "indices": 0, "mode": mode,
                                       "material": 0}]}],
           "nodes": [dict(mesh=0, **(node or {}))], "scenes": [{"nodes": [0]}], "scene": 0}
    if glb:
        doc["buffers"] = [{"byteLength": len(blob)}]
        js = json.dumps(doc).encode()
        js += b" " * (-len(js) % 4)
        bb = blob + b"\0" * (-len(blob) % 4)
        body = (struct.pack("<II", len(js), 0x4E4F534A) + js +
                struct.pack("<II", len(bb), 0x004E4942) + bb)
        out = tmp / "quad.glb"
        out.write_bytes(struct.pack("<III", 0x46546C67, 2, 12 + len(body)) + body)
        return out
    if data_uri:
        doc["buffers"] = [{"byteLength": len(blob),
                           "uri": "data:application/octet-stream;base64," +
                                  base64.b64encode(blob).decode()}]
    else:
        (tmp / "quad.bin").write_bytes(blob)
        doc["buffers"] = [{"byteLength": len(blob), "uri": "quad.bin"}]
    out = tmp / "quad.gltf"
    out.write_text(json.dumps(doc))
    return out


def _zup(p):
    x, y, z = p
    return (x, -z, y)


def test_gltf_triangles_and_yup_conversion(tmp_path):
    m = load_gltf(_quad_gltf(tmp_path))
    assert m.num_triangles == 2 and m.num_vertices == 4
    np.testing.assert_allclose(m.positions[2], _zup((1, 1, 0)))
    np.testing.assert_allclose(m.normals[0], _zup((0, 0, 1)))
    np.testing.assert_array_equal(m.triangles, [[0, 1, 2], [0, 2, 3]])


@pytest.mark.parametrize("kw", [dict(glb=True), dict(data_uri=True), dict(stride=16)])
def test_gltf_buffer_forms_agree(tmp_path, kw):
    a = load_gltf(_quad_gltf(tmp_path))
    b = load_gltf(_quad_gltf(tmp_path, **kw))
    np.testing.assert_array_equal(a.positions, b.positions)
    np.testing.assert_array_equal(a.triangles, b.triangles)


def test_gltf_strip_fan_and_double_sided(tmp_path):
    strip = load_gltf(_quad_gltf(tmp_path, mode=5))
    fan = load_gltf(_quad_gltf(tmp_path, mode=6))
    base = load_gltf(_quad_gltf(tmp_path))
    for m in (strip, fan):
        # same covered square, counter-clockwise seen from +normal
        n = np.cross(m.positions[m.triangles[:, 1]] - m.positions[m.triangles[:, 0]],
                     m.positions[m.triangles[:, 2]] - m.positions[m.triangles[:, 0]])
        assert (n @ base.normals[0] > 0).all() and len(m.triangles) == 2
    two = load_gltf(_quad_gltf(tmp_path, double_sided=True))
    assert two.num_triangles == 4
    np.testing.assert_array_equal(two.triangles[2:], two.triangles[:2, ::-1])


def test_gltf_node_transform_and_missing_normals(tmp_path):
    node = {"translation": [1.0, 2.0, 3.0], "rotation": [0.0, 0.0, 0.7071067811865476,
                                                          0.7071067811865476],
            "scale": [2.0, 2.0, 2.0]}
    m = load_gltf(_quad_gltf(tmp_path, node=node, normals=False))
    # vertex (1, 0, 0): scale -> (2, 0, 0), 90 deg about z -> (0, 2, 0), + t -> (1, 4, 3)
    assert any(np.allclose(p, _zup((1, 4, 3))) for p in m.positions)
    assert np.allclose(m.normals, np.array(_zup((0, 0, 1))))  # face normal of the quad


def _write_egg(tmp: Path, body: str) -> Path:
    p = tmp / "m.egg"
    p.write_text(body)
    return p


def test_egg_pool_polygons_groups(tmp_path):
    egg = """<CoordinateSystem> { Z-Up }
<Comment> { "a quad and a pentagon" }
<VertexPool> vp {
  <Vertex> 0 { 0 0 0 <Normal> { 0 0 1 } }
  <Vertex> 1 { 1 0 0 <Normal> { 0 0 1 } }
  <Vertex> 2 { 1 1 0 <Normal> { 0 0 1 } }
  <Vertex> 3 { 0 1 0 <Normal> { 0 0 1 } }
  <Vertex> 4 { 0 0 0 }
  <Vertex> 5 { 2 0 0 }
  <Vertex> 6 { 3 1 0 }
  <Vertex> 7 { 1 2 0 }
  <Vertex> 8 { -1 1 0 }
}
<Group> g {
  <Transform> { <Translate> { 0 0 5 } }
  <Polygon> { <VertexRef> { 0 1 2 3 <Ref> { vp } } }
  <Polygon> { <Normal> { 0 0 -1 } <VertexRef> { 4 5 6 7 8 <Ref> { vp } } }
}
"""
    m = load_egg(_write_egg(tmp_path, egg))
    assert m.num_triangles == 2 + 3
    assert np.allclose(m.positions[:, 2], 5.0)
    # quad corners keep their vertex normals, the pentagon takes the polygon normal
    assert np.allclose(m.normals[m.triangles[:2]].reshape(-1, 3), [0, 0, 1])
    assert np.allclose(m.normals[m.triangles[2:]].reshape(-1, 3), [0, 0, -1])


def test_egg_yup_and_matrix(tmp_path):
    egg = """<CoordinateSystem> { Y-Up }
<VertexPool> vp { <Vertex> 0 { 0 0 0 } <Vertex> 1 { 1 0 0 } <Vertex> 2 { 0 1 0 } }
<Group> g {
  <Transform> { <Matrix4> { 1 0 0 0  0 1 0 0  0 0 1 0  4 5 6 1 } }
  <Polygon> { <VertexRef> { 0 1 2 <Ref> { vp } } }
}
"""
    m = load_egg(_write_egg(tmp_path, egg))
    # egg row-vector matrix: translation (4, 5, 6); then Y-up -> Z-up
    assert any(np.allclose(p, _zup((4, 5, 6))) for p in m.positions)
    assert np.allclose(m.normals, np.array(_zup((0, 0, 1))))  # face normal


def test_load_model_resolves_files_and_builtins(tmp_path, monkeypatch):
    _quad_gltf(tmp_path)
    monkeypatch.setattr(meshmod, "MODEL_PATH", [str(tmp_path)])
    assert meshmod.load_model("quad.gltf").num_triangles == 2
    assert meshmod.load_model("models/cone").num_triangles == 64
    with pytest.raises(FileNotFoundError):
        meshmod.load_model("models/missing.gltf")


@pytest.mark.skipif(not REF_MODELS.exists(), reason="reference model files not present")
def test_reference_cylinder_gltf():
    m = load_gltf(REF_MODELS / "cylinder" / "scene.gltf")
    # 226 vertices and 960 indices (320 triangles), double-sided material
    assert m.num_vertices == 226 and m.num_triangles == 640
    ext = np.sort(m.positions.max(0) - m.positions.min(0))
    np.testing.assert_allclose(ext, [100.0, 100.0, 200.0], rtol=1e-6)
    assert np.allclose(np.linalg.norm(m.normals, axis=1), 1.0)


@pytest.mark.skipif(not REF_MODELS.exists(), reason="reference model files not present")
def test_reference_cone_egg():
    m = load_egg(REF_MODELS / "cone.egg")
    text = (REF_MODELS / "cone.egg").read_text()
    assert m.num_triangles >= text.count("<Polygon>")
    ext = m.positions.max(0) - m.positions.min(0)
    np.testing.assert_allclose(np.sort(ext), [2.0, 2.0, 2.0], atol=1e-5)
    # outward: face normals agree with the corner normals
    P, T = m.positions, m.triangles
    fn = np.cross(P[T[:, 1]] - P[T[:, 0]], P[T[:, 2]] - P[T[:, 0]])
    vn = m.normals[T].sum(1)
    assert ((fn * vn).sum(1) > 0).mean() > 0.95


def write_egg(m, path: Path) -> Path:
    """A mesh as an egg file (vertex pool with normals, one polygon per triangle)."""
    lines = ["<CoordinateSystem> { Z-Up }", "<VertexPool> vp {"]
    for i, (p, n) in enumerate(zip(m.positions.tolist(), m.normals.tolist())):
        lines.append(f"  <Vertex> {i} {{ {p[0]!r} {p[1]!r} {p[2]!r} "
                     f"<Normal> {{ {n[0]!r} {n[1]!r} {n[2]!r} }} }}")
    lines.append("}")
    for t in m.triangles:
        lines.append(f"<Polygon> {{ <VertexRef> {{ {t[0]} {t[1]} {t[2]} <Ref> {{ vp }} }} }}")
    path.write_text("\n".join(lines) + "\n")
    return path


def test_egg_round_trip_of_a_builtin(tmp_path):
    src = meshmod.uv_sphere(8, 4)
    m = load_egg(write_egg(src, tmp_path / "s.egg"))
    assert m.num_triangles == src.num_triangles
    np.testing.assert_allclose(m.positions[m.triangles], src.positions[src.triangles])
    np.testing.assert_allclose(m.normals[m.triangles], src.normals[src.triangles])


# --- model_hpr: Panda3D's setHpr convention (reference node.py:262-268) -------------
# Heading turns about +z, pitch about +x, roll about +y (right-handed, degrees),
# applied roll first, then pitch, then heading.  Unit vectors under quarter
# turns land on the axes below; compositions show the order.
@pytest.mark.parametrize("hpr,v,want", [
    ((90, 0, 0), (0, 1, 0), (-1, 0, 0)),   # heading 90: forward (+y) turns to -x
    ((90, 0, 0), (1, 0, 0), (0, 1, 0)),
    ((0, 90, 0), (0, 1, 0), (0, 0, 1)),    # pitch 90: forward points up
    ((0, 90, 0), (0, 0, 1), (0, -1, 0)),
    ((0, 0, 90), (1, 0, 0), (0, 0, -1)),   # roll 90 about forward: right side down
    ((0, 0, 90), (0, 0, 1), (1, 0, 0)),
    ((90, 90, 0), (0, 1, 0), (0, 0, 1)),   # pitch, then heading
    ((90, 90, 0), (1, 0, 0), (0, 1, 0)),
    ((0, 90, 90), (1, 0, 0), (0, 1, 0)),   # roll, then pitch
    ((90, 0, 90), (0, 0, 1), (0, 1, 0)),   # roll, then heading
    ((180, 0, 0), (1, 2, 3), (-1, -2, 3)),
    ((0, 0, 0), (1, 2, 3), (1, 2, 3)),
])
def test_model_hpr_quarter_turns(hpr, v, want):
    from pybatchrender.renderer.mesh import Mesh, bake

    P = np.array([v, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0)], dtype=np.float64)
    N = np.array([v, v, v], dtype=np.float64) / np.linalg.norm(v)
    m = bake(Mesh(P, N, np.array([[0, 1, 2]]), "probe"), model_hpr=hpr)
    np.testing.assert_allclose(m.positions[0], want, atol=1e-12)
    # normals turn with the mesh (a rotation is its own inverse transpose)
    np.testing.assert_allclose(m.normals[0], np.array(want) / np.linalg.norm(want), atol=1e-12)


def test_model_hpr_then_scale_order():
    """Scale is applied before the rotation (node.py:224-268): a box scaled
    along x and then turned by heading 90 is long along y."""
    from pybatchrender.renderer.mesh import bake, box

    m = bake(box(), model_scale=(4.0, 1.0, 1.0), model_hpr=(90, 0, 0))
    ext = m.positions.max(0) - m.positions.min(0)
    np.testing.assert_allclose(ext, [1.0, 4.0, 1.0], atol=1e-12)
