"""Model files: glTF 2.0 (.gltf / .glb) and Panda3D egg (.egg) readers.

The reference hands model paths to Panda3D's ``loader.loadModel``
(pybatchrender/renderer/node.py:61-62) and flattens the loaded scene graph
into one vertex pool (``flattenStrong``, node.py:65); the Steering env loads
``models/cone.egg`` and ``models/cylinder/scene.gltf``
(envs/steering/config.py:56-59).  These readers produce the same flattened
triangle soup as a ``Mesh`` (float64 positions and normals, counter-clockwise
triangles), in Panda3D's Z-up frame:

glTF 2.0 (Khronos specification)
  - buffers from the external ``uri`` next to the file, a base64 ``data:``
    URI, or the BIN chunk of a .glb;
  - accessors of FLOAT / UNSIGNED_BYTE / UNSIGNED_SHORT / UNSIGNED_INT
    components with any ``byteStride``; indices optional;
  - primitive modes TRIANGLES, TRIANGLE_STRIP (alternate winding fixed) and
    TRIANGLE_FAN; other modes (points, lines) carry no triangles;
  - the node hierarchy of the default scene: ``matrix`` (column-major) or
    T * R * S, accumulated from the roots; normals by the inverse transpose;
  - Y-up to Z-up: (x, y, z) -> (x, -z, y), a +90 degree turn about x (the
    conversion Panda3D's glTF loader applies between the two coordinate
    systems);
  - ``doubleSided`` materials: each triangle is also emitted with the
    opposite winding and the same normals, which draws the back side the way
    a two-sided GL polygon does (the shader does not flip normals);
  - a missing NORMAL attribute: the face normal of each triangle.

Panda3D egg (Panda3D's egg syntax documentation)
  - ``<VertexPool>`` entries with ``<Normal>``; ``<Polygon>`` with
    ``<VertexRef> { i j k ... <Ref> { pool } }`` and an optional polygon
    ``<Normal>``; convex polygons split as fans from their first vertex;
  - ``<Group>`` nesting with ``<Transform>`` (``<Matrix4>`` in egg's
    row-vector convention, ``<Translate>``, ``<Scale>``, ``<RotX/Y/Z>``,
    ``<Rotate>`` in degrees), applied in the order written;
  - ``<CoordinateSystem> { Y-Up }`` converted to Z-up as for glTF
    (Z-up is the default);
  - vertex normal, else polygon normal, else the face normal.
"""
from __future__ import annotations

import base64
import json
import math
import struct
from pathlib import Path

import numpy as np

from .mesh import Mesh

_YUP_TO_ZUP = np.array([[1.0, 0.0, 0.0], [0.0, 0.0, -1.0], [0.0, 1.0, 0.0]])


def _face_normals(P: np.ndarray, T: np.ndarray) -> np.ndarray:
    n = np.cross(P[T[:, 1]] - P[T[:, 0]], P[T[:, 2]] - P[T[:, 0]])
    ln = np.linalg.norm(n, axis=1, keepdims=True)
    return np.where(ln > 0, n / np.where(ln > 0, ln, 1.0), 0.0)


def _flat(P: np.ndarray, T: np.ndarray) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Un-indexed copy of triangles T with face normals."""
    fn = _face_normals(P, T)
    P2 = P[T].reshape(-1, 3)
    N2 = np.repeat(fn, 3, axis=0)
    return P2, N2, np.arange(len(P2), dtype=np.int64).reshape(-1, 3)


def _normalize(N: np.ndarray) -> np.ndarray:
    ln = np.linalg.norm(N, axis=1, keepdims=True)
    return np.where(ln > 0, N / np.where(ln > 0, ln, 1.0), N)


# --- glTF ---------------------------------------------------------------------
_COMP = {5120: ("b", 1), 5121: ("B", 1), 5122: ("h", 2), 5123: ("H", 2), 5125: ("I", 4),
         5126: ("f", 4)}
_NCOMP = {"SCALAR": 1, "VEC2": 2, "VEC3": 3, "VEC4": 4, "MAT4": 16}


def _accessor(doc: dict, bufs: list[bytes], idx: int) -> np.ndarray:
    acc = doc["accessors"][idx]
    fmt, size = _COMP[acc["componentType"]]
    nc = _NCOMP[acc["type"]]
    count = int(acc["count"])
    if "bufferView" not in acc:  # all zeros by the specification
        return np.zeros((count, nc), dtype=np.float64)
    view = doc["bufferViews"][acc["bufferView"]]
    data = bufs[view["buffer"]]
    base = int(view.get("byteOffset", 0)) + int(acc.get("byteOffset", 0))
    stride = int(view.get("byteStride", 0)) or size * nc
    dt = np.dtype("<" + fmt)
    out = np.empty((count, nc), dtype=np.float64)
    if count == 0:
        return out
    if base + stride * (count - 1) + size * nc > len(data):
        raise ValueError(f"glTF accessor {idx} runs past its buffer")
    for k in range(nc):
        out[:, k] = np.ndarray((count,), dtype=dt, buffer=data, offset=base + k * size,
                               strides=(stride,))
    return out


def _node_matrix(node: dict) -> np.ndarray:
    if "matrix" in node:
        return np.array(node["matrix"], dtype=np.float64).reshape(4, 4).T  # column-major
    M = np.eye(4)
    if "scale" in node:
        M = np.diag([*map(float, node["scale"]), 1.0]) @ M
    if "rotation" in node:
        x, y, z, w = map(float, node["rotation"])
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                      [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                      [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
        Rh = np.eye(4)
        Rh[:3, :3] = R
        M = Rh @ M
    if "translation" in node:
        Th = np.eye(4)
        Th[:3, 3] = [float(v) for v in node["translation"]]
        M = Th @ M
    return M


def _load_buffers(doc: dict, base_dir: Path, glb_bin: bytes | None) -> list[bytes]:
    out = []
    for k, b in enumerate(doc.get("buffers", [])):
        uri = b.get("uri")
        if uri is None:
            if glb_bin is None or k != 0:
                raise ValueError("glTF buffer without uri outside a .glb")
            out.append(glb_bin)
        elif uri.startswith("data:"):
            out.append(base64.b64decode(uri.split(",", 1)[1]))
        else:
            out.append((base_dir / uri).read_bytes())
    return out


def load_gltf(path: str | Path, yup_to_zup: bool = True) -> Mesh:
    path = Path(path)
    glb_bin = None
    if path.suffix.lower() == ".glb":
        blob = path.read_bytes()
        magic, _, length = struct.unpack_from("<III", blob, 0)
        if magic != 0x46546C67:
            raise ValueError(f"{path}: not a binary glTF file")
        off, doc = 12, None
        while off < length:
            clen, ctype = struct.unpack_from("<II", blob, off)
            chunk = blob[off + 8: off + 8 + clen]
            if ctype == 0x4E4F534A:
                doc = json.loads(chunk.decode("utf-8"))
            elif ctype == 0x004E4942:
                glb_bin = chunk
            off += 8 + clen
        if doc is None:
            raise ValueError(f"{path}: no JSON chunk")
    else:
        doc = json.loads(path.read_text())
    bufs = _load_buffers(doc, path.parent, glb_bin)
    materials = doc.get("materials", [])
    Ps, Ns, Ts = [], [], []
    nv = 0

    def emit(P, N, T):
        nonlocal nv
        Ps.append(P)
        Ns.append(N)
        Ts.append(T + nv)
        nv += len(P)

    def visit(ni: int, parent: np.ndarray):
        node = doc["nodes"][ni]
        M = parent @ _node_matrix(node)
        if "mesh" in node:
            A = M[:3, :3]
            nA = np.linalg.inv(A).T if abs(np.linalg.det(A)) > 0 else A
            for prim in doc["meshes"][node["mesh"]]["primitives"]:
                mode = prim.get("mode", 4)
                if mode not in (4, 5, 6):
                    continue
                P = _accessor(doc, bufs, prim["attributes"]["POSITION"])[:, :3]
                if "indices" in prim:
                    idx = _accessor(doc, bufs, prim["indices"])[:, 0].astype(np.int64)
                else:
                    idx = np.arange(len(P), dtype=np.int64)
                if mode == 4:
                    T = idx[: len(idx) // 3 * 3].reshape(-1, 3)
                elif mode == 5:  # strip: every other triangle reversed
                    T = np.array([(idx[i], idx[i + 1], idx[i + 2]) if i % 2 == 0 else
                                  (idx[i + 1], idx[i], idx[i + 2]) for i in range(len(idx) - 2)],
                                 dtype=np.int64).reshape(-1, 3)
                else:  # fan
                    T = np.array([(idx[0], idx[i], idx[i + 1]) for i in range(1, len(idx) - 1)],
                                 dtype=np.int64).reshape(-1, 3)
                Pw = P @ A.T + M[:3, 3]
                if "NORMAL" in prim["attributes"]:
                    Nw = _normalize(_accessor(doc, bufs, prim["attributes"]["NORMAL"])[:, :3] @ nA.T)
                    Pp, Np, Tp = Pw, Nw, T
                else:
                    Pp, Np, Tp = _flat(Pw, T)
                emit(Pp, Np, Tp)
                mat = materials[prim["material"]] if "material" in prim else {}
                if mat.get("doubleSided", False):  # same vertices, opposite winding
                    Ts.append(Tp[:, ::-1] + (nv - len(Pp)))
        for c in node.get("children", []):
            visit(c, M)

    scene = doc.get("scenes", [{}])[doc.get("scene", 0)] if doc.get("scenes") else None
    roots = scene.get("nodes", []) if scene else list(range(len(doc.get("nodes", []))))
    for r in roots:
        visit(r, np.eye(4))
    if not Ps:
        return Mesh(np.zeros((0, 3)), np.zeros((0, 3)), np.zeros((0, 3)), path.name)
    P = np.concatenate(Ps)
    N = np.concatenate(Ns)
    T = np.concatenate(Ts)
    if yup_to_zup:
        P = P @ _YUP_TO_ZUP.T
        N = N @ _YUP_TO_ZUP.T
    return Mesh(P, N, T, path.name)


# --- egg ----------------------------------------------------------------------
def _egg_tokens(text: str):
    i, n = 0, len(text)
    while i < n:
        c = text[i]
        if c.isspace():
            i += 1
        elif c == "/" and text.startswith("//", i):
            j = text.find("\n", i)
            i = n if j < 0 else j + 1
        elif c == "/" and text.startswith("/*", i):
            j = text.find("*/", i + 2)
            i = n if j < 0 else j + 2
        elif c in "{}":
            yield c
            i += 1
        elif c == '"':
            j = i + 1
            while j < n and text[j] != '"':
                j += 2 if text[j] == "\\" else 1
            yield ("str", text[i + 1:j])
            i = j + 1
        elif c == "<":
            j = text.find(">", i)
            yield ("tag", text[i + 1:j])
            i = j + 1
        else:
            j = i
            while j < n and not text[j].isspace() and text[j] not in '{}"<':
                j += 1
            yield ("word", text[i:j])
            i = j


def _egg_parse(tokens) -> list:
    """Nested entries: (tag, name, children) where children mixes words and
    entries."""
    out = []
    stack = [out]
    pending = None
    for t in tokens:
        if isinstance(t, tuple) and t[0] == "tag":
            pending = [t[1], None]
        elif t == "{":
            tag, name = pending if pending else ("", None)
            entry = (tag, name, [])
            stack[-1].append(entry)
            stack.append(entry[2])
            pending = None
        elif t == "}":
            stack.pop()
        elif pending is not None:
            pending[1] = t[1]
        else:
            stack[-1].append(t[1])
    return out


def _words(children) -> list[str]:
    return [c for c in children if isinstance(c, str)]


def _sub(children, tag: str):
    return [c for c in children if isinstance(c, tuple) and c[0].lower() == tag.lower()]


def _egg_transform(entry) -> np.ndarray:
    """<Transform> { ... } -> 4x4 column-vector matrix, components in order."""
    M = np.eye(4)
    for c in entry[2]:
        if not isinstance(c, tuple):
            continue
        tag = c[0].lower()
        v = [float(w) for w in _words(c[2])]
        X = np.eye(4)
        if tag == "matrix4":
            X = np.array(v, dtype=np.float64).reshape(4, 4).T  # egg rows are row vectors
        elif tag == "translate":
            X[:3, 3] = v[:3]
        elif tag == "scale":
            X = np.diag([v[0], v[0], v[0], 1.0]) if len(v) == 1 else np.diag([*v[:3], 1.0])
        elif tag in ("rotx", "roty", "rotz", "rotate"):
            if tag == "rotate":
                ang, axis = v[0], np.array(v[1:4])
            else:
                ang, axis = v[0], np.eye(3)["xyz".index(tag[-1])]
            a = math.radians(ang)
            k = axis / np.linalg.norm(axis)
            K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
            X[:3, :3] = np.eye(3) + math.sin(a) * K + (1 - math.cos(a)) * K @ K
        M = X @ M
    return M


def load_egg(path: str | Path) -> Mesh:
    path = Path(path)
    tree = _egg_parse(_egg_tokens(path.read_text()))
    yup = any(isinstance(e, tuple) and e[0].lower() == "coordinatesystem" and
              "y-up" in " ".join(_words(e[2])).lower() for e in tree)
    pools: dict[str, dict[int, tuple]] = {}

    def collect_pools(entries):
        for e in entries:
            if not isinstance(e, tuple):
                continue
            if e[0].lower() == "vertexpool":
                verts = {}
                for v in _sub(e[2], "Vertex"):
                    xyz = [float(w) for w in _words(v[2])[:3]]
                    nrm = _sub(v[2], "Normal")
                    nvec = [float(w) for w in _words(nrm[0][2])[:3]] if nrm else None
                    verts[int(v[1])] = (xyz, nvec)
                pools[e[1]] = verts
            collect_pools(e[2])

    collect_pools(tree)
    Ps, Ns = [], []

    def walk(entries, M):
        for e in entries:
            if not isinstance(e, tuple):
                continue
            tag = e[0].lower()
            if tag == "group":
                G = M
                for tr in _sub(e[2], "Transform"):
                    G = M @ _egg_transform(tr)
                walk(e[2], G)
            elif tag == "polygon":
                refs = _sub(e[2], "VertexRef")
                if not refs:
                    continue
                ids = [int(float(w)) for w in _words(refs[0][2])]
                pref = _sub(refs[0][2], "Ref")
                pool = pools[_words(pref[0][2])[0]] if pref else next(iter(pools.values()))
                pn = _sub(e[2], "Normal")
                poly_n = [float(w) for w in _words(pn[0][2])[:3]] if pn else None
                pts = np.array([pool[i][0] for i in ids], dtype=np.float64)
                A = M[:3, :3]
                nA = np.linalg.inv(A).T if abs(np.linalg.det(A)) > 0 else A
                pts = pts @ A.T + M[:3, 3]
                for k in range(1, len(ids) - 1):
                    tri = [0, k, k + 1]
                    P = pts[tri]
                    fn = np.cross(P[1] - P[0], P[2] - P[0])
                    ln = np.linalg.norm(fn)
                    fn = fn / ln if ln > 0 else fn
                    for q, t in enumerate(tri):
                        vn = pool[ids[t]][1]
                        if vn is not None:
                            nrm = np.array(vn) @ nA.T
                        elif poly_n is not None:
                            nrm = np.array(poly_n) @ nA.T
                        else:
                            nrm = fn
                        Ps.append(P[q])
                        Ns.append(nrm)
            else:
                walk(e[2], M)

    walk(tree, np.eye(4))
    if not Ps:
        return Mesh(np.zeros((0, 3)), np.zeros((0, 3)), np.zeros((0, 3)), path.name)
    P = np.array(Ps)
    N = _normalize(np.array(Ns))
    if yup:
        P = P @ _YUP_TO_ZUP.T
        N = N @ _YUP_TO_ZUP.T
    T = np.arange(len(P), dtype=np.int64).reshape(-1, 3)
    return _dedupe(Mesh(P, N, T, path.name))


def _dedupe(m: Mesh) -> Mesh:
    """Merge corners with identical position and normal (keeps first-use order)."""
    if not m.num_triangles:
        return m
    key = np.concatenate([m.positions, m.normals], axis=1)
    uniq, first, inv = np.unique(key, axis=0, return_index=True, return_inverse=True)
    order = np.argsort(first)
    remap = np.empty_like(order)
    remap[order] = np.arange(len(order))
    rows = uniq[order]
    return Mesh(rows[:, :3], rows[:, 3:], remap[inv.reshape(-1)].reshape(-1, 3), m.name)


def load_file(path: str | Path) -> Mesh:
    """Mesh of a .gltf / .glb / .egg file."""
    p = Path(path)
    suf = p.suffix.lower()
    if suf in (".gltf", ".glb"):
        return load_gltf(p)
    if suf == ".egg":
        return load_egg(p)
    raise ValueError(f"{p}: unsupported model format '{suf}' (.gltf, .glb, .egg)")
