"""Meshes: procedural built-ins and model baking.

The reference loads Panda3D models and bakes the node's model transform into
the vertices at ``add_node`` time (pybatchrender/renderer/node.py:61-71:
loadModel, _apply_model_scale :214-260, _apply_model_hpr :262-268,
flattenStrong, pivot_to_rel :313-341).  Panda3D's built-in ``models/box`` is
not available here, so the built-ins are generated:

    models/box       unit cube [0,1]^3, 24 vertices (4 per face, face
                     normals), 12 counter-clockwise-outward triangles
    models/quad      unit square [0,1]^2 at z = 0, normal +z, 2 triangles
    models/sphere    UV sphere of radius 0.5, 16 segments x 8 rings,
                     smooth normals, 224 triangles
    models/cylinder  radius 0.5, height 1 along z, 32 segments: smooth
                     sides (64 triangles) + flat caps (64), 128 triangles
    models/cone      radius 0.5, height 1 along z, 32 segments, 64 triangles

Paths ending in .gltf / .glb / .egg are read from disk by ``model_io``.

Baking runs in float64 on the host and is stored as float32 (one upload at
``setup_environment``).
"""
from __future__ import annotations

import math
from collections.abc import Sequence
from dataclasses import dataclass

import numpy as np


@dataclass
class Mesh:
    positions: np.ndarray  # [V, 3]
    normals: np.ndarray    # [V, 3]
    triangles: np.ndarray  # [T, 3] int32, counter-clockwise seen from outside
    name: str = "mesh"

    def __post_init__(self) -> None:
        self.positions = np.asarray(self.positions, dtype=np.float64).reshape(-1, 3)
        self.normals = np.asarray(self.normals, dtype=np.float64).reshape(-1, 3)
        self.triangles = np.asarray(self.triangles, dtype=np.int32).reshape(-1, 3)
        if self.positions.shape != self.normals.shape:
            raise ValueError("positions and normals must have the same shape")
        if self.triangles.size and (
            self.triangles.min() < 0 or self.triangles.max() >= len(self.positions)
        ):
            raise ValueError(f"mesh {self.name}: triangle index out of range")

    @property
    def num_vertices(self) -> int:
        return int(self.positions.shape[0])

    @property
    def num_triangles(self) -> int:
        return int(self.triangles.shape[0])

    def is_flat(self) -> bool:
        """True if every triangle's three vertex normals are equal in float32."""
        if not self.num_triangles:
            return True
        n = self.normals.astype(np.float32)[self.triangles]  # [T, 3, 3]
        return bool(np.all(n[:, 0] == n[:, 1]) and np.all(n[:, 0] == n[:, 2]))

    def vertex_block(self) -> np.ndarray:
        """[V, 6] float32 rows px py pz nx ny nz, the device vertex layout."""
        return np.concatenate([self.positions, self.normals], axis=1).astype(np.float32)


def box() -> Mesh:
    pos, nrm, tris = [], [], []
    faces = [
        ((1, 0, 0), (0, 1, 0), (0, 0, 1)),
        ((-1, 0, 0), (0, 0, 1), (0, 1, 0)),
        ((0, 1, 0), (0, 0, 1), (1, 0, 0)),
        ((0, -1, 0), (1, 0, 0), (0, 0, 1)),
        ((0, 0, 1), (1, 0, 0), (0, 1, 0)),
        ((0, 0, -1), (0, 1, 0), (1, 0, 0)),
    ]
    for n, u, v in faces:
        n, u, v = (np.array(a, dtype=np.float64) for a in (n, u, v))
        c = 0.5 + 0.5 * n
        base = len(pos)
        for su, sv in ((-1, -1), (1, -1), (1, 1), (-1, 1)):
            pos.append(c + 0.5 * su * u + 0.5 * sv * v)
            nrm.append(n)
        tris += [(base, base + 1, base + 2), (base, base + 2, base + 3)]
    return Mesh(np.array(pos), np.array(nrm), np.array(tris), "box")


def quad() -> Mesh:
    pos = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0)]
    nrm = [(0, 0, 1)] * 4
    return Mesh(np.array(pos, float), np.array(nrm, float), np.array([(0, 1, 2), (0, 2, 3)]), "quad")


def uv_sphere(segments: int = 16, rings: int = 8, radius: float = 0.5) -> Mesh:
    pos = [(0.0, 0.0, radius)]
    for k in range(1, rings):
        th = math.pi * k / rings
        for s in range(segments):
            ph = 2 * math.pi * s / segments
            pos.append((radius * math.sin(th) * math.cos(ph), radius * math.sin(th) * math.sin(ph),
                        radius * math.cos(th)))
    pos.append((0.0, 0.0, -radius))
    pos = np.array(pos)
    nrm = pos / np.linalg.norm(pos, axis=1, keepdims=True)
    tris = []
    south = len(pos) - 1

    def ring(k: int, s: int) -> int:  # k in [1, rings-1]
        return 1 + (k - 1) * segments + (s % segments)

    for s in range(segments):
        tris.append((0, ring(1, s), ring(1, s + 1)))
    for k in range(1, rings - 1):
        for s in range(segments):
            a, b = ring(k, s), ring(k, s + 1)
            c, d = ring(k + 1, s), ring(k + 1, s + 1)
            tris += [(a, c, d), (a, d, b)]
    for s in range(segments):
        tris.append((south, ring(rings - 1, s + 1), ring(rings - 1, s)))
    return Mesh(pos, nrm, np.array(tris), "sphere")


def cylinder(segments: int = 32, radius: float = 0.5, height: float = 1.0) -> Mesh:
    pos, nrm, tris = [], [], []
    h0, h1 = -0.5 * height, 0.5 * height
    for s in range(segments):  # side rings: bottom s, top s
        ph = 2 * math.pi * s / segments
        c, sn = math.cos(ph), math.sin(ph)
        pos += [(radius * c, radius * sn, h0), (radius * c, radius * sn, h1)]
        nrm += [(c, sn, 0.0), (c, sn, 0.0)]
    for s in range(segments):
        b0, t0 = 2 * s, 2 * s + 1
        b1, t1 = 2 * ((s + 1) % segments), 2 * ((s + 1) % segments) + 1
        tris += [(b0, b1, t1), (b0, t1, t0)]
    for z, nz in ((h1, 1.0), (h0, -1.0)):  # caps: centre + ring
        centre = len(pos)
        pos.append((0.0, 0.0, z))
        nrm.append((0.0, 0.0, nz))
        first = len(pos)
        for s in range(segments):
            ph = 2 * math.pi * s / segments
            pos.append((radius * math.cos(ph), radius * math.sin(ph), z))
            nrm.append((0.0, 0.0, nz))
        for s in range(segments):
            a, b = first + s, first + (s + 1) % segments
            tris.append((centre, a, b) if nz > 0 else (centre, b, a))
    return Mesh(np.array(pos), np.array(nrm), np.array(tris), "cylinder")


def cone(segments: int = 32, radius: float = 0.5, height: float = 1.0) -> Mesh:
    """Cone along +z: base disc at z = -height/2, apex at +height/2; side
    normals of the slanted surface per ring vertex, flat base."""
    pos, nrm, tris = [], [], []
    h0, h1 = -0.5 * height, 0.5 * height
    slope = radius / height
    k = 1.0 / math.sqrt(1.0 + slope * slope)
    for s in range(segments):  # one apex per segment so each side face has its own normal
        ph0 = 2 * math.pi * s / segments
        ph1 = 2 * math.pi * (s + 1) / segments
        phm = 0.5 * (ph0 + ph1)
        base = len(pos)
        pos += [(radius * math.cos(ph0), radius * math.sin(ph0), h0),
                (radius * math.cos(ph1), radius * math.sin(ph1), h0), (0.0, 0.0, h1)]
        nrm += [(k * math.cos(ph0), k * math.sin(ph0), k * slope),
                (k * math.cos(ph1), k * math.sin(ph1), k * slope),
                (k * math.cos(phm), k * math.sin(phm), k * slope)]
        tris.append((base, base + 1, base + 2))
    centre = len(pos)
    pos.append((0.0, 0.0, h0))
    nrm.append((0.0, 0.0, -1.0))
    first = len(pos)
    for s in range(segments):
        ph = 2 * math.pi * s / segments
        pos.append((radius * math.cos(ph), radius * math.sin(ph), h0))
        nrm.append((0.0, 0.0, -1.0))
    for s in range(segments):
        tris.append((centre, first + (s + 1) % segments, first + s))
    return Mesh(np.array(pos), np.array(nrm), np.array(tris), "cone")


_BUILTINS = {
    "models/cone": cone,
    "cone": cone,
    "models/box": box,
    "box": box,
    "models/quad": quad,
    "quad": quad,
    "models/sphere": uv_sphere,
    "sphere": uv_sphere,
    "models/cylinder": cylinder,
    "cylinder": cylinder,
}


MODEL_PATH: list[str] = []  # extra directories searched for model files


def load_model(path: str) -> Mesh:
    """Mesh for a model path (node.py:61-62 loadModel).

    An existing .gltf / .glb / .egg file (as given, or under a directory of
    ``MODEL_PATH`` or the ``PBR_MODEL_PATH`` environment variable) is read by
    ``model_io``; otherwise built-in names resolve to procedural meshes."""
    import os
    from pathlib import Path

    from . import model_io

    key = str(path)
    if Path(key).suffix.lower() in (".gltf", ".glb", ".egg"):
        dirs = [""] + MODEL_PATH + [d for d in os.environ.get("PBR_MODEL_PATH", "").split(
            os.pathsep) if d]
        for d in dirs:
            cand = Path(d) / key if d else Path(key)
            if cand.is_file():
                return model_io.load_file(cand)
    for suffix in (".egg", ".bam", ".egg.pz"):
        if key.endswith(suffix) and key[: -len(suffix)] in _BUILTINS:
            key = key[: -len(suffix)]
    if key in _BUILTINS:
        return _BUILTINS[key]()
    raise FileNotFoundError(
        f"model '{path}' is neither a readable .gltf/.glb/.egg file nor a built-in mesh "
        f"({sorted(set(_BUILTINS))}); pass a Mesh instance or add its directory to "
        "MODEL_PATH / PBR_MODEL_PATH"
    )


def _coerce_scale(scale) -> tuple[tuple[float, float, float], bool] | None:
    """node.py:180-201: number or length-1/3 sequence -> (xyz, uniform)."""
    if scale is None:
        return None
    if isinstance(scale, (int, float)):
        v = float(scale)
        return (v, v, v), True
    if isinstance(scale, Sequence) or isinstance(scale, np.ndarray):
        seq = tuple(float(v) for v in scale)
        if len(seq) == 0:
            raise ValueError("model_scale sequence cannot be empty")
        if len(seq) == 1:
            return (seq[0],) * 3, True
        if len(seq) != 3:
            raise ValueError("model_scale sequence must have length 1 or 3")
        return seq, False
    raise TypeError("model_scale must be a number, a sequence of length 1 or 3, or None")


def panda_hpr_matrix(hpr_deg: Sequence[float]) -> np.ndarray:
    """Rotation Panda3D's NodePath.setHpr applies (Z-up, right-handed):
    roll about y, then pitch about x, then heading about z, in degrees."""
    h, p, r = (math.radians(float(a)) for a in hpr_deg)
    ch, sh, cp, sp, cr, sr = math.cos(h), math.sin(h), math.cos(p), math.sin(p), math.cos(r), math.sin(r)
    rz = np.array([[ch, -sh, 0], [sh, ch, 0], [0, 0, 1]])
    rx = np.array([[1, 0, 0], [0, cp, -sp], [0, sp, cp]])
    ry = np.array([[cr, 0, sr], [0, 1, 0], [-sr, 0, cr]])
    return rz @ rx @ ry


def bake(mesh: Mesh, model_scale=None, model_hpr=None, model_scale_units: str = "relative",
         pivot_relative_point=None) -> Mesh:
    """Model transform baked into vertices, as add_node does (node.py:61-67).

    Scale first (relative, or absolute against the longest tight-bounds edge),
    then the model HPR, then the pivot: the bounds-relative point
    ``pivot_relative_point`` becomes the origin (node.py:313-341)."""
    P = mesh.positions.copy()
    N = mesh.normals.copy()
    s = np.ones(3)
    coerced = _coerce_scale(model_scale)
    if coerced is not None:
        values, uniform = coerced
        if model_scale_units == "relative":
            s = np.array(values, dtype=np.float64)
        elif model_scale_units == "absolute":
            dims = P.max(0) - P.min(0) if len(P) else np.zeros(3)
            longest = max(float(dims.max()) if len(P) else 0.0, 1e-8)
            if uniform:
                s = np.full(3, values[0] / longest)
            else:
                comp = []
                for target, src in zip(values, dims):
                    denom = src if abs(src) >= 1e-8 else longest
                    if abs(denom) < 1e-8:
                        denom = 1.0
                    comp.append(target / denom)
                s = np.array(comp)
        else:
            raise ValueError(f"Unknown model_scale_units '{model_scale_units}'")
    R = np.eye(3)
    if model_hpr is not None:
        if not isinstance(model_hpr, (Sequence, np.ndarray)):
            raise TypeError("model_hpr must be a sequence of length 3 or None")
        hpr = tuple(float(v) for v in model_hpr)
        if len(hpr) != 3:
            raise ValueError("model_hpr sequence must have length 3")
        R = panda_hpr_matrix(hpr)
    A = R @ np.diag(s)
    P = P @ A.T
    if np.linalg.det(A) != 0:
        N = N @ np.linalg.inv(A)  # inverse-transpose for normals
        ln = np.linalg.norm(N, axis=1, keepdims=True)
        N = np.where(ln > 0, N / np.where(ln > 0, ln, 1), N)
    if pivot_relative_point is not None and len(P):
        a, b = P.min(0), P.max(0)
        rel = np.array([float(v) for v in pivot_relative_point])
        P = P - (a + (b - a) * rel)
    return Mesh(P, N, mesh.triangles.copy(), mesh.name)
