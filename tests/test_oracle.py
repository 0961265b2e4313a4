"""Pinning the CPU oracle (the checker) against facts outside itself.

* the reference's own Panda3D frames (notebook tiles, tests/golden);
* an exact rational point-in-triangle test with the top-left rule;
* watertightness and full-coverage invariants of the fill rule;
* closed forms: a ground plane seen by a horizontal camera (near-plane and
  guard-band clipping, depth), and the Lambert + ambient colour of a face.
"""
from __future__ import annotations

import json
import math
from fractions import Fraction
from pathlib import Path

import numpy as np
import pytest
import torch
from hypothesis import given, settings
from hypothesis import strategies as st

from oracle import oracle
from pybatchrender import PBRConfig, PBRRenderer
from pybatchrender.envs.cartpole import CartPoleConfig, CartPoleRenderer

GOLDEN = Path(__file__).resolve().parent / "golden"


# --- the reference's own frames --------------------------------------------------
@pytest.fixture(scope="module")
def notebook():
    meta = json.loads((GOLDEN / "notebook_tiles.json").read_text())
    tiles = np.load(GOLDEN / "notebook_tiles.npz")
    return meta, tiles


@pytest.fixture(scope="module")
def reference_scene_frames(notebook):
    meta, _ = notebook
    r = CartPoleRenderer(CartPoleConfig(num_scenes=meta["num_scenes"], device="cpu"))
    st_ = torch.zeros(meta["num_scenes"], 4)
    for k, v in meta["initial_states_printed"].items():
        st_[int(k)] = torch.tensor(v)
    r._step(st_)
    desc, keep = r.scene_descriptor()
    return oracle.render(desc, keep, scenes=(0, 3))


def test_config_matches_notebook_setup(notebook):
    meta, _ = notebook
    cfg = CartPoleConfig(num_scenes=meta["num_scenes"], device="cpu")
    assert list(cfg.tiles) == meta["tiles"]
    assert list(cfg.window_resolution) == meta["window_resolution"]


@pytest.mark.parametrize("k", [0, 1, 2])
def test_notebook_tiles_coverage_and_colour(notebook, reference_scene_frames, k):
    """Reference frames (Panda3D on an L4) of scenes with printed states:
    >=99% of pixels within 2/255, mismatches only where colours sit on a
    rounding boundary; the drawn-pixel set is identical."""
    _, tiles = notebook
    ref = tiles["initial"][k].astype(int)
    mine = reference_scene_frames["pixels"][k].permute(1, 2, 0).numpy().astype(int)
    diff = np.abs(mine - ref).max(-1)
    assert (diff <= 2).mean() >= 0.99
    assert diff.max() <= 1
    bg = np.array([105, 105, 105])
    assert ((ref != bg).any(-1) == (mine != bg).any(-1)).all()


def test_background_is_reference_grey(notebook, reference_scene_frames):
    _, tiles = notebook
    corner = tiles["initial"][:, :4, :4].reshape(-1, 3)
    assert (corner == 105).all()
    assert (reference_scene_frames["pixels"][:3, :, :4, :4] == 105).all()


def test_ambient_only_face_closed_form(reference_scene_frames):
    """Cart faces facing away from the light show round(255 * c * ambient)."""
    c0 = np.array([0.6, 0.8, 1.0])  # cart colour of scene 0 (t = 0)
    expect = np.floor(255 * c0 * np.array([0.2, 0.2, 0.25]) + 0.5).astype(int)
    px = reference_scene_frames["pixels"][0].permute(1, 2, 0).reshape(-1, 3).numpy().astype(int)
    assert (px == expect).all(-1).sum() >= 30


# --- fill rule ---------------------------------------------------------------------
def exact_coverage(X, Y, W, H):
    """Rational point-in-triangle with the top-left rule (GL window coords)."""
    x = [Fraction(v, 256) for v in X]
    y = [Fraction(v, 256) for v in Y]
    area = (x[1] - x[0]) * (y[2] - y[0]) - (x[2] - x[0]) * (y[1] - y[0])
    mask = np.zeros((H, W), np.uint8)
    if area <= 0:
        return mask
    for j in range(H):
        for i in range(W):
            px, py = Fraction(2 * i + 1, 2), Fraction(2 * j + 1, 2)
            ok = True
            for k in range(3):
                a, b = (k + 1) % 3, (k + 2) % 3
                e = (x[b] - x[a]) * (py - y[a]) - (y[b] - y[a]) * (px - x[a])
                dx, dy = x[b] - x[a], y[b] - y[a]
                top_left = dy < 0 or (dy == 0 and dx < 0)
                if e < 0 or (e == 0 and not top_left):
                    ok = False
            if ok:
                mask[H - 1 - j, i] = 1
    return mask


coord = st.integers(min_value=-4 * 256, max_value=12 * 256)
snapped = st.integers(min_value=-4, max_value=12).map(lambda v: v * 256 + 128)  # pixel centres


@settings(max_examples=150, deadline=None)
@given(st.lists(st.one_of(coord, snapped), min_size=6, max_size=6))
def test_fill_rule_matches_exact_rational(v):
    X, Y = v[:3], v[3:]
    mask, n = oracle.coverage(X, Y, 8, 8)
    ref = exact_coverage(X, Y, 8, 8)
    assert (mask == ref).all()
    assert n == int(ref.sum())


@settings(max_examples=100, deadline=None)
@given(st.lists(st.integers(min_value=-2 * 256, max_value=18 * 256), min_size=8, max_size=8))
def test_shared_edge_is_watertight(v):
    """Two triangles sharing an edge (a convex or concave quad split on a
    diagonal): no pixel is covered twice, and for a convex quad their union
    equals the coverage of the quad computed by the rational test."""
    X = v[0::2]
    Y = v[1::2]
    t1 = ([X[0], X[1], X[2]], [Y[0], Y[1], Y[2]])
    t2 = ([X[0], X[2], X[3]], [Y[0], Y[2], Y[3]])
    m1, _ = oracle.coverage(*t1, 16, 16)
    m2, _ = oracle.coverage(*t2, 16, 16)
    assert not (m1 & m2).any()


def test_full_tile_quad_covers_every_pixel():
    W = H = 16
    X = [0, W * 256, W * 256, 0]
    Y = [0, 0, H * 256, H * 256]
    m1, _ = oracle.coverage([X[0], X[1], X[2]], [Y[0], Y[1], Y[2]], W, H)
    m2, _ = oracle.coverage([X[0], X[2], X[3]], [Y[0], Y[2], Y[3]], W, H)
    assert ((m1 + m2) == 1).all()


def test_clockwise_is_culled():
    m, n = oracle.coverage([0, 0, 2048], [0, 2048, 0], 8, 8)
    assert n == 0 and not m.any()


# --- closed forms through the full render path ---------------------------------------
def _single_node_renderer(model, res=(64, 64), **node_kw):
    r = PBRRenderer(PBRConfig(num_scenes=1, tile_resolution=res, device="cpu"))
    node = r.add_node(model, instances_per_scene=1, **node_kw)
    return r, node


def test_ground_plane_clipping_and_depth_closed_form():
    """Horizontal camera over an 2000 x 2000 ground: the quad crosses the near
    plane and the guard band.  Rows below the horizon are covered iff the
    pixel ray meets the ground before the far plane; depth matches the
    window depth of that hit point."""
    W = H = 64
    near, far = 0.1, 100.0
    r, g = _single_node_renderer("models/quad", (W, H), shared_across_scenes=True,
                                 model_scale=(2000.0, 2000.0, 1.0),
                                 model_pivot_relative_point=(0.5, 0.5, 0.0))
    r.add_camera(fov_y_deg=90.0, z_near=near, z_far=far)
    r._pbr_cam.set_positions_and_lookat(torch.tensor([0.0, 0.0, 1.0]), torch.tensor([0.0, 1.0, 1.0]))
    r.setup_environment()
    desc, keep = r.scene_descriptor()
    out = oracle.render(desc, keep)
    vis, depth = out["vis"][0].numpy(), out["depth"][0].numpy()
    for yo in range(H):
        y_ndc = 1.0 - (2 * yo + 1) / H
        for i in range(W):
            x_ndc = (2 * i + 1) / W - 1.0
            if y_ndc < 0 and 1.0 / -y_ndc <= far * 0.999 and abs(x_ndc / -y_ndc) < 999:
                t = 1.0 / -y_ndc  # eye depth of the ground hit
                z_ndc = (far + near) / (far - near) - 2 * far * near / ((far - near) * t)
                assert vis[yo, i] >= 0, (yo, i)
                assert abs(depth[yo, i] - (0.5 * z_ndc + 0.5)) < 2e-5
            elif y_ndc > 0:
                assert vis[yo, i] == -1


def test_lambert_ambient_closed_form():
    """A unit box face seen head-on: colour = c * (amb + max(n.L, 0) * col)."""
    amb, ldir, lcol = (0.1, 0.15, 0.2), (0.3, -0.8, 0.5), (0.9, 0.7, 0.6)
    col = (0.7, 0.45, 0.3, 1.0)
    r, node = _single_node_renderer("models/box", (32, 32), model_pivot_relative_point=(0.5, 0.5, 0.5))
    node.set_colors(torch.tensor([col]))
    r.add_camera(fov_y_deg=40.0)
    r._pbr_cam.set_positions_and_lookat(torch.tensor([0.0, -4.0, 0.0]), torch.zeros(3))
    r.add_light(ambient=amb, dir_dir=ldir, dir_col=lcol, strength=1.0)
    r.setup_environment()
    desc, keep = r.scene_descriptor()
    out = oracle.render(desc, keep, dtype="float32")
    px = out["pixels"][0, :, 16, 16].numpy()
    L = np.array(ldir) / np.linalg.norm(ldir)
    ndl = max(float(np.dot([0.0, -1.0, 0.0], L)), 0.0)
    expect = np.array(col[:3]) * (np.array(amb) + ndl * np.array(lcol))
    assert np.allclose(px, expect, rtol=1e-6, atol=1e-7)
    assert out["vis"][0, 16, 16] >= 0


def test_strength_mixes_towards_unlit():
    r, node = _single_node_renderer("models/box", (16, 16), model_pivot_relative_point=(0.5, 0.5, 0.5))
    node.set_colors(torch.tensor([[0.5, 0.5, 0.5, 1.0]]))
    r.add_camera(fov_y_deg=40.0)
    r._pbr_cam.set_positions_and_lookat(torch.tensor([0.0, -4.0, 0.0]), torch.zeros(3))
    r.add_light(ambient=(0.2, 0.2, 0.2), dir_dir=(0, -1, 0), dir_col=(1, 1, 1), strength=0.25)
    r.setup_environment()
    desc, keep = r.scene_descriptor()
    px = oracle.render(desc, keep, dtype="float32")["pixels"][0, 0, 8, 8].item()
    light = 0.2 + 1.0 * 1.0
    assert math.isclose(px, 0.5 * (0.75 + 0.25 * light), rel_tol=1e-6)


def test_bf16_and_unorm8_conversions():
    r, node = _single_node_renderer("models/box", (16, 16), model_pivot_relative_point=(0.5, 0.5, 0.5))
    node.set_colors(torch.tensor([[0.3, 0.6, 0.9, 0.5]]))
    r.add_camera(fov_y_deg=40.0)
    r._pbr_cam.set_positions_and_lookat(torch.tensor([0.0, -4.0, 0.0]), torch.zeros(3))
    r.setup_environment()
    desc, keep = r.scene_descriptor()
    f = oracle.render(desc, keep, channels=4, dtype="float32")["pixels"]
    b = oracle.render(desc, keep, channels=4, dtype="bfloat16")["pixels"]
    u = oracle.render(desc, keep, channels=4, dtype="uint8")["pixels"]
    assert torch.equal(b, f.to(torch.bfloat16))
    assert torch.equal(u, torch.floor(f.clamp(0, 1) * 255 + 0.5).to(torch.uint8))
    assert f[0, 3, 8, 8].item() == 0.5  # alpha = instance colour alpha
