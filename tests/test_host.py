"""Host-side logic on the CPU: config rules, scene math, node/camera state,
descriptor packing, the registry and the env loop without pixels.

Closed forms restated from the reference: rotation R = Rz(h) Ry(p) Rx(r)
(shader_context.py:48-84), GL perspective (camera.py:150-166), look-at rows
(camera.py:201-240), atlas tiling (config.py:61-126), instance rows
(node.py:45-46, basic.vert:25-28).
"""
from __future__ import annotations

import math

import numpy as np
import pytest
import torch

import pybatchrender as pbr
from pybatchrender import PBRConfig, PBRRenderer
from pybatchrender.renderer import mesh as meshes
from pybatchrender.renderer.shader_context import PBRShaderContext


# --- config -----------------------------------------------------------------------
@pytest.mark.parametrize("n,tiles", [(1, (1, 1)), (4, (2, 2)), (5, (3, 2)), (4096, (64, 64)),
                                     (4098, (65, 64)), (16384, (128, 128))])
def test_tiles_from_num_scenes(n, tiles):
    cfg = PBRConfig(num_scenes=n, tile_resolution=(64, 64), device="cpu")
    assert cfg.tiles == tiles
    assert cfg.window_resolution == (64 * tiles[0], 64 * tiles[1])


def test_config_errors():
    with pytest.raises(ValueError):
        PBRConfig(tiles=(2, 2), tile_resolution=(64, 64), window_resolution=(100, 128), device="cpu")
    with pytest.raises(ValueError):
        PBRConfig(num_scenes=10, tiles=(2, 2), device="cpu")
    with pytest.raises(ValueError):
        PBRConfig(num_scenes=1, device="tpu")
    with pytest.raises(ValueError):
        PBRConfig(num_scenes=1, device="cpu", output_dtype="int16")
    cfg = PBRConfig.from_config({"num_scenes": 3, "device": "cpu"}, tile_resolution=(32, 16))
    assert cfg.window_resolution == (64, 32)


# --- math closed forms ---------------------------------------------------------------
def test_rotation_convention():
    h, p, r = 0.3, -0.7, 1.1
    R = PBRShaderContext._rotation_mats_from_hpr(torch.tensor([[h, p, r]])).numpy()[0]
    cz, sz, cy, sy, cx, sx = math.cos(h), math.sin(h), math.cos(p), math.sin(p), math.cos(r), math.sin(r)
    Rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    Ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    Rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    assert np.allclose(R, Rz @ Ry @ Rx, atol=1e-6)
    # pitch alone rotates +z towards +x (the pole tilt of CartPole)
    Rp = PBRShaderContext._rotation_mats_from_hpr(torch.tensor([[0.0, 0.5, 0.0]])).numpy()[0]
    assert np.allclose(Rp @ [0, 0, 1], [math.sin(0.5), 0, math.cos(0.5)], atol=1e-6)


def test_projection_and_view():
    r = PBRRenderer(PBRConfig(num_scenes=6, tile_resolution=(64, 32), device="cpu"))
    cam = r.add_camera(fov_y_deg=60.0, z_near=0.5, z_far=20.0)
    P = cam.P_k44.numpy()
    aspect = r.cfg.window_resolution[0] / r.cfg.window_resolution[1]  # whole atlas (camera.py:154)
    f = 1 / math.tan(math.radians(30))
    assert np.isclose(P[0, 0], f / aspect) and np.isclose(P[1, 1], f)
    assert np.isclose(P[2, 2], 20.5 / -19.5) and np.isclose(P[2, 3], 20.0 / -19.5) and P[3, 2] == -1
    eye = torch.tensor([[1.0, -5.0, 2.0]]).repeat(6, 1)
    eye[:, 0] += torch.arange(6)
    cam.set_positions_and_lookat(eye, torch.zeros(3))
    V = cam.V_k44.numpy()
    for k in range(6):
        R3 = V[k, :3, :3]
        assert np.allclose(R3 @ R3.T, np.eye(3), atol=1e-5)
        e = np.append(eye[k].numpy(), 1.0)
        assert np.allclose(V[k] @ e, [0, 0, 0, 1], atol=1e-5)  # eye -> origin
        tgt = V[k] @ np.array([0, 0, 0, 1.0])
        assert tgt[2] < 0 and abs(tgt[0]) < 1e-5 and abs(tgt[1]) < 1e-5  # target on -z axis
    assert np.allclose(cam.VP_k44.numpy(), np.einsum("ij,bjk->bik", P, V), atol=1e-5)


def test_set_projection_takes_effect_at_next_view_update():
    r = PBRRenderer(PBRConfig(num_scenes=1, device="cpu"))
    cam = r.add_camera()
    vp0 = cam.get_vp()
    cam.set_projection(fov_y_deg=30.0)
    assert torch.equal(cam.get_vp(), vp0)
    cam.set_eye(cam.get_eye())
    assert not torch.equal(cam.get_vp(), vp0)


# --- meshes ---------------------------------------------------------------------------
@pytest.mark.parametrize("name,nv,nt", [("models/box", 24, 12), ("models/quad", 4, 2),
                                        ("models/sphere", 114, 224), ("models/cylinder", 130, 128)])
def test_builtin_meshes_are_outward_ccw(name, nv, nt):
    m = meshes.load_model(name)
    assert (m.num_vertices, m.num_triangles) == (nv, nt)
    P, N, T = m.positions, m.normals, m.triangles
    face = np.cross(P[T[:, 1]] - P[T[:, 0]], P[T[:, 2]] - P[T[:, 0]])
    assert (np.einsum("ij,ij->i", face, N[T].mean(1)) > 0).all()


def test_box_is_flat_sphere_is_not():
    assert meshes.box().is_flat()
    assert not meshes.uv_sphere().is_flat()


def test_bake_scale_and_pivot():
    m = meshes.bake(meshes.box(), model_scale=(1.2, 0.8, 0.5), pivot_relative_point=(0.5, 0.5, 0.05))
    assert np.allclose(m.positions.min(0), [-0.6, -0.4, -0.025])
    assert np.allclose(m.positions.max(0), [0.6, 0.4, 0.475])
    a = meshes.bake(meshes.box(), model_scale=2.0, model_scale_units="absolute")
    assert np.allclose(a.positions.max(0) - a.positions.min(0), [2, 2, 2])
    b = meshes.bake(meshes.box(), model_scale=(4.0, 1.0, 2.0), model_scale_units="absolute")
    assert np.allclose(b.positions.max(0) - b.positions.min(0), [4, 1, 2])
    with pytest.raises(ValueError):
        meshes.bake(meshes.box(), model_scale=(1.0, 2.0))


# --- nodes and the descriptor ----------------------------------------------------------
def test_node_setters_match_reference_formula():
    r = PBRRenderer(PBRConfig(num_scenes=3, device="cpu"))
    node = r.add_node("models/box", instances_per_scene=2)
    assert node.buf_instances == 6
    pos = torch.arange(18, dtype=torch.float32).reshape(6, 3)
    hpr = torch.rand(6, 3)
    node.set_positions(pos, lazy=True)
    node.set_hprs(hpr, lazy=True)
    node.set_scales(torch.full((6,), 2.0))
    M = node.transforms_b44
    R = PBRShaderContext._rotation_mats_from_hpr(hpr)
    assert torch.allclose(M[:, :3, :3], R * 2.0)
    assert torch.equal(M[:, :3, 3], pos)
    assert torch.equal(M[:, 3], torch.tensor([0.0, 0, 0, 1]).expand(6, 4))
    mats = M.clone()
    node.set_transforms(mats)
    assert torch.allclose(node.transforms_b44, mats, atol=1e-6)


def test_shared_node_rows_and_counts_in_descriptor():
    r = PBRRenderer(PBRConfig(num_scenes=4, device="cpu"))
    shared = r.add_node("models/quad", instances_per_scene=3, shared_across_scenes=True)
    per = r.add_node("models/sphere", instances_per_scene=5)
    per.set_instance_counts(torch.tensor([0, 5, 9, 2]))
    r.setup_environment()
    d, keep = r.scene_descriptor()
    assert d.num_nodes == 2 and d.num_scenes == 4
    assert (d.nodes[0].shared, d.nodes[0].instances_per_scene, d.nodes[0].flat) == (1, 3, 1)
    assert (d.nodes[1].shared, d.nodes[1].instances_per_scene, d.nodes[1].flat) == (0, 5, 0)
    assert d.meshes[1].num_triangles == 224 and d.meshes[1].index_offset == 2
    assert per.instance_counts.tolist() == [0, 5, 5, 2]
    assert d.nodes[1].counts == per.instance_counts.data_ptr()
    assert d.strength == 0.0  # no light -> unlit
    assert tuple(round(v, 4) for v in d.background) == (0.41, 0.41, 0.41, 1.0)


def test_descriptor_cache_follows_every_input():
    """scene_descriptor reuses its last build only while every buffer it
    points to is the same tensor and the light / background are unchanged;
    each call returns its own copy."""
    r = PBRRenderer(PBRConfig(num_scenes=3, device="cpu"))
    box = r.add_node("models/box", instances_per_scene=2)
    r.add_light(strength=0.5)
    r.setup_environment()
    d1, k1 = r.scene_descriptor()
    d1.strength = 7.0  # the caller's copy only
    k1.append(torch.zeros(1))
    d2, k2 = r.scene_descriptor()
    assert d2.strength == 0.5 and len(k2) == len(k1) - 1
    assert bytes(d2) == bytes(r._build_scene_descriptor()[0])
    # light, background, camera (new VP tensor), node buffers, counts, geometry
    r._pbr_light.set_strength(0.25)
    assert r.scene_descriptor()[0].strength == 0.25
    r.set_background_color(0.1, 0.2, 0.3)
    assert round(r.scene_descriptor()[0].background[2], 5) == 0.3
    r._pbr_cam.set_positions(torch.tensor([1.0, -9.0, 2.0]))
    assert r.scene_descriptor()[0].view_proj == r._pbr_cam.VP_k44.data_ptr()
    box.set_colors(torch.rand(6, 4))
    assert r.scene_descriptor()[0].nodes[0].colors == box.colors_b4.data_ptr()
    box.set_instance_counts(torch.tensor([1, 2, 0]))
    d, _ = r.scene_descriptor()
    assert d.nodes[0].counts == box.instance_counts.data_ptr()
    assert r._desc_cache[3] == r.slot_budget() == 3 * 12
    r.add_node("models/quad", instances_per_scene=1, shared_across_scenes=True)
    r.setup_environment()
    assert r.scene_descriptor()[0].num_nodes == 2
    # in-place updates keep the descriptor (it holds the pointers)
    before = r._desc_cache
    box.colors_b4.mul_(0.5)
    r.scene_descriptor()
    assert r._desc_cache is before


def test_too_few_rows_rejected():
    r = PBRRenderer(PBRConfig(num_scenes=4, device="cpu"))
    node = r.add_node("models/box", instances_per_scene=1)
    node.transforms_b44 = node.transforms_b44[:2].clone()
    with pytest.raises(ValueError):
        r.scene_descriptor()


def test_render_on_cpu_device_fails_loudly():
    r = PBRRenderer(PBRConfig(num_scenes=2, device="cpu"))
    r.add_node("models/box", instances_per_scene=1)
    r.setup_environment()
    with pytest.raises(RuntimeError, match="GPU"):
        r.step()


def test_atlas_rearrange_roundtrip():
    r = PBRRenderer(PBRConfig(num_scenes=5, tile_resolution=(4, 3), device="cpu"))
    img = torch.arange(5 * 3 * 3 * 4, dtype=torch.int32).reshape(5, 3, 3, 4)
    cols, rows = r.cfg.tiles
    full = torch.zeros((cols * rows, 3, 3, 4), dtype=img.dtype)
    full[:5] = img
    atlas = full.view(rows, cols, 3, 3, 4).permute(0, 3, 1, 4, 2).reshape(rows * 3, cols * 4, 3)
    assert torch.equal(r._rearrange_img(atlas), img)


# --- registry and env --------------------------------------------------------------------
def test_registry_surface():
    assert "CartPole-v0" in pbr.envs.list_envs()
    E, R, C = pbr.envs.get_env_classes("CartPole-v0")
    assert R.__name__ == "CartPoleRenderer" and C.__name__ == "CartPoleConfig"
    with pytest.raises(ValueError):
        pbr.envs.make("Nope-v0")


def test_cartpole_env_without_pixels_on_cpu():
    env = pbr.envs.make("CartPole-v0", num_scenes=32, device="cpu", render=False)
    td = env.reset()
    assert td["observation"].shape == (32, 4)
    for _ in range(5):
        td["action"] = env.action_spec.rand()
        td = env.step(td)["next"]
    assert td["reward"].shape == (32, 1) and "pixels" not in td


def test_env_render_on_cpu_raises():
    """No CPU fallback: the reference's env drops `pixels` when rendering
    fails (cartpole/env.py:171-176,191-196); here the renderer's error
    reaches the caller, for reset and for step."""
    env = pbr.envs.make("CartPole-v0", num_scenes=8, device="cpu")
    with pytest.raises(RuntimeError, match="needs a GPU"):
        env.reset()
    env = pbr.envs.make("Steering-v0", num_scenes=4, device="cpu")
    with pytest.raises(RuntimeError, match="needs a GPU"):
        env.reset()
    env = pbr.envs.make("CartPole-v0", num_scenes=8, device="cpu", render=False)
    td = env.reset()
    env.render = True
    td["action"] = env.action_spec.rand()
    with pytest.raises(RuntimeError, match="needs a GPU"):
        env.step(td)


def test_cartpole_dynamics_closed_form():
    env = pbr.envs.make("CartPole-v0", num_scenes=2, device="cpu", render=False)
    obs = torch.tensor([[0.1, -0.2, 0.05, 0.3], [0.0, 0.0, 0.0, 0.0]])
    nxt = env._dynamics(obs, torch.tensor([1, 0]))
    x, xd, th, thd = obs[0].tolist()
    g, mc, mp, l, F, tau = 9.8, 1.0, 0.1, 0.5, 10.0, 0.02
    temp = (F + mp * l * thd ** 2 * math.sin(th)) / (mc + mp)
    tacc = (g * math.sin(th) - math.cos(th) * temp) / (l * (4 / 3 - mp * math.cos(th) ** 2 / (mc + mp)))
    xacc = temp - mp * l * tacc * math.cos(th) / (mc + mp)
    want = [x + tau * xd, xd + tau * xacc, th + tau * thd, thd + tau * tacc]
    assert np.allclose(nxt[0].numpy(), want, atol=1e-6)


def test_autotune_candidates_and_signature():
    from pybatchrender import autotune
    from pybatchrender.envs.cartpole import CartPoleConfig, CartPoleGroundRenderer

    r = CartPoleGroundRenderer(CartPoleConfig(num_scenes=8, device="cpu"))
    cands = autotune.candidates(r)
    assert any(c["kernel"] == "scene" for c in cands)  # 26 slots: the scene kernel qualifies
    tiles = [c for c in cands if c["kernel"] == "tiles"]
    assert {c.get("tile_size") for c in tiles} == {0, 8, 16, 32, 64}
    assert {c.get("bin_capacity") for c in tiles if not c["persistent"]} == {0, 1024, 4096}
    sig = autotune.signature(r)
    assert '"N": 8' in sig and '"W": 64' in sig
    autotune._apply(r, {"kernel": "tiles", "tile_size": 8, "persistent": False,
                        "bin_capacity": 4096})
    o = r.render_opts()
    assert o.kernel == 3 and o.tile_size == 8 and o.persistent == 0 and o.bin_capacity == 4096
    # 8-pixel tiles only while every tile can test every slot cheaply
    from pybatchrender import workloads
    big = workloads.uneven_batch(num_scenes=4, resolution=32, max_instances=64, device="cpu")
    assert 8 not in {c.get("tile_size") for c in autotune.candidates(big)}


def test_camera_setters_closed_forms():
    """The basis setters (camera.py:268-344): right x forward gives the up
    hint; a yaw of 90 deg about +z turns the default forward +y to -x."""
    r = PBRRenderer(PBRConfig(num_scenes=2, device="cpu"))
    cam = r.add_camera()
    assert torch.allclose(cam.get_forward(), torch.tensor([0.0, 1.0, 0.0]).expand(2, 3))
    assert torch.allclose(cam.get_up(), torch.tensor([0.0, 0.0, 1.0]).expand(2, 3))
    assert torch.allclose(cam.get_right(), torch.tensor([1.0, 0.0, 0.0]).expand(2, 3))
    cam.set_right(torch.tensor([0.0, 0.0, -1.0]))  # roll the camera: up becomes +x
    assert torch.allclose(cam.get_right(), torch.tensor([0.0, 0.0, -1.0]).expand(2, 3), atol=1e-6)
    assert torch.allclose(cam.get_up(), torch.tensor([1.0, 0.0, 0.0]).expand(2, 3), atol=1e-6)
    cam.set_hprs(torch.tensor([[math.pi / 2, 0.0, 0.0], [0.0, 0.0, 0.0]]))
    assert torch.allclose(cam.get_forward()[0], torch.tensor([-1.0, 0.0, 0.0]), atol=1e-6)
    assert torch.allclose(cam.get_forward()[1], torch.tensor([0.0, 1.0, 0.0]), atol=1e-6)
    cam.set_eye(torch.tensor([1.0, 2.0, 3.0]))
    cam.look_at(torch.tensor([1.0, 2.0, 10.0]) * 0 + torch.tensor([4.0, 6.0, 3.0]))
    assert torch.allclose(cam.get_forward(), torch.tensor([0.6, 0.8, 0.0]).expand(2, 3), atol=1e-6)
    V = cam.V_k44[0]
    assert torch.allclose(V @ torch.tensor([1.0, 2.0, 3.0, 1.0]), torch.tensor([0.0, 0, 0, 1]),
                          atol=1e-6)
    with pytest.raises(NotImplementedError):
        cam.set_positions(torch.zeros(3), keep_lookat=True)
    with pytest.raises(ValueError):
        cam.set_eye(torch.zeros(3, 3))  # three rows for two scenes


@pytest.mark.parametrize("kw,tiles,tile,window,n", [
    (dict(num_scenes=10), (4, 3), (64, 64), (256, 192), 10),
    (dict(num_scenes=4098, tile_resolution=(64, 64)), (65, 64), (64, 64), (4160, 4096), 4098),
    (dict(tiles=(2, 3), window_resolution=(100, 90)), (2, 3), (50, 30), (100, 90), 6),
    (dict(tile_resolution=(40, 20), window_resolution=(120, 60)), (3, 3), (40, 20), (120, 60), 9),
    (dict(tile_resolution=(40, 20)), (1, 1), (40, 20), (40, 20), 1),
    (dict(window_resolution=(33, 17)), (1, 1), (33, 17), (33, 17), 1),
    (dict(), (1, 1), (64, 64), (64, 64), 1),
    (dict(tiles=5), (3, 2), (64, 64), (192, 128), 6),
])
def test_atlas_layout_rules(kw, tiles, tile, window, n):
    """Atlas layout (reference config.py:61-126): grid ceil(sqrt(n)) columns,
    window = tiles x tile, defaults 64 x 64 tiles and a 1 x 1 grid."""
    c = PBRConfig(device="cpu", **kw)
    assert tuple(c.tiles) == tiles and tuple(c.tile_resolution) == tile
    assert tuple(c.window_resolution) == window and c.num_scenes == n
    assert c.batch_inner_dim == tiles[0] * tiles[1]


def test_atlas_layout_errors():
    with pytest.raises(ValueError):
        PBRConfig(device="cpu", tiles=(2, 2), tile_resolution=(8, 8), window_resolution=(16, 17))
    with pytest.raises(ValueError):
        PBRConfig(device="cpu", num_scenes=5, tiles=(2, 2))
    with pytest.raises(ValueError):
        PBRConfig(device="tpu")
    c = PBRConfig.from_config({"num_scenes": 3}, device="cpu")
    assert c.num_scenes == 3 and c.device == "cpu"
    from pybatchrender.envs.cartpole import CartPoleConfig
    cc = PBRConfig.from_config(CartPoleConfig(num_scenes=2, device="cpu"), max_steps=7)
    assert type(cc) is CartPoleConfig and cc.max_steps == 7 and cc.num_scenes == 2
