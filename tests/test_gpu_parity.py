"""HIP renderer vs the CPU oracle, bit for bit (coverage, depth, pixels).

Every case renders on cuda:0 through libpbr_hip.so (PBRRenderer.render ->
pbr_render) and on the CPU through oracle/liboracle_pbr.so from the very same
descriptor tensors.  Integer outputs (visibility, uint8/bf16 pixels) and the
float depth / float pixels are compared bitwise.
"""
from __future__ import annotations

import json
import math
from pathlib import Path

import numpy as np
import pytest
import torch

from oracle import oracle
from pybatchrender_amd import PBRConfig, PBRRenderer, workloads
from pybatchrender_amd.envs.cartpole import (CartPoleConfig, CartPoleGroundRenderer,
                                             CartPoleRenderer)

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"


def _bits(t: torch.Tensor) -> torch.Tensor:
    t = t.detach().cpu().contiguous()
    if t.dtype == torch.float32:
        return t.view(torch.int32)
    if t.dtype == torch.bfloat16:
        return t.view(torch.int16)
    return t


def check_parity(r: PBRRenderer, label: str) -> dict:
    print(f"[parity] {label}: N={r.num_scenes} {r.cfg.tile_resolution} {r.cfg.output_dtype} "
          f"C={r.cfg.num_channels}", flush=True)
    got = r.render(depth=True, vis=True)
    torch.cuda.synchronize()
    desc, keep = r.scene_descriptor()
    ref = oracle.render(desc, keep, channels=int(r.cfg.num_channels), dtype=r.cfg.output_dtype)
    gv, rv = got["vis"].cpu(), ref["vis"]
    bad = (gv != rv).sum().item()
    assert bad == 0, f"{label}: {bad} visibility mismatches"
    assert torch.equal(_bits(got["depth"]), _bits(ref["depth"])), f"{label}: depth differs"
    gp, rp = _bits(got["pixels"]), _bits(ref["pixels"])
    assert torch.equal(gp, rp), f"{label}: {(gp != rp).sum().item()} pixel values differ"
    return {"got": got, "ref": ref}


def ground_states(n: int, seed: int = 0) -> torch.Tensor:
    g = torch.Generator("cpu").manual_seed(seed)
    st = torch.zeros(n, 4)
    st[:, 0] = torch.empty(n).uniform_(-2.4, 2.4, generator=g)
    st[:, 2] = torch.empty(n).uniform_(-0.21, 0.21, generator=g)
    return st


def test_cartpole_reference_scene_notebook_states(cuda):
    cfg = CartPoleConfig(num_scenes=4098, device="cuda")
    r = CartPoleRenderer(cfg)
    meta = json.loads((GOLDEN / "notebook_tiles.json").read_text())
    st = torch.zeros(4098, 4)
    g = torch.Generator("cpu").manual_seed(1)
    st[:, 0] = torch.empty(4098).uniform_(-2.0, 2.0, generator=g)
    st[:, 2] = torch.empty(4098).uniform_(-0.5, 0.5, generator=g)
    for k, v in meta["initial_states_printed"].items():
        st[int(k)] = torch.tensor(v)
    r._step(st.cuda())
    out = check_parity(r, "cartpole reference scene")
    tiles = np.load(GOLDEN / "notebook_tiles.npz")["initial"]
    for k in range(3):
        mine = out["got"]["pixels"][k].permute(1, 2, 0).cpu().numpy().astype(int)
        diff = np.abs(mine - tiles[k].astype(int)).max(-1)
        assert (diff <= 2).mean() >= 0.99


@pytest.mark.parametrize("n", [1, 256, 4096])
def test_cartpole_ground_scene(cuda, n):
    r = CartPoleGroundRenderer(CartPoleConfig(num_scenes=n, device="cuda"))
    r._step(ground_states(n).cuda())
    check_parity(r, f"cartpole ground N={n}")


@pytest.mark.parametrize("dtype,channels", [("float32", 3), ("bfloat16", 3), ("uint8", 4),
                                            ("uint8", 1), ("float32", 4), ("bfloat16", 2)])
def test_many_cubes_dtypes(cuda, dtype, channels):
    r = workloads.many_cubes(num_scenes=32, resolution=128, instances=64, seed=3,
                             output_dtype=dtype, num_channels=channels)
    check_parity(r, f"many cubes {dtype} C={channels}")


@pytest.mark.parametrize("res", [(17, 13), (1, 1), (84, 84), (200, 40), (64, 96), (8, 300)])
def test_resolutions(cuda, res):
    gen = torch.Generator("cpu").manual_seed(7)
    c = PBRConfig(num_scenes=24, tile_resolution=res, device="cuda")
    r = PBRRenderer(c)
    node = r.add_node("models/box", instances_per_scene=16)
    node.set_transforms(workloads.random_transforms(24 * 16, gen))
    node.set_colors(torch.rand(24 * 16, 4, generator=gen))
    r.add_camera()
    r._pbr_cam.set_positions_and_lookat(torch.tensor([0.0, -7.0, 2.0]), torch.zeros(3))
    r.add_light()
    r.setup_environment()
    check_parity(r, f"resolution {res}")


def test_smooth_meshes_uneven_counts(cuda):
    r = workloads.uneven_batch(num_scenes=96, resolution=84, max_instances=64, seed=5)
    check_parity(r, "uneven spheres/cylinders/boxes")


@pytest.mark.parametrize("rows", [8, 16, 24, 64])
def test_region_rows(cuda, rows):
    r = workloads.many_cubes(num_scenes=16, resolution=64, instances=24, seed=9,
                             output_dtype="uint8", output_depth=False, region_rows=rows)
    check_parity(r, f"region rows {rows}")


def test_persistent_queue(cuda):
    r = workloads.uneven_batch(num_scenes=200, resolution=64, max_instances=48, seed=11,
                               persistent=True)
    check_parity(r, "persistent work queue")
    check_parity(r, "persistent work queue (second launch)")


def test_near_plane_and_guard_band_clipping(cuda):
    """Geometry that crosses the near plane and extends far past the view."""
    n = 16
    c = PBRConfig(num_scenes=n, tile_resolution=(64, 64), device="cuda")
    r = PBRRenderer(c)
    ground = r.add_node("models/quad", instances_per_scene=1, shared_across_scenes=True,
                        model_scale=(4000.0, 4000.0, 1.0), model_pivot_relative_point=(0.5, 0.5, 0.0))
    ground.set_colors(torch.tensor([[0.3, 0.6, 0.3, 1.0]]))
    big = r.add_node("models/box", instances_per_scene=2, model_pivot_relative_point=(0.5, 0.5, 0.5))
    gen = torch.Generator("cpu").manual_seed(2)
    M = workloads.random_transforms(n * 2, gen, pos_range=1.0, scale_range=(1.5, 4.0))
    big.set_transforms(M)
    big.set_colors(torch.rand(n * 2, 4, generator=gen))
    r.add_camera(fov_y_deg=70.0, z_near=0.5, z_far=50.0)
    eye = torch.zeros(n, 3)
    eye[:, 0] = torch.linspace(-1, 1, n)
    eye[:, 2] = 1.0
    r._pbr_cam.set_positions_and_lookat(eye, torch.tensor([0.0, 10.0, 0.5]))
    r.add_light()
    r.setup_environment()
    check_parity(r, "near/far/guard clipping")


def test_empty_and_zero_instances(cuda):
    c = PBRConfig(num_scenes=5, tile_resolution=(32, 32), device="cuda")
    r = PBRRenderer(c)
    r.set_background_color(0.1, 0.2, 0.3, 1.0)
    r.setup_environment()
    out = check_parity(r, "no nodes")
    assert (out["got"]["vis"] == -1).all()
    node = r.add_node("models/box", instances_per_scene=4)
    node.set_instance_counts(torch.zeros(5, dtype=torch.int32))
    r.setup_environment()
    check_parity(r, "zero active instances")


def test_unlit_without_light(cuda):
    """No add_light -> colour only (reference: lightingStrength stays 0)."""
    c = PBRConfig(num_scenes=3, tile_resolution=(32, 32), device="cuda")
    r = PBRRenderer(c)
    node = r.add_node("models/box", instances_per_scene=1, model_pivot_relative_point=(0.5, 0.5, 0.5))
    node.set_colors(torch.tensor([[0.2, 0.4, 0.6, 1.0]]).expand(3, 4))
    r.add_camera()
    r._pbr_cam.set_positions_and_lookat(torch.tensor([2.0, -3.0, 1.5]), torch.zeros(3))
    r.setup_environment()
    out = check_parity(r, "unlit")
    px = out["got"]["pixels"].cpu()
    vis = out["got"]["vis"].cpu()
    covered = vis >= 0
    for ch, v in enumerate((0.2, 0.4, 0.6)):
        assert (px[:, ch][covered] == math.floor(v * 255 + 0.5)).all()


def test_env_step_renders(cuda):
    import pybatchrender_amd as pbr

    env = pbr.envs.make("CartPole-v0", num_scenes=64)
    td = env.reset()
    assert td["pixels"].shape == (64, 3, 64, 64) and td["pixels"].is_cuda
    td["action"] = env.action_spec.rand()
    td = env.step(td)
    assert td["next", "pixels"].dtype == torch.uint8
