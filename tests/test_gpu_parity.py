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
from pybatchrender import PBRConfig, PBRRenderer, workloads
from pybatchrender.envs.cartpole import (CartPoleConfig, CartPoleGroundRenderer,
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


def check_parity(r: PBRRenderer, label: str, depth: bool = True, vis: bool = True,
                 got: dict | None = None) -> dict:
    """Render (or take `got`, a finished render of the current state) and
    compare with the oracle bitwise.  depth/vis False request no such output,
    which selects the kernel variants that do not carry them."""
    print(f"[parity] {label}: N={r.num_scenes} {r.cfg.tile_resolution} {r.cfg.output_dtype} "
          f"C={r.cfg.num_channels} depth={depth} vis={vis}", flush=True)
    if got is None:
        got = r.render(depth=depth, vis=vis)
    torch.cuda.synchronize()
    desc, keep = r.scene_descriptor()
    ref = oracle.render(desc, keep, channels=int(r.cfg.num_channels), dtype=r.cfg.output_dtype)
    if vis:
        gv, rv = got["vis"].cpu(), ref["vis"]
        bad = (gv != rv).sum().item()
        assert bad == 0, f"{label}: {bad} visibility mismatches"
    else:
        assert "vis" not in got
    if depth:
        assert torch.equal(_bits(got["depth"]), _bits(ref["depth"])), f"{label}: depth differs"
    else:
        assert "depth" not in got
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
                             output_dtype="uint8", output_depth=False, region_rows=rows,
                             kernel="chunked")
    check_parity(r, f"region rows {rows}")


@pytest.mark.parametrize("kernel", ["tiles", "chunked"])
def test_persistent_queue(cuda, kernel):
    r = workloads.uneven_batch(num_scenes=200, resolution=64, max_instances=48, seed=11,
                               persistent=True, kernel=kernel)
    check_parity(r, f"persistent work queue ({kernel})")
    check_parity(r, f"persistent work queue ({kernel}, second launch)")


@pytest.mark.parametrize("ts", [8, 16, 32, 64])
@pytest.mark.parametrize("res", [(84, 84), (128, 72), (50, 30)])
def test_tile_sizes(cuda, ts, res):
    gen = torch.Generator("cpu").manual_seed(13)
    c = PBRConfig(num_scenes=20, tile_resolution=res, device="cuda", kernel="tiles",
                  tile_size=ts, output_depth=True)
    r = PBRRenderer(c)
    box = r.add_node("models/box", instances_per_scene=20)
    box.set_transforms(workloads.random_transforms(20 * 20, gen))
    box.set_colors(torch.rand(20 * 20, 4, generator=gen))
    sph = r.add_node("models/sphere", instances_per_scene=6)
    sph.set_transforms(workloads.random_transforms(20 * 6, gen, scale_range=(0.3, 2.5)))
    sph.set_colors(torch.rand(20 * 6, 4, generator=gen))
    sph.set_instance_counts(torch.randint(0, 7, (20,), generator=gen))
    r.add_camera()
    r._pbr_cam.set_positions_and_lookat(torch.tensor([0.0, -7.0, 2.0]), torch.zeros(3))
    r.add_light()
    r.setup_environment()
    check_parity(r, f"tiles {ts} at {res}")


@pytest.mark.parametrize("cap,ts", [(1024, 0), (1024, 8), (4096, 32), (16384, 16)])
def test_tile_bin_capacities(cuda, cap, ts):
    """LDS bins of 1024 entries (emptied after every round of box tests) up
    to 16384: up to 256 instances per scene fill the bins of a tile many
    times over."""
    r = workloads.uneven_batch(num_scenes=24, resolution=64, max_instances=256, seed=4,
                               kernel="tiles", tile_size=ts, bin_capacity=cap)
    check_parity(r, f"bin capacity {cap}, tile {ts}")


def test_tiles_uneven_batch_full_size_slice(cuda):
    """The uneven-batch workload's scene mix at its real resolution."""
    r = workloads.uneven_batch(num_scenes=64, resolution=84, max_instances=256, seed=0,
                               kernel="tiles", persistent=True)
    check_parity(r, "uneven 84x84 up to 256 instances")


@pytest.mark.parametrize("persistent", [False, True])
def test_tiles_many_tiny_and_empty_scenes(cuda, persistent):
    """Per-scene counts of 0, 1 or 2 boxes over 700 scenes, with runs of
    empty scenes: one 256-slot setup step crosses many scenes (and every
    scene boundary of a run of zeros), so the tile setup's walk from the
    previous slot's scene and its binary-search fallback both run."""
    n = 700
    c = PBRConfig(num_scenes=n, tile_resolution=(24, 24), device="cuda", kernel="tiles",
                  persistent=persistent)
    r = PBRRenderer(c)
    # 8 boxes (96 triangle slots) per scene at most; counts leave most of
    # them unused
    box = r.add_node("models/box", instances_per_scene=8, model_pivot_relative_point=(0.5, 0.5, 0.5))
    gen = torch.Generator("cpu").manual_seed(11)
    counts = torch.randint(0, 3, (n,), generator=gen, dtype=torch.int32)
    counts[100:180] = 0                     # a run of empty scenes
    counts[400:420] = 8                     # a few full ones
    box.set_instance_counts(counts)
    box.set_transforms(workloads.random_transforms(n * 8, gen, pos_range=1.0, scale_range=(0.5, 1.2)))
    box.set_colors(torch.rand(n * 8, 4, generator=gen))
    r.add_camera(fov_y_deg=60.0, z_near=0.1, z_far=50.0)
    r._pbr_cam.set_positions_and_lookat(torch.tensor([[0.0, -5.0, 1.0]]).repeat(n, 1),
                                        torch.tensor([0.0, 0.0, 0.0]))
    r.add_light()
    r.setup_environment()
    check_parity(r, f"700 tiny and empty scenes (persistent={persistent})")


@pytest.mark.parametrize("kernel", ["tiles", "chunked"])
def test_smooth_geometry_through_the_near_plane(cuda, kernel):
    """Spheres and boxes straddling the near plane: clipped smooth sub-triangles."""
    n = 12
    c = PBRConfig(num_scenes=n, tile_resolution=(48, 48), device="cuda", kernel=kernel,
                  output_dtype="float32", output_depth=True)
    r = PBRRenderer(c)
    sph = r.add_node("models/sphere", instances_per_scene=3)
    gen = torch.Generator("cpu").manual_seed(4)
    M = workloads.random_transforms(n * 3, gen, pos_range=1.2, scale_range=(2.0, 4.0))
    sph.set_transforms(M)
    sph.set_colors(torch.rand(n * 3, 4, generator=gen))
    r.add_camera(fov_y_deg=80.0, z_near=0.8, z_far=40.0)
    eye = torch.zeros(n, 3)
    eye[:, 1] = -1.5 + 0.1 * torch.arange(n)
    r._pbr_cam.set_positions_and_lookat(eye, torch.tensor([0.0, 5.0, 0.0]))
    r.add_light()
    r.setup_environment()
    check_parity(r, f"smooth near-plane clipping ({kernel})")


def _clipping_scene(kernel: str, n: int = 16):
    """A 4000-unit ground quad past the guard band and boxes around the camera
    that cross the near plane; returns the renderer and the box node."""
    c = PBRConfig(num_scenes=n, tile_resolution=(64, 64), device="cuda", kernel=kernel)
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
    return r, big


@pytest.mark.parametrize("kernel", ["tiles", "chunked", "rows"])
def test_near_plane_and_guard_band_clipping(cuda, kernel):
    """Geometry that crosses the near plane and extends far past the view."""
    r, _ = _clipping_scene(kernel)
    check_parity(r, f"near/far/guard clipping ({kernel})")


def test_graph_replays_with_clipped_geometry(cuda):
    """Replays of a captured row-span render whose triangles straddle the near
    plane and the guard band (their row spans come from the out-of-line
    clipper, which reads the launch's descriptor from the kernel arguments)
    match eager renders and the oracle, with new box matrices written in
    place between replays."""
    r, big = _clipping_scene("rows")
    g = r.capture_step(depth=True, vis=True)
    gen = torch.Generator("cpu").manual_seed(5)
    for k in range(3):
        M = workloads.random_transforms(r.num_scenes * 2, gen, pos_range=1.0, scale_range=(1.5, 4.0))
        big.transforms_b44.copy_(M.reshape(-1, 4, 4).to(big.transforms_b44.device))
        got = {key: v.clone() for key, v in g().items()}
        torch.cuda.synchronize()
        out = check_parity(r, f"graph replay {k} with clipped geometry")
        for key in ("pixels", "depth", "vis"):
            assert torch.equal(_bits(got[key]), _bits(out["got"][key])), f"replay {k}: {key}"


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
    import pybatchrender as pbr

    env = pbr.envs.make("CartPole-v0", num_scenes=64)
    td = env.reset()
    assert td["pixels"].shape == (64, 3, 64, 64) and td["pixels"].is_cuda
    td["action"] = env.action_spec.rand()
    td = env.step(td)
    assert td["next", "pixels"].dtype == torch.uint8


# --- the scene kernel (<= 64 triangle slots per scene) against the oracle ---------
def _small_scene(n, res, seed, model="models/box", instances=5, **cfg):
    gen = torch.Generator("cpu").manual_seed(seed)
    r = PBRRenderer(PBRConfig(num_scenes=n, tile_resolution=res, device="cuda", **cfg))
    node = r.add_node(model, instances_per_scene=instances)
    node.set_transforms(workloads.random_transforms(n * instances, gen, pos_range=2.0,
                                                    scale_range=(0.3, 1.2)))
    node.set_colors(torch.rand(n * instances, 4, generator=gen))
    r.add_camera()
    eye = torch.stack([torch.rand(n, generator=gen) * 4 - 2, torch.full((n,), -6.0),
                       torch.rand(n, generator=gen) * 3], 1)
    r._pbr_cam.set_positions_and_lookat(eye, torch.zeros(3))
    r.add_light()
    r.setup_environment()
    return r


@pytest.mark.parametrize("bw", [16, 32, 64])
@pytest.mark.parametrize("per", [0, 1, 3])
def test_scene_kernel_bins(cuda, bw, per):
    r = _small_scene(40, (64, 48), 21, kernel="scene", bin_width=bw, bins_per_item=per)
    check_parity(r, f"scene kernel bw={bw} per={per}")


@pytest.mark.parametrize("dtype,channels", [("float32", 4), ("bfloat16", 3), ("uint8", 2)])
def test_scene_kernel_smooth_dtypes(cuda, dtype, channels):
    from pybatchrender.renderer import mesh

    low = mesh.uv_sphere(segments=6, rings=4)  # 36 smooth triangles
    r = _small_scene(24, (40, 40), 22, model=low, instances=1, kernel="scene",
                     output_dtype=dtype, num_channels=channels)
    check_parity(r, f"scene kernel smooth {dtype}")


@pytest.mark.parametrize("res", [(17, 13), (1, 1), (84, 84), (130, 70)])
def test_scene_kernel_resolutions(cuda, res):
    r = _small_scene(12, res, 23, kernel="scene")
    check_parity(r, f"scene kernel {res}")


def test_scene_kernel_heavy_clipping_multi_pass(cuda):
    """Large boxes around the camera: most triangles cross the near plane,
    so a scene yields more than 64 records and is rasterised in passes."""
    n = 8
    gen = torch.Generator("cpu").manual_seed(24)
    r = PBRRenderer(PBRConfig(num_scenes=n, tile_resolution=(48, 48), device="cuda",
                              kernel="scene"))
    node = r.add_node("models/box", instances_per_scene=5, model_pivot_relative_point=(0.5, 0.5, 0.5))
    node.set_transforms(workloads.random_transforms(n * 5, gen, pos_range=1.5,
                                                    scale_range=(3.0, 8.0)))
    node.set_colors(torch.rand(n * 5, 4, generator=gen))
    r.add_camera(z_near=0.3)
    r._pbr_cam.set_positions_and_lookat(torch.zeros(3), torch.tensor([0.3, 1.0, 0.2]))
    r.add_light()
    r.setup_environment()
    check_parity(r, "scene kernel clipping")


@pytest.mark.parametrize("kernel", ["scene", "chunked", "tiles", "rows"])
def test_cartpole_both_kernels(cuda, kernel):
    r = CartPoleGroundRenderer(CartPoleConfig(num_scenes=512, device="cuda", kernel=kernel))
    r._step(ground_states(512, 5).cuda())
    check_parity(r, f"cartpole ground {kernel}")


# --- the row-span kernel (<= 64 slots, <= 128 columns) against the oracle -----------
@pytest.mark.parametrize("res", [(64, 64), (17, 13), (1, 1), (84, 84), (128, 70), (64, 200)])
def test_rows_kernel_resolutions(cuda, res):
    r = _small_scene(12, res, 31, kernel="rows")
    check_parity(r, f"rows kernel {res}")


@pytest.mark.parametrize("cap", [8, 16, 64])
def test_rows_kernel_short_lists(cuda, cap):
    """A short row list forces many bands and, at 8 entries, single rows whose
    entries are built in record chunks with the pixel state carried over."""
    r = _small_scene(16, (64, 64), 32, kernel="rows", row_list_cap=cap)
    check_parity(r, f"rows kernel list cap {cap}")


@pytest.mark.parametrize("res,cap", [((64, 64), 0), ((64, 64), 8), ((64, 64), 16),
                                     ((33, 17), 64), ((128, 70), 0), ((64, 64), 200)])
@pytest.mark.parametrize("dtype", ["uint8", "bfloat16", "float32"])
def test_rows_kernel_small_scenes(cuda, res, cap, dtype):
    """Two boxes (24 slots): flat scenes of at most 32 slots and 64 columns
    whose list cap is at most 128 run the 32-record form of the kernel (short
    lists: many bands and record chunks); 128 columns or a 200-entry cap
    take the 64-record form."""
    r = _small_scene(20, res, 36, instances=2, kernel="rows", output_dtype=dtype,
                     row_list_cap=cap)
    check_parity(r, f"rows kernel 24 slots {res} cap {cap} {dtype}")
    check_parity(r, f"rows kernel 24 slots {res} cap {cap} {dtype} colour only", depth=False,
                 vis=False)


@pytest.mark.parametrize("dtype", ["uint8", "float32"])
def test_rows_kernel_small_scenes_instance_counts(cuda, dtype):
    """The 32-record form with per-scene instance counts (0-2 boxes of one
    node and 0-1 of a shared quad: at most 26 slots, some scenes empty), then
    counts edited in place."""
    n = 40
    gen = torch.Generator("cpu").manual_seed(38)
    r = PBRRenderer(PBRConfig(num_scenes=n, tile_resolution=(64, 48), device="cuda",
                              kernel="rows", output_dtype=dtype))
    a = r.add_node("models/box", instances_per_scene=2)
    a.set_transforms(workloads.random_transforms(n * 2, gen, pos_range=2.0))
    a.set_colors(torch.rand(n * 2, 4, generator=gen))
    a.set_instance_counts(torch.randint(0, 3, (n,), generator=gen))
    b = r.add_node("models/quad", instances_per_scene=1, shared_across_scenes=True)
    b.set_transforms(workloads.random_transforms(1, gen, pos_range=1.0))
    b.set_colors(torch.rand(1, 4, generator=gen))
    b.set_instance_counts(torch.randint(0, 2, (n,), generator=gen))
    r.add_camera()
    eye = torch.stack([torch.rand(n, generator=gen) * 4 - 2, torch.full((n,), -6.0),
                       torch.rand(n, generator=gen) * 3], 1)
    r._pbr_cam.set_positions_and_lookat(eye, torch.zeros(3))
    r.add_light()
    r.setup_environment()
    check_parity(r, f"rows kernel 26-slot node counts {dtype}")
    check_parity(r, f"rows kernel counts {dtype} colour only", depth=False, vis=False)
    a.instance_counts[::3] = 2
    b.instance_counts[1::2] = 0
    check_parity(r, f"rows kernel counts edited in place {dtype}")


@pytest.mark.parametrize("cap", [0, 8])
def test_rows_kernel_small_scene_clipping(cuda, cap):
    """The 32-record form with triangles through the near plane (two large
    boxes around the camera, 24 slots)."""
    n = 8
    gen = torch.Generator("cpu").manual_seed(37)
    r = PBRRenderer(PBRConfig(num_scenes=n, tile_resolution=(48, 48), device="cuda",
                              kernel="rows", row_list_cap=cap))
    node = r.add_node("models/box", instances_per_scene=2, model_pivot_relative_point=(0.5, 0.5, 0.5))
    node.set_transforms(workloads.random_transforms(n * 2, gen, pos_range=1.5,
                                                    scale_range=(3.0, 8.0)))
    node.set_colors(torch.rand(n * 2, 4, generator=gen))
    r.add_camera(z_near=0.3)
    r._pbr_cam.set_positions_and_lookat(torch.zeros(3), torch.tensor([0.3, 1.0, 0.2]))
    r.add_light()
    r.setup_environment()
    check_parity(r, f"rows kernel 24-slot clipping cap {cap}")
    check_parity(r, f"rows kernel 24-slot clipping cap {cap} colour only", depth=False, vis=False)


@pytest.mark.parametrize("dtype,channels", [("float32", 4), ("bfloat16", 3), ("uint8", 2),
                                            ("uint8", 1)])
def test_rows_kernel_smooth_dtypes(cuda, dtype, channels):
    from pybatchrender.renderer import mesh

    low = mesh.uv_sphere(segments=6, rings=4)  # 36 smooth triangles
    r = _small_scene(24, (40, 40), 33, model=low, instances=1, kernel="rows",
                     output_dtype=dtype, num_channels=channels)
    check_parity(r, f"rows kernel smooth {dtype} C={channels}")


def test_rows_kernel_overlapping_full_frame(cuda):
    """Thirty frame-filling quads, some at exactly equal depth: every row
    holds dozens of entries (many bands) and ties go to the earlier slot."""
    n = 6
    r = PBRRenderer(PBRConfig(num_scenes=n, tile_resolution=(64, 64), device="cuda",
                              kernel="rows"))
    node = r.add_node("models/quad", instances_per_scene=30, model_scale=(8.0, 8.0, 1.0),
                      model_pivot_relative_point=(0.5, 0.5, 0.0))
    depth = torch.tensor([float(k // 2) * 0.2 for k in range(30)])  # pairs share a depth
    M = torch.eye(4).repeat(n * 30, 1, 1)
    M[:, :3, :3] = torch.tensor([[1.0, 0.0, 0.0], [0.0, 0.0, -1.0], [0.0, 1.0, 0.0]])  # face -y
    M[:, 1, 3] = depth.repeat(n)
    node.set_transforms(M)
    gen = torch.Generator("cpu").manual_seed(34)
    node.set_colors(torch.rand(n * 30, 4, generator=gen))
    r.add_camera()
    eye = torch.zeros(n, 3)
    eye[:, 1] = -3.0 - 0.5 * torch.arange(n, dtype=torch.float32)
    eye[:, 0] = 0.05 * torch.arange(n, dtype=torch.float32)
    eye[:, 2] = 0.2
    r._pbr_cam.set_positions_and_lookat(eye, torch.tensor([0.0, 10.0, 0.0]))
    r.add_light()
    r.setup_environment()
    out = check_parity(r, "rows kernel overlapping quads")
    assert (out["got"]["vis"] >= 0).float().mean() > 0.9


@pytest.mark.parametrize("cap", [0, 8])
def test_rows_kernel_heavy_clipping(cuda, cap):
    """Large boxes around the camera: most triangles cross the near plane and
    are re-derived into sub-triangles where they are drawn."""
    n = 8
    gen = torch.Generator("cpu").manual_seed(35)
    r = PBRRenderer(PBRConfig(num_scenes=n, tile_resolution=(48, 48), device="cuda",
                              kernel="rows", row_list_cap=cap))
    node = r.add_node("models/box", instances_per_scene=5, model_pivot_relative_point=(0.5, 0.5, 0.5))
    node.set_transforms(workloads.random_transforms(n * 5, gen, pos_range=1.5,
                                                    scale_range=(3.0, 8.0)))
    node.set_colors(torch.rand(n * 5, 4, generator=gen))
    r.add_camera(z_near=0.3)
    r._pbr_cam.set_positions_and_lookat(torch.zeros(3), torch.tensor([0.3, 1.0, 0.2]))
    r.add_light()
    r.setup_environment()
    check_parity(r, f"rows kernel clipping cap {cap}")


def test_rows_kernel_smooth_through_the_near_plane(cuda):
    """Low-poly spheres around the camera: clipped smooth sub-triangles are
    shaded from their re-derived vertices."""
    from pybatchrender.renderer import mesh

    n = 10
    low = mesh.uv_sphere(segments=6, rings=4)
    c = PBRConfig(num_scenes=n, tile_resolution=(48, 48), device="cuda", kernel="rows",
                  output_dtype="float32", output_depth=True)
    r = PBRRenderer(c)
    sph = r.add_node(low, instances_per_scene=1)
    gen = torch.Generator("cpu").manual_seed(36)
    M = workloads.random_transforms(n, gen, pos_range=0.5, scale_range=(2.0, 4.0))
    sph.set_transforms(M)
    sph.set_colors(torch.rand(n, 4, generator=gen))
    r.add_camera(fov_y_deg=80.0, z_near=0.8, z_far=40.0)
    eye = torch.zeros(n, 3)
    eye[:, 1] = -1.0 + 0.1 * torch.arange(n)
    r._pbr_cam.set_positions_and_lookat(eye, torch.tensor([0.0, 5.0, 0.0]))
    r.add_light()
    r.setup_environment()
    check_parity(r, "rows kernel smooth near-plane clipping")


def test_rows_kernel_per_triangle_setup(cuda):
    """Meshes with many more vertices than their triangles use (here 400
    vertices, 12 triangles) overflow the vertex table; such batches set up
    each triangle on its own lane."""
    from pybatchrender.renderer import mesh

    box = mesh.box()
    pad = 400 - box.num_vertices
    rng = np.random.default_rng(38)
    pos = np.concatenate([box.positions, rng.uniform(-1, 1, (pad, 3))])
    nrm = np.concatenate([box.normals, np.tile([[0.0, 0.0, 1.0]], (pad, 1))])
    tris = box.triangles + 0
    order = rng.permutation(len(pos))  # scatter the used vertices through the table
    inv = np.argsort(order)
    sparse = mesh.Mesh(pos[order], nrm[order], inv[tris], name="sparse_box")
    r = _small_scene(10, (64, 64), 38, model=sparse, instances=3, kernel="rows")
    check_parity(r, "rows kernel per-triangle setup")


def test_auto_beyond_rows_limits(cuda):
    """Tiles wider than 128 pixels: AUTO takes the scene kernel instead."""
    r = _small_scene(6, (130, 40), 37)
    check_parity(r, "auto at 130 columns")


# --- pose kernel and graph-captured steps -------------------------------------------
def test_pose_kernel_matches_host_formula(cuda):
    n = 300
    g = torch.Generator("cpu").manual_seed(31)
    vals = {k: torch.rand(n, generator=g) * 4 - 2 for k in "xyzhpr"}
    vals["s"] = torch.rand(n, generator=g) + 0.5
    rg = PBRRenderer(PBRConfig(num_scenes=n, device="cuda"))
    rc = PBRRenderer(PBRConfig(num_scenes=n, device="cpu"))
    ng = rg.add_node("models/box", instances_per_scene=1)
    nc = rc.add_node("models/box", instances_per_scene=1)
    st = torch.stack([vals[k] for k in "xyzhprs"], 1)  # strided columns on the GPU
    stg = st.cuda()
    ng.set_pose(**{k: stg[:, i] for i, k in enumerate("xyzhprs")})
    nc.set_pose(**{k: st[:, i] for i, k in enumerate("xyzhprs")})
    torch.cuda.synchronize()
    assert torch.allclose(ng.transforms_b44.cpu(), nc.transforms_b44, atol=2e-6)
    assert torch.allclose(ng.rot3_b33.cpu(), nc.rot3_b33, atol=2e-6)
    ng.set_pose(x=1.5, p=stg[:, 4])  # constants + one column
    nc.set_pose(x=1.5, p=st[:, 4])
    assert torch.allclose(ng.transforms_b44.cpu(), nc.transforms_b44, atol=2e-6)
    # a one-element tensor (broadcast), a longer column (first rows), float64
    # and host tensors (converted), then the output buffers replaced
    long = torch.rand(n + 17, generator=g) - 0.5
    ng.set_pose(x=torch.tensor([0.75], device="cuda"), y=long.cuda(), h=st[:, 3].double().cuda(),
                r=st[:, 5])
    nc.set_pose(x=0.75, y=long[:n], h=st[:, 3], r=st[:, 5])
    assert torch.allclose(ng.transforms_b44.cpu(), nc.transforms_b44, atol=2e-6)
    ng.transforms_b44 = torch.zeros(1, device="cuda")
    ng.set_pose(z=stg[:, 2])
    nc.set_pose(z=st[:, 2])
    assert ng.transforms_b44.shape == (n, 4, 4)
    assert torch.allclose(ng.transforms_b44.cpu(), nc.transforms_b44, atol=2e-6)


@pytest.mark.parametrize("scene", ["ground", "reference"])
def test_graphed_step_replays_match_eager(cuda, scene):
    n = 1024
    cls = CartPoleGroundRenderer if scene == "ground" else CartPoleRenderer
    r = cls(CartPoleConfig(num_scenes=n, device="cuda"))
    states = [ground_states(n, seed).cuda() for seed in range(4)]
    g = r.capture_step(states[0], depth=True, vis=True)
    for k, stt in enumerate(states + states[:1]):
        got = {key: v.clone() for key, v in g(stt).items()}
        torch.cuda.synchronize()
        r._step(stt)  # eager work between replays
        ref = r.render(depth=True, vis=True)
        torch.cuda.synchronize()
        for key in ("pixels", "depth", "vis"):
            assert torch.equal(_bits(got[key]), _bits(ref[key])), f"replay {k}: {key}"
    desc, keep = r.scene_descriptor()
    o = oracle.render(desc, keep)
    assert torch.equal(g.outputs()["vis"].cpu(), o["vis"])


def test_batched_poses_match_single(cuda):
    n = 777
    g = torch.Generator("cpu").manual_seed(32)
    st = (torch.rand(n, 4, generator=g) * 2 - 1).cuda()
    r = PBRRenderer(PBRConfig(num_scenes=n, device="cuda"))
    a = r.add_node("models/box", instances_per_scene=1)
    b = r.add_node("models/box", instances_per_scene=2)
    r.set_poses([(a, dict(x=st[:, 0], p=st[:, 2])), (b, dict(y=1.0, h=0.3, s=2.0))])
    ta, tb = a.transforms_b44.clone(), b.transforms_b44.clone()
    a.set_pose(x=st[:, 0], p=st[:, 2])
    b.set_pose(y=1.0, h=0.3, s=2.0)
    torch.cuda.synchronize()
    assert torch.equal(ta, a.transforms_b44) and torch.equal(tb, b.transforms_b44)


def test_autotune_picks_a_shape_that_matches_the_oracle(cuda, tmp_path):
    from pybatchrender import autotune
    # boxes and spheres, at most 8 each: 1888 slots per scene, so 8-pixel
    # tiles are among the candidates
    r = workloads.uneven_batch(num_scenes=48, resolution=84, max_instances=8, seed=21,
                               models=("models/box", "models/sphere"))
    cache = tmp_path / "tune.json"
    rep = autotune.autotune(r, launches=2, cache_path=cache)
    assert not rep["cached"] and len(rep["results"]) >= 16
    tried = [c["cand"] for c in rep["results"]]
    assert any(c.get("tile_size") == 8 for c in tried)
    assert {c.get("bin_capacity") for c in tried if c["kernel"] == "tiles"} >= {1024, 4096}
    print(f"[autotune] best {rep['best']} {rep['us']} us", flush=True)
    check_parity(r, f"autotuned {rep['best']}")
    rep2 = autotune.autotune(r, cache_path=cache)
    assert rep2["cached"] and rep2["best"] == rep["best"]


def test_meshes_loaded_from_files(cuda, tmp_path):
    """Nodes built from an egg and a glTF file on disk render like any mesh."""
    from pybatchrender.renderer import mesh as meshmod
    from tests.test_models import _quad_gltf, write_egg
    egg = write_egg(meshmod.cone(16), tmp_path / "cone.egg")
    gltf = _quad_gltf(tmp_path, double_sided=True)
    n = 10
    c = PBRConfig(num_scenes=n, tile_resolution=(48, 48), device="cuda")
    r = PBRRenderer(c)
    gen = torch.Generator("cpu").manual_seed(31)
    a = r.add_node(str(egg), instances_per_scene=3, model_scale=1.5)
    a.set_transforms(workloads.random_transforms(n * 3, gen, pos_range=1.5, scale_range=(0.8, 1.6)))
    a.set_colors(torch.rand(n * 3, 4, generator=gen))
    b = r.add_node(str(gltf), instances_per_scene=2, model_pivot_relative_point=(0.5, 0.5, 0.5))
    b.set_transforms(workloads.random_transforms(n * 2, gen, pos_range=1.0, scale_range=(1.0, 2.0)))
    b.set_colors(torch.rand(n * 2, 4, generator=gen))
    r.add_camera()
    r._pbr_cam.set_positions_and_lookat(torch.tensor([0.0, -6.0, 2.0]), torch.zeros(3))
    r.add_light()
    r.setup_environment()
    check_parity(r, "meshes from egg / glTF files")
