"""Synthetic batch workloads beyond CartPole, built through the public API.

many_cubes     N scenes of `instances` boxes: positions U(-3,3)^3, rotations
               from normalised N(0,1)^4 quaternions, uniform scales U(0.1,0.5),
               colours U(0,1)^3; one camera pose (eye (0,-8,3) -> origin, fov
               55) yawed per scene by U(0, 2 pi); one directional light.
uneven_batch   N scenes whose instance count is round(exp(U(0, ln max))) in
               [1, max]; each instance draws box / UV sphere (16x8) /
               cylinder (32) uniformly, so three nodes carry per-scene counts
               (PBRNode.set_instance_counts).  Transforms/colours/camera as
               many_cubes.
differentiable N scenes of `instances` instances, half boxes (flat) and half
               UV spheres (16x8, smooth) at positions U(-1.8, 1.8)^3, scales
               U(0.5, 1), camera eye (0, -5, 1.8) yawed per scene (about half
               of each frame covered), float32 output, plus the perturbed
               parameters a target image is rendered from: colours
               + N(0, 0.05), light direction rotated by an N(0, 5 deg) angle
               about a uniformly random axis, model-space vertex positions
               + N(0, 0.01).
All random numbers come from torch.Generator('cpu').manual_seed(seed).
"""
from __future__ import annotations

import math

import torch

from .config import PBRConfig
from .renderer.renderer import PBRRenderer

LIGHT = dict(ambient=(0.2, 0.2, 0.2), dir_dir=(-0.3, 0.5, -0.8), dir_col=(0.8, 0.8, 0.8),
             strength=1.0)


def random_rotations(n: int, gen: torch.Generator) -> torch.Tensor:
    q = torch.randn(n, 4, generator=gen, dtype=torch.float64)
    q = q / q.norm(dim=1, keepdim=True)
    w, x, y, z = q.unbind(1)
    R = torch.stack([
        1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w),
        2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w),
        2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y),
    ], dim=1).reshape(n, 3, 3)
    return R


def random_transforms(n: int, gen: torch.Generator, pos_range: float = 3.0,
                      scale_range=(0.1, 0.5)) -> torch.Tensor:
    R = random_rotations(n, gen)
    p = (torch.rand(n, 3, generator=gen, dtype=torch.float64) * 2 - 1) * pos_range
    s = torch.rand(n, 1, 1, generator=gen, dtype=torch.float64) * (scale_range[1] - scale_range[0]) + scale_range[0]
    M = torch.zeros(n, 4, 4, dtype=torch.float64)
    M[:, :3, :3] = R * s
    M[:, :3, 3] = p
    M[:, 3, 3] = 1.0
    return M.to(torch.float32)


def _yaw_camera(r: PBRRenderer, n: int, gen: torch.Generator, fov: float = 55.0,
                eye=(0.0, -8.0, 3.0)) -> None:
    yaw = torch.rand(n, generator=gen, dtype=torch.float64) * 2 * math.pi
    base = torch.tensor(eye, dtype=torch.float64)
    eye = torch.stack([base[0] * torch.cos(yaw) - base[1] * torch.sin(yaw),
                       base[0] * torch.sin(yaw) + base[1] * torch.cos(yaw),
                       base[2].expand(n)], dim=1).to(torch.float32)
    r.add_camera(fov_y_deg=fov)
    r._pbr_cam.set_positions_and_lookat(eye, torch.zeros(3))


def many_cubes(num_scenes: int = 1024, resolution: int = 128, instances: int = 64, seed: int = 0,
               device: str = "cuda", output_dtype: str = "float32", output_depth: bool = True,
               model: str = "models/box", **cfg) -> PBRRenderer:
    gen = torch.Generator("cpu").manual_seed(seed)
    c = PBRConfig(num_scenes=num_scenes, tile_resolution=(resolution, resolution), device=device,
                  output_dtype=output_dtype, output_depth=output_depth, **cfg)
    r = PBRRenderer(c)
    node = r.add_node(model, instances_per_scene=instances)
    total = num_scenes * instances
    node.set_transforms(random_transforms(total, gen))
    col = torch.ones(total, 4)
    col[:, :3] = torch.rand(total, 3, generator=gen)
    node.set_colors(col)
    _yaw_camera(r, num_scenes, gen)
    r.add_light(**LIGHT)
    r.setup_environment()
    return r


def uneven_batch(num_scenes: int = 8192, resolution: int = 84, max_instances: int = 256,
                 seed: int = 0, device: str = "cuda", output_dtype: str = "uint8",
                 models=("models/box", "models/sphere", "models/cylinder"), uniform: bool = False,
                 **cfg) -> PBRRenderer:
    """``uniform=True`` gives every scene the batch's mean instance count (the
    same draws otherwise): the uniform reference for per-triangle cost."""
    gen = torch.Generator("cpu").manual_seed(seed)
    c = PBRConfig(num_scenes=num_scenes, tile_resolution=(resolution, resolution), device=device,
                  output_dtype=output_dtype, **cfg)
    r = PBRRenderer(c)
    u = torch.rand(num_scenes, generator=gen, dtype=torch.float64)
    total = torch.round(torch.exp(u * math.log(max_instances))).clamp(1, max_instances).long()
    if uniform:
        total = torch.full_like(total, int(torch.round(total.double().mean()).item()))
    choice = torch.randint(0, len(models), (num_scenes, max_instances), generator=gen)
    active = torch.arange(max_instances)[None, :] < total[:, None]
    for k, m in enumerate(models):
        node = r.add_node(m, instances_per_scene=max_instances)
        rows = num_scenes * max_instances
        node.set_transforms(random_transforms(rows, gen))
        col = torch.ones(rows, 4)
        col[:, :3] = torch.rand(rows, 3, generator=gen)
        node.set_colors(col)
        node.set_instance_counts(((choice == k) & active).sum(1).to(torch.int32))
    _yaw_camera(r, num_scenes, gen)
    r.add_light(**LIGHT)
    r.setup_environment()
    return r


def _rotate(v: torch.Tensor, axis: torch.Tensor, angle: float) -> torch.Tensor:
    """Rodrigues rotation of vector v about the unit axis by angle (radians)."""
    c, s_ = math.cos(angle), math.sin(angle)
    return v * c + torch.cross(axis, v, dim=0) * s_ + axis * torch.dot(axis, v) * (1.0 - c)


def differentiable(num_scenes: int = 512, resolution: int = 128, instances: int = 32,
                   seed: int = 0, device: str = "cuda", sphere=(16, 8), **cfg):
    """Returns (renderer, target): the renderer's own colours, vertices and
    light are the parameters being fitted; ``target`` is a diff.SceneParams
    (no gradients) of the perturbed parameters the target image is rendered
    from (``diff.render(renderer, target)``)."""
    from . import diff
    from .renderer.mesh import uv_sphere

    gen = torch.Generator("cpu").manual_seed(seed)
    c = PBRConfig(num_scenes=num_scenes, tile_resolution=(resolution, resolution), device=device,
                  output_dtype="float32", **cfg)
    r = PBRRenderer(c)
    n_box = instances // 2
    n_sph = instances - n_box
    for model, ips in (("models/box", n_box), (uv_sphere(*sphere), n_sph)):
        if ips == 0:
            continue
        node = r.add_node(model, instances_per_scene=ips)
        rows = num_scenes * ips
        node.set_transforms(random_transforms(rows, gen, pos_range=1.8, scale_range=(0.5, 1.0)))
        col = torch.ones(rows, 4)
        col[:, :3] = torch.rand(rows, 3, generator=gen) * 0.8 + 0.1
        node.set_colors(col)
    _yaw_camera(r, num_scenes, gen, eye=(0.0, -5.0, 1.8))
    # strength below 1: every light entry is differentiable (clamp(strength,
    # 0, 1) has a kink at 1)
    r.add_light(**{**LIGHT, "strength": 0.9})
    r.setup_environment()

    p = diff.scene_params(r, requires_grad=False)
    cols = [t + (torch.randn(t.shape, generator=gen) * 0.05).to(t.device) for t in p.colors]
    verts = p.vertices.clone()
    off = torch.randn(verts.shape[0], 3, generator=gen) * 0.01
    verts[:, :3] += off.to(verts.device)
    light = p.light.clone()
    axis = torch.randn(3, generator=gen, dtype=torch.float64)
    axis = axis / axis.norm()
    angle = math.radians(float(torch.randn(1, generator=gen).item()) * 5.0)
    light[:3] = _rotate(light[:3].double(), axis, angle).to(torch.float32)
    return r, diff.SceneParams(cols, verts, light)
