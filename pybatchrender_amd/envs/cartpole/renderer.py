"""CartPole scenes.

CartPoleRenderer is the reference scene (pybatchrender/envs/cartpole/
renderer.py:10-138): rail + cart + pole boxes, camera (5,5,2) looking at the
origin, default light.  Its ``_step`` maps states [B, 4] (x, x_dot, theta,
theta_dot) to cart/pole instance transforms on the device; the reference's
``.detach().cpu()`` host round trip (:112) is gone.

CartPoleGroundRenderer is the cart + pole + ground-quad scene the
throughput figures are measured on (64x64 tiles, uint8 RGB).  Values the
workload description leaves open are fixed here:
    cart   box scaled (0.4, 0.2, 0.2), centre (x, 0, 0.1): rests on z = 0
    pole   box scaled (0.05, 0.05, 1.0), pivot at its bottom centre, hinged
           at the cart's top centre (x, 0, 0.2), rotated Ry(theta)
    ground quad 10 x 10 centred at the origin, z = 0, shared by all scenes
    colours cart (0.9, 0.3, 0.2), pole (0.95, 0.85, 0.3), ground (0.35, 0.55, 0.3)
    camera fov 60 deg, eye (0, -6, 1.5) -> origin, near 0.05, far 100
    light  dir (-1, -1, -2) (normalised in the kernel), colour 0.8, ambient 0.2
    background 0.41 grey (the reference's observed clear colour)
"""
from __future__ import annotations

import math

import torch

from ...config import PBRConfig
from ...renderer.renderer import PBRRenderer


def _state_rows(state_batch, n: int, device) -> torch.Tensor:
    """[B, 4] float32 on the device, padded with its last row or truncated to n
    (cartpole/renderer.py:112-122)."""
    st = torch.as_tensor(state_batch, dtype=torch.float32).detach()
    if st.device != device:
        st = st.to(device, non_blocking=True)
    st = st.reshape(-1, st.shape[-1])
    B = int(st.shape[0])
    if B < n:
        st = torch.cat([st, st[-1:].expand(n - B, -1)], dim=0)
    elif B > n:
        st = st[:n]
    return st


class CartPoleRenderer(PBRRenderer):
    # _step only poses nodes: step() poses and renders in one launch
    fuse_poses = True

    def __init__(self, cfg: PBRConfig | dict | None = None):
        super().__init__(cfg)
        ips = 1
        self.rail_size = (6.0, 0.05, 0.05)
        self.cart_size = (1.2, 0.8, 0.5)
        self.pole_size = (0.1, 0.1, 2.0)
        self.rail_pos_color = (0.2, 0.2, 0.2, 1.0)
        self.cart_pos_color_range = ((0.6, 0.8, 1.0, 1.0), (1.0, 0.6, 0.8, 1.0))
        self.pole_pos_color = (1.0, 0.7, 0.2, 1.0)

        self.rail = self.add_node("models/box", model_pivot_relative_point=(0.5, 0.5, 0.5),
                                  model_scale=self.rail_size, instances_per_scene=ips,
                                  shared_across_scenes=True)
        self.cart = self.add_node("models/box", model_pivot_relative_point=(0.5, 0.5, 0.5),
                                  model_scale=self.cart_size, instances_per_scene=ips,
                                  shared_across_scenes=False)
        self.pole = self.add_node("models/box", model_pivot_relative_point=(0.5, 0.5, 0.05),
                                  model_scale=self.pole_size, instances_per_scene=ips,
                                  shared_across_scenes=False)

        n = int(self.cfg.num_scenes)
        dev = self.device
        self.rail_pos = torch.zeros((1, ips, 3), dtype=torch.float32, device=dev)
        self.cart_pos = torch.zeros((n, ips, 3), dtype=torch.float32, device=dev)
        self.pole_pos = self.cart_pos.clone() + torch.tensor(
            [0, (self.cart_size[1] + self.pole_size[1]) * 0.5, 0], dtype=torch.float32, device=dev)
        self.rail.set_positions(self.rail_pos)
        self.cart.set_positions(self.cart_pos)
        self.pole.set_positions(self.pole_pos)

        self.rail_base_color = torch.ones((1, ips, 4), dtype=torch.float32) * torch.tensor(
            self.rail_pos_color, dtype=torch.float32)
        start_col = torch.tensor(self.cart_pos_color_range[0], dtype=torch.float32)
        end_col = torch.tensor(self.cart_pos_color_range[1], dtype=torch.float32)
        t = torch.linspace(0.0, 1.0, steps=n, dtype=torch.float32)
        self.cart_base_color = start_col.unsqueeze(0) + (end_col - start_col).unsqueeze(0) * t.unsqueeze(1)
        self.pole_base_color = torch.ones((n, ips, 4), dtype=torch.float32) * torch.tensor(
            self.pole_pos_color, dtype=torch.float32)
        self.rail.set_colors(self.rail_base_color)
        self.cart.set_colors(self.cart_base_color)
        self.pole.set_colors(self.pole_base_color)

        self.pole_hpr = torch.zeros((n, ips, 3), dtype=torch.float32, device=dev)
        self.pole_hpr[:, :, 1] = math.pi * 0.5
        self.pole.set_hprs(self.pole_hpr)

        self.pole_y = (self.cart_size[1] + self.pole_size[1]) * 0.5

        self.add_camera()
        self._pbr_cam.set_positions(torch.tensor([5, 5, 2], dtype=torch.float32))
        self._pbr_cam.look_at(torch.tensor([0, 0, 0], dtype=torch.float32))
        self.add_light()
        self.setup_environment()

    def _step(self, state_batch: torch.Tensor | None = None):
        """State -> cart position and pole position + pitch (renderer.py:98-138):
        cart M = T(x, 0, 0), pole M = T(x, 0.45, 0) Ry(theta), one pose launch
        for both nodes reading the state columns in place."""
        if state_batch is None:
            return
        st = _state_rows(state_batch, int(self.cfg.num_scenes), self.device)
        self.set_poses(self.pose_spec(st))

    def pose_spec(self, st: torch.Tensor) -> list:
        """(node, pose components) of a [N, 4] state, read in place."""
        x = st[:, 0]
        return [(self.cart, dict(x=x, s=self.cart.scale_b11.reshape(-1))),
                (self.pole, dict(x=x, y=self.pole_y, p=st[:, 2],
                                 s=self.pole.scale_b11.reshape(-1)))]


class CartPoleGroundRenderer(PBRRenderer):
    # _step only poses nodes: step() poses and renders in one launch
    fuse_poses = True

    CART_SCALE = (0.4, 0.2, 0.2)
    POLE_SCALE = (0.05, 0.05, 1.0)
    GROUND_SIZE = 10.0
    CART_COLOR = (0.9, 0.3, 0.2, 1.0)
    POLE_COLOR = (0.95, 0.85, 0.3, 1.0)
    GROUND_COLOR = (0.35, 0.55, 0.3, 1.0)
    EYE = (0.0, -6.0, 1.5)
    FOV_Y_DEG = 60.0
    LIGHT_DIR = (-1.0, -1.0, -2.0)
    LIGHT_COL = (0.8, 0.8, 0.8)
    AMBIENT = (0.2, 0.2, 0.2)

    def __init__(self, cfg: PBRConfig | dict | None = None):
        super().__init__(cfg)
        n = int(self.cfg.num_scenes)
        dev = self.device
        self.ground = self.add_node("models/quad", instances_per_scene=1, shared_across_scenes=True,
                                    model_scale=(self.GROUND_SIZE, self.GROUND_SIZE, 1.0),
                                    model_pivot_relative_point=(0.5, 0.5, 0.0))
        self.cart = self.add_node("models/box", instances_per_scene=1, model_scale=self.CART_SCALE,
                                  model_pivot_relative_point=(0.5, 0.5, 0.5))
        self.pole = self.add_node("models/box", instances_per_scene=1, model_scale=self.POLE_SCALE,
                                  model_pivot_relative_point=(0.5, 0.5, 0.0))
        self.ground.set_colors(torch.tensor(self.GROUND_COLOR).reshape(1, 4))
        self.cart.set_colors(torch.tensor(self.CART_COLOR).reshape(1, 4).expand(n, 4))
        self.pole.set_colors(torch.tensor(self.POLE_COLOR).reshape(1, 4).expand(n, 4))
        self.cart_pos = torch.zeros((n, 3), dtype=torch.float32, device=dev)
        self.cart_pos[:, 2] = 0.5 * self.CART_SCALE[2]
        self.pole_pos = torch.zeros((n, 3), dtype=torch.float32, device=dev)
        self.pole_pos[:, 2] = self.CART_SCALE[2]
        self.pole_hpr = torch.zeros((n, 3), dtype=torch.float32, device=dev)
        self.cart.set_positions(self.cart_pos)
        self.pole.set_positions(self.pole_pos)
        self.add_camera(fov_y_deg=self.FOV_Y_DEG)
        self._pbr_cam.set_positions(torch.tensor(self.EYE, dtype=torch.float32))
        self._pbr_cam.look_at(torch.tensor([0.0, 0.0, 0.0]))
        self.add_light(ambient=self.AMBIENT, dir_dir=self.LIGHT_DIR, dir_col=self.LIGHT_COL,
                       strength=1.0)
        self.setup_environment()

    def _step(self, state_batch: torch.Tensor | None = None):
        """State (x, x_dot, theta, theta_dot) -> cart T(x, 0, 0.1) and pole
        T(x, 0, 0.2) Ry(theta): one pose launch for both nodes."""
        if state_batch is None:
            return
        st = _state_rows(state_batch, int(self.cfg.num_scenes), self.device)
        self.set_poses(self.pose_spec(st))

    def pose_spec(self, st: torch.Tensor) -> list:
        """(node, pose components) of a [N, 4] state, read in place."""
        x = st[:, 0]
        return [(self.cart, dict(x=x, z=0.5 * self.CART_SCALE[2])),
                (self.pole, dict(x=x, z=self.CART_SCALE[2], p=st[:, 2]))]
