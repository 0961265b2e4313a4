"""SteeringRenderer (pybatchrender/envs/steering/renderer.py:20-271).

Scene per env: the player box, two border rails that follow the player along
the track, up to ``number_of_obstacles`` obstacle spheres (red, or gold when
flagged), a camera that follows the player at ``camera_eye_offset`` looking
down +y, black background.  As in the reference, the cone flag of an
obstacle does not change what is drawn (every obstacle is a sphere).

Model paths resolve like the reference's ``_resolve_model_path`` (:89-107):
a file that exists (as given, or under ``mesh.MODEL_PATH`` /
``PBR_MODEL_PATH``) is read by the glTF / egg readers; otherwise the
built-in equivalents are used (``models/smiley`` -> UV sphere,
``models/cone.egg`` -> cone, ``models/cylinder/scene.gltf`` -> cylinder; all
are scaled to absolute dimensions, so only their shapes matter).

``_step`` keeps everything on the device: player, rails and camera eye are
written from the observation tensor in place (the reference copies the state
to the host, :222-271), and the player colour switches to the crash colour
while the grace period runs.
"""
from __future__ import annotations

import torch

from ...config import PBRConfig
from ...renderer import mesh as meshmod
from ...renderer.renderer import PBRRenderer

_FALLBACK = {"models/smiley": "models/sphere", "models/cone.egg": "models/cone",
             "models/cylinder/scene.gltf": "models/cylinder"}


def resolve_model(path: str):
    """A Mesh for a configured model path (file if found, built-in otherwise)."""
    try:
        return meshmod.load_model(path)
    except FileNotFoundError:
        if path in _FALLBACK:
            return meshmod.load_model(_FALLBACK[path])
        raise


class SteeringRenderer(PBRRenderer):
    # _step reads no posed node's buffers after posing (it sets colours and
    # the camera): step() poses and renders in one launch
    fuse_poses = True

    def __init__(self, cfg: PBRConfig | dict | None = None) -> None:
        super().__init__(cfg)
        n = int(self.cfg.num_scenes)
        dev = self.device
        self._num_scenes = n
        self._setup_called = False
        c = self.cfg
        self._color_player = torch.tensor(c.player_color, dtype=torch.float32, device=dev)
        self._color_red = torch.tensor(c.red_obstacle_color, dtype=torch.float32, device=dev)
        self._color_gold = torch.tensor(c.gold_obstacle_color, dtype=torch.float32, device=dev)
        self._color_border = torch.tensor(c.edge_color, dtype=torch.float32, device=dev)
        self._color_crash = torch.tensor(getattr(c, "crash_player_color", c.player_color),
                                         dtype=torch.float32, device=dev)
        bg = c.background_color
        self.set_background_color(float(bg[0]), float(bg[1]), float(bg[2]), float(bg[3]))
        self.sphere_node = None
        self.cone_node = None
        self._camera_eye_offset = torch.tensor(c.camera_eye_offset, dtype=torch.float32,
                                               device=dev).reshape(1, 3)
        self._build_player()
        self._build_borders()
        self.add_camera(z_far=float(c.z_far), z_near=float(c.z_near), fov_y_deg=float(c.fov_y_deg))
        self.add_light(ambient=c.ambient_light, dir_dir=c.directional_light_dir)

    def _build_player(self) -> None:
        c = self.cfg
        self.player_node = self.add_node(
            resolve_model(c.player_model), instances_per_scene=1,
            model_scale=c.player_dimensions, model_hpr=c.player_model_hpr,
            model_scale_units="absolute", model_pivot_relative_point=c.player_pivot_relative_point,
            shared_across_scenes=False, name="player")
        self.player_node.set_pose()
        self.player_node.set_colors(self._color_player.reshape(1, 4).expand(self._num_scenes, 4))

    def _build_borders(self) -> None:
        c = self.cfg
        border = resolve_model(c.border_model)
        half = float(c.lane_width) * 0.5
        z = float(c.rail_offset[2])
        self.left_border = self.add_node(border, instances_per_scene=1, model_scale=c.rail_dimensions,
                                         model_hpr=c.border_model_hpr, model_scale_units="absolute",
                                         shared_across_scenes=False, name="left_border")
        self.right_border = self.add_node(border, instances_per_scene=1,
                                          model_scale=c.rail_dimensions,
                                          model_hpr=c.border_model_hpr,
                                          model_scale_units="absolute",
                                          shared_across_scenes=False, name="right_border")
        col = self._color_border.reshape(1, 4).expand(self._num_scenes, 4)
        self.left_border.set_pose(x=-half, z=z)
        self.right_border.set_pose(x=half, z=z)
        self.left_border.set_colors(col)
        self.right_border.set_colors(col)

    def build_obstacles(self, obstacles: torch.Tensor | None) -> None:
        """Obstacle spheres from a [B, N, 4] tensor of (x, y, gold, cone)
        (renderer.py:163-218); the first call finishes the scene setup."""
        if obstacles is None or obstacles.numel() == 0:
            if not self._setup_called:
                self.setup_environment()
                self._setup_called = True
            return
        obs = obstacles.to(self.device, torch.float32)
        B, N = int(obs.shape[0]), int(obs.shape[1])
        if self.sphere_node is None or self.sphere_node.instances_per_scene != N:
            if self.sphere_node is not None:
                self._pbr_nodes.remove(self.sphere_node)
            c = self.cfg
            self.sphere_node = self.add_node(
                resolve_model(c.obstacle_sphere_model), instances_per_scene=N,
                model_scale=c.obstacle_dimensions, model_hpr=c.obstacle_model_hpr,
                model_scale_units="absolute",
                model_pivot_relative_point=c.obstacle_pivot_relative_point,
                shared_across_scenes=False, name="obstacles")
            self._mark_geometry_dirty()
        rows = obs[: self._num_scenes].reshape(-1, 4)
        self.sphere_node.set_pose(x=rows[:, 0].contiguous(), y=rows[:, 1].contiguous())
        gold = (rows[:, 2] >= 0.5).unsqueeze(-1)
        self.sphere_node.set_colors(torch.where(gold, self._color_gold, self._color_red))
        if not self._setup_called:
            self.setup_environment()
            self._setup_called = True
        elif self._geometry is None:
            self._build_geometry()

    def _step(self, state_batch: torch.Tensor | None = None) -> None:
        """Observation [B, >=4] -> player, rails, player colour, camera eye."""
        if state_batch is None:
            return
        n = self._num_scenes
        st = torch.as_tensor(state_batch, dtype=torch.float32).detach()
        if st.device != self.device:
            st = st.to(self.device, non_blocking=True)
        st = st.reshape(-1, st.shape[-1])
        if st.shape[0] < n:
            st = torch.cat([st, st[-1:].expand(n - st.shape[0], -1)], 0)
        st = st[:n]
        px, py = st[:, 0].contiguous(), st[:, 1].contiguous()
        half = float(self.cfg.lane_width) * 0.5
        z = float(self.cfg.rail_offset[2])
        self.set_poses([(self.player_node, dict(x=px, y=py)),
                        (self.left_border, dict(x=-half, y=py, z=z)),
                        (self.right_border, dict(x=half, y=py, z=z))])
        grace = (st[:, 3:4] > 0.5) if st.shape[1] > 3 else torch.zeros_like(st[:, :1], dtype=torch.bool)
        self.player_node.set_colors(torch.where(grace, self._color_crash, self._color_player))
        if self._pbr_cam is not None:
            eye = torch.stack([px, py, torch.zeros_like(px)], 1) + self._camera_eye_offset
            self._pbr_cam.set_positions(eye)
