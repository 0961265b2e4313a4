"""PBRCam: per-scene view matrices and a shared perspective projection.

Mirrors pybatchrender/renderer/camera.py:10-433 with device tensors:
    P   GL perspective, aspect = whole-atlas W/H (camera.py:150-166, :154)
    V   look-at rows (s, u, -f), translation (-s.e, -u.e, f.e) (:201-240)
    VP  = P @ V per scene, [N, 4, 4] (:252-262)
The tile buffer of the atlas (:110-135) has no use without an atlas: scene k
renders straight into out[k].  Interactive base-camera sync (:346-428) is
not carried over.
"""
from __future__ import annotations

import math
from typing import Literal

import torch

from .shader_context import PBRShaderContext


class PBRCam(PBRShaderContext):
    def __init__(
        self,
        showbase,
        num_scenes: int,
        cols: int | None = None,
        rows: int | None = None,
        backend: Literal["loop", "instanced"] = "instanced",
        fov_y_deg: float = 55.0,
        z_near: float = 0.05,
        z_far: float = 100.0,
        auto_tiles: bool = True,
        fixed_projection: bool = True,
    ) -> None:
        super().__init__(showbase, backend=backend)
        self.device = torch.device(getattr(showbase, "device", "cpu"))
        self.num_scenes = int(num_scenes)
        self.fov_y_deg = float(fov_y_deg)
        self.z_near = float(z_near)
        self.z_far = float(z_far)
        self.fixed_projection = fixed_projection
        self.cols = cols
        self.rows = rows
        self._set_tiles()
        dev = self.device
        K = self.num_scenes
        self.eye_k3 = torch.zeros((K, 3), dtype=torch.float32, device=dev)
        self.forward_k3 = torch.zeros((K, 3), dtype=torch.float32, device=dev)
        self.up_k3 = torch.zeros((K, 3), dtype=torch.float32, device=dev)
        self.hpr_k3 = torch.zeros((K, 3), dtype=torch.float32, device=dev)
        self.target_k3 = torch.zeros((K, 3), dtype=torch.float32, device=dev)
        self.V_k44 = torch.eye(4, dtype=torch.float32, device=dev)[None].repeat(K, 1, 1)
        self.VP_k44 = torch.zeros((K, 4, 4), dtype=torch.float32, device=dev)
        self._proj_cache_key = None
        self._proj_cache: torch.Tensor | None = None
        self.P_k44 = self._get_projection()
        self._update_view(
            eye_k3=torch.tensor([0.0, -12.0, 0.0]),
            forward_k3=torch.tensor([0.0, 1.0, 0.0]),
            up_k3=torch.tensor([0.0, 0.0, 1.0]),
        )
        self._update_vp()
        self.sync_from_base_cam = False
        self._register_self()

    # --- tiles / projection ------------------------------------------------
    def _set_tiles(self) -> None:
        """Atlas columns/rows (camera.py:110-135); they only fix the aspect."""
        if self.num_scenes is None and (self.cols is None or self.rows is None):
            raise ValueError("num_scenes or (cols and rows) must be provided")
        if self.cols is None and self.rows is None:
            self.cols = math.ceil(math.sqrt(self.num_scenes))
        if self.rows is None:
            self.rows = math.ceil(self.num_scenes / self.cols)
        if self.num_scenes is None:
            self.num_scenes = self.cols * self.rows

    def tile_rects(self) -> torch.Tensor:
        """(u0, u1, v0, v1) of every scene's atlas tile (camera.py:123-134)."""
        K = int(self.num_scenes)
        idx = torch.arange(K, dtype=torch.float32)
        cols, rows = float(self.cols), float(self.rows)
        col = torch.remainder(idx, cols)
        row = torch.floor(idx / cols)
        return torch.stack(
            [col / cols, (col + 1.0) / cols, 1.0 - (row + 1.0) / rows, 1.0 - row / rows], dim=1
        )

    def _aspect(self) -> float:
        w, h = self.base.cfg.window_resolution
        return max(1e-6, float(w) / max(1, h))

    def _get_projection(self) -> torch.Tensor:
        if self.fixed_projection and self._proj_cache is not None:
            return self._proj_cache
        aspect = self._aspect()
        key = (float(self.fov_y_deg), float(self.z_near), float(self.z_far), float(aspect))
        if self._proj_cache_key != key or self._proj_cache is None:
            f = 1.0 / math.tan(math.radians(self.fov_y_deg) * 0.5)
            P = torch.zeros((4, 4), dtype=torch.float32)
            P[0, 0] = f / aspect
            P[1, 1] = f
            P[2, 2] = (self.z_far + self.z_near) / (self.z_near - self.z_far)
            P[2, 3] = (2.0 * self.z_far * self.z_near) / (self.z_near - self.z_far)
            P[3, 2] = -1.0
            self._proj_cache_key = key
            self._proj_cache = P.to(self.device)
        return self._proj_cache

    def set_projection(self, fov_y_deg: float | None = None, z_near: float | None = None,
                       z_far: float | None = None) -> None:
        """Changes take effect at the next view update, as in camera.py:168-181."""
        if fov_y_deg is not None:
            self.fov_y_deg = float(fov_y_deg)
        if z_near is not None:
            self.z_near = float(z_near)
        if z_far is not None:
            self.z_far = float(z_far)
        self._proj_cache_key = None
        self._proj_cache = None

    # --- view ----------------------------------------------------------------
    def _ensure_kx3(self, arr, name: str) -> torch.Tensor:
        a = torch.as_tensor(arr, dtype=torch.float32)
        if a.device != self.device:
            a = a.to(self.device)
        if a.ndim == 1:
            a = a.unsqueeze(0)
        if a.shape[-1] != 3:
            raise ValueError(f"{name} must have shape (K,3) or (3,), got {tuple(a.shape)}")
        if a.shape[0] == 1 and self.num_scenes > 1:
            a = a.repeat(self.num_scenes, 1)
        if a.shape[0] != self.num_scenes:
            raise ValueError(f"{name} first dim must be K={self.num_scenes}, got {a.shape[0]}")
        return a

    @staticmethod
    def _normalize(v: torch.Tensor) -> torch.Tensor:
        n = torch.linalg.norm(v, dim=-1, keepdim=True)
        return v / n.clamp_min(1e-8)

    def _update_view(self, eye_k3=None, forward_k3=None, up_k3=None) -> None:
        basis_changed = (forward_k3 is not None) or (up_k3 is not None)
        eye_changed = eye_k3 is not None
        if not basis_changed and not eye_changed:
            return
        if basis_changed:
            if forward_k3 is not None:
                f = type(self)._normalize(self._ensure_kx3(forward_k3, "forward_k3"))
            else:
                f = -self.V_k44[:, 2, 0:3]
            if up_k3 is not None:
                up_in = type(self)._normalize(self._ensure_kx3(up_k3, "up_k3"))
            else:
                up_in = self.V_k44[:, 1, 0:3]
            s = type(self)._normalize(torch.cross(f, up_in, dim=-1))
            u = torch.cross(s, f, dim=-1)
            self.V_k44[:, 0, 0:3] = s
            self.V_k44[:, 1, 0:3] = u
            self.V_k44[:, 2, 0:3] = -f
        else:
            s = self.V_k44[:, 0, 0:3]
            u = self.V_k44[:, 1, 0:3]
            f = -self.V_k44[:, 2, 0:3]
        if eye_changed:
            self.eye_k3 = self._ensure_kx3(eye_k3, "eye_k3")
        self.V_k44[:, 0, 3] = -(s * self.eye_k3).sum(-1)
        self.V_k44[:, 1, 3] = -(u * self.eye_k3).sum(-1)
        self.V_k44[:, 2, 3] = (f * self.eye_k3).sum(-1)

    @staticmethod
    def _fwd_up_from_hpr(hpr_k3: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor]:
        R = PBRShaderContext._rotation_mats_from_hpr(hpr_k3)
        return R[:, :, 1].clone(), R[:, :, 2].clone()  # R @ (0,1,0), R @ (0,0,1)

    def _update_vp(self) -> None:
        P = self._get_projection()
        self.P_k44 = P
        self.VP_k44 = torch.matmul(P.unsqueeze(0), self.V_k44).contiguous()

    def get_vp(self) -> torch.Tensor:
        return self.VP_k44.clone()

    # --- public API (camera.py:268-344) ---------------------------------------
    def look_at(self, target_k3, lazy: bool = False) -> None:
        self._update_view(forward_k3=self._fwd_from_lookat(target_k3))
        if not lazy:
            self._update_vp()

    def set_positions(self, eye_k3, keep_lookat: bool = False, lazy: bool = False) -> None:
        if keep_lookat:
            raise NotImplementedError(
                "Target can't be reconsturcted from eye and forward, needs to be stored in memory then"
            )
        self.set_eye(eye_k3, lazy=lazy)

    def set_positions_and_lookat(self, eye_k3, target_k3, lazy: bool = False) -> None:
        eye_k3 = self._ensure_kx3(eye_k3, "eye_k3")
        target_k3 = self._ensure_kx3(target_k3, "target_k3")
        self._update_view(eye_k3=eye_k3, forward_k3=self._fwd_from_lookat(target_k3, eye_k3))
        if not lazy:
            self._update_vp()

    def _fwd_from_lookat(self, target_k3, eye_k3=None) -> torch.Tensor:
        eye = eye_k3 if eye_k3 is not None else self.eye_k3
        return self._ensure_kx3(target_k3, "target_k3") - self._ensure_kx3(eye, "eye_k3")

    def set_hprs(self, hpr_k3, lazy: bool = False) -> None:
        fwd, up = self._fwd_up_from_hpr(self._ensure_kx3(hpr_k3, "hpr_k3"))
        self._update_view(forward_k3=fwd, up_k3=up)
        if not lazy:
            self._update_vp()

    def set_eye(self, eye_k3, lazy: bool = False) -> None:
        self._update_view(eye_k3=eye_k3)
        if not lazy:
            self._update_vp()

    def set_forward(self, forward_k3, lazy: bool = False) -> None:
        self._update_view(forward_k3=forward_k3)
        if not lazy:
            self._update_vp()

    def set_up(self, up_k3, lazy: bool = False) -> None:
        self._update_view(up_k3=up_k3)
        if not lazy:
            self._update_vp()

    def set_right(self, right_k3, lazy: bool = False) -> None:
        right_k3 = type(self)._normalize(self._ensure_kx3(right_k3, "right_k3"))
        f = -self.V_k44[:, 2, 0:3]
        u = torch.cross(right_k3, f, dim=-1)
        self._update_view(forward_k3=f, up_k3=u)
        if not lazy:
            self._update_vp()

    def get_eye(self) -> torch.Tensor:
        return self.eye_k3.clone()

    def get_forward(self) -> torch.Tensor:
        return (-self.V_k44[:, 2, 0:3]).clone()

    def get_up(self) -> torch.Tensor:
        return self.V_k44[:, 1, 0:3].clone()

    def get_right(self) -> torch.Tensor:
        return self.V_k44[:, 0, 0:3].clone()

    def attach(self, node) -> None:
        """Nodes read the camera at render time; nothing to bind."""

    def attach_all(self) -> None:
        """Nodes read the camera at render time; nothing to bind."""

    def _register_self(self) -> None:
        self.base._pbr_cam = self
