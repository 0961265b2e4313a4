"""PBRCam: per-scene view matrices and a shared perspective projection.

Same public API and the same formulas as pybatchrender/renderer/camera.py
(10-433), on device tensors:
    P   GL perspective, f = 1 / tan(fov / 2), aspect = whole-atlas W / H
        (camera.py:150-166, the aspect at :154)
    V   rows (s, u, -f, [0 0 0 1]) with translation (-s.e, -u.e, f.e), where
        f = normalize(forward), s = normalize(f x up), u = s x f (:201-240)
    VP  = P @ V per scene, [N, 4, 4] (:252-262)

The camera keeps each scene's basis (right, up, forward) and eye as [N, 3]
tensors and composes V from them whenever one changes.  The atlas tiling
(:110-135) only fixes the projection aspect here: scene k renders straight
into out[k].  Interactive base-camera sync (:346-428) is not carried over.
"""
from __future__ import annotations

import math
from typing import Literal

import torch

from .shader_context import PBRShaderContext


def _unit(v: torch.Tensor) -> torch.Tensor:
    """Rows of v scaled to unit length (rows shorter than 1e-8 are divided by 1e-8)."""
    return v / torch.linalg.norm(v, dim=-1, keepdim=True).clamp_min(1e-8)


def _per_scene(x, scenes: int, device: torch.device, what: str) -> torch.Tensor:
    """A (3,), (1, 3) or (scenes, 3) value as a float32 [scenes, 3] tensor."""
    t = torch.as_tensor(x, dtype=torch.float32).to(device)
    t = t.reshape(1, -1) if t.ndim == 1 else t
    if t.ndim != 2 or t.shape[1] != 3:
        raise ValueError(f"{what}: expected 3-vectors, one per scene or one for all; "
                         f"got shape {tuple(t.shape)}")
    if t.shape[0] == 1:
        return t.expand(scenes, 3).clone()
    if t.shape[0] != scenes:
        raise ValueError(f"{what}: {t.shape[0]} rows for {scenes} scenes")
    return t


def _look_basis(forward: torch.Tensor, up: torch.Tensor):
    """(right, up, forward) of a camera looking along `forward` with `up` as
    the up hint: f = normalize(forward), s = normalize(f x normalize(up)),
    u = s x f."""
    f = _unit(forward)
    s = _unit(torch.cross(f, _unit(up), dim=-1))
    return s, torch.cross(s, f, dim=-1), f


def _perspective(fov_y_deg: float, near: float, far: float, aspect: float) -> torch.Tensor:
    """The GL perspective matrix (gluPerspective) in float32."""
    f = 1.0 / math.tan(math.radians(fov_y_deg) * 0.5)
    P = torch.zeros((4, 4), dtype=torch.float32)
    P[0, 0], P[1, 1] = f / aspect, f
    P[2, 2] = (far + near) / (near - far)
    P[2, 3] = (2.0 * far * near) / (near - far)
    P[3, 2] = -1.0
    return P


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
        K, dev = self.num_scenes, self.device
        self.hpr_k3 = torch.zeros((K, 3), dtype=torch.float32, device=dev)
        self.target_k3 = torch.zeros((K, 3), dtype=torch.float32, device=dev)
        self._proj: tuple | None = None  # (fov, near, far, aspect) key and P
        self.P_k44 = self._get_projection()
        # default pose (renderer.py:131-135 / camera.py:56-58): at (0, -12, 0)
        # looking along +y with +z up
        self.eye_k3 = _per_scene((0.0, -12.0, 0.0), K, dev, "eye")
        self._right, self._up, self._fwd = _look_basis(
            _per_scene((0.0, 1.0, 0.0), K, dev, "forward"),
            _per_scene((0.0, 0.0, 1.0), K, dev, "up"))
        self.V_k44 = self._compose_view()
        self.VP_k44 = torch.matmul(self.P_k44.unsqueeze(0), self.V_k44).contiguous()
        self.sync_from_base_cam = False
        self._register_self()

    # --- tiles / projection ------------------------------------------------
    def _set_tiles(self) -> None:
        """Atlas grid (camera.py:110-135): cols = ceil(sqrt(N)) and rows =
        ceil(N / cols) unless given; here it only fixes the aspect."""
        if self.num_scenes is None and (self.cols is None or self.rows is None):
            raise ValueError("give num_scenes, or both cols and rows")
        if self.cols is None and self.rows is None:
            self.cols = math.ceil(math.sqrt(self.num_scenes))
        if self.rows is None:
            self.rows = math.ceil(self.num_scenes / self.cols)
        if self.num_scenes is None:
            self.num_scenes = self.cols * self.rows

    def tile_rects(self) -> torch.Tensor:
        """(u0, u1, v0, v1) of every scene's atlas tile (camera.py:123-134),
        scene 0 in the top-left corner."""
        k = torch.arange(int(self.num_scenes), dtype=torch.float32)
        cols, rows = float(self.cols), float(self.rows)
        c, r = torch.remainder(k, cols), torch.floor(k / cols)
        return torch.stack([c / cols, (c + 1.0) / cols, 1.0 - (r + 1.0) / rows, 1.0 - r / rows], 1)

    def _aspect(self) -> float:
        w, h = self.base.cfg.window_resolution
        return max(1e-6, float(w) / max(1, h))

    def _get_projection(self) -> torch.Tensor:
        """P for the current field of view and clip planes, rebuilt when one
        of them (or the aspect) changed; frozen after the first build when
        the projection is fixed."""
        if self.fixed_projection and self._proj is not None:
            return self._proj[1]
        key = (self.fov_y_deg, self.z_near, self.z_far, self._aspect())
        if self._proj is None or self._proj[0] != key:
            self._proj = (key, _perspective(*key).to(self.device))
        return self._proj[1]

    def set_projection(self, fov_y_deg: float | None = None, z_near: float | None = None,
                       z_far: float | None = None) -> None:
        """New projection parameters; VP picks them up at the next view
        update (camera.py:168-181)."""
        for name, v in (("fov_y_deg", fov_y_deg), ("z_near", z_near), ("z_far", z_far)):
            if v is not None:
                setattr(self, name, float(v))
        self._proj = None

    # --- basis and view matrix -------------------------------------------------
    def _ensure_kx3(self, arr, name: str) -> torch.Tensor:
        return _per_scene(arr, self.num_scenes, self.device, name)

    @staticmethod
    def _normalize(v: torch.Tensor) -> torch.Tensor:
        return _unit(v)

    def _compose_view(self) -> torch.Tensor:
        """V from the basis and the eye: rows s, u, -f with translations
        -s.e, -u.e, f.e, then (0, 0, 0, 1)."""
        s, u, f, e = self._right, self._up, self._fwd, self.eye_k3
        K = s.shape[0]
        top = torch.stack([
            torch.cat([s, -(s * e).sum(-1, keepdim=True)], 1),
            torch.cat([u, -(u * e).sum(-1, keepdim=True)], 1),
            torch.cat([-f, (f * e).sum(-1, keepdim=True)], 1),
        ], 1)
        last = torch.tensor([0.0, 0.0, 0.0, 1.0], device=s.device).expand(K, 1, 4)
        return torch.cat([top, last], 1)

    def _update_view(self, eye_k3=None, forward_k3=None, up_k3=None) -> None:
        """Re-aim (a new forward and/or up; the other one is kept) and/or
        move the eye, then recompose V.  VP is left to _update_vp."""
        if forward_k3 is None and up_k3 is None and eye_k3 is None:
            return
        if forward_k3 is not None or up_k3 is not None:
            fwd = self._fwd if forward_k3 is None else self._ensure_kx3(forward_k3, "forward")
            up = self._up if up_k3 is None else self._ensure_kx3(up_k3, "up")
            self._right, self._up, self._fwd = _look_basis(fwd, up)
        if eye_k3 is not None:
            self.eye_k3 = self._ensure_kx3(eye_k3, "eye")
        self.V_k44 = self._compose_view()

    def _update_vp(self) -> None:
        self.P_k44 = self._get_projection()
        self.VP_k44 = torch.matmul(self.P_k44.unsqueeze(0), self.V_k44).contiguous()

    def _refresh(self, lazy: bool) -> None:
        if not lazy:
            self._update_vp()

    @staticmethod
    def _fwd_up_from_hpr(hpr_k3: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor]:
        """Forward (+y) and up (+z) of a camera turned by R(h, p, r): the
        second and third columns of R (shader_context.py:48-84)."""
        R = PBRShaderContext._rotation_mats_from_hpr(hpr_k3)
        return R[:, :, 1].contiguous(), R[:, :, 2].contiguous()

    def get_vp(self) -> torch.Tensor:
        return self.VP_k44.clone()

    # --- public API (camera.py:268-344) ---------------------------------------
    def look_at(self, target_k3, lazy: bool = False) -> None:
        self._update_view(forward_k3=self._fwd_from_lookat(target_k3))
        self._refresh(lazy)

    def set_positions(self, eye_k3, keep_lookat: bool = False, lazy: bool = False) -> None:
        if keep_lookat:
            # the camera keeps a direction, not a target point, so the point
            # it looked at is unknown after a move
            raise NotImplementedError(
                "keep_lookat=True needs the look-at target, which the camera does not store; "
                "use set_positions_and_lookat(eye, target)")
        self.set_eye(eye_k3, lazy=lazy)

    def set_positions_and_lookat(self, eye_k3, target_k3, lazy: bool = False) -> None:
        eye = self._ensure_kx3(eye_k3, "eye")
        self._update_view(eye_k3=eye, forward_k3=self._fwd_from_lookat(target_k3, eye))
        self._refresh(lazy)

    def _fwd_from_lookat(self, target_k3, eye_k3=None) -> torch.Tensor:
        eye = self.eye_k3 if eye_k3 is None else self._ensure_kx3(eye_k3, "eye")
        return self._ensure_kx3(target_k3, "target") - eye

    def set_hprs(self, hpr_k3, lazy: bool = False) -> None:
        fwd, up = self._fwd_up_from_hpr(self._ensure_kx3(hpr_k3, "hpr"))
        self._update_view(forward_k3=fwd, up_k3=up)
        self._refresh(lazy)

    def set_eye(self, eye_k3, lazy: bool = False) -> None:
        self._update_view(eye_k3=eye_k3)
        self._refresh(lazy)

    def set_forward(self, forward_k3, lazy: bool = False) -> None:
        self._update_view(forward_k3=forward_k3)
        self._refresh(lazy)

    def set_up(self, up_k3, lazy: bool = False) -> None:
        self._update_view(up_k3=up_k3)
        self._refresh(lazy)

    def set_right(self, right_k3, lazy: bool = False) -> None:
        """Keep the forward direction and choose up so that the camera's
        right axis becomes normalize(right): up hint = right x forward."""
        right = _unit(self._ensure_kx3(right_k3, "right"))
        self._update_view(forward_k3=self._fwd, up_k3=torch.cross(right, self._fwd, dim=-1))
        self._refresh(lazy)

    def get_eye(self) -> torch.Tensor:
        return self.eye_k3.clone()

    def get_forward(self) -> torch.Tensor:
        return self._fwd.clone()

    def get_up(self) -> torch.Tensor:
        return self._up.clone()

    def get_right(self) -> torch.Tensor:
        return self._right.clone()

    def attach(self, node) -> None:
        """Nodes read the camera at render time; nothing to bind."""

    def attach_all(self) -> None:
        """Nodes read the camera at render time; nothing to bind."""

    def _register_self(self) -> None:
        self.base._pbr_cam = self
