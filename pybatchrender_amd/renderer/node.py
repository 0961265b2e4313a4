"""PBRNode: an instanced mesh with per-instance transforms and colours.

Mirrors pybatchrender/renderer/node.py:12-342.  State that the reference kept
as CPU tensors and uploaded into GL buffer textures every step
(``_upload_mat`` node.py:110-114, ``set_colors`` :156-161) is kept as device
tensors the kernel reads in place:

    transforms_b44  [buf_instances, 4, 4]  M = [[R*s, p], [0, 0, 0, 1]]
    colors_b4       [buf_instances, 4]     RGBA (white by default)

with ``buf_instances = ips`` for nodes shared across scenes and ``N * ips``
otherwise (node.py:45-46).  Instance i of scene s reads row
``shared ? i : s * ips + i`` (basic.vert:25-28).

Extension: ``set_instance_counts(counts[N])`` draws only the first counts[s]
instances in scene s (uneven batches); the reference always draws all.
"""
from __future__ import annotations

import warnings
from collections.abc import Sequence
from typing import Literal

import torch

from .mesh import Mesh, bake, load_model
from .shader_context import PBRShaderContext


# Pose components a node takes when a pose leaves them out.
_POSE_DEFAULTS = {"x": 0.0, "y": 0.0, "z": 0.0, "h": 0.0, "p": 0.0, "r": 0.0, "s": 1.0}


class PBRNode(PBRShaderContext):
    def __init__(
        self,
        showbase,
        model_path: "str | Mesh | None",
        num_scenes: int = 1,
        instances_per_scene: int = 1,
        texture=None,
        model_pivot_relative_point: tuple[float, float, float] | None = None,
        model_scale: float | Sequence[float] | None = None,
        model_hpr: tuple[float, float, float] | Sequence[float] | None = None,
        model_scale_units: Literal["relative", "absolute"] = "relative",
        positions: torch.Tensor | None = None,
        hprs: torch.Tensor | None = None,
        scales: torch.Tensor | None = None,
        colors: torch.Tensor | None = None,
        backend: Literal["loop", "instanced"] = "instanced",
        shared_across_scenes: bool = False,
        parent=None,
        name: str | None = None,
    ) -> None:
        if positions is not None or colors is not None or scales is not None or hprs is not None:
            # node.py:36-40: the reference ignores these with a warning
            warnings.warn(
                "initializing positions, colors, scales are not implemented yet, they will be ignored"
            )
        if parent is not None:
            raise NotImplementedError("PBRNode parent is not implemented (node.py:29)")
        super().__init__(showbase, backend=backend)
        self.device = torch.device(getattr(showbase, "device", "cpu"))
        self._pose_cache = None  # GPU pose descriptor of the output buffers (_pose_inputs_gpu)
        self.num_scenes = int(num_scenes)
        self.instances_per_scene = int(instances_per_scene)
        if self.instances_per_scene < 0:
            raise ValueError("instances_per_scene must be >= 0")
        self.shared_across = bool(shared_across_scenes)
        self.total_instances = self.num_scenes * self.instances_per_scene
        self.buf_instances = self.instances_per_scene if self.shared_across else self.total_instances
        self.model_pivot_relative_point = model_pivot_relative_point
        self.model_scale = model_scale
        self.model_hpr = model_hpr
        self.model_scale_units = model_scale_units
        self.name = name or (model_path if isinstance(model_path, str) else "mesh")
        self.instance_counts: torch.Tensor | None = None
        self.instance_count_total = 0
        self._counts_version = 0

        if model_path is not None and (not isinstance(model_path, str) or model_path):
            raw = model_path if isinstance(model_path, Mesh) else load_model(model_path)
            self.mesh = bake(raw, model_scale, model_hpr, model_scale_units, model_pivot_relative_point)
            self.has_geometry = True
        else:
            self.mesh = None
            self.has_geometry = False

        self.set_texture(texture)
        self._register_self()

        if self.has_geometry:
            B = self.buf_instances
            dev = self.device
            self.transforms_b44 = torch.eye(4, dtype=torch.float32, device=dev).repeat(B, 1, 1)
            self.rot3_b33 = torch.eye(3, dtype=torch.float32, device=dev).repeat(B, 1, 1)
            self.scale_b11 = torch.ones((B, 1, 1), dtype=torch.float32, device=dev)
            self.colors_b4 = torch.ones((B, 4), dtype=torch.float32, device=dev)
            self._upload_current_transforms()

    # --- per-step setters (node.py:116-178) -------------------------------
    def _as_dev(self, x, shape_tail: int) -> torch.Tensor:
        t = torch.as_tensor(x, dtype=torch.float32)
        if t.device != self.device:
            t = t.to(self.device, non_blocking=True)
        return t.reshape(-1, shape_tail) if shape_tail else t

    def _upload_current_transforms(self) -> None:
        if not self.has_geometry:
            return
        self.transforms_b44[:, 0:3, 0:3] = self.rot3_b33 * self.scale_b11

    def set_positions(self, pos_si3: torch.Tensor, lazy: bool = False) -> None:
        if not self.has_geometry:
            return
        pos = self._as_dev(pos_si3, 3)
        self.transforms_b44[:, 0:3, 3] = pos[: self.buf_instances]
        if not lazy:
            self._upload_current_transforms()

    def set_hprs(self, hpr_si3: torch.Tensor, lazy: bool = False) -> None:
        if not self.has_geometry:
            return
        hpr = self._as_dev(hpr_si3, 3)
        self.rot3_b33[:, :, :] = type(self)._rotation_mats_from_hpr(hpr[: self.buf_instances])
        if not lazy:
            self._upload_current_transforms()

    def set_scales(self, scale_si1: torch.Tensor | float, lazy: bool = False) -> None:
        if not self.has_geometry:
            return
        if isinstance(scale_si1, (float, int)):
            s = torch.full((self.buf_instances, 1, 1), float(scale_si1), dtype=torch.float32,
                           device=self.device)
        else:
            s = self._as_dev(scale_si1, 0).reshape(-1, 1, 1)
        self.scale_b11 = s[: self.buf_instances]
        if not lazy:
            self._upload_current_transforms()

    def set_colors(self, col_si4: torch.Tensor) -> None:
        if not self.has_geometry:
            return
        col = self._as_dev(col_si4, 4)[: self.buf_instances]
        if col.shape[0] == self.buf_instances or col.shape[0] == 1:
            self.colors_b4[:] = col
        else:  # a shorter upload only replaces the first rows
            self.colors_b4[: col.shape[0]] = col

    def set_transforms(self, mats_b44: torch.Tensor) -> None:
        """Full per-instance matrices (node.py:163-178): the reference splits
        them into rot3/scale (mean row norm) and re-multiplies on upload."""
        if not self.has_geometry:
            return
        mats = self._as_dev(mats_b44, 0).reshape(-1, 4, 4)
        self.transforms_b44 = mats.clone()
        self.rot3_b33 = self.transforms_b44[:, 0:3, 0:3].clone()
        col0 = torch.linalg.norm(self.rot3_b33[:, 0, :], dim=1)
        col1 = torch.linalg.norm(self.rot3_b33[:, 1, :], dim=1)
        col2 = torch.linalg.norm(self.rot3_b33[:, 2, :], dim=1)
        s = (col0 + col1 + col2) / 3.0
        s = torch.clamp(s, min=1e-8)
        self.scale_b11 = s.reshape(-1, 1, 1).to(torch.float32)
        self.rot3_b33 = self.rot3_b33 / self.scale_b11
        self._upload_current_transforms()

    def set_pose(self, x=0.0, y=0.0, z=0.0, h=0.0, p=0.0, r=0.0, s=1.0) -> None:
        """Set every buffer row's pose in one pass:
        M = [[R(h, p, r) * s, (x, y, z)], [0, 0, 0, 1]], R = Rz(h) Ry(p) Rx(r).

        The fused form of set_positions + set_hprs + set_scales + upload
        (node.py:116-154).  Each argument is a number or a 1-D tensor with a
        value per row (strided views such as ``state[:, 0]`` are read in
        place).  On the GPU this is one HIP launch (pbr_set_pose);
        PBRRenderer.set_poses poses several nodes in one launch."""
        prep = self._pose_inputs(x=x, y=y, z=z, h=h, p=p, r=r, s=s)
        if prep is None:
            return
        if self.device.type == "cuda":
            from .. import _native

            desc, keep = prep
            _native.set_poses([desc], torch.cuda.current_stream(self.device).cuda_stream)
            del keep
            return
        self._pose_host(prep)

    def _pose_inputs(self, **comps):
        """Output buffers sized for the pose, and the pose descriptor (GPU) or
        the full per-row component tensors (CPU).  Components not given take
        the identity pose's values (0, and 1 for s)."""
        if not self.has_geometry:
            return None
        if self.device.type == "cuda":
            return self._pose_inputs_gpu(comps)
        comps = {**_POSE_DEFAULTS, **comps}
        B = self.buf_instances
        dev = self.device
        if self.transforms_b44.shape != (B, 4, 4) or not self.transforms_b44.is_contiguous():
            self.transforms_b44 = torch.eye(4, dtype=torch.float32, device=dev).repeat(B, 1, 1)
        if self.rot3_b33.shape != (B, 3, 3) or not self.rot3_b33.is_contiguous():
            self.rot3_b33 = torch.empty((B, 3, 3), dtype=torch.float32, device=dev)
        if self.scale_b11.shape != (B, 1, 1) or not self.scale_b11.is_contiguous():
            self.scale_b11 = torch.empty((B, 1, 1), dtype=torch.float32, device=dev)
        vals = {}
        for k, v in comps.items():
            if isinstance(v, torch.Tensor):
                t = v if (v.dtype == torch.float32 and v.device == dev) else v.to(dev, torch.float32)
                t = t.reshape(-1) if t.dim() != 1 else t
                if t.numel() < B:
                    if t.numel() != 1:
                        raise ValueError(f"pose component {k}: {t.numel()} values for {B} rows")
                    t = t.expand(B)
                vals[k] = t[:B]
            else:
                vals[k] = float(v)
        if dev.type != "cuda":
            return {k: (v if isinstance(v, torch.Tensor) else torch.full((B,), v, device=dev))
                    for k, v in vals.items()}
        from .. import _abi

        d = _abi.pbr_pose_desc()
        d.transforms = self.transforms_b44.data_ptr()
        d.rot3 = self.rot3_b33.data_ptr()
        d.scale = self.scale_b11.data_ptr()
        d.rows = B
        keep = []
        for k, v in vals.items():
            c = getattr(d, k)
            if isinstance(v, torch.Tensor):
                c.src, c.stride, c.value = v.data_ptr(), v.stride(0), 0.0
                keep.append(v)
            else:
                c.src, c.stride, c.value = None, 0, v
        return d, keep

    def _pose_inputs_gpu(self, comps: dict):
        """_pose_inputs on a GPU device: the pose descriptor and the tensors
        it points to.  The output-buffer part of the descriptor is kept per
        node and reused while the buffers are the same objects (a renderer
        step poses every dynamic node, so this runs per step)."""
        from .. import _abi

        B = self.buf_instances
        pc = self._pose_cache
        if (pc is None or pc[0] is not self.transforms_b44 or pc[1] is not self.rot3_b33
                or pc[2] is not self.scale_b11 or pc[3] != B):
            dev = self.device
            if self.transforms_b44.shape != (B, 4, 4) or not self.transforms_b44.is_contiguous():
                self.transforms_b44 = torch.eye(4, dtype=torch.float32, device=dev).repeat(B, 1, 1)
            if self.rot3_b33.shape != (B, 3, 3) or not self.rot3_b33.is_contiguous():
                self.rot3_b33 = torch.empty((B, 3, 3), dtype=torch.float32, device=dev)
            if self.scale_b11.shape != (B, 1, 1) or not self.scale_b11.is_contiguous():
                self.scale_b11 = torch.empty((B, 1, 1), dtype=torch.float32, device=dev)
            t = _abi.pbr_pose_desc()  # the identity pose: components 0, s = 1
            t.s.value = 1.0
            t.transforms = self.transforms_b44.data_ptr()
            t.rot3 = self.rot3_b33.data_ptr()
            t.scale = self.scale_b11.data_ptr()
            t.rows = B
            pc = self._pose_cache = (self.transforms_b44, self.rot3_b33, self.scale_b11, B, t)
        d = _abi.pbr_pose_desc.from_buffer_copy(pc[4])
        dev_index = self.device.index
        keep = []
        for k, v in comps.items():
            c = getattr(d, k)
            if isinstance(v, torch.Tensor):
                if v.dtype is not torch.float32 or not v.is_cuda or v.get_device() != dev_index:
                    v = v.to(self.device, torch.float32)
                if v.dim() != 1:
                    v = v.reshape(-1)
                n = v.numel()
                if n < B:
                    if n != 1:
                        raise ValueError(f"pose component {k}: {n} values for {B} rows")
                    v = v.expand(B)
                elif n > B:
                    v = v[:B]
                c.src, c.stride, c.value = v.data_ptr(), v.stride(0), 0.0
                keep.append(v)
            else:
                c.src, c.stride, c.value = None, 0, float(v)
        return d, keep

    def _pose_host(self, full: dict) -> None:
        B = self.buf_instances
        R = type(self)._rotation_mats_from_hpr(torch.stack([full["h"], full["p"], full["r"]], 1))
        self.rot3_b33[:] = R
        self.scale_b11[:] = full["s"].reshape(B, 1, 1)
        self.transforms_b44[:, 0:3, 0:3] = R * self.scale_b11
        self.transforms_b44[:, 0:3, 3] = torch.stack([full["x"], full["y"], full["z"]], 1)
        self.transforms_b44[:, 3] = torch.tensor([0.0, 0.0, 0.0, 1.0], device=self.device)

    def set_instance_counts(self, counts: torch.Tensor | None) -> None:
        """Active instances per scene ([N] int, clamped to [0, ips]); None = all."""
        if counts is None:
            self.instance_counts = None
            self.instance_count_total = 0
            return
        c = torch.as_tensor(counts).to(device=self.device, dtype=torch.int32).reshape(-1)
        if c.numel() != self.num_scenes:
            raise ValueError(f"counts must have {self.num_scenes} entries, got {c.numel()}")
        self.instance_counts = c.clamp(0, self.instances_per_scene).contiguous()
        # the batch's instance total sizes the tile path's record pool (one sync here)
        self.instance_count_total = int(self.instance_counts.sum().item())
        self._counts_version = self.instance_counts._version

    def counts_total(self) -> int:
        """Sum of the per-scene instance counts (each clamped to [0, ips]),
        recomputed with one device sync if the counts tensor was written in
        place since the last time."""
        c = self.instance_counts
        if c is None:
            return self.num_scenes * self.instances_per_scene
        if c._version != self._counts_version:
            self.instance_count_total = int(c.clamp(0, self.instances_per_scene).sum().item())
            self._counts_version = c._version
        return self.instance_count_total

    # --- misc API of the reference ----------------------------------------
    def set_texture(self, texture=None) -> None:
        if texture is not None and texture is not False:
            raise NotImplementedError("textured nodes are not supported by the HIP renderer")
        self.use_texture = False

    def reparent_to(self, parent) -> None:
        raise NotImplementedError("PBRNode.reparent_to is not implemented yet")

    def pivot_to_rel(self, relative_point=None) -> None:
        """Re-centre the baked mesh on a bounds-relative point (node.py:313-341)."""
        if relative_point is None or self.mesh is None:
            return
        self.mesh = bake(self.mesh, pivot_relative_point=relative_point)
        self.base._mark_geometry_dirty()

    def _register_self(self) -> None:
        nodes = self.base._pbr_nodes
        if self not in nodes:
            nodes.append(self)

    @property
    def flat(self) -> bool:
        return self.mesh.is_flat() if self.mesh is not None else True
