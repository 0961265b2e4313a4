"""Shared scene math of nodes, camera and light (shader_context.py:10-114).

The reference's Panda3D plumbing (shader objects, buffer textures, uniform
broadcast) is gone: scene state lives in device tensors that the kernel reads
directly.  The math is kept operation for operation.
"""
from __future__ import annotations

from typing import Literal

import torch


class PBRShaderContext:
    def __init__(self, showbase, backend: Literal["instanced", "loop"] = "instanced") -> None:
        self.base = showbase
        if backend == "loop":
            raise NotImplementedError(
                "Loop backend is not implemented yet. Use instanced backend instead."
            )
        self.backend = backend

    @staticmethod
    def _pack_columns(mat_batch: torch.Tensor) -> torch.Tensor:
        """(B,4,4) -> column-packed (shader_context.py:43-45).  Kept for API
        parity; the kernel reads row-major matrices, so nothing calls it."""
        return mat_batch.transpose(1, 2).to(torch.float32)

    @staticmethod
    def _rotation_mats_from_hpr(hpr_b3: torch.Tensor) -> torch.Tensor:
        """R = Rz(H) @ Ry(P) @ Rx(R) from radians (shader_context.py:48-84).

        Note: this is the reference's own convention, not Panda3D's HPR."""
        hpr_b3 = hpr_b3.to(torch.float32)
        h, p, r = hpr_b3[:, 0], hpr_b3[:, 1], hpr_b3[:, 2]
        ch, sh = torch.cos(h), torch.sin(h)
        cp, sp = torch.cos(p), torch.sin(p)
        cr, sr = torch.cos(r), torch.sin(r)
        # the product Rz @ Ry @ Rx written out (no batched GEMM for 3x3s)
        chsp, shsp = ch * sp, sh * sp
        R = torch.stack([
            ch * cp, chsp * sr - sh * cr, chsp * cr + sh * sr,
            sh * cp, shsp * sr + ch * cr, shsp * cr - ch * sr,
            -sp, cp * sr, cp * cr,
        ], 1)
        return R.view(-1, 3, 3)

    def _register_self(self) -> None:  # pragma: no cover - abstract in the reference
        raise NotImplementedError
