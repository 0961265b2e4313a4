"""ctypes front end of the CPU oracle (oracle/liboracle_pbr.so).

TEST INFRASTRUCTURE ONLY: the checker of the HIP renderer.  Imported by
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg; never by
the pybatchrender_amd package.

``render(desc, keep, ...)`` takes the pbr_scene_desc a PBRRenderer built for
the GPU (plus the tensors it points to), swaps every device pointer for a
host copy of the same tensor and renders all scenes on the CPU.
"""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np
import torch

from pybatchrender import _abi

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle_pbr.so"
_LIB = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def library() -> ctypes.CDLL:
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            build()
        lib = ctypes.CDLL(str(LIB_PATH))
        lib.pbr_oracle_render.argtypes = [ctypes.POINTER(_abi.pbr_scene_desc),
                                          ctypes.POINTER(_abi.pbr_frame_out), ctypes.c_int32,
                                          ctypes.c_int32]
        lib.pbr_oracle_render.restype = ctypes.c_int
        lib.pbr_oracle_render_mt.argtypes = [ctypes.POINTER(_abi.pbr_scene_desc),
                                             ctypes.POINTER(_abi.pbr_frame_out), ctypes.c_int32]
        lib.pbr_oracle_render_mt.restype = ctypes.c_int
        lib.pbr_oracle_coverage.argtypes = [ctypes.POINTER(ctypes.c_int32),
                                            ctypes.POINTER(ctypes.c_int32), ctypes.c_int32,
                                            ctypes.c_int32, ctypes.c_void_p]
        lib.pbr_oracle_coverage.restype = ctypes.c_int
        _LIB = lib
    return _LIB


_DT = {"uint8": (torch.uint8, _abi.PBR_DTYPE_U8), "bfloat16": (torch.bfloat16, _abi.PBR_DTYPE_BF16),
       "float32": (torch.float32, _abi.PBR_DTYPE_F32)}


class HostScene:
    """A host copy of a GPU scene descriptor (pointers swapped to CPU tensors)."""

    def __init__(self, desc: _abi.pbr_scene_desc, keep: list[torch.Tensor]):
        by_ptr = {t.data_ptr(): t for t in keep}
        self.tensors: list[torch.Tensor] = []
        self.desc = _abi.pbr_scene_desc.from_buffer_copy(desc)

        def host(ptr):
            if not ptr:
                return None
            t = by_ptr[ptr].detach().to("cpu").contiguous()
            self.tensors.append(t)
            return t.data_ptr()

        self.desc.vertices = host(desc.vertices)
        self.desc.indices = host(desc.indices)
        self.desc.view_proj = host(desc.view_proj)
        for k in range(desc.num_nodes):
            src, dst = desc.nodes[k], self.desc.nodes[k]
            dst.transforms = host(src.transforms)
            dst.colors = host(src.colors)
            dst.counts = host(src.counts)


def render(desc: _abi.pbr_scene_desc, keep: list[torch.Tensor], channels: int = 3,
           dtype: str = "uint8", threads: int = 0, scenes: tuple[int, int] | None = None) -> dict:
    """Render on the CPU; returns {'pixels', 'depth', 'vis'} CPU tensors."""
    hs = HostScene(desc, keep)
    N, H, W = desc.num_scenes, desc.height, desc.width
    tdt, code = _DT[dtype]
    pixels = torch.empty((N, channels, H, W), dtype=tdt)
    depth = torch.empty((N, H, W), dtype=torch.float32)
    vis = torch.empty((N, H, W), dtype=torch.int32)
    fo = _abi.pbr_frame_out()
    fo.color, fo.depth, fo.vis = pixels.data_ptr(), depth.data_ptr(), vis.data_ptr()
    fo.channels, fo.dtype = channels, code
    lib = library()
    if scenes is not None:
        rc = lib.pbr_oracle_render(ctypes.byref(hs.desc), ctypes.byref(fo), scenes[0], scenes[1])
    else:
        import os
        nt = threads or len(os.sched_getaffinity(0))
        rc = lib.pbr_oracle_render_mt(ctypes.byref(hs.desc), ctypes.byref(fo), nt)
    if rc != 0:
        raise RuntimeError(f"oracle render failed with status {rc}")
    return {"pixels": pixels, "depth": depth, "vis": vis}


def render_renderer(renderer, threads: int = 0) -> dict:
    """Oracle frame of a PBRRenderer's current state (same descriptor as the GPU)."""
    desc, keep = renderer.scene_descriptor()
    return render(desc, keep, channels=int(renderer.cfg.num_channels),
                  dtype=renderer.cfg.output_dtype, threads=threads)


def coverage(X, Y, W: int, H: int) -> tuple[np.ndarray, int]:
    """Coverage mask (row 0 = top) of one fixed-point window-space triangle."""
    xs = (ctypes.c_int32 * 3)(*[int(v) for v in X])
    ys = (ctypes.c_int32 * 3)(*[int(v) for v in Y])
    mask = np.zeros((H, W), dtype=np.uint8)
    n = library().pbr_oracle_coverage(xs, ys, W, H, mask.ctypes.data)
    return mask, n
