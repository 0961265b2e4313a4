"""Loader for libpbr_hip.so, the HIP renderer behind the C ABI.

There is no fallback: if the library is missing or the device is not a GPU,
rendering raises.  Build it with ``python -c "import __graft_entry__ as g;
g.build()"`` (or ``pybatchrender_amd._build.build_hip()``).
"""
from __future__ import annotations

import ctypes
from pathlib import Path

from . import _abi

LIB_PATH = Path(__file__).resolve().parent / "libpbr_hip.so"
_LIB: ctypes.CDLL | None = None


def library() -> ctypes.CDLL:
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is not built; run pybatchrender_amd._build.build_hip() first"
            )
        lib = _abi.declare(ctypes.CDLL(str(LIB_PATH)))
        _abi.check_layout(lib)
        _LIB = lib
    return _LIB


def stream_ptr(device) -> int:
    """Raw HIP stream handle of torch's current stream on `device` (a
    torch.device of type cuda)."""
    import torch

    idx = device.index if device.index is not None else torch.cuda.current_device()
    raw = getattr(torch._C, "_cuda_getCurrentRawStream", None)
    if raw is not None:
        return raw(idx)
    return torch.cuda.current_stream(device).cuda_stream


def last_error() -> str:
    msg = library().pbr_last_error()
    return msg.decode() if msg else ""


def render(desc: _abi.pbr_scene_desc, out: _abi.pbr_frame_out, opts: _abi.pbr_render_opts,
           workspace_ptr: int, workspace_bytes: int, stream_ptr: int) -> None:
    rc = library().pbr_render(
        ctypes.byref(desc), ctypes.byref(out), ctypes.byref(opts),
        ctypes.c_void_p(workspace_ptr or None), workspace_bytes, ctypes.c_void_p(stream_ptr or None),
    )
    if rc != _abi.PBR_OK:
        raise RuntimeError(f"pbr_render failed (status {rc}): {last_error()}")


def render_posed(desc: _abi.pbr_scene_desc, out: _abi.pbr_frame_out, opts: _abi.pbr_render_opts,
                 pose_descs: list, workspace_ptr: int, workspace_bytes: int,
                 stream_ptr: int) -> None:
    """pbr_render_posed: pbr_set_poses(pose_descs) then pbr_render, one launch
    when the row-span kernel can apply the poses itself."""
    arr = (_abi.pbr_pose_desc * len(pose_descs))(*pose_descs)
    rc = library().pbr_render_posed(
        ctypes.byref(desc), ctypes.byref(out), ctypes.byref(opts), arr, len(pose_descs),
        ctypes.c_void_p(workspace_ptr or None), workspace_bytes, ctypes.c_void_p(stream_ptr or None),
    )
    if rc != _abi.PBR_OK:
        raise RuntimeError(f"pbr_render_posed failed (status {rc}): {last_error()}")


def workspace_bytes(desc: _abi.pbr_scene_desc, opts: _abi.pbr_render_opts) -> int:
    n = ctypes.c_size_t(0)
    rc = library().pbr_workspace_bytes(ctypes.byref(desc), ctypes.byref(opts), ctypes.byref(n))
    if rc != _abi.PBR_OK:
        raise RuntimeError(f"pbr_workspace_bytes failed (status {rc}): {last_error()}")
    return int(n.value)


def choose_kernel(desc: _abi.pbr_scene_desc, opts: _abi.pbr_render_opts) -> int:
    """PBR_KERNEL_* that pbr_render runs for this descriptor and options."""
    k = ctypes.c_int32(0)
    rc = library().pbr_choose_kernel(ctypes.byref(desc), ctypes.byref(opts), ctypes.byref(k))
    if rc != _abi.PBR_OK:
        raise RuntimeError(f"pbr_choose_kernel failed (status {rc}): {last_error()}")
    return int(k.value)


def set_pose(desc: _abi.pbr_pose_desc, stream_ptr: int) -> None:
    rc = library().pbr_set_pose(ctypes.byref(desc), ctypes.c_void_p(stream_ptr or None))
    if rc != _abi.PBR_OK:
        raise RuntimeError(f"pbr_set_pose failed (status {rc}): {last_error()}")


def set_poses(descs: list, stream_ptr: int) -> None:
    arr = (_abi.pbr_pose_desc * len(descs))(*descs)
    rc = library().pbr_set_poses(arr, len(descs), ctypes.c_void_p(stream_ptr or None))
    if rc != _abi.PBR_OK:
        raise RuntimeError(f"pbr_set_poses failed (status {rc}): {last_error()}")


def backward_workspace_bytes(desc: _abi.pbr_scene_desc, channels: int, grads: _abi.pbr_grads) -> int:
    n = ctypes.c_size_t(0)
    rc = library().pbr_backward_workspace_bytes(ctypes.byref(desc), channels, ctypes.byref(grads),
                                                ctypes.byref(n))
    if rc != _abi.PBR_OK:
        raise RuntimeError(f"pbr_backward_workspace_bytes failed (status {rc}): {last_error()}")
    return int(n.value)


def render_backward(desc: _abi.pbr_scene_desc, vis_ptr: int, grad_ptr: int, channels: int,
                    grads: _abi.pbr_grads, workspace_ptr: int, workspace_bytes: int,
                    stream_ptr: int) -> None:
    rc = library().pbr_render_backward(
        ctypes.byref(desc), ctypes.c_void_p(vis_ptr or None), ctypes.c_void_p(grad_ptr or None),
        channels, ctypes.byref(grads), ctypes.c_void_p(workspace_ptr or None), workspace_bytes,
        ctypes.c_void_p(stream_ptr or None))
    if rc != _abi.PBR_OK:
        raise RuntimeError(f"pbr_render_backward failed (status {rc}): {last_error()}")


def cartpole_step(params: _abi.pbr_cartpole_params, batch: _abi.pbr_cartpole_batch, poses: list,
                  stream_ptr: int) -> None:
    arr = (_abi.pbr_pose_desc * max(len(poses), 1))(*poses)
    rc = library().pbr_cartpole_step(ctypes.byref(params), ctypes.byref(batch), arr, len(poses),
                                     ctypes.c_void_p(stream_ptr or None))
    if rc != _abi.PBR_OK:
        raise RuntimeError(f"pbr_cartpole_step failed (status {rc}): {last_error()}")


def cartpole_reset(params: _abi.pbr_cartpole_params, batch: _abi.pbr_cartpole_batch,
                   stream_ptr: int) -> None:
    rc = library().pbr_cartpole_reset(ctypes.byref(params), ctypes.byref(batch),
                                      ctypes.c_void_p(stream_ptr or None))
    if rc != _abi.PBR_OK:
        raise RuntimeError(f"pbr_cartpole_reset failed (status {rc}): {last_error()}")
