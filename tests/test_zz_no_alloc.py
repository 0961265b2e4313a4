"""pbr_render_noalloc (the library's own scratch memory) against pbr_render.

Each case renders a batch through both entry points via the C ABI and
compares colour, depth and visibility bitwise; then a smaller batch of the
same kind, so the second no-allocation call runs on scratch memory that
still holds the first call's records and queue counters.  Cases cover every
kernel family that uses scratch memory (tile path with and without the
persistent queue and per-scene instance counts, the scene kernel, the
persistent chunked kernel) and the row-span kernel, which uses none."""
from __future__ import annotations

import ctypes

import pytest
import torch

from pybatchrender import _abi, _native, workloads
from pybatchrender.envs.cartpole import CartPoleConfig, CartPoleGroundRenderer

pytestmark = pytest.mark.gpu

_DT = {"uint8": (torch.uint8, _abi.PBR_DTYPE_U8), "bfloat16": (torch.bfloat16, _abi.PBR_DTYPE_BF16),
       "float32": (torch.float32, _abi.PBR_DTYPE_F32)}


def _bits(t: torch.Tensor) -> torch.Tensor:
    t = t.detach().cpu().contiguous()
    return t.view(torch.int32) if t.dtype == torch.float32 else (
        t.view(torch.int16) if t.dtype == torch.bfloat16 else t)


def _render(r, noalloc: bool) -> dict:
    lib = _native.library()
    N, C, H, W = r.output_shape()
    tdt, code = _DT[r.cfg.output_dtype]
    dev = r.device
    out = {"pixels": torch.full((N, C, H, W), 7, dtype=tdt, device=dev),
           "depth": torch.full((N, H, W), 7.0, dtype=torch.float32, device=dev),
           "vis": torch.full((N, H, W), 7, dtype=torch.int32, device=dev)}
    desc, keep = r.scene_descriptor()
    opts = r.render_opts()
    fo = _abi.pbr_frame_out()
    fo.color, fo.depth, fo.vis = (out["pixels"].data_ptr(), out["depth"].data_ptr(),
                                  out["vis"].data_ptr())
    fo.channels, fo.dtype = C, code
    stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    if noalloc:
        rc = lib.pbr_render_noalloc(ctypes.byref(desc), ctypes.byref(fo), ctypes.byref(opts), stream)
    else:
        nbytes = _native.workspace_bytes(desc, opts)
        ws = torch.empty(max(nbytes, 256), dtype=torch.uint8, device=dev)
        rc = lib.pbr_render(ctypes.byref(desc), ctypes.byref(fo), ctypes.byref(opts),
                            ctypes.c_void_p(ws.data_ptr()), nbytes, stream)
    assert rc == _abi.PBR_OK, _native.last_error()
    torch.cuda.synchronize()
    del keep
    return out


def _same(a: dict, b: dict, label: str) -> None:
    for key in ("pixels", "depth", "vis"):
        assert torch.equal(_bits(a[key]), _bits(b[key])), f"{label}: {key} differs"


def _cartpole(n, **cfg):
    g = torch.Generator("cpu").manual_seed(n)
    st = torch.zeros(n, 4)
    st[:, 0] = torch.empty(n).uniform_(-2.4, 2.4, generator=g)
    st[:, 2] = torch.empty(n).uniform_(-0.21, 0.21, generator=g)
    r = CartPoleGroundRenderer(CartPoleConfig(num_scenes=n, device="cuda", **cfg))
    r._step(st.cuda())
    return r


CASES = {
    "tiles persistent, per-scene counts": lambda n: workloads.uneven_batch(
        num_scenes=n, resolution=84, max_instances=48, seed=n, kernel="tiles", persistent=True),
    "tiles per tile, per-scene counts": lambda n: workloads.uneven_batch(
        num_scenes=n, resolution=64, max_instances=32, seed=n, kernel="tiles", tile_size=16),
    "tiles, fixed counts, float32 + depth": lambda n: workloads.many_cubes(
        num_scenes=n, resolution=96, instances=24, seed=n, kernel="tiles"),
    "scene kernel": lambda n: _cartpole(n, kernel="scene"),
    "chunked persistent": lambda n: workloads.many_cubes(
        num_scenes=n, resolution=64, instances=16, seed=n, kernel="chunked", persistent=True,
        output_dtype="bfloat16"),
    "row-span kernel (no scratch)": lambda n: _cartpole(n, kernel="rows"),
}


@pytest.mark.parametrize("case", list(CASES))
def test_noalloc_matches_render_on_two_calls(cuda, case):
    for n in (160, 37):  # the second call is the smaller batch
        print(f"[noalloc] {case}: N={n}", flush=True)
        r = CASES[case](n)
        a = _render(r, noalloc=True)
        b = _render(r, noalloc=False)
        _same(a, b, f"{case} N={n}")


def test_noalloc_rejects_batches_beyond_its_scratch(cuda):
    r = workloads.many_cubes(num_scenes=2048, resolution=128, instances=64, seed=1, kernel="tiles")
    desc, keep = r.scene_descriptor()
    opts = r.render_opts()
    need = _native.workspace_bytes(desc, opts)
    assert need > _abi.PBR_NOALLOC_SCRATCH_BYTES
    N, C, H, W = r.output_shape()
    px = torch.zeros((N, C, H, W), dtype=torch.float32, device="cuda")
    fo = _abi.pbr_frame_out()
    fo.color, fo.channels, fo.dtype = px.data_ptr(), C, _abi.PBR_DTYPE_F32
    lib = _native.library()
    rc = lib.pbr_render_noalloc(ctypes.byref(desc), ctypes.byref(fo), ctypes.byref(opts),
                                ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == _abi.PBR_ERR_WORKSPACE
    assert "scratch" in _native.last_error() and str(need) in _native.last_error()
    torch.cuda.synchronize()
    assert not px.any()  # nothing was enqueued
    del keep
