"""Throughput of PBRRenderer.step on one MI355X.

Workload: the 64x64 CartPole scene with cart, pole and a shared ground quad
(CartPoleGroundRenderer), 4096 scenes, uint8 RGB output [4096, 3, 64, 64].
A step is one renderer.step(state): one row-span launch (pbr_render_posed)
that evaluates the cart/pole instance transforms of each scene from the
state [4096, 4], writes them to the nodes' buffers and renders (setup, row
lists and pixel walk of every scene).
States are seeded (x ~ U(-2.4, 2.4), theta ~ U(-0.21, 0.21)), one batch per
step, and resident on the GPU before timing starts.

Printed JSON line (one, from this single process):
  value        scene-frames/s over `--steps` timed renderer.step calls
               (synchronize on both sides); `graphed_step` is the same step
               captured once as a hipGraph and replayed (it adds the copy of
               each state batch into the graph's static input)
  roofline     the step's launch (one row-span kernel that also poses the
               nodes): algorithmic bytes per call (uint8 RGB pixels + 80 B
               per dynamic instance + 64 B VP per scene, + per posed row the
               rot3 / scale outputs and the pose components read) / average
               time of the call from HIP events around back-to-back calls on
               the launch stream, vs 8 TB/s; `render_only` is the render
               call of the same posed scene without the poses
  cpu_baseline the CPU oracle (a C restatement of the same render path)
               rendering a 256-scene sample of the same workload on all host
               cores: median render (p10/p90 beside it) and a one-thread
               median
  fwd_bwd      the same workload with float32 output: plain forward vs
               forward with the visibility buffer + pbr_render_backward
               (colour, light and vertex gradients), C-ABI launches timed
               with HIP events, and the autograd step through diff.render
  fwd_bwd_differentiable  the same pair on the differentiable-render
               workload: 512 scenes 128x128 float32, 16 boxes + 16 UV
               spheres per scene, backward of an L2 loss to a target image
               rendered from perturbed parameters
  env_loop     16384-scene on-device CartPole loop (pbr_cartpole_step +
               render, bf16 RGB + depth), one hipGraph replay per step
  env_step     CartPole-v0 (reference scene) env.step at the same size:
               dynamics + auto-reset + render, scene-frames/s
  tile_workloads  many-cubes (1024 scenes 128x128 float32 RGB + depth) and
               the uneven persistent batch (8192 scenes 84x84, 1-256
               instances) vs its uniform twin: render-call time (HIP
               events), HBM rate and the uneven/uniform ns-per-triangle ratio
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
PUBLISHED_ENV_STEP_L4 = 980562.0  # reference notebook, 4096 scenes 64x64, NVIDIA L4


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def seeded_states(n: int, batches: int, seed: int = 0):
    import torch

    g = torch.Generator("cpu").manual_seed(seed)
    st = torch.zeros(batches, n, 4)
    st[:, :, 0] = torch.empty(batches, n).uniform_(-2.4, 2.4, generator=g)
    st[:, :, 2] = torch.empty(batches, n).uniform_(-0.21, 0.21, generator=g)
    return st


def algorithmic_bytes(r) -> int:
    """Pixels (and depth) written + the instance rows drawn (80 B each) + VP
    (64 B) per scene; shared rows, meshes and the light are read once."""
    N, C, H, W = r.output_shape()
    esize = {"uint8": 1, "bfloat16": 2, "float32": 4}[r.cfg.output_dtype]
    per_scene = H * W * C * esize + 64
    if r.cfg.output_depth:
        per_scene += H * W * 4
    inst = 0
    for n in r._drawn_nodes():
        if n.shared_across:
            continue
        if n.instance_counts is not None:
            inst += int(n.instance_counts.sum().item())
        else:
            inst += N * n.instances_per_scene
    return N * per_scene + 80 * inst


def triangles_drawn(r) -> int:
    t = 0
    for n in r._drawn_nodes():
        if n.instance_counts is not None:
            cnt = int(n.instance_counts.sum().item())
        else:
            cnt = n.instances_per_scene * (1 if n.shared_across else r.num_scenes)
        t += cnt * n.mesh.num_triangles
    return t


def config_renderer(name: str, persistent: bool):
    """The renderer of a named workload for --kernel-only timing."""
    from pybatchrender import workloads

    if name == "many-cubes":
        return workloads.many_cubes(persistent=persistent)
    if name == "uneven":
        return workloads.uneven_batch(persistent=persistent)
    if name == "uneven-uniform":
        return workloads.uneven_batch(persistent=persistent, uniform=True)
    raise ValueError(name)


def step_poses(r, state) -> list:
    """The pose descriptors renderer.step(state) hands to its render call
    (renderers with fuse_poses), with the tensors they point to; [] for
    renderers that render without poses."""
    if state is None or not getattr(r, "fuse_poses", False):
        return []
    r._pending_poses = []
    try:
        r._step(state)
        return list(r._pending_poses)
    finally:
        r._pending_poses = None


def pose_bytes(poses: list) -> int:
    """Algorithmic bytes a fused step adds to the render's: per posed row the
    rot3 (36 B) and scale (4 B) outputs and 4 B per component read from a
    source (the 64-byte transforms row, which the render counts as read, is
    written instead)."""
    total = 0
    for _, (d, _keep) in poses:
        n_src = sum(1 for k in "xyzhprs" if getattr(d, k).src)
        total += d.rows * ((36 if d.rot3 else 0) + (4 if d.scale else 0) + 4 * n_src)
    return total


def kernel_time_us(r, launches: int = 200, state=None) -> float:
    """Average render-kernel duration from HIP events bracketing back-to-back
    launches of one prebuilt descriptor on the current stream: the launch
    renderer.step(state) makes (pbr_render_posed with the step's poses for
    fuse_poses renderers), or the plain render without a state."""
    import torch

    from pybatchrender import _abi, _native

    N, C, H, W = r.output_shape()
    tdt, code = {"uint8": (torch.uint8, _abi.PBR_DTYPE_U8),
                 "bfloat16": (torch.bfloat16, _abi.PBR_DTYPE_BF16),
                 "float32": (torch.float32, _abi.PBR_DTYPE_F32)}[r.cfg.output_dtype]
    out = torch.empty((N, C, H, W), dtype=tdt, device="cuda")
    depth = torch.empty((N, H, W), dtype=torch.float32, device="cuda") if r.cfg.output_depth else None
    desc, keep = r.scene_descriptor()
    fo = _abi.pbr_frame_out()
    fo.color, fo.channels, fo.dtype = out.data_ptr(), C, code
    fo.depth = depth.data_ptr() if depth is not None else None
    opts = r.render_opts()
    lib = _native.library()
    ws_bytes = _native.workspace_bytes(desc, opts)
    ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device="cuda")
    stream = torch.cuda.current_stream()
    sp = ctypes.c_void_p(stream.cuda_stream)
    poses = step_poses(r, state)
    if poses:
        arr = (_abi.pbr_pose_desc * len(poses))(*[d for _, (d, _k) in poses])
        fn = lib.pbr_render_posed
        args = (ctypes.byref(desc), ctypes.byref(fo), ctypes.byref(opts), arr, len(poses),
                ctypes.c_void_p(ws.data_ptr()), ws_bytes, sp)
    else:
        fn = lib.pbr_render
        args = (ctypes.byref(desc), ctypes.byref(fo), ctypes.byref(opts),
                ctypes.c_void_p(ws.data_ptr()), ws_bytes, sp)
    for _ in range(20):
        fn(*args)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(launches):
        rc = fn(*args)
        if rc:
            raise RuntimeError(_native.last_error())
    e1.record(stream)
    e1.synchronize()
    del keep, poses
    return e0.elapsed_time(e1) * 1000.0 / launches


def time_launches_us(launch, launches: int = 200, warm: int = 20) -> float:
    """Average duration of `launch()` (which enqueues on torch's current
    stream) from HIP events around back-to-back calls."""
    import torch

    stream = torch.cuda.current_stream()
    for _ in range(warm):
        launch()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(launches):
        launch()
    e1.record(stream)
    e1.synchronize()
    return e0.elapsed_time(e1) * 1000.0 / launches


def _fwd_bwd_launch_times(r, gout, launches: int) -> dict:
    """Average durations (us) of the plain float32 forward (colour only), the
    forward with the visibility buffer and pbr_render_backward of `gout`
    (gradients for every node's colours, the light and the vertex pool), all
    C-ABI launches back to back on one stream, timed with HIP events."""
    import torch

    from pybatchrender import _abi, _native

    N, C, H, W = r.output_shape()
    desc, keep = r.scene_descriptor()
    opts = r.render_opts()
    lib = _native.library()
    sp = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    out = torch.empty((N, C, H, W), dtype=torch.float32, device="cuda")
    vis = torch.empty((N, H, W), dtype=torch.int32, device="cuda")
    ws_bytes = _native.workspace_bytes(desc, opts)
    ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device="cuda")
    plain = _abi.pbr_frame_out()
    plain.color, plain.channels, plain.dtype = out.data_ptr(), C, _abi.PBR_DTYPE_F32
    withvis = _abi.pbr_frame_out()
    withvis.color, withvis.vis, withvis.channels, withvis.dtype = (out.data_ptr(), vis.data_ptr(),
                                                                   C, _abi.PBR_DTYPE_F32)

    def render(fo):
        rc = lib.pbr_render(ctypes.byref(desc), ctypes.byref(fo), ctypes.byref(opts),
                            ctypes.c_void_p(ws.data_ptr()), ws_bytes, sp)
        if rc:
            raise RuntimeError(_native.last_error())

    g = _abi.pbr_grads()
    gcols = [torch.zeros(nd.buf_instances, 4, device="cuda") for nd in r._geometry["nodes"]]
    gverts = torch.zeros_like(r._geometry["vertices"])
    glight = torch.zeros(10, device="cuda")
    for j, t in enumerate(gcols):
        g.colors[j] = t.data_ptr()
    g.vertices, g.light = gverts.data_ptr(), glight.data_ptr()
    gout = gout.to(torch.float32).contiguous()
    bws_bytes = _native.backward_workspace_bytes(desc, C, g)
    bws = torch.empty(max(bws_bytes, 1), dtype=torch.uint8, device="cuda")

    def backward():
        rc = lib.pbr_render_backward(ctypes.byref(desc), ctypes.c_void_p(vis.data_ptr()),
                                     ctypes.c_void_p(gout.data_ptr()), C, ctypes.byref(g),
                                     ctypes.c_void_p(bws.data_ptr()), bws_bytes, sp)
        if rc:
            raise RuntimeError(_native.last_error())

    t_plain = time_launches_us(lambda: render(plain), launches)
    t_fwd = time_launches_us(lambda: render(withvis), launches)
    t_bwd = time_launches_us(backward, launches)
    del keep
    return {"plain": t_plain, "fwd": t_fwd, "bwd": t_bwd}


def _fwd_bwd_record(r, t: dict, workload: str) -> dict:
    N, C, H, W = r.output_shape()
    hw = H * W
    inst = 0
    for nd in r._drawn_nodes():
        if not nd.shared_across:
            inst += N * nd.instances_per_scene
    params = 80 * inst + 64 * N
    b_plain = N * hw * C * 4 + params
    b_fwd = b_plain + N * hw * 4
    b_bwd = N * (hw * C * 4 + hw * 4) + params
    return {
        "workload": workload,
        "plain_fwd_us": round(t["plain"], 2),
        "fwd_with_vis_us": round(t["fwd"], 2),
        "bwd_us": round(t["bwd"], 2),
        "fwd_bwd_over_plain_fwd": round((t["fwd"] + t["bwd"]) / t["plain"], 3),
        "fwd_bwd_scene_frames_per_s": round(N / ((t["fwd"] + t["bwd"]) * 1e-6), 1),
        "plain_fwd_scene_frames_per_s": round(N / (t["plain"] * 1e-6), 1),
        "hbm_gbs": {"plain_fwd": round(b_plain / t["plain"] / 1e3, 1),
                    "fwd_with_vis": round(b_fwd / t["fwd"] / 1e3, 1),
                    "bwd": round(b_bwd / t["bwd"] / 1e3, 1)},
        "hbm_frac": {"plain_fwd": round(b_plain / t["plain"] / 1e3 / HBM_PEAK_GBS, 4),
                     "fwd_with_vis": round(b_fwd / t["fwd"] / 1e3 / HBM_PEAK_GBS, 4),
                     "bwd": round(b_bwd / t["bwd"] / 1e3 / HBM_PEAK_GBS, 4)},
    }


def _autograd_step_ms(step, iters: int = 20) -> float:
    import torch

    for _ in range(5):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(iters):
        step()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / iters * 1e3


def fwd_bwd_leg(n: int, state, launches: int = 100) -> dict:
    """Differentiable render of the CartPole workload with float32 output:
    plain forward (colour only) vs forward with the visibility buffer plus
    pbr_render_backward (gradients for instance colours, light and vertices),
    all through the C ABI on one stream; and the same pair through autograd
    (pybatchrender.diff)."""
    import torch

    from pybatchrender import diff
    from pybatchrender.envs.cartpole import CartPoleConfig, CartPoleGroundRenderer

    r = CartPoleGroundRenderer(CartPoleConfig(num_scenes=n, device="cuda",
                                              output_dtype="float32"))
    r.step(state)
    N, C, H, W = r.output_shape()
    gout = torch.randn((N, C, H, W), device="cuda")
    rec = _fwd_bwd_record(r, _fwd_bwd_launch_times(r, gout, launches),
                          f"CartPole cart+pole+ground quad, {N} scenes {W}x{H}, float32 RGB")
    p = diff.scene_params(r)

    def step():
        diff.render(r, p).backward(gout)

    rec["autograd_step_ms"] = round(_autograd_step_ms(step), 3)
    rec["note"] = ("C-ABI launches back to back on one stream, HIP events; autograd_step = "
                   "diff.render + backward incl. parameter copies and gradient zeroing")
    return rec


def differentiable_leg(n: int = 512, launches: int = 50) -> dict:
    """The differentiable-render workload: 128 x 128 float32, 32 instances
    per scene (16 boxes, 16 UV spheres), L2 loss to a target rendered from
    perturbed colours, light direction and vertices.  Plain forward vs
    forward with visibility + backward of the loss gradient 2 (img -
    target), C-ABI launches; and one autograd step of the loss."""
    import torch

    from pybatchrender import diff, workloads

    r, tgt = workloads.differentiable(num_scenes=n)
    with torch.no_grad():
        target = diff.render(r, tgt)
        p = diff.scene_params(r, requires_grad=False)
        gout = 2.0 * (diff.render(r, p) - target)
    N, C, H, W = r.output_shape()
    rec = _fwd_bwd_record(r, _fwd_bwd_launch_times(r, gout, launches),
                          f"differentiable render, {N} scenes {W}x{H} float32, 16 boxes + 16 "
                          "UV spheres (16x8) per scene, L2 loss to a perturbed-parameter target")
    p = diff.scene_params(r)

    def step():
        for t in p.tensors():
            t.grad = None
        ((diff.render(r, p) - target) ** 2).sum().backward()

    rec["autograd_step_ms"] = round(_autograd_step_ms(step, iters=10), 3)
    rec["kernel"] = render_kernel_name(r)
    return rec


def env_loop_leg(n: int = 16384, steps: int = 1000) -> dict:
    """On-device CartPole env loop: per step one hipGraph replay of
    {pbr_cartpole_step (dynamics, auto-reset, poses), setup, raster}, bf16 RGB
    + float32 depth, random actions drawn on the device.  Per-step time from
    HIP events between consecutive replays; throughput from the wall clock."""
    import torch

    from pybatchrender.envs.cartpole import CartPoleConfig, CartPoleGroundRenderer
    from pybatchrender.envs.cartpole.loop import CartPoleLoop

    cfg = CartPoleConfig(num_scenes=n, device="cuda", output_dtype="bfloat16", output_depth=True,
                         theta_threshold_deg=math.degrees(0.21), x_threshold=2.4, max_steps=500,
                         init_x_range=(-0.05, 0.05), init_x_dot_range=(-0.05, 0.05),
                         init_theta_range_deg=(-math.degrees(0.05), math.degrees(0.05)),
                         init_theta_dot_range_deg=(-math.degrees(0.05), math.degrees(0.05)))
    loop = CartPoleLoop(CartPoleGroundRenderer(cfg), cfg, actions="random", graph=True, seed=0)
    for _ in range(20):
        loop.step()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
    dones = torch.zeros(n, dtype=torch.int64, device="cuda")
    t0 = time.perf_counter()
    ev[0].record(stream)
    for k in range(steps):
        loop.graph.replay()
        ev[k + 1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    per = sorted(ev[k].elapsed_time(ev[k + 1]) * 1000.0 for k in range(steps))
    dones += loop.done.to(torch.int64)
    hw = 64 * 64
    step_bytes = n * (hw * (3 * 2 + 4) + 2 * 80 + 64 + 16 + 32)
    return {
        "workload": f"{n} CartPole scenes 64x64 bf16 RGB + fp32 depth, device dynamics "
                    "(dt 0.02, force +-10 N, Bernoulli(0.5) actions, reset |theta|>0.21 or "
                    "|x|>2.4), one hipGraph per step",
        "steps": steps,
        "scene_frames_per_s": round(n * steps / wall, 1),
        "step_us_p50": round(per[len(per) // 2], 2),
        "step_us_p99": round(per[min(len(per) - 1, int(len(per) * 0.99))], 2),
        "step_us_mean": round(sum(per) / len(per), 2),
        "hbm_gbs_at_p50": round(step_bytes / per[len(per) // 2] / 1e3, 1),
        "episodes_ending_last_step": int(dones.sum().item()),
    }


def tile_workloads_leg(launches: int = 30) -> dict:
    """The many-cubes batch (1024 scenes 128x128, float32 RGB + depth, 64
    boxes each) and the uneven persistent batch (8192 scenes 84x84 uint8,
    1-256 instances of box / sphere / cylinder) against its uniform twin
    (every scene at the mean instance count): render-call time from HIP
    events, algorithmic HBM rate, and ns per triangle."""
    import torch

    from pybatchrender import workloads

    out = {}
    r = workloads.many_cubes()
    kus = kernel_time_us(r, launches=launches)
    nb = algorithmic_bytes(r)
    out["many_cubes"] = {
        "workload": "1024 scenes 128x128 float32 RGB + float32 depth, 64 random boxes each",
        "kernel": render_kernel_name(r), "render_us": round(kus, 2),
        "scene_frames_per_s": round(r.num_scenes / kus * 1e6, 1),
        "bytes_per_launch": nb, "achieved_gbs": round(nb / kus / 1e3, 1),
        "frac": round(nb / kus / 1e3 / HBM_PEAK_GBS, 4),
        "triangles": triangles_drawn(r)}
    del r
    torch.cuda.empty_cache()
    res = {}
    for name, uniform in (("uneven", False), ("uniform", True)):
        r = workloads.uneven_batch(persistent=True, uniform=uniform)
        kus = kernel_time_us(r, launches=max(5, launches // 3))
        tri = triangles_drawn(r)
        nb = algorithmic_bytes(r)
        res[name] = {"render_us": round(kus, 2), "triangles": tri,
                     "ns_per_triangle": round(kus * 1e3 / tri, 5),
                     "scene_frames_per_s": round(r.num_scenes / kus * 1e6, 1),
                     "achieved_gbs": round(nb / kus / 1e3, 1), "kernel": render_kernel_name(r)}
        del r
        torch.cuda.empty_cache()
    out["uneven_persistent"] = {
        "workload": "8192 scenes 84x84 uint8, log-uniform 1-256 instances of box / UV sphere / "
                    "cylinder, one persistent (scene, tile) work-queue launch; uniform: same "
                    "draws with every scene at the mean instance count",
        **res,
        "uneven_over_uniform_per_triangle": round(res["uneven"]["ns_per_triangle"] /
                                                  res["uniform"]["ns_per_triangle"], 3)}
    return out


def render_kernel_name(r) -> str:
    """The kernel family pbr_render runs for this renderer (pbr_choose_kernel)."""
    return r._chosen_kernel()


KERNEL_LABEL = {"rows": "rows_kernel (setup + lists + row walk, one launch)",
                "scene": "setup_kernel + raster_kernel",
                "tiles": "tile_prep_kernel + tile_setup_kernel + tile_raster_kernel",
                "chunked": "chunked_kernel"}


def measured_traffic(kernel: str) -> int | None:
    """HBM bytes per render call of this workload from the committed counter
    profile of the same kernel (FETCH_SIZE and WRITE_SIZE scaled by a copy of
    known size)."""
    prof = ROOT / "profiles" / "render_cartpole_ground_4096x64x64_u8.md"
    if not prof.exists():
        return None
    traffic, prof_kernel = None, None
    for line in prof.read_text().splitlines():
        if line.startswith("traffic_bytes_per_call:"):
            traffic = int(line.split(":")[1].strip())
        if line.startswith("kernel_family:"):
            prof_kernel = line.split(":")[1].strip()
    return traffic if prof_kernel == kernel else None


def copy_calibration(mib: int = 256, reps: int = 5) -> None:
    """A plain device copy of known size: its FETCH_SIZE / WRITE_SIZE counter
    values calibrate the render kernel's traffic figures."""
    import torch

    src = torch.empty(mib << 20, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": mib << 20, "reps": reps}), flush=True)


def _oracle_render_times(lib, desc, fo, threads: int, min_seconds: float,
                         min_iters: int = 20) -> list[float]:
    """Seconds per oracle render of `desc` on `threads` threads: 3 warm-up
    renders, then at least `min_iters` timed ones and `min_seconds` of work."""
    for _ in range(3):
        lib.pbr_oracle_render_mt(ctypes.byref(desc), ctypes.byref(fo), threads)
    times, start = [], time.perf_counter()
    while True:
        t0 = time.perf_counter()
        lib.pbr_oracle_render_mt(ctypes.byref(desc), ctypes.byref(fo), threads)
        t1 = time.perf_counter()
        times.append(t1 - t0)
        if t1 - start >= min_seconds and len(times) >= min_iters:
            return times


def cpu_baseline(r, sample_scenes: int = 256, min_seconds: float = 8.0,
                 single_seconds: float = 3.0) -> dict:
    """The CPU oracle on a bounded sample: the first `sample_scenes` scenes,
    timed per render on every host core the process may use (the reported
    value, from the median render) and on one thread."""
    from oracle import oracle

    desc, keep = r.scene_descriptor()
    hs = oracle.HostScene(desc, keep)
    hs.desc.num_scenes = sample_scenes
    import numpy as np
    import torch

    from pybatchrender import _abi

    N, C, H, W = sample_scenes, 3, hs.desc.height, hs.desc.width
    pix = torch.empty((N, C, H, W), dtype=torch.uint8)
    fo = _abi.pbr_frame_out()
    fo.color, fo.channels, fo.dtype = pix.data_ptr(), C, _abi.PBR_DTYPE_U8
    threads = len(os.sched_getaffinity(0))
    lib = oracle.library()
    t_mt = np.array(_oracle_render_times(lib, hs.desc, fo, threads, min_seconds))
    t_1 = np.array(_oracle_render_times(lib, hs.desc, fo, 1, single_seconds, min_iters=5))
    p10, p50, p90 = np.percentile(t_mt, [10, 50, 90])
    return {"value": float(N / p50), "unit": "scene-frames/s", "cores": threads, "kind": "port",
            "p10_p90": [round(float(N / p90), 1), round(float(N / p10), 1)],
            "single_thread": round(N / float(np.median(t_1)), 1),
            "sample": f"{N} scenes of the same workload (first {N} states); median of "
                      f"{len(t_mt)} renders on {threads} threads ({t_mt.sum():.1f} s) after 3 "
                      f"warm-up renders; single_thread: median of {len(t_1)} renders on 1 "
                      f"thread; C oracle built -O3 -march=x86-64-v3 -ffp-contract=off"}


def env_step_rate(n: int, steps: int) -> float:
    import torch

    import pybatchrender as pbr

    env = pbr.envs.make("CartPole-v0", num_scenes=n)
    td = env.reset()
    for _ in range(3):
        td["action"] = env.action_spec.rand()
        td = env.step(td)["next"]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        td["action"] = env.action_spec.rand()
        td = env.step(td)["next"]
    torch.cuda.synchronize()
    return n * steps / (time.perf_counter() - t0)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # 1000 steps of about 28 us: a timed region long enough that one host
    # hiccup does not move the per-step figure
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--scenes", type=int, default=4096)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-env", action="store_true")
    ap.add_argument("--no-bwd", action="store_true")
    ap.add_argument("--no-loop", action="store_true")
    ap.add_argument("--no-tiles", action="store_true",
                    help="skip the many-cubes and uneven-batch legs")
    ap.add_argument("--loop-scenes", type=int, default=16384)
    ap.add_argument("--loop-steps", type=int, default=1000)
    ap.add_argument("--kernel-only", action="store_true",
                    help="only the back-to-back render launches (for counter profiles)")
    ap.add_argument("--kernel", default="auto", choices=["auto", "scene", "tiles", "chunked"],
                    help="kernel family of the render call (PBRConfig.kernel)")
    ap.add_argument("--bwd-only", action="store_true",
                    help="only the fwd+bwd leg (for kernel profiles of the backward)")
    ap.add_argument("--diff-only", action="store_true",
                    help="only the differentiable-render workload's fwd+bwd leg")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="with --kernel-only: also copy 256 MiB device-to-device 5 times")
    ap.add_argument("--config", default="cartpole",
                    choices=["cartpole", "many-cubes", "uneven", "uneven-uniform"],
                    help="with --kernel-only: the workload.  cartpole: 4096 CartPole scenes "
                         "64x64 uint8; many-cubes: 1024 scenes 128x128 float32 RGB + depth, 64 "
                         "random boxes each; uneven: 8192 scenes 84x84 uint8, 1-256 instances "
                         "(log-uniform) of box / sphere / cylinder; uneven-uniform: the same "
                         "draws with every scene at the mean instance count")
    ap.add_argument("--persistent", action="store_true",
                    help="with --config other than cartpole: persistent work queue")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this renderer shards scenes with no "
              "communication, so --gpus N > 1 is not run here.", file=sys.stderr)
        return 2

    import torch

    from pybatchrender.envs.cartpole import CartPoleConfig, CartPoleGroundRenderer

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        return 1
    if args.kernel_only and args.config != "cartpole":
        log(f"[bench] kernel timing only: {args.config}")
        rc_ = config_renderer(args.config, args.persistent)
        kus = kernel_time_us(rc_, launches=args.steps)
        tri = triangles_drawn(rc_)
        nb = algorithmic_bytes(rc_)
        print(json.dumps({"config": args.config, "persistent": bool(args.persistent),
                          "kernel": render_kernel_name(rc_), "kernel_us": round(kus, 2),
                          "scene_frames_per_s": round(rc_.num_scenes / kus * 1e6, 1),
                          "triangles": tri, "ns_per_triangle": round(kus * 1e3 / tri, 4),
                          "bytes_per_launch": nb, "gbs": round(nb / kus / 1e3, 1)}), flush=True)
        if args.copy_calibration:
            copy_calibration()
        return 0
    n = args.scenes
    r = CartPoleGroundRenderer(CartPoleConfig(num_scenes=n, device="cuda", kernel=args.kernel))
    states = seeded_states(n, args.warmup + args.steps).cuda()
    if args.kernel_only:
        log("[bench] kernel timing only")
        r.step(states[0])
        kus = kernel_time_us(r, launches=args.steps, state=states[0])
        nb = algorithmic_bytes(r) + pose_bytes(step_poses(r, states[0]))
        print(json.dumps({"kernel_us": kus, "bytes_per_launch": nb,
                          "render_only_us": kernel_time_us(r, launches=args.steps)}), flush=True)
        if args.copy_calibration:
            copy_calibration()
        return 0
    if args.diff_only:
        log("[bench] differentiable-render workload only")
        print(json.dumps(differentiable_leg()), flush=True)
        return 0
    if args.bwd_only:
        log("[bench] fwd+bwd leg only")
        print(json.dumps(fwd_bwd_leg(n, states[0])), flush=True)
        if args.copy_calibration:
            copy_calibration()
        return 0
    log(f"[bench] eager: warmup {args.warmup} steps, N={n}")
    for i in range(args.warmup):
        r.step(states[i])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        r.step(states[args.warmup + i])
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    value = n * args.steps / el

    log("[bench] graphed steps")
    g = r.capture_step(states[0])
    for i in range(args.warmup):
        g(states[i])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        g(states[args.warmup + i])
    torch.cuda.synchronize()
    graphed = n * args.steps / (time.perf_counter() - t0)

    log("[bench] kernel timing")
    # the step's one launch: pose + render (pbr_render_posed)
    kus = kernel_time_us(r, state=states[0])
    nbytes = algorithmic_bytes(r) + pose_bytes(step_poses(r, states[0]))
    achieved = nbytes / (kus * 1e-6) / 1e9
    render_us = kernel_time_us(r)
    render_bytes = algorithmic_bytes(r)

    rec = {
        "metric": "scene-frames/s, 4096-env CartPole 64x64, 1 GPU; fwd+bwd vs plain fwd; HBM %",
        "value": round(value, 1),
        "unit": "scene-frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(el * 1000.0 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {"workload": "CartPole cart+pole+ground quad, 64x64 uint8 RGB, renderer.step",
                   "num_scenes": n, "tile": [64, 64], "channels": 3,
                   "tris_per_scene": 26, "parallelism": "none (1 GPU)"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(render_kernel_name(r)),
                     "kernel": KERNEL_LABEL[render_kernel_name(r)] + ", posing the step's nodes",
                     "kernel_us": round(kus, 3), "bytes_per_launch": nbytes,
                     "render_only": {"kernel_us": round(render_us, 3),
                                     "bytes_per_launch": render_bytes,
                                     "frac": round(render_bytes / render_us / 1e3 / HBM_PEAK_GBS,
                                                   4)}},
    }
    rec["graphed_step"] = {"scene_frames_per_s": round(graphed, 1),
                           "note": "PBRRenderer.capture_step replay incl. state copy"}
    if not args.no_bwd:
        log("[bench] fwd+bwd (differentiable render)")
        rec["fwd_bwd"] = fwd_bwd_leg(n, states[0])
    if not args.no_bwd:
        log("[bench] fwd+bwd (differentiable-render workload)")
        rec["fwd_bwd_differentiable"] = differentiable_leg()
    if not args.no_loop:
        log("[bench] on-device env loop")
        rec["env_loop"] = env_loop_leg(args.loop_scenes, args.loop_steps)
    if not args.no_tiles:
        log("[bench] many-cubes and uneven persistent batches")
        rec["tile_workloads"] = tile_workloads_leg()
    if not args.no_env:
        log("[bench] env.step CartPole-v0")
        es = env_step_rate(n, 50)
        rec["env_step"] = {"scene_frames_per_s": round(es, 1),
                           "workload": "pbr.envs.make('CartPole-v0') env.step, reference scene",
                           "published_l4": PUBLISHED_ENV_STEP_L4,
                           "ratio_to_published": round(es / PUBLISHED_ENV_STEP_L4, 3)}
    if not args.no_cpu:
        log("[bench] cpu baseline (oracle)")
        rec["cpu_baseline"] = cpu_baseline(r, min_seconds=args.cpu_seconds)
    print(json.dumps(rec), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
