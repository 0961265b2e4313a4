"""CartPoleLoop: the CartPole env step and its frame entirely on the device.

The reference steps the env with torch ops and renders through Panda3D
(envs/cartpole/env.py:178-218 -> renderer.py:98-138 -> PBRRenderer.step).
Here one env step of every scene is three HIP launches with no host round
trip, captured once as a hipGraph and replayed per step:

    pbr_cartpole_step   dynamics, termination, reward, auto-reset
                        (env.py:111-139, 198-206) and the cart / pole
                        instance matrices of the pre-reset state
    setup + raster      the frame of that state (pbr_render)

``step()`` returns the graph's static outputs (overwritten by the next step):
observation [N, 4] (after reset), reward [N, 1], done [N, 1] bool,
step_count [N], pixels [N, C, H, W] (the renderer's output dtype) and, when
asked for, depth [N, H, W].  Actions are either drawn on the device
(Bernoulli(0.5), ``actions="random"``) or copied from the caller's [N] int64
tensor into the graph's static input (``actions="given"``).

Random numbers come from a counter hash per scene (SplitMix64 of the seed,
the scene index and the scene's draw count), not torch's generator, so the
draws differ from the reference's ``torch.manual_seed`` stream; dynamics,
termination and reset rules are the same.
"""
from __future__ import annotations

import math

import torch

from ... import _abi, _native
from .config import CartPoleConfig


def cartpole_params(cfg: CartPoleConfig, seed: int | None = None) -> _abi.pbr_cartpole_params:
    p = _abi.pbr_cartpole_params()
    p.gravity, p.masscart, p.masspole, p.length = (float(cfg.gravity), float(cfg.masscart),
                                                   float(cfg.masspole), float(cfg.length))
    p.force_mag, p.tau = float(cfg.force_mag), float(cfg.tau)
    p.theta_threshold = float(cfg.theta_threshold_deg * 2 * math.pi / 360.0)
    p.x_threshold = float(cfg.x_threshold)
    p.max_steps = int(cfg.max_steps)
    p.auto_reset = 1 if cfg.auto_reset else 0
    th = sorted(math.radians(v) for v in cfg.init_theta_range_deg)
    thd = sorted(math.radians(v) for v in cfg.init_theta_dot_range_deg)
    lo_hi = [tuple(cfg.init_x_range), tuple(cfg.init_x_dot_range), tuple(th), tuple(thd)]
    p.init_lo[:] = [float(a) for a, _ in lo_hi]
    p.init_hi[:] = [float(b) for _, b in lo_hi]
    p.seed = int(cfg.seed if seed is None else seed) & 0xFFFFFFFFFFFFFFFF
    return p


class CartPoleLoop:
    def __init__(self, renderer, cfg: CartPoleConfig | None = None, actions: str = "random",
                 depth: bool | None = None, graph: bool = True, seed: int | None = None):
        if renderer.device.type != "cuda":
            raise RuntimeError("CartPoleLoop runs on the GPU")
        if actions not in ("random", "given"):
            raise ValueError("actions must be 'random' or 'given'")
        self.r = renderer
        self.cfg = cfg if cfg is not None else renderer.cfg
        dev = renderer.device
        N = int(renderer.num_scenes)
        self.N = N
        self.params = cartpole_params(self.cfg, seed)
        self.state = torch.zeros((N, 4), dtype=torch.float32, device=dev)
        self.next_raw = torch.zeros((N, 4), dtype=torch.float32, device=dev)
        self.step_count = torch.zeros(N, dtype=torch.int64, device=dev)
        self.reward = torch.zeros(N, dtype=torch.float32, device=dev)
        self.done = torch.zeros(N, dtype=torch.uint8, device=dev)
        self.rng_counter = torch.zeros(N, dtype=torch.int64, device=dev)  # uint64 bits
        self.action = torch.zeros(N, dtype=torch.int64, device=dev) if actions == "given" else None
        self.batch = _abi.pbr_cartpole_batch()
        b = self.batch
        b.state, b.next_raw = self.state.data_ptr(), self.next_raw.data_ptr()
        b.step_count, b.reward = self.step_count.data_ptr(), self.reward.data_ptr()
        b.done, b.rng_counter = self.done.data_ptr(), self.rng_counter.data_ptr()
        b.action = self.action.data_ptr() if self.action is not None else None
        b.num_scenes = N
        # pose descriptors of the pre-reset state (sources read next_raw in place)
        self._pose_keep = []
        self.poses = []
        for node, kw in renderer.pose_spec(self.next_raw):
            full = {"x": 0.0, "y": 0.0, "z": 0.0, "h": 0.0, "p": 0.0, "r": 0.0, "s": 1.0, **kw}
            prep = node._pose_inputs(**full)
            if prep is None:
                continue
            desc, keep = prep
            if desc.rows != N:
                raise ValueError(f"node {node.name}: {desc.rows} pose rows for {N} scenes")
            self.poses.append(desc)
            # the sources and the output buffers the descriptor points to
            self._pose_keep.append((keep, node.transforms_b44, node.rot3_b33, node.scale_b11))
        if len(self.poses) > 8:
            raise ValueError("at most 8 posed nodes")
        Nn, C, H, W = renderer.output_shape()
        from ...renderer.renderer import _DTYPES
        tdtype, _ = _DTYPES[renderer.cfg.output_dtype]
        self.pixels = torch.empty((Nn, C, H, W), dtype=tdtype, device=dev)
        if depth is None:
            depth = bool(renderer.cfg.output_depth)
        self.depth = torch.empty((N, H, W), dtype=torch.float32, device=dev) if depth else None
        # the loop's own workspace (the graph keeps using it whatever the
        # renderer later renders eagerly) and the descriptor's tensors, held
        # for as long as the loop may replay its graph
        self.workspace = torch.empty(max(renderer.workspace_bytes(), 256), dtype=torch.uint8,
                                     device=dev)
        self.graph = None
        if graph:
            self._capture()
        self._keep = list(renderer._desc_cache[2])
        # episodes (and the random streams) start at the reset
        self.state.zero_()
        self.step_count.zero_()
        self.rng_counter.zero_()
        self.reset()

    # --- launches ----------------------------------------------------------
    def _stream(self) -> int:
        return torch.cuda.current_stream(self.r.device).cuda_stream

    def _body(self) -> None:
        _native.cartpole_step(self.params, self.batch, self.poses, self._stream())
        self.r.render(out=self.pixels, depth=self.depth if self.depth is not None else False,
                      workspace=self.workspace)

    def _capture(self, warmup: int = 2) -> None:
        dev = self.r.device
        side = torch.cuda.Stream(dev)
        side.wait_stream(torch.cuda.current_stream(dev))
        with torch.cuda.stream(side):
            for _ in range(warmup):
                self._body()
        torch.cuda.current_stream(dev).wait_stream(side)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self._body()

    def reset(self) -> dict:
        """Fresh states for all scenes (env.py:153-176) and their frame."""
        _native.cartpole_reset(self.params, self.batch, self._stream())
        self.r._step(self.state)
        self.r.render(out=self.pixels, depth=self.depth if self.depth is not None else False,
                      workspace=self.workspace)
        self.done.zero_()
        return self.outputs()

    def step(self, action: torch.Tensor | None = None) -> dict:
        if action is not None:
            if self.action is None:
                raise ValueError("this loop draws its own actions (actions='random')")
            a = action.reshape(-1)
            if a.data_ptr() != self.action.data_ptr():
                self.action.copy_(a, non_blocking=True)
        elif self.action is not None:
            raise ValueError("actions='given' needs an action tensor")
        if self.graph is not None:
            self.graph.replay()
        else:
            self._body()
        return self.outputs()

    def outputs(self) -> dict:
        out = {"observation": self.state, "reward": self.reward.unsqueeze(-1),
               "done": self.done.bool().unsqueeze(-1), "step_count": self.step_count,
               "pixels": self.pixels}
        if self.depth is not None:
            out["depth"] = self.depth
        return out
