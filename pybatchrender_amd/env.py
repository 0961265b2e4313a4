"""PBREnv: TorchRL environment base that renders through PBRRenderer.

Mirrors pybatchrender/env.py:18-95 (specs, render_pixels).  The PNG/GIF grid
utilities and the multi-process ParallelEnv factory of the reference are not
part of this package.
"""
from __future__ import annotations

import torch

from ._rl import Categorical, Composite, EnvBase, TensorDict, Unbounded
from .config import PBRConfig


class PBREnv(EnvBase):
    def __init__(self, renderer, cfg: PBRConfig | None = None, device="cpu",
                 batch_size: torch.Size = torch.Size([128]), **kwargs) -> None:
        super().__init__(device=torch.device(device), batch_size=batch_size, **kwargs)
        self._renderer = renderer
        self.cfg = cfg if cfg is not None else renderer.cfg

    def set_default_specs(self, *, direct_obs_dim: int | None = None, actions: int | None = None,
                          with_pixels: bool = False, pixels_only: bool = False,
                          discrete_actions: bool = True) -> None:
        """env.py:33-82: pixels uint8 [N, C, H, W] (+ observation, action, reward, done)."""
        assert with_pixels or not pixels_only, "Cannot set both with_pixels and pixels_only to True."
        bs = self.batch_size if self.batch_size != torch.Size([]) else torch.Size([1])
        fields = {}
        if with_pixels:
            C = int(self.cfg.num_channels)
            W = int(self.cfg.tile_resolution[0])
            H = int(self.cfg.tile_resolution[1])
            fields["pixels"] = Unbounded(shape=bs + torch.Size([C, H, W]), dtype=torch.uint8,
                                         device=self.device)
        if not pixels_only:
            if direct_obs_dim is None:
                raise ValueError("direct_obs_dim must be provided when pixels_only=False.")
            fields["observation"] = Unbounded(shape=bs + torch.Size([int(direct_obs_dim)]),
                                              dtype=torch.float32, device=self.device)
        self.observation_spec = Composite(**fields, shape=bs)
        if actions is None:
            raise ValueError("actions must be provided.")
        if discrete_actions:
            self.action_spec = Categorical(n=int(actions), shape=bs, dtype=torch.long,
                                           device=self.device)
        else:
            self.action_spec = Unbounded(shape=bs + torch.Size([int(actions)]),
                                         dtype=torch.float32, device=self.device)
        self.reward_spec = Unbounded(shape=bs + torch.Size([1]), dtype=torch.float32,
                                     device=self.device)
        self.done_spec = Unbounded(shape=bs + torch.Size([1]), dtype=torch.bool, device=self.device)

    def _step(self, tensordict: TensorDict) -> TensorDict:  # pragma: no cover - abstract
        raise NotImplementedError

    def _reset(self, tensordict: TensorDict | None = None) -> TensorDict:  # pragma: no cover
        raise NotImplementedError

    def render_pixels(self, obs: torch.Tensor | None = None) -> torch.Tensor:
        if self._renderer is None:
            raise RuntimeError(
                "Renderer is not initialized. Construct env with a renderer and pass it to PBREnv.")
        return self._renderer.step(obs)
