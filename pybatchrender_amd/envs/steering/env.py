"""SteeringEnv (pybatchrender/envs/steering/env.py:18-456).

Same state, dynamics, collision rule, rewards, termination and (partial)
auto-reset as the reference, all as device tensor ops with no host
synchronisation.  The one structural change: the reference walks the
obstacles a player passed in a data-dependent ``while`` loop with a host
``.any()`` per iteration (:239-287); obstacles sit at increasing y, so the
obstacles passed in a step are exactly those with index >= next_obstacle_idx
and y <= player_y, and they are tested in one masked pass.
"""
from __future__ import annotations

import random

import torch

from ..._rl import Bounded, Categorical, Composite, TensorDict, Unbounded
from ...env import PBREnv
from .config import SteeringConfig


class SteeringEnv(PBREnv):
    def __init__(self, renderer, cfg: SteeringConfig | dict | None = None, **cfg_overrides):
        if cfg is None:
            cfg = renderer.cfg
        super().__init__(renderer=renderer, cfg=cfg, device=torch.device(cfg.device),
                         batch_size=torch.Size([cfg.num_scenes]))
        self.speed_initial = float(cfg.speed_initial)
        self.minimal_speed = float(cfg.minimal_speed)
        self.speed_increment = float(cfg.speed_increment)
        self.speed_decrement = float(cfg.speed_decrement)
        self.grace_period = float(cfg.grace_period)
        self.steering_speed = float(cfg.steering_speed)
        self.tau = float(cfg.tau)
        self.lane_width = float(cfg.lane_width)
        self.player_half_width = float(cfg.player_dimensions[0]) * 0.5
        self.obstacle_half_width = float(cfg.obstacle_dimensions[0]) * 0.5
        self.collision_threshold = self.player_half_width + self.obstacle_half_width
        self.rail_half_width = float(cfg.rail_dimensions[0]) * 0.5
        limit = self.lane_width * 0.5 - self.player_half_width - self.rail_half_width
        self.x_min, self.x_max = -limit, limit
        self.obstacle_seed = cfg.obstacle_seed
        self.prob_gold = float(cfg.prob_gold)
        self.prob_cone = float(cfg.prob_cone)
        self.distance_to_first_obstacle = float(cfg.distance_to_first_obstacle)
        self.number_of_obstacles = int(cfg.number_of_obstacles)
        self.obstacle_y_spacing = float(cfg.obstacle_y_spacing)
        self.obstacle_x_positions = list(cfg.obstacle_x_position)
        self.max_steps = int(cfg.max_steps)
        self.auto_reset = bool(cfg.auto_reset)
        self.wheel_sensitivity = float(cfg.wheel_sensitivity)
        self.collision_penalty = float(cfg.collision_penalty)
        self.terminate_on_collision = bool(cfg.terminate_on_collision)
        self.terminate_on_track_end = bool(cfg.terminate_on_track_end)
        self.action_low = float(cfg.action_low)
        self.action_high = float(cfg.action_high)
        self.seed_val = int(cfg.seed) if cfg.seed is not None else 0
        self.render_enabled = bool(cfg.render)
        self._obstacles = None
        self._player_x = self._player_y = self._forward_speed = None
        self._grace_remaining = self._next_obstacle_idx = None
        self._prev_player_x = self._prev_player_y = None
        self._setup_specs(int(cfg.direct_obs_dim), int(cfg.action_n),
                          cfg.action_type == "discrete")
        if self.seed_val is not None:
            self.set_seed(self.seed_val)

    def _setup_specs(self, obs_dim: int, action_n: int, discrete: bool) -> None:
        """env.py:111-150."""
        bs = self.batch_size if self.batch_size != torch.Size([]) else torch.Size([1])
        fields = {}
        if self.render_enabled:
            C = int(self.cfg.num_channels)
            W, H = int(self.cfg.tile_resolution[0]), int(self.cfg.tile_resolution[1])
            fields["pixels"] = Unbounded(shape=bs + torch.Size([C, H, W]), dtype=torch.uint8,
                                         device=self.device)
        fields["observation"] = Unbounded(shape=bs + torch.Size([obs_dim]), dtype=torch.float32,
                                          device=self.device)
        self.observation_spec = Composite(**fields, shape=bs)
        if discrete:
            self.action_spec = Categorical(n=action_n, shape=bs, dtype=torch.long,
                                           device=self.device)
        else:
            self.action_spec = Bounded(low=self.action_low, high=self.action_high,
                                       shape=bs + torch.Size([action_n]), dtype=torch.float32,
                                       device=self.device)
        self.reward_spec = Unbounded(shape=bs + torch.Size([1]), dtype=torch.float32,
                                     device=self.device)
        self.done_spec = Unbounded(shape=bs + torch.Size([1]), dtype=torch.bool, device=self.device)

    def _generate_obstacles(self) -> torch.Tensor:
        """[B, N, 4] (x, y, gold, cone) (env.py:152-183), same draws in the same order."""
        B, N = int(self.cfg.num_scenes), self.number_of_obstacles
        obstacles = torch.zeros(B, N, 4, dtype=torch.float32, device=self.device)
        x_choices = torch.tensor(self.obstacle_x_positions, dtype=torch.float32, device=self.device)
        x_idx = torch.randint(0, len(x_choices), (B, N), device=self.device)
        obstacles[:, :, 0] = x_choices[x_idx]
        y_offsets = torch.arange(N, dtype=torch.float32, device=self.device)
        obstacles[:, :, 1] = y_offsets * self.obstacle_y_spacing + self.distance_to_first_obstacle
        obstacles[:, :, 2] = (torch.rand(B, N, device=self.device) < self.prob_gold).float()
        obstacles[:, :, 3] = (torch.rand(B, N, device=self.device) < self.prob_cone).float()
        return obstacles

    def _get_observation(self) -> torch.Tensor:
        """[player_x, player_y, speed, grace_active, next_obs_x, next_obs_y] (env.py:185-221)."""
        B = int(self.cfg.num_scenes)
        next_x = torch.zeros(B, device=self.device)
        next_y = torch.zeros(B, device=self.device)
        if self._obstacles is not None:
            max_obs = self._obstacles.shape[1]
            valid = self._next_obstacle_idx < max_obs
            idx = torch.clamp(self._next_obstacle_idx, max=max_obs - 1)
            nxt = self._obstacles[torch.arange(B, device=self.device), idx]
            next_x = torch.where(valid, nxt[:, 0], torch.zeros_like(next_x))
            next_y = torch.where(valid, nxt[:, 1], self._player_y + 1000.0)
        grace = (self._grace_remaining > 0).float()
        return torch.stack([self._player_x, self._player_y, self._forward_speed, grace,
                            next_x, next_y], dim=-1)

    def _check_collisions(self) -> tuple[torch.Tensor, torch.Tensor]:
        """Collisions and passages of this step (env.py:223-287) in one masked
        pass: passed obstacles are k >= next_idx with y_k <= player_y."""
        B = int(self.cfg.num_scenes)
        if self._obstacles is None or self._obstacles.numel() == 0:
            z = torch.zeros(B, dtype=torch.int32, device=self.device)
            return z, z.clone()
        N = self._obstacles.shape[1]
        k = torch.arange(N, device=self.device).unsqueeze(0)
        oy = self._obstacles[:, :, 1]
        ox = self._obstacles[:, :, 0]
        passed = (k >= self._next_obstacle_idx.unsqueeze(1)) & (oy <= self._player_y.unsqueeze(1))
        prev_y, curr_y = self._prev_player_y.unsqueeze(1), self._player_y.unsqueeze(1)
        denom = curr_y - prev_y
        denom = torch.where(denom.abs() < 1e-6, torch.ones_like(denom), denom)
        t = (oy - prev_y) / denom
        prev_x, curr_x = self._prev_player_x.unsqueeze(1), self._player_x.unsqueeze(1)
        x_at = prev_x + (curr_x - prev_x) * t
        hit = passed & ((x_at - ox).abs() <= self.collision_threshold)
        passages = passed.sum(1).to(torch.int32)
        collisions = hit.sum(1).to(torch.int32)
        self._next_obstacle_idx = self._next_obstacle_idx + passages.to(torch.long)
        return collisions, passages

    def _dynamics(self, action: torch.Tensor):
        """One step of player physics (env.py:289-337)."""
        self._prev_player_x = self._player_x.clone()
        self._prev_player_y = self._player_y.clone()
        self._grace_remaining = torch.clamp(self._grace_remaining - self.tau, min=0.0)
        self._player_y = self._player_y + self._forward_speed * self.tau
        steering = action.float() if action.dim() == 1 else action.squeeze(-1).float()
        lateral = steering * self._forward_speed / self.wheel_sensitivity * self.tau
        self._player_x = torch.clamp(self._player_x + lateral, self.x_min, self.x_max)
        collisions, passages = self._check_collisions()
        crashes = (collisions > 0) & (self._grace_remaining <= 0)
        self._grace_remaining = torch.where(
            crashes, torch.full_like(self._grace_remaining, self.grace_period),
            self._grace_remaining)
        self._forward_speed = self._forward_speed + passages.float() * self.speed_increment
        self._forward_speed = self._forward_speed - crashes.float() * self.speed_decrement
        self._forward_speed = torch.clamp(self._forward_speed, min=self.minimal_speed)
        reached_end = self._next_obstacle_idx >= self.number_of_obstacles
        return crashes, reached_end

    def _termination(self, crashes, reached_end, step_count) -> torch.Tensor:
        done = torch.zeros_like(step_count, dtype=torch.bool)
        if self.terminate_on_collision:
            done = done | (crashes > 0)
        if self.terminate_on_track_end:
            done = done | reached_end
        done = done | (step_count >= self.max_steps - 1)
        return done.unsqueeze(-1)

    def _reward(self, crashes, done) -> torch.Tensor:
        return (crashes.float() * self.collision_penalty).unsqueeze(-1)

    def _set_seed(self, seed: int) -> None:
        torch.manual_seed(int(seed))
        random.seed(int(seed))

    def _reset(self, tensordict: TensorDict | None = None) -> TensorDict:
        """env.py:381-414."""
        bs = self.batch_size if self.batch_size != torch.Size([]) else torch.Size([1])
        B = int(bs[0]) if len(bs) > 0 else 1
        self._obstacles = self._generate_obstacles()
        self._renderer.build_obstacles(self._obstacles)
        self._player_x = torch.empty(B, device=self.device).uniform_(self.x_min, self.x_max)
        self._player_y = torch.zeros(B, device=self.device)
        self._prev_player_x = torch.zeros(B, device=self.device)
        self._prev_player_y = torch.zeros(B, device=self.device)
        self._forward_speed = torch.full((B,), self.speed_initial, device=self.device)
        self._grace_remaining = torch.zeros(B, device=self.device)
        self._next_obstacle_idx = torch.zeros(B, dtype=torch.long, device=self.device)
        fields = {"observation": self._get_observation(),
                  "step_count": torch.zeros(*bs, dtype=torch.long, device=self.device),
                  "done": torch.zeros(*bs, 1, dtype=torch.bool, device=self.device)}
        if self.render_enabled:
            fields["pixels"] = self.render_pixels(fields["observation"])
        return TensorDict(fields, batch_size=self.batch_size)

    @torch.no_grad()
    def _step(self, tensordict: TensorDict) -> TensorDict:
        """env.py:416-456.  Auto-reset keeps the reference's behaviour: done
        envs restart their step count but keep their state (the reference's
        partial reset is a TODO there).  A render error propagates (the
        reference drops the pixels key and carries on)."""
        obs = tensordict.get("observation", None)
        if obs is None:
            td0 = self._reset()
            step_count = td0["step_count"]
        else:
            step_count = tensordict.get("step_count", None)
            if step_count is None:
                step_count = torch.zeros_like(obs[..., 0], dtype=torch.long, device=self.device)
        action = tensordict["action"].to(self.device)
        crashes, reached_end = self._dynamics(action)
        next_obs = self._get_observation()
        done = self._termination(crashes, reached_end, step_count)
        reward = self._reward(crashes, done)
        pixels = self.render_pixels(next_obs) if self.render_enabled else None
        if self.auto_reset:
            next_step_count = torch.where(done.squeeze(-1), torch.zeros_like(step_count),
                                          step_count + 1)
        else:
            next_step_count = step_count + 1
        out = {"observation": next_obs, "reward": reward, "done": done,
               "step_count": next_step_count}
        if pixels is not None:
            out["pixels"] = pixels
        return TensorDict(out, batch_size=self.batch_size)
