"""Closed-form pins of the Steering step (reference envs/steering/env.py:234-337).

Each scenario is built from exactly representable numbers (tau = 1/16 s,
speed 16 units/s: one unit of track per step), so its outcome follows from
the game's rules by hand, independent of any restatement of the step:

  * obstacles k sit at y = d0 + k * spacing (env.py:182-184): after driving
    straight to y the player has passed floor((y - d0) / spacing) + 1 of them;
  * a player driving on an obstacle's line hits every obstacle; a hit is a
    crash only once the grace period has run out, so with grace 1 s crashes
    come every 16 steps;
  * the hit test uses x interpolated at the obstacle's y between the
    previous and new positions (env.py:268-281): a diagonal move can miss
    although its end point is within the threshold, and hit although both
    end points are outside it;
  * x is clamped to the road; speed never drops below the minimum.

Every scenario runs through the oracle's scene loop (oracle/steering_ref.py)
and through the vectorised env step (SteeringEnv._dynamics) on the CPU.
"""
from __future__ import annotations

import math

import pytest
import torch

from oracle import steering_ref
from pybatchrender.envs.steering import SteeringConfig, SteeringEnv, SteeringRenderer

TAU = 1.0 / 16.0


def _env(**cfg) -> SteeringEnv:
    cfg = {"minimal_speed": 1.0, **cfg}  # below the scenarios' 16 units/s
    c = SteeringConfig(num_scenes=1, device="cpu", render=False, tau=TAU, **cfg)
    env = SteeringEnv(renderer=SteeringRenderer(c), cfg=c)
    env.reset()
    return env


def _params(env) -> dict:
    return {k: getattr(env, k) for k in ("tau", "wheel_sensitivity", "x_min", "x_max",
                                          "collision_threshold", "grace_period",
                                          "speed_increment", "speed_decrement",
                                          "minimal_speed")}


class Both:
    """One scene stepped by the oracle and by the env from the same state."""

    def __init__(self, env, obstacles, x, speed, y=0.0):
        self.env, self.p = env, _params(env)
        self.obs = obstacles
        n = len(obstacles)
        env._obstacles = torch.zeros(1, n, 4)
        for k, (ox, oy) in enumerate(obstacles):
            env._obstacles[0, k, 0], env._obstacles[0, k, 1] = ox, oy
        env.number_of_obstacles = n
        env._player_x = torch.tensor([x])
        env._player_y = torch.tensor([y])
        env._forward_speed = torch.tensor([speed])
        env._grace_remaining = torch.zeros(1)
        env._next_obstacle_idx = torch.zeros(1, dtype=torch.long)
        self.s = {"x": x, "y": y, "speed": speed, "grace": 0.0, "next": 0}
        self.ref_obs = [(ox, oy, 0.0, 0.0) for ox, oy in obstacles]

    def step(self, action: float = 0.0) -> dict:
        self.s = steering_ref.step_scene(self.s, self.ref_obs, action, self.p)
        crashes, _ = self.env._dynamics(torch.tensor([[action]]))
        e = self.env
        got = {"x": e._player_x.item(), "y": e._player_y.item(), "speed": e._forward_speed.item(),
               "next": int(e._next_obstacle_idx.item()), "crash": bool(crashes.item())}
        return {"ref": self.s, "env": got}


def _agree(r: dict) -> dict:
    ref, env = r["ref"], r["env"]
    for k in ("x", "y", "speed"):
        assert env[k] == pytest.approx(ref[k], abs=1e-5), k
    assert env["next"] == ref["next"] and env["crash"] == ref["crash"]
    return ref


def test_passages_follow_the_obstacle_spacing():
    env = _env(speed_increment=0.0)
    d0, sp = 5.5, 3.0
    obstacles = [(11.0, d0 + k * sp) for k in range(30)]
    b = Both(env, obstacles, x=-5.0, speed=16.0)  # far from the x = 11 line
    for k in range(1, 80):
        s = _agree(b.step())
        y = float(k)  # one unit per step, exact
        assert s["y"] == y
        want = 0 if y < d0 else min(30, math.floor((y - d0) / sp) + 1)
        assert s["next"] == want, k
        assert not s["crash"]


def test_speed_grows_by_the_increment_per_passage():
    env = _env(speed_increment=0.25)
    obstacles = [(11.0, 2.0 + 2.0 * k) for k in range(40)]
    b = Both(env, obstacles, x=0.0, speed=16.0)
    for _ in range(60):
        s = _agree(b.step())
        assert s["speed"] == 16.0 + 0.25 * s["next"]


def test_crashes_every_grace_period_on_an_obstacle_line():
    env = _env(speed_increment=0.0, speed_decrement=0.0, grace_period=1.0)
    obstacles = [(1.0, 4.0 * (k + 1)) for k in range(25)]  # passed at steps 4, 8, 12, ...
    b = Both(env, obstacles, x=1.0, speed=16.0)
    crashes = []
    for k in range(1, 100):
        s = _agree(b.step())
        if s["crash"]:
            crashes.append(k)
    # grace 1 s = 16 steps: a crash at step 4, then the first passage 16
    # steps later, ...; the passages in between hit inside the grace period
    assert crashes == [4, 20, 36, 52, 68, 84]


def test_an_obstacle_line_off_the_threshold_never_crashes():
    env = _env(speed_increment=0.0, grace_period=1.0)
    thr = env.collision_threshold
    obstacles = [(1.0, 4.0 * (k + 1)) for k in range(25)]
    b = Both(env, obstacles, x=1.0 + thr + 0.5, speed=16.0)
    assert not any(_agree(b.step())["crash"] for _ in range(99))


@pytest.mark.parametrize("x0,x1,hit", [
    (-5.0, 7.0, True),    # both ends 6 away from x = 1, crossing at x = 1
    (9.0, 3.0, False),    # ends at 3 (within 2 of x = 1), crossing at 6
])
def test_hit_is_decided_at_the_obstacle_y(x0, x1, hit):
    env = _env(speed_increment=0.0, speed_decrement=0.0)
    # y goes 3 -> 4 in step 4; the obstacle at y = 3.5 is crossed mid-step
    b = Both(env, [(1.0, 3.5), (1.0, 500.0)], x=x0, speed=16.0)
    for _ in range(3):
        assert not _agree(b.step())["crash"]
    # lateral move = action * speed / wheel_sensitivity * tau
    action = (x1 - x0) * env.wheel_sensitivity / (16.0 * TAU)
    s = _agree(b.step(action))
    assert s["x"] == x1 and s["next"] == 1
    assert s["crash"] is hit


def test_x_is_clamped_to_the_road_and_speed_to_its_minimum():
    env = _env(speed_decrement=100.0, grace_period=0.0625, minimal_speed=18.0)
    obstacles = [(env.x_max, 1.0 + k) for k in range(10)]
    b = Both(env, obstacles, x=0.0, speed=16.0)
    s = _agree(b.step(1e6))  # a huge right turn
    assert s["x"] == env.x_max
    s = _agree(b.step(1e6))
    assert s["crash"] and s["speed"] == env.minimal_speed
