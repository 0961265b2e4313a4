"""Steering env (N4): the vectorised device step against the scene-by-scene
restatement of the reference's loop, the registry entry, and the rendered
frames against the CPU oracle."""
from __future__ import annotations

import pytest
import torch

import pybatchrender as pbr
from oracle import steering_ref
from pybatchrender.envs.steering import SteeringConfig, SteeringEnv, SteeringRenderer


def _params(env) -> dict:
    return {k: getattr(env, k) for k in ("tau", "wheel_sensitivity", "x_min", "x_max",
                                          "collision_threshold", "grace_period",
                                          "speed_increment", "speed_decrement",
                                          "minimal_speed")}


def test_steering_registered():
    assert "Steering-v0" in pbr.envs.list_envs()
    env_cls, r_cls, c_cls = pbr.envs.get_env_classes("Steering-v0")
    assert env_cls is SteeringEnv and r_cls is SteeringRenderer and c_cls is SteeringConfig


def test_steering_step_matches_scene_loop():
    """Big steering inputs, dense obstacles: passages, hits, grace periods."""
    B = 16
    cfg = SteeringConfig(num_scenes=B, device="cpu", render=False, number_of_obstacles=40,
                         obstacle_y_spacing=3.0, distance_to_first_obstacle=2.0, seed=3)
    env = SteeringEnv(renderer=SteeringRenderer(cfg), cfg=cfg)
    td = env.reset()
    obstacles = env._obstacles.tolist()
    p = _params(env)
    scenes = [{"x": env._player_x[b].item(), "y": env._player_y[b].item(),
               "speed": env._forward_speed[b].item(), "grace": 0.0, "next": 0}
              for b in range(B)]
    g = torch.Generator().manual_seed(0)
    crashes_total = 0
    for k in range(120):
        a = (torch.rand(B, 1, generator=g) * 2 - 1) * 327.68
        td["action"] = a
        nxt = env.step(td)["next"]
        for b in range(B):
            scenes[b] = steering_ref.step_scene(scenes[b], obstacles[b], float(a[b, 0]), p)
            o = nxt["observation"][b]
            assert abs(o[0].item() - scenes[b]["x"]) < 1e-3, (k, b)
            assert abs(o[1].item() - scenes[b]["y"]) < 1e-3 * max(1.0, scenes[b]["y"])
            assert abs(o[2].item() - scenes[b]["speed"]) < 1e-3, (k, b)
            assert (o[3].item() > 0) == (scenes[b]["grace"] > 0), (k, b)
            assert nxt["reward"][b, 0].item() == (-1.0 if scenes[b]["crash"] else 0.0), (k, b)
            assert env._next_obstacle_idx[b].item() == scenes[b]["next"]
            crashes_total += scenes[b]["crash"]
        td = nxt
    assert crashes_total > 5


def test_observation_layout_and_specs():
    cfg = SteeringConfig(num_scenes=4, device="cpu", render=False)
    env = SteeringEnv(renderer=SteeringRenderer(cfg), cfg=cfg)
    td = env.reset()
    o = td["observation"]
    assert o.shape == (4, 6)
    assert torch.all(o[:, 1] == 0) and torch.all(o[:, 2] == cfg.speed_initial)
    assert torch.all(o[:, 5] == cfg.distance_to_first_obstacle)  # first obstacle's y
    a = env.action_spec.rand()
    assert a.shape == (4, 1) and a.abs().max() <= 327.68


@pytest.mark.gpu
def test_steering_frames_match_oracle(cuda):
    from oracle import oracle
    env = pbr.envs.make("Steering-v0", num_scenes=24, device="cuda", number_of_obstacles=30,
                        obstacle_y_spacing=4.0, distance_to_first_obstacle=6.0)
    td = env.reset()
    assert td["pixels"].shape == (24, 3, 64, 64)
    for _ in range(15):
        td["action"] = env.action_spec.rand()
        td = env.step(td)["next"]
    torch.cuda.synchronize()
    r = env._renderer
    got = r.render(depth=True, vis=True)
    desc, keep = r.scene_descriptor()
    ref = oracle.render(desc, keep, channels=3, dtype="uint8")
    assert torch.equal(got["vis"].cpu(), ref["vis"])
    assert torch.equal(got["pixels"].cpu(), ref["pixels"])
    assert torch.equal(td["pixels"].cpu(), ref["pixels"])
    # obstacles are visible in front of the player in most scenes
    assert (ref["vis"] >= 0).float().mean() > 0.01
