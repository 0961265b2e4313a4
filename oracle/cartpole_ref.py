"""CPU restatement of the fused CartPole env step (numpy, float32).

TEST INFRASTRUCTURE ONLY: the checker of pbr_cartpole_step /
pbr_cartpole_reset.  Imported by tests/ only, never by the package.

Follows pybatchrender/envs/cartpole/env.py: _dynamics :111-129 (float32,
same operation order), _termination :131-139, _reward :141-143, auto-reset
:198-206; reset sampling :95-109 with the device's counter-hash draws
(SplitMix64 of the scene key and the scene's draw counter, five draws per
step: action, then x, x_dot, theta, theta_dot).

Pins (tests/test_env_loop.py): SplitMix64 against its published first
outputs from state 0; the float32 dynamics against an independent float64
statement of the cart-pole equations (Barto, Sutton and Anderson 1983, in
the form of env.py:111-129) and its closed form at theta = theta_dot = 0.
"""
from __future__ import annotations

import numpy as np

M64 = np.uint64(0xFFFFFFFFFFFFFFFF)
GOLDEN = np.uint64(0x9E3779B97F4A7C15)


def splitmix64(z: np.ndarray) -> np.ndarray:
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = z + GOLDEN
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def scene_key(seed: int, scenes: np.ndarray) -> np.ndarray:
    return splitmix64(np.uint64(seed) ^ splitmix64(np.asarray(scenes, dtype=np.uint64)))


def uniform01(key: np.ndarray, counter: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        bits = splitmix64(key + np.asarray(counter, dtype=np.uint64)) >> np.uint64(40)
    return bits.astype(np.float32) * np.float32(1.0 / 16777216.0)


def sample_state(p: dict, key, counter) -> np.ndarray:
    cols = []
    for k in range(4):
        u = uniform01(key, np.asarray(counter, dtype=np.uint64) + np.uint64(k))
        lo, hi = np.float32(p["init_lo"][k]), np.float32(p["init_hi"][k])
        cols.append(lo + u * (hi - lo))
    return np.stack(cols, 1).astype(np.float32)


def dynamics(p: dict, s: np.ndarray, action: np.ndarray) -> np.ndarray:
    f32 = np.float32
    total_mass = f32(float(f32(p["masscart"])) + float(f32(p["masspole"])))
    pml = f32(float(f32(p["masspole"])) * float(f32(p["length"])))
    x, xd, th, thd = (s[:, k].astype(np.float32) for k in range(4))
    force = np.where(action == 1, f32(p["force_mag"]), -f32(p["force_mag"])).astype(np.float32)
    c = np.cos(th).astype(np.float32)
    sn = np.sin(th).astype(np.float32)
    temp = (force + pml * (thd * thd) * sn) / total_mass
    thetaacc = (f32(p["gravity"]) * sn - c * temp) / (
        f32(p["length"]) * (f32(4.0 / 3.0) - f32(p["masspole"]) * (c * c) / total_mass))
    xacc = temp - pml * thetaacc * c / total_mass
    tau = f32(p["tau"])
    return np.stack([x + tau * xd, xd + tau * xacc, th + tau * thd, thd + tau * thetaacc],
                    1).astype(np.float32)


def step(p: dict, state, step_count, counter, action=None) -> dict:
    """One fused step of every scene; ``action`` None draws Bernoulli(0.5)."""
    N = state.shape[0]
    key = scene_key(p["seed"], np.arange(N))
    counter = np.asarray(counter, dtype=np.uint64)
    if action is None:
        action = (uniform01(key, counter) < np.float32(0.5)).astype(np.int64)
    raw = dynamics(p, state, np.asarray(action))
    done = ((np.abs(raw[:, 0]) > np.float32(p["x_threshold"]))
            | (np.abs(raw[:, 2]) > np.float32(p["theta_threshold"]))
            | (np.asarray(step_count) >= p["max_steps"] - 1))
    nxt = raw.copy()
    nsc = np.asarray(step_count) + 1
    if p["auto_reset"]:
        fresh = sample_state(p, key, counter + np.uint64(1))
        nxt[done] = fresh[done]
        nsc = np.where(done, 0, nsc)
    return {"next_raw": raw, "state": nxt, "done": done, "step_count": nsc,
            "counter": counter + np.uint64(5), "action": np.asarray(action)}


def reset(p: dict, counter) -> dict:
    N = len(counter)
    key = scene_key(p["seed"], np.arange(N))
    counter = np.asarray(counter, dtype=np.uint64)
    return {"state": sample_state(p, key, counter), "counter": counter + np.uint64(4)}


def dynamics_f64(p: dict, s: np.ndarray, action: np.ndarray) -> np.ndarray:
    """Independent float64 statement of the cart-pole Euler step."""
    g, mc, mp, l = p["gravity"], p["masscart"], p["masspole"], p["length"]
    F = np.where(action == 1, p["force_mag"], -p["force_mag"])
    x, xd, th, thd = (s[:, k].astype(np.float64) for k in range(4))
    m = mc + mp
    # pole angular acceleration and cart acceleration of the rigid-pole model
    a_th = (g * np.sin(th) + np.cos(th) * (-F - mp * l * thd ** 2 * np.sin(th)) / m) / (
        l * (4.0 / 3.0 - mp * np.cos(th) ** 2 / m))
    a_x = (F + mp * l * (thd ** 2 * np.sin(th) - a_th * np.cos(th))) / m
    tau = p["tau"]
    return np.stack([x + tau * xd, xd + tau * a_x, th + tau * thd, thd + tau * a_th], 1)


def params_dict(pp) -> dict:
    """The ctypes pbr_cartpole_params as a dict."""
    return {"gravity": pp.gravity, "masscart": pp.masscart, "masspole": pp.masspole,
            "length": pp.length, "force_mag": pp.force_mag, "tau": pp.tau,
            "theta_threshold": pp.theta_threshold, "x_threshold": pp.x_threshold,
            "max_steps": pp.max_steps, "auto_reset": bool(pp.auto_reset),
            "init_lo": list(pp.init_lo), "init_hi": list(pp.init_hi), "seed": int(pp.seed)}
