"""Scene-by-scene restatement of the Steering env step (plain Python floats).

TEST INFRASTRUCTURE ONLY: the checker of envs/steering/env.py's vectorised
step.  Follows pybatchrender/envs/steering/env.py:289-337 (_dynamics) and the
obstacle loop of _check_collisions (:239-287) literally: while the next
obstacle's y is at or behind the player, interpolate the player's x at that
y between the previous and current positions, count a hit within the
collision threshold, advance.  Float64 arithmetic, one scene at a time.

Pinned by closed forms (tests/test_steering_pins.py): passages =
floor((y - d0) / spacing) + 1 when driving straight, speed + increment per
passage, crashes exactly one grace period apart on an obstacle line, the
hit decided at the obstacle's y (a diagonal move through the obstacle hits
with both end points outside the threshold and misses with its end point
inside it), x clamped to the road and speed to its minimum.
"""
from __future__ import annotations


def step_scene(s: dict, obstacles: list, action: float, p: dict) -> dict:
    """One step of one scene.  ``s``: x, y, speed, grace, next; obstacles:
    [(x, y, gold, cone), ...]; returns the new state plus crash / end flags."""
    s = dict(s)
    prev_x, prev_y = s["x"], s["y"]
    s["grace"] = max(s["grace"] - p["tau"], 0.0)
    s["y"] = s["y"] + s["speed"] * p["tau"]
    lateral = action * s["speed"] / p["wheel_sensitivity"] * p["tau"]
    s["x"] = min(max(s["x"] + lateral, p["x_min"]), p["x_max"])
    collisions = passages = 0
    while s["next"] < len(obstacles) and obstacles[s["next"]][1] <= s["y"]:
        ox, oy = obstacles[s["next"]][0], obstacles[s["next"]][1]
        s["next"] += 1
        passages += 1
        denom = s["y"] - prev_y
        if abs(denom) < 1e-6:
            denom = 1.0
        t = (oy - prev_y) / denom
        x_at = prev_x + (s["x"] - prev_x) * t
        if abs(x_at - ox) <= p["collision_threshold"]:
            collisions += 1
    crash = collisions > 0 and s["grace"] <= 0
    if crash:
        s["grace"] = p["grace_period"]
    s["speed"] = s["speed"] + passages * p["speed_increment"]
    s["speed"] = s["speed"] - (p["speed_decrement"] if crash else 0.0)
    s["speed"] = max(s["speed"], p["minimal_speed"])
    s["crash"] = crash
    s["end"] = s["next"] >= len(obstacles)
    return s
