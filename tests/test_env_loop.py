"""Fused CartPole env step (pbr_cartpole_step) and CartPoleLoop vs the numpy
oracle (oracle/cartpole_ref.py)."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

from oracle import cartpole_ref as ref
from oracle import oracle
from pybatchrender import _abi, _native
from pybatchrender.envs.cartpole import CartPoleConfig, CartPoleGroundRenderer
from pybatchrender.envs.cartpole.loop import CartPoleLoop, cartpole_params


def _params(**kw) -> dict:
    cfg = CartPoleConfig(num_scenes=4, device="cpu", **kw)
    return ref.params_dict(cartpole_params(cfg))


# --- CPU: the oracle ---------------------------------------------------------

def test_splitmix64_published_outputs():
    """SplitMix64 from state 0: the generator's first three outputs."""
    state = np.array([0], dtype=np.uint64)
    outs = []
    for _ in range(3):
        outs.append(int(ref.splitmix64(state)[0]))
        with np.errstate(over="ignore"):
            state = state + ref.GOLDEN
    assert outs == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_uniform_draws_cover_the_unit_interval():
    key = ref.scene_key(7, np.arange(1))
    u = ref.uniform01(np.repeat(key, 200000), np.arange(200000))
    assert u.min() >= 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 0.005 and abs(u.var() - 1 / 12) < 0.002


def test_dynamics_match_float64_equations():
    p = _params()
    rng = np.random.default_rng(0)
    s = np.stack([rng.uniform(-2, 2, 1000), rng.uniform(-1, 1, 1000),
                  rng.uniform(-0.5, 0.5, 1000), rng.uniform(-0.3, 0.3, 1000)], 1).astype(np.float32)
    a = rng.integers(0, 2, 1000)
    got = ref.dynamics(p, s, a).astype(np.float64)
    want = ref.dynamics_f64(p, s, a)
    assert np.abs(got - want).max() < 2e-6


def test_dynamics_closed_form_at_rest():
    """theta = theta_dot = 0: x_acc = F/m - m_p l a_th / m, a_th = -F/m / (l (4/3 - m_p/m))."""
    p = _params()
    s = np.zeros((2, 4), dtype=np.float32)
    out = ref.dynamics(p, s, np.array([1, 0]))
    m = 1.1
    a_th = -(10.0 / m) / (0.5 * (4.0 / 3.0 - 0.1 / m))
    a_x = 10.0 / m - 0.05 * a_th / m
    np.testing.assert_allclose(out[0], [0.0, 0.02 * a_x, 0.0, 0.02 * a_th], rtol=1e-6)
    np.testing.assert_allclose(out[1], [0.0, -0.02 * a_x, 0.0, -0.02 * a_th], rtol=1e-6)


def test_cartpole_entry_points_validate_without_gpu():
    lib = _native.library()
    p = _abi.pbr_cartpole_params()
    b = _abi.pbr_cartpole_batch()
    b.num_scenes = 4
    assert lib.pbr_cartpole_step(ctypes.byref(p), ctypes.byref(b), None, 0, None) == \
        _abi.PBR_ERR_INVALID
    assert "null state" in _native.last_error()
    b.state = b.step_count = b.rng_counter = 64
    b.next_raw = 72  # misaligned
    assert lib.pbr_cartpole_reset(ctypes.byref(p), ctypes.byref(b), None) == _abi.PBR_ERR_INVALID
    assert "aligned" in _native.last_error()
    b.next_raw = 128
    pose = _abi.pbr_pose_desc()
    pose.rows, pose.transforms = 8, 256
    assert lib.pbr_cartpole_step(ctypes.byref(p), ctypes.byref(b), ctypes.byref(pose), 1, None) \
        == _abi.PBR_ERR_INVALID
    assert "exceed" in _native.last_error()


# --- GPU ------------------------------------------------------------------------

def _loop(n: int, graph: bool, actions: str = "random", seed: int = 3, **cfgkw):
    cfg = CartPoleConfig(num_scenes=n, device="cuda", theta_threshold_deg=12.0, **cfgkw)
    r = CartPoleGroundRenderer(cfg)
    return CartPoleLoop(r, cfg, actions=actions, graph=graph, seed=seed)


def _host(t: torch.Tensor) -> np.ndarray:
    return t.detach().cpu().numpy()


@pytest.mark.gpu
@pytest.mark.parametrize("actions", ["random", "given"])
def test_fused_step_matches_oracle(cuda, actions):
    """Each step of the device loop against the oracle stepped from the same
    state: states within float32 rounding, draws and resets exact."""
    lp = _loop(512, graph=False, actions=actions, max_steps=40)
    p = ref.params_dict(lp.params)
    r0 = ref.reset(p, np.zeros(512, dtype=np.uint64))
    np.testing.assert_array_equal(_host(lp.state), r0["state"])
    g = torch.Generator("cpu").manual_seed(0)
    n_done = 0
    for k in range(60):
        st, sc = _host(lp.state).copy(), _host(lp.step_count).copy()
        ctr = _host(lp.rng_counter).astype(np.uint64)
        act = None
        if actions == "given":
            act = torch.randint(0, 2, (512,), generator=g)
            out = lp.step(act.cuda())
            act = act.numpy()
        else:
            out = lp.step()
        torch.cuda.synchronize()
        want = ref.step(p, st, sc, ctr, act)
        raw = _host(lp.next_raw)
        np.testing.assert_allclose(raw, want["next_raw"], rtol=1e-5, atol=1e-6)
        # termination decided on values the oracle agrees with (away from the thresholds)
        margin = (np.abs(np.abs(want["next_raw"][:, 0]) - p["x_threshold"]) > 1e-4) & \
                 (np.abs(np.abs(want["next_raw"][:, 2]) - p["theta_threshold"]) > 1e-4)
        d = _host(out["done"]).reshape(-1)
        assert (d[margin] == want["done"][margin]).all(), k
        ok = margin & (d == want["done"])
        np.testing.assert_array_equal(_host(out["step_count"])[ok], want["step_count"][ok])
        reset_rows = ok & want["done"]
        np.testing.assert_array_equal(_host(lp.state)[reset_rows], want["state"][reset_rows])
        np.testing.assert_array_equal(_host(lp.rng_counter).astype(np.uint64), want["counter"])
        n_done += int(want["done"].sum())
    assert n_done > 50  # resets were exercised (max_steps 40, 12 degree limit)
    assert (_host(lp.reward) == 1.0).all()


@pytest.mark.gpu
def test_fused_step_frame_matches_oracle_render(cuda):
    """The pixels of a step are the frame of the pre-reset state: bit-exact
    against the CPU oracle rendering that state."""
    lp = _loop(256, graph=True)
    for _ in range(7):
        out = lp.step()
    torch.cuda.synchronize()
    pix = out["pixels"].cpu()
    raw = lp.next_raw.clone()
    lp.r._step(raw)
    desc, keep = lp.r.scene_descriptor()
    want = oracle.render(desc, keep, channels=3, dtype=lp.r.cfg.output_dtype)
    assert torch.equal(pix, want["pixels"])


@pytest.mark.gpu
def test_graph_replays_equal_eager_steps(cuda):
    a = _loop(1024, graph=True, output_dtype="bfloat16", output_depth=True)
    b = _loop(1024, graph=False, output_dtype="bfloat16", output_depth=True)
    for k in range(25):
        oa, ob = a.step(), b.step()
        torch.cuda.synchronize()
        for key in ("observation", "done", "step_count", "pixels", "depth"):
            assert torch.equal(oa[key], ob[key]), (k, key)


@pytest.mark.gpu
def test_random_actions_are_fair_coins(cuda):
    lp = _loop(4096, graph=True, max_steps=1000)
    x0 = lp.state.clone()
    lp.step()
    torch.cuda.synchronize()
    # x_dot moves by +-tau * x_acc; the sign of the change gives the action
    dv = (lp.next_raw[:, 1] - x0[:, 1]).cpu()
    frac = (dv > 0).float().mean().item()
    assert 0.47 < frac < 0.53
