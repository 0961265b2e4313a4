"""The render calls the benchmark times, each bit-exact against the oracle.

The row-span kernel is compiled per output format: uint8 without a
visibility output keeps packed colours in its row lists (the 4096-scene
CartPole headline), while uint8 with visibility, bfloat16 and float32 keep
record references and resolve colours from the record table.  Every
(dtype, depth, visibility) combination is rendered here on the flat CartPole
ground scene and compared with the oracle bitwise, including exactly the
calls the benchmark reports:
  * renderer.step at 4096 scenes, uint8, no depth, no visibility (eager and
    graph-captured);
  * the float32 plain forward (no visibility) and the forward with the
    visibility buffer behind the fwd+bwd ratio, at 4096 scenes;
  * the 16384-scene bf16 RGB + float32 depth frame of the on-device env loop;
  * the frames CartPole-v0's env.step returns (reference scene).
"""
from __future__ import annotations

import pytest
import torch

import pybatchrender as pbr
from oracle import oracle
from pybatchrender.envs.cartpole import CartPoleConfig, CartPoleGroundRenderer
from pybatchrender.envs.cartpole.loop import CartPoleLoop
from tests.test_gpu_parity import _bits, check_parity, ground_states

pytestmark = pytest.mark.gpu


def _ground(n: int, **cfg) -> CartPoleGroundRenderer:
    r = CartPoleGroundRenderer(CartPoleConfig(num_scenes=n, device="cuda", kernel="rows", **cfg))
    assert r.cfg.kernel == "rows"
    return r


@pytest.mark.parametrize("vis", [False, True])
@pytest.mark.parametrize("depth", [False, True])
@pytest.mark.parametrize("dtype", ["uint8", "bfloat16", "float32"])
def test_rows_kernel_flat_variants(cuda, dtype, depth, vis):
    r = _ground(512, output_dtype=dtype)
    r._step(ground_states(512, 17).cuda())
    check_parity(r, f"rows flat {dtype} depth={depth} vis={vis}", depth=depth, vis=vis)


@pytest.mark.parametrize("dtype,channels", [("uint8", 1), ("uint8", 4), ("bfloat16", 1),
                                            ("bfloat16", 4), ("float32", 1), ("float32", 2)])
def test_rows_kernel_flat_channel_counts(cuda, dtype, channels):
    r = _ground(300, output_dtype=dtype, num_channels=channels)
    r._step(ground_states(300, 18).cuda())
    check_parity(r, f"rows flat {dtype} C={channels}", depth=False, vis=False)
    check_parity(r, f"rows flat {dtype} C={channels} + depth", depth=True, vis=False)


def test_headline_step_4096_uint8(cuda):
    """renderer.step as the benchmark calls it (AUTO kernel choice, uint8,
    no depth, no visibility), three state batches."""
    r = CartPoleGroundRenderer(CartPoleConfig(num_scenes=4096, device="cuda"))
    for seed in range(3):
        pix = r.step(ground_states(4096, seed).cuda())
        check_parity(r, f"headline step seed {seed}", depth=False, vis=False,
                     got={"pixels": pix})


def test_headline_graphed_step_4096_uint8(cuda):
    """The same step captured as a hipGraph: each replay's frame is the
    oracle's frame of that replay's states."""
    r = CartPoleGroundRenderer(CartPoleConfig(num_scenes=4096, device="cuda"))
    states = [ground_states(4096, seed).cuda() for seed in range(3)]
    g = r.capture_step(states[0])
    for k, st in enumerate(states + states[:1]):
        out = g(st)
        check_parity(r, f"graphed headline step {k}", depth=False, vis=False,
                     got={"pixels": out["pixels"].clone()})


@pytest.mark.parametrize("vis", [False, True])
def test_float32_forward_4096(cuda, vis):
    """The plain forward (no visibility) and the forward with the visibility
    buffer that the fwd+bwd ratio divides, at the benchmark's size."""
    r = _ground(4096, output_dtype="float32")
    r._step(ground_states(4096, 0).cuda())
    check_parity(r, f"float32 forward vis={vis}", depth=False, vis=vis)


def test_env_loop_frame_bf16_depth_16384(cuda):
    """The on-device env loop at the benchmark's configuration (16384 scenes,
    bf16 RGB + float32 depth, one graph replay per step): the frame and depth
    of a step are the oracle's render of that step's pre-reset states."""
    import math

    cfg = CartPoleConfig(num_scenes=16384, device="cuda", output_dtype="bfloat16",
                         output_depth=True, theta_threshold_deg=math.degrees(0.21),
                         x_threshold=2.4, max_steps=500)
    lp = CartPoleLoop(CartPoleGroundRenderer(cfg), cfg, actions="random", graph=True, seed=3)
    for k in range(12):
        out = lp.step()
        if k in (0, 11):
            torch.cuda.synchronize()
            pix, dep = out["pixels"].clone(), out["depth"].clone()
            lp.r._step(lp.next_raw)
            ref = oracle.render_renderer(lp.r)
            assert torch.equal(_bits(pix), _bits(ref["pixels"])), f"step {k}: bf16 pixels"
            assert torch.equal(_bits(dep), _bits(ref["depth"])), f"step {k}: depth"
            print(f"[env loop] step {k}: bf16 + depth frame bit-exact", flush=True)


def test_cartpole_env_step_frames(cuda):
    """pbr.envs.make('CartPole-v0').step returns the frame of the pre-reset
    states (reference envs/cartpole/env.py:178-218): every frame of several
    steps (with auto-resets) equals the oracle's render of those states, and
    the reset frame equals the oracle's render of the initial states."""
    env = pbr.envs.make("CartPole-v0", num_scenes=256, device="cuda", max_steps=6)
    td = env.reset()
    torch.cuda.synchronize()
    env._renderer._step(td["observation"])
    ref = oracle.render_renderer(env._renderer)
    assert torch.equal(td["pixels"].cpu(), ref["pixels"]), "reset frame"
    g = torch.Generator("cpu").manual_seed(4)
    resets = 0
    for k in range(10):
        td["action"] = torch.randint(0, 2, (256,), generator=g).cuda()
        obs = td["observation"].clone()
        nxt = env.step(td)["next"]
        torch.cuda.synchronize()
        raw = env._dynamics(obs, td["action"])
        env._renderer._step(raw)
        ref = oracle.render_renderer(env._renderer)
        assert torch.equal(nxt["pixels"].cpu(), ref["pixels"]), f"step {k}"
        resets += int(nxt["done"].sum().item())
        td = nxt
    assert resets > 0  # auto-resets happened between the checked frames
