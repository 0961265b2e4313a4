"""The reference's import surface: ``import pybatchrender as pbr``.

Names and submodule paths of the reference (pybatchrender/__init__.py:1-70,
pybatchrender/renderer/*.py, pybatchrender/envs/*) resolve to this
implementation, as the very same module and class objects."""
from __future__ import annotations

import importlib

import pytest
import torch

REFERENCE_ALL = ["PBRConfig", "PBREnv", "PBRRenderer", "PBRNode", "PBRCam", "PBRLight",
                 "PBRShaderContext", "GPUFrameGrabber", "CPUFrameGrabber", "FrameGrabber", "envs",
                 "__version__"]

SUBMODULES = ["config", "env", "renderer", "renderer.renderer", "renderer.node",
              "renderer.camera", "renderer.light", "renderer.shader_context", "envs",
              "envs.cartpole", "envs.cartpole.env", "envs.cartpole.renderer",
              "envs.cartpole.config", "envs.steering", "envs.steering.env",
              "envs.steering.renderer", "envs.steering.config"]


def test_top_level_names():
    import pybatchrender as pbr

    assert sorted(pbr.__all__) == sorted(REFERENCE_ALL)
    for name in REFERENCE_ALL:
        assert hasattr(pbr, name), name
    for name in ("PBRConfig", "PBREnv", "PBRRenderer", "PBRNode", "PBRCam", "PBRLight",
                 "PBRShaderContext"):
        assert getattr(pbr, name) is not None, name
    # no frame grabber: the kernels write the final tensor
    assert pbr.GPUFrameGrabber is None and pbr.FrameGrabber is None


@pytest.mark.parametrize("sub", SUBMODULES)
def test_submodules_are_the_implementation(sub):
    alias = importlib.import_module(f"pybatchrender.{sub}")
    real = importlib.import_module(f"pybatchrender_amd.{sub}")
    assert alias is real


def test_classes_shared_between_names():
    import pybatchrender as pbr
    from pybatchrender.envs.cartpole.env import CartPoleEnv
    from pybatchrender.renderer.renderer import PBRRenderer

    import pybatchrender_amd

    assert PBRRenderer is pybatchrender_amd.PBRRenderer is pbr.PBRRenderer
    env = pbr.envs.make("CartPole-v0", num_scenes=4, render=False)
    assert isinstance(env, CartPoleEnv)
    assert "CartPole-v0" in pybatchrender_amd.envs.list_envs()


def test_make_cartpole_with_defaults():
    """The reference's own call: make('CartPole-v0', num_scenes=4)."""
    import pybatchrender as pbr

    env = pbr.envs.make("CartPole-v0", num_scenes=4)
    assert env.batch_size == torch.Size([4])
    assert tuple(env.observation_spec["pixels"].shape) == (4, 3, 64, 64)


def test_unknown_submodule_is_an_import_error():
    with pytest.raises(ModuleNotFoundError):
        importlib.import_module("pybatchrender.no_such_module")
