"""Environment registry (pybatchrender/envs/__init__.py:38-231).

``make(name, **overrides)`` builds Config(**overrides) -> Renderer(cfg) ->
Env(renderer=..., cfg=...) exactly like the reference.  Unlike Panda3D's
ShowBase, renderers are ordinary objects, so several may live in one process.
Third-party packages register through the same entry-point group,
``pybatchrender.envs``.  ``make_parallel`` (multi-process ParallelEnv) is not
provided.
"""
from __future__ import annotations

import warnings

_ENV_REGISTRY: dict = {}


def register(name: str, env_cls, renderer_cls, config_cls) -> None:
    if name in _ENV_REGISTRY:
        warnings.warn(f"Environment '{name}' is already registered. Overwriting.")
    _ENV_REGISTRY[name] = (env_cls, renderer_cls, config_cls)


def unregister(name: str) -> None:
    _ENV_REGISTRY.pop(name, None)


def list_envs() -> list[str]:
    return sorted(_ENV_REGISTRY.keys())


def is_registered(name: str) -> bool:
    return name in _ENV_REGISTRY


def get_env_classes(name: str):
    if name not in _ENV_REGISTRY:
        raise ValueError(f"Unknown environment: '{name}'. Available: {list_envs()}")
    return _ENV_REGISTRY[name]


def make(name: str, **config_overrides):
    if name not in _ENV_REGISTRY:
        raise ValueError(f"Unknown environment: '{name}'. Available: {list_envs()}")
    env_cls, renderer_cls, config_cls = _ENV_REGISTRY[name]
    cfg = config_cls(**config_overrides)
    renderer = renderer_cls(cfg)
    return env_cls(renderer=renderer, cfg=cfg)


def _register_builtins() -> None:
    from .cartpole import (CartPoleConfig, CartPoleEnv, CartPoleGroundRenderer,
                           CartPoleRenderer)
    register("CartPole-v0", CartPoleEnv, CartPoleRenderer, CartPoleConfig)
    register("CartPoleGround-v0", CartPoleEnv, CartPoleGroundRenderer, CartPoleConfig)
    from .steering import SteeringConfig, SteeringEnv, SteeringRenderer
    register("Steering-v0", SteeringEnv, SteeringRenderer, SteeringConfig)


def _register_from_entry_points() -> None:
    try:
        from importlib.metadata import entry_points
        for ep in entry_points(group="pybatchrender.envs"):
            try:
                fn = ep.load()
                if callable(fn):
                    fn()
            except Exception as e:  # pragma: no cover - third-party plugins
                warnings.warn(f"Failed to load environment entry point '{ep.name}': {e}")
    except Exception:  # pragma: no cover
        pass


_register_builtins()
_register_from_entry_points()

__all__ = ["register", "unregister", "list_envs", "is_registered", "get_env_classes", "make"]
