"""CartPole environment (pybatchrender/envs/cartpole)."""
from .config import CartPoleConfig
from .env import CartPoleEnv
from .renderer import CartPoleGroundRenderer, CartPoleRenderer

__all__ = ["CartPoleConfig", "CartPoleEnv", "CartPoleRenderer", "CartPoleGroundRenderer"]
