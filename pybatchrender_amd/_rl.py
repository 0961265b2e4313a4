"""TorchRL / TensorDict bindings with a minimal stand-in when they are absent.

The reference's environments subclass ``torchrl.envs.EnvBase`` and exchange
``tensordict.TensorDict`` objects (pybatchrender/env.py:1-82).  When torchrl
is installed it is used as is.  Otherwise the small classes below provide the
subset the environments of this package use: nested-key TensorDicts, the
``reset``/``step`` protocol with a ``"next"`` sub-dict, and specs with
``rand()``/``zero()``.
"""
from __future__ import annotations

import torch

try:  # pragma: no cover - exercised only where torchrl is installed
    from tensordict import TensorDict  # type: ignore
    from torchrl.data.tensor_specs import (Bounded, Categorical, Composite,  # type: ignore
                                           Unbounded)
    from torchrl.envs import EnvBase  # type: ignore

    HAVE_TORCHRL = True
except Exception:  # torchrl / tensordict not installed
    HAVE_TORCHRL = False

    class TensorDict(dict):
        """Dict of tensors with a batch size and tuple (nested) keys."""

        def __init__(self, source=None, batch_size=None, device=None):
            super().__init__()
            self.batch_size = torch.Size(batch_size if batch_size is not None else [])
            for k, v in (source or {}).items():
                self[k] = v

        def __getitem__(self, key):
            if isinstance(key, tuple):
                cur = self
                for k in key:
                    cur = dict.__getitem__(cur, k)
                return cur
            return dict.__getitem__(self, key)

        def __setitem__(self, key, value):
            if isinstance(key, tuple):
                cur = self
                for k in key[:-1]:
                    if k not in cur:
                        dict.__setitem__(cur, k, TensorDict(batch_size=self.batch_size))
                    cur = dict.__getitem__(cur, k)
                dict.__setitem__(cur, key[-1], value)
            else:
                dict.__setitem__(self, key, value)

        def get(self, key, default=None):
            try:
                return self[key]
            except KeyError:
                return default

        def set(self, key, value):
            self[key] = value
            return self

        def clone(self):
            out = TensorDict(batch_size=self.batch_size)
            for k, v in self.items():
                out[k] = v.clone() if hasattr(v, "clone") else v
            return out

    class _Spec:
        def __init__(self, shape, dtype=torch.float32, device=None):
            self.shape = torch.Size(shape)
            self.dtype = dtype
            self.device = torch.device(device) if device is not None else torch.device("cpu")

        def zero(self):
            return torch.zeros(self.shape, dtype=self.dtype, device=self.device)

    class Unbounded(_Spec):
        def rand(self):
            if self.dtype.is_floating_point:
                return torch.randn(self.shape, dtype=self.dtype, device=self.device)
            return torch.randint(0, 256, self.shape, dtype=self.dtype, device=self.device)

    class Categorical(_Spec):
        def __init__(self, n, shape, dtype=torch.long, device=None):
            super().__init__(shape, dtype, device)
            self.n = int(n)

        def rand(self):
            return torch.randint(0, self.n, self.shape, dtype=self.dtype, device=self.device)

    class Bounded(_Spec):
        def __init__(self, low, high, shape, dtype=torch.float32, device=None):
            super().__init__(shape, dtype, device)
            self.low, self.high = float(low), float(high)

        def rand(self):
            u = torch.rand(self.shape, dtype=self.dtype, device=self.device)
            return self.low + (self.high - self.low) * u

    class Composite(dict):
        def __init__(self, shape=None, **fields):
            super().__init__(fields)
            self.shape = torch.Size(shape if shape is not None else [])

        def zero(self):
            return TensorDict({k: v.zero() for k, v in self.items()}, batch_size=self.shape)

    class EnvBase:
        """reset()/step() protocol of torchrl.envs.EnvBase, without checks."""

        def __init__(self, device="cpu", batch_size=torch.Size([]), **kwargs):
            self.device = torch.device(device)
            self.batch_size = torch.Size(batch_size)

        def reset(self, tensordict=None):
            return self._reset(tensordict)

        def step(self, tensordict):
            nxt = self._step(tensordict)
            tensordict["next"] = nxt
            return tensordict

        def rand_action(self, tensordict=None):
            td = tensordict if tensordict is not None else TensorDict(batch_size=self.batch_size)
            td["action"] = self.action_spec.rand()
            return td

        def set_seed(self, seed=None, static_seed=False):
            if seed is not None:
                self._set_seed(int(seed))
            return seed

        def close(self):
            pass

__all__ = ["TensorDict", "EnvBase", "Composite", "Unbounded", "Categorical", "Bounded",
           "HAVE_TORCHRL"]
