"""pybatchrender: the reference's import name for this renderer.

Code written for the reference says ``import pybatchrender as pbr`` and
imports its submodules by their full names (``pybatchrender.renderer.node``,
``pybatchrender.envs.cartpole.env`` ...).  The implementation lives in
``pybatchrender_amd``; this package gives it the reference's names without a
second copy of anything: every ``pybatchrender.X`` module *is*
``pybatchrender_amd.X`` (the same module object, so classes, the env registry
and ``isinstance`` checks are shared between the two names).

The top level lists the reference's names (pybatchrender/__init__.py:1-70)
with the same guarded imports: a name whose import fails is None.  There is
no frame grabber (the kernels write the final [N, C, H, W] tensor), so
GPUFrameGrabber, CPUFrameGrabber and FrameGrabber are None and
GPU_AVAILABLE is False, as in the reference when its grabber cannot load.
"""
from __future__ import annotations

import importlib
import importlib.abc
import importlib.util
import sys

_IMPL = "pybatchrender_amd"


class _SameModule(importlib.abc.Loader):
    """Loader whose module is the already imported implementation module."""

    def __init__(self, module):
        self._module = module

    def create_module(self, spec):
        return self._module

    def exec_module(self, module) -> None:
        return None


class _AliasFinder(importlib.abc.MetaPathFinder):
    """Resolves ``pybatchrender.<sub>`` to ``pybatchrender_amd.<sub>``.

    It sits first on sys.meta_path: the implementation packages' __path__
    would otherwise let the path finder load their files a second time under
    the alias name, as distinct modules."""

    def find_spec(self, fullname, path=None, target=None):
        if not fullname.startswith(__name__ + "."):
            return None
        real = _IMPL + fullname[len(__name__):]
        try:
            if importlib.util.find_spec(real) is None:
                return None
        except ModuleNotFoundError:
            return None
        module = importlib.import_module(real)
        return importlib.util.spec_from_loader(fullname, _SameModule(module),
                                               is_package=hasattr(module, "__path__"))


if not any(isinstance(f, _AliasFinder) for f in sys.meta_path):
    sys.meta_path.insert(0, _AliasFinder())

_impl = importlib.import_module(_IMPL)
__version__ = _impl.__version__

from pybatchrender_amd.config import PBRConfig  # noqa: E402,F401

try:
    from pybatchrender_amd.env import PBREnv  # noqa: F401
except Exception:  # pragma: no cover - mirrors the reference's optional TorchRL
    PBREnv = None  # type: ignore

try:
    from pybatchrender_amd.renderer.renderer import PBRRenderer  # noqa: F401
except Exception:  # pragma: no cover
    PBRRenderer = None  # type: ignore

try:
    from pybatchrender_amd.renderer.node import PBRNode  # noqa: F401
except Exception:  # pragma: no cover
    PBRNode = None  # type: ignore

try:
    from pybatchrender_amd.renderer.camera import PBRCam  # noqa: F401
except Exception:  # pragma: no cover
    PBRCam = None  # type: ignore

try:
    from pybatchrender_amd.renderer.light import PBRLight  # noqa: F401
except Exception:  # pragma: no cover
    PBRLight = None  # type: ignore

try:
    from pybatchrender_amd.renderer.shader_context import PBRShaderContext  # noqa: F401
except Exception:  # pragma: no cover
    PBRShaderContext = None  # type: ignore

GPUFrameGrabber = None
CPUFrameGrabber = None
FrameGrabber = None
GPU_AVAILABLE = False

from pybatchrender_amd import envs, renderer  # noqa: E402,F401

# the package's own submodules by their reference names
for _sub in ("config", "env", "renderer", "envs"):
    sys.modules.setdefault(f"{__name__}.{_sub}", getattr(_impl, _sub))
del _sub

__all__ = [
    "PBRConfig", "PBREnv", "PBRRenderer", "PBRNode", "PBRCam", "PBRLight", "PBRShaderContext",
    "GPUFrameGrabber", "CPUFrameGrabber", "FrameGrabber", "envs", "__version__",
]
