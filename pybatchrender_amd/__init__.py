"""pybatchrender_amd: MI355X-native batch renderer with the pybatchrender API.

``import pybatchrender_amd as pbr`` exposes the reference's surface
(pybatchrender/__init__.py:1-70): PBRConfig, PBRRenderer, PBRNode, PBRCam,
PBRLight, PBRShaderContext, PBREnv and the ``envs`` registry.  Rendering runs
in hand-written HIP kernels (libpbr_hip.so) that write [N, C, H, W] tensors
directly; there is no frame grabber, so the reference's GPUFrameGrabber /
CPUFrameGrabber names are None.
"""
__version__ = "0.1.0"

from .config import PBRConfig  # noqa: F401
from .env import PBREnv  # noqa: F401
from .renderer import PBRCam, PBRLight, PBRNode, PBRRenderer, PBRShaderContext  # noqa: F401

GPUFrameGrabber = None
CPUFrameGrabber = None
FrameGrabber = None
GPU_AVAILABLE = False

from . import envs  # noqa: E402,F401

__all__ = [
    "PBRConfig", "PBREnv", "PBRRenderer", "PBRNode", "PBRCam", "PBRLight", "PBRShaderContext",
    "GPUFrameGrabber", "CPUFrameGrabber", "FrameGrabber", "envs", "__version__",
]
