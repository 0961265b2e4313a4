"""Renderer core: PBRRenderer and its node / camera / light helpers."""
from .camera import PBRCam
from .light import PBRLight
from .mesh import Mesh
from .node import PBRNode
from .renderer import PBRRenderer
from .shader_context import PBRShaderContext

__all__ = ["PBRRenderer", "PBRNode", "PBRCam", "PBRLight", "PBRShaderContext", "Mesh"]
