"""Differentiable rendering: gradients of the colour image with respect to the
instance colours, the light and the mesh vertices, visibility held fixed.

The reference renders through Panda3D/GL and has no backward; this module adds
one on top of the same scene API.  ``render(renderer, params)`` draws the
renderer's current scenes with the parameter tensors in ``params`` substituted
for the renderer's own, and returns a float32 ``[N, C, H, W]`` image that
autograd can differentiate:

    params = diff.scene_params(renderer)          # leaf tensors, requires_grad
    img = diff.render(renderer, params)
    loss = ((img - target) ** 2).mean()
    loss.backward()                               # params.colors[j].grad, ...

Forward: ``pbr_render`` with float32 output and the triangle-id (visibility)
buffer.  Backward: ``pbr_render_backward`` reuses that buffer, recomputes the
interpolation weights per pixel and scatters the gradient to

    colors[j]   [buf_instances, 4]  per-instance RGBA of drawn node j
    vertices    [V, 6]              model-space position and normal of every
                                    vertex of the renderer's mesh pool
    light       [10]                light_dir (3, as given, unnormalised),
                                    light_col (3), ambient (3), strength

The backward differentiates the shading of basic.frag:31-38 with
perspective-correct weights taken from the unsnapped window positions, so the
vertex-position gradient is that of a smooth function (flat-shaded nodes get
none: their colour does not depend on positions while visibility is fixed).
Instance transforms and the camera are inputs, not parameters: the forward
keeps a copy of both for its backward.
"""
from __future__ import annotations

from dataclasses import dataclass

import torch

from . import _abi, _native

__all__ = ["SceneParams", "scene_params", "render", "light_vector"]


@dataclass
class SceneParams:
    colors: list[torch.Tensor]
    vertices: torch.Tensor
    light: torch.Tensor

    def tensors(self) -> list[torch.Tensor]:
        return [self.vertices, self.light, *self.colors]


def light_vector(renderer) -> torch.Tensor:
    """The renderer's light as the 10-vector the backward differentiates."""
    L = renderer._pbr_light
    if L is None:  # the no-light quirk: unlit (strength 0)
        return torch.zeros(10, dtype=torch.float32)
    return torch.tensor([*L.dir_dir, *L.dir_col, *L.ambient, float(L.strength)],
                        dtype=torch.float32)


def scene_params(renderer, requires_grad: bool = True) -> SceneParams:
    """Leaf copies of the renderer's colours, vertex pool and light.  The
    light lives on the CPU (its values are launch arguments); colours and
    vertices on the renderer's device."""
    if renderer._geometry is None:
        renderer._build_geometry()
    g = renderer._geometry
    cols = [n.colors_b4.detach().to(torch.float32).contiguous().clone() for n in g["nodes"]]
    verts = g["vertices"].detach().clone()
    light = light_vector(renderer)
    out = SceneParams(cols, verts, light)
    for t in out.tensors():
        t.requires_grad_(requires_grad)
    return out


def _descriptor(renderer, vertices, light, colors):
    desc, keep = renderer.scene_descriptor()
    g = renderer._geometry
    if vertices.shape != g["vertices"].shape or vertices.dtype != torch.float32:
        raise ValueError(f"vertices must be float32 {tuple(g['vertices'].shape)}")
    if not vertices.is_contiguous() or vertices.device != renderer.device:
        raise ValueError(f"vertices must be contiguous on {renderer.device}")
    desc.vertices = vertices.data_ptr()
    keep.append(vertices)
    if len(colors) != desc.num_nodes:
        raise ValueError(f"expected {desc.num_nodes} colour tensors, got {len(colors)}")
    for j, c in enumerate(colors):
        node = g["nodes"][j]
        if c.shape != (node.buf_instances, 4) or c.dtype != torch.float32 or not c.is_contiguous():
            raise ValueError(f"colors[{j}] must be contiguous float32 ({node.buf_instances}, 4)")
        if c.device != renderer.device:
            raise ValueError(f"colors[{j}] must be on {renderer.device}")
        desc.nodes[j].colors = c.data_ptr()
        keep.append(c)
    lv = [float(v) for v in light.detach().reshape(-1).tolist()]
    if len(lv) != 10:
        raise ValueError("light must have 10 entries")
    desc.light_dir[:] = lv[0:3]
    desc.light_col[:] = lv[3:6]
    desc.ambient[:] = lv[6:9]
    desc.strength = lv[9]
    # transforms and view-projection are read again by the backward: freeze them
    for j in range(desc.num_nodes):
        tr = g["nodes"][j].transforms_b44.detach().to(torch.float32).contiguous().clone()
        desc.nodes[j].transforms = tr.data_ptr()
        keep.append(tr)
    vp = renderer._pbr_cam.VP_k44.detach().to(torch.float32).reshape(-1, 4, 4).contiguous().clone()
    desc.view_proj = vp.data_ptr()
    keep.append(vp)
    return desc, keep


class _DiffRender(torch.autograd.Function):
    @staticmethod
    def forward(ctx, renderer, vertices, light, *colors):
        if renderer.device.type != "cuda":
            raise RuntimeError("differentiable rendering needs the HIP renderer on a GPU device")
        desc, keep = _descriptor(renderer, vertices, light, colors)
        N, C, H, W = renderer.output_shape()
        out = torch.empty((N, C, H, W), dtype=torch.float32, device=renderer.device)
        vis = torch.empty((N, H, W), dtype=torch.int32, device=renderer.device)
        fo = _abi.pbr_frame_out()
        fo.color, fo.depth, fo.vis = out.data_ptr(), None, vis.data_ptr()
        fo.channels, fo.dtype = C, _abi.PBR_DTYPE_F32
        opts = renderer.render_opts()
        ws_bytes = _native.workspace_bytes(desc, opts)
        ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=renderer.device)
        stream = torch.cuda.current_stream(renderer.device).cuda_stream
        _native.render(desc, fo, opts, ws.data_ptr() if ws_bytes else 0, ws_bytes, stream)
        ctx.desc, ctx.keep, ctx.vis = desc, keep + [ws], vis
        ctx.shapes = (C, [c.shape for c in colors], vertices.shape, light.device)
        ctx.mark_non_differentiable(vis)
        return out

    @staticmethod
    def backward(ctx, grad_out):
        C, col_shapes, vshape, light_dev = ctx.shapes
        desc = ctx.desc
        dev = ctx.vis.device
        need = ctx.needs_input_grad
        g = _abi.pbr_grads()
        g_verts = torch.zeros(vshape, dtype=torch.float32, device=dev) if need[1] else None
        g_light = torch.zeros(10, dtype=torch.float32, device=dev) if need[2] else None
        g_cols = [torch.zeros(s, dtype=torch.float32, device=dev) if need[3 + j] else None
                  for j, s in enumerate(col_shapes)]
        if g_verts is not None:
            g.vertices = g_verts.data_ptr()
        if g_light is not None:
            g.light = g_light.data_ptr()
        for j, t in enumerate(g_cols):
            if t is not None:
                g.colors[j] = t.data_ptr()
        gout = grad_out.detach().to(torch.float32).contiguous()
        ws_bytes = _native.backward_workspace_bytes(desc, C, g)
        ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=dev)
        stream = torch.cuda.current_stream(dev).cuda_stream
        _native.render_backward(desc, ctx.vis.data_ptr(), gout.data_ptr(), C, g,
                                ws.data_ptr() if ws_bytes else 0, ws_bytes, stream)
        if g_light is not None:
            g_light = g_light.to(light_dev)
        return (None, g_verts, g_light, *g_cols)


def render(renderer, params: SceneParams) -> torch.Tensor:
    """Differentiable float32 [N, C, H, W] image of the renderer's scenes."""
    return _DiffRender.apply(renderer, params.vertices, params.light, *params.colors)
