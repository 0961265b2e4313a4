"""Dense PyTorch reference of the differentiable shading (float64, CPU).

TEST INFRASTRUCTURE ONLY: the checker of pbr_render_backward.  Imported by
tests/ and bench.py's checks; never by the pybatchrender_amd package.

``DenseScene(desc, keep)`` takes a pbr_scene_desc and the tensors it points
to (as PBRRenderer.scene_descriptor returns them) and copies the inputs that
are not parameters (instance transforms, view-projection, indices, mesh
table, per-scene instance counts) to float64 CPU tensors.
``DenseScene.shade(vis, colors, vertices, light, channels)`` is the function
the backward differentiates, written with plain tensor ops so autograd gives
its gradient and central differences can be taken in float64:

  pixel (s, y, x) with vis = slot >= 0 belongs to triangle t of instance i of
  node j (slots run node by node, instance by instance, triangle by triangle:
  basic.vert:25-28 and the instanced draw order of node.py:45-46,68);
  clip_k = VP_s (M_row v_k), n_k = normalize(mat3(M_row) n_obj_k)
  (basic.vert:43,53);
  flat node:   n = n_0
  smooth node: n = sum_k u_k n_k, u_k = sgn(w0 w1 w2) (h_a x h_b).(px, py, 1),
               h = ((x + w) W/2, (y + w) H/2, w), (px, py) = (x + 1/2,
               H - 1 - y + 1/2) GL window pixel centre;
  rgb = rgba * ((1 - s) + (amb + max(normalize(n).L^, 0) col) s),
  alpha = rgba_3 (basic.frag:31-38), s = clamp(strength, 0, 1);
  background pixels keep the clear colour and carry no gradient.
"""
from __future__ import annotations

import torch

from pybatchrender import _abi


class DenseScene:
    def __init__(self, desc: _abi.pbr_scene_desc, keep: list[torch.Tensor]):
        by_ptr = {t.data_ptr(): t for t in keep}

        def host(ptr, dtype=torch.float64):
            return by_ptr[ptr].detach().to("cpu").to(dtype)

        self.N, self.W, self.H = int(desc.num_scenes), int(desc.width), int(desc.height)
        self.background = torch.tensor(list(desc.background), dtype=torch.float64)
        self.VP = host(desc.view_proj).reshape(-1, 4, 4)
        self.indices = host(desc.indices, torch.int64).reshape(-1, 3)
        self.nodes = []
        for j in range(desc.num_nodes):
            nd = desc.nodes[j]
            me = desc.meshes[nd.mesh]
            counts = None
            if nd.counts:
                counts = host(nd.counts, torch.int64).reshape(-1).clamp(0, nd.instances_per_scene)
            self.nodes.append(dict(
                transforms=host(nd.transforms).reshape(-1, 4, 4),
                ips=int(nd.instances_per_scene), shared=bool(nd.shared), flat=bool(nd.flat),
                voff=int(me.vertex_offset), toff=int(me.index_offset), ntri=int(me.num_triangles),
                counts=counts))

    def light_vector(self, desc: _abi.pbr_scene_desc) -> torch.Tensor:
        return torch.tensor([*desc.light_dir, *desc.light_col, *desc.ambient, desc.strength],
                            dtype=torch.float64)

    def _slot_base(self) -> torch.Tensor:
        """[N, nodes + 1] first slot of each node per scene."""
        cols = [torch.zeros(self.N, dtype=torch.int64)]
        for nd in self.nodes:
            cnt = nd["counts"] if nd["counts"] is not None else torch.full((self.N,), nd["ips"])
            cols.append(cols[-1] + cnt * nd["ntri"])
        return torch.stack(cols, 1)

    def shade(self, vis: torch.Tensor, colors: list[torch.Tensor], vertices: torch.Tensor,
              light: torch.Tensor, channels: int, ndl_out: torch.Tensor | None = None) -> torch.Tensor:
        """Image [N, C, H, W] (float64).  If ``ndl_out`` ([N, H, W]) is given,
        normalize(n).L^ of every covered pixel is written into it."""
        N, H, W, C = self.N, self.H, self.W, channels
        vis = vis.to("cpu").to(torch.int64)
        light = light.to("cpu", torch.float64)
        vertices = vertices.to("cpu", torch.float64)
        Ld = light[0:3]
        Lh = Ld / torch.linalg.norm(Ld)
        lcol, amb = light[3:6], light[6:9]
        s = torch.clamp(light[9], 0.0, 1.0)
        img = self.background[:C].reshape(1, C, 1, 1).expand(N, C, H, W).clone()
        base = self._slot_base()
        sc, yy, xx = torch.nonzero(vis >= 0, as_tuple=True)
        slot = vis[sc, yy, xx]
        for j, nd in enumerate(self.nodes):
            lo, hi = base[sc, j], base[sc, j + 1]
            m = (slot >= lo) & (slot < hi)
            if not bool(m.any()):
                continue
            ps, py_, px_ = sc[m], yy[m], xx[m]
            local = slot[m] - lo[m]
            inst, tri = local // nd["ntri"], local % nd["ntri"]
            row = inst if nd["shared"] else ps * nd["ips"] + inst
            M = nd["transforms"][row]                                   # [P,4,4]
            VP = self.VP[ps]                                            # [P,4,4]
            vidx = nd["voff"] + self.indices[nd["toff"] + tri]          # [P,3]
            v = vertices[vidx]                                          # [P,3,6]
            wp = torch.einsum("pij,pkj->pki", M[:, :, :3], v[:, :, :3]) + M[:, None, :, 3]
            clip = torch.einsum("pij,pkj->pki", VP, wp)                 # [P,3,4]
            u = torch.einsum("pij,pkj->pki", M[:, :3, :3], v[:, :, 3:6])
            nw = u / torch.linalg.norm(u, dim=2, keepdim=True)          # [P,3,3]
            if nd["flat"]:
                n = nw[:, 0]
            else:
                x, y, w = clip[..., 0], clip[..., 1], clip[..., 3]
                h = torch.stack([(x + w) * (W / 2), (y + w) * (H / 2), w], 2)   # [P,3,3]
                sg = torch.where(w.prod(1) < 0, -1.0, 1.0).to(torch.float64)
                P = torch.stack([px_.to(torch.float64) + 0.5,
                                 (H - 1 - py_).to(torch.float64) + 0.5,
                                 torch.ones_like(px_, dtype=torch.float64)], 1)
                n = 0
                for k in range(3):
                    ck = torch.cross(h[:, (k + 1) % 3], h[:, (k + 2) % 3], dim=1) * sg[:, None]
                    uk = (ck * P).sum(1)
                    n = n + uk[:, None] * nw[:, k]
            nn = n / torch.linalg.norm(n, dim=1, keepdim=True)
            ndl_raw = (nn * Lh).sum(1)
            if ndl_out is not None:
                ndl_out[ps, py_, px_] = ndl_raw.detach()
            ndl = torch.where(ndl_raw > 0, ndl_raw, torch.zeros_like(ndl_raw))
            lit = amb[None, :] + ndl[:, None] * lcol[None, :]
            l = (1.0 - s) + lit * s
            rgba = colors[j].to("cpu").to(torch.float64)[row]
            vals = torch.cat([rgba[:, :3] * l, rgba[:, 3:4]], 1)[:, :C]
            idx = (ps[:, None].expand(-1, C), torch.arange(C)[None, :].expand(len(ps), C),
                   py_[:, None].expand(-1, C), px_[:, None].expand(-1, C))
            img = img.index_put(idx, vals)
        return img


def interior_mask(scene: DenseScene, vis: torch.Tensor, colors, vertices, light, channels: int,
                  margin: float = 0.02) -> torch.Tensor:
    """Covered pixels whose 3x3 neighbourhood shows the same triangle and whose
    |normalize(n).L^| exceeds ``margin`` (away from silhouettes and from the
    kink of max(., 0), so a small parameter step keeps the function smooth)."""
    vis = vis.to("cpu").to(torch.int64)
    ndl = torch.zeros(vis.shape, dtype=torch.float64)
    with torch.no_grad():
        scene.shade(vis, colors, vertices, light, channels, ndl_out=ndl)
    N, H, W = vis.shape
    pad = torch.nn.functional.pad(vis, (1, 1, 1, 1), value=-2)
    same = torch.ones_like(vis, dtype=torch.bool)
    for dy in (0, 1, 2):
        for dx in (0, 1, 2):
            same &= pad[:, dy:dy + H, dx:dx + W] == vis
    return same & (vis >= 0) & (ndl.abs() > margin)


def loss_fn(scene: DenseScene, vis, gout64, colors, vertices, light, channels):
    """<gout, shade(...)> summed over covered pixels (background is constant)."""
    img = scene.shade(vis, colors, vertices, light, channels)
    return (img * gout64).sum()
