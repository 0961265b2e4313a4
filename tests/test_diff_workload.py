"""The differentiable-render workload: 128 x 128 float32 frames of 32
instances per scene (16 flat boxes, 16 smooth 16x8 UV spheres), fitted to a
target image rendered from perturbed parameters (colours + N(0, 0.05), light
direction rotated by N(0, 5 deg), vertex offsets N(0, 0.01)) with a
per-pixel L2 loss.

CPU: the perturbations have the stated distributions.
GPU: on an 8-scene slice at the full resolution and instance count, the HIP
gradient of the L2 loss equals the float64 dense reference's autograd
gradient of the same vector-Jacobian product (oracle/diff_ref.py) within
2e-3 of each tensor's scale, and eps = 1e-3 central differences of the
float64 product on interior pixels within 1e-3 relative."""
from __future__ import annotations

import math

import pytest
import torch

from oracle.diff_ref import DenseScene, interior_mask, loss_fn
from pybatchrender import diff, workloads


def test_target_perturbations_have_the_stated_distributions():
    r, tgt = workloads.differentiable(num_scenes=64, device="cpu", seed=2)
    p = diff.scene_params(r, requires_grad=False)
    dc = torch.cat([(a - b).reshape(-1) for a, b in zip(tgt.colors, p.colors)])
    assert abs(dc.mean().item()) < 0.005 and abs(dc.std().item() - 0.05) < 0.003
    dv = tgt.vertices - p.vertices
    assert (dv[:, 3:] == 0).all()                        # normals are not perturbed
    assert abs(dv[:, :3].std().item() - 0.01) < 0.0015
    # the light direction keeps its length and turns by a few degrees; the
    # rest of the light vector is unchanged
    a, b = p.light[:3].double(), tgt.light[:3].double()
    assert abs(a.norm().item() - b.norm().item()) < 1e-6
    ang = math.degrees(math.acos(min(1.0, (a @ b).item() / (a.norm() * b.norm()).item())))
    assert 0.0 < ang < 20.0
    assert torch.equal(tgt.light[3:], p.light[3:])
    # angles over many seeds: |N(0, 5 deg)| has mean 5 * sqrt(2 / pi) = 3.99 deg
    angs = []
    for seed in range(40):
        r, t = workloads.differentiable(num_scenes=1, instances=2, device="cpu", seed=seed)
        a = diff.light_vector(r)[:3].double()
        b = t.light[:3].double()
        angs.append(math.degrees(math.acos(min(1.0, (a @ b).item() / (a.norm() * b.norm()).item()))))
    assert 2.8 < sum(angs) / len(angs) < 5.2


def _fd(fn, t: torch.Tensor, idx, eps: float = 1e-3) -> float:
    with torch.no_grad():
        old = t[idx].item()
        t[idx] = old + eps
        lp = fn().item()
        t[idx] = old - eps
        lm = fn().item()
        t[idx] = old
    return (lp - lm) / (2 * eps)


@pytest.mark.gpu
def test_differentiable_workload_slice(cuda):
    """The backward of the L2 loss is the vector-Jacobian product of the
    shading with the loss gradient g = 2 (img - target) of the rendered
    image.  The reference takes the same product, <g, shade(theta)>, through
    float64 autograd (and central differences), with g from the kernel's
    image: the kernel shades smooth nodes with 1/256-pixel snapped weights,
    the reference with continuous ones, so residuals computed separately
    would differ by up to 2e-3 per pixel, which is not a backward error."""
    n = 8
    r, tgt = workloads.differentiable(num_scenes=n, seed=0)
    print(f"[diff workload] {n} scenes 128x128, slots/scene "
          f"{sum(nd.instances_per_scene * nd.mesh.num_triangles for nd in r._drawn_nodes())}",
          flush=True)
    with torch.no_grad():
        target = diff.render(r, tgt)
    p = diff.scene_params(r)
    img = diff.render(r, p)
    ((img - target) ** 2).sum().backward()
    torch.cuda.synchronize()
    got = {"light": p.light.grad.cpu().double(), "vertices": p.vertices.grad.cpu().double(),
           **{f"color{j}": c.grad.cpu().double() for j, c in enumerate(p.colors)}}
    g64 = (2.0 * (img.detach() - target)).cpu().double()

    desc, keep = r.scene_descriptor()
    ds = DenseScene(desc, keep)
    vis = r.render(vis=True)["vis"].cpu()
    assert (vis >= 0).float().mean() > 0.2
    cols = [c.detach().cpu().double().requires_grad_(True) for c in p.colors]
    verts = p.vertices.detach().cpu().double().requires_grad_(True)
    lt = p.light.detach().double().requires_grad_(True)
    loss_fn(ds, vis, g64, cols, verts, lt, 3).backward()
    want = {"light": lt.grad, "vertices": verts.grad,
            **{f"color{j}": c.grad for j, c in enumerate(cols)}}
    for k in want:
        a, b = got[k], want[k]
        scale = b.abs().max().item()
        err = (a - b).abs().max().item()
        print(f"[diff workload] {k}: scale {scale:.3e} max err {err:.3e}", flush=True)
        assert scale > 0 and err <= 2e-3 * scale, k
        big = b.abs() > 0.05 * scale
        assert ((a - b).abs()[big] <= 2e-3 * b.abs()[big] + 1e-5 * scale).all(), k

    # central differences of the float64 product on interior pixels
    m = interior_mask(ds, vis, [c.detach() for c in cols], verts.detach(), lt.detach(), 3)
    assert m.sum() > 10000
    gm = g64 * m[:, None]
    for t in (*p.colors, p.vertices, p.light):
        t.grad = None
    img = diff.render(r, p)
    img.backward(gm.to(img.device, torch.float32))
    torch.cuda.synchronize()
    got = {"light": p.light.grad.cpu(), "vertices": p.vertices.grad.cpu(),
           **{f"color{j}": c.grad.cpu() for j, c in enumerate(p.colors)}}

    def f():
        return loss_fn(ds, vis, gm, cols, verts, lt, 3)

    voff, nv = r._geometry["meshes"][1][:2]
    top = (torch.argsort(got["vertices"][voff:voff + nv, :3].abs().amax(1), descending=True)[:3]
           + voff).tolist()
    box_top = int(torch.argmax(got["vertices"][:24, 3:].abs().amax(1)).item())
    checks = [("light", lt, (i,)) for i in range(10)]
    checks += [("vertices", verts, (v, k)) for v in top for k in range(6)]
    checks += [("vertices", verts, (box_top, k)) for k in range(3, 6)]
    for j in range(2):
        g = got[f"color{j}"]
        k = int(torch.argmax(g.abs().sum(1)).item())
        checks += [(f"color{j}", cols[j], (k, ch)) for ch in range(3)]
    n_sig = 0
    for name, t, idx in checks:
        a, fd = got[name][idx].item(), _fd(f, t, idx)
        print(f"[fd] {name}{idx}: kernel {a:+.6e} fd {fd:+.6e}", flush=True)
        if abs(fd) > 1e-4:
            n_sig += 1
            assert abs(a - fd) <= 1e-3 * abs(fd), (name, idx, a, fd)
        else:
            assert abs(a) < 1e-3, (name, idx, a, fd)
    assert n_sig >= 25
