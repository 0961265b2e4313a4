"""Differentiable rendering: pbr_render_backward vs a dense float64 PyTorch
reference (oracle/diff_ref.py) and vs central finite differences.

CPU tests pin the reference itself: its image equals the CPU oracle's on every
covered pixel (to float32 rounding for flat nodes, to the sub-pixel snap for
smooth ones), and its autograd gradient equals eps=1e-3 central differences
on interior pixels.  GPU tests compare the HIP backward with that gradient.
"""
from __future__ import annotations

import pytest
import torch

from oracle import oracle
from oracle.diff_ref import DenseScene, interior_mask, loss_fn
from pybatchrender import PBRConfig, PBRRenderer, diff
from pybatchrender.renderer.mesh import uv_sphere
from pybatchrender.workloads import random_transforms


def mixed_scene(device: str, N: int = 3, res: int = 48, seed: int = 0, channels: int = 3,
                strength: float = 0.9, sphere=(8, 4), sphere_ips: int = 2,
                box_ips: int = 2) -> PBRRenderer:
    """Boxes (flat), UV spheres (smooth) and a shared ground quad (flat)."""
    g = torch.Generator("cpu").manual_seed(seed)
    r = PBRRenderer(PBRConfig(num_scenes=N, tile_resolution=(res, res), device=device,
                              output_dtype="float32", num_channels=channels))
    box = r.add_node("models/box", instances_per_scene=box_ips)
    sph = r.add_node(uv_sphere(*sphere), instances_per_scene=sphere_ips)
    gnd = r.add_node("models/quad", instances_per_scene=1, shared_across_scenes=True)
    for nd, ip in ((box, box_ips), (sph, sphere_ips)):
        nd.set_transforms(random_transforms(N * ip, g, pos_range=1.2, scale_range=(1.0, 1.6)))
        col = torch.rand(N * ip, 4, generator=g) * 0.8 + 0.1
        nd.set_colors(col)
    Mg = torch.eye(4).reshape(1, 4, 4).clone()
    Mg[0, 0, 0] = Mg[0, 1, 1] = 8.0
    Mg[0, 2, 3] = -1.5
    gnd.set_transforms(Mg)
    gnd.set_colors(torch.tensor([[0.3, 0.6, 0.3, 0.8]]))
    r.add_camera(fov_y_deg=50.0)
    eye = torch.tensor([[0.5, -5.0, 2.5]]).expand(N, 3)
    r._pbr_cam.set_positions_and_lookat(eye, torch.zeros(3))
    r.add_light(ambient=(0.2, 0.2, 0.25), dir_dir=(0.4, -0.6, 0.7), dir_col=(0.9, 0.8, 0.7),
                strength=strength)
    r.setup_environment()
    return r


def flat_scene(device: str, N: int = 5, res: int = 48, seed: int = 0, channels: int = 3,
               box_ips: int = 3) -> PBRRenderer:
    """Boxes and a shared ground quad: every node shades flat."""
    g = torch.Generator("cpu").manual_seed(seed)
    r = PBRRenderer(PBRConfig(num_scenes=N, tile_resolution=(res, res), device=device,
                              output_dtype="float32", num_channels=channels))
    box = r.add_node("models/box", instances_per_scene=box_ips)
    gnd = r.add_node("models/quad", instances_per_scene=1, shared_across_scenes=True)
    box.set_transforms(random_transforms(N * box_ips, g, pos_range=1.2, scale_range=(1.0, 1.6)))
    box.set_colors(torch.rand(N * box_ips, 4, generator=g) * 0.8 + 0.1)
    Mg = torch.eye(4).reshape(1, 4, 4).clone()
    Mg[0, 0, 0] = Mg[0, 1, 1] = 8.0
    Mg[0, 2, 3] = -1.5
    gnd.set_transforms(Mg)
    gnd.set_colors(torch.tensor([[0.3, 0.6, 0.3, 0.8]]))
    r.add_camera(fov_y_deg=50.0)
    eye = torch.tensor([[0.5, -5.0, 2.5]]).expand(N, 3)
    r._pbr_cam.set_positions_and_lookat(eye, torch.zeros(3))
    r.add_light(ambient=(0.2, 0.2, 0.25), dir_dir=(0.4, -0.6, 0.7), dir_col=(0.9, 0.8, 0.7),
                strength=0.9)
    r.setup_environment()
    assert all(r._geometry["flat"])
    return r


def _node_masks(ds: DenseScene, vis: torch.Tensor) -> list[torch.Tensor]:
    base = ds._slot_base()
    v = vis.long()
    return [(v >= base[:, j].reshape(-1, 1, 1)) & (v < base[:, j + 1].reshape(-1, 1, 1))
            for j in range(len(ds.nodes))]


# --- CPU: the reference itself ------------------------------------------------

def test_dense_reference_matches_oracle_image():
    r = mixed_scene("cpu")
    desc, keep = r.scene_descriptor()
    ref = oracle.render(desc, keep, channels=3, dtype="float32")
    ds = DenseScene(desc, keep)
    p = diff.scene_params(r, requires_grad=False)
    img = ds.shade(ref["vis"], p.colors, p.vertices, ds.light_vector(desc), 3)
    d = (img - ref["pixels"].double()).abs().amax(1)
    masks = _node_masks(ds, ref["vis"])
    assert masks[0].any() and masks[1].any() and masks[2].any()
    # flat nodes: the same formula, float32 vs float64 only
    assert d[masks[0]].max() < 1e-6 and d[masks[2]].max() < 1e-6
    # smooth node: continuous vs 1/256-pixel snapped interpolation weights
    assert d[masks[1]].max() < 2e-3
    assert d[ref["vis"] < 0].max() == 0


def _fd(fn, t: torch.Tensor, idx, eps: float = 1e-3) -> float:
    with torch.no_grad():
        old = t[idx].item()
        t[idx] = old + eps
        lp = fn().item()
        t[idx] = old - eps
        lm = fn().item()
        t[idx] = old
    return (lp - lm) / (2 * eps)


def test_dense_reference_gradient_matches_finite_differences():
    r = mixed_scene("cpu")
    desc, keep = r.scene_descriptor()
    vis = oracle.render(desc, keep, channels=3, dtype="float32")["vis"]
    ds = DenseScene(desc, keep)
    p = diff.scene_params(r, requires_grad=False)
    light = ds.light_vector(desc)
    m = interior_mask(ds, vis, p.colors, p.vertices, light, 3)
    assert m.sum() > 500
    gout = torch.randn(3, 3, 48, 48, generator=torch.Generator().manual_seed(1),
                       dtype=torch.float64) * m[:, None]
    cols = [c.double().clone().requires_grad_(True) for c in p.colors]
    verts = p.vertices.double().clone().requires_grad_(True)
    lt = light.clone().requires_grad_(True)
    loss_fn(ds, vis, gout, cols, verts, lt, 3).backward()

    def f():
        return loss_fn(ds, vis, gout, cols, verts, lt, 3)

    voff, nv = r._geometry["meshes"][1][:2]
    top = (torch.argsort(verts.grad[voff:voff + nv, :3].abs().amax(1), descending=True)[:3]
           + voff).tolist()
    checks = [(lt, (i,)) for i in range(10)]
    checks += [(verts, (v, k)) for v in top for k in range(6)]
    checks += [(cols[0], (0, 0)), (cols[1], (1, 2)), (cols[2], (0, 1))]
    n_sig = 0
    for t, idx in checks:
        a, fd = t.grad[idx].item(), _fd(f, t, idx)
        if abs(fd) > 1e-4:
            n_sig += 1
            assert abs(a - fd) <= 1e-3 * abs(fd), (idx, a, fd)
        else:
            assert abs(a) < 1e-3, (idx, a, fd)
    assert n_sig >= 20


def test_finite_differences_catch_a_planted_sign_error():
    """A reference with the light direction's sign flipped in its gradient
    path fails the same finite-difference check."""
    r = mixed_scene("cpu", N=1)
    desc, keep = r.scene_descriptor()
    vis = oracle.render(desc, keep, channels=3, dtype="float32")["vis"]
    ds = DenseScene(desc, keep)
    p = diff.scene_params(r, requires_grad=False)
    light = ds.light_vector(desc)
    m = interior_mask(ds, vis, p.colors, p.vertices, light, 3)
    gout = torch.randn(1, 3, 48, 48, generator=torch.Generator().manual_seed(2),
                       dtype=torch.float64) * m[:, None]
    lt = light.clone().requires_grad_(True)

    def planted(L):
        # d/dL of the flipped term is negated while its value is unchanged
        L2 = torch.cat([2 * L[:3].detach() - L[:3], L[3:]])
        return loss_fn(ds, vis, gout, p.colors, p.vertices, L2, 3)

    planted(lt).backward()
    fd = _fd(lambda: loss_fn(ds, vis, gout, p.colors, p.vertices, lt, 3), lt, (0,))
    assert abs(fd) > 1e-3
    assert abs(lt.grad[0].item() - fd) > 1e-3 * abs(fd)


# --- GPU: the HIP backward ------------------------------------------------------

def _kernel_and_reference(r: PBRRenderer, seed: int = 3, interior_only: bool = False):
    C = int(r.cfg.num_channels)
    p = diff.scene_params(r)
    img = diff.render(r, p)
    N, _, H, W = img.shape
    desc, keep = r.scene_descriptor()
    ds = DenseScene(desc, keep)
    vis = r.render(vis=True)["vis"].cpu()
    gout = torch.randn(N, C, H, W, generator=torch.Generator().manual_seed(seed))
    if interior_only:
        m = interior_mask(ds, vis, [c.detach() for c in p.colors], p.vertices.detach(),
                          p.light.detach(), C)
        gout = gout * m[:, None]
    (img * gout.to(img.device)).sum().backward()
    torch.cuda.synchronize()
    cols = [c.detach().cpu().double().requires_grad_(True) for c in p.colors]
    verts = p.vertices.detach().cpu().double().requires_grad_(True)
    lt = p.light.detach().double().requires_grad_(True)
    loss_fn(ds, vis, gout.double(), cols, verts, lt, C).backward()
    got = {"light": p.light.grad.cpu(), "vertices": p.vertices.grad.cpu(),
           **{f"color{j}": c.grad.cpu() for j, c in enumerate(p.colors)}}
    want = {"light": lt.grad, "vertices": verts.grad,
            **{f"color{j}": c.grad for j, c in enumerate(cols)}}
    return got, want, (ds, vis, gout, cols, verts, lt, C)


def _assert_close(got: dict, want: dict, rel: float = 2e-3, label: str = ""):
    for k in want:
        a, b = got[k].double(), want[k]
        scale = b.abs().max().item()
        err = (a - b).abs()
        print(f"[diff] {label} {k}: scale {scale:.3e} max err {err.max().item():.3e}", flush=True)
        assert err.max().item() <= rel * scale + 1e-6, (label, k, err.max().item(), scale)
        big = b.abs() > 0.05 * scale
        if big.any():
            assert (err[big] <= rel * b.abs()[big] + 1e-5 * scale).all(), (label, k)


@pytest.mark.gpu
def test_backward_matches_dense_reference_mixed_scene(cuda):
    r = mixed_scene("cuda")
    got, want, _ = _kernel_and_reference(r)
    _assert_close(got, want, label="mixed")
    assert got["vertices"].abs().sum() > 0 and got["light"].abs().sum() > 0


@pytest.mark.gpu
@pytest.mark.parametrize("channels", [1, 4])
def test_backward_channels(cuda, channels):
    r = mixed_scene("cuda", channels=channels, seed=5)
    got, want, _ = _kernel_and_reference(r, seed=4)
    _assert_close(got, want, label=f"C={channels}")
    if channels == 4:
        assert got["color0"][:, 3].abs().sum() > 0


@pytest.mark.gpu
def test_backward_many_slots_compact_passes(cuda):
    """2688 slots per scene: the visible slots are numbered through the LDS
    bitmap and processed in several passes."""
    r = mixed_scene("cuda", N=2, res=64, sphere=(16, 8), sphere_ips=12, seed=7)
    got, want, _ = _kernel_and_reference(r, seed=5)
    _assert_close(got, want, label="compact")


@pytest.mark.gpu
def test_backward_strength_clamp_and_uneven_counts(cuda):
    r = mixed_scene("cuda", N=4, strength=1.5, seed=9)
    counts = torch.tensor([0, 1, 2, 1], dtype=torch.int32)
    r._pbr_nodes[1].set_instance_counts(counts)
    r._mark_geometry_dirty()
    r.setup_environment()
    got, want, _ = _kernel_and_reference(r, seed=6)
    _assert_close(got, want, label="clamp+uneven")
    assert got["light"][9].item() == 0.0  # clamp(strength) passes no gradient above 1


@pytest.mark.gpu
def test_backward_cartpole_ground_scene(cuda):
    from pybatchrender.envs.cartpole import CartPoleConfig, CartPoleGroundRenderer
    r = CartPoleGroundRenderer(CartPoleConfig(num_scenes=16, device="cuda",
                                              output_dtype="float32"))
    g = torch.Generator("cpu").manual_seed(0)
    st = torch.zeros(16, 4)
    st[:, 0] = torch.empty(16).uniform_(-2.4, 2.4, generator=g)
    st[:, 2] = torch.empty(16).uniform_(-0.21, 0.21, generator=g)
    r._step(st.cuda())
    got, want, _ = _kernel_and_reference(r, seed=8)
    _assert_close(got, want, label="cartpole")


@pytest.mark.gpu
@pytest.mark.parametrize("channels", [1, 2, 3, 4])
def test_backward_flat_scene(cuda, channels):
    """Every node flat: the persistent flat kernel."""
    r = flat_scene("cuda", channels=channels, seed=channels)
    got, want, _ = _kernel_and_reference(r, seed=10 + channels)
    _assert_close(got, want, label=f"flat C={channels}")
    assert got["vertices"].abs().sum() > 0 and got["color0"].abs().sum() > 0


@pytest.mark.gpu
def test_backward_flat_uneven_counts(cuda):
    r = flat_scene("cuda", N=6, seed=11, box_ips=4)
    r._pbr_nodes[0].set_instance_counts(torch.tensor([0, 4, 1, 3, 2, 4], dtype=torch.int32))
    r._mark_geometry_dirty()
    r.setup_environment()
    got, want, _ = _kernel_and_reference(r, seed=12)
    _assert_close(got, want, label="flat uneven")


def _abi_backward(r: PBRRenderer, gout: torch.Tensor, rows: int | None = None) -> dict:
    """pbr_render_backward through the C ABI with a workspace of `rows`
    partial rows (None: the size pbr_backward_workspace_bytes asks for)."""
    from pybatchrender import _abi, _native
    p = diff.scene_params(r, requires_grad=False)
    desc, keep = diff._descriptor(r, p.vertices, p.light, p.colors)
    vis = r.render(vis=True)["vis"].contiguous()
    C = gout.shape[1]
    gcols = [torch.zeros_like(c) for c in p.colors]
    gverts = torch.zeros_like(p.vertices)
    glight = torch.zeros(10, device="cuda")
    g = _abi.pbr_grads()
    for j, t in enumerate(gcols):
        g.colors[j] = t.data_ptr()
    g.vertices, g.light = gverts.data_ptr(), glight.data_ptr()
    full = _native.backward_workspace_bytes(desc, C, g)
    nbytes = full
    if rows is not None:
        shared = sum(nd.instances_per_scene for nd in desc.nodes[:desc.num_nodes] if nd.shared)
        nbytes = rows * 4 * (10 + 6 * p.vertices.shape[0] + 4 * shared)
        assert nbytes <= full
    ws = torch.empty(max(full, 1), dtype=torch.uint8, device="cuda")
    go = gout.to("cuda").contiguous()
    stream = torch.cuda.current_stream().cuda_stream
    _native.render_backward(desc, vis.data_ptr(), go.data_ptr(), C, g, ws.data_ptr(), nbytes,
                            stream)
    torch.cuda.synchronize()
    del keep
    return {"light": glight.cpu(), "vertices": gverts.cpu(),
            **{f"color{j}": t.cpu() for j, t in enumerate(gcols)}}


@pytest.mark.gpu
@pytest.mark.parametrize("rows", [1, 3])
def test_backward_flat_small_workspace(cuda, rows):
    """A workspace of `rows` partial rows runs that many workgroups, each
    walking several scenes through both halves of its accumulator table."""
    r = flat_scene("cuda", N=11, res=40, seed=13)
    N, C, H, W = r.output_shape()
    gout = torch.randn(N, C, H, W, generator=torch.Generator().manual_seed(14))
    want = _abi_backward(r, gout)
    got = _abi_backward(r, gout, rows=rows)
    _assert_close(got, {k: v.double() for k, v in want.items()}, rel=1e-4,
                  label=f"{rows} rows")
    # and the full-size call against the float64 reference
    ds = DenseScene(*r.scene_descriptor())
    vis = r.render(vis=True)["vis"].cpu()
    p = diff.scene_params(r, requires_grad=False)
    cols = [c.detach().cpu().double().requires_grad_(True) for c in p.colors]
    verts = p.vertices.detach().cpu().double().requires_grad_(True)
    lt = p.light.detach().double().requires_grad_(True)
    loss_fn(ds, vis, gout.double(), cols, verts, lt, C).backward()
    ref = {"light": lt.grad, "vertices": verts.grad,
           **{f"color{j}": c.grad for j, c in enumerate(cols)}}
    _assert_close(got, ref, label=f"{rows} rows vs reference")


@pytest.mark.gpu
def test_backward_matches_finite_differences(cuda):
    """HIP gradient vs eps=1e-3 central differences of the float64 function
    on interior pixels, within 1e-3 relative."""
    r = mixed_scene("cuda", seed=0)
    got, _, (ds, vis, gout, cols, verts, lt, C) = _kernel_and_reference(r, seed=1,
                                                                         interior_only=True)
    g64 = gout.double()

    def f():
        return loss_fn(ds, vis, g64, cols, verts, lt, C)

    voff, nv = r._geometry["meshes"][1][:2]
    top = (torch.argsort(got["vertices"][voff:voff + nv, :3].abs().amax(1), descending=True)[:3]
           + voff).tolist()
    checks = [("light", lt, (i,)) for i in range(10)]
    checks += [("vertices", verts, (v, k)) for v in top for k in range(6)]
    checks += [("color0", cols[0], (0, 0)), ("color1", cols[1], (1, 2)),
               ("color2", cols[2], (0, 1))]
    n_sig = 0
    for name, t, idx in checks:
        a, fd = got[name][idx].item(), _fd(f, t, idx)
        print(f"[fd] {name}{idx}: kernel {a:+.6e} fd {fd:+.6e}", flush=True)
        if abs(fd) > 1e-4:
            n_sig += 1
            assert abs(a - fd) <= 1e-3 * abs(fd), (name, idx, a, fd)
        else:
            assert abs(a) < 1e-3, (name, idx, a, fd)
    assert n_sig >= 20
