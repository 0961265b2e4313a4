"""pbr_render_posed: one renderer step whose poses the render launch applies
itself (renderer.py:377-389 step = _step + render), against the two launches
it stands for (pbr_set_poses, then pbr_render) and against the oracle."""
from __future__ import annotations

import pytest
import torch

from oracle import oracle
from pybatchrender import PBRConfig, PBRRenderer, _abi, _native, workloads
from pybatchrender.envs.cartpole import CartPoleConfig, CartPoleGroundRenderer, CartPoleRenderer
from tests.test_gpu_parity import _bits, check_parity, ground_states

pytestmark = pytest.mark.gpu


def _kernel_names(fn) -> list[str]:
    """GPU kernels `fn` launches, from torch's profiler (empty when it
    records none on this platform)."""
    from torch.profiler import ProfilerActivity, profile
    with profile(activities=[ProfilerActivity.CUDA]) as prof:
        fn()
        torch.cuda.synchronize()
    return [e.name for e in prof.events()
            if e.device_type == torch.autograd.DeviceType.CUDA and "kernel" in e.name]


def _node_buffers(r: PBRRenderer) -> list[torch.Tensor]:
    out = []
    for n in r._drawn_nodes():
        out.append(n.transforms_b44)
        for name in ("rot3_b33", "scale_b11"):
            t = getattr(n, name, None)
            if isinstance(t, torch.Tensor):
                out.append(t)
    return out


def _compare_steps(fused: PBRRenderer, plain: PBRRenderer, states: list, label: str) -> None:
    """Step both renderers through `states`; pixels and every node buffer
    (the posed rows included) must agree bitwise, and the fused frames must
    match the oracle."""
    assert fused.fuse_poses and not plain.fuse_poses
    for k, st in enumerate(states):
        pf = fused.step(st).clone()
        pp = plain.step(st).clone()
        torch.cuda.synchronize()
        assert torch.equal(_bits(pf), _bits(pp)), f"{label} step {k}: pixels differ"
        for a, b in zip(_node_buffers(fused), _node_buffers(plain)):
            assert torch.equal(_bits(a), _bits(b)), f"{label} step {k}: node buffer differs"
    check_parity(fused, f"{label} fused step", depth=False, vis=False, got={"pixels": pf})


@pytest.mark.parametrize("n", [1, 257, 4096])
def test_cartpole_step_fused_matches_two_launches(cuda, n):
    """The headline step: cart and pole posed from the state columns inside
    the 32-record row-span launch."""
    fused = CartPoleGroundRenderer(CartPoleConfig(num_scenes=n, device="cuda"))
    plain = CartPoleGroundRenderer(CartPoleConfig(num_scenes=n, device="cuda"))
    plain.fuse_poses = False
    states = [ground_states(n, seed).cuda() for seed in (3, 4, 5)]
    _compare_steps(fused, plain, states, f"cartpole ground N={n}")


@pytest.mark.parametrize("dtype,depth", [("uint8", False), ("bfloat16", True)])
def test_cartpole_step_fused_at_16384_scenes(cuda, dtype, depth):
    """The largest CartPole batch the benchmarks run (the env-loop size), in
    the headline format and the env loop's bf16 + depth format."""
    n = 16384
    cfg = dict(num_scenes=n, device="cuda", output_dtype=dtype, output_depth=depth)
    fused = CartPoleGroundRenderer(CartPoleConfig(**cfg))
    plain = CartPoleGroundRenderer(CartPoleConfig(**cfg))
    plain.fuse_poses = False
    states = [ground_states(n, seed).cuda() for seed in (60, 61)]
    _compare_steps(fused, plain, states, f"cartpole ground N={n} {dtype} depth={depth}")


def test_cartpole_reference_scene_step_fused(cuda):
    n = 130
    fused = CartPoleRenderer(CartPoleConfig(num_scenes=n, device="cuda"))
    plain = CartPoleRenderer(CartPoleConfig(num_scenes=n, device="cuda"))
    plain.fuse_poses = False
    states = [ground_states(n, seed).cuda() for seed in (6, 7)]
    _compare_steps(fused, plain, states, "cartpole reference scene")


def test_fused_step_launches_one_kernel(cuda):
    n = 512
    fused = CartPoleGroundRenderer(CartPoleConfig(num_scenes=n, device="cuda"))
    plain = CartPoleGroundRenderer(CartPoleConfig(num_scenes=n, device="cuda"))
    plain.fuse_poses = False
    st = ground_states(n, 8).cuda()
    fused.step(st)
    plain.step(st)
    kf = _kernel_names(lambda: fused.step(st))
    kp = _kernel_names(lambda: plain.step(st))
    print(f"[posed] fused {kf} / two launches {kp}", flush=True)
    if not kp:
        pytest.skip("the profiler recorded no kernels on this platform")
    assert any("pose_kernel" in k for k in kp)
    assert not any("pose_kernel" in k for k in kf)
    assert sum("rows_kernel" in k for k in kf) == 1


class _PosedScene(PBRRenderer):
    """Boxes (two per scene) and an optional low-poly sphere posed from a
    [N, 7 * (ips)] state, over a shared ground quad."""
    fuse_poses = True

    def __init__(self, n, res, sphere: bool, panels: int = 0, boxes: int = 2, **cfg):
        super().__init__(PBRConfig(num_scenes=n, tile_resolution=res, device="cuda",
                                   kernel="rows", **cfg))
        gen = torch.Generator("cpu").manual_seed(40)
        ground = self.add_node("models/quad", instances_per_scene=1, shared_across_scenes=True,
                               model_scale=(40.0, 40.0, 1.0),
                               model_pivot_relative_point=(0.5, 0.5, 0.0))
        ground.set_colors(torch.tensor([[0.3, 0.6, 0.3, 1.0]]))
        self.nbox = boxes
        self.boxes = self.add_node("models/box", instances_per_scene=boxes,
                                   model_pivot_relative_point=(0.5, 0.5, 0.5))
        self.boxes.set_colors(torch.rand(n * boxes, 4, generator=gen))
        self.sphere = None
        if sphere:
            from pybatchrender.renderer import mesh
            self.sphere = self.add_node(mesh.uv_sphere(segments=6, rings=4), instances_per_scene=1)
            self.sphere.set_colors(torch.rand(n, 4, generator=gen))
        self.panels = []  # more posed nodes: unshared quads, one per scene
        for _ in range(panels):
            q = self.add_node("models/quad", instances_per_scene=1,
                              model_pivot_relative_point=(0.5, 0.5, 0.0))
            q.set_colors(torch.rand(n, 4, generator=gen))
            self.panels.append(q)
        self.add_camera(fov_y_deg=70.0, z_near=0.4, z_far=60.0)
        self._pbr_cam.set_positions_and_lookat(torch.zeros(3), torch.tensor([0.2, 1.0, 0.1]))
        self.add_light()
        self.setup_environment()

    def _step(self, st: torch.Tensor) -> None:
        n = self.num_scenes
        b = st[:, :7].reshape(n, 7).repeat_interleave(self.nbox, 0).contiguous()
        for k in range(1, self.nbox):
            b[k::self.nbox, 0] += 1.5 * k
        poses = [(self.boxes, {k: b[:, i] for i, k in enumerate("xyzhprs")})]
        if self.sphere is not None:
            poses.append((self.sphere, dict(x=st[:, 7], y=st[:, 8], z=st[:, 9], s=st[:, 10])))
        for k, q in enumerate(self.panels):
            poses.append((q, dict(x=st[:, 9] + k, y=st[:, 8] + 1.0, p=st[:, 3], h=1.2, s=2.0)))
        self.set_poses(poses)


def _posed_states(n: int, seed: int, near: bool) -> torch.Tensor:
    """Positions in front of the camera (or around it, so boxes cross the
    near plane), angles, scales."""
    g = torch.Generator("cpu").manual_seed(seed)
    st = torch.zeros(n, 11)
    st[:, 0] = torch.empty(n).uniform_(-2.0, 2.0, generator=g)
    st[:, 1] = torch.empty(n).uniform_(0.0, 2.0, generator=g) if near else \
        torch.empty(n).uniform_(4.0, 9.0, generator=g)
    st[:, 2] = torch.empty(n).uniform_(-1.0, 1.0, generator=g)
    st[:, 3:6] = torch.empty(n, 3).uniform_(-3.0, 3.0, generator=g)
    st[:, 6] = torch.empty(n).uniform_(2.0, 5.0, generator=g) if near else \
        torch.empty(n).uniform_(0.5, 1.5, generator=g)
    st[:, 7] = torch.empty(n).uniform_(-1.5, 1.5, generator=g)
    st[:, 8] = torch.empty(n).uniform_(0.5, 6.0, generator=g)
    st[:, 9] = torch.empty(n).uniform_(-1.0, 1.0, generator=g)
    st[:, 10] = torch.empty(n).uniform_(0.5, 2.5, generator=g)
    return st


@pytest.mark.parametrize("sphere,dtype", [(False, "uint8"), (False, "float32"),
                                          (True, "uint8"), (True, "bfloat16"),
                                          (True, "float32")])
@pytest.mark.parametrize("near", [False, True])
def test_posed_boxes_and_spheres(cuda, sphere, dtype, near):
    """The 32-record form (boxes only, 26 slots) and the 64-record smooth form
    (boxes + sphere, 62 slots), with boxes through the near plane (their
    clipped triangles are re-derived from the posed rows read back after the
    pose phase)."""
    n = 33
    fused = _PosedScene(n, (48, 40), sphere, output_dtype=dtype)
    plain = _PosedScene(n, (48, 40), sphere, output_dtype=dtype)
    plain.fuse_poses = False
    states = [_posed_states(n, seed, near).cuda() for seed in (50, 51)]
    label = f"posed sphere={sphere} {dtype} near={near}"
    _compare_steps(fused, plain, states, label)
    # depth and visibility too (the forms that keep records per pixel)
    got = fused._step_and_render((states[0],), {}, depth=True, vis=True)
    check_parity(fused, label + " depth/vis", got=got)


@pytest.mark.parametrize("res", [(48, 48), (96, 40)])
def test_posed_flat_64_record_form(cuda, res):
    """Flat scenes over 32 slots (three boxes per scene + ground = 38) run the
    64-record form, with 4-pixel lanes up to 64 columns and 8-pixel lanes
    past them; near-plane crossings included."""
    n = 19
    fused = _PosedScene(n, res, False, boxes=3)
    plain = _PosedScene(n, res, False, boxes=3)
    plain.fuse_poses = False
    states = [_posed_states(n, 54, True).cuda(), _posed_states(n, 55, False).cuda()]
    _compare_steps(fused, plain, states, f"posed flat 64-record form {res}")
    got = fused._step_and_render((states[0],), {}, depth=True, vis=True)
    check_parity(fused, f"posed flat 64-record form {res} depth/vis", got=got)


@pytest.mark.parametrize("sphere", [False, True])
def test_three_posed_nodes(cuda, sphere):
    """Three poses: one launch in the 64-record form (four waves, one pose
    each; boxes + sphere + a panel + ground = 64 slots), the two launches in
    the 32-record form (two waves; boxes + two panels + ground)."""
    n = 21
    panels = 1 if sphere else 2
    fused = _PosedScene(n, (40, 48), sphere, panels=panels)
    plain = _PosedScene(n, (40, 48), sphere, panels=panels)
    plain.fuse_poses = False
    states = [_posed_states(n, seed, False).cuda() for seed in (52, 53)]
    _compare_steps(fused, plain, states, f"three posed nodes sphere={sphere}")
    names = _kernel_names(lambda: fused.step(states[0]))
    print(f"[posed] three poses, sphere={sphere}: {names}", flush=True)
    if names:
        assert any("pose_kernel" in k for k in names) == (not sphere)


def test_graph_replays_of_fused_step(cuda):
    """A captured fused step replayed with new states, eager two-launch
    steps in between."""
    n = 300
    r = CartPoleGroundRenderer(CartPoleConfig(num_scenes=n, device="cuda"))
    plain = CartPoleGroundRenderer(CartPoleConfig(num_scenes=n, device="cuda"))
    plain.fuse_poses = False
    states = [ground_states(n, seed).cuda() for seed in range(10, 14)]
    g = r.capture_step(states[0], depth=True, vis=True)
    for k, st in enumerate(states + states[:1]):
        got = {key: v.clone() for key, v in g(st).items()}
        plain.step(st, return_pixels=False)
        ref = plain.render(depth=True, vis=True)
        torch.cuda.synchronize()
        for key in ("pixels", "depth", "vis"):
            assert torch.equal(_bits(got[key]), _bits(ref[key])), f"replay {k}: {key}"
        for a, b in zip(_node_buffers(r), _node_buffers(plain)):
            assert torch.equal(_bits(a), _bits(b)), f"replay {k}: node buffer differs"


def _abi_call(r: PBRRenderer, poses: list, fused: bool) -> torch.Tensor:
    """pbr_render_posed (fused=True) or pbr_set_poses + pbr_render with the
    descriptors of `poses` ((node, components) pairs)."""
    desc, keep = r.scene_descriptor()
    N, C, H, W = r.output_shape()
    out = torch.zeros((N, C, H, W), dtype=torch.uint8, device="cuda")
    fo = _abi.pbr_frame_out()
    fo.color, fo.channels, fo.dtype = out.data_ptr(), C, _abi.PBR_DTYPE_U8
    opts = r.render_opts()
    prepared = [node._pose_inputs(**{"x": 0.0, "y": 0.0, "z": 0.0, "h": 0.0, "p": 0.0,
                                     "r": 0.0, "s": 1.0, **kw}) for node, kw in poses]
    descs = [p[0] for p in prepared]
    ws_bytes = _native.workspace_bytes(desc, opts)
    ws = torch.empty(max(ws_bytes, 256), dtype=torch.uint8, device="cuda")
    stream = _native.stream_ptr(r.device)
    if fused:
        _native.render_posed(desc, fo, opts, descs, ws.data_ptr(), ws_bytes, stream)
    else:
        _native.set_poses(descs, stream)
        _native.render(desc, fo, opts, ws.data_ptr(), ws_bytes, stream)
    torch.cuda.synchronize()
    del keep, prepared
    return out


def test_render_posed_cases_it_does_not_fuse(cuda):
    """Poses the row-span launch does not apply itself (a shared node, a
    node of another scene, a pose of fewer rows than the node holds, a node
    drawn with per-scene instance counts, the tile path) take the two
    launches: same frames and rows as calling them."""
    n = 40
    gen = torch.Generator("cpu").manual_seed(41)
    mk = lambda: _PosedScene(n, (32, 32), False)  # noqa: E731
    a, b, other = mk(), mk(), mk()
    ground_a = a._drawn_nodes()[0]
    ground_b = b._drawn_nodes()[0]
    x = (torch.rand(2 * n, generator=gen) * 2 - 1).cuda()
    cases = [
        lambda r, g: [(r.boxes, dict(x=x, y=4.0, s=0.8)), (g, dict(z=-0.5))],   # shared node
        lambda r, g: [(r.boxes, dict(x=x, y=5.0)), (other.boxes, dict(y=5.0, h=0.4))],
    ]
    for k, case in enumerate(cases):
        fa = _abi_call(a, case(a, ground_a), fused=True)
        fb = _abi_call(b, case(b, ground_b), fused=False)
        assert torch.equal(fa, fb), f"case {k}: pixels differ"
        for ta, tb in zip(_node_buffers(a), _node_buffers(b)):
            assert torch.equal(_bits(ta), _bits(tb)), f"case {k}: node buffer differs"
    # a pose descriptor that covers only part of the node's rows
    desc_a = a.boxes._pose_inputs(x=x, y=4.0, z=0.0, h=0.0, p=0.0, r=0.0, s=1.0)
    assert desc_a[0].rows == 2 * n
    da, keep_a = a.scene_descriptor()
    db, keep_b = b.scene_descriptor()
    outs = []
    for r, dd, fused in ((a, da, True), (b, db, False)):
        pd = r.boxes._pose_inputs(x=x, y=4.0, z=0.0, h=0.0, p=0.0, r=0.0, s=1.0)
        pd[0].rows = n  # the first half of the node's rows
        N, C, H, W = r.output_shape()
        out = torch.zeros((N, C, H, W), dtype=torch.uint8, device="cuda")
        fo = _abi.pbr_frame_out()
        fo.color, fo.channels, fo.dtype = out.data_ptr(), C, _abi.PBR_DTYPE_U8
        opts = r.render_opts()
        stream = _native.stream_ptr(r.device)
        if fused:
            _native.render_posed(dd, fo, opts, [pd[0]], 0, 0, stream)
        else:
            _native.set_poses([pd[0]], stream)
            _native.render(dd, fo, opts, 0, 0, stream)
        torch.cuda.synchronize()
        outs.append((out, r.boxes.transforms_b44.clone()))
        del pd
    assert torch.equal(outs[0][0], outs[1][0]) and torch.equal(outs[0][1], outs[1][1])
    del keep_a, keep_b, desc_a
    # a posed node drawn with per-scene instance counts
    cnt = torch.randint(0, 3, (n,), generator=gen)
    c_a, c_b = mk(), mk()
    c_a.boxes.set_instance_counts(cnt)
    c_b.boxes.set_instance_counts(cnt)
    xb = (torch.rand(2 * n, generator=gen) * 2 - 1).cuda()
    f1 = _abi_call(c_a, [(c_a.boxes, dict(x=xb, y=4.5, h=0.7))], fused=True)
    f2 = _abi_call(c_b, [(c_b.boxes, dict(x=xb, y=4.5, h=0.7))], fused=False)
    assert torch.equal(f1, f2) and torch.equal(c_a.boxes.transforms_b44, c_b.boxes.transforms_b44)
    # the tile path: 8 boxes per scene (96 slots)
    t_a = workloads.many_cubes(num_scenes=8, resolution=64, instances=8, seed=3,
                               output_dtype="uint8", output_depth=False)
    t_b = workloads.many_cubes(num_scenes=8, resolution=64, instances=8, seed=3,
                               output_dtype="uint8", output_depth=False)
    na, nb = t_a._drawn_nodes()[0], t_b._drawn_nodes()[0]
    xs = torch.linspace(-1, 1, na.buf_instances, device="cuda")
    f1 = _abi_call(t_a, [(na, dict(x=xs, y=6.0))], fused=True)
    f2 = _abi_call(t_b, [(nb, dict(x=xs, y=6.0))], fused=False)
    assert torch.equal(f1, f2) and torch.equal(na.transforms_b44, nb.transforms_b44)
