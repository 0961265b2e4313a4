"""The C-ABI library loads on a host without a GPU and exports its header."""
import ctypes
import re
from pathlib import Path

import pytest

import pybatchrender as pbr
from pybatchrender import _abi, _native

ROOT = Path(__file__).resolve().parent.parent


def declared_functions() -> set[str]:
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names |= set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(pbr_[a-z_0-9]+)\s*\(", text, re.M))
    return names


def test_header_declares_the_exports():
    assert declared_functions() == set(_abi.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(_native.LIB_PATH))
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_struct_layout_matches_header():
    lib = _native.library()
    _abi.check_layout(lib)
    assert lib.pbr_abi_version() == _abi.PBR_ABI_VERSION


def test_invalid_descriptor_is_rejected_without_gpu():
    """Validation runs before any HIP call, so errors surface on any host."""
    lib = _native.library()
    d = _abi.pbr_scene_desc()
    d.num_scenes, d.width, d.height = 1, 0, 64
    fo = _abi.pbr_frame_out()
    fo.channels, fo.dtype = 3, _abi.PBR_DTYPE_U8
    rc = lib.pbr_render(ctypes.byref(d), ctypes.byref(fo), None, None, 0, None)
    assert rc == _abi.PBR_ERR_INVALID
    assert "resolution" in _native.last_error()
    d.width = 64
    fo.channels = 7
    assert lib.pbr_render(ctypes.byref(d), ctypes.byref(fo), None, None, 0, None) == _abi.PBR_ERR_INVALID
    assert "channels" in _native.last_error()
    fo.channels = 3
    fo.color = d.view_proj = 64  # never dereferenced: validation fails first
    d.num_nodes = 1
    d.num_meshes = 0
    assert lib.pbr_render(ctypes.byref(d), ctypes.byref(fo), None, None, 0, None) == _abi.PBR_ERR_INVALID
    assert "mesh" in _native.last_error()


def test_backward_validation_without_gpu():
    lib = _native.library()
    d = _abi.pbr_scene_desc()
    d.num_scenes, d.width, d.height = 2, 32, 32
    d.view_proj = 64  # never dereferenced: validation fails first
    g = _abi.pbr_grads()
    rc = lib.pbr_render_backward(ctypes.byref(d), None, None, 3, None, None, 0, None)
    assert rc == _abi.PBR_ERR_INVALID and "gradient descriptor" in _native.last_error()
    rc = lib.pbr_render_backward(ctypes.byref(d), None, None, 5, ctypes.byref(g), None, 0, None)
    assert rc == _abi.PBR_ERR_INVALID and "channels" in _native.last_error()
    rc = lib.pbr_render_backward(ctypes.byref(d), None, None, 3, ctypes.byref(g), None, 0, None)
    assert rc == _abi.PBR_ERR_INVALID and "visibility" in _native.last_error()
    g.light = 3  # misaligned
    rc = lib.pbr_render_backward(ctypes.byref(d), 64, 64, 3, ctypes.byref(g), None, 0, None)
    assert rc == _abi.PBR_ERR_INVALID and "aligned" in _native.last_error()
    g.light = None
    rc = lib.pbr_render_backward(ctypes.byref(d), 64, 64, 3, ctypes.byref(g), None, 0, None)
    assert rc == _abi.PBR_ERR_WORKSPACE
    # one partial row of 10 light floats per scene, rounded to 256 bytes
    assert _native.backward_workspace_bytes(d, 3, g) == 256
    # every node flat: one row per persistent workgroup, at most 2048 of them
    d.num_scenes = 5000
    assert _native.backward_workspace_bytes(d, 3, g) == 2048 * 40
    # ... and any workspace holding one row is accepted; less is refused
    rc = lib.pbr_render_backward(ctypes.byref(d), 64, 64, 3, ctypes.byref(g), 256, 39, None)
    assert rc == _abi.PBR_ERR_WORKSPACE and "workspace" in _native.last_error()


def test_workspace_size():
    d = _abi.pbr_scene_desc()
    o = _abi.pbr_render_opts()
    assert _native.workspace_bytes(d, o) == 0
    d.num_scenes, d.width, d.height = 3, 32, 32
    o.persistent = 1
    o.kernel = _abi.PBR_KERNEL_CHUNKED
    assert _native.workspace_bytes(d, o) == 256  # the queue head
    # tile path: header, scene offsets, per-scene visible counts, then per
    # slot box + visible slot id + record + colour + 4 shade floats
    o.kernel = _abi.PBR_KERNEL_TILES
    d.num_nodes, d.num_meshes = 1, 1
    d.meshes[0].num_triangles = 100
    d.nodes[0].instances_per_scene = 2
    d.nodes[0].flat = 1  # flat meshes keep 4 shade floats per record (smooth: 16)
    budget = 3 * 200
    want = 256 + 256 + 256 + sum((b + 255) // 256 * 256 for b in (8 * budget, 4 * budget,
                                                                 64 * budget, 4 * budget,
                                                                 16 * budget))
    assert _native.workspace_bytes(d, o) == want
    o.slot_budget = 50
    assert _native.workspace_bytes(d, o) < want


def test_rows_kernel_needs_no_workspace():
    d = _abi.pbr_scene_desc()
    o = _abi.pbr_render_opts()
    d.num_scenes, d.width, d.height = 4096, 64, 64
    d.num_nodes, d.num_meshes = 1, 1
    d.meshes[0].num_triangles = 12
    d.nodes[0].instances_per_scene = 2
    assert _native.workspace_bytes(d, o) == 0  # AUTO picks the row-span kernel
    o.kernel = _abi.PBR_KERNEL_ROWS
    assert _native.workspace_bytes(d, o) == 0
    o.kernel = _abi.PBR_KERNEL_SCENE
    assert _native.workspace_bytes(d, o) > 0


def test_rows_list_cap_is_validated():
    with pytest.raises(ValueError):
        pbr.PBRConfig(num_scenes=1, device="cpu", row_list_cap=4)
    with pytest.raises(ValueError):
        pbr.PBRConfig(num_scenes=1, device="cpu", row_list_cap=512)
    assert pbr.PBRConfig(num_scenes=1, device="cpu", row_list_cap=8).row_list_cap == 8


def test_kernel_choice_without_gpu():
    """pbr_choose_kernel resolves AUTO on the host (no GPU call)."""
    from pybatchrender.envs.cartpole import CartPoleConfig, CartPoleGroundRenderer
    from pybatchrender import workloads

    r = CartPoleGroundRenderer(CartPoleConfig(num_scenes=8, device="cpu"))
    assert r._chosen_kernel() == "rows"
    r = CartPoleGroundRenderer(CartPoleConfig(num_scenes=8, device="cpu", tile_resolution=(130, 64)))
    assert r._chosen_kernel() == "scene"
    r = CartPoleGroundRenderer(CartPoleConfig(num_scenes=8, device="cpu", persistent=True))
    assert r._chosen_kernel() == "tiles"
    assert workloads.many_cubes(num_scenes=4, device="cpu")._chosen_kernel() == "tiles"
    r = workloads.many_cubes(num_scenes=4, device="cpu", kernel="chunked")
    assert r._chosen_kernel() == "chunked"


def test_slot_budget_follows_in_place_count_edits():
    """Counts written in place through torch change the tile path's pool size
    (the budget is re-derived from the counts, not cached past the edit)."""
    from pybatchrender import workloads

    r = workloads.uneven_batch(num_scenes=16, resolution=32, max_instances=8, seed=1, device="cpu")
    nodes = r._drawn_nodes()
    b0 = r.slot_budget()
    before = [int(n.instance_counts.sum()) for n in nodes]
    nodes[0].instance_counts.fill_(nodes[0].instances_per_scene)
    b1 = r.slot_budget()
    assert b1 - b0 == (16 * nodes[0].instances_per_scene - before[0]) * nodes[0].mesh.num_triangles
    r.scene_descriptor()
    assert r._cached_opts().slot_budget == b1
    nodes[1].instance_counts[:4] = 0
    b2 = r.slot_budget()
    assert b2 < b1
    r.scene_descriptor()
    assert r._cached_opts().slot_budget == b2


def test_render_posed_validation_without_gpu():
    """pbr_render_posed checks the scene, the poses and the options before it
    enqueues anything (pbr_set_poses + pbr_render's errors, in that order)."""
    lib = _native.library()
    d = _abi.pbr_scene_desc()
    d.num_scenes, d.width, d.height = 1, 0, 64
    fo = _abi.pbr_frame_out()
    fo.channels, fo.dtype = 3, _abi.PBR_DTYPE_U8
    rc = lib.pbr_render_posed(ctypes.byref(d), ctypes.byref(fo), None, None, 0, None, 0, None)
    assert rc == _abi.PBR_ERR_INVALID and "resolution" in _native.last_error()
    d.width = 64
    fo.color = d.view_proj = 64  # never dereferenced: validation fails first
    poses = (_abi.pbr_pose_desc * 9)()
    rc = lib.pbr_render_posed(ctypes.byref(d), ctypes.byref(fo), None, poses, 9, None, 0, None)
    assert rc == _abi.PBR_ERR_INVALID and "pose count" in _native.last_error()
    poses[0].rows, poses[0].transforms = 1, 8  # not 16-byte aligned
    rc = lib.pbr_render_posed(ctypes.byref(d), ctypes.byref(fo), None, poses, 1, None, 0, None)
    assert rc == _abi.PBR_ERR_INVALID and "aligned" in _native.last_error()
    opts = _abi.pbr_render_opts()
    opts.kernel = 9
    rc = lib.pbr_render_posed(ctypes.byref(d), ctypes.byref(fo), ctypes.byref(opts), None, 0,
                              None, 0, None)
    assert rc == _abi.PBR_ERR_INVALID and "unknown kernel" in _native.last_error()
