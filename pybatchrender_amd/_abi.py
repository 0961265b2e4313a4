"""ctypes mirror of include/pbr_abi.h (the C ABI of libpbr_hip.so).

The structs here must match the header byte for byte; ``check_layout``
compares them with the facts the library reports through ``pbr_abi_layout``.
"""
from __future__ import annotations

import ctypes as C

PBR_ABI_VERSION = 4
PBR_MAX_NODES = 16
PBR_MAX_MESHES = 16
PBR_MAX_RESOLUTION = 4096
PBR_NOALLOC_SCRATCH_BYTES = 64 << 20

PBR_OK = 0
PBR_ERR_INVALID = 1
PBR_ERR_LAUNCH = 2
PBR_ERR_WORKSPACE = 3
PBR_ERR_UNSUPPORTED = 4

PBR_DTYPE_U8 = 0
PBR_DTYPE_BF16 = 1
PBR_DTYPE_F32 = 2


class pbr_mesh(C.Structure):
    _fields_ = [
        ("vertex_offset", C.c_int32),
        ("num_vertices", C.c_int32),
        ("index_offset", C.c_int32),
        ("num_triangles", C.c_int32),
    ]


class pbr_node(C.Structure):
    _fields_ = [
        ("transforms", C.c_void_p),
        ("colors", C.c_void_p),
        ("counts", C.c_void_p),
        ("mesh", C.c_int32),
        ("instances_per_scene", C.c_int32),
        ("shared", C.c_int32),
        ("flat", C.c_int32),
    ]


class pbr_scene_desc(C.Structure):
    _fields_ = [
        ("num_scenes", C.c_int32),
        ("width", C.c_int32),
        ("height", C.c_int32),
        ("num_nodes", C.c_int32),
        ("num_meshes", C.c_int32),
        ("reserved0", C.c_int32),
        ("vertices", C.c_void_p),
        ("indices", C.c_void_p),
        ("view_proj", C.c_void_p),
        ("light_dir", C.c_float * 3),
        ("light_col", C.c_float * 3),
        ("ambient", C.c_float * 3),
        ("strength", C.c_float),
        ("background", C.c_float * 4),
        ("meshes", pbr_mesh * PBR_MAX_MESHES),
        ("nodes", pbr_node * PBR_MAX_NODES),
    ]


class pbr_frame_out(C.Structure):
    _fields_ = [
        ("color", C.c_void_p),
        ("depth", C.c_void_p),
        ("vis", C.c_void_p),
        ("channels", C.c_int32),
        ("dtype", C.c_int32),
    ]


PBR_KERNEL_AUTO = 0
PBR_KERNEL_CHUNKED = 1
PBR_KERNEL_SCENE = 2
PBR_KERNEL_TILES = 3
PBR_KERNEL_ROWS = 4


class pbr_render_opts(C.Structure):
    _fields_ = [
        ("region_rows", C.c_int32),
        ("persistent", C.c_int32),
        ("workgroups", C.c_int32),
        ("kernel", C.c_int32),
        ("bin_width", C.c_int32),
        ("bins_per_item", C.c_int32),
        ("lane_pixels", C.c_int32),
        ("tile_size", C.c_int32),
        ("slot_budget", C.c_int64),
        ("row_list_cap", C.c_int32),
        ("bin_capacity", C.c_int32),
    ]


class pbr_pose_component(C.Structure):
    _fields_ = [
        ("src", C.c_void_p),
        ("stride", C.c_int64),
        ("value", C.c_float),
        ("reserved", C.c_int32),
    ]


class pbr_pose_desc(C.Structure):
    _fields_ = [
        ("transforms", C.c_void_p),
        ("rot3", C.c_void_p),
        ("scale", C.c_void_p),
        ("rows", C.c_int32),
        ("reserved", C.c_int32),
        ("x", pbr_pose_component),
        ("y", pbr_pose_component),
        ("z", pbr_pose_component),
        ("h", pbr_pose_component),
        ("p", pbr_pose_component),
        ("r", pbr_pose_component),
        ("s", pbr_pose_component),
    ]


class pbr_grads(C.Structure):
    _fields_ = [
        ("colors", C.c_void_p * PBR_MAX_NODES),
        ("vertices", C.c_void_p),
        ("light", C.c_void_p),
    ]


class pbr_cartpole_params(C.Structure):
    _fields_ = [
        ("gravity", C.c_float),
        ("masscart", C.c_float),
        ("masspole", C.c_float),
        ("length", C.c_float),
        ("force_mag", C.c_float),
        ("tau", C.c_float),
        ("theta_threshold", C.c_float),
        ("x_threshold", C.c_float),
        ("max_steps", C.c_int32),
        ("auto_reset", C.c_int32),
        ("init_lo", C.c_float * 4),
        ("init_hi", C.c_float * 4),
        ("seed", C.c_uint64),
    ]


class pbr_cartpole_batch(C.Structure):
    _fields_ = [
        ("state", C.c_void_p),
        ("next_raw", C.c_void_p),
        ("step_count", C.c_void_p),
        ("action", C.c_void_p),
        ("reward", C.c_void_p),
        ("done", C.c_void_p),
        ("rng_counter", C.c_void_p),
        ("num_scenes", C.c_int32),
        ("reserved", C.c_int32),
    ]


LIGHT_GRAD_FIELDS =("light_dir", "light_dir", "light_dir", "light_col", "light_col", "light_col",
                     "ambient", "ambient", "ambient", "strength")


def expected_layout() -> list[int]:
    """The facts ``pbr_abi_layout`` reports, computed from the mirror."""
    return [
        C.sizeof(pbr_mesh),
        C.sizeof(pbr_node),
        pbr_node.mesh.offset,
        pbr_node.flat.offset,
        C.sizeof(pbr_scene_desc),
        pbr_scene_desc.vertices.offset,
        pbr_scene_desc.light_dir.offset,
        pbr_scene_desc.background.offset,
        pbr_scene_desc.meshes.offset,
        pbr_scene_desc.nodes.offset,
        C.sizeof(pbr_frame_out),
        pbr_frame_out.channels.offset,
        C.sizeof(pbr_render_opts),
        pbr_render_opts.kernel.offset,
        pbr_render_opts.bins_per_item.offset,
        pbr_render_opts.slot_budget.offset,
        pbr_render_opts.row_list_cap.offset,
        pbr_render_opts.bin_capacity.offset,
        C.sizeof(pbr_pose_component),
        C.sizeof(pbr_pose_desc),
        pbr_pose_desc.x.offset,
        pbr_pose_desc.s.offset,
        C.sizeof(pbr_grads),
        pbr_grads.vertices.offset,
        C.sizeof(pbr_cartpole_params),
        pbr_cartpole_params.seed.offset,
        C.sizeof(pbr_cartpole_batch),
        pbr_cartpole_batch.num_scenes.offset,
    ]


# Exported symbols of libpbr_hip.so, as declared in include/pbr_abi.h.
EXPORTS = (
    "pbr_render",
    "pbr_render_noalloc",
    "pbr_render_backward",
    "pbr_backward_workspace_bytes",
    "pbr_set_pose",
    "pbr_set_poses",
    "pbr_render_posed",
    "pbr_cartpole_step",
    "pbr_cartpole_reset",
    "pbr_workspace_bytes",
    "pbr_choose_kernel",
    "pbr_last_error",
    "pbr_abi_version",
    "pbr_abi_layout",
)


def declare(lib: C.CDLL) -> C.CDLL:
    """Attach argtypes/restype to the library's entry points."""
    lib.pbr_render.argtypes = [
        C.POINTER(pbr_scene_desc),
        C.POINTER(pbr_frame_out),
        C.POINTER(pbr_render_opts),
        C.c_void_p,
        C.c_size_t,
        C.c_void_p,
    ]
    lib.pbr_render.restype = C.c_int
    lib.pbr_render_noalloc.argtypes = [
        C.POINTER(pbr_scene_desc),
        C.POINTER(pbr_frame_out),
        C.POINTER(pbr_render_opts),
        C.c_void_p,
    ]
    lib.pbr_render_noalloc.restype = C.c_int
    lib.pbr_render_backward.argtypes = [
        C.POINTER(pbr_scene_desc),
        C.c_void_p,
        C.c_void_p,
        C.c_int32,
        C.POINTER(pbr_grads),
        C.c_void_p,
        C.c_size_t,
        C.c_void_p,
    ]
    lib.pbr_render_backward.restype = C.c_int
    lib.pbr_backward_workspace_bytes.argtypes = [
        C.POINTER(pbr_scene_desc),
        C.c_int32,
        C.POINTER(pbr_grads),
        C.POINTER(C.c_size_t),
    ]
    lib.pbr_backward_workspace_bytes.restype = C.c_int
    lib.pbr_set_pose.argtypes = [C.POINTER(pbr_pose_desc), C.c_void_p]
    lib.pbr_set_pose.restype = C.c_int
    lib.pbr_set_poses.argtypes = [C.POINTER(pbr_pose_desc), C.c_int32, C.c_void_p]
    lib.pbr_set_poses.restype = C.c_int
    lib.pbr_render_posed.argtypes = [
        C.POINTER(pbr_scene_desc),
        C.POINTER(pbr_frame_out),
        C.POINTER(pbr_render_opts),
        C.POINTER(pbr_pose_desc),
        C.c_int32,
        C.c_void_p,
        C.c_size_t,
        C.c_void_p,
    ]
    lib.pbr_render_posed.restype = C.c_int
    lib.pbr_cartpole_step.argtypes = [C.POINTER(pbr_cartpole_params), C.POINTER(pbr_cartpole_batch),
                                      C.POINTER(pbr_pose_desc), C.c_int32, C.c_void_p]
    lib.pbr_cartpole_step.restype = C.c_int
    lib.pbr_cartpole_reset.argtypes = [C.POINTER(pbr_cartpole_params),
                                       C.POINTER(pbr_cartpole_batch), C.c_void_p]
    lib.pbr_cartpole_reset.restype = C.c_int
    lib.pbr_workspace_bytes.argtypes = [
        C.POINTER(pbr_scene_desc),
        C.POINTER(pbr_render_opts),
        C.POINTER(C.c_size_t),
    ]
    lib.pbr_workspace_bytes.restype = C.c_int
    lib.pbr_choose_kernel.argtypes = [
        C.POINTER(pbr_scene_desc),
        C.POINTER(pbr_render_opts),
        C.POINTER(C.c_int32),
    ]
    lib.pbr_choose_kernel.restype = C.c_int
    lib.pbr_last_error.argtypes = []
    lib.pbr_last_error.restype = C.c_char_p
    lib.pbr_abi_version.argtypes = []
    lib.pbr_abi_version.restype = C.c_int
    lib.pbr_abi_layout.argtypes = [C.POINTER(C.c_int64), C.c_int32]
    lib.pbr_abi_layout.restype = C.c_int
    return lib


def check_layout(lib: C.CDLL) -> None:
    """Raise if the library's struct layout differs from this mirror."""
    n = lib.pbr_abi_layout(None, 0)
    buf = (C.c_int64 * n)()
    lib.pbr_abi_layout(buf, n)
    got = list(buf)
    want = expected_layout()
    if got != want:
        raise RuntimeError(f"pbr ABI layout mismatch: library {got}, binding {want}")
    if lib.pbr_abi_version() != PBR_ABI_VERSION:
        raise RuntimeError("pbr ABI version mismatch")
