"""PBRConfig: the reference's configuration dataclass, same fields and rules.

Restates pybatchrender/config.py:16-198.  Atlas tiling (``process_resolution``)
and device choice (``process_device``) keep the reference's semantics because
the projection aspect depends on the atlas (camera.py:154).  Fields that only
configured Panda3D (PRC data, backend, clip_camera, ...) are kept so configs
written for the reference still construct; they have no effect here.

New fields (defaults keep the reference's behaviour):
    output_dtype   'uint8' (reference), 'bfloat16' or 'float32' colour output
    output_depth   also produce a [N, H, W] float32 window-depth tensor
    region_rows    chunked kernel: output rows per work item (0 = automatic)
    persistent     chunked kernel over a persistent (scene, region) work queue
    kernel         'auto', 'rows' (setup and exact per-row spans in one
                   launch: at most 64 triangle slots per scene and tiles at
                   most 128 pixels wide), 'scene' (one wave per scene, at most
                   64 triangle slots per scene), 'tiles' (any triangle count:
                   one setup pass, then one workgroup per scene tile) or
                   'chunked' (any triangle count, setup repeated per row strip)
    tile_size      tile kernels: 8, 16, 32 or 64 pixel square tiles (0 =
                   full-width bands of at most 8192 pixels)
    bin_capacity   tile kernels: entries of each LDS bin of a tile before it
                   is rasterised and emptied, 1024..16384 (0 = 1904)
    row_list_cap   rows kernel: row-list entries per band, 8..256 (0 = 256)
    bin_width      scene kernel: 16, 32 or 64 pixel-wide bins (0 = automatic)
    bins_per_item  scene kernel: bins walked by one work item (0 = automatic)
    lane_pixels    scene kernel: 4 or 8 consecutive pixels per lane (0 = automatic)
"""
from __future__ import annotations

import logging
import math
from dataclasses import asdict, dataclass
from typing import TypeVar

import torch

T = TypeVar("T", bound="PBRConfig")


@dataclass
class PBRConfig:
    offscreen: bool = True

    num_scenes: int | None = None
    tiles: tuple[int, int] | int | None = None
    tile_resolution: tuple[int, int] | None = None
    window_resolution: tuple[int, int] | None = None

    num_channels: int = 3
    batch_inner_dim: int | None = None

    clip_camera: tuple[float, float] = (3.0, 500.0)
    min_objects: int = 50
    max_objects: int = 100
    device: str | None = None
    panda3d_backend: str | None = "arm"
    log_level: int = logging.DEBUG
    extra_prc_file_data: str = (
        "audio-library-name null\n"
        "textures-power-2 none\n"
        "sync-video 0\n"
    )

    render_mode: str = "rgb_array"
    dt: float = 1 / 60
    warmup_steps: int = 5

    report_fps: bool = True
    report_fps_interval: float = 1.0

    manual_camera_control: bool = False
    cuda_gl_interop: bool = True

    interactive: bool = False

    # TorchRL env-related
    direct_obs_dim: int | None = None
    action_n: int | None = None
    action_type: str = "discrete"
    max_steps: int = 500
    auto_reset: bool = True

    num_workers: int = 1

    # MI355X renderer options
    output_dtype: str = "uint8"
    output_depth: bool = False
    region_rows: int = 0
    persistent: bool = False
    kernel: str = "auto"
    bin_width: int = 0
    bins_per_item: int = 0
    lane_pixels: int = 0
    tile_size: int = 0
    row_list_cap: int = 0
    bin_capacity: int = 0

    def process_resolution(self) -> None:
        """Atlas layout (reference config.py:61-126).  An int ``tiles`` (or
        ``num_scenes`` when ``tiles`` is None) becomes a grid of ceil(sqrt(n))
        columns; then ``window_resolution = tiles x tile_resolution`` fills in
        whichever of the three is missing (64 x 64 tiles and a 1 x 1 grid
        when two or more are).  ``num_scenes`` must fit the grid and
        ``batch_inner_dim`` equals its size."""
        if self.tiles is None and self.num_scenes is not None:
            self.tiles = self.num_scenes
        if isinstance(self.tiles, int):
            self.tiles = _grid_for(self.tiles)
        self.tiles, self.tile_resolution, self.window_resolution = _layout(
            self.tiles, self.tile_resolution, self.window_resolution)
        cells = self.tiles[0] * self.tiles[1]
        if self.num_scenes is None:
            self.num_scenes = cells
        elif self.num_scenes > cells:
            raise ValueError(f"{self.num_scenes} scenes do not fit a {self.tiles[0]} x "
                             f"{self.tiles[1]} atlas")
        if self.batch_inner_dim is None:
            self.batch_inner_dim = cells
        elif self.batch_inner_dim != cells:
            raise ValueError(f"batch_inner_dim {self.batch_inner_dim} differs from the atlas "
                             f"size {cells}")

    def process_device(self) -> None:
        """Device rules (reference config.py:128-145): 'cpu', 'cuda' (the
        ROCm GPU here) or 'mps'; None picks the GPU when one is present."""
        if self.device is None:
            self.device = "cuda" if torch.cuda.is_available() else "cpu"
            return
        present = {"cpu": lambda: True, "cuda": torch.cuda.is_available,
                   "mps": torch.backends.mps.is_available}
        if self.device not in present:
            raise ValueError(f"device must be one of {sorted(present)}, not {self.device!r}")
        if not present[self.device]():
            raise RuntimeError(f"device {self.device!r} was asked for but is not available")

    def process_output(self) -> None:
        if self.output_dtype not in ("uint8", "bfloat16", "float32"):
            raise ValueError(f"Invalid output_dtype: {self.output_dtype}")
        if not 1 <= int(self.num_channels) <= 4:
            raise ValueError("num_channels must be in [1, 4]")
        if self.region_rows < 0 or self.region_rows % 8:
            raise ValueError("region_rows must be 0 or a positive multiple of 8")
        if self.kernel not in ("auto", "chunked", "scene", "tiles", "rows"):
            raise ValueError(f"Invalid kernel: {self.kernel}")
        if self.tile_size not in (0, 8, 16, 32, 64):
            raise ValueError("tile_size must be 0, 8, 16, 32 or 64")
        if self.bin_capacity != 0 and not 1024 <= self.bin_capacity <= 16384:
            raise ValueError("bin_capacity must be 0 or within [1024, 16384]")
        if self.row_list_cap != 0 and not 8 <= self.row_list_cap <= 256:
            raise ValueError("row_list_cap must be 0 or within [8, 256]")
        if self.bin_width not in (0, 16, 32, 64):
            raise ValueError("bin_width must be 0, 16, 32 or 64")
        if self.lane_pixels not in (0, 4, 8):
            raise ValueError("lane_pixels must be 0, 4 or 8")

    def __post_init__(self) -> None:
        self.process_resolution()
        self.process_device()
        self.process_output()

    @classmethod
    def from_config(cls: type[T], cfg: T | dict | None = None, **overrides) -> T:
        """A config from another one of this class (its own class is kept,
        so subclasses survive), from a dict of fields, or from defaults;
        keyword overrides win (reference config.py:152-174)."""
        if isinstance(cfg, cls):
            target, fields = type(cfg), asdict(cfg)
        else:
            target, fields = cls, (dict(cfg) if isinstance(cfg, dict) else {})
        return target(**{**fields, **overrides})


_DEFAULT_TILE = (64, 64)


def _grid_for(n: int) -> tuple[int, int]:
    """Columns and rows of the smallest near-square grid holding n cells."""
    cols = math.ceil(math.sqrt(n))
    return cols, math.ceil(n / cols)


def _layout(tiles, tile, window):
    """(tiles, tile_resolution, window_resolution) with window = tiles x tile,
    filling in what is missing."""
    given = [v is not None for v in (tiles, tile, window)]
    mul = lambda a, b: (a[0] * b[0], a[1] * b[1])  # noqa: E731
    div = lambda a, b: (a[0] // b[0], a[1] // b[1])  # noqa: E731
    if all(given):
        if tuple(window) != mul(tiles, tile):
            raise ValueError(f"window_resolution {tuple(window)} is not tiles {tuple(tiles)} "
                             f"times tile_resolution {tuple(tile)}")
        return tiles, tile, window
    if sum(given) == 2:
        if window is None:
            return tiles, tile, mul(tiles, tile)
        if tile is None:
            return tiles, div(window, tiles), window
        return div(window, tile), tile, window
    if tiles is not None:
        return tiles, _DEFAULT_TILE, mul(tiles, _DEFAULT_TILE)
    one = tuple(tile if tile is not None else (window if window is not None else _DEFAULT_TILE))
    return (1, 1), one, one
