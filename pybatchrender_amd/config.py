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
    tile_size      tile kernels: 16, 32 or 64 pixel square tiles (0 = 32)
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

    def process_resolution(self) -> None:
        """Atlas tiling rules of config.py:61-126 (cols = ceil(sqrt(N)))."""
        if self.tiles is None and self.num_scenes is not None:
            self.tiles = self.num_scenes

        if isinstance(self.tiles, int):
            cols = math.ceil(math.sqrt(self.tiles))
            rows = math.ceil(self.tiles / cols)
            self.tiles = (cols, rows)

        num_nones = sum(
            1 for x in [self.tiles, self.tile_resolution, self.window_resolution] if x is None
        )

        if num_nones == 0:
            if (
                self.window_resolution[0] != self.tiles[0] * self.tile_resolution[0]
                or self.window_resolution[1] != self.tiles[1] * self.tile_resolution[1]
            ):
                raise ValueError(
                    "window_resolution must equal tiles * tile_resolution. "
                    f"Got window_resolution={self.window_resolution}, tiles={self.tiles}, "
                    f"tile_resolution={self.tile_resolution}."
                )
        elif num_nones > 1:
            default_tiles = (1, 1)
            default_tile_resolution = (64, 64)
            if self.tiles is not None:
                self.tile_resolution = default_tile_resolution
                self.window_resolution = (64 * self.tiles[0], 64 * self.tiles[1])
            elif self.tile_resolution is not None:
                self.tiles = default_tiles
                self.window_resolution = (self.tile_resolution[0], self.tile_resolution[1])
            elif self.window_resolution is not None:
                self.tiles = default_tiles
                self.tile_resolution = (self.window_resolution[0], self.window_resolution[1])
            else:
                self.tiles = default_tiles
                self.tile_resolution = default_tile_resolution
                self.window_resolution = (
                    default_tile_resolution[0] * default_tiles[0],
                    default_tile_resolution[1] * default_tiles[1],
                )
        else:
            if self.tiles is None:
                self.tiles = (
                    self.window_resolution[0] // self.tile_resolution[0],
                    self.window_resolution[1] // self.tile_resolution[1],
                )
            elif self.tile_resolution is None:
                self.tile_resolution = (
                    self.window_resolution[0] // self.tiles[0],
                    self.window_resolution[1] // self.tiles[1],
                )
            elif self.window_resolution is None:
                self.window_resolution = (
                    self.tiles[0] * self.tile_resolution[0],
                    self.tiles[1] * self.tile_resolution[1],
                )

        if self.num_scenes is None:
            self.num_scenes = self.tiles[0] * self.tiles[1]
        elif self.num_scenes > self.tiles[0] * self.tiles[1]:
            raise ValueError(f"{self.num_scenes=} can't fit into {self.tiles=}")

        if self.batch_inner_dim is None:
            self.batch_inner_dim = self.tiles[0] * self.tiles[1]
        elif self.batch_inner_dim != self.tiles[0] * self.tiles[1]:
            raise ValueError("batch_inner_dim must equal tiles[0] * tiles[1]")

    def process_device(self) -> None:
        """Device rules of config.py:128-145 ('cuda' is the ROCm GPU)."""
        if self.device is not None and self.device not in ["cpu", "cuda", "mps"]:
            raise ValueError(f"Invalid device: {self.device}")
        if self.device is None:
            self.device = "cuda" if torch.cuda.is_available() else "cpu"
        if self.device == "cuda" and not torch.cuda.is_available():
            raise RuntimeError("device is set to CUDA but CUDA is not available")
        elif self.device == "mps" and not torch.backends.mps.is_available():
            raise RuntimeError("device is set to MPS but MPS is not available")

    def process_output(self) -> None:
        if self.output_dtype not in ("uint8", "bfloat16", "float32"):
            raise ValueError(f"Invalid output_dtype: {self.output_dtype}")
        if not 1 <= int(self.num_channels) <= 4:
            raise ValueError("num_channels must be in [1, 4]")
        if self.region_rows < 0 or self.region_rows % 8:
            raise ValueError("region_rows must be 0 or a positive multiple of 8")
        if self.kernel not in ("auto", "chunked", "scene", "tiles", "rows"):
            raise ValueError(f"Invalid kernel: {self.kernel}")
        if self.tile_size not in (0, 16, 32, 64):
            raise ValueError("tile_size must be 0, 16, 32 or 64")
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
        """Merge rules of config.py:152-174."""
        if isinstance(cfg, cls):
            cfg_dict = asdict(cfg)
            cls = cfg.__class__
        elif isinstance(cfg, dict):
            cfg_dict = dict(cfg)
        else:
            cfg_dict = {}
        if overrides:
            cfg_dict.update(overrides)
        return cls(**cfg_dict)

    def build_prc(self) -> str:
        """PRC text the reference fed Panda3D (config.py:176-191); informational."""
        if self.window_resolution is None:
            self.process_resolution()
        prc = [
            f"window-type {'offscreen' if self.offscreen else 'onscreen'}\n",
            f"win-size {self.window_resolution[0]} {self.window_resolution[1]}\n",
        ]
        prc.append("gl-version 3 2\n" if self.panda3d_backend == "arm" else "threading-model Cull/Draw\n")
        prc.append(self.extra_prc_file_data)
        return "".join(prc)

    def __repr__(self):
        s = f"{type(self).__name__}(\n"
        for key, value in asdict(self).items():
            s += f"    {key}: {value!r}\n"
        return s + ")"
