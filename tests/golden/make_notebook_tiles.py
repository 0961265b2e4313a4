"""Extract the reference's own rendered CartPole tiles from its notebook.

Source: examples/notebooks/cartpole_benchmark.ipynb of the reference,
cell "Visualize Initial State" (output PNG) and cell "Visualize Final State".
Each PNG is a 768x768 nearest-neighbour x3 upscale of a 4x4 grid of 64x64
tiles built by env.py:_make_grid_frame (scene k at grid cell (k // 4, k % 4)).

The render setup printed by the notebook: N = 4098 scenes, tiles (65, 64),
window 4160 x 4096, device cuda, NVIDIA L4, panda3d 1.10.16.  The states of
scenes 0..2 at reset are printed to 4 decimals (x, x_dot, theta, theta_dot).

Writes tests/golden/notebook_tiles.npz (uint8 [16, 64, 64, 3] for the
initial and the final grid) and tests/golden/notebook_tiles.json (metadata).
Run once in a container that has the reference mounted; the outputs are
committed so tests never need the reference.
"""
from __future__ import annotations

import base64
import io
import json
import sys
from pathlib import Path

import numpy as np
from PIL import Image

NOTEBOOK = Path("/root/reference/examples/notebooks/cartpole_benchmark.ipynb")
OUT = Path(__file__).resolve().parent


def grid_from_cell(nb: dict, cell: int) -> np.ndarray:
    for o in nb["cells"][cell].get("outputs", []):
        data = o.get("data", {})
        if "image/png" in data:
            raw = data["image/png"]
            raw = raw if isinstance(raw, str) else "".join(raw)
            img = np.array(Image.open(io.BytesIO(base64.b64decode(raw))).convert("RGB"))
            small = img[::3, ::3]
            up = np.repeat(np.repeat(small, 3, 0), 3, 1)
            if not (up == img).all():
                raise ValueError(f"cell {cell}: image is not an exact x3 upscale")
            return small
    raise ValueError(f"cell {cell} has no PNG output")


def tiles(grid: np.ndarray) -> np.ndarray:
    out = np.zeros((16, 64, 64, 3), dtype=np.uint8)
    for k in range(16):
        r, c = k // 4, k % 4
        out[k] = grid[r * 64:(r + 1) * 64, c * 64:(c + 1) * 64]
    return out


def main() -> int:
    nb = json.loads(NOTEBOOK.read_text())
    init = tiles(grid_from_cell(nb, 14))
    final = tiles(grid_from_cell(nb, 20))
    np.savez_compressed(OUT / "notebook_tiles.npz", initial=init, final=final)
    meta = {
        "source": "examples/notebooks/cartpole_benchmark.ipynb cells 14 (initial) and 20 (final)",
        "num_scenes": 4098,
        "tiles": [65, 64],
        "tile_resolution": [64, 64],
        "window_resolution": [4160, 4096],
        "background_uint8": [105, 105, 105],
        "initial_states_printed": {
            "0": [-0.4038, 0.9445, -0.5032, 0.1505],
            "1": [0.0667, 0.5819, 0.2862, -0.1914],
            "2": [-1.9003, -0.0619, 0.3823, -0.1472],
        },
        "state_source": "notebook cell 12 output, 4 decimals",
    }
    (OUT / "notebook_tiles.json").write_text(json.dumps(meta, indent=2) + "\n")
    print("wrote", OUT / "notebook_tiles.npz")
    return 0


if __name__ == "__main__":
    sys.exit(main())
