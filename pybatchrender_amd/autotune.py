"""Autotuner for the render launch shape.

The reference has a single GL path; the HIP renderer has several launch
shapes whose best choice depends on the batch (scenes, resolution, triangle
counts, how uneven they are, output dtype).  ``autotune(renderer)`` times the
candidate shapes on the renderer's current scenes, keeps the fastest in the
renderer's config and returns the measurements:

  row-span kernel (<= 64 triangle slots per scene, <= 128 columns)
  scene kernel (<= 64 triangle slots per scene): bin width 16 / 32 / 64,
      4 or 8 pixels per lane, bins per work item 0 (auto) / 1 / 2 / 4
  tile kernels: tile 0 (full-width bands) / 8 / 16 / 32 / 64 pixels (8 only
      for scenes of at most 2048 triangle slots: every tile tests every
      slot's box), LDS bin capacity 1024 / 1904 / 4096 entries, and the
      persistent (scene, tile) queue at 4 / 8 / 16 workgroups per CU

Results are cached per batch signature in memory (and in ``cache_path`` as
JSON when one is given), so a later renderer with the same signature reuses
the winner without timing again.  Every candidate renders the same image
(the kernels are bit-identical), so tuning never changes the output.
"""
from __future__ import annotations

import json
import time
from pathlib import Path

import torch

_FIELDS = ("kernel", "tile_size", "persistent", "bin_width", "bins_per_item", "lane_pixels",
           "bin_capacity")
_CACHE: dict[str, dict] = {}


def signature(renderer) -> str:
    """Batch facts the best launch shape depends on."""
    nodes = renderer._drawn_nodes()
    tris = [(n.instances_per_scene, n.mesh.num_triangles, n.instance_counts is not None,
             bool(n.shared_across), bool(n.flat)) for n in nodes]
    W, H = renderer.cfg.tile_resolution
    return json.dumps({"N": renderer.num_scenes, "W": int(W), "H": int(H),
                       "dtype": renderer.cfg.output_dtype, "C": int(renderer.cfg.num_channels),
                       "depth": bool(renderer.cfg.output_depth), "nodes": tris,
                       "budget": renderer.slot_budget()}, sort_keys=True)


def max_slots(renderer) -> int:
    return sum(n.instances_per_scene * n.mesh.num_triangles for n in renderer._drawn_nodes())


def candidates(renderer) -> list[dict]:
    """Launch shapes valid for this batch."""
    out = []
    W = int(renderer.cfg.tile_resolution[0])
    if max_slots(renderer) <= 64 and W <= 128:
        out.append({"kernel": "rows"})
    if max_slots(renderer) <= 64:
        for bw in (16, 32, 64):
            for px in (4, 8):
                for per in (0, 1, 2, 4):
                    out.append({"kernel": "scene", "bin_width": bw, "lane_pixels": px,
                                "bins_per_item": per})
    sizes = (0, 8, 16, 32, 64) if max_slots(renderer) <= 2048 else (0, 16, 32, 64)
    for ts in sizes:
        for cap in (1024, 0, 4096):
            out.append({"kernel": "tiles", "tile_size": ts, "persistent": False,
                        "bin_capacity": cap})
        for wpc in (4, 8, 16):
            out.append({"kernel": "tiles", "tile_size": ts, "persistent": True,
                        "workgroups_per_cu": wpc})
    return out


def _apply(renderer, cand: dict) -> None:
    cfg = renderer.cfg
    cfg.kernel = cand.get("kernel", "auto")
    cfg.tile_size = int(cand.get("tile_size", 0))
    cfg.persistent = bool(cand.get("persistent", False))
    cfg.bin_width = int(cand.get("bin_width", 0))
    cfg.bins_per_item = int(cand.get("bins_per_item", 0))
    cfg.lane_pixels = int(cand.get("lane_pixels", 0))
    cfg.bin_capacity = int(cand.get("bin_capacity", 0))
    wpc = int(cand.get("workgroups_per_cu", 0))
    if wpc:
        cus = torch.cuda.get_device_properties(renderer.device).multi_processor_count
        renderer.workgroups = wpc * cus
    else:
        renderer.workgroups = 0


def _time(renderer, out, depth, launches: int) -> float:
    renderer.render(out=out, depth=depth)
    torch.cuda.synchronize()
    s = torch.cuda.current_stream(renderer.device)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(launches):
        renderer.render(out=out, depth=depth)
    e1.record(s)
    e1.synchronize()
    return e0.elapsed_time(e1) * 1000.0 / launches


def autotune(renderer, launches: int = 5, cache_path: str | Path | None = None,
             cands: list[dict] | None = None, verbose: bool = False) -> dict:
    """Pick the fastest launch shape for the renderer's current batch and
    keep it in ``renderer.cfg``.  Returns {'best', 'us', 'results', 'cached'}."""
    if renderer.device.type != "cuda":
        raise RuntimeError("autotune needs the HIP renderer on a GPU")
    key = signature(renderer)
    disk = {}
    if cache_path is not None and Path(cache_path).exists():
        disk = json.loads(Path(cache_path).read_text())
    hit = _CACHE.get(key) or disk.get(key)
    if hit is not None:
        _apply(renderer, hit["best"])
        return {**hit, "cached": True}
    N, C, H, W = renderer.output_shape()
    from .renderer.renderer import _DTYPES
    out = torch.empty((N, C, H, W), dtype=_DTYPES[renderer.cfg.output_dtype][0],
                      device=renderer.device)
    depth = (torch.empty((N, H, W), dtype=torch.float32, device=renderer.device)
             if renderer.cfg.output_depth else False)
    saved = {f: getattr(renderer.cfg, f) for f in _FIELDS}
    saved_wg = getattr(renderer, "workgroups", 0)
    results = []
    for cand in (cands if cands is not None else candidates(renderer)):
        _apply(renderer, cand)
        t0 = time.perf_counter()
        us = _time(renderer, out, depth, launches)
        results.append({"cand": cand, "us": round(us, 2)})
        if verbose:
            print(f"[autotune] {cand} {us:.1f} us ({time.perf_counter() - t0:.2f} s)", flush=True)
    for f, v in saved.items():
        setattr(renderer.cfg, f, v)
    renderer.workgroups = saved_wg
    best = min(results, key=lambda r: r["us"])
    rec = {"best": best["cand"], "us": best["us"], "results": results}
    _CACHE[key] = rec
    if cache_path is not None:
        disk[key] = rec
        Path(cache_path).write_text(json.dumps(disk, indent=1))
    _apply(renderer, best["cand"])
    renderer._workspace = None  # the chosen shape may need a different workspace size
    return {**rec, "cached": False}
