"""Graph capture of the tile path, per-scene instance counts edited in place,
and the memory a captured graph owns.

  * the tile path's record pool follows counts written in place through
    torch (eager renders stay bit-exact against the oracle);
  * counts written behind torch's back (an alias outside its version
    counter) overflow the pool: the frame drops the slots past it, and
    last_render_overflowed() reports it;
  * the persistent tile render of an uneven batch captured as a hipGraph:
    replays with transforms and (smaller) counts written in place between
    them equal eager renders and the oracle; counts grown past the captured
    total raise instead of dropping triangles;
  * a captured graph owns its workspace and keeps the tensors it reads: a
    destroyed renderer and reallocated memory do not change its replays.
"""
from __future__ import annotations

import pytest
import torch

from pybatchrender import workloads
from tests.test_gpu_parity import _bits, check_parity

pytestmark = pytest.mark.gpu


def _uneven(n: int, persistent: bool, seed: int = 3):
    return workloads.uneven_batch(num_scenes=n, resolution=64, max_instances=24, seed=seed,
                                  kernel="tiles", persistent=persistent)


@pytest.mark.parametrize("persistent", [False, True])
def test_counts_grown_in_place_resize_the_pool(cuda, persistent):
    r = _uneven(96, persistent)
    check_parity(r, "uneven before the edit")
    node = r._drawn_nodes()[1]
    before = r.slot_budget()
    node.instance_counts.copy_(torch.full_like(node.instance_counts, node.instances_per_scene))
    assert r.slot_budget() > before
    check_parity(r, "counts grown in place")
    assert not r.last_render_overflowed()
    node.instance_counts[::2] = 0  # shrink part of it again, in place
    check_parity(r, "counts shrunk in place")


def test_overflow_flag_reports_counts_written_behind_torch(cuda):
    r = _uneven(64, persistent=True)
    check_parity(r, "uneven")
    assert not r.last_render_overflowed()
    node = r._drawn_nodes()[2]
    # an alias of the counts storage with its own version counter: writes
    # through it are invisible to the renderer's budget bookkeeping
    alias = torch.empty(0, dtype=torch.int32, device="cuda").set_(
        node.instance_counts.untyped_storage(), 0, node.instance_counts.shape)
    alias.fill_(node.instances_per_scene)
    r.render()
    torch.cuda.synchronize()
    assert r.last_render_overflowed()
    # the supported way: set_instance_counts (or an in-place torch write)
    node.set_instance_counts(alias.clone())
    check_parity(r, "counts set again")
    assert not r.last_render_overflowed()


def test_graphed_persistent_tile_render_replays(cuda):
    """The uneven batch's persistent (scene, tile) render in a hipGraph."""
    r = _uneven(160, persistent=True, seed=5)
    g = r.capture_step(depth=True, vis=True)
    gen = torch.Generator("cpu").manual_seed(9)
    nodes = r._drawn_nodes()
    full = [n.instance_counts.clone() for n in nodes]
    for k in range(4):
        for n, f in zip(nodes, full):
            M = workloads.random_transforms(n.transforms_b44.shape[0], gen)
            n.transforms_b44.copy_(M.to(n.transforms_b44.device))
            # in place, never above the counts at capture
            n.instance_counts.copy_(
                (f.cpu() - torch.randint(0, 3, f.shape, generator=gen)).clamp(min=0).to(f.device))
        got = {key: v.clone() for key, v in g().items()}
        torch.cuda.synchronize()
        assert not g.overflowed()
        out = check_parity(r, f"graph replay {k} of the persistent tile render")
        for key in ("pixels", "depth", "vis"):
            assert torch.equal(_bits(got[key]), _bits(out["got"][key])), f"replay {k}: {key}"
    for n, f in zip(nodes, full):
        n.instance_counts.copy_(f)
    g()  # back at the captured total: fine
    nodes[0].instance_counts.fill_(nodes[0].instances_per_scene)  # past it
    with pytest.raises(RuntimeError, match="capture the step again"):
        g()


def test_graph_owns_its_memory(cuda):
    r = _uneven(80, persistent=True, seed=7)
    g = r.capture_step(depth=True, vis=True)
    first = {key: v.clone() for key, v in g().items()}
    torch.cuda.synchronize()
    r.destroy()  # drops the renderer's workspace and descriptor cache
    junk = [torch.full((1 << 20,), 0x7f, dtype=torch.uint8, device="cuda") for _ in range(64)]
    for _ in range(3):
        again = g()
        torch.cuda.synchronize()
        for key in ("pixels", "depth", "vis"):
            assert torch.equal(_bits(again[key]), _bits(first[key])), key
    del junk
