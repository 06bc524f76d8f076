"""Ruler spawn (the reference's `spawn`, sparse_ruling_set.cpp:186-231):
results with spawn are bit-exact against the oracle and identical to the
ranking without it, on adversarial inputs (very long gaps between rulers,
spacing up to 4096; forests of short lists; locality-biased lists), in every
launch mode; the per-level trace shows the spawned rulers and a shorter
longest chain."""
import numpy as np
import pytest

import oracle_lib as o
from test_gpu_rank import dev, host, random_forest_lists, say

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("spacing", [64, 512, 4096])
@pytest.mark.parametrize("n", [100003, 1 << 21])
@pytest.mark.parametrize("launch", ["auto", "device", "host"])
def test_spawn_long_gaps(ctx, spacing, n, launch):
    say("spawn long gaps", spacing, n, launch)
    succ, w = o.gen_list(n, 17 + spacing, True)
    root, rank = ctx.rank_list(dev(succ), dev(w), ruler_spacing=spacing, spawn="on", launch=launch)
    er, ek = o.list_closed_form(n, 17 + spacing, w)
    assert (host(root) == er).all() and (host(rank) == ek).all()


@pytest.mark.parametrize("nlists", [2, 300, 20000])
@pytest.mark.parametrize("weighted", [True, False])
def test_spawn_forests(ctx, nlists, weighted):
    say("spawn forests", nlists, weighted)
    n = 400009
    succ, w = random_forest_lists(n, nlists, nlists + 3, weighted=weighted)
    cr, ck = o.sequential_walk(succ, w)
    for spacing in (0, 32, 1024):
        root, rank = ctx.rank_list(dev(succ), dev(w), ruler_spacing=spacing, spawn="on")
        assert (host(root) == cr).all() and (host(rank) == ck).all(), spacing


def test_spawn_locality_biased(ctx):
    n = 1 << 20
    say("spawn local", n)
    succ, w = o.gen_list(n, 5, True, local=True)
    root, rank = ctx.rank_list(dev(succ), dev(w), ruler_spacing=256, spawn="on")
    er, ek = o.sequential_walk(succ, w)
    assert (host(root) == er).all() and (host(rank) == ek).all()


def test_spawn_segmented(ctx):
    """A segmented batch whose levels spawn (forced)."""
    import torch
    rng = np.random.default_rng(3)
    sizes = np.floor(2.0 ** rng.uniform(8, 16, size=300)).astype(np.uint64)
    off = np.zeros(len(sizes) + 1, np.uint64)
    off[1:] = np.cumsum(sizes)
    seeds = rng.integers(1, 1 << 62, size=len(sizes), dtype=np.uint64)
    say("spawn segmented", int(off[-1]))
    d_off = torch.from_numpy(off).cuda()
    succ, w = ctx.gen_lists_segmented(d_off, torch.from_numpy(seeds).cuda(), None, weighted=True)
    r0, k0 = ctx.rank_lists_segmented(succ, d_off, w, spawn="off")
    r1, k1 = ctx.rank_lists_segmented(succ, d_off, w, spawn="on", ruler_spacing=64)
    assert torch.equal(r0.view(torch.int32), r1.view(torch.int32)) and torch.equal(k0, k1)
    # every segment against the oracle's walk
    hs, hw, hr, hk = host(succ), host(w), host(r1), host(k1)
    for s in range(0, len(sizes), 37):
        a, b = int(off[s]), int(off[s + 1])
        cr, ck = o.sequential_walk(hs[a:b].astype(np.uint64), hw[a:b])
        assert (hr[a:b] == cr).all() and (hk[a:b] == ck).all()


def test_spawn_trace_and_default(ctx):
    """The trace reports the spawned rulers; spawn shortens the longest walk
    of a level with very long gaps; and the default (auto) spawns on a
    large list with 256 or more elements per ruler."""
    n = 1 << 22
    say("spawn trace", n)
    succ, w = ctx.gen_list(n, 44, weighted=True)
    out = {}
    for mode in ("off", "on"):
        ctx.trace(True)
        root, rank = ctx.rank_list(succ, w, ruler_spacing=2048, spawn=mode)
        tr = ctx.trace_read()
        ctx.trace(False)
        out[mode] = (tr[0], host(root), host(rank))
    t_off, t_on = out["off"][0], out["on"][0]
    assert t_off["spawned"] == 0 and t_on["spawned"] > 0
    assert t_on["chain_length_max"] < t_off["chain_length_max"]
    assert (out["off"][1] == out["on"][1]).all() and (out["off"][2] == out["on"][2]).all()
    er, ek = o.list_closed_form(n, 44, host(w))
    assert (out["on"][1] == er).all() and (out["on"][2] == ek).all()
    ctx.trace(True)
    ctx.rank_list(succ, w, ruler_spacing=256)
    tr = ctx.trace_read()
    ctx.trace(False)
    assert tr[0]["spawned"] > 0


def test_spawn_errors(ctx):
    """Invalid inputs still fail cleanly with spawn on: a cycle, an element
    with two predecessors (validated), an out-of-range id."""
    import kascade_amd
    n = 50000
    succ, w = o.gen_list(n, 8, True)
    cyc = succ.copy()
    tail = int(np.nonzero(succ == np.arange(n, dtype=np.uint64))[0][0])
    cyc[tail] = succ[0] if succ[0] != tail else 1
    wc = w.copy()
    wc[tail] = 5
    with pytest.raises(kascade_amd.KascadeError) as e:
        ctx.rank_list(dev(cyc), dev(wc), ruler_spacing=512, spawn="on")
    assert e.value.status in (kascade_amd.KC_ECYCLE, kascade_amd.KC_EINVAL)
    bad = succ.copy()
    bad[3] = n + 5
    with pytest.raises(kascade_amd.KascadeError) as e:
        ctx.rank_list(dev(bad), dev(w), ruler_spacing=512, spawn="on")
    assert e.value.status == kascade_amd.KC_EINVAL
    # the context is still usable
    root, rank = ctx.rank_list(dev(succ), dev(w), ruler_spacing=512, spawn="on")
    er, ek = o.list_closed_form(n, 8, w)
    assert (host(root) == er).all() and (host(rank) == ek).all()


@pytest.mark.parametrize("n", [100003, 1 << 21, (1 << 22) + 5])
@pytest.mark.parametrize("spacing", [64, 256, 2048])
@pytest.mark.parametrize("launch", ["auto", "host"])
def test_spawn_trees(ctx, n, spacing, launch):
    """Spawn on Euler tours: candidates are moved onto upward edges (the only
    tour elements that carry claims).  Wide and narrow tours, linked and
    sorted builds, against the closed form."""
    say("spawn trees", n, spacing, launch)
    parent = o.gen_tree(n, 7 + spacing)
    er, ed = o.tree_closed_form(n, 7 + spacing)
    for build in ("linked", "sorted"):
        root, depth = ctx.root_tree(dev(parent), tour_build=build, ruler_spacing=spacing, spawn="on",
                                    launch=launch)
        assert (host(root) == er).all() and (host(depth) == ed).all(), build


@pytest.mark.parametrize("trees", [2, 3000])
def test_spawn_tree_forests(ctx, trees):
    from test_gpu_narrow import fast_forest
    n = (1 << 21) + 17
    say("spawn tree forests", trees)
    parent = fast_forest(n, trees, trees + 5)
    er, ed = o.euler_tour_rooting(parent)
    for spacing in (0, 128, 1024):
        root, depth = ctx.root_tree(dev(parent), ruler_spacing=spacing, spawn="on")
        assert (host(root) == er).all() and (host(depth) == ed).all(), spacing
