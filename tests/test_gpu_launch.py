"""Device-sized levels and HIP-graph replay: every launch mode (device-sized,
graph-replayed, host-sized) gives the same bit-exact ranks and roots; a
repeated call is captured once and replayed; an input whose unreached
prefix outgrows the device-sized capacity is redone host-sized, exactly."""
import numpy as np
import pytest

import oracle_lib as o
from test_gpu_rank import dev, host, random_forest_lists, say

pytestmark = pytest.mark.gpu

MODES = ["device", "host", "auto"]


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("n", [10000, 300007, 1 << 22])
def test_launch_modes_single_list(ctx, mode, n):
    say("launch list", mode, n)
    succ, w = o.gen_list(n, 5 + n, True)
    er, ek = o.list_closed_form(n, 5 + n, w)
    for algo in ("srs", "pd"):
        root, rank = ctx.rank_list(dev(succ), dev(w), algorithm=algo, launch=mode)
        assert (host(root) == er).all() and (host(rank) == ek).all(), algo


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("nlists,spacing", [(30, 0), (3000, 0), (3000, 64), (30000, 0), (40, 1024)])
def test_launch_modes_forests(ctx, mode, nlists, spacing):
    """Forests put many head prefixes into the unreached sub-problem."""
    say("launch forest", mode, nlists, spacing)
    succ, w = random_forest_lists(200000, nlists, 3 + nlists)
    er, ek = o.sequential_walk(succ, w)
    root, rank = ctx.rank_list(dev(succ), dev(w), launch=mode, ruler_spacing=spacing)
    assert (host(root) == er).all() and (host(rank) == ek).all()


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("build", ["linked", "sorted"])
def test_launch_modes_trees(ctx, mode, build):
    say("launch trees", mode, build)
    from test_gpu_tree import random_forest
    for parent in (o.gen_tree(300000, 4), random_forest(50000, 500, 9)):
        er, ed = o.euler_tour_rooting(parent)
        for algo in ("srs", "pd"):
            root, depth = ctx.root_tree(dev(parent), tour_build=build, algorithm=algo, launch=mode)
            assert (host(root) == er).all() and (host(depth) == ed).all(), algo


@pytest.mark.parametrize("mode", MODES)
def test_launch_modes_segmented(ctx, mode):
    say("launch segmented", mode)
    rng = np.random.default_rng(77)
    sizes = list(rng.integers(1, 30000, size=40)) + [0, 1, 9000]
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    succ = np.empty(int(off[-1]), np.uint32)
    w = np.empty(int(off[-1]), np.int64)
    exp_root = np.empty(int(off[-1]), np.uint32)
    exp_rank = np.empty(int(off[-1]), np.int64)
    for s, (a, b) in enumerate(zip(off[:-1], off[1:])):
        if b == a:
            continue
        ls, lw = o.gen_list(int(b - a), 900 + s, True)
        succ[a:b] = ls
        w[a:b] = lw
        r, k = o.sequential_walk(ls, lw)
        exp_root[a:b] = r
        exp_rank[a:b] = k
    root, rank = ctx.rank_lists_segmented(dev(succ), dev(off), dev(w), launch=mode)
    assert (host(root) == exp_root).all() and (host(rank) == exp_rank).all()


def test_graph_capture_and_replay(ctx):
    """The same call repeated: the first runs directly, the second is captured
    and launched, later ones replay the graph.  Outputs are cleared before
    every call and checked after it; other work runs between the calls."""
    import torch
    import kascade_amd
    say("graph replay")
    other = kascade_amd.Context(0)
    for n, tree in ((1 << 20, False), (300000, True)):
        if tree:
            parent = dev(o.gen_tree(n, 6))
            er, ek = o.tree_closed_form(n, 6)
        else:
            succ_h, w_h = o.gen_list(n, 6, True)
            succ, w = dev(succ_h), dev(w_h)
            er, ek = o.list_closed_form(n, 6, w_h)
        root = torch.empty(n, dtype=torch.uint64, device="cuda")
        rank = torch.empty(n, dtype=torch.int64, device="cuda")
        before = ctx.launch_stats()
        for rep in range(6):
            root.fill_(7)
            rank.fill_(-7)
            if tree:
                ctx.root_tree(parent, root, rank)
            else:
                ctx.rank_list(succ, w, root, rank)
            assert (host(root) == er).all() and (host(rank) == ek).all(), (n, rep)
            # unrelated work between replays: torch kernels and another context
            s2, w2 = o.gen_list(5000 + rep, rep, True)
            r2, k2 = other.rank_list(dev(s2), dev(w2))
            e2r, e2k = o.list_closed_form(5000 + rep, rep, w2)
            assert (host(r2) == e2r).all() and (host(k2) == e2k).all()
            torch.cuda.synchronize()
        after = ctx.launch_stats()
        assert after["graph_captures"] - before["graph_captures"] >= 1, (before, after)
        assert after["graph_replays"] - before["graph_replays"] >= 3, (before, after)
    other.close()


def splitmix64(x):
    m = (1 << 64) - 1
    x = (x + 0x9E3779B97F4A7C15) & m
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & m
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & m
    return x ^ (x >> 31)


def hash32(x):
    m = (1 << 32) - 1
    x ^= x >> 16
    x = (x * 0x7FEB352D) & m
    x ^= x >> 15
    x = (x * 0x846CA68B) & m
    x ^= x >> 16
    return x


def test_overflow_redone_host_sized(ctx):
    """White-box: with ruler spacing 2 the level-0 rulers are one of each
    pair {2b, 2b+1} (device.hpp ruler_of_block, seeded by n).  A list that
    starts with 6000 non-rulers has an unreached prefix beyond the
    device-sized capacity (2 K lists + 24 K + 4096 = 4148), so the call is
    redone host-sized; the result is still exact."""
    say("overflow redo")
    n = 40000
    seed = splitmix64(0xC0FFEE + n) & 0xFFFFFFFF
    rulers = np.array([2 * b + (hash32(b ^ seed) & 1) for b in range(n // 2)], np.int64)
    nonrulers = np.arange(0, n, 2, dtype=np.int64) + (1 - (rulers - np.arange(0, n, 2)))
    prefix = nonrulers[:6000]
    rng = np.random.default_rng(1)
    rest = np.setdiff1d(np.arange(n, dtype=np.int64), prefix)
    order = np.concatenate([prefix, rng.permutation(rest)])
    succ = np.empty(n, np.uint64)
    succ[order[:-1]] = order[1:]
    succ[order[-1]] = order[-1]
    w = rng.integers(1, 101, size=n).astype(np.int64)
    w[order[-1]] = 0
    er, ek = o.sequential_walk(succ, w)
    before = ctx.launch_stats()["host_redos"]
    for mode in ("device", "auto"):
        root, rank = ctx.rank_list(dev(succ), dev(w), ruler_spacing=2, launch=mode)
        assert (host(root) == er).all() and (host(rank) == ek).all(), mode
    assert ctx.launch_stats()["host_redos"] - before == 2
    root, rank = ctx.rank_list(dev(succ), dev(w), ruler_spacing=2, launch="host")
    assert (host(root) == er).all() and (host(rank) == ek).all()


@pytest.mark.parametrize("mode", MODES)
def test_launch_modes_edges(ctx, mode):
    """Boundaries of the device-sized plan: the one-workgroup solver's limit
    (8192 / 8193 records), forests made only of singletons or of 2-element
    lists (every element a tail or a head), and tiny trees."""
    import numpy as np
    say("launch edges", mode)
    for n in (8191, 8192, 8193, 8194, 16385):
        succ, w = o.gen_list(n, n, True)
        er, ek = o.list_closed_form(n, n, w)
        root, rank = ctx.rank_list(dev(succ), dev(w), launch=mode)
        assert (host(root) == er).all() and (host(rank) == ek).all(), n
    n = 50000
    single = np.arange(n, dtype=np.uint64)
    root, rank = ctx.rank_list(dev(single), None, launch=mode)
    assert (host(root) == single).all() and (host(rank) == 0).all()
    pairs = np.arange(n, dtype=np.uint64)
    pairs[0::2] += 1                                   # 2i -> 2i+1, 2i+1 tail
    w = np.zeros(n, np.int64)
    w[0::2] = 7
    er, ek = o.sequential_walk(pairs, w)
    root, rank = ctx.rank_list(dev(pairs), dev(w), launch=mode, ruler_spacing=4)
    assert (host(root) == er).all() and (host(rank) == ek).all()
    for nv in (1, 2, 3, 4097, 4098):
        parent = o.gen_tree(nv, nv)
        er, ed = o.tree_closed_form(nv, nv)
        root, depth = ctx.root_tree(dev(parent), launch=mode)
        assert (host(root) == er).all() and (host(depth) == ed).all(), nv
