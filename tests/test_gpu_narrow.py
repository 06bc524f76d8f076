"""Narrow (8-byte) level-0 records: single-list rankings from 2^22 elements
pack their top level into {u32 successor, i32 weight} records and claim
with i32 distances.  Weights that do not fit, distances that outgrow int32
and id spaces beyond 2^30 fall back to 16-byte records; every result is
bit-exact against the oracle either way."""
import numpy as np
import pytest

import oracle_lib as o
from test_gpu_rank import dev, host, random_forest_lists, say

pytestmark = pytest.mark.gpu


def check(ctx, succ, w, **cfg):
    root, rank = ctx.rank_list(dev(succ), dev(w) if w is not None else None, **cfg)
    if w is None:
        w = (succ != np.arange(len(succ), dtype=np.uint64)).astype(np.int64)
    er, ek = o.sequential_walk(succ, w)
    assert (host(root) == er).all(), cfg
    assert (host(rank) == ek).all(), cfg


@pytest.mark.parametrize("n", [1 << 22, (1 << 22) + 12345, 1 << 23])
@pytest.mark.parametrize("launch", ["auto", "device"])
def test_narrow_random_lists(ctx, n, launch):
    say("narrow random", n, launch)
    succ, w = o.gen_list(n, n + 7, True)
    before = ctx.record_stats()
    for _ in range(3):  # the third call replays a captured graph
        root, rank = ctx.rank_list(dev(succ), dev(w), launch=launch)
        er, ek = o.list_closed_form(n, n + 7, w)
        assert (host(root) == er).all() and (host(rank) == ek).all()
    after = ctx.record_stats()
    assert after["narrow_calls"] - before["narrow_calls"] == 3
    assert after["narrow_redos"] == before["narrow_redos"]


@pytest.mark.parametrize("weight", [1 << 30, -(1 << 30), (1 << 31) - 1, -(1 << 31), 1 << 31, -(1 << 31) - 1])
def test_narrow_weight_limits(ctx, weight):
    """Weights at and beyond the int32 limits: the first four fit the narrow
    records but their claim distances overflow after a step or two (the call
    is redone with wide records), the last two do not fit at all."""
    n = 1 << 22
    say("narrow weights", weight)
    succ, _ = o.gen_list(n, 3, False)
    w = np.full(n, weight, np.int64)
    w[succ == np.arange(n, dtype=np.uint64)] = 0
    before = ctx.record_stats()
    check(ctx, succ, w)
    after = ctx.record_stats()
    fits = -(1 << 31) <= weight < (1 << 31)
    assert after["narrow_redos"] - before["narrow_redos"] == (1 if fits else 0)
    assert after["narrow_calls"] == before["narrow_calls"]


def test_narrow_unit_and_signed_weights(ctx):
    n = 1 << 22
    say("narrow unit/signed")
    succ, _ = o.gen_list(n, 5, False)
    check(ctx, succ, None)
    rng = np.random.default_rng(5)
    w = rng.integers(-1000, 1001, size=n).astype(np.int64)
    w[succ == np.arange(n, dtype=np.uint64)] = 0
    check(ctx, succ, w)


@pytest.mark.parametrize("nlists", [3, 5000, 200000])
def test_narrow_forests(ctx, nlists):
    n = (1 << 22) + 3
    say("narrow forest", nlists)
    succ, w = random_forest_lists(n, nlists, nlists + 1)
    check(ctx, succ, w)


def test_narrow_spawn_and_spacings(ctx):
    n = 1 << 22
    say("narrow spawn")
    succ, w = o.gen_list(n, 9, True)
    for spacing, spawn in ((16, "off"), (512, "on"), (4096, "on")):
        check(ctx, succ, w, ruler_spacing=spacing, spawn=spawn)


def test_narrow_host_launch_and_errors(ctx):
    """Host-sized levels use wide records; errors are reported as on the
    wide path and the context stays usable."""
    import kascade_amd
    n = 1 << 22
    say("narrow errors")
    succ, w = o.gen_list(n, 11, True)
    check(ctx, succ, w, launch="host")
    cyc = np.roll(np.arange(n, dtype=np.uint64), -1)
    with pytest.raises(kascade_amd.KascadeError) as e:
        ctx.rank_list(dev(cyc), None, algorithm="srs")
    assert e.value.status == kascade_amd.KC_ECYCLE
    bad = succ.copy()
    bad[5] = n + 1
    with pytest.raises(kascade_amd.KascadeError) as e:
        ctx.rank_list(dev(bad), dev(w))
    assert e.value.status == kascade_amd.KC_EINVAL
    wt = w.copy()
    wt[succ == np.arange(n, dtype=np.uint64)] = 7
    with pytest.raises(kascade_amd.KascadeError) as e:
        ctx.rank_list(dev(succ), dev(wt))
    assert e.value.status == kascade_amd.KC_EINVAL
    check(ctx, succ, w)


def fast_forest(n, trees, seed):
    """A forest of `trees` random recursive trees (each vertex's parent is
    uniform among the earlier vertices of its tree), relabelled by a seeded
    permutation; vectorised for large n."""
    rng = np.random.default_rng(seed)
    cuts = np.sort(rng.choice(np.arange(1, n), size=trees - 1, replace=False))
    start = np.zeros(n, np.int64)
    start[cuts] = cuts
    start = np.maximum.accumulate(start)
    v = np.arange(n, dtype=np.int64)
    parent = start + np.floor(rng.random(n) * (v - start)).astype(np.int64)
    parent[v == start] = v[v == start]
    sigma = rng.permutation(n)
    out = np.empty(n, np.uint64)
    out[sigma] = sigma[parent]
    return out


@pytest.mark.parametrize("build", ["linked", "sorted"])
@pytest.mark.parametrize("n", [1 << 21, (1 << 22) + 7])
def test_narrow_trees(ctx, build, n):
    """Tree rooting from 2^21 vertices ranks an 8-byte tour."""
    say("narrow tree", build, n)
    parent = ctx.gen_tree(n, 31)
    before = ctx.record_stats()
    root, depth = ctx.root_tree(parent, tour_build=build)
    after = ctx.record_stats()
    assert after["narrow_calls"] - before["narrow_calls"] == 1
    er, ed = o.tree_closed_form(n, 31)
    assert (host(root) == er).all() and (host(depth) == ed).all()


@pytest.mark.parametrize("trees", [7, 3000])
def test_narrow_tree_forests(ctx, trees):
    n = (1 << 21) + 5
    say("narrow tree forest", trees)
    parent = fast_forest(n, trees, trees)
    before = ctx.record_stats()
    for build in ("linked", "sorted"):
        root, depth = ctx.root_tree(dev(parent), tour_build=build)
        er, ed = o.euler_tour_rooting(parent)
        assert (host(root) == er).all() and (host(depth) == ed).all(), build
    after = ctx.record_stats()
    assert after["narrow_calls"] - before["narrow_calls"] == 2
    assert after["narrow_redos"] == before["narrow_redos"]


def test_narrow_tree_cycle(ctx):
    """A 2-cycle (a root made the child of its own child) beside 19 proper
    trees fails with KC_ECYCLE, and the context ranks a tree afterwards."""
    import kascade_amd
    n = 1 << 21
    say("narrow tree cycle")
    parent = fast_forest(n, 20, 3)
    v = np.arange(n, dtype=np.uint64)
    r = int(np.nonzero(parent == v)[0][0])
    parent[r] = int(np.nonzero((parent == r) & (v != r))[0][0])
    with pytest.raises(kascade_amd.KascadeError) as e:
        ctx.root_tree(dev(parent))
    assert e.value.status == kascade_amd.KC_ECYCLE
    parent = fast_forest(n, 20, 4)
    root, depth = ctx.root_tree(dev(parent))
    er, ed = o.euler_tour_rooting(parent)
    assert (host(root) == er).all() and (host(depth) == ed).all()


def test_records_wide_knob(ctx):
    """kc_rank_cfg.records = WIDE keeps every top level at 16-byte records
    (lists, trees, segmented batches); the results are those of AUTO, which
    uses 8-byte records for all three."""
    import torch
    say("records knob")
    n = 1 << 22
    succ, w = o.gen_list(n, 41, True)
    er, ek = o.list_closed_form(n, 41, w)
    parent = ctx.gen_tree(n, 43)
    tr, td = o.tree_closed_form(n, 43)
    sizes = np.full(64, (1 << 16) + 3, np.uint64)
    off = np.zeros(len(sizes) + 1, np.uint64)
    off[1:] = np.cumsum(sizes)
    seeds = torch.from_numpy(np.arange(1, len(sizes) + 1, dtype=np.uint64)).to("cuda")
    ssucc, sw = ctx.gen_lists_segmented(dev(off), seeds, None, weighted=True)
    for records, narrow in [("wide", 0), ("auto", 3)]:
        before = ctx.record_stats()
        root, rank = ctx.rank_list(dev(succ), dev(w), records=records)
        assert (host(root) == er).all() and (host(rank) == ek).all(), records
        root, depth = ctx.root_tree(parent, records=records)
        assert (host(root) == tr).all() and (host(depth) == td).all(), records
        sroot, srank = ctx.rank_lists_segmented(ssucc, dev(off), sw, records=records)
        if records == "wide":
            wide_seg = (host(sroot).copy(), host(srank).copy())
        else:
            assert (host(sroot) == wide_seg[0]).all() and (host(srank) == wide_seg[1]).all()
        after = ctx.record_stats()
        assert after["narrow_calls"] - before["narrow_calls"] == narrow, records
    with pytest.raises(Exception):
        ctx.rank_list(dev(succ), dev(w), records=2)
