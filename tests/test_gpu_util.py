"""GPU parity of the successor utilities (is_list, LeafInfo, reverse_list) and
of the benchmark statistics (compute_stats) through the C ABI, against the CPU
oracle's restatements."""
import numpy as np
import pytest

import oracle_lib as o
from test_gpu_rank import dev, host, random_forest_lists, say

pytestmark = pytest.mark.gpu


def random_function(n, seed):
    rng = np.random.default_rng(seed)
    return rng.integers(0, n, size=n).astype(np.uint64)


CASES = [(1, 1), (2, 1), (300, 7), (5000, 1), (65537, 400), (1 << 20, 3)]


@pytest.mark.parametrize("n,nlists", CASES)
def test_is_list_and_leaf_info(ctx, n, nlists):
    say("is_list/leaf_info", n, nlists)
    for succ in (random_forest_lists(n, nlists, n + nlists)[0], random_function(n, n),
                 np.arange(n, dtype=np.uint64)):
        assert ctx.is_list(dev(succ)) == o.is_list(succ)
        flags, k = ctx.leaf_info(dev(succ))
        ef, ek = o.leaf_info(succ)
        assert k == ek and (host(flags) == ef).all()
        # host arrays are staged
        assert ctx.is_list(succ) == o.is_list(succ)


@pytest.mark.parametrize("n,nlists", CASES)
@pytest.mark.parametrize("weighted", [True, False])
def test_reverse_list(ctx, n, nlists, weighted):
    say("reverse_list", n, nlists, weighted)
    succ, w = random_forest_lists(n, nlists, 7 * n + nlists)
    dist = w if weighted else None
    pred, dpred, roots = ctx.reverse_list(dev(succ), None if dist is None else dev(dist))
    ep, ed, er = o.reverse_list(succ, dist)
    assert (host(pred) == ep).all() and (host(dpred) == ed).all() and (host(roots) == er).all()


def test_util_errors(ctx):
    import kascade_amd
    say("util errors")
    bad = np.array([0, 5, 1], np.uint64)
    for call in (lambda: ctx.is_list(dev(bad)), lambda: ctx.leaf_info(dev(bad)),
                 lambda: ctx.reverse_list(dev(bad))):
        with pytest.raises(kascade_amd.KascadeError) as e:
            call()
        assert e.value.status == kascade_amd.KC_EINVAL
    with pytest.raises(kascade_amd.KascadeError) as e:      # two predecessors
        ctx.reverse_list(dev(np.array([2, 2, 2], np.uint64)))
    assert e.value.status == kascade_amd.KC_EINVAL
    # the context stays usable after errors
    assert ctx.is_list(dev(np.array([1, 2, 2], np.uint64)))


@pytest.mark.parametrize("level", ["basic", "extensive-reduced-output", "extensive"])
@pytest.mark.parametrize("n,nlists", [(1, 1), (64, 64), (3000, 1), (70000, 900), (1 << 20, 40)])
def test_compute_stats(ctx, level, n, nlists):
    say("compute_stats", level, n, nlists)
    succ, w = random_forest_lists(n, nlists, n + 3 * nlists)
    root, rank = o.sequential_walk(succ, w)
    got = ctx.compute_stats(dev(root), dev(rank), level)
    assert got == o.compute_stats(root, rank, level)


def test_compute_stats_on_device_outputs(ctx):
    """Statistics of a device ranking: a generated list and a generated tree."""
    import torch
    say("compute_stats device outputs")
    n = 1 << 22
    succ, w = ctx.gen_list(n, 77, weighted=True)
    root, rank = ctx.rank_list(succ, w)
    got = ctx.compute_stats(root, rank, "extensive")
    er, ek = o.list_closed_form(n, 77, host(w))
    assert got == o.compute_stats(er, ek, "extensive")
    parent = ctx.gen_tree(n, 78)
    troot, depth = ctx.root_tree(parent)
    got = ctx.compute_stats(troot, depth, "extensive-reduced-output")
    er, ed = o.tree_closed_form(n, 78)
    assert got == o.compute_stats(er, ed, "extensive-reduced-output")
    torch.cuda.synchronize()


def test_compute_stats_segmented_forest(ctx):
    """Many contiguous trees (a segmented batch's layout) and singletons."""
    say("compute_stats segments")
    rng = np.random.default_rng(5)
    sizes = rng.integers(1, 3000, size=700)
    root = np.concatenate([np.full(s, off + s - 1, np.uint64) for s, off in zip(sizes, np.cumsum(sizes) - sizes)])
    rank = rng.integers(-50, 10**6, size=len(root)).astype(np.int64)
    rank[root == np.arange(len(root))] = 0
    got = ctx.compute_stats(dev(root), dev(rank), "extensive")
    assert got == o.compute_stats(root, rank, "extensive")


def test_measurement_probes(ctx):
    """The bench's access-pattern probes run at a small size and report
    positive rates whose time matches the count; invalid sizes and patterns
    are rejected."""
    import kascade_amd
    say("probes")
    n = 1 << 16
    g, ms = ctx.probe_random_gather(n, 2)
    assert g > 0 and ms > 0
    g, ms = ctx.probe_random_gather(n, 2, mode=1)
    assert g > 0 and ms > 0
    r, ms = ctx.probe_table_gather(n, 1 << 12, 2)
    assert r > 0 and ms > 0
    for pat in ctx.RECORD_PATTERNS:
        r, ms = ctx.probe_record_access(pat, n, 2)
        assert r > 0 and ms > 0 and abs(r * ms * 1e-3 - n) < 1e-6 * n, pat
    for bad in [lambda: ctx.probe_record_access("walk_claim", n + 1, 1),
                lambda: ctx.probe_record_access("walk_claim8", 1 << 31, 1),
                lambda: ctx._check(ctx._lib.kc_probe_record_access(ctx.handle, n, 7, 1, None, None)),
                lambda: ctx.probe_random_gather(n, 1, mode=4)]:
        with pytest.raises(kascade_amd.KascadeError):
            bad()
