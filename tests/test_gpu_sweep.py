"""Randomised parity sweep through the C ABI: seeded list forests, tree
forests and segmented batches of random shapes, sizes and weights, ranked
with random algorithms, launch modes and ruler spacings; every output bit-exact
against the oracle."""
import numpy as np
import pytest

import oracle_lib as o

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def host(t):
    return t.cpu().numpy()


def list_forest(rng, n, nlists, wlo, whi, local):
    """nlists lists over [0, n): a random permutation cut at random places, or
    (local) the identity order cut into runs whose order is shuffled."""
    order = np.arange(n) if local else rng.permutation(n)
    cuts = np.sort(rng.choice(np.arange(1, n), size=nlists - 1, replace=False)) if nlists > 1 else []
    succ = np.arange(n, dtype=np.uint64)
    w = np.zeros(n, np.int64)
    for seg in np.split(order, cuts):
        if len(seg) > 1:
            succ[seg[:-1]] = seg[1:]
            w[seg[:-1]] = rng.integers(wlo, whi + 1, size=len(seg) - 1)
    return succ, w


def tree_forest(rng, n):
    """Roots, random-recursive attachments, path steps and hubs, mixed at
    random rates, then relabelled."""
    p_root, p_path, p_hub = rng.uniform(0, 0.05), rng.uniform(0, 0.9), rng.uniform(0, 0.5)
    parent = np.arange(n, dtype=np.uint64)
    hub = 0
    for v in range(1, n):
        u = rng.random()
        if u < p_root:
            continue
        if u < p_root + p_path * (1 - p_root):
            parent[v] = v - 1
        elif rng.random() < p_hub:
            parent[v] = hub
        else:
            parent[v] = rng.integers(0, v)
        if rng.random() < 0.001:
            hub = v
    sigma = rng.permutation(n)
    out = np.empty(n, np.uint64)
    out[sigma] = sigma[parent]
    return out


@pytest.mark.parametrize("seed", range(24))
def test_list_sweep(ctx, seed):
    rng = np.random.default_rng(1000 + seed)
    n = int(2 ** rng.uniform(0, 18))
    nlists = int(max(1, min(n, 2 ** rng.uniform(0, np.log2(n) + 1))))
    wlo = int(rng.choice([1, 0, -50]))
    succ, w = list_forest(rng, n, nlists, wlo, 100, bool(rng.random() < 0.3))
    kw = dict(algorithm=str(rng.choice(["auto", "srs", "pd"])), launch=str(rng.choice(["auto", "device", "host"])),
              ruler_spacing=int(rng.choice([0, 0, 2, 4, 16, 128])), ruler_spacing_deep=int(rng.choice([0, 2, 8])))
    say("list sweep", seed, n, nlists, wlo, kw)
    root, rank = ctx.rank_list(dev(succ), dev(w), **kw)
    er, ek = o.sequential_walk(succ, w)
    assert (host(root) == er).all() and (host(rank) == ek).all()


@pytest.mark.parametrize("seed", range(16))
def test_tree_sweep(ctx, seed):
    rng = np.random.default_rng(2000 + seed)
    n = int(2 ** rng.uniform(0, 16))
    parent = tree_forest(rng, n)
    kw = dict(tour_build=str(rng.choice(["auto", "linked", "sorted"])),
              algorithm=str(rng.choice(["auto", "srs", "pd"])), launch=str(rng.choice(["auto", "device", "host"])),
              ruler_spacing=int(rng.choice([0, 0, 4, 64])))
    say("tree sweep", seed, n, kw)
    root, depth = ctx.root_tree(dev(parent), **kw)
    er, ed = o.euler_tour_rooting_mt(parent, 4)
    assert (host(root) == er).all() and (host(depth) == ed).all()


@pytest.mark.parametrize("seed", range(8))
def test_segmented_sweep(ctx, seed):
    rng = np.random.default_rng(3000 + seed)
    nseg = int(rng.integers(1, 400))
    sizes = np.where(rng.random(nseg) < 0.1, 0, np.floor(2 ** rng.uniform(0, 13, size=nseg))).astype(np.uint64)
    off = np.zeros(nseg + 1, np.uint64)
    off[1:] = np.cumsum(sizes)
    total = int(off[-1])
    succ = np.empty(total, np.uint32)
    w = np.empty(total, np.int64)
    gsucc = np.empty(total, np.uint64)
    for s in range(nseg):
        a, b = int(off[s]), int(off[s + 1])
        if b == a:
            continue
        ls, lw = list_forest(rng, b - a, 1, -20, 100, bool(rng.random() < 0.3))
        succ[a:b] = ls.astype(np.uint32)
        w[a:b] = lw
        gsucc[a:b] = ls + np.uint64(a)
    kw = dict(algorithm=str(rng.choice(["auto", "srs", "pd"])), launch=str(rng.choice(["auto", "device", "host"])))
    say("segmented sweep", seed, nseg, total, kw)
    if total == 0:
        return
    root, rank = ctx.rank_lists_segmented(dev(succ), dev(off), dev(w), **kw)
    er, ek = o.sequential_walk(gsucc, w)
    seg_base = np.repeat(off[:-1], sizes.astype(np.int64))
    assert (host(root).astype(np.uint64) == er - seg_base).all() and (host(rank) == ek).all()
