"""Local contraction (the reference's use_local_contraction / local
preprocessing: pointer_doubling_generic.hpp:471-543,
sparse_ruling_set_detail/local_contraction.hpp:28-104): runs of consecutive
successors are collapsed before ranking and expanded afterwards.  Every
result is bit-exact against the oracle's sequential walk with contraction
forced on, forced off and left to the sample, on inputs whose runs cross the
1024-element tiles, start and end at tile edges, hold tails, form cycles,
and in segmented batches; the trace shows the contracted level size."""
import numpy as np
import pytest

import oracle_lib as o
from test_gpu_rank import dev, host, say

pytestmark = pytest.mark.gpu
ALGOS = ["pointer_doubling", "sparse_ruling_set"]
MODES = ["on", "off", "auto"]


def blocky_forest(n, seed, max_block, nlists=1, signed=False):
    """Blocks of consecutive ids (lengths 1..max_block) chained in a random
    block order, cut into nlists lists; random weights (signed or not)."""
    rng = np.random.default_rng(seed)
    lens = []
    left = n
    while left:
        b = int(min(left, rng.integers(1, max_block + 1)))
        lens.append(b)
        left -= b
    starts = np.concatenate(([0], np.cumsum(lens)[:-1]))
    order = rng.permutation(len(lens))
    seq = np.concatenate([np.arange(starts[b], starts[b] + lens[b]) for b in order]).astype(np.uint64)
    cuts = np.sort(rng.choice(np.arange(1, n), size=nlists - 1, replace=False)) if nlists > 1 else []
    succ = np.arange(n, dtype=np.uint64)
    w = np.zeros(n, np.int64)
    for part in np.split(seq, cuts):
        succ[part[:-1]] = part[1:]
        lo, hi = (-1000, 1001) if signed else (1, 101)
        w[part[:-1]] = rng.integers(lo, hi, size=len(part) - 1)
    return succ, w


def check(ctx, succ, w, **cfg):
    root, rank = ctx.rank_list(dev(succ), dev(w) if w is not None else None, **cfg)
    if w is None:  # unit weights, 0 on tails
        w = (succ != np.arange(len(succ), dtype=np.uint64)).astype(np.int64)
    er, ek = o.sequential_walk(succ, w)
    assert (host(root) == er).all(), cfg
    assert (host(rank) == ek).all(), cfg


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("algo", ALGOS)
@pytest.mark.parametrize("n", [4096, 4097, 8191, 12289, 100003])
def test_contract_local_lists(ctx, mode, algo, n):
    say("contract local", mode, algo, n)
    succ, w = o.gen_list(n, n + 3, True, local=True)
    check(ctx, succ, w, algorithm=algo, local_contraction=mode)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("max_block,nlists,signed", [(3, 1, False), (5000, 1, True), (20000, 7, False),
                                                     (300, 2000, True), (9000, 40000, False)])
def test_contract_blocky_forests(ctx, mode, max_block, nlists, signed):
    n = 300007
    say("contract blocky", mode, max_block, nlists, signed)
    succ, w = blocky_forest(n, max_block + nlists, max_block, nlists, signed)
    for algo in ALGOS:
        check(ctx, succ, w, algorithm=algo, local_contraction=mode)


@pytest.mark.parametrize("mode", MODES)
def test_contract_paths_and_singletons(ctx, mode):
    """The sorted path (one run per tile), the reversed path (no runs),
    all singletons (every element a tail), unit weights."""
    say("contract paths", mode)
    for n in (4096, 70001):
        up = np.minimum(np.arange(1, n + 1), n - 1).astype(np.uint64)
        root, rank = ctx.rank_list(dev(up), None, local_contraction=mode)
        assert (host(root) == n - 1).all() and (host(rank) == np.arange(n - 1, -1, -1)).all()
        down = np.maximum(np.arange(-1, n - 1), 0).astype(np.uint64)
        root, rank = ctx.rank_list(dev(down), None, local_contraction=mode)
        assert (host(root) == 0).all() and (host(rank) == np.arange(n)).all()
        ident = np.arange(n, dtype=np.uint64)
        root, rank = ctx.rank_list(dev(ident), None, local_contraction=mode)
        assert (host(root) == ident).all() and (host(rank) == 0).all()
        # tails in the middle of runs: 0..k-1 -> k (a tail), k+1.. -> n-1
        k = n // 3
        s = np.minimum(np.arange(1, n + 1), n - 1).astype(np.uint64)
        s[k] = k
        check(ctx, s, None, local_contraction=mode)


def test_contract_random_list_forced(ctx):
    """A random list has almost no runs: forced contraction still gives the
    oracle's results (the contracted list is about as long as the input)."""
    n = 1 << 20
    say("contract random forced", n)
    succ, w = o.gen_list(n, 91, True)
    root, rank = ctx.rank_list(dev(succ), dev(w), local_contraction="on")
    er, ek = o.list_closed_form(n, 91, w)
    assert (host(root) == er).all() and (host(rank) == ek).all()


@pytest.mark.parametrize("launch", ["auto", "device", "host"])
def test_contract_large_local(ctx, launch):
    n = 1 << 24
    say("contract large", launch)
    succ, w = ctx.gen_list(n, 12, weighted=True, local=True)
    r_on, k_on = ctx.rank_list(succ, w, local_contraction="on", launch=launch)
    r_off, k_off = ctx.rank_list(succ, w, local_contraction="off", launch=launch)
    import torch
    assert torch.equal(r_on, r_off) and torch.equal(k_on, k_off)
    er, ek = o.sequential_walk(host(succ), host(w))
    assert (host(r_on) == er).all() and (host(k_on) == ek).all()


def test_contract_trace_level_size(ctx):
    """With contraction on, the first traced level ranks one element per
    run: n / 10 (+ tile cuts) for blocks of 10."""
    n = 1 << 20
    succ, w = o.gen_list(n, 4, True, local=True)
    for mode, lo, hi in (("on", n // 10, n // 10 + n // 1024 + 2), ("off", n, n)):
        ctx.trace(True)
        ctx.rank_list(dev(succ), dev(w), local_contraction=mode)
        tr = ctx.trace_read()
        ctx.trace(False)
        assert lo <= tr[0]["n"] <= hi, (mode, tr[0]["n"])


def test_contract_errors(ctx):
    """A cycle inside one run, a cycle across tiles, an element with two
    predecessors and an out-of-range id fail cleanly with contraction on."""
    import kascade_amd
    for n in (3, 5000, 50000):
        cyc = np.roll(np.arange(n, dtype=np.uint64), -1)
        for mode in MODES:
            with pytest.raises(kascade_amd.KascadeError) as e:
                ctx.rank_list(dev(cyc), None, local_contraction=mode)
            assert e.value.status == kascade_amd.KC_ECYCLE, (n, mode)
    # a cycle on 100..199 beside a proper list 0..99 -> 99
    n = 9000
    s = np.minimum(np.arange(1, n + 1), n - 1).astype(np.uint64)
    s[99] = 99
    s[199] = 100
    with pytest.raises(kascade_amd.KascadeError) as e:
        ctx.rank_list(dev(s), None, local_contraction="on")
    assert e.value.status == kascade_amd.KC_ECYCLE
    # in-degree 2 (validated) and a range error
    tree = np.minimum(np.arange(1, n + 1), n - 1).astype(np.uint64)
    tree[5000] = 10
    with pytest.raises(kascade_amd.KascadeError) as e:
        ctx.rank_list(dev(tree), None, local_contraction="on", validate=True)
    assert e.value.status == kascade_amd.KC_EINVAL
    bad = np.minimum(np.arange(1, n + 1), n - 1).astype(np.uint64)
    bad[7] = n + 3
    with pytest.raises(kascade_amd.KascadeError) as e:
        ctx.rank_list(dev(bad), None, local_contraction="on")
    assert e.value.status == kascade_amd.KC_EINVAL
    with pytest.raises(kascade_amd.KascadeError) as e:
        ctx.rank_list(dev(bad), None, local_contraction=7)
    assert e.value.status == kascade_amd.KC_EINVAL
    check(ctx, *o.gen_list(20000, 5, True, local=True), local_contraction="on")


@pytest.mark.parametrize("frac_local", [0.3, 1.0])
def test_contract_segmented(ctx, frac_local):
    import torch
    rng = np.random.default_rng(11)
    nseg = 400
    sizes = np.floor(2.0 ** rng.uniform(4, 15, size=nseg)).astype(np.uint64)
    off = np.zeros(nseg + 1, np.uint64)
    off[1:] = np.cumsum(sizes)
    seeds = rng.integers(1, 1 << 62, size=nseg, dtype=np.uint64)
    local = (rng.random(nseg) < frac_local).astype(np.uint8)
    say("contract segmented", frac_local, int(off[-1]))
    d_off = torch.from_numpy(off).cuda()
    succ, w = ctx.gen_lists_segmented(d_off, torch.from_numpy(seeds).cuda(), torch.from_numpy(local).cuda(),
                                      weighted=True)
    res = {m: ctx.rank_lists_segmented(succ, d_off, w, local_contraction=m) for m in MODES}
    for m in ("on", "auto"):
        assert torch.equal(res[m][0].view(torch.int32), res["off"][0].view(torch.int32)), m
        assert torch.equal(res[m][1], res["off"][1]), m
    hs, hw, hr, hk = host(succ), host(w), host(res["on"][0]), host(res["on"][1])
    for s in range(nseg):
        a, b = int(off[s]), int(off[s + 1])
        cr, ck = o.sequential_walk(hs[a:b].astype(np.uint64), hw[a:b])
        assert (hr[a:b] == cr).all() and (hk[a:b] == ck).all(), s


@pytest.mark.parametrize("n", [1 << 22, 1 << 24])
def test_contract_auto_decision(ctx, n):
    """AUTO contracts a locality-biased list and not a random one, both when
    it samples the packed records (below 2^24) and the caller's successor
    array (from 2^24); the results match the oracle either way."""
    say("contract auto", n)
    for local, contracted in ((True, True), (False, False)):
        succ, w = o.gen_list(n, 21, True, local=local)  # C generator: about a second at 2^24
        ctx.trace(True)
        root, rank = ctx.rank_list(dev(succ), dev(w))
        tr = ctx.trace_read()
        ctx.trace(False)
        assert (tr[0]["n"] < n // 5) == contracted, (local, tr[0]["n"])
        er, ek = o.sequential_walk(succ, w)
        assert (host(root) == er).all() and (host(rank) == ek).all()
