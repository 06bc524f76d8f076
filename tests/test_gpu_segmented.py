"""GPU parity of batched segmented ranking (kc_rank_lists_segmented): every
segment bit-exact against the oracle's closed form for that segment."""
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


def make_batch(sizes, seed, local_frac=0.3):
    rng = np.random.default_rng(seed)
    off = np.zeros(len(sizes) + 1, np.uint64)
    off[1:] = np.cumsum(sizes)
    seeds = rng.integers(1, 1 << 62, size=len(sizes), dtype=np.uint64)
    local = (rng.random(len(sizes)) < local_frac).astype(np.uint8)
    return off, seeds, local


def expected(off, seeds, local, weighted):
    roots, ranks = [], []
    for s in range(len(seeds)):
        n = int(off[s + 1] - off[s])
        if n == 0:
            continue
        succ, w = o.gen_list(n, int(seeds[s]), weighted, local=bool(local[s]))
        r, k = o.sequential_walk(succ, w)
        roots.append(r.astype(np.uint32))
        ranks.append(k)
    return np.concatenate(roots), np.concatenate(ranks)


@pytest.mark.parametrize("weighted", [True, False])
def test_segmented_generator_matches_cpu(ctx, weighted):
    say("seg gen", weighted)
    sizes = [1, 2, 10, 999, 4096, 1, 5000]
    off, seeds, local = make_batch(sizes, 1)
    succ, w = ctx.gen_lists_segmented(dev(off), dev(seeds), dev(local), weighted=weighted)
    succ, w = host(succ), host(w)
    for s in range(len(sizes)):
        a, b = int(off[s]), int(off[s + 1])
        es, ew = o.gen_list(b - a, int(seeds[s]), weighted, local=bool(local[s]))
        assert (succ[a:b] == es).all() and (w[a:b] == ew).all(), s


@pytest.mark.parametrize("algo", ["srs", "pd"])
@pytest.mark.parametrize("weighted", [True, False])
def test_segmented_small(ctx, algo, weighted):
    say("seg small", algo, weighted)
    rng = np.random.default_rng(7)
    sizes = list(rng.integers(1, 5000, size=60)) + [0, 1, 0, 4097]
    off, seeds, local = make_batch(sizes, 2)
    succ, w = ctx.gen_lists_segmented(dev(off), dev(seeds), dev(local), weighted=weighted)
    root, rank = ctx.rank_lists_segmented(succ, dev(off), w if weighted else None, algorithm=algo)
    er, ek = expected(off, seeds, local, weighted)
    assert (host(root) == er).all() and (host(rank) == ek).all()


def test_segmented_log_uniform(ctx):
    """A reduced segmented batch: 2000 lists, sizes log-uniform in [2^8, 2^16], 30 %
    locality-biased."""
    say("seg log-uniform")
    rng = np.random.default_rng(11)
    sizes = np.floor(2.0 ** rng.uniform(8, 16, size=2000)).astype(np.uint64)
    off, seeds, local = make_batch(sizes, 3)
    succ, w = ctx.gen_lists_segmented(dev(off), dev(seeds), dev(local), weighted=True)
    root, rank = ctx.rank_lists_segmented(succ, dev(off), w)
    er, ek = expected(off, seeds, local, True)
    assert (host(root) == er).all() and (host(rank) == ek).all()


def test_segmented_host_arrays_and_errors(ctx):
    import kascade_amd
    say("seg host/errors")
    off = np.array([0, 3, 3, 5], np.uint64)
    succ = np.array([1, 2, 2, 0, 0], np.uint32)      # [0->1->2], [], [1->0]
    root, rank = ctx.rank_lists_segmented(succ, off, None)
    assert root.tolist() == [2, 2, 2, 0, 0] and rank.tolist() == [2, 1, 0, 0, 1]
    bad = np.array([1, 3, 2, 0, 0], np.uint32)       # 3 is outside segment 0
    with pytest.raises(kascade_amd.KascadeError) as e:
        ctx.rank_lists_segmented(bad, off, None)
    assert e.value.status == kascade_amd.KC_EINVAL


@pytest.mark.parametrize("algo", ["srs", "pd"])
def test_segmented_many_tiny_lists(ctx, algo):
    """200,000 segments of 0 to 20 elements: most lists hold no ruler, so
    nearly every element goes through the unreached-prefix sub-problem."""
    say("seg many tiny", algo)
    rng = np.random.default_rng(23)
    sizes = rng.integers(0, 21, size=200000).astype(np.uint64)
    off = np.zeros(len(sizes) + 1, np.uint64)
    off[1:] = np.cumsum(sizes)
    total = int(off[-1])
    succ = np.empty(total, np.uint32)
    w = rng.integers(-5, 50, size=total).astype(np.int64)
    gsucc = np.empty(total, np.uint64)
    for s in np.flatnonzero(sizes):
        a, b = int(off[s]), int(off[s + 1])
        perm = rng.permutation(b - a)
        loc = np.arange(b - a, dtype=np.uint64)
        loc[perm[:-1]] = perm[1:]
        succ[a:b] = loc.astype(np.uint32)
        gsucc[a:b] = loc + np.uint64(a)
        w[a + perm[-1]] = 0
    root, rank = ctx.rank_lists_segmented(dev(succ), dev(off), dev(w), algorithm=algo)
    er, ek = o.sequential_walk(gsucc, w)
    base = np.repeat(off[:-1], sizes.astype(np.int64))
    assert (host(root).astype(np.uint64) == er - base).all() and (host(rank) == ek).all()


def host_batch(sizes, seed, weighted=True):
    """Host arrays of a segmented batch (local u32 successors) and each
    segment's oracle result."""
    off = np.zeros(len(sizes) + 1, np.uint64)
    off[1:] = np.cumsum(sizes)
    succ = np.empty(int(off[-1]), np.uint32)
    w = np.empty(int(off[-1]), np.int64)
    for s, n in enumerate(sizes):
        a = int(off[s])
        ss, ww = o.gen_list(int(n), seed + s, weighted)
        succ[a:a + n] = ss.astype(np.uint32)
        w[a:a + n] = ww
    return off, succ, w


def host_expected(off, succ, w):
    roots, ranks = [], []
    for s in range(len(off) - 1):
        a, b = int(off[s]), int(off[s + 1])
        if a == b:
            continue
        r, k = o.sequential_walk(succ[a:b].astype(np.uint64), w[a:b])
        roots.append(r.astype(np.uint32))
        ranks.append(k)
    return np.concatenate(roots), np.concatenate(ranks)


@pytest.mark.parametrize("weighted", [True, False])
def test_segmented_narrow_group(ctx, weighted):
    """A batch of 2^22 elements or more (below 2^30) is ranked with 8-byte
    records."""
    say("seg narrow group", weighted)
    rng = np.random.default_rng(23)
    sizes = np.floor(2.0 ** rng.uniform(10, 16, size=1500)).astype(np.uint64)
    off, seeds, local = make_batch(sizes, 5)
    assert off[-1] >= 1 << 22
    succ, w = ctx.gen_lists_segmented(dev(off), dev(seeds), dev(local), weighted=weighted)
    before = ctx.record_stats()
    root, rank = ctx.rank_lists_segmented(succ, dev(off), w if weighted else None)
    after = ctx.record_stats()
    er, ek = expected(off, seeds, local, weighted)
    assert (host(root) == er).all() and (host(rank) == ek).all()
    assert after["narrow_calls"] - before["narrow_calls"] == 1


@pytest.mark.parametrize("weight", [1 << 33, 1 << 29])
def test_segmented_narrow_fallbacks(ctx, weight):
    """One segment with weights beyond int32 (the batch is packed wide), or
    whose claim distances outgrow int32 (the narrow ranking is redone
    wide): results unchanged."""
    say("seg narrow fallback", weight)
    rng = np.random.default_rng(29)
    sizes = [int(x) for x in np.floor(2.0 ** rng.uniform(12, 17, size=120))]
    off, succ, w = host_batch(sizes, 300)
    assert off[-1] >= 1 << 22
    a, b = int(off[7]), int(off[8])
    w[a:b] = np.where(w[a:b] != 0, np.int64(weight), np.int64(0))
    before = ctx.record_stats()
    root, rank = ctx.rank_lists_segmented(dev(succ), dev(off), dev(w))
    after = ctx.record_stats()
    er, ek = host_expected(off, succ, w)
    assert (host(root) == er).all() and (host(rank) == ek).all()
    assert after["narrow_calls"] == before["narrow_calls"]
    if weight < (1 << 31):
        assert after["narrow_redos"] - before["narrow_redos"] == 1


def test_segmented_narrow_large(ctx):
    """About 9e8 elements (below 2^30 ids): one ranking with 8-byte records;
    every element checked through the local identities."""
    import torch
    from test_gpu_fullsize import segmented_identities_hold
    rng = np.random.default_rng(31)
    sizes = np.floor(2.0 ** rng.uniform(16, 22, size=900)).astype(np.uint64)
    off = np.zeros(len(sizes) + 1, np.uint64)
    off[1:] = np.cumsum(sizes)
    assert (1 << 29) < off[-1] < (1 << 30)
    say("seg narrow large", int(off[-1]))
    seeds = rng.integers(1, 1 << 62, size=len(sizes), dtype=np.uint64)
    d_off = torch.from_numpy(off).to("cuda")
    succ, w = ctx.gen_lists_segmented(d_off, torch.from_numpy(seeds).to("cuda"), None, weighted=True)
    before = ctx.record_stats()
    root, rank = ctx.rank_lists_segmented(succ, d_off, w)
    after = ctx.record_stats()
    assert after["narrow_calls"] - before["narrow_calls"] == 1
    assert segmented_identities_hold(succ, d_off, root, rank, w)
    del succ, w, root, rank
    ctx.trim()
    torch.cuda.empty_cache()


def test_segmented_segment_beyond_narrow_ids(ctx):
    """A batch with one segment beyond 2^30 elements between two small ones:
    some successors lie 2^29 or more away from their element, more than an
    8-byte record's relative successor holds, so the batch is ranked with
    16-byte records."""
    import torch
    from test_gpu_fullsize import segmented_identities_hold
    sizes = [10, (1 << 30) + 5, 1000]
    off = np.zeros(len(sizes) + 1, np.uint64)
    off[1:] = np.cumsum(sizes)
    say("seg huge segment", int(off[-1]))
    d_off = torch.from_numpy(off).to("cuda")
    seeds = torch.tensor([3, 5, 7], dtype=torch.uint64, device="cuda")
    succ, w = ctx.gen_lists_segmented(d_off, seeds, None, weighted=False)
    del w
    before = ctx.record_stats()
    root, rank = ctx.rank_lists_segmented(succ, d_off, None)
    after = ctx.record_stats()
    assert after["narrow_calls"] == before["narrow_calls"]
    assert segmented_identities_hold(succ, d_off, root, rank)
    del succ, root, rank
    ctx.trim()
    torch.cuda.empty_cache()


def test_segmented_narrow_beyond_2e30(ctx):
    """A batch of more than 2^30 elements whose segments are all below 2^21
    is still ranked with 8-byte records (successors stored relative to their
    element); every element checked through the local identities."""
    import torch
    from test_gpu_fullsize import segmented_identities_hold
    rng = np.random.default_rng(37)
    sizes = np.floor(2.0 ** rng.uniform(16, 21, size=2200)).astype(np.uint64)
    off = np.zeros(len(sizes) + 1, np.uint64)
    off[1:] = np.cumsum(sizes)
    assert off[-1] > (1 << 30)
    say("seg narrow beyond 2^30", int(off[-1]))
    seeds = rng.integers(1, 1 << 62, size=len(sizes), dtype=np.uint64)
    d_off = torch.from_numpy(off).to("cuda")
    succ, w = ctx.gen_lists_segmented(d_off, torch.from_numpy(seeds).to("cuda"), None, weighted=True)
    before = ctx.record_stats()
    root, rank = ctx.rank_lists_segmented(succ, d_off, w)
    after = ctx.record_stats()
    assert after["narrow_calls"] - before["narrow_calls"] == 1
    assert segmented_identities_hold(succ, d_off, root, rank, w)
    del succ, w, root, rank
    ctx.trim()
    torch.cuda.empty_cache()


def test_segmented_narrow_offsets_near_2e29(ctx):
    """One random list of 2^29 - 3 elements between two small ones: its
    successors lie up to nearly 2^29 before or after their element, the
    most an 8-byte record's relative successor holds; the batch is ranked
    with 8-byte records and every element checked."""
    import torch
    from test_gpu_fullsize import segmented_identities_hold
    sizes = [10, (1 << 29) - 3, 1000]
    off = np.zeros(len(sizes) + 1, np.uint64)
    off[1:] = np.cumsum(sizes)
    say("seg offsets near 2^29", int(off[-1]))
    d_off = torch.from_numpy(off).to("cuda")
    seeds = torch.tensor([11, 13, 17], dtype=torch.uint64, device="cuda")
    succ, w = ctx.gen_lists_segmented(d_off, seeds, None, weighted=False)
    del w
    loc = succ[10:10 + sizes[1]].view(torch.int32).to(torch.int64) & 0xFFFFFFFF
    delta = loc - torch.arange(sizes[1], device="cuda", dtype=torch.int64)
    assert int(delta.max()) > (1 << 29) - (1 << 20) and int(delta.min()) < -(1 << 29) + (1 << 20)
    del loc, delta
    before = ctx.record_stats()
    root, rank = ctx.rank_lists_segmented(succ, d_off, None)
    after = ctx.record_stats()
    assert after["narrow_calls"] - before["narrow_calls"] == 1
    assert segmented_identities_hold(succ, d_off, root, rank)
    del succ, root, rank
    ctx.trim()
    torch.cuda.empty_cache()


def test_segmented_tuning_table(ctx):
    """Segmented batches have their own tuning table (classes count a
    batch's elements; below its first class the list table applies): it
    changes the knobs of segmented rankings only, never a result."""
    say("segmented tuning")
    builtin = ctx.tuning_get(1 << 31, kind="segmented")
    assert (builtin["spacing_top"], builtin["spacing_deep"], builtin["chase_blocks_per_cu"]) == (256, 16, 2)
    knobs = ("spacing_top", "spacing_deep", "chase_blocks_per_cu", "chase_chunk")
    below, as_list = ctx.tuning_get(1 << 20, kind="segmented"), ctx.tuning_get(1 << 20)
    assert [below[k] for k in knobs] == [as_list[k] for k in knobs]
    list_before = ctx.tuning_get(1 << 22)
    rng = np.random.default_rng(43)
    sizes = [int(x) for x in np.floor(2.0 ** rng.uniform(8, 14, size=900))]
    off, succ, w = host_batch(sizes, 44)
    er, ek = host_expected(off, succ, w)
    try:
        for rows in ([(12, 4, 4, 1, 64)], [(12, 512, 32, 1, 1024)], [(30, 8, 8, 8, 64)]):
            ctx.tuning_set(rows, kind="segmented")
            assert ctx.tuning_get(1 << 22) == list_before
            root, rank = ctx.rank_lists_segmented(dev(succ), dev(off), dev(w))
            assert (host(root) == er).all() and (host(rank) == ek).all(), rows
    finally:
        ctx.tuning_set([], kind="segmented")
    assert ctx.tuning_get(1 << 31, kind="segmented")["table_rows"] == 0
