"""Full-size parity through size-independent properties, checked on the
device.  For a valid list, rank[v] = w[v] + rank[succ[v]] and root[v] =
root[succ[v]] off the tails, rank = 0 and root = self on them: by induction
along each list these local identities pin every output, so they are a
complete check at sizes the CPU oracle would take minutes for.  Trees: the
same with w = 1 (depth) over parent pointers, and for a segmented batch with
each local successor taken inside its segment.  Sizes: the 2^27-element
weighted list and the 2^26-vertex tree of the headline workloads, a
2^31-element list (ids and ranks beyond 32 bits) and the 20,000-list
segmented batch (about 2.5e9 elements)."""
import pytest

from test_gpu_rank import say

pytestmark = pytest.mark.gpu


def local_identities_hold(succ, w, root, rank):
    import torch
    si = succ.view(torch.int64)
    ri = root.view(torch.int64)
    ok = True
    step = 1 << 27
    for a in range(0, succ.numel(), step):
        b = min(a + step, succ.numel())
        s = si[a:b]
        idx = torch.arange(a, b, device=succ.device, dtype=torch.int64)
        tail = s == idx
        ww = torch.ones_like(s) if w is None else w[a:b]
        ok &= bool(torch.all(torch.where(tail, rank[a:b] == 0, rank[a:b] == ww + rank[s])).item())
        ok &= bool(torch.all(torch.where(tail, ri[a:b] == idx, ri[a:b] == ri[s])).item())
    return ok


def planted_error_is_caught(succ, w, root, rank):
    """The check itself must fail on a wrong rank (one element off by one)."""
    rank[succ.numel() // 3] += 1
    bad = local_identities_hold(succ, w, root, rank)
    rank[succ.numel() // 3] -= 1
    return not bad


@pytest.mark.parametrize("n", [1 << 27, 1 << 31])
def test_list_full_size(ctx, n):
    import torch
    say("full-size list", n)
    succ, w = ctx.gen_list(n, 27, weighted=True)
    root = torch.empty(n, dtype=torch.uint64, device="cuda")
    rank = torch.empty(n, dtype=torch.int64, device="cuda")
    ctx.rank_list(succ, w, root, rank)
    assert local_identities_hold(succ, w, root, rank)
    assert planted_error_is_caught(succ, w, root, rank)
    # one list: every root is the same tail
    assert bool(torch.all(root == root[0]).item())
    del succ, w, root, rank
    ctx.trim()
    torch.cuda.empty_cache()


def test_tree_full_size(ctx):
    import torch
    n = 1 << 26
    say("full-size tree", n)
    parent = ctx.gen_tree(n, 26)
    root = torch.empty(n, dtype=torch.uint64, device="cuda")
    depth = torch.empty(n, dtype=torch.int64, device="cuda")
    ctx.root_tree(parent, root, depth)
    assert local_identities_hold(parent, None, root, depth)
    assert planted_error_is_caught(parent, None, root, depth)
    del parent, root, depth
    ctx.trim()
    torch.cuda.empty_cache()


def segmented_identities_hold(succ, off, root, rank, w=None):
    """Segment s owns [off[s], off[s + 1]); a local successor j of element v is
    the global element off[s] + j.  Off the tails: rank[v] = w[v] + rank[g]
    (w = 1 when no weights are given) and root[v] = root[g] (both local, same
    segment); on a tail: rank 0, root = its own local id."""
    import torch
    off64 = off.view(torch.int64)
    starts = off64[:-1].contiguous()
    ok = True
    step = 1 << 27
    total = succ.numel()
    for a in range(0, total, step):
        b = min(a + step, total)
        idx = torch.arange(a, b, device=succ.device, dtype=torch.int64)
        seg = torch.searchsorted(starts, idx, right=True) - 1
        base = off64[seg]
        size = off64[seg + 1] - base
        loc = succ[a:b].view(torch.int32).to(torch.int64) & 0xFFFFFFFF
        ok &= bool(torch.all(loc < size).item())
        g = base + loc
        tail = g == idx
        rt = root[a:b].view(torch.int32).to(torch.int64) & 0xFFFFFFFF
        rg = root.view(torch.int32)[g].to(torch.int64) & 0xFFFFFFFF
        ww = 1 if w is None else w[a:b]
        ok &= bool(torch.all(torch.where(tail, rank[a:b] == 0, rank[a:b] == ww + rank[g])).item())
        ok &= bool(torch.all(torch.where(tail, rt == idx - base, rt == rg)).item())
    return ok


def test_segmented_full_size(ctx):
    """20,000 lists, sizes log-uniform in [2^8, 2^20] (about 2.5e9 elements),
    30 % locality-biased, unit weights, one segmented call; every element
    checked through the local identities."""
    import numpy as np
    import torch
    rng = np.random.default_rng(2027)
    sizes = np.floor(2.0 ** rng.uniform(8, 20, size=20000)).astype(np.uint64)
    off = np.zeros(len(sizes) + 1, np.uint64)
    off[1:] = np.cumsum(sizes)
    seeds = rng.integers(1, 1 << 62, size=len(sizes), dtype=np.uint64)
    local = (rng.random(len(sizes)) < 0.3).astype(np.uint8)
    say("full-size segmented", int(off[-1]))
    d_off = torch.from_numpy(off).to("cuda")
    succ, w = ctx.gen_lists_segmented(d_off, torch.from_numpy(seeds).to("cuda"),
                                      torch.from_numpy(local).to("cuda"), weighted=False)
    del w
    root, rank = ctx.rank_lists_segmented(succ, d_off, None)
    assert segmented_identities_hold(succ, d_off, root, rank)
    k = succ.numel() // 3
    rank[k] += 1
    assert not segmented_identities_hold(succ, d_off, root, rank)
    del succ, root, rank
    ctx.trim()
    torch.cuda.empty_cache()
