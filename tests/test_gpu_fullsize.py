"""Full-size parity through size-independent properties, checked on the
device.  For a valid list, rank[v] = w[v] + rank[succ[v]] and root[v] =
root[succ[v]] off the tails, rank = 0 and root = self on them: by induction
along each list these local identities pin every output, so they are a
complete check at sizes the CPU oracle would take minutes for.  Trees: the
same with w = 1 (depth) over parent pointers.  Sizes: the 2^27-element
weighted list and the 2^26-vertex tree of the headline workloads, and a
2^31-element list (ids and ranks beyond 32 bits)."""
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
