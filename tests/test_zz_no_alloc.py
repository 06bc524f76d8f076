"""kc_rank_list_noalloc: ranking with no device allocation (scratch from the
library's static array), checked against kc_rank_list on the same inputs,
bit for bit, on calls in a row (the second smaller, so it runs over the
first call's leftovers), on forests up to the maximum size (a forest of
singletons needs the most scratch), and on its limits (n above the maximum,
host arrays)."""
import numpy as np
import pytest

import oracle_lib as o

pytestmark = pytest.mark.gpu

MAX_N = 1 << 20


def say(*a):
    print(*a, flush=True)


def forest(n, nlists, seed):
    """A forest of nlists random lists over n ids, weights in [1, 100]."""
    rng = np.random.default_rng(seed)
    perm = rng.permutation(n).astype(np.uint64)
    end = np.zeros(n, bool)
    if nlists > 1:
        end[rng.choice(np.arange(1, n), size=nlists - 1, replace=False) - 1] = True
    end[n - 1] = True
    succ = np.empty(n, np.uint64)
    w = np.empty(n, np.int64)
    succ[perm] = np.where(end, perm, np.roll(perm, -1))
    w[perm] = np.where(end, 0, rng.integers(1, 101, size=n))
    return succ, w


def check_against_rank_list(ctx, succ, w, **cfg):
    import torch
    n = len(succ)
    ds = torch.from_numpy(succ).cuda()
    dw = torch.from_numpy(w).cuda() if w is not None else None
    root = torch.full((n,), -1, dtype=torch.int64, device="cuda").view(torch.uint64)
    rank = torch.full((n,), -1, dtype=torch.int64, device="cuda")
    ctx.rank_list_noalloc(ds, dw, root, rank, **cfg)
    er, ek = ctx.rank_list(ds, dw, **cfg)
    assert torch.equal(root.view(torch.int64), er.view(torch.int64))
    assert torch.equal(rank, ek)
    return root, rank


def test_two_calls_match_rank_list(ctx):
    """Two calls in a row, the second with a smaller input."""
    for n, seed in ((MAX_N, 1), (300007, 2)):
        say("noalloc list", n)
        succ, w = o.gen_list(n, seed, True)
        root, rank = check_against_rank_list(ctx, succ, w)
        er, ek = o.list_closed_form(n, seed, w)
        assert (root.cpu().numpy() == er).all() and (rank.cpu().numpy() == ek).all()


@pytest.mark.parametrize("algo", ["srs", "pd"])
@pytest.mark.parametrize("nlists", [1, 3, 1000, 60000, MAX_N // 2, MAX_N])
def test_forests_at_max_size(ctx, algo, nlists):
    say("noalloc forest", algo, nlists)
    succ, w = forest(MAX_N, nlists, nlists)
    check_against_rank_list(ctx, succ, w, algorithm=algo)


def test_unit_weights_and_small_sizes(ctx):
    for n in (1, 2, 8192, 8193, 65536):
        say("noalloc small", n)
        succ, _ = o.gen_list(n, 40 + n, False)
        check_against_rank_list(ctx, succ, None)


def test_no_device_memory_allocated(ctx):
    """Device memory in use is the same before and after the calls."""
    import torch
    n = MAX_N
    succ, w = o.gen_list(n, 9, True)
    ds, dw = torch.from_numpy(succ).cuda(), torch.from_numpy(w).cuda()
    root = torch.empty(n, dtype=torch.uint64, device="cuda")
    rank = torch.empty(n, dtype=torch.int64, device="cuda")
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    scratch0 = ctx.scratch_bytes()
    for _ in range(3):
        ctx.rank_list_noalloc(ds, dw, root, rank)
    torch.cuda.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    assert free1 == free0
    assert ctx.scratch_bytes() == scratch0
    er, ek = o.list_closed_form(n, 9, w)
    assert (root.cpu().numpy() == er).all() and (rank.cpu().numpy() == ek).all()


def test_limits(ctx):
    import torch
    import kascade_amd
    n = MAX_N + 1
    succ = torch.arange(n, dtype=torch.int64, device="cuda").view(torch.uint64)
    root = torch.empty(n, dtype=torch.uint64, device="cuda")
    rank = torch.empty(n, dtype=torch.int64, device="cuda")
    with pytest.raises(kascade_amd.KascadeError) as e:
        ctx.rank_list_noalloc(succ, None, root, rank)
    assert e.value.status == kascade_amd.KC_EINVAL
    hs = np.arange(16, dtype=np.uint64)
    with pytest.raises(kascade_amd.KascadeError) as e:
        ctx.rank_list_noalloc(hs, None, np.empty(16, np.uint64), np.empty(16, np.int64))
    assert e.value.status == kascade_amd.KC_EINVAL
    # the context still works afterwards
    succ, w = o.gen_list(5000, 3, True)
    check_against_rank_list(ctx, succ, w)
