"""Parity of the five workload configurations at their stated sizes.

- Single random list, n = 2^20, unit weights, pointer doubling: every output
  against the closed form and against the oracle's restatement of the
  reference's sync pointer doubling on 4 emulated ranks.
- Single random list, n = 2^27, int64 weights in [1, 100], pointer doubling
  (the sparse ruling set at this size is covered in test_gpu_fullsize.py):
  every output through the local identities, which pin a valid list's ranks.
- Random tree, 2^26 vertices, with the radix-sorted tour build: the same.
- The serving stream: 60 % list jobs (weights in [1, 100]) and 40 % tree
  jobs, sizes log-uniform in [2^10, 2^24], generated on the device, through
  the coalesced server; every job against its closed form on the host, and
  equal to the same job run alone.
- Stream ordering: inputs written by torch kernels on torch's stream and
  ranked with no explicit synchronisation.
"""
import numpy as np
import pytest

import oracle_lib as o
from test_gpu_fullsize import local_identities_hold, planted_error_is_caught

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def host(t):
    return t.cpu().numpy()


def test_unit_list_2e20_pointer_doubling(ctx):
    """2^20 elements from a seeded random permutation, unit weights
    (set_initial_ranking_state, list_ranking.cpp:60-80), ranked by sync
    pointer doubling."""
    import torch
    n = 1 << 20
    say("unit list pd", n)
    succ, _ = o.gen_list(n, 20, False)
    d_succ = torch.from_numpy(succ).cuda()
    root, rank = ctx.rank_list(d_succ, None, algorithm="pd")
    root, rank = host(root), host(rank)
    ones = np.where(succ == np.arange(n, dtype=np.uint64), 0, 1).astype(np.int64)
    er, ek = o.list_closed_form(n, 20, ones)
    assert (root == er).all() and (rank == ek).all()
    init_root, init_rank = o.initial_ranking_state(succ)
    assert (init_rank == ones).all()
    cr, ck, rounds = o.pointer_doubling(init_root, init_rank, 4)
    assert (root == cr).all() and (rank == ck).all()
    assert rounds == 20


def test_weighted_list_2e27_pointer_doubling(ctx):
    import torch
    n = 1 << 27
    say("weighted list pd", n)
    succ, w = ctx.gen_list(n, 2027, weighted=True)
    root = torch.empty(n, dtype=torch.uint64, device="cuda")
    rank = torch.empty(n, dtype=torch.int64, device="cuda")
    ctx.rank_list(succ, w, root, rank, algorithm="pd")
    assert local_identities_hold(succ, w, root, rank)
    assert planted_error_is_caught(succ, w, root, rank)
    assert bool(torch.all(root == root[0]).item())
    del succ, w, root, rank
    ctx.trim()
    torch.cuda.empty_cache()


def test_tree_2e26_sorted_tour(ctx):
    import torch
    n = 1 << 26
    say("tree sorted tour", n)
    parent = ctx.gen_tree(n, 26)
    root = torch.empty(n, dtype=torch.uint64, device="cuda")
    depth = torch.empty(n, dtype=torch.int64, device="cuda")
    ctx.root_tree(parent, root, depth, tour_build="sorted")
    assert local_identities_hold(parent, None, root, depth)
    assert planted_error_is_caught(parent, None, root, depth)
    del parent, root, depth
    ctx.trim()
    torch.cuda.empty_cache()


def serving_stream(count, seed, lo=10, hi=24):
    """Job kinds, sizes and seeds of the serving stream; the two largest
    sizes are forced so that batches near their 2^26-element cap and a job
    ranked alone in its batch both occur."""
    rng = np.random.default_rng(seed)
    jobs = []
    for k in range(count):
        n = int(2 ** rng.uniform(lo, hi))
        kind = "list" if rng.random() < 0.6 else "tree"
        jobs.append((kind, n, 5000 + k))
    jobs[3] = ("list", 1 << hi, jobs[3][2])
    jobs[7] = ("tree", 1 << hi, jobs[7][2])
    return jobs


def test_serving_stream_full_sizes(ctx):
    import torch
    import kascade_amd
    jobs = serving_stream(80, 11)
    total = sum(n for _, n, _ in jobs)
    say("serving stream", len(jobs), "jobs", total, "elements")
    inputs = []
    for kind, n, seed in jobs:
        if kind == "list":
            s, w = ctx.gen_list(n, seed, weighted=True)
        else:
            s, w = ctx.gen_tree(n, seed), None
        inputs.append((s, w))
    torch.cuda.synchronize()
    sv = kascade_amd.Server(ctx, num_slots=3)
    tickets = []
    for (kind, n, _), (s, w) in zip(jobs, inputs):
        ro = torch.empty(n, dtype=torch.uint64, device="cuda")
        ra = torch.empty(n, dtype=torch.int64, device="cuda")
        t, _, _ = sv.submit(kind, s, w, ro, ra)
        tickets.append((t, ro, ra))
    for t, _, _ in tickets:
        sv.wait(t)
    st = sv.stats()
    sv.close()
    assert st["jobs_completed"] == len(jobs) and st["jobs_failed"] == 0
    assert st["batches_launched"] < len(jobs)
    say("batches", st["batches_launched"], "reruns", st["batches_rerun"])
    for (kind, n, seed), (s, w), (t, ro, ra) in zip(jobs, inputs, tickets):
        r, k = host(ro), host(ra)
        if kind == "list":
            er, ek = o.list_closed_form(n, seed, host(w))
        else:
            er, ek = o.tree_closed_form(n, seed)
        assert (r == er).all() and (k == ek).all(), (kind, n, seed)
        if n >= 1 << 22:
            # the same job alone (one job per launch) is identical
            if kind == "list":
                ar, ak = ctx.rank_list(s, w)
            else:
                ar, ak = ctx.root_tree(s)
            assert torch.equal(ar.view(torch.int64), ro.view(torch.int64)) and torch.equal(ak, ra)
    del inputs, tickets
    ctx.trim()
    torch.cuda.empty_cache()


def test_inputs_from_torch_kernels_without_synchronize(ctx):
    """The list is built by torch kernels on torch's current stream and the
    outputs are pre-filled by torch, then ranked with no synchronisation in
    between: the call must wait for that work (it runs after an event wait on
    torch's stream)."""
    import torch
    n = 1 << 24
    say("torch-produced inputs", n)
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    for _ in range(3):
        perm = torch.randperm(n, device="cuda", generator=g)
        succ = torch.empty(n, dtype=torch.int64, device="cuda")
        succ[perm[:-1]] = perm[1:]
        succ[perm[-1]] = perm[-1]
        w = (perm % 100 + 1).to(torch.int64)
        w[perm[-1]] = 0
        root = torch.full((n,), -1, dtype=torch.int64, device="cuda").view(torch.uint64)
        rank = torch.full((n,), -7, dtype=torch.int64, device="cuda")
        ctx.rank_list(succ.view(torch.uint64), w, root, rank)
        # expected: suffix sums of the weights in list order, computed by torch
        wl = w[perm]
        suffix = torch.flip(torch.cumsum(torch.flip(wl, [0]), 0), [0])
        exp = torch.empty_like(rank)
        exp[perm] = suffix
        assert torch.equal(rank, exp)
        assert bool(torch.all(root.view(torch.int64) == perm[-1]).item())
