"""The per-level trace (the counterpart of RulerTrace): its counts must be
consistent with each other and with the input, level by level."""
import math

import numpy as np
import pytest

import oracle_lib as o
from test_gpu_rank import dev, host, random_forest_lists, say

pytestmark = pytest.mark.gpu


def check_levels(trace, n):
    srs = [t for t in trace if t["kind"] == "srs"]
    assert srs and srs[0]["level"] == 0 and srs[0]["n"] == n
    for t in srs:
        assert t["rulers"] == -(-t["n"] // t["spacing"])                   # one ruler per block of K ids
        chains = round(t["chain_length_avg"] * t["rulers"])
        assert chains + t["tail_rulers"] + t["unreached"] == t["n"]      # every element counted once
        assert 1 <= t["chain_length_min"] <= t["chain_length_avg"] <= t["chain_length_max"]
    # the ruler sublist of level L is the n of the next main-recursion entry
    main = [t for t in trace if t["subproblem"] == 0]
    for a, b in zip(main, main[1:]):
        if a["kind"] == "srs":
            assert b["n"] == a["rulers"] + a["tail_rulers"] and b["level"] == a["level"] + 1


@pytest.mark.parametrize("n", [10000, 1 << 20, 1 << 23])
def test_trace_single_list(ctx, n):
    say("trace list", n)
    succ, w = ctx.gen_list(n, 3, weighted=True)
    ctx.trace(True)
    root, rank = ctx.rank_list(succ, w, algorithm="srs")
    tr = ctx.trace_read()
    ctx.trace(False)
    check_levels(tr, n)
    assert tr[0]["tail_rulers"] <= 1                                     # one list: one tail
    er, ek = o.list_closed_form(n, 3, host(w))
    assert (host(root) == er).all() and (host(rank) == ek).all()


def test_trace_forest_and_unclaimed(ctx):
    say("trace forest")
    succ, w = random_forest_lists(200000, 3000, 5)
    ctx.trace(True)
    ctx.rank_list(dev(succ), dev(w), algorithm="srs")
    tr = ctx.trace_read()
    ctx.trace(False)
    check_levels(tr, 200000)
    assert tr[0]["unreached"] > 0 and any(t["subproblem"] == 1 for t in tr)


@pytest.mark.parametrize("n", [2, 1000, 1 << 20])
def test_trace_pointer_doubling_rounds(ctx, n):
    """Jacobi doubling on one list: an element at distance d is done after the
    first round k with 2^k >= d, so max(1, ceil(log2(n - 1))) rounds."""
    say("trace pd", n)
    succ, w = ctx.gen_list(n, 8, weighted=False)
    ctx.trace(True)
    ctx.rank_list(succ, None, algorithm="pd")
    tr = ctx.trace_read()
    ctx.trace(False)
    pd = [t for t in tr if t["kind"] == "pd"]
    if n <= 8192:                                  # the one-workgroup solver takes small inputs
        assert [t["kind"] for t in tr] == ["small"]
    else:
        assert len(pd) == 1 and pd[0]["rounds"] == max(1, math.ceil(math.log2(n - 1)))


def test_trace_off_records_nothing(ctx):
    succ, w = ctx.gen_list(100000, 1, weighted=True)
    ctx.trace(False)
    ctx.rank_list(succ, w)
    assert ctx.trace_read() == []
