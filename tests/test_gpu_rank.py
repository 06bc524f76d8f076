"""GPU parity of list ranking through the C ABI: every output bit-exact against
the CPU oracle (sequential walk, closed forms, PD rounds, pack KAT)."""
import json
import os
import sys

import numpy as np
import pytest

import oracle_lib as o

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def say(*a):
    print(*a, flush=True)


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def host(t):
    return t.cpu().numpy()


def random_forest_lists(n, nlists, seed, weighted=True):
    rng = np.random.default_rng(seed)
    perm = rng.permutation(n)
    cuts = np.sort(rng.choice(np.arange(1, n), size=nlists - 1, replace=False)) if nlists > 1 else []
    succ = np.arange(n, dtype=np.uint64)
    w = np.zeros(n, np.int64)
    for seg in np.split(perm, cuts):
        if len(seg) > 1:
            succ[seg[:-1]] = seg[1:]
            w[seg[:-1]] = rng.integers(1, 101, size=len(seg) - 1) if weighted else 1
    return succ, w


@pytest.mark.parametrize("n", [1, 2, 17, 1000, 65537])
@pytest.mark.parametrize("weighted,local", [(True, False), (False, False), (True, True)])
def test_gen_list_matches_cpu(ctx, n, weighted, local):
    say("gen_list", n, weighted, local)
    s, w = ctx.gen_list(n, 1234 + n, weighted=weighted, local=local)
    es, ew = o.gen_list(n, 1234 + n, weighted, local=local)
    assert (host(s) == es).all() and (host(w) == ew).all()


@pytest.mark.parametrize("n", [1, 2, 33, 4099, 100000])
def test_gen_tree_matches_cpu(ctx, n):
    say("gen_tree", n)
    p = ctx.gen_tree(n, 77 + n)
    assert (host(p) == o.gen_tree(n, 77 + n)).all()


ALGOS = ["pd", "srs"]


@pytest.mark.parametrize("algo", ALGOS)
@pytest.mark.parametrize("n", [1, 2, 3, 100, 4096, 4097, 8192, 8193, 10000, 1 << 16, 300007])
def test_rank_random_list(ctx, algo, n):
    say("rank", algo, n)
    succ, w = o.gen_list(n, 99 + n, True)
    root, rank = ctx.rank_list(dev(succ), dev(w), algorithm=algo)
    er, ek = o.list_closed_form(n, 99 + n, w)
    assert (host(root) == er).all()
    assert (host(rank) == ek).all()


@pytest.mark.parametrize("algo", ALGOS)
@pytest.mark.parametrize("kind", ["signed", "zero", "wide"])
@pytest.mark.parametrize("n,nlists", [(100003, 1), (1 << 20, 1), (200000, 5000)])
def test_rank_weight_ranges(ctx, algo, kind, n, nlists):
    """Weights are any int64 (rank_t): signed, all zero, and wide (|w| up to
    2^40, |rank| up to about 2^60); tails carry weight 0."""
    say("weights", algo, kind, n, nlists)
    rng = np.random.default_rng(n + nlists)
    succ, _ = random_forest_lists(n, nlists, n + 7, weighted=False)
    if kind == "signed":
        w = rng.integers(-1000, 1001, size=n, dtype=np.int64)
    elif kind == "zero":
        w = np.zeros(n, np.int64)
    else:
        w = rng.integers(-(1 << 40), (1 << 40) + 1, size=n, dtype=np.int64)
    w[succ == np.arange(n, dtype=np.uint64)] = 0
    root, rank = ctx.rank_list(dev(succ), dev(w), algorithm=algo)
    er, ek = o.sequential_walk(succ, w)
    assert (host(root) == er).all() and (host(rank) == ek).all()


@pytest.mark.parametrize("nlists", [1, 2])
@pytest.mark.parametrize("weight", [2047, 2048, 2049, -2048, -2049])
def test_rank_near_int32_edge(ctx, weight, nlists):
    """Constant weights whose ranks end just inside or just beyond int32
    (|weight| x 2^20 around 2^31): the rulers' ranks are gathered from a
    4-byte table (one list) or a packed (root, rank) table (two lists) only
    while every ruler rank fits; either way the results are exact."""
    say("int32 edge", weight, nlists)
    n = (1 << 20) * nlists + 5
    succ, _ = random_forest_lists(n, nlists, 4242 + nlists, weighted=False)
    w = np.where(succ == np.arange(n, dtype=np.uint64), 0, weight).astype(np.int64)
    root, rank = ctx.rank_list(dev(succ), dev(w), algorithm="srs")
    er, ek = o.sequential_walk(succ, w)
    assert (host(root) == er).all() and (host(rank) == ek).all()


@pytest.mark.parametrize("spacing,deep", [(2, 2), (4, 16), (64, 4), (256, 32), (4096, 8)])
def test_rank_spacings(ctx, spacing, deep):
    n = 200003
    say("spacing", spacing, deep)
    succ, w = o.gen_list(n, 5, True)
    root, rank = ctx.rank_list(dev(succ), dev(w), algorithm="srs", ruler_spacing=spacing,
                               ruler_spacing_deep=deep)
    er, ek = o.list_closed_form(n, 5, w)
    assert (host(root) == er).all() and (host(rank) == ek).all()


@pytest.mark.parametrize("algo", ALGOS)
@pytest.mark.parametrize("n,nlists", [(10, 10), (5000, 1), (5000, 2500), (100000, 7), (100000, 30000),
                                      (300000, 300)])
def test_rank_forest(ctx, algo, n, nlists):
    say("forest", algo, n, nlists)
    succ, w = random_forest_lists(n, nlists, n + nlists)
    root, rank = ctx.rank_list(dev(succ), dev(w), algorithm=algo)
    er, ek = o.sequential_walk(succ, w)
    assert (host(root) == er).all() and (host(rank) == ek).all()


@pytest.mark.parametrize("algo", ALGOS)
def test_rank_local_list(ctx, algo):
    n = 500009
    say("local", algo)
    succ, w = o.gen_list(n, 8, True, local=True)
    root, rank = ctx.rank_list(dev(succ), dev(w), algorithm=algo)
    er, ek = o.sequential_walk(succ, w)
    assert (host(root) == er).all() and (host(rank) == ek).all()


def test_identity_and_sorted_lists(ctx):
    say("identity")
    n = 70000
    succ = np.arange(n, dtype=np.uint64)            # all singletons
    root, rank = ctx.rank_list(dev(succ), None)
    assert (host(root) == succ).all() and (host(rank) == 0).all()
    succ = np.minimum(np.arange(1, n + 1), n - 1).astype(np.uint64)   # 0->1->...->n-1
    root, rank = ctx.rank_list(dev(succ), None)
    assert (host(root) == n - 1).all() and (host(rank) == np.arange(n - 1, -1, -1)).all()
    succ = np.maximum(np.arange(-1, n - 1), 0).astype(np.uint64)      # n-1 -> ... -> 0
    root, rank = ctx.rank_list(dev(succ), None)
    assert (host(root) == 0).all() and (host(rank) == np.arange(n)).all()


def test_unit_weights_match_initial_state(ctx):
    say("unit")
    n = 123457
    succ, _ = o.gen_list(n, 3, False)
    r0, k0 = ctx.set_initial_ranking_state(dev(succ))
    er0, ek0 = o.initial_ranking_state(succ)
    assert (host(r0) == er0).all() and (host(k0) == ek0).all()
    root, rank = ctx.rank_list(dev(succ), None)
    er, ek = o.sequential_walk(succ, ek0)
    assert (host(root) == er).all() and (host(rank) == ek).all()


@pytest.mark.parametrize("algo", ALGOS)
def test_in_place_and_host_arrays(ctx, algo):
    say("inplace", algo)
    n = 77777
    succ, w = o.gen_list(n, 21, True)
    er, ek = o.list_closed_form(n, 21, w)
    s, r = dev(succ), dev(w)
    (ctx.pointer_doubling if algo == "pd" else ctx.sparse_ruling_set)(s, r)
    assert (host(s) == er).all() and (host(r) == ek).all()
    hs, hr = succ.copy(), w.copy()                      # host arrays are staged
    (ctx.pointer_doubling if algo == "pd" else ctx.sparse_ruling_set)(hs, hr)
    assert (hs == er).all() and (hr == ek).all()


@pytest.mark.parametrize("n", [1, 9, 1023, 70001])
def test_pd_rounds_match_oracle(ctx, n):
    say("pd_rounds", n)
    succ, w = random_forest_lists(n, max(1, n // 1000), n)
    for k in [0, 1, 2, 5, 9, 17]:
        s, r = ctx.pd_rounds(dev(succ), dev(w), k)
        es, er, _ = o.pd_rounds(succ, w, k)
        assert (host(s) == es).all() and (host(r) == er).all(), k


def test_pack_kat(ctx):
    say("pack kat")
    kat = json.load(open(os.path.join(GOLDEN, "pack_kat.json")))
    off = kat["offsets"]
    succ = np.concatenate([np.array(x, np.uint64) for x in kat["succ_per_rank"]])
    rank = np.concatenate([np.array(x, np.int64) for x in kat["rank_per_rank"]])
    active = np.concatenate([np.array(a, np.uint64) + off[r] for r, a in enumerate(kat["active_local_per_rank"])])
    ps, pr = ctx.pack(dev(succ), dev(rank), dev(active))
    assert host(ps).tolist() == sum(kat["expected_packed_succ_per_rank"], [])
    assert host(pr).tolist() == sum(kat["expected_packed_rank_per_rank"], [])
    ps_h = np.full(len(active), 6, np.uint64)
    pr_h = np.array([6 - i - 2 * r for r, a in enumerate(kat["active_local_per_rank"]) for i in range(len(a))],
                    np.int64)
    s, k = dev(succ), dev(rank)
    ctx.unpack(dev(ps_h), dev(pr_h), dev(active), s, k)
    assert host(s).tolist() == sum(kat["expected_succ_after_unpack_per_rank"], [])
    assert host(k).tolist() == sum(kat["expected_rank_after_unpack_per_rank"], [])


@pytest.mark.parametrize("algo", ALGOS)
def test_errors(ctx, algo):
    import kascade_amd
    say("errors", algo)
    bad = np.array([1, 5, 2], np.uint64)                     # id out of range
    with pytest.raises(kascade_amd.KascadeError) as e:
        ctx.rank_list(dev(bad), None, algorithm=algo)
    assert e.value.status == kascade_amd.KC_EINVAL
    succ = np.array([1, 2, 2], np.uint64)
    w = np.array([1, 1, 5], np.int64)                       # tail weight must be 0
    with pytest.raises(kascade_amd.KascadeError) as e:
        ctx.rank_list(dev(succ), dev(w), algorithm=algo)
    assert e.value.status == kascade_amd.KC_EINVAL
    for n in [3, 50000]:                                    # a cycle without a root
        cyc = np.roll(np.arange(n, dtype=np.uint64), -1)
        with pytest.raises(kascade_amd.KascadeError) as e:
            ctx.rank_list(dev(cyc), None, algorithm=algo)
        assert e.value.status == kascade_amd.KC_ECYCLE, n
    tree = np.array([0, 0, 0, 1], np.uint64)                 # in-degree 2: not a list
    with pytest.raises(kascade_amd.KascadeError) as e:
        ctx.rank_list(dev(tree), None, algorithm=algo, validate=True)
    assert e.value.status == kascade_amd.KC_EINVAL
    # more records than 32-bit kernel grids can cover: rejected before any
    # array is touched
    import ctypes
    small = dev(np.zeros(4, np.uint64))
    p = ctypes.c_void_p(small.data_ptr())
    assert ctx._lib.kc_rank_list(ctx.handle, p, None, ctypes.c_uint64(1 << 32), p, p, None,
                                 None) == kascade_amd.KC_EINVAL
    # the context stays usable
    succ, w = o.gen_list(1000, 1, True)
    root, rank = ctx.rank_list(dev(succ), dev(w), algorithm=algo)
    er, ek = o.list_closed_form(1000, 1, w)
    assert (host(rank) == ek).all()


def test_pd_ranks_trees(ctx):
    """Pointer doubling ranks in-trees too (depth), like the reference's."""
    say("pd tree")
    n = 50000
    parent = o.gen_tree(n, 3)
    root, rank = ctx.rank_list(dev(parent), None, algorithm="pd")
    er, ed = o.tree_closed_form(n, 3)
    assert (host(root) == er).all() and (host(rank) == ed).all()


@pytest.mark.parametrize("n", [1 << 22, 1 << 24])
def test_rank_large_device_generated(ctx, n):
    """Device-generated list against the closed form computed on the host."""
    say("large", n)
    s, w = ctx.gen_list(n, 42, weighted=True)
    root, rank = ctx.rank_list(s, w, algorithm="srs")
    er, ek = o.list_closed_form(n, 42, host(w))
    assert (host(root) == er).all() and (host(rank) == ek).all()


def test_cpp_dropin_header():
    """The C++ drop-in layer (include/kascade_amd.hpp): pack KAT, identity
    sanity check, AbstractAlgorithm adapters against GatherRank, errors."""
    import subprocess
    here = os.path.dirname(__file__)
    exe = os.path.join(here, "cpp", "dropin_test")
    inputs = [os.path.join(here, "cpp", "dropin_test.cpp"),
              os.path.join(here, "..", "include", "kascade_amd.hpp"),
              os.path.join(here, "..", "include", "kascade_amd.h")]
    if not os.path.exists(exe) or os.path.getmtime(exe) < max(os.path.getmtime(f) for f in inputs):
        subprocess.run(["sh", os.path.join(here, "cpp", "build.sh")], check=True)
    say("cpp dropin")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr


def test_tuning_table(ctx):
    """A per-size-class tuning table changes the knobs, never the result."""
    say("tuning")
    before = ctx.tuning_get(1 << 20)
    try:
        ctx.tuning_set([(10, 4, 4, 1, 64), (18, 128, 8, 2, 1024)])
        assert ctx.tuning_get(1 << 12)["spacing_top"] == 4
        assert ctx.tuning_get(1 << 20)["spacing_top"] == 128
        for n in [10000, 300000]:
            succ, w = o.gen_list(n, 17 + n, True)
            root, rank = ctx.rank_list(dev(succ), dev(w))
            er, ek = o.list_closed_form(n, 17 + n, w)
            assert (host(root) == er).all() and (host(rank) == ek).all()
    finally:
        ctx.tuning_set([])
    assert ctx.tuning_get(1 << 20)["spacing_top"] in (8, 16, 32, 64, 128)
    assert before["small_n"] == 8192


@pytest.mark.parametrize("launch", ["device", "host"])
@pytest.mark.parametrize("nlists", [30, 60000])
@pytest.mark.parametrize("spacing,spawn", [(4, "off"), (256, "on"), (1024, "off")])
def test_forest_propagation_shapes(ctx, launch, nlists, spacing, spawn):
    """Propagation runs one element per thread when few waves meet an
    unreached list prefix (30 lists), and strides a resident grid, one
    counter atomic per wave and 32 strides, when many do (60,000 short
    lists).  Both forests ranked with dense and sparse rulers, with and
    without spawn, host- and device-sized; results against the sequential
    walk."""
    say("propagation shapes", launch, nlists, spacing, spawn)
    n = 3_000_017
    succ, w = random_forest_lists(n, nlists, 99 + nlists, weighted=True)
    root, rank = ctx.rank_list(dev(succ), dev(w), algorithm="srs", ruler_spacing=spacing, spawn=spawn,
                               launch=launch)
    er, ek = o.sequential_walk(succ, w)
    assert (host(root) == er).all() and (host(rank) == ek).all()
