"""GPU parity of the graph inputs (the reference's "bfs" / "eulertour" input
processing): seeded edge generators, CSR, and the BFS forest, whose roots and
depths are unique and must equal the oracle's restatement of
generate_bfs_tree."""
import numpy as np
import pytest

import oracle_lib as o
from test_gpu_rank import dev, host, say

pytestmark = pytest.mark.gpu

CASES = [("gnm", 1, 0), ("gnm", 2, 1), ("gnm", 1000, 300), ("gnm", 1 << 16, 1 << 18), ("rmat", 1 << 14, 1 << 17),
         ("rmat", 1 << 18, 1 << 21), ("gnm", 1 << 20, 1 << 19)]


def forest_root_depth(parent):
    n = len(parent)
    w = (parent != np.arange(n, dtype=np.uint64)).astype(np.int64)
    return o.sequential_walk(parent, w)


@pytest.mark.parametrize("kind,n,m", CASES)
def test_graph_bfs_forest(ctx, kind, n, m):
    say("graph", kind, n, m)
    seed = 17 + n + m
    src, dst = ctx.gen_graph(kind, n, m, seed)
    es, ed = o.gen_graph(kind, n, m, seed)
    assert (host(src) == es).all() and (host(dst) == ed).all()
    xadj, adj = ctx.build_csr(src, dst, n)
    ex, ea = o.csr(es, ed, n)
    hx, ha = host(xadj), host(adj)
    assert (hx == ex).all()
    for v in np.random.default_rng(1).choice(n, size=min(n, 300), replace=False):   # same neighbour multisets
        assert sorted(ha[hx[v]:hx[v + 1]].tolist()) == ea[ex[v]:ex[v + 1]].tolist()
    parent = host(ctx.bfs_forest(xadj, adj))
    er, edp = forest_root_depth(o.bfs_forest(ex, ea))
    r, d = forest_root_depth(parent)
    assert (r == er).all() and (d == edp).all()
    # every parent is a neighbour one level up
    nonroot = np.nonzero(parent != np.arange(n, dtype=np.uint64))[0]
    for v in nonroot[:2000]:
        p = int(parent[v])
        assert p in ha[hx[v]:hx[v + 1]].tolist() and d[p] == d[v] - 1


def test_graph_forest_to_tree_rooting_and_tour(ctx):
    """The reference's input pipelines end to end on the device: BFS forest ->
    Euler-tour rooting ("bfs" processing), and BFS forest -> Euler tour ->
    list ranking ("eulertour" processing)."""
    say("graph pipelines")
    n, m = 1 << 20, 3 << 19
    src, dst = ctx.gen_graph("gnm", n, m, 5)
    xadj, adj = ctx.build_csr(src, dst, n)
    parent = ctx.bfs_forest(xadj, adj)
    root, depth = ctx.root_tree(parent)
    es, ed = o.gen_graph("gnm", n, m, 5)
    ex, ea = o.csr(es, ed, n)
    er, edp = forest_root_depth(o.bfs_forest(ex, ea))
    assert (host(root) == er).all() and (host(depth) == edp).all()
    ts, tw = ctx.compute_euler_tour(parent)
    es, ew = o.compute_euler_tour(host(parent))
    assert (host(ts) == es).all() and (host(tw) == ew).all()
    troot, trank = ctx.rank_list(ts, tw)
    cr, ck = o.sequential_walk(es, ew)
    assert (host(troot) == cr).all() and (host(trank) == ck).all()


def test_bfs_forest_rejects_asymmetric_adjacency(ctx):
    import kascade_amd
    say("graph errors")
    xadj = np.array([0, 0, 1], np.uint64)          # 1 -> 0 only: vertex 1 is never reached from 0
    adj = np.array([0], np.uint64)
    with pytest.raises(kascade_amd.KascadeError) as e:
        ctx.bfs_forest(dev(xadj), dev(adj))
    assert e.value.status == kascade_amd.KC_EINVAL


@pytest.mark.parametrize("n,m", [(0, 0), (1, 3), (2, 1), (500, 1500), (1 << 16, 1 << 18), (1 << 19, 1 << 22)])
def test_rgg2d_matches_oracle(ctx, n, m):
    """The device random geometric graph equals the oracle's edge for edge and
    in order (the oracle is pinned by an all-pairs check); its BFS forest
    gives the oracle's roots and depths (the submission suites' rgg2d +
    eulertour pipeline at its per-rank size, 2^19 vertices and 2^22 edges)."""
    say("rgg2d", n, m)
    seed = 11 + n
    src, dst = ctx.gen_rgg2d(n, m, seed)
    es, ed = o.gen_rgg2d(n, m, seed)
    assert (host(src) == es).all() and (host(dst) == ed).all()
    if n < 2:
        return
    xadj, adj = ctx.build_csr(src, dst, n)
    ex, ea = o.csr(es, ed, n)
    parent = ctx.bfs_forest(xadj, adj)
    er, edp = forest_root_depth(o.bfs_forest(ex, ea))
    r, d = forest_root_depth(host(parent))
    assert (r == er).all() and (d == edp).all()
    if n >= 1 << 16:
        root, depth = ctx.root_tree(parent)
        assert (host(root) == er).all() and (host(depth) == edp).all()


def test_rgg2d_host_arrays_and_capacity(ctx):
    import ctypes
    import kascade_amd
    say("rgg2d host")
    n, m = 3000, 9000
    es, ed = o.gen_rgg2d(n, m, 4)
    cnt = ctypes.c_uint64()
    lib = kascade_amd.load()
    assert lib.kc_gen_rgg2d(ctx.handle, n, m, 4, None, None, 0, ctypes.byref(cnt), None) == kascade_amd.KC_OK
    assert cnt.value == len(es)
    src = np.empty(len(es), np.uint64)
    dst = np.empty(len(es), np.uint64)
    assert lib.kc_gen_rgg2d(ctx.handle, n, m, 4, src.ctypes.data, dst.ctypes.data, len(es), ctypes.byref(cnt),
                            None) == kascade_amd.KC_OK
    assert (src == es).all() and (dst == ed).all()
    small = np.empty(10, np.uint64)
    assert lib.kc_gen_rgg2d(ctx.handle, n, m, 4, small.ctypes.data, small.ctypes.data, 10, ctypes.byref(cnt),
                            None) == kascade_amd.KC_EINVAL
