"""Pins the CPU oracle (oracle/oracle.c) against checks that do not restate it:
closed forms from the generators' construction order, brute-force walks,
BFS depths, the reference's pack known-answer test, and hand-made cases."""
import json
import os

import numpy as np
import pytest

import oracle_lib as o

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def walk_brute(succ, w):
    """Definition of rank: follow succ to the fixed point, summing weights."""
    n = len(succ)
    root = np.empty(n, np.uint64)
    rank = np.empty(n, np.int64)
    for v in range(n):
        cur, acc, steps = v, 0, 0
        while int(succ[cur]) != cur:
            acc += int(w[cur])
            cur = int(succ[cur])
            steps += 1
            assert steps <= n, "cycle"
        root[v] = cur
        rank[v] = acc
    return root, rank


def random_forest_lists(n, nlists, seed, weighted=True):
    rng = np.random.default_rng(seed)
    perm = rng.permutation(n)
    cuts = np.sort(rng.choice(np.arange(1, n), size=nlists - 1, replace=False)) if nlists > 1 else []
    succ = np.arange(n, dtype=np.uint64)
    w = np.zeros(n, np.int64)
    for seg in np.split(perm, cuts):
        for a, b in zip(seg[:-1], seg[1:]):
            succ[a] = b
            w[a] = rng.integers(1, 101) if weighted else 1
    return succ, w


@pytest.mark.parametrize("n", [1, 2, 3, 5, 64, 1000, 4097])
@pytest.mark.parametrize("weighted", [False, True])
def test_generator_is_one_list_and_closed_form(n, weighted):
    succ, w = o.gen_list(n, 11 + n, weighted)
    assert succ.max() < n
    heads = np.setdiff1d(np.arange(n), succ[succ != np.arange(n)])
    assert len(heads) == 1                       # exactly one head
    tails = np.nonzero(succ == np.arange(n))[0]
    assert len(tails) == 1 and w[tails[0]] == 0  # one tail, weight 0
    if weighted and n > 1:
        nt = w[succ != np.arange(n)]
        assert nt.min() >= 1 and nt.max() <= 100
    root, rank = o.list_closed_form(n, 11 + n, w)
    if n <= 1000:
        br, bk = walk_brute(succ, w)
        assert (br == root).all() and (bk == rank).all()


def test_local_generator():
    n = 1003
    succ, w = o.gen_list(n, 5, True, local=True)
    br, bk = walk_brute(succ, w)
    assert len(set(br.tolist())) == 1                       # a single list
    assert np.mean(succ[:-1] == np.arange(1, n)) > 0.85     # mostly i -> i+1


def test_perm_is_bijection_and_inverse():
    for n in [1, 2, 7, 100, 1025]:
        p = [o.lib().orc_perm(i, n, 99) for i in range(n)]
        assert sorted(p) == list(range(n))
        assert all(o.lib().orc_perm_inv(p[i], n, 99) == i for i in range(n))


@pytest.mark.parametrize("n,nlists", [(1, 1), (10, 10), (50, 3), (2000, 17), (3000, 1)])
def test_sequential_walk_matches_definition(n, nlists):
    succ, w = random_forest_lists(n, nlists, n + nlists)
    r, k = o.sequential_walk(succ, w)
    br, bk = walk_brute(succ, w)
    assert (r == br).all() and (k == bk).all()


def test_sequential_walk_hand_case():
    # 3 -> 0 -> 2 -> 1 (tail), weights 5, 7, 11; and an isolated 4
    succ = np.array([2, 1, 1, 0, 4], np.uint64)
    w = np.array([7, 0, 11, 5, 0], np.int64)
    r, k = o.sequential_walk(succ, w)
    assert r.tolist() == [1, 1, 1, 1, 4]
    assert k.tolist() == [18, 0, 11, 23, 0]


def test_initial_state():
    succ = np.array([1, 1, 0, 3], np.uint64)
    root, rank = o.initial_ranking_state(succ)
    assert root.tolist() == [1, 1, 0, 3] and rank.tolist() == [1, 0, 1, 0]


@pytest.mark.parametrize("p", [1, 4, 7])
@pytest.mark.parametrize("n,nlists", [(1, 1), (5, 2), (1000, 1), (4096, 40), (20000, 3)])
def test_pointer_doubling_ranks(p, n, nlists):
    succ, w = random_forest_lists(n, nlists, 3 * n + p)
    r, k, rounds = o.pointer_doubling(succ, w, p)
    er, ek = o.sequential_walk(succ, w)
    assert (r == er).all() and (k == ek).all()


def test_pointer_doubling_local_lists_contract():
    # locality-biased lists exercise local contraction / uncontraction
    succ, w = o.gen_list(30011, 3, True, local=True)
    r, k, _ = o.pointer_doubling(succ, w, 4)
    root, rank = o.sequential_walk(succ, w)
    assert (r == root).all() and (k == rank).all()


@pytest.mark.parametrize("n", [1, 9, 200, 1023])
def test_pd_rounds_closed_form(n):
    """After k Jacobi rounds: succ_k(v) = 2^k-th successor clamped at the root,
    rank_k(v) = sum of the 2^k weights passed (brute-force stepping), and the
    root flag is set once the distance to the root is below 2^k."""
    succ, w = random_forest_lists(n, max(1, n // 100), n)
    s64 = succ.astype(np.int64)
    is_root = s64 == np.arange(n)
    # distance to the root by stepping
    dist = np.zeros(n, np.int64)
    cur = np.arange(n)
    for _ in range(n):
        moving = ~is_root[cur]
        dist += moving
        cur = s64[cur]
    cur = np.arange(n)
    acc = np.zeros(n, np.int64)
    steps = 0
    for k in range(0, 12):
        while steps < (1 << k):
            acc += w[cur]          # w is 0 on roots, succ[root] = root
            cur = s64[cur]
            steps += 1
        s, r, _ = o.pd_rounds(succ, w, k)
        flagged = (s >> np.uint64(63)).astype(bool)
        assert ((s & np.uint64((1 << 63) - 1)).astype(np.int64) == cur).all(), k
        assert (r == acc).all(), k
        assert (flagged == (dist < (1 << k))).all(), k


def test_pack_kat():
    kat = json.load(open(os.path.join(GOLDEN, "pack_kat.json")))
    off = kat["offsets"]
    succ = np.concatenate([np.array(x, np.uint64) for x in kat["succ_per_rank"]])
    rank = np.concatenate([np.array(x, np.int64) for x in kat["rank_per_rank"]])
    active = np.concatenate([np.array(a, np.uint64) + off[r] for r, a in enumerate(kat["active_local_per_rank"])])
    ps, pr = o.pack(succ, rank, active)
    assert ps.tolist() == sum(kat["expected_packed_succ_per_rank"], [])
    assert pr.tolist() == sum(kat["expected_packed_rank_per_rank"], [])
    # simulated ranking: root 6, rank 6 - i - 2r
    ps[:] = 6
    pos = 0
    for r, a in enumerate(kat["active_local_per_rank"]):
        for i in range(len(a)):
            pr[pos] = 6 - i - 2 * r
            pos += 1
    o.unpack(ps, pr, active, succ, rank)
    assert succ.tolist() == sum(kat["expected_succ_after_unpack_per_rank"], [])
    assert rank.tolist() == sum(kat["expected_rank_after_unpack_per_rank"], [])


def test_pack_sanity_identity():
    succ = np.arange(1, 13, dtype=np.uint64)
    succ[-1] = 11
    rank = np.ones(12, np.int64)
    rank[-1] = 0
    ps, pr = o.pack(succ, rank, np.arange(12, dtype=np.uint64))
    assert (ps == succ).all() and (pr == rank).all()


def bfs_depth(parent):
    n = len(parent)
    children = [[] for _ in range(n)]
    roots = []
    for v in range(n):
        p = int(parent[v])
        (roots.append(v) if p == v else children[p].append(v))
    depth = np.full(n, -1, np.int64)
    root = np.zeros(n, np.uint64)
    for r in roots:
        depth[r] = 0
        root[r] = r
        stack = [r]
        while stack:
            x = stack.pop()
            for c in children[x]:
                depth[c] = depth[x] + 1
                root[c] = r
                stack.append(c)
    return root, depth


@pytest.mark.parametrize("n", [1, 2, 3, 10, 257, 5000])
def test_tree_generator_and_euler_tour(n):
    parent = o.gen_tree(n, 4 + n)
    tr, td = o.tree_closed_form(n, 4 + n)
    br, bd = bfs_depth(parent)
    assert (tr == br).all() and (td == bd).all()
    er, ed = o.euler_tour_rooting(parent)
    assert (er == br).all() and (ed == bd).all()


def test_euler_tour_forest_and_paths():
    rng = np.random.default_rng(0)
    # forest: several roots, a star, a path, isolated roots
    n = 600
    parent = np.arange(n, dtype=np.uint64)
    for v in range(1, 200):
        parent[v] = rng.integers(0, v)          # tree rooted at 0
    for v in range(201, 300):
        parent[v] = 200                         # star rooted at 200
    for v in range(301, 400):
        parent[v] = v - 1                       # path rooted at 300
    # 400..599 isolated roots
    sigma = rng.permutation(n)
    relab = np.empty(n, np.uint64)
    relab[sigma] = sigma[parent]
    br, bd = bfs_depth(relab)
    er, ed = o.euler_tour_rooting(relab)
    assert (er == br).all() and (ed == bd).all()


@pytest.mark.parametrize("p", [1, 3, 8])
@pytest.mark.parametrize("n", [1, 2, 3, 10, 257, 5000, 100003])
def test_euler_tour_rooting_threads(p, n):
    """rank_via_euler_tour on p threads (parallel build, ranking by the
    sparse ruling set on p emulated ranks) gives the BFS roots and depths."""
    parent = o.gen_tree(n, 9 + n)
    tr, td = o.tree_closed_form(n, 9 + n)
    er, ed = o.euler_tour_rooting_mt(parent, p)
    assert (er == tr).all() and (ed == td).all()
    if n <= 5000:
        br, bd = bfs_depth(parent)
        assert (er == br).all() and (ed == bd).all()


@pytest.mark.parametrize("p", [1, 4])
def test_euler_tour_rooting_threads_shapes(p):
    """Forest, star, deep path and isolated roots, relabelled."""
    rng = np.random.default_rng(p)
    n = 6000
    parent = np.arange(n, dtype=np.uint64)
    for v in range(1, 2000):
        parent[v] = rng.integers(0, v)
    for v in range(2001, 3000):
        parent[v] = 2000
    for v in range(3001, 5000):
        parent[v] = v - 1
    sigma = rng.permutation(n)
    relab = np.empty(n, np.uint64)
    relab[sigma] = sigma[parent]
    br, bd = bfs_depth(relab)
    er, ed = o.euler_tour_rooting_mt(relab, p)
    assert (er == br).all() and (ed == bd).all()


@pytest.mark.parametrize("p", [1, 2, 4, 7])
@pytest.mark.parametrize("n,nlists", [(1, 1), (10, 1), (100, 3), (5000, 40), (20000, 1), (30000, 3000)])
def test_sparse_ruling_set_ranks(p, n, nlists):
    """The distributed sparse-ruling-set restatement on p emulated ranks gives
    the sequential walk's (root, rank) on list forests."""
    succ, w = random_forest_lists(n, nlists, 11 * n + p)
    r, k = o.sparse_ruling_set(succ, w, p)
    er, ek = o.sequential_walk(succ, w)
    assert (r == er).all() and (k == ek).all()


def test_sparse_ruling_set_generated_lists():
    for local in (False, True):
        succ, w = o.gen_list(100003, 9, True, local=local)
        r, k = o.sparse_ruling_set(succ, w, 4)
        er, ek = o.sequential_walk(succ, w)
        assert (r == er).all() and (k == ek).all()


# ---- successor utilities and statistics ------------------------------------
def _indegree(succ):
    n = len(succ)
    nonroot = succ != np.arange(n, dtype=np.uint64)
    return np.bincount(succ[nonroot].astype(np.int64), minlength=n)


def _random_function(n, seed):
    """Successor arrays with in-degree > 1 somewhere (random functions)."""
    rng = np.random.default_rng(seed)
    return rng.integers(0, n, size=n).astype(np.uint64)


@pytest.mark.parametrize("seed", range(6))
def test_is_list_and_leaf_info_brute_force(seed):
    for n in (1, 2, 7, 300, 5000):
        for succ in (random_forest_lists(n, max(1, n // 50), seed + n)[0], _random_function(n, seed + n),
                     np.arange(n, dtype=np.uint64)):
            indeg = _indegree(succ)
            assert o.is_list(succ) == bool(indeg.max(initial=0) <= 1)
            flags, k = o.leaf_info(succ)
            assert (flags == (indeg == 0)).all() and k == int((indeg == 0).sum())


def test_is_list_hand_cases():
    assert o.is_list(np.array([1, 2, 2], np.uint64))          # 0 -> 1 -> 2 (tail)
    assert not o.is_list(np.array([2, 2, 2], np.uint64))      # 0 and 1 both point at 2
    assert o.is_list(np.array([0, 1, 2], np.uint64))          # singletons
    flags, k = o.leaf_info(np.array([1, 2, 2], np.uint64))
    assert list(flags) == [1, 0, 0] and k == 1                # only the head
    flags, k = o.leaf_info(np.array([0], np.uint64))
    assert list(flags) == [1] and k == 1                      # a lone root counts as a leaf


@pytest.mark.parametrize("n,nlists", [(1, 1), (9, 2), (1000, 1), (5000, 37)])
def test_reverse_list_invariants(n, nlists):
    succ, w = random_forest_lists(n, nlists, 3 * n + nlists)
    pred, dpred, roots = o.reverse_list(succ, w)
    ids = np.arange(n, dtype=np.uint64)
    # brute force: pred[succ[u]] = u for non-roots, else (self, 0)
    ep, ed = ids.copy(), np.zeros(n, np.int64)
    nonroot = succ != ids
    ep[succ[nonroot]] = ids[nonroot]
    ed[succ[nonroot]] = w[nonroot]
    assert (pred == ep).all() and (dpred == ed).all()
    assert (roots == np.nonzero(succ == ids)[0]).all()        # ascending original roots
    # reversing twice gives the list back
    succ2, w2, roots2 = o.reverse_list(pred, dpred)
    assert (succ2 == succ).all() and (w2 == w).all()
    assert set(roots2.tolist()) == set(np.nonzero(_indegree(succ) == 0)[0].tolist())
    # rank along the list + rank along the reversed list = the list's total
    r, k = o.sequential_walk(succ, w)
    rr, kr = o.sequential_walk(pred, dpred)
    head_of = {int(t): int(h) for h, t in zip(rr, r)}          # reversed root = original head
    total = np.array([k[head_of[int(t)]] for t in r], np.int64)
    assert (k + kr == total).all()


def test_reverse_list_rejects_second_predecessor():
    with pytest.raises(ValueError):
        o.reverse_list(np.array([2, 2, 2], np.uint64))


def _stats_brute(root, rank, level):
    n = len(root)
    out = {"basic_stats": {"num_trees": int((root == np.arange(n)).sum()), "max_rank": int(rank.max()),
                           "avg_rank": float(rank.sum()) / n}}
    if level == "basic":
        return out
    trees = {}
    for r, k in zip(root.tolist(), rank.tolist()):
        t = trees.setdefault(r, [0, 0, -2**63])
        t[0] += 1
        t[1] += k
        t[2] = max(t[2], k)
    nt = {r: t for r, t in trees.items() if t[0] > 1}
    zs = sum(t[0] for t in nt.values())
    zr = sum(t[1] for t in nt.values())
    triv = len(trees) - len(nt)
    ext = {"num_trees": len(trees), "num_nontrivial_trees": len(nt), "nontrivial_size_sum": zs,
           "nontrivial_rank_sum": zr, "max_size": max(t[0] for t in trees.values()),
           "avg_size": (zs + triv) / len(trees), "nontrivial_avg_size": zs / len(nt) if nt else 0.0,
           "nontrivial_avg_rank": zr / zs if zs else 0.0}
    if level == "extensive":
        ext["per_tree_stats"] = [{"root": r, "size": t[0], "max_rank": t[2], "rank_sum": t[1]}
                                 for r, t in sorted(nt.items())]
    out["extensive_stats"] = ext
    return out


@pytest.mark.parametrize("level", ["basic", "extensive-reduced-output", "extensive"])
@pytest.mark.parametrize("n,nlists", [(1, 1), (50, 50), (3000, 1), (3000, 200)])
def test_compute_stats_brute_force(level, n, nlists):
    succ, w = random_forest_lists(n, nlists, n + nlists)
    root, rank = o.sequential_walk(succ, w)
    assert o.compute_stats(root, rank, level) == _stats_brute(root, rank, level)


def test_compute_stats_closed_form():
    """One unit-weight list of n: ranks n-1..0, so avg (n-1)/2, rank sum n(n-1)/2."""
    n = 1 << 16
    succ, w = o.gen_list(n, 5, False)
    root, rank = o.sequential_walk(succ, w)
    st = o.compute_stats(root, rank, "extensive")
    assert st["basic_stats"] == {"num_trees": 1, "max_rank": n - 1, "avg_rank": (n - 1) / 2}
    e = st["extensive_stats"]
    assert e["per_tree_stats"] == [{"root": int(root[0]), "size": n, "max_rank": n - 1,
                                    "rank_sum": n * (n - 1) // 2}]
    assert e["avg_size"] == n and e["nontrivial_avg_rank"] == (n - 1) / 2
    # a tree: depths from the generator's closed form
    parent = o.gen_tree(5000, 3)
    er, ed = o.tree_closed_form(5000, 3)
    e = o.compute_stats(er, ed, "extensive")["extensive_stats"]
    assert e["per_tree_stats"][0]["rank_sum"] == int(ed.sum()) and e["max_size"] == 5000


# ---- graph inputs (BFS forest) ----------------------------------------------
def _bfs_brute(xadj, adj):
    """Component minimum and BFS distance from it, by queue BFS."""
    from collections import deque
    n = len(xadj) - 1
    root = np.full(n, -1, np.int64)
    depth = np.zeros(n, np.int64)
    for s in range(n):
        if root[s] >= 0:
            continue
        root[s], depth[s] = s, 0
        q = deque([s])
        while q:
            u = q.popleft()
            for v in adj[xadj[u]:xadj[u + 1]].tolist():
                if root[v] < 0:
                    root[v], depth[v] = s, depth[u] + 1
                    q.append(v)
    return root.astype(np.uint64), depth


@pytest.mark.parametrize("kind", ["gnm", "rmat"])
@pytest.mark.parametrize("n,m", [(1, 0), (1, 3), (2, 1), (50, 20), (1000, 600), (1000, 5000), (4096, 4096)])
def test_bfs_forest_matches_queue_bfs(kind, n, m):
    src, dst = o.gen_graph(kind, n, m, n + 7 * m)
    assert (src < n).all() and (dst < n).all()
    if n > 1:
        assert (src != dst).all()
    xadj, adj = o.csr(src, dst, n)
    parent = o.bfs_forest(xadj, adj)
    er, ed = _bfs_brute(xadj, adj)
    w = (parent != np.arange(n, dtype=np.uint64)).astype(np.int64)
    r, d = o.sequential_walk(parent, w)          # a forest: root and depth
    assert (r == er).all() and (d == ed).all()
    for v in range(n):                           # parents are neighbours one level up
        p = int(parent[v])
        if p != v:
            assert p in adj[xadj[v]:xadj[v + 1]].tolist() and ed[p] == ed[v] - 1


def test_rmat_is_skewed():
    """R-MAT concentrates edges on low ids: quadrant (0, 0) is chosen with
    probability 0.57 per bit, so vertex 0's degree far exceeds the mean."""
    n, m = 1 << 12, 1 << 16
    src, dst = o.gen_graph("rmat", n, m, 3)
    deg = np.bincount(np.concatenate([src, dst]).astype(np.int64), minlength=n)
    assert deg[0] > 20 * deg.mean()
    g_src, g_dst = o.gen_graph("gnm", n, m, 3)
    gdeg = np.bincount(np.concatenate([g_src, g_dst]).astype(np.int64), minlength=n)
    assert gdeg.max() < 4 * gdeg.mean()


def _rgg2d_points(n, seed):
    """The generator's definition of point i (31-bit fixed point)."""
    m64 = (1 << 64) - 1

    def sm(x):
        x = (x + 0x9E3779B97F4A7C15) & m64
        x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & m64
        x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & m64
        return x ^ (x >> 31)

    xs = [sm(seed ^ (((2 * i + 1) * 0xA0761D6478BD642F) & m64)) >> 33 for i in range(n)]
    ys = [sm(seed ^ (((2 * i + 2) * 0xE7037ED1A0B428DB) & m64)) >> 33 for i in range(n)]
    return np.array(xs, np.int64), np.array(ys, np.int64)


@pytest.mark.parametrize("n,m", [(0, 0), (1, 5), (2, 1), (50, 40), (700, 3000), (1500, 100000)])
def test_rgg2d_grid_matches_brute_force(n, m):
    """The grid neighbour search finds exactly the pairs within the radius
    (all-pairs check on the same points), each once with i < j, grouped by i."""
    seed = 17 + n
    src, dst = o.gen_rgg2d(n, m, seed)
    r = o.rgg2d_radius(n, m)
    exp = set()
    if n >= 2 and r > 0:
        x, y = _rgg2d_points(n, seed)
        dx = x[:, None] - x[None, :]
        dy = y[:, None] - y[None, :]
        close = (dx * dx + dy * dy) <= r * r
        iu, ju = np.nonzero(np.triu(close, 1))
        exp = set(zip(iu.tolist(), ju.tolist()))
    got = list(zip(src.tolist(), dst.tolist()))
    assert len(got) == len(set(got)) == len(exp)
    assert set(got) == exp
    assert all(a < b for a, b in got)
    assert list(src) == sorted(src)


def test_rgg2d_edge_count_near_target():
    n, m = 4000, 20000
    src, _ = o.gen_rgg2d(n, m, 3)
    # pairs near the border lose part of their disc: expect a little under m
    assert 0.85 * m < len(src) <= 1.02 * m


# compute_euler_tour (eulertour.cpp:379-392) on one PE, worked by hand from
# reverse_rooted_tree (edges (i, p), (p, i) per non-root i, in order of i),
# make_graph (counting sort by source, each bucket filled back to front) and
# compute_euler_tour_impl (succ[(u_prev, v)] = (v, u), element ids in CSR
# order, the wrap at a root to its dummy):
#   parent [0, 0, 1]: CSR 0:[1] 1:[2, 0] 2:[1]; elements (0,1) (1,2) (1,0)
#   (2,1), dummy of 0 at 4; tour (0,1)->(1,2)->(2,1)->(1,0)->dummy.
#   parent [1, 1, 1, 0]: CSR 0:[3, 1] 1:[2, 0] 2:[1] 3:[0] (children above
#   the vertex, its parent, children below it); elements (0,3) (0,1) (1,2)
#   (1,0) (2,1) (3,0), dummy of 1 at 6; tour (1,2)->(2,1)->(1,0)->(0,3)->
#   (3,0)->(0,1)->dummy.
#   parent [0, 0, 2]: vertex 2 is a childless root: its slot comes after
#   the dummy of 0.
TOUR_CASES = [
    ([0, 0, 1], [1, 3, 4, 2, 4], [-1, -1, 1, 1, 0]),
    ([1, 1, 1, 0], [5, 6, 4, 0, 3, 1, 6], [-1, 1, -1, -1, 1, 1, 0]),
    ([0, 0, 2], [1, 2, 2, 3], [-1, 1, 0, 0]),
    ([0, 1], [0, 1], [0, 0]),
    ([0], [0], [0]),
]


@pytest.mark.parametrize("parent,succ,w", TOUR_CASES)
def test_compute_euler_tour_hand_cases(parent, succ, w):
    ts, tw = o.compute_euler_tour(np.array(parent, np.uint64))
    assert ts.tolist() == succ and tw.tolist() == w


def tour_order_brute(parent):
    """The CSR neighbour order by the rule make_graph produces: for vertex v,
    children above v in descending order, then its parent, then children
    below v in descending order."""
    n = len(parent)
    kids = [[] for _ in range(n)]
    for c in range(n):
        if parent[c] != c:
            kids[int(parent[c])].append(c)
    nb = []
    for v in range(n):
        above = sorted([c for c in kids[v] if c > v], reverse=True)
        below = sorted([c for c in kids[v] if c < v], reverse=True)
        nb.append(above + ([int(parent[v])] if parent[v] != v else []) + below)
    return nb


@pytest.mark.parametrize("seed", range(6))
def test_compute_euler_tour_invariants(seed):
    """On random forests: the element order is the CSR rule above (edge ids
    by source, then neighbour position); each tree's tour visits every
    element of the tree once and ends at its root's dummy; ranking the tour
    with its weights gives every vertex's depth at its upward edge."""
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 400))
    parent = np.arange(n, dtype=np.uint64)
    for v in range(n):
        if rng.random() < 0.9 and v > 0:
            parent[v] = rng.integers(0, v)
    sigma = rng.permutation(n)
    relab = np.empty(n, np.uint64)
    relab[sigma] = sigma[parent]
    ts, tw = o.compute_euler_tour(relab)
    nb = tour_order_brute(relab)
    roots = [v for v in range(n) if relab[v] == v]
    E = sum(len(x) for x in nb)
    assert len(ts) == E + len(roots) and E == 2 * (n - len(roots))
    # edge ids in CSR order; succ by the cyclic rule with the root wrap
    eid, off = {}, [0] * (n + 1)
    for v in range(n):
        off[v + 1] = off[v] + len(nb[v])
        for i, u in enumerate(nb[v]):
            eid[(v, u)] = off[v] + i
    with_kids = [v for v in roots if nb[v]]
    dummy = {v: E + k for k, v in enumerate(with_kids)}
    for (x, y), e in eid.items():
        i = (nb[y].index(x) + 1) % len(nb[y])
        exp = dummy[y] if (relab[y] == y and i == 0) else eid[(y, nb[y][i])]
        assert ts[e] == exp
        assert tw[e] == (1 if relab[x] == y else -1)
    assert (ts[E:] == np.arange(E, E + len(roots))).all() and (tw[E:] == 0).all()
    # ranking the tour: rank of (v, parent v) is depth(v)
    r, k = o.sequential_walk(ts, tw)
    br, bd = bfs_depth(relab)
    for v in range(n):
        if relab[v] != v:
            e = eid[(v, int(relab[v]))]
            assert k[e] == bd[v] and r[e] == dummy[int(br[v])]
