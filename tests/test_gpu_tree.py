"""GPU parity of Euler-tour tree rooting through the C ABI: (root, depth) per
vertex bit-exact against the generator's closed form, BFS and the oracle's
restatement of rank_via_euler_tour."""
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


def relabel_forest(seed):
    rng = np.random.default_rng(seed)
    n = 5000
    parent = np.arange(n, dtype=np.uint64)
    for v in range(1, 1500):
        parent[v] = rng.integers(0, v)            # random recursive tree at 0
    for v in range(1501, 2500):
        parent[v] = 1500                          # star at 1500
    for v in range(2501, 4000):
        parent[v] = v - 1                         # path at 2500 (depth 1499)
    for v in range(4001, 4500):
        parent[v] = 4000 + (v - 4001) // 2        # binary-ish tree at 4000
    # 4500..4999 isolated roots
    sigma = rng.permutation(n)
    out = np.empty(n, np.uint64)
    out[sigma] = sigma[parent]
    return out


@pytest.mark.parametrize("build", ["linked", "sorted"])
@pytest.mark.parametrize("algo", ["srs", "pd"])
@pytest.mark.parametrize("n", [1, 2, 3, 100, 2049, 4097, 100003, 1 << 20])
def test_root_random_tree(ctx, algo, n, build):
    say("tree", algo, n, build)
    parent = o.gen_tree(n, 31 + n)
    root, depth = ctx.root_tree(dev(parent), algorithm=algo, tour_build=build)
    er, ed = o.tree_closed_form(n, 31 + n)
    assert (host(root) == er).all() and (host(depth) == ed).all()


@pytest.mark.parametrize("build", ["linked", "sorted"])
@pytest.mark.parametrize("algo", ["srs", "pd"])
def test_root_forest_shapes(ctx, algo, build):
    say("forest shapes", algo, build)
    parent = relabel_forest(3)
    root, depth = ctx.root_tree(dev(parent), algorithm=algo, tour_build=build)
    er, ed = o.euler_tour_rooting(parent)
    assert (host(root) == er).all() and (host(depth) == ed).all()


def test_in_place_and_host(ctx):
    say("tree in place")
    n = 54321
    parent = o.gen_tree(n, 5)
    er, ed = o.tree_closed_form(n, 5)
    s = dev(parent)
    r = dev(np.ones(n, np.int64))
    ctx.rank_via_euler_tour(s, r)
    assert (host(s) == er).all() and (host(r) == ed).all()
    hs, hr = parent.copy(), np.zeros(n, np.int64)
    ctx.rank_via_euler_tour(hs, hr)
    assert (hs == er).all() and (hr == ed).all()


@pytest.mark.parametrize("seed", [9, 10])
def test_tour_structure(ctx, seed):
    """compute_euler_tour returns the reference's tour element for element
    (the oracle's restatement of compute_euler_tour on one PE), and ranking
    it gives every vertex's depth at its upward edge."""
    say("tour structure", seed)
    parent = relabel_forest(seed)
    n = len(parent)
    ts, tw = ctx.compute_euler_tour(dev(parent))
    ts, tw = host(ts), host(tw)
    es, ew = o.compute_euler_tour(parent)
    roots = np.count_nonzero(parent == np.arange(n))
    assert len(ts) == 2 * (n - roots) + roots
    assert (ts == es).all() and (tw == ew).all()
    # ranked as a list on the GPU, the tour gives the sequential walk's ranks
    r, k = ctx.rank_list(dev(ts), dev(tw))
    cr, ck = o.sequential_walk(es, ew)
    assert (host(r) == cr).all() and (host(k) == ck).all()


@pytest.mark.parametrize("n", [1, 2, 3, 100, 70000])
def test_tour_random_trees(ctx, n):
    say("tour random tree", n)
    parent = o.gen_tree(n, 31 + n)
    ts, tw = ctx.compute_euler_tour(dev(parent))
    es, ew = o.compute_euler_tour(parent)
    assert (host(ts) == es).all() and (host(tw) == ew).all()


def test_tree_errors(ctx):
    import kascade_amd
    say("tree errors")
    with pytest.raises(kascade_amd.KascadeError) as e:
        ctx.root_tree(dev(np.array([0, 7, 0], np.uint64)))
    assert e.value.status == kascade_amd.KC_EINVAL
    for build in ["linked", "sorted"]:
        for n in [2, 3, 9000]:                               # parent cycle, no root
            cyc = np.roll(np.arange(n, dtype=np.uint64), 1)
            with pytest.raises(kascade_amd.KascadeError) as e:
                ctx.root_tree(dev(cyc), tour_build=build)
            assert e.value.status == kascade_amd.KC_ECYCLE, (n, build)


@pytest.mark.parametrize("n", [1 << 22])
def test_root_large_device_generated(ctx, n):
    say("tree large", n)
    parent = ctx.gen_tree(n, 1234)
    root, depth = ctx.root_tree(parent)
    er, ed = o.tree_closed_form(n, 1234)
    assert (host(root) == er).all() and (host(depth) == ed).all()


def random_forest(n, trees, seed):
    """A forest of `trees` random recursive trees of random sizes (plus the
    isolated roots that size 1 gives), relabelled by a seeded permutation."""
    rng = np.random.default_rng(seed)
    cuts = np.sort(rng.choice(np.arange(1, n), size=trees - 1, replace=False))
    starts = np.concatenate([[0], cuts])
    parent = np.arange(n, dtype=np.uint64)
    for a, b in zip(starts, np.concatenate([cuts, [n]])):
        for v in range(a + 1, b):
            parent[v] = rng.integers(a, v)
    sigma = rng.permutation(n)
    out = np.empty(n, np.uint64)
    out[sigma] = sigma[parent]
    return out


@pytest.mark.parametrize("build", ["linked", "sorted"])
@pytest.mark.parametrize("spacing", [0, 2, 4, 64])
@pytest.mark.parametrize("trees", [1, 37, 3000])
def test_root_random_forest_spacings(ctx, build, spacing, trees):
    """Every tour starts at a chased head: with small ruler spacings many heads
    are also computed rulers, and with many trees most tours are short."""
    say("random forest", build, spacing, trees)
    parent = random_forest(60000, trees, 100 + trees)
    root, depth = ctx.root_tree(dev(parent), tour_build=build, ruler_spacing=spacing)
    er, ed = o.euler_tour_rooting(parent)
    assert (host(root) == er).all() and (host(depth) == ed).all()


@pytest.mark.parametrize("build", ["linked", "sorted"])
@pytest.mark.parametrize("cycle", [2, 3, 50, 5000])
def test_tree_cycle_beside_trees(ctx, build, cycle):
    """A rootless component (a cycle with trees hanging off it) next to valid
    trees is reported as a cycle, whether or not a ruler falls on it."""
    import kascade_amd
    say("cycle beside trees", build, cycle)
    n = 40000
    parent = random_forest(n, 20, 7).astype(np.uint64)
    rng = np.random.default_rng(cycle)
    extra = cycle + 300
    ids = np.arange(n, n + extra, dtype=np.uint64)
    p2 = np.empty(extra, np.uint64)
    p2[:cycle] = np.roll(ids[:cycle], 1)                       # the cycle
    for k in range(cycle, extra):
        p2[k] = ids[rng.integers(0, k)]                        # hanging trees
    full = np.concatenate([parent, p2])
    sigma = rng.permutation(len(full))
    out = np.empty(len(full), np.uint64)
    out[sigma] = sigma[full.astype(np.int64)]
    for spacing in (0, 2, 1024):
        with pytest.raises(kascade_amd.KascadeError) as e:
            ctx.root_tree(dev(out), tour_build=build, ruler_spacing=spacing)
        assert e.value.status == kascade_amd.KC_ECYCLE, (build, cycle, spacing)


def test_tree_tuning_table(ctx):
    """Tree rooting has its own tuning table (its classes count vertices): it
    changes the knobs of tree rooting only, never a result."""
    say("tree tuning")
    list_before = ctx.tuning_get(1 << 14)
    try:
        ctx.tuning_set([(10, 4, 4, 1, 64), (16, 256, 8, 2, 1024)], kind="tree")
        assert ctx.tuning_get(1 << 12, kind="tree")["spacing_top"] == 4
        assert ctx.tuning_get(1 << 17, kind="tree")["spacing_top"] == 256
        assert ctx.tuning_get(1 << 14) == list_before
        for n in [6000, 100003]:
            parent = o.gen_tree(n, 31 + n)
            root, depth = ctx.root_tree(dev(parent))
            er, ed = o.tree_closed_form(n, 31 + n)
            assert (host(root) == er).all() and (host(depth) == ed).all()
    finally:
        ctx.tuning_set([], kind="tree")
    assert ctx.tuning_get(1 << 26, kind="tree")["table_rows"] == 0


def test_tuning_file_kinds(tmp_path, monkeypatch):
    """A table file holds list rows, "tree" rows and "seg" rows; malformed
    rows are skipped."""
    import kascade_amd
    say("tuning file kinds")
    f = tmp_path / "tuning.txt"
    f.write_text("# comment\n10 8 8 2 64\n20 64 8 2 64   # 1.0\n"
                 "tree 10 16 8 1 64\ntree 20 32 16 4 256\nnot a row\ntree 12 3 8 1 64\n"
                 "seg 24 128 8 3 256\nsegs 25 64 8 2 64\nseg 26 3 8 2 64\n")
    monkeypatch.setenv("KASCADE_AMD_TUNING", str(f))
    c = kascade_amd.Context(0)
    try:
        lst = c.tuning_get(1 << 21)
        assert (lst["spacing_top"], lst["table_rows"]) == (64, 2)
        t = c.tuning_get(1 << 21, kind="tree")
        assert (t["spacing_top"], t["spacing_deep"], t["chase_blocks_per_cu"], t["chase_chunk"],
                t["table_rows"]) == (32, 16, 4, 256, 2)
        assert c.tuning_get(1 << 11, kind="tree")["spacing_top"] == 16
        s = c.tuning_get(1 << 25, kind="segmented")
        assert (s["spacing_top"], s["spacing_deep"], s["chase_blocks_per_cu"], s["chase_chunk"],
                s["table_rows"]) == (128, 8, 3, 256, 1)
        # below the first segmented class: the list table
        assert c.tuning_get(1 << 21, kind="segmented")["spacing_top"] == 64
    finally:
        c.close()


def hub_tree(n, seed, hub_share):
    """Vertex 0 is the root; a share of the vertices hang directly off it,
    the rest off a random earlier vertex; relabelled by a seeded permutation."""
    rng = np.random.default_rng(seed)
    v = np.arange(n)
    parent = np.floor(rng.random(n) * v).astype(np.int64)
    parent[rng.random(n) < hub_share] = 0
    parent[0] = 0
    sigma = rng.permutation(n)
    out = np.empty(n, np.uint64)
    out[sigma] = sigma[parent]
    return out


@pytest.mark.parametrize("build", ["linked", "sorted"])
@pytest.mark.parametrize("share", [1.0, 0.5, 0.05])
def test_root_high_degree(ctx, build, share):
    """High-degree parents (a star, a hub holding half the vertices, a
    light hub): the linked build inserts the children of one wave that share
    a parent with one atomic exchange per group."""
    n = (1 << 20) + 3
    say("high degree", build, share)
    parent = hub_tree(n, 41, share)
    er, ed = o.euler_tour_rooting(parent)
    root, depth = ctx.root_tree(dev(parent), tour_build=build)
    assert (host(root) == er).all() and (host(depth) == ed).all()


def test_server_star_tree(ctx):
    """A star tree inside a coalesced serving batch (the batch's linked tour
    build groups the star's children too)."""
    import torch
    import kascade_amd
    say("server star")
    jobs = [hub_tree(300000, 5, 1.0), hub_tree(200000, 6, 0.5), o.gen_tree(100000, 7)]
    sv = kascade_amd.Server(ctx, num_slots=1)
    arrays = []
    for p in jobs:
        s = dev(p)
        ro = torch.empty(len(p), dtype=torch.uint64, device="cuda")
        ra = torch.empty(len(p), dtype=torch.int64, device="cuda")
        arrays.append((s, ro, ra))
    torch.cuda.synchronize()
    tickets = [sv.submit("tree", s, None, ro, ra)[0] for s, ro, ra in arrays]
    for t in tickets:
        sv.wait(t)
    assert sv.stats()["batches_launched"] < len(jobs)
    for p, (s, ro, ra) in zip(jobs, arrays):
        er, ed = o.euler_tour_rooting(p)
        assert (host(ro) == er).all() and (host(ra) == ed).all()
    sv.close()
