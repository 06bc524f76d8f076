"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module.  It checks the product; it is never the
product path.
"""
import ctypes
import os
import subprocess

import numpy as np

_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_LIB_PATH = os.path.join(_ROOT, "oracle", "liboracle.so")
_lib = None

u64p = ctypes.POINTER(ctypes.c_uint64)
i64p = ctypes.POINTER(ctypes.c_int64)


def build():
    subprocess.run(["make", "-s", "-C", os.path.join(_ROOT, "oracle")], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        c = ctypes
        sig = {
            "orc_splitmix64": (c.c_uint64, [c.c_uint64]),
            "orc_perm": (c.c_uint64, [c.c_uint64, c.c_uint64, c.c_uint64]),
            "orc_perm_inv": (c.c_uint64, [c.c_uint64, c.c_uint64, c.c_uint64]),
            "orc_gen_list": (None, [c.c_uint64, c.c_uint64, c.c_int, u64p, i64p]),
            "orc_gen_list_local": (None, [c.c_uint64, c.c_uint64, c.c_int, u64p, i64p]),
            "orc_list_closed_form": (None, [c.c_uint64, c.c_uint64, i64p, u64p, i64p]),
            "orc_gen_tree": (None, [c.c_uint64, c.c_uint64, u64p]),
            "orc_tree_closed_form": (None, [c.c_uint64, c.c_uint64, u64p, i64p]),
            "orc_local_pointer_chasing": (None, [u64p, i64p, c.c_uint64]),
            "orc_set_initial_ranking_state": (None, [u64p, u64p, i64p, c.c_uint64]),
            "orc_pd_rounds": (c.c_int, [u64p, i64p, c.c_uint64, c.c_int]),
            "orc_pointer_doubling": (c.c_int, [u64p, i64p, c.c_uint64, c.c_int]),
            "orc_pack": (c.c_int, [u64p, i64p, u64p, c.c_uint64, u64p, i64p]),
            "orc_unpack": (None, [u64p, i64p, u64p, c.c_uint64, u64p, i64p]),
            "orc_euler_tour_rooting": (c.c_int, [u64p, i64p, c.c_uint64]),
            "orc_compute_euler_tour": (c.c_int64, [u64p, c.c_uint64, u64p, i64p]),
            "orc_euler_tour_rooting_mt": (c.c_int, [u64p, i64p, c.c_uint64, c.c_int]),
            "orc_sparse_ruling_set": (c.c_int, [u64p, i64p, c.c_uint64, c.c_int, c.c_int]),
            "orc_is_list": (c.c_int, [u64p, c.c_uint64]),
            "orc_leaf_info": (c.c_int64, [u64p, c.c_uint64, c.c_void_p]),
            "orc_reverse_list": (c.c_int64, [u64p, c.c_void_p, c.c_uint64, u64p, i64p, c.c_void_p]),
            "orc_compute_stats": (c.c_int64, [u64p, i64p, c.c_uint64, c.c_int, c.c_void_p, c.c_void_p,
                                              c.c_void_p, c.c_void_p, c.c_void_p]),
            "orc_gen_graph": (None, [c.c_int, c.c_uint64, c.c_uint64, c.c_uint64, u64p, u64p]),
            "orc_bfs_forest": (None, [u64p, u64p, c.c_uint64, u64p]),
            "orc_rgg2d_radius": (c.c_uint64, [c.c_uint64, c.c_uint64]),
            "orc_gen_rgg2d": (c.c_uint64, [c.c_uint64, c.c_uint64, c.c_uint64, u64p, u64p, c.c_uint64]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p64(a):
    assert a.flags["C_CONTIGUOUS"]
    if a.dtype == np.uint64:
        return a.ctypes.data_as(u64p)
    assert a.dtype == np.int64
    return a.ctypes.data_as(i64p)


def gen_list(n, seed, weighted=True, local=False):
    succ = np.empty(n, np.uint64)
    w = np.empty(n, np.int64)
    f = lib().orc_gen_list_local if local else lib().orc_gen_list
    f(n, seed, int(weighted), _p64(succ), _p64(w))
    return succ, w


def list_closed_form(n, seed, w):
    root = np.empty(n, np.uint64)
    rank = np.empty(n, np.int64)
    lib().orc_list_closed_form(n, seed, _p64(np.ascontiguousarray(w)), _p64(root), _p64(rank))
    return root, rank


def gen_tree(n, seed):
    parent = np.empty(n, np.uint64)
    lib().orc_gen_tree(n, seed, _p64(parent))
    return parent


def tree_closed_form(n, seed):
    root = np.empty(n, np.uint64)
    depth = np.empty(n, np.int64)
    lib().orc_tree_closed_form(n, seed, _p64(root), _p64(depth))
    return root, depth


def sequential_walk(succ, w):
    """local_pointer_chasing on copies; returns (root, rank)."""
    s = np.array(succ, dtype=np.uint64, copy=True)
    r = np.array(w, dtype=np.int64, copy=True)
    lib().orc_local_pointer_chasing(_p64(s), _p64(r), len(s))
    return s, r


def initial_ranking_state(succ):
    succ = np.ascontiguousarray(succ, dtype=np.uint64)
    root = np.empty_like(succ)
    rank = np.empty(len(succ), np.int64)
    lib().orc_set_initial_ranking_state(_p64(succ), _p64(root), _p64(rank), len(succ))
    return root, rank


def pd_rounds(succ, w, rounds):
    s = np.array(succ, dtype=np.uint64, copy=True)
    r = np.array(w, dtype=np.int64, copy=True)
    k = lib().orc_pd_rounds(_p64(s), _p64(r), len(s), rounds)
    return s, r, k


def pointer_doubling(succ, w, p=4):
    s = np.array(succ, dtype=np.uint64, copy=True)
    r = np.array(w, dtype=np.int64, copy=True)
    k = lib().orc_pointer_doubling(_p64(s), _p64(r), len(s), p)
    assert k >= 0
    return s, r, k


def pack(succ, rank, active):
    active = np.ascontiguousarray(active, dtype=np.uint64)
    ps = np.empty(len(active), np.uint64)
    pr = np.empty(len(active), np.int64)
    rc = lib().orc_pack(_p64(np.ascontiguousarray(succ, dtype=np.uint64)),
                        _p64(np.ascontiguousarray(rank, dtype=np.int64)),
                        _p64(active), len(active), _p64(ps), _p64(pr))
    assert rc == 0
    return ps, pr


def unpack(psucc, prank, active, succ, rank):
    lib().orc_unpack(_p64(np.ascontiguousarray(psucc, dtype=np.uint64)),
                     _p64(np.ascontiguousarray(prank, dtype=np.int64)),
                     _p64(np.ascontiguousarray(active, dtype=np.uint64)), len(active),
                     _p64(succ), _p64(rank))


def euler_tour_rooting(parent):
    s = np.array(parent, dtype=np.uint64, copy=True)
    r = np.empty(len(s), np.int64)
    rc = lib().orc_euler_tour_rooting(_p64(s), _p64(r), len(s))
    assert rc == 0
    return s, r


def compute_euler_tour(parent):
    """compute_euler_tour on one PE: (tour successors, tour weights)."""
    parent = np.ascontiguousarray(parent, dtype=np.uint64)
    n = len(parent)
    ts = np.empty(max(2 * n, 1), np.uint64)
    tw = np.empty(max(2 * n, 1), np.int64)
    k = lib().orc_compute_euler_tour(_p64(parent), n, _p64(ts), _p64(tw))
    assert k >= 0
    return ts[:k], tw[:k]


def euler_tour_rooting_mt(parent, p=4):
    """rank_via_euler_tour on p threads: parallel tour build, ranking by the
    sparse ruling set on p emulated ranks."""
    s = np.array(parent, dtype=np.uint64, copy=True)
    r = np.empty(len(s), np.int64)
    rc = lib().orc_euler_tour_rooting_mt(_p64(s), _p64(r), len(s), p)
    assert rc == 0
    return s, r


def sparse_ruling_set(succ, w, p=4, levels=2):
    """The reference's distributed sparse ruling set on p emulated ranks."""
    s = np.array(succ, dtype=np.uint64, copy=True)
    r = np.array(w, dtype=np.int64, copy=True)
    rc = lib().orc_sparse_ruling_set(_p64(s), _p64(r), len(s), p, levels)
    assert rc == 0
    return s, r


def is_list(succ):
    succ = np.ascontiguousarray(succ, dtype=np.uint64)
    rc = lib().orc_is_list(_p64(succ), len(succ))
    assert rc >= 0
    return bool(rc)


def leaf_info(succ):
    succ = np.ascontiguousarray(succ, dtype=np.uint64)
    flags = np.empty(len(succ), np.uint8)
    k = lib().orc_leaf_info(_p64(succ), len(succ), flags.ctypes.data)
    assert k >= 0
    return flags, int(k)


def reverse_list(succ, dist=None):
    succ = np.ascontiguousarray(succ, dtype=np.uint64)
    n = len(succ)
    pred = np.empty(n, np.uint64)
    dpred = np.empty(n, np.int64)
    roots = np.empty(max(n, 1), np.uint64)
    d = None if dist is None else np.ascontiguousarray(dist, dtype=np.int64)
    k = lib().orc_reverse_list(_p64(succ), None if d is None else d.ctypes.data, n, _p64(pred), _p64(dpred),
                               roots.ctypes.data)
    if k < 0:
        raise ValueError("not a list")
    return pred, dpred, roots[:k]


class _Stats(ctypes.Structure):
    _fields_ = [("num_trees", ctypes.c_uint64), ("max_rank", ctypes.c_int64), ("avg_rank", ctypes.c_double),
                ("num_nontrivial_trees", ctypes.c_uint64), ("nontrivial_size_sum", ctypes.c_uint64),
                ("nontrivial_rank_sum", ctypes.c_int64), ("max_size", ctypes.c_uint64),
                ("avg_size", ctypes.c_double), ("nontrivial_avg_size", ctypes.c_double),
                ("nontrivial_avg_rank", ctypes.c_double)]


STATS_LEVELS = {"none": 1, "basic": 2, "extensive-reduced-output": 3, "extensive": 4}


def compute_stats(root, rank, level="basic"):
    """Same dict shape as kascade_amd.Context.compute_stats."""
    root = np.ascontiguousarray(root, dtype=np.uint64)
    rank = np.ascontiguousarray(rank, dtype=np.int64)
    n = len(root)
    lv = STATS_LEVELS[level]
    st = _Stats()
    tr, ts = np.empty(max(n, 1), np.uint64), np.empty(max(n, 1), np.uint64)
    tm, tsum = np.empty(max(n, 1), np.int64), np.empty(max(n, 1), np.int64)
    k = lib().orc_compute_stats(_p64(root), _p64(rank), n, lv, ctypes.byref(st), tr.ctypes.data, ts.ctypes.data,
                                tm.ctypes.data, tsum.ctypes.data)
    assert k >= 0
    out = {}
    if lv >= 2:
        out["basic_stats"] = {"num_trees": st.num_trees, "max_rank": st.max_rank, "avg_rank": st.avg_rank}
    if lv >= 3:
        ext = {f: getattr(st, f) for f in ("num_trees", "num_nontrivial_trees", "nontrivial_size_sum",
                                           "nontrivial_rank_sum", "max_size", "avg_size", "nontrivial_avg_size",
                                           "nontrivial_avg_rank")}
        if lv == 4:
            ext["per_tree_stats"] = [{"root": int(tr[i]), "size": int(ts[i]), "max_rank": int(tm[i]),
                                      "rank_sum": int(tsum[i])} for i in range(k)]
        out["extensive_stats"] = ext
    return out


GRAPH_KINDS = {"gnm": 0, "rmat": 1}


def gen_graph(kind, n, m, seed):
    src = np.empty(max(m, 1), np.uint64)
    dst = np.empty(max(m, 1), np.uint64)
    lib().orc_gen_graph(GRAPH_KINDS[kind], n, m, seed, _p64(src), _p64(dst))
    return src[:m], dst[:m]


def gen_rgg2d(n, m, seed):
    """The oracle's random geometric graph: (src, dst) in its emission order."""
    cnt = lib().orc_gen_rgg2d(n, m, seed, None, None, 0)
    src = np.empty(max(cnt, 1), np.uint64)
    dst = np.empty(max(cnt, 1), np.uint64)
    lib().orc_gen_rgg2d(n, m, seed, _p64(src), _p64(dst), cnt)
    return src[:cnt], dst[:cnt]


def rgg2d_radius(n, m):
    return lib().orc_rgg2d_radius(n, m)


def csr(src, dst, n):
    """Undirected CSR with each vertex's neighbours in ascending order."""
    a = np.concatenate([src, dst]).astype(np.int64)
    b = np.concatenate([dst, src]).astype(np.uint64)
    order = np.lexsort((b, a))
    xadj = np.zeros(n + 1, np.uint64)
    np.cumsum(np.bincount(a, minlength=n), out=xadj[1:])
    return xadj, np.ascontiguousarray(b[order])


def bfs_forest(xadj, adj):
    n = len(xadj) - 1
    parent = np.empty(n, np.uint64)
    adj = np.ascontiguousarray(adj, dtype=np.uint64) if len(adj) else np.zeros(1, np.uint64)
    lib().orc_bfs_forest(_p64(np.ascontiguousarray(xadj, dtype=np.uint64)), _p64(adj), n, _p64(parent))
    return parent
