"""Python binding of libkascade_amd (ctypes over the C ABI in include/kascade_amd.h).

Arrays may be torch tensors (device or host) or numpy arrays (host); host arrays
are staged by the library.  Every call raises KascadeError on a non-OK status.
The HIP library must be built (python -m kascade_amd._build); importing this
module on a machine without it raises ImportError: there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

__all__ = ["Context", "KascadeError", "Server", "lib_path", "load"]

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libkascade_amd.so")
_lib = None

KC_OK, KC_EINVAL, KC_ECYCLE, KC_ENOMEM, KC_EDEVICE, KC_EBUDGET, KC_EAGAIN, KC_ENOTFOUND = range(8)
STATUS_NAMES = {0: "KC_OK", 1: "KC_EINVAL", 2: "KC_ECYCLE", 3: "KC_ENOMEM", 4: "KC_EDEVICE",
                5: "KC_EBUDGET", 6: "KC_EAGAIN", 7: "KC_ENOTFOUND"}
ALGORITHMS = {"auto": 0, "pointer_doubling": 1, "pd": 1, "sparse_ruling_set": 2, "srs": 2}


class KascadeError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class RankCfg(ctypes.Structure):
    _fields_ = [("algorithm", ctypes.c_int32), ("ruler_spacing", ctypes.c_uint32),
                ("ruler_spacing_deep", ctypes.c_uint32), ("max_levels", ctypes.c_uint32),
                ("validate", ctypes.c_uint32), ("chase_blocks_per_cu", ctypes.c_uint32),
                ("chase_chunk", ctypes.c_uint32), ("launch", ctypes.c_uint32), ("spawn", ctypes.c_uint32),
                ("local_contraction", ctypes.c_uint32), ("records", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32)]


class TreeCfg(ctypes.Structure):
    _fields_ = [("rank", RankCfg), ("tour_build", ctypes.c_int32), ("reserved", ctypes.c_uint32 * 3)]


TOUR_BUILDS = {"auto": 0, "sorted": 1, "linked": 2}


class Job(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("reserved", ctypes.c_uint32), ("n", ctypes.c_uint64),
                ("succ", ctypes.c_void_p), ("weight", ctypes.c_void_p),
                ("root_out", ctypes.c_void_p), ("rank_out", ctypes.c_void_p)]


class ServerCfg(ctypes.Structure):
    _fields_ = [("mem_budget_bytes", ctypes.c_uint64), ("max_batch_jobs", ctypes.c_uint32),
                ("num_slots", ctypes.c_uint32), ("max_batch_elements", ctypes.c_uint64),
                ("one_job_per_launch", ctypes.c_uint32), ("tour_build", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32 * 2)]


class TraceLevel(ctypes.Structure):
    _fields_ = [("level", ctypes.c_int32), ("kind", ctypes.c_int32), ("subproblem", ctypes.c_int32),
                ("rounds", ctypes.c_uint32), ("n", ctypes.c_uint64), ("spacing", ctypes.c_uint64),
                ("rulers", ctypes.c_uint64), ("tail_rulers", ctypes.c_uint64), ("unreached", ctypes.c_uint64),
                ("chain_length_min", ctypes.c_uint64), ("chain_length_max", ctypes.c_uint64),
                ("chain_length_avg", ctypes.c_double), ("spawned", ctypes.c_uint64)]


TRACE_KINDS = ("srs", "pd", "small")


class Stats(ctypes.Structure):
    _fields_ = [("num_trees", ctypes.c_uint64), ("max_rank", ctypes.c_int64), ("avg_rank", ctypes.c_double),
                ("num_nontrivial_trees", ctypes.c_uint64), ("nontrivial_size_sum", ctypes.c_uint64),
                ("nontrivial_rank_sum", ctypes.c_int64), ("max_size", ctypes.c_uint64),
                ("avg_size", ctypes.c_double), ("nontrivial_avg_size", ctypes.c_double),
                ("nontrivial_avg_rank", ctypes.c_double)]


# StatsLevel names (bench/detail/serialization.hpp) -> kc_stats_level
STATS_LEVELS = {"none": 1, "basic": 2, "extensive-reduced-output": 3, "extensive": 4}


class ServerStats(ctypes.Structure):
    _fields_ = [("jobs_completed", ctypes.c_uint64), ("elements_completed", ctypes.c_uint64),
                ("batches_launched", ctypes.c_uint64), ("p50_latency_ms", ctypes.c_double),
                ("p99_latency_ms", ctypes.c_double), ("mean_latency_ms", ctypes.c_double),
                ("jobs_per_s", ctypes.c_double), ("elements_per_s", ctypes.c_double),
                ("peak_device_bytes", ctypes.c_uint64), ("batches_rerun", ctypes.c_uint64),
                ("jobs_failed", ctypes.c_uint64), ("scratch_bytes", ctypes.c_uint64)]


def lib_path():
    return _LIB_PATH


# name -> (restype, argtypes); every symbol declared in include/kascade_amd.h
_P = ctypes.c_void_p
_U64 = ctypes.c_uint64
_SIGS = {
    "kc_context_create": (ctypes.c_int, [ctypes.c_int, _U64, ctypes.POINTER(_P)]),
    "kc_context_destroy": (None, [_P]),
    "kc_last_error": (ctypes.c_char_p, [_P]),
    "kc_context_scratch_bytes": (_U64, [_P]),
    "kc_context_trim": (ctypes.c_int, [_P]),
    "kc_release_cached_scratch": (ctypes.c_int, [ctypes.c_int]),
    "kc_context_launch_stats": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_uint64)]),
    "kc_context_record_stats": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_uint64)]),
    "kc_profile_enable": (ctypes.c_int, [_P, ctypes.c_int]),
    "kc_profile_reset": (ctypes.c_int, [_P]),
    "kc_profile_read": (ctypes.c_uint32, [_P, ctypes.c_uint32, ctypes.POINTER(ctypes.c_char_p),
                                          ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_uint64)]),
    "kc_tuning_set": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_uint32), ctypes.c_uint32]),
    "kc_tuning_get": (ctypes.c_uint32, [_P, _U64, ctypes.POINTER(ctypes.c_uint32)]),
    "kc_tuning_set_kind": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.POINTER(ctypes.c_uint32), ctypes.c_uint32]),
    "kc_tuning_get_kind": (ctypes.c_uint32, [_P, ctypes.c_int32, _U64, ctypes.POINTER(ctypes.c_uint32)]),
    "kc_set_initial_ranking_state": (ctypes.c_int, [_P, _P, _U64, _P, _P, _P]),
    "kc_rank_list": (ctypes.c_int, [_P, _P, _P, _U64, _P, _P, ctypes.POINTER(RankCfg), _P]),
    "kc_rank_list_noalloc": (ctypes.c_int, [_P, _P, _P, _U64, _P, _P, ctypes.POINTER(RankCfg), _P]),
    "kc_pointer_doubling": (ctypes.c_int, [_P, ctypes.POINTER(RankCfg), _P, _P, _U64, _P]),
    "kc_sparse_ruling_set": (ctypes.c_int, [_P, ctypes.POINTER(RankCfg), _P, _P, _U64, _P]),
    "kc_pointer_doubling_rounds": (ctypes.c_int, [_P, _P, _P, _U64, ctypes.c_uint32, _P, _P, _P]),
    "kc_pack": (ctypes.c_int, [_P, _P, _P, _U64, _P, _U64, _P, _P, _P]),
    "kc_unpack": (ctypes.c_int, [_P, _P, _P, _P, _U64, _P, _P, _U64, _P]),
    "kc_rank_lists_segmented": (ctypes.c_int, [_P, _P, _P, _U64, _P, _P, _P, ctypes.POINTER(RankCfg), _P]),
    "kc_root_tree": (ctypes.c_int, [_P, _P, _U64, _P, _P, ctypes.POINTER(TreeCfg), _P]),
    "kc_rank_via_euler_tour": (ctypes.c_int, [_P, ctypes.POINTER(TreeCfg), _P, _P, _U64, _P]),
    "kc_compute_euler_tour": (ctypes.c_int, [_P, _P, _U64, _P, _P, ctypes.POINTER(_U64), _P]),
    "kc_gen_list": (ctypes.c_int, [_P, _U64, _U64, ctypes.c_int, ctypes.c_int, _P, _P, _P]),
    "kc_gen_tree": (ctypes.c_int, [_P, _U64, _U64, _P, _P]),
    "kc_gen_lists_segmented": (ctypes.c_int, [_P, _P, _U64, _P, _P, ctypes.c_int, _P, _P, _P]),
    "kc_probe_random_gather": (ctypes.c_int, [_P, _U64, ctypes.c_uint32, ctypes.POINTER(ctypes.c_double),
                                               ctypes.POINTER(ctypes.c_double)]),
    "kc_probe_random_gather_mode": (ctypes.c_int, [_P, _U64, ctypes.c_uint32, ctypes.c_int,
                                                    ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
    "kc_probe_table_gather": (ctypes.c_int, [_P, _U64, _U64, ctypes.c_uint32, ctypes.POINTER(ctypes.c_double),
                                              ctypes.POINTER(ctypes.c_double)]),
    "kc_probe_record_access": (ctypes.c_int, [_P, _U64, ctypes.c_int, ctypes.c_uint32,
                                               ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
    "kc_is_list": (ctypes.c_int, [_P, _P, _U64, ctypes.POINTER(ctypes.c_int32), _P]),
    "kc_leaf_info": (ctypes.c_int, [_P, _P, _U64, _P, ctypes.POINTER(_U64), _P]),
    "kc_reverse_list": (ctypes.c_int, [_P, _P, _P, _U64, _P, _P, _P, _U64, ctypes.POINTER(_U64), _P]),
    "kc_compute_stats": (ctypes.c_int, [_P, ctypes.c_int32, _P, _P, _U64, ctypes.POINTER(Stats), _P, _P, _P, _P,
                                        _U64, _P]),
    "kc_gen_graph": (ctypes.c_int, [_P, ctypes.c_int32, _U64, _U64, _U64, _P, _P, _P]),
    "kc_gen_rgg2d": (ctypes.c_int, [_P, _U64, _U64, _U64, _P, _P, _U64, ctypes.POINTER(_U64), _P]),
    "kc_build_csr": (ctypes.c_int, [_P, _P, _P, _U64, _U64, _P, _P, _P]),
    "kc_bfs_forest": (ctypes.c_int, [_P, _P, _P, _U64, _P, _P]),
    "kc_trace_enable": (ctypes.c_int, [_P, ctypes.c_int]),
    "kc_trace_read": (ctypes.c_uint32, [_P, ctypes.POINTER(TraceLevel), ctypes.c_uint32]),
    "kc_server_create": (ctypes.c_int, [_P, ctypes.POINTER(ServerCfg), ctypes.POINTER(_P)]),
    "kc_server_destroy": (None, [_P]),
    "kc_submit": (ctypes.c_int, [_P, ctypes.POINTER(Job), ctypes.POINTER(_U64)]),
    "kc_poll": (ctypes.c_int, [_P, _U64]),
    "kc_wait": (ctypes.c_int, [_P, _U64]),
    "kc_server_stats_get": (ctypes.c_int, [_P, ctypes.POINTER(ServerStats)]),
    "kc_server_reset_stats": (ctypes.c_int, [_P]),
}


def load():
    """Loads the HIP library; raises ImportError if it has not been built.

    torch is imported first when it is installed: the library then binds to
    the HIP runtime torch already loaded, so the process has one runtime and
    device tensors can be passed straight through."""
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise ImportError(f"{_LIB_PATH} is missing: run `python -m kascade_amd._build`")
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(_LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a, dtype):
    """Data pointer of a torch tensor or numpy array of the given numpy dtype."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        if a.dtype != dtype:
            raise TypeError(f"expected {np.dtype(dtype)}, got {a.dtype}")
        if not a.flags["C_CONTIGUOUS"]:
            raise ValueError("array must be contiguous")
        return a.ctypes.data
    import torch  # torch tensors (device or host)
    tmap = {np.uint64: torch.uint64, np.int64: torch.int64, np.uint32: torch.uint32,
            np.uint8: torch.uint8}
    if a.dtype != tmap[np.dtype(dtype).type]:
        raise TypeError(f"expected {tmap[np.dtype(dtype).type]}, got {a.dtype}")
    if not a.is_contiguous():
        raise ValueError("tensor must be contiguous")
    return a.data_ptr()


def _st(stream):
    """The stream argument of a call.  None becomes torch's current stream
    when torch has initialised the GPU, so that the library's work is ordered
    after the torch kernels that produced its inputs (and, the call being
    synchronous, before the ones queued after it).  torch's default stream is
    HIP's null stream, which the C ABI does not take (NULL selects the
    context's stream): its queued work is waited for on the host, and the
    call runs on the context's stream.  Without torch on the GPU, None selects
    the context's stream."""
    if stream is not None:
        return stream
    import sys
    torch = sys.modules.get("torch")
    if torch is None or not torch.cuda.is_initialized():
        return None
    cur = torch.cuda.current_stream()
    if cur.cuda_stream == 0:
        _wait_for(cur)
        return None
    return cur.cuda_stream


def _wait_for(stream):
    """Waits on the host for the work queued so far on a torch stream, and no
    other: synchronising HIP's null stream itself may also wait for the work
    of every other stream on the device (the server's, for one)."""
    import torch
    ev = torch.cuda.Event()
    ev.record(stream)
    ev.synchronize()


def _numel(a):
    return a.size if isinstance(a, np.ndarray) else a.numel()


LAUNCHES = {"auto": 0, "device": 1, "host": 2}
JOB_KINDS = {"list": 0, "tree": 1}  # kc_job_kind
TUNING_KINDS = {"list": 0, "tree": 1, "segmented": 2}  # kc_tuning_kind


SPAWNS = {"auto": 0, "off": 1, "on": 2}  # kc_spawn
CONTRACTIONS = {"auto": 0, "off": 1, "on": 2}  # kc_contract
RECORDS = {"auto": 0, "wide": 1}  # kc_records


def _cfg(algorithm="auto", ruler_spacing=0, ruler_spacing_deep=0, max_levels=0, validate=False,
         chase_blocks_per_cu=0, chase_chunk=0, launch="auto", spawn="auto", local_contraction="auto",
         records="auto"):
    c = RankCfg()
    c.records = RECORDS[records] if isinstance(records, str) else int(records)
    c.local_contraction = (CONTRACTIONS[local_contraction] if isinstance(local_contraction, str)
                           else int(local_contraction))
    c.spawn = SPAWNS[spawn] if isinstance(spawn, str) else int(spawn)
    c.launch = LAUNCHES[launch] if isinstance(launch, str) else int(launch)
    c.chase_blocks_per_cu = chase_blocks_per_cu
    c.chase_chunk = chase_chunk
    c.algorithm = ALGORITHMS[algorithm] if isinstance(algorithm, str) else int(algorithm)
    c.ruler_spacing = ruler_spacing
    c.ruler_spacing_deep = ruler_spacing_deep
    c.max_levels = max_levels
    c.validate = int(bool(validate))
    return c


def _like(a, dtype, n=None):
    n = _numel(a) if n is None else n
    if isinstance(a, np.ndarray):
        return np.empty(n, dtype)
    import torch
    tmap = {np.uint64: torch.uint64, np.int64: torch.int64, np.uint32: torch.uint32, np.uint8: torch.uint8}
    return torch.empty(n, dtype=tmap[np.dtype(dtype).type], device=a.device)


class Context:
    """One device context (kc_context_create).  Not thread-safe, like a communicator."""

    def __init__(self, device=0, mem_budget_bytes=0):
        self._lib = load()
        h = ctypes.c_void_p()
        self._check(self._lib.kc_context_create(device, mem_budget_bytes, ctypes.byref(h)), None)
        self.handle = h
        self.device = device

    def _check(self, status, handle=True):
        if status != KC_OK:
            msg = ""
            if handle is not None and getattr(self, "handle", None):
                raw = self._lib.kc_last_error(self.handle)
                msg = raw.decode() if raw else ""
            raise KascadeError(status, msg)

    def close(self):
        if getattr(self, "handle", None):
            self._lib.kc_context_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- list ranking ---------------------------------------------------
    def rank_list(self, succ, weight=None, root_out=None, rank_out=None, stream=None, **cfg):
        n = _numel(succ)
        root_out = _like(succ, np.uint64) if root_out is None else root_out
        rank_out = _like(succ, np.int64) if rank_out is None else rank_out
        c = _cfg(**cfg)
        self._check(self._lib.kc_rank_list(self.handle, _ptr(succ, np.uint64), _ptr(weight, np.int64), n,
                                           _ptr(root_out, np.uint64), _ptr(rank_out, np.int64),
                                           ctypes.byref(c), _st(stream)))
        return root_out, rank_out

    NOALLOC_MAX_N = 1 << 20  # KC_NOALLOC_MAX_N

    def rank_list_noalloc(self, succ, weight, root_out, rank_out, stream=None, **cfg):
        """kc_rank_list_noalloc: rank_list with no device allocation (scratch
        from the library's static array); device tensors only, caller-provided
        outputs, n at most NOALLOC_MAX_N."""
        c = _cfg(**cfg)
        self._check(self._lib.kc_rank_list_noalloc(self.handle, _ptr(succ, np.uint64), _ptr(weight, np.int64),
                                                   _numel(succ), _ptr(root_out, np.uint64),
                                                   _ptr(rank_out, np.int64), ctypes.byref(c), _st(stream)))
        return root_out, rank_out

    def pointer_doubling(self, succ_array, rank_array, stream=None, **cfg):
        """In place: succ -> root, rank (weights) -> rank (pointer_doubling.hpp:18-23)."""
        c = _cfg(**cfg)
        self._check(self._lib.kc_pointer_doubling(self.handle, ctypes.byref(c), _ptr(succ_array, np.uint64),
                                                  _ptr(rank_array, np.int64), _numel(succ_array), _st(stream)))

    def sparse_ruling_set(self, succ_array, rank_array, stream=None, **cfg):
        """In place, as kascade::sparse_ruling_set (sparse_ruling_set.hpp:13-18)."""
        c = _cfg(**cfg)
        self._check(self._lib.kc_sparse_ruling_set(self.handle, ctypes.byref(c), _ptr(succ_array, np.uint64),
                                                   _ptr(rank_array, np.int64), _numel(succ_array), _st(stream)))

    def set_initial_ranking_state(self, succ, root_out=None, rank_out=None, stream=None):
        root_out = _like(succ, np.uint64) if root_out is None else root_out
        rank_out = _like(succ, np.int64) if rank_out is None else rank_out
        self._check(self._lib.kc_set_initial_ranking_state(self.handle, _ptr(succ, np.uint64), _numel(succ),
                                                           _ptr(root_out, np.uint64), _ptr(rank_out, np.int64),
                                                           _st(stream)))
        return root_out, rank_out

    # ---- per-level trace (RulerTrace) --------------------------------------------
    def trace(self, on=True):
        self._check(self._lib.kc_trace_enable(self.handle, int(bool(on))))

    def trace_read(self):
        k = self._lib.kc_trace_read(self.handle, None, 0)
        arr = (TraceLevel * max(1, k))()
        self._lib.kc_trace_read(self.handle, arr, k)
        out = []
        for t in arr[:k]:
            d = {f: getattr(t, f) for f, _ in TraceLevel._fields_}
            d["kind"] = TRACE_KINDS[t.kind]
            out.append(d)
        return out

    # ---- successor utilities and statistics -------------------------------------
    def is_list(self, succ, stream=None):
        """kascade::is_list: no element has two predecessors."""
        out = ctypes.c_int32()
        self._check(self._lib.kc_is_list(self.handle, _ptr(succ, np.uint64), _numel(succ), ctypes.byref(out), _st(stream)))
        return bool(out.value)

    def leaf_info(self, succ, stream=None):
        """kascade::LeafInfo: (is_leaf flags as uint8, number of leaves)."""
        flags = _like(succ, np.uint8)
        cnt = ctypes.c_uint64()
        self._check(self._lib.kc_leaf_info(self.handle, _ptr(succ, np.uint64), _numel(succ), _ptr(flags, np.uint8),
                                           ctypes.byref(cnt), _st(stream)))
        return flags, cnt.value

    def reverse_list(self, succ, dist_to_succ=None, stream=None):
        """kascade::reverse_list: (pred, dist_to_pred, original roots ascending)."""
        n = _numel(succ)
        pred = _like(succ, np.uint64)
        dpred = _like(succ, np.int64)
        roots = _like(succ, np.uint64)
        nr = ctypes.c_uint64()
        self._check(self._lib.kc_reverse_list(self.handle, _ptr(succ, np.uint64), _ptr(dist_to_succ, np.int64), n,
                                              _ptr(pred, np.uint64), _ptr(dpred, np.int64), _ptr(roots, np.uint64),
                                              n, ctypes.byref(nr), _st(stream)))
        return pred, dpred, roots[:nr.value]

    def compute_stats(self, root, rank, level="basic", stream=None):
        """compute_stats (bench/detail/stats.cpp): a dict shaped like the
        reference's serialized Stats, plus the per-tree list for "extensive"."""
        lv = STATS_LEVELS[level] if isinstance(level, str) else int(level)
        n = _numel(root)
        st = Stats()
        cap = n if lv == 4 else 0
        tr = ts = tm = tsum = None
        if cap:
            tr, ts = np.empty(cap, np.uint64), np.empty(cap, np.uint64)
            tm, tsum = np.empty(cap, np.int64), np.empty(cap, np.int64)
        self._check(self._lib.kc_compute_stats(self.handle, lv, _ptr(root, np.uint64), _ptr(rank, np.int64), n,
                                               ctypes.byref(st), _ptr(tr, np.uint64), _ptr(ts, np.uint64),
                                               _ptr(tm, np.int64), _ptr(tsum, np.int64), cap, _st(stream)))
        out = {}
        if lv >= 2:
            out["basic_stats"] = {"num_trees": st.num_trees, "max_rank": st.max_rank, "avg_rank": st.avg_rank}
        if lv >= 3:
            ext = {k: getattr(st, k) for k in ("num_trees", "num_nontrivial_trees", "nontrivial_size_sum",
                                               "nontrivial_rank_sum", "max_size", "avg_size",
                                               "nontrivial_avg_size", "nontrivial_avg_rank")}
            if lv == 4:
                k = st.num_nontrivial_trees
                ext["per_tree_stats"] = [{"root": int(tr[i]), "size": int(ts[i]), "max_rank": int(tm[i]),
                                          "rank_sum": int(tsum[i])} for i in range(k)]
            out["extensive_stats"] = ext
        return out

    # ---- graph inputs ----------------------------------------------------------------
    GRAPH_KINDS = {"gnm": 0, "rmat": 1}

    def gen_graph(self, kind, n, m, seed, stream=None):
        """m seeded random undirected edges on n vertices (device tensors)."""
        import torch
        src = torch.empty(m, dtype=torch.uint64, device="cuda")
        dst = torch.empty(m, dtype=torch.uint64, device="cuda")
        self._check(self._lib.kc_gen_graph(self.handle, self.GRAPH_KINDS[kind], n, m, seed, _ptr(src, np.uint64),
                                           _ptr(dst, np.uint64), _st(stream)))
        return src, dst

    def gen_rgg2d(self, n, m, seed, stream=None):
        """Random geometric graph with about m edges on n points (device tensors)."""
        import torch
        cnt = ctypes.c_uint64()
        self._check(self._lib.kc_gen_rgg2d(self.handle, n, m, seed, None, None, 0, ctypes.byref(cnt), _st(stream)))
        src = torch.empty(max(cnt.value, 1), dtype=torch.uint64, device="cuda")
        dst = torch.empty(max(cnt.value, 1), dtype=torch.uint64, device="cuda")
        self._check(self._lib.kc_gen_rgg2d(self.handle, n, m, seed, _ptr(src, np.uint64), _ptr(dst, np.uint64),
                                           cnt.value, ctypes.byref(cnt), _st(stream)))
        return src[:cnt.value], dst[:cnt.value]

    def build_csr(self, src, dst, n, stream=None):
        m = _numel(src)
        xadj = _like(src, np.uint64, n + 1)
        adj = _like(src, np.uint64, 2 * m)
        self._check(self._lib.kc_build_csr(self.handle, _ptr(src, np.uint64), _ptr(dst, np.uint64), n, m,
                                           _ptr(xadj, np.uint64), _ptr(adj, np.uint64), _st(stream)))
        return xadj, adj

    def bfs_forest(self, xadj, adj, stream=None):
        """generate_bfs_tree: the BFS forest's parent array."""
        n = _numel(xadj) - 1
        parent = _like(xadj, np.uint64, n)
        self._check(self._lib.kc_bfs_forest(self.handle, _ptr(xadj, np.uint64), _ptr(adj, np.uint64), n,
                                            _ptr(parent, np.uint64), _st(stream)))
        return parent

    def pd_rounds(self, succ, weight, rounds, stream=None):
        so = _like(succ, np.uint64)
        ro = _like(succ, np.int64)
        self._check(self._lib.kc_pointer_doubling_rounds(self.handle, _ptr(succ, np.uint64), _ptr(weight, np.int64),
                                                         _numel(succ), rounds, _ptr(so, np.uint64),
                                                         _ptr(ro, np.int64), _st(stream)))
        return so, ro

    def pack(self, succ, rank, active, stream=None):
        na = _numel(active)
        ps = _like(active, np.uint64, na)
        pr = _like(active, np.int64, na)
        self._check(self._lib.kc_pack(self.handle, _ptr(succ, np.uint64), _ptr(rank, np.int64), _numel(succ),
                                      _ptr(active, np.uint64), na, _ptr(ps, np.uint64), _ptr(pr, np.int64),
                                      _st(stream)))
        return ps, pr

    def unpack(self, psucc, prank, active, succ, rank, stream=None):
        self._check(self._lib.kc_unpack(self.handle, _ptr(psucc, np.uint64), _ptr(prank, np.int64),
                                        _ptr(active, np.uint64), _numel(active), _ptr(succ, np.uint64),
                                        _ptr(rank, np.int64), _numel(succ), _st(stream)))

    def rank_lists_segmented(self, succ_local, seg_off, weight=None, root_out=None, rank_out=None,
                             stream=None, **cfg):
        total = _numel(succ_local)
        root_out = _like(succ_local, np.uint32) if root_out is None else root_out
        rank_out = _like(succ_local, np.int64, total) if rank_out is None else rank_out
        c = _cfg(**cfg)
        self._check(self._lib.kc_rank_lists_segmented(self.handle, _ptr(succ_local, np.uint32),
                                                       _ptr(seg_off, np.uint64), _numel(seg_off) - 1,
                                                       _ptr(weight, np.int64), _ptr(root_out, np.uint32),
                                                       _ptr(rank_out, np.int64), ctypes.byref(c), _st(stream)))
        return root_out, rank_out

    # ---- trees --------------------------------------------------------------
    def root_tree(self, parent, root_out=None, depth_out=None, stream=None, tour_build="auto", **cfg):
        root_out = _like(parent, np.uint64) if root_out is None else root_out
        depth_out = _like(parent, np.int64) if depth_out is None else depth_out
        t = TreeCfg()
        t.rank = _cfg(**cfg)
        t.tour_build = TOUR_BUILDS[tour_build]
        self._check(self._lib.kc_root_tree(self.handle, _ptr(parent, np.uint64), _numel(parent),
                                           _ptr(root_out, np.uint64), _ptr(depth_out, np.int64),
                                           ctypes.byref(t), _st(stream)))
        return root_out, depth_out

    def rank_via_euler_tour(self, succ_array, rank_array, stream=None, tour_build="auto", **cfg):
        """In place: parents -> roots, rank -> depth (eulertour.hpp:22-27)."""
        t = TreeCfg()
        t.rank = _cfg(**cfg)
        t.tour_build = TOUR_BUILDS[tour_build]
        self._check(self._lib.kc_rank_via_euler_tour(self.handle, ctypes.byref(t), _ptr(succ_array, np.uint64),
                                                     _ptr(rank_array, np.int64), _numel(succ_array), _st(stream)))

    def compute_euler_tour(self, parent, stream=None):
        """kascade::compute_euler_tour in the reference's element order:
        (tour successors, tour weights), 2(n - roots) + roots entries."""
        n = _numel(parent)
        s = _like(parent, np.uint64, 2 * n)
        w = _like(parent, np.int64, 2 * n)
        ln = ctypes.c_uint64()
        self._check(self._lib.kc_compute_euler_tour(self.handle, _ptr(parent, np.uint64), n,
                                                    _ptr(s, np.uint64), _ptr(w, np.int64), ctypes.byref(ln),
                                                    _st(stream)))
        return s[:ln.value], w[:ln.value]

    # ---- synthetic inputs --------------------------------------------------------
    def gen_list(self, n, seed, weighted=True, local=False, device="cuda", stream=None):
        succ, w = _alloc_pair(n, device)
        self._check(self._lib.kc_gen_list(self.handle, n, seed, int(weighted), int(local),
                                          _ptr(succ, np.uint64), _ptr(w, np.int64), _st(stream)))
        return succ, w

    def gen_tree(self, n, seed, device="cuda", stream=None):
        parent, _ = _alloc_pair(n, device, weights=False)
        self._check(self._lib.kc_gen_tree(self.handle, n, seed, _ptr(parent, np.uint64), _st(stream)))
        return parent

    def gen_lists_segmented(self, seg_off, seeds, local_flags=None, weighted=True, device="cuda",
                            stream=None):
        total = int(seg_off[-1])
        if device is None or device == "cpu":
            succ = np.empty(total, np.uint32)
            w = np.empty(total, np.int64)
        else:
            import torch
            succ = torch.empty(total, dtype=torch.uint32, device=device)
            w = torch.empty(total, dtype=torch.int64, device=device)
        self._check(self._lib.kc_gen_lists_segmented(self.handle, _ptr(seg_off, np.uint64), _numel(seg_off) - 1,
                                                     _ptr(seeds, np.uint64),
                                                     _ptr(local_flags, np.uint8) if local_flags is not None else None,
                                                     int(weighted), _ptr(succ, np.uint32), _ptr(w, np.int64),
                                                     _st(stream)))
        return succ, w

    # ---- phase timing -----------------------------------------------------------
    def profile(self, on=True):
        self._check(self._lib.kc_profile_enable(self.handle, int(on)))

    def profile_reset(self):
        self._check(self._lib.kc_profile_reset(self.handle))

    def profile_read(self):
        """{phase: (total_ms, launches)} accumulated since the last reset."""
        k = self._lib.kc_profile_read(self.handle, 0, None, None, None)
        names = (ctypes.c_char_p * max(k, 1))()
        ms = (ctypes.c_double * max(k, 1))()
        cnt = (ctypes.c_uint64 * max(k, 1))()
        k = self._lib.kc_profile_read(self.handle, k, names, ms, cnt)
        return {names[i].decode(): (ms[i], cnt[i]) for i in range(k)}

    def probe_random_gather(self, n=1 << 27, reps=5, mode=0):
        g = ctypes.c_double()
        ms = ctypes.c_double()
        self._check(self._lib.kc_probe_random_gather_mode(self.handle, n, reps, mode, ctypes.byref(g),
                                                          ctypes.byref(ms)))
        return g.value, ms.value

    def probe_table_gather(self, n, m, reps=5):
        g = ctypes.c_double()
        ms = ctypes.c_double()
        self._check(self._lib.kc_probe_table_gather(self.handle, n, m, reps, ctypes.byref(g), ctypes.byref(ms)))
        return g.value, ms.value

    RECORD_PATTERNS = ("read", "write", "read_write", "walk_read", "walk_claim", "walk_read8", "walk_claim8")

    def probe_record_access(self, pattern, n=1 << 27, reps=5):
        """Records per second and ms for one access pattern of 16-byte records
        (see RECORD_PATTERNS; the "8" walks use 8-byte records) over a random
        cycle of n records."""
        r = ctypes.c_double()
        ms = ctypes.c_double()
        self._check(self._lib.kc_probe_record_access(self.handle, n, self.RECORD_PATTERNS.index(pattern), reps,
                                                     ctypes.byref(r), ctypes.byref(ms)))
        return r.value, ms.value

    # ---- tuning ---------------------------------------------------------------
    def tuning_set(self, rows, kind="list"):
        """rows: iterable of (log2n, spacing_top, spacing_deep, chase_blocks_per_cu, chase_chunk);
        kind "list" (classes count elements), "tree" (classes count vertices) or
        "segmented" (classes count a batch's elements)."""
        flat = [int(x) for r in rows for x in r]
        arr = (ctypes.c_uint32 * max(1, len(flat)))(*flat)
        self._check(self._lib.kc_tuning_set_kind(self.handle, TUNING_KINDS[kind], arr, len(flat) // 5))

    def tuning_get(self, n, kind="list"):
        out = (ctypes.c_uint32 * 5)()
        rows = self._lib.kc_tuning_get_kind(self.handle, TUNING_KINDS[kind], n, out)
        return {"spacing_top": out[0], "spacing_deep": out[1], "chase_blocks_per_cu": out[2],
                "chase_chunk": out[3], "small_n": out[4], "table_rows": rows}

    def scratch_bytes(self):
        return self._lib.kc_context_scratch_bytes(self.handle)

    def trim(self):
        self._check(self._lib.kc_context_trim(self.handle))

    def launch_stats(self):
        """{'graph_captures', 'graph_replays', 'host_redos'} since the context was created."""
        out = (ctypes.c_uint64 * 3)()
        self._check(self._lib.kc_context_launch_stats(self.handle, out))
        return {"graph_captures": out[0], "graph_replays": out[1], "host_redos": out[2]}

    def record_stats(self):
        """{'narrow_calls', 'narrow_redos'}: list rankings whose top level used
        8-byte records, and those redone with 16-byte records."""
        out = (ctypes.c_uint64 * 2)()
        self._check(self._lib.kc_context_record_stats(self.handle, out))
        return {"narrow_calls": out[0], "narrow_redos": out[1]}


def release_cached_scratch(device=0):
    """Frees the scratch chunks that destroyed contexts left in the device's cache."""
    st = load().kc_release_cached_scratch(device)
    if st != KC_OK:
        raise KascadeError(st, "kc_release_cached_scratch")


def _alloc_pair(n, device, weights=True):
    if device is None or device == "cpu":
        return np.empty(n, np.uint64), (np.empty(n, np.int64) if weights else None)
    import torch
    return (torch.empty(n, dtype=torch.uint64, device=device),
            torch.empty(n, dtype=torch.int64, device=device) if weights else None)


class Server:
    """Continuous-batching request server (kc_server_*); a job's arrays may be host or device memory."""

    def __init__(self, ctx, mem_budget_bytes=200 << 30, max_batch_jobs=0, num_slots=0,
                 max_batch_elements=0, one_job_per_launch=False, tour_build="auto"):
        self._ctx = ctx
        self._lib = ctx._lib
        cfg = ServerCfg()
        cfg.mem_budget_bytes = mem_budget_bytes
        cfg.max_batch_jobs = max_batch_jobs
        cfg.num_slots = num_slots
        cfg.max_batch_elements = max_batch_elements
        cfg.one_job_per_launch = int(bool(one_job_per_launch))
        cfg.tour_build = TOUR_BUILDS[tour_build]
        h = ctypes.c_void_p()
        ctx._check(self._lib.kc_server_create(ctx.handle, ctypes.byref(cfg), ctypes.byref(h)))
        self.handle = h
        self._keep = {}

    def submit(self, kind, succ, weight=None, root_out=None, rank_out=None):
        n = _numel(succ)
        root_out = np.empty(n, np.uint64) if root_out is None else root_out
        rank_out = np.empty(n, np.int64) if rank_out is None else rank_out
        # the server's slots run on streams of their own: device arrays must
        # be complete before the job is handed over
        if any(getattr(a, "is_cuda", False) for a in (succ, weight, root_out, rank_out)):
            import torch
            _wait_for(torch.cuda.current_stream())
        j = Job()
        j.kind = JOB_KINDS[kind] if isinstance(kind, str) else int(kind)
        j.n = n
        j.succ = _ptr(succ, np.uint64)
        j.weight = _ptr(weight, np.int64)
        j.root_out = _ptr(root_out, np.uint64)
        j.rank_out = _ptr(rank_out, np.int64)
        t = ctypes.c_uint64()
        st = self._lib.kc_submit(self.handle, ctypes.byref(j), ctypes.byref(t))
        if st != KC_OK:
            raise KascadeError(st, "submit failed")
        self._keep[t.value] = (succ, weight, root_out, rank_out)
        return t.value, root_out, rank_out

    def poll(self, ticket):
        st = self._lib.kc_poll(self.handle, ticket)
        if st == KC_EAGAIN:
            return False
        self._keep.pop(ticket, None)
        if st != KC_OK:
            raise KascadeError(st, f"job {ticket} failed")
        return True

    def wait(self, ticket):
        st = self._lib.kc_wait(self.handle, ticket)
        self._keep.pop(ticket, None)
        if st != KC_OK:
            raise KascadeError(st, f"job {ticket} failed")

    def stats(self):
        s = ServerStats()
        st = self._lib.kc_server_stats_get(self.handle, ctypes.byref(s))
        if st != KC_OK:
            raise KascadeError(st, "stats")
        return {k: getattr(s, k) for k, _ in ServerStats._fields_}

    def reset_stats(self):
        self._lib.kc_server_reset_stats(self.handle)

    def close(self):
        if getattr(self, "handle", None):
            self._lib.kc_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
