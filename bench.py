"""Benchmark: elements ranked per second for one random weighted list of
n = 2^27 elements on one MI355X (sparse ruling set through the C ABI).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N] [--algorithm srs|pd]
    python bench.py --workload tree|segmented|serve

Inputs: a seeded random list (successors of a Feistel permutation, int64
weights uniform in [1,100], 0 on the tail) generated on the device and
resident in HBM before the timed region.  One step = one kc_rank_list call
(pack, ruler chase, recursion, propagation), synchronous.  The last step's
output is compared bit for bit with the generator's closed form.  Prints one
JSON line.  The random-gather ceiling of the device is measured in the same
run (u64 gather through a random permutation of 2^27) and the dominant
kernel is reported against both it and the HBM peak.  The CPU baseline is
the oracle's restatement of the reference's MPI path (sparse ruling set and
sync pointer doubling on emulated ranks, one thread per rank, plus the
sequential walk on one core), timed on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
# The level-0 chase's measured HBM traffic per element (rocprofv3 FETCH_SIZE and
# WRITE_SIZE, scaled by a copy kernel of known size) and the configuration it
# was measured in, as the committed profile summary states them.
CHASE_PROFILE = os.path.join("profiles", "chase_propagate_list_2e27_weighted.md")
BYTES_PER_ELEMENT = 32  # succ 8 in + weight 8 in + root 8 out + rank 8 out
# The chase is credited with the ranking's compulsory bytes per element (succ
# and weight in, root and rank out: 32 B), whatever record width it moves them
# in: 16-byte records, or the narrow 8-byte level-0 records.
CHASE_BYTES_PER_ELEMENT = 32


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=["list", "tree", "segmented", "serve", "unit-list"], default="list",
                    help="list: one random weighted list (the headline); tree: random rooted tree "
                         "(Euler tour); segmented: 20,000 lists of log-uniform size in one batch; "
                         "serve: the mixed serving stream alone; unit-list: the 2^20 unit-weight list "
                         "section alone")
    ap.add_argument("--serve-mode", choices=["all", "batched", "serial"], default="all",
                    help="with --workload serve: every mode, or only the coalescing server (batched) "
                         "or only the one-job-per-launch baseline (serial), each at capacity")
    ap.add_argument("--n", type=int, default=0, help="elements (list: 2^27, tree: 2^26 by default)")
    ap.add_argument("--segments", type=int, default=20000)
    ap.add_argument("--tour-build", choices=["auto", "linked", "sorted"], default="auto")
    ap.add_argument("--seed", type=int, default=2027)
    ap.add_argument("--algorithm", default="srs")
    ap.add_argument("--spacing", type=int, default=0)
    ap.add_argument("--spacing-deep", type=int, default=0)
    ap.add_argument("--spawn", choices=["auto", "off", "on"], default="auto", help="ruler spawn (kc_spawn)")
    ap.add_argument("--records", choices=["auto", "wide"], default="auto", help="top-level record width (kc_records)")
    ap.add_argument("--chase-blocks", type=int, default=0,
                    help="resident chase workgroups per CU (0: the tuning table's)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=1 << 25,
                    help="elements in the CPU baseline's sample (trees: a quarter as many vertices); "
                         "about 1-2 s on 16 threads, 15-30 s of CPU work")
    ap.add_argument("--no-check", action="store_true", help="skip the bit-exact check of the last step")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-host-buffers", action="store_true",
                    help="skip the host-buffer (PCIe-inclusive) rates of the headline ranking")
    ap.add_argument("--no-unit-list", action="store_true",
                    help="skip the 2^20 unit-weight list section (GPU beside the CPU path on 4 ranks)")
    ap.add_argument("--serve-jobs", type=int, default=300)
    ap.add_argument("--serve-slots", type=int, default=3)
    return ap.parse_args()


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def profiled_traffic(phase):
    """(bytes per element, {key: value} of the profiled configuration) for
    `phase` from the profile summary, or (None, {}) when it has none."""
    per_elem, conf = None, {}
    try:
        with open(os.path.join(ROOT, CHASE_PROFILE)) as f:
            for line in f:
                tok = line.split()
                if len(tok) == 3 and tok[0] == "traffic_bytes_per_element" and tok[1] == phase:
                    per_elem = float(tok[2])
                if len(tok) >= 2 and tok[0] == "profiled_config" and tok[1] == phase:
                    conf = dict(t.split("=", 1) for t in tok[2:])
    except OSError:
        pass
    return per_elem, conf


# One GPU's share of the host: the MI355X hosts this is measured on allot 16
# CPUs to each GPU (OMP_NUM_THREADS=16), whatever os.cpu_count() shows.
CORE_CAP = 16
CORE_CAP_NOTE = ("threads capped at 16, the host CPU share of one GPU on this machine; "
                 "the reference's own runs use one MPI rank per core")


def host_cores():
    """Cores this process may run on (the box's CPU share), capped at CORE_CAP."""
    try:
        c = len(os.sched_getaffinity(0))
    except AttributeError:
        c = os.cpu_count() or 1
    return max(1, min(CORE_CAP, c))


def cpu_baseline(n_sample, seed):
    """The oracle's restatement of the reference's CPU path on emulated MPI
    ranks (one thread per rank): the sparse ruling set in the submission
    configuration (sparse_ruling_set.cpp:335-393) on all host cores as the
    value; the same on 4 ranks, sync pointer doubling
    (pointer_doubling_generic.hpp:608-700) and the sequential walk
    (local_pointer_chasing.cpp:18-59, one core) beside it."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as o
    ranks = host_cores()
    succ, w = o.gen_list(n_sample, seed, True)
    er, ek = o.list_closed_form(n_sample, seed, w)

    def run(fn):
        t0 = time.perf_counter()
        res = fn()
        dt = time.perf_counter() - t0
        return dt, bool((res[0] == er).all() and (res[1] == ek).all())

    dt, ok = run(lambda: o.sparse_ruling_set(succ, w, ranks))
    dt4, ok4 = run(lambda: o.sparse_ruling_set(succ, w, 4))
    dt_pd, ok_pd = run(lambda: o.pointer_doubling(succ, w, ranks))
    dt_seq, ok_seq = run(lambda: o.sequential_walk(succ, w))
    rate = lambda t: n_sample / t
    return {"value": rate(dt), "unit": "elements/s", "cores": ranks, "cores_note": CORE_CAP_NOTE, "kind": "port",
            "cpu_model": cpu_model(),
            "sample": f"random weighted list n={n_sample} (2^{n_sample.bit_length() - 1}) of the same generator: "
                      f"sparse ruling set (ultimate-sanders f=2, spawn, post-invert, local contraction, "
                      f"2 rounds + PD) on {ranks} emulated ranks, {dt:.2f} s",
            "srs_4_ranks_elements_per_s": rate(dt4),
            "pd_elements_per_s": rate(dt_pd),
            "sequential_walk_1_core_elements_per_s": rate(dt_seq),
            "bit_exact_vs_closed_form": ok and ok4 and ok_pd and ok_seq}


def host_buffer_rates(ctx, succ, w, cfg, reps=2):
    """The same ranking with every array in host memory (pinned, then
    pageable): the library stages inputs and outputs over PCIe, 32 B per
    element.  Reported beside the headline, never as it."""
    import torch
    n = succ.numel()
    out = {}
    for kind in ("pinned", "pageable"):
        pin = kind == "pinned"
        hs, hw = succ.cpu(), w.cpu()
        hr = torch.empty(n, dtype=torch.uint64, pin_memory=pin)
        hk = torch.empty(n, dtype=torch.int64, pin_memory=pin)
        if pin:
            hs, hw = hs.pin_memory(), hw.pin_memory()
        ctx.rank_list(hs, hw, hr, hk, **cfg)
        t0 = time.perf_counter()
        for _ in range(reps):
            ctx.rank_list(hs, hw, hr, hk, **cfg)
        dt = (time.perf_counter() - t0) / reps
        out[kind] = {"ms": dt * 1e3, "elements_per_s": n / dt, "pcie_gbps": 32 * n / dt / 1e9}
        del hs, hw, hr, hk
    return out


def unit_list_bench(ctx, seed, reps=20):
    """One random list of 2^20 elements, unit weights.  GPU sparse ruling
    set and pointer doubling beside the oracle's restatements of the reference's
    sync pointer doubling and sparse ruling set on 4 emulated MPI ranks; all
    outputs compared."""
    import numpy as np
    import torch
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as o
    n = 1 << 20
    succ_h, w_h = o.gen_list(n, seed, False)
    succ = torch.from_numpy(succ_h).cuda()
    root = torch.empty(n, dtype=torch.uint64, device="cuda")
    rank = torch.empty(n, dtype=torch.int64, device="cuda")
    out = {"n": n, "weights": "unit (set_initial_ranking_state)"}
    for algo in ("srs", "pd"):
        ctx.rank_list(succ, None, root, rank, algorithm=algo)
        torch.cuda.synchronize()
        t = []
        for _ in range(reps):
            t0 = time.perf_counter()
            ctx.rank_list(succ, None, root, rank, algorithm=algo)
            torch.cuda.synchronize()
            t.append(time.perf_counter() - t0)
        out[f"gpu_{algo}_ms"] = 1e3 * sorted(t)[len(t) // 2]
    t0 = time.perf_counter()
    cr, ck, rounds = o.pointer_doubling(succ_h, w_h, 4)
    out["cpu_pd_4ranks_ms"] = 1e3 * (time.perf_counter() - t0)
    out["cpu_pd_rounds"] = rounds
    t0 = time.perf_counter()
    sr, sk = o.sparse_ruling_set(succ_h, w_h, 4)
    out["cpu_srs_4ranks_ms"] = 1e3 * (time.perf_counter() - t0)
    gr, gk = root.cpu().numpy(), rank.cpu().numpy()
    out["bit_exact_gpu_vs_cpu"] = bool((gr == cr).all() and (gk == ck).all()
                                       and (gr == sr).all() and (gk == sk).all())
    return out


def serve_bench(ctx, jobs, seed, slots, mode="all"):
    """Mixed serving stream (60 % weighted random lists, 40 % random rooted trees, sizes
    log-uniform in [2^10, 2^24], device-resident arrays).

    Baseline: one job per launch, one job at a time (submit, rank, wait, next).
    Also measured: one job per launch on `slots` concurrent streams, a
    stronger baseline.  Capacities (all jobs submitted at once), then Poisson
    arrivals at 0.8 x the serial capacity through every mode, then the
    coalescing server at 3 x that offered load.  Each server is measured at
    steady state, after one untimed pass over the pool.  Every batched result
    must equal the one-job-per-launch result bit for bit."""
    import numpy as np
    import torch
    import kascade_amd

    rng = np.random.default_rng(seed)
    pool = []
    for k in range(jobs):
        n = int(2 ** rng.uniform(10, 24))
        if rng.random() < 0.6:
            s, w = ctx.gen_list(n, seed * 7919 + k, weighted=True)
            pool.append(("list", s, w))
        else:
            pool.append(("tree", ctx.gen_tree(n, seed * 7919 + k), None))
    outs = {m: [(torch.empty(p[1].numel(), dtype=torch.uint64, device="cuda"),
                 torch.empty(p[1].numel(), dtype=torch.int64, device="cuda")) for p in pool]
            for m in ("one", "batched")}
    torch.cuda.synchronize()

    def run(mode, arrivals, nslots=slots):
        sv = kascade_amd.Server(ctx, one_job_per_launch=(mode == "one"), num_slots=nslots)
        # steady state: one untimed pass over the pool sizes the slots' scratch
        warm = [sv.submit(kind, s, w, *outs[mode][k])[0] for k, (kind, s, w) in enumerate(pool)]
        for t in warm:
            sv.wait(t)
        sv.reset_stats()
        t0 = time.perf_counter()
        tickets = []
        for k, (kind, s, w) in enumerate(pool):
            if arrivals is not None:
                delay = t0 + arrivals[k] - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
            tickets.append(sv.submit(kind, s, w, *outs[mode][k])[0])
        for t in tickets:
            sv.wait(t)
        st = sv.stats()
        sv.close()
        return st

    if mode != "all":
        # one mode at capacity, three timed passes (for profiling)
        name = "one" if mode == "serial" else "batched"
        passes = [run(name, None, 1 if mode == "serial" else slots) for _ in range(3)]
        return {"jobs": jobs, "mode": mode, "capacity_jobs_per_s": [p["jobs_per_s"] for p in passes]}
    cap_serial = run("one", None, 1)["jobs_per_s"]
    cap_one = run("one", None)["jobs_per_s"]
    cap_batched = run("batched", None)["jobs_per_s"]
    lam = 0.8 * cap_serial
    arr = np.cumsum(rng.exponential(1.0 / lam, size=jobs))
    serial = run("one", arr, 1)
    one = run("one", arr)
    batched = run("batched", arr)
    arr3 = arr / 3.0
    batched3 = run("batched", arr3)
    same = all(bool(torch.equal(outs["one"][k][0], outs["batched"][k][0]) and
                    torch.equal(outs["one"][k][1], outs["batched"][k][1])) for k in range(jobs))
    keep = ("jobs_per_s", "p50_latency_ms", "p99_latency_ms", "mean_latency_ms", "batches_launched")
    return {
        "jobs": jobs, "mix": "60% list (int64 weights [1,100]) / 40% tree, n log-uniform in [2^10, 2^24], "
                             "arrays resident in HBM",
        "capacity_one_job_per_launch_serial_jobs_per_s": cap_serial,
        "capacity_one_job_per_launch_concurrent_jobs_per_s": cap_one,
        "concurrent_streams": slots,
        "capacity_batched_jobs_per_s": cap_batched,
        "capacity_ratio": cap_batched / cap_serial,
        "capacity_ratio_vs_concurrent_one_job": cap_batched / cap_one,
        "offered_load_jobs_per_s": lam,
        "one_job_per_launch_serial": {k: serial[k] for k in keep},
        "one_job_per_launch_concurrent": {k: one[k] for k in keep},
        "batched": {k: batched[k] for k in keep},
        "batched_at_3x_load": {"offered_load_jobs_per_s": 3 * lam, **{k: batched3[k] for k in keep}},
        "bit_exact_alone_vs_mix": same,
    }


def timed(ctx, fn, steps, warmup):
    """Seconds per step over `steps` calls after `warmup` untimed ones, then
    the per-phase device times (HIP events) from a second pass of `steps`
    calls.  The timed pass runs as a user would (repeated calls replay their
    captured HIP graph); the profiled pass launches kernel by kernel so that
    every phase can be bracketed by events."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ctx.profile(True)
    ctx.profile_reset()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    phases = ctx.profile_read()
    ctx.profile(False)
    return dt / steps, phases


def tree_main(args, ctx):
    """Random rooted tree: parent of generation index i uniform in [0, i),
    relabelled by a seeded permutation; Euler-tour rooting."""
    import torch
    n = args.n or (1 << 26)
    parent = ctx.gen_tree(n, args.seed)
    root = torch.empty(n, dtype=torch.uint64, device="cuda")
    depth = torch.empty(n, dtype=torch.int64, device="cuda")
    tcfg = dict(spawn=args.spawn, records=args.records, ruler_spacing=args.spacing, ruler_spacing_deep=args.spacing_deep,
                chase_blocks_per_cu=args.chase_blocks)
    t, phases = timed(ctx, lambda: ctx.root_tree(parent, root, depth, tour_build=args.tour_build, **tcfg), args.steps,
                      args.warmup)
    out = {"metric": "vertices rooted/s", "value": n / t, "unit": "vertices/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": t * 1e3, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "int64",
           "data": "synthetic: random recursive tree, seeded relabelling",
           "config": {"workload": f"random rooted tree n={n}, Euler-tour rooting", "n": n, "seed": args.seed,
                      "tour_build": args.tour_build},
           "effective_gbps": 24 * n / t / 1e9,
           "phases_ms_per_step": {k: v[0] / args.steps for k, v in sorted(phases.items())},
           "launch": ctx.launch_stats()}
    if not args.no_check:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as o
        er, ed = o.tree_closed_form(n, args.seed)
        out["bit_exact_vs_closed_form"] = bool((root.cpu().numpy() == er).all() and (depth.cpu().numpy() == ed).all())
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = tree_cpu_baseline(min(args.cpu_sample // 4, n), args.seed)
    print(json.dumps(out), flush=True)


def tree_cpu_baseline(n_sample, seed):
    """The oracle's restatement of the reference's rank_via_euler_tour
    (eulertour.cpp:337-377) on all host cores: tour build and linking on one
    thread per core, ranking by the sparse ruling set on as many emulated ranks
    (experiments/eulertour-sparse-ruling-set.suite.yaml:86-92), on a bounded
    sample of the same tree generator.  Beside it: the one-core restatement
    (hash-table linking, GatherChase ranking) on a quarter of the sample."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as o
    ranks = host_cores()
    parent = o.gen_tree(n_sample, seed)
    er, ed = o.tree_closed_form(n_sample, seed)
    t0 = time.perf_counter()
    r, d = o.euler_tour_rooting_mt(parent, ranks)
    dt = time.perf_counter() - t0
    n1 = max(1, n_sample // 4)
    p1 = o.gen_tree(n1, seed)
    e1r, e1d = o.tree_closed_form(n1, seed)
    t0 = time.perf_counter()
    r1, d1 = o.euler_tour_rooting(p1)
    dt1 = time.perf_counter() - t0
    return {"value": n_sample / dt, "unit": "vertices/s", "cores": ranks, "cores_note": CORE_CAP_NOTE, "kind": "port", "cpu_model": cpu_model(),
            "sample": f"random rooted tree n={n_sample} of the same generator, Euler-tour rooting on {ranks} threads "
                      f"(parallel tour build and linking, sparse ruling set on {ranks} emulated ranks, map back), "
                      f"{dt:.2f} s",
            "one_core_vertices_per_s": n1 / dt1,
            "one_core_sample": f"n={n1}, hash-table linking and sequential walk, {dt1:.2f} s",
            "bit_exact_vs_closed_form": bool((r == er).all() and (d == ed).all() and (r1 == e1r).all()
                                             and (d1 == e1d).all())}


def segmented_main(args, ctx):
    """20,000 lists, sizes log-uniform in [2^8, 2^20], 30 % locality-biased,
    u32 local ids, unit weights, ranked in one segmented call."""
    import numpy as np
    import torch
    rng = np.random.default_rng(args.seed)
    sizes = np.floor(2.0 ** rng.uniform(8, 20, size=args.segments)).astype(np.uint64)
    off = np.zeros(args.segments + 1, np.uint64)
    off[1:] = np.cumsum(sizes)
    seeds = rng.integers(1, 1 << 62, size=args.segments, dtype=np.uint64)
    local = (rng.random(args.segments) < 0.3).astype(np.uint8)
    total = int(off[-1])
    d_off = torch.from_numpy(off).cuda()
    succ, _ = ctx.gen_lists_segmented(d_off, torch.from_numpy(seeds).cuda(), torch.from_numpy(local).cuda(),
                                      weighted=False)
    root = torch.empty(total, dtype=torch.uint32, device="cuda")
    rank = torch.empty(total, dtype=torch.int64, device="cuda")
    scfg = dict(spawn=args.spawn, records=args.records, ruler_spacing=args.spacing, ruler_spacing_deep=args.spacing_deep,
                chase_blocks_per_cu=args.chase_blocks)
    rs0 = ctx.record_stats()
    t, phases = timed(ctx, lambda: ctx.rank_lists_segmented(succ, d_off, None, root, rank, **scfg), args.steps,
                      args.warmup)
    rs1 = ctx.record_stats()
    knobs = ctx.tuning_get(total, kind="segmented")
    out = {"metric": "elements ranked/s (segmented batch)", "value": total / t, "unit": "elements/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": t * 1e3,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int64",
           "data": "synthetic: seeded random permutations, 30% locality-biased (blocks of 10)",
           "config": {"workload": f"{args.segments} lists log-uniform in [2^8, 2^20], one segmented call",
                      "segments": args.segments, "total_elements": total, "seed": args.seed},
           "effective_gbps": (16 * total + 8 * args.segments) / t / 1e9,
           "phases_ms_per_step": {k: v[0] / args.steps for k, v in sorted(phases.items())},
           "launch": ctx.launch_stats(),
           # 8 when every warm-up, timed and profiled call used 8-byte level-0 records
           "level0_record_bytes": 8 if rs1["narrow_calls"] - rs0["narrow_calls"] == args.warmup + 2 * args.steps
           else 16,
           "tuning": {k: knobs[k] for k in ("spacing_top", "spacing_deep", "chase_blocks_per_cu", "chase_chunk")}}
    if not args.no_check:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as o
        ok = True
        for s in rng.choice(args.segments, size=min(40, args.segments), replace=False):
            a, b = int(off[s]), int(off[s + 1])
            es, ew = o.gen_list(b - a, int(seeds[s]), False, local=bool(local[s]))
            er, ek = o.sequential_walk(es, ew)
            ok &= bool((root[a:b].cpu().numpy() == er).all() and (rank[a:b].cpu().numpy() == ek).all())
        out["bit_exact_sampled_segments"] = ok
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = segmented_cpu_baseline(off, seeds, local, args.cpu_sample)
    print(json.dumps(out), flush=True)


def segmented_cpu_baseline(off, seeds, local, max_elements):
    """The oracle's restatement of the reference's sparse ruling set
    (sparse_ruling_set.cpp:335-393, submission configuration) on emulated MPI
    ranks, ranking the batch's first segments (at most max_elements elements
    in all) as one forest with global ids, as the reference would rank the
    whole batch; compared with the sequential walk."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as o
    k = 0
    while k < len(seeds) and int(off[k + 1]) <= max_elements:
        k += 1
    k = max(k, 1)
    total = int(off[k])
    succ = np.empty(total, np.uint64)
    w = np.empty(total, np.int64)
    for s in range(k):
        a, b = int(off[s]), int(off[s + 1])
        ls, lw = o.gen_list(b - a, int(seeds[s]), False, local=bool(local[s]))
        succ[a:b] = ls + np.uint64(a)
        w[a:b] = lw
    ranks = host_cores()
    t0 = time.perf_counter()
    r, rk = o.sparse_ruling_set(succ, w, ranks)
    dt = time.perf_counter() - t0
    er, ek = o.sequential_walk(succ, w)
    return {"value": total / dt, "unit": "elements/s", "cores": ranks, "cores_note": CORE_CAP_NOTE, "kind": "port", "cpu_model": cpu_model(),
            "sample": f"the batch's first {k} segments ({total} elements) as one forest: sparse ruling set "
                      f"(ultimate-sanders f=2, spawn, post-invert, local contraction, 2 rounds + PD) on {ranks} "
                      f"emulated ranks, {dt:.2f} s",
            "bit_exact_vs_sequential_walk": bool((r == er).all() and (rk == ek).all())}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)
    import torch
    import kascade_amd

    ctx = kascade_amd.Context(0)
    if args.workload == "tree":
        return tree_main(args, ctx)
    if args.workload == "segmented":
        return segmented_main(args, ctx)
    if args.workload == "unit-list":
        print(json.dumps({"metric": "elements ranked/s @ n=2^20, unit weights", "n_gpus": 1,
                          "unit_list_2e20": unit_list_bench(ctx, args.seed)}), flush=True)
        return
    if args.workload == "serve":
        out = {"metric": "serving capacity (jobs/s) and p50/p99 latency", "n_gpus": 1,
               "serve": serve_bench(ctx, args.serve_jobs, args.seed, args.serve_slots, args.serve_mode)}
        print(json.dumps(out), flush=True)
        return
    n = args.n or (1 << 27)
    succ, w = ctx.gen_list(n, args.seed, weighted=True)
    root = torch.empty(n, dtype=torch.uint64, device="cuda")
    rank = torch.empty(n, dtype=torch.int64, device="cuda")
    cfg = dict(algorithm=args.algorithm, ruler_spacing=args.spacing, ruler_spacing_deep=args.spacing_deep,
               chase_blocks_per_cu=args.chase_blocks, spawn=args.spawn, records=args.records)
    rs0 = ctx.record_stats()
    dt, phases = timed(ctx, lambda: ctx.rank_list(succ, w, root, rank, **cfg), args.steps, args.warmup)
    rs1 = ctx.record_stats()
    ms_per_step = dt * 1e3
    value = n / dt
    # level-0 record width of the timed calls (every one narrow, or none)
    narrow_calls = rs1["narrow_calls"] - rs0["narrow_calls"]
    record_bytes = 8 if narrow_calls > 0 else 16

    # dominant kernel and its roofline (HIP events on the library's stream)
    kernel_phases = {k: v for k, v in phases.items() if k not in ("unclaimed",)}
    dom = max(kernel_phases, key=lambda k: kernel_phases[k][0])
    dom_ms = phases[dom][0] / max(1, phases[dom][1])
    if dom.startswith("chase_ruler_L0"):
        alg_bytes = CHASE_BYTES_PER_ELEMENT * n
    elif dom.startswith("ruler_propagation_L0") or dom == "pack":
        alg_bytes = 32 * n
    else:
        alg_bytes = BYTES_PER_ELEMENT * n
    achieved = alg_bytes / (dom_ms * 1e-3) / 1e9
    whole_gbps = BYTES_PER_ELEMENT * n / (ms_per_step * 1e-3) / 1e9
    traffic_per_elem, traffic_conf = profiled_traffic(dom) if n == 1 << 27 else (None, {})
    traffic = traffic_per_elem * n if traffic_per_elem is not None else None

    check = None
    if not args.no_check:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as o
        er, ek = o.list_closed_form(n, args.seed, w.cpu().numpy())
        check = bool((root.cpu().numpy() == er).all() and (rank.cpu().numpy() == ek).all())

    out = {
        "metric": "elements ranked/s @ n=2^27",
        "value": value,
        "unit": "elements/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic: seeded random permutation list, int64 weights uniform in [1,100], tail weight 0",
        "config": {"workload": f"single random list n={n}, int64 weights, {args.algorithm} "
                               f"(uint64 successors, weighted ranks bit-exact)",
                   "n": n, "algorithm": args.algorithm, "seed": args.seed},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBPS,
                     "traffic": traffic,
                     "kernel_ms": dom_ms, "algorithmic_bytes": alg_bytes,
                     "level0_record_bytes": record_bytes},
        "whole_call_effective_gbps": whole_gbps,
        "phases_ms_per_step": {k: v[0] / args.steps for k, v in sorted(phases.items())},
        "launch": ctx.launch_stats(),
    }
    if check is not None:
        out["bit_exact_vs_closed_form"] = check
    # per-level trace of one more call (RulerTrace counterpart)
    ctx.trace(True)
    ctx.rank_list(succ, w, root, rank, **cfg)
    out["ruler_trace"] = ctx.trace_read()
    ctx.trace(False)
    if traffic is not None:
        # the profiled configuration must be the one this run used
        lvl0 = out["ruler_trace"][0]
        spawn_on = lvl0.get("spawned", 0) > 0
        blocks = args.chase_blocks or ctx.tuning_get(n)["chase_blocks_per_cu"]
        matches = (str(lvl0.get("spacing")) == traffic_conf.get("spacing")
                   and ("on" if spawn_on else "off") == traffic_conf.get("spawn")
                   and str(record_bytes) == traffic_conf.get("records")
                   and str(blocks) == traffic_conf.get("blocks"))
        out["roofline"]["traffic_source"] = CHASE_PROFILE
        out["roofline"]["traffic_config_matches"] = matches
        if not matches:
            out["roofline"]["traffic"] = None
            out["roofline"]["traffic_note"] = (f"profile measured {traffic_conf}, this run used spacing "
                                               f"{lvl0.get('spacing')} spawn {'on' if spawn_on else 'off'} "
                                               f"records {record_bytes} blocks {blocks}")
            traffic = None
    if not args.no_host_buffers:
        out["host_buffers"] = host_buffer_rates(ctx, succ, w, cfg)
    del succ, w, root, rank
    torch.cuda.empty_cache()
    # random-gather ceiling: dst[i] = src[idx[i]], idx a random permutation of
    # 2^27, credited 24 B per element (idx 8 + src 8 + dst 8)
    g_gather, g_ms = ctx.probe_random_gather(1 << 27, 5)
    out["gather_ceiling"] = {"gbps": g_gather, "ms": g_ms, "random_accesses_per_s": (1 << 27) / (g_ms * 1e-3),
                             "probe": "u64 gather through a random permutation of 2^27, 24 B/element"}
    out["record_access_ceilings"] = {
        pat: {"records_per_s": r, "ms": ms}
        for pat in ("read", "read_write", "walk_read", "walk_claim", "walk_read8", "walk_claim8")
        for r, ms in [ctx.probe_record_access(pat, 1 << 27, 3)]}
    # the chase phase against its own pattern's floor: walks of 2^19 chains
    # of 256 steps claiming every record, over records of the level-0 width
    floor = out["record_access_ceilings"]["walk_claim8" if record_bytes == 8 else "walk_claim"]
    if n == 1 << 27:
        out["roofline"]["walk_claim_floor_ms"] = floor["ms"]
        out["roofline"]["frac_of_walk_claim_floor"] = floor["ms"] / dom_ms
    out["roofline"]["frac_of_gather_ceiling"] = achieved / g_gather
    out["whole_call_frac_of_gather_ceiling"] = whole_gbps / g_gather
    out["whole_call_frac_of_hbm_peak"] = whole_gbps / HBM_PEAK_GBPS
    if traffic is not None:
        out["roofline"]["amplification"] = traffic / alg_bytes
        out["roofline"]["traffic_gbps"] = traffic / (dom_ms * 1e-3) / 1e9
    if not args.no_unit_list:
        out["unit_list_2e20"] = unit_list_bench(ctx, args.seed)
    if not args.no_serve:
        out["serve"] = serve_bench(ctx, args.serve_jobs, args.seed, args.serve_slots)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(min(args.cpu_sample, n), args.seed)
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
