"""Per-size-class autotuner for the list-ranking kernels (runs on the GPU).

    python -m kascade_amd.autotune [--kind list|tree|segmented] [--classes 10,12,...,27]
                                   [--out kascade_amd/tuning_table.txt]

For every size class c it generates a seeded random weighted list of 2^c
elements (at least 12288) on the device, with --kind tree a random rooted
tree of 2^c vertices (at least 6144; its Euler tour has twice as many
records), or with --kind segmented a batch of seeded random lists of
2^8..2^20 elements each (log-uniform, 30 % locality-biased, unit weights)
with about 2^c elements in all, and picks, by coordinate descent over the
knobs

    ruler spacing at the top level, ruler spacing below it,
    resident chase workgroups per CU, rulers per work-counter atomic,

starting from the knobs the context currently uses for that size, the
configuration with the lowest median time over `reps` calls.  Every
candidate's output must equal the first configuration's output bit for bit
(outputs are unique, so any difference is a bug and aborts the sweep).  The
table is written as text (tree rows prefixed with "tree", segmented-batch
rows with "seg"; rows already in the file for the other kinds or for classes
not tuned in this run are kept); contexts created afterwards load it from
next to the library.
"""
import argparse
import os
import statistics
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
KNOBS = {
    "spacing_top": [4, 8, 16, 32, 64, 128, 256, 512],
    "spacing_deep": [2, 4, 8, 16, 32],
    "chase_blocks_per_cu": [1, 2, 3, 4, 8],
    "chase_chunk": [64, 256, 1024],
}
START = {"spacing_top": 32, "spacing_deep": 16, "chase_blocks_per_cu": 4, "chase_chunk": 256}


def measure(run, root, rank, cfg, reps):
    import torch
    kw = dict(ruler_spacing=cfg["spacing_top"], ruler_spacing_deep=cfg["spacing_deep"],
              chase_blocks_per_cu=cfg["chase_blocks_per_cu"], chase_chunk=cfg["chase_chunk"])
    run(root, rank, kw)
    torch.cuda.synchronize()
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        run(root, rank, kw)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    return statistics.median(times)


# Rankings of at most 8192 records go to the one-workgroup solver, so the
# small classes are timed where the sparse ruling set starts (a tree of n
# vertices has a tour of 2n records).
MIN_TUNED_N = {"list": 12288, "tree": 6144, "segmented": 1 << 22}
PREFIX = {"list": "", "tree": "tree ", "segmented": "seg "}


def segmented_batch(ctx, total, seed):
    """Offsets of lists of 2^8..2^20 elements (log-uniform, seeded) adding up
    to at least `total`, 30 % of them locality-biased, unit weights (the
    composition of the 20,000-list batch benchmark), and the batch's
    successors."""
    import numpy as np
    import torch
    rng = np.random.default_rng(seed)
    sizes = []
    while sum(sizes) < total:
        sizes.append(int(2.0 ** rng.uniform(8, 20)))
    off = np.zeros(len(sizes) + 1, np.uint64)
    off[1:] = np.cumsum(np.array(sizes, np.uint64))
    seeds = rng.integers(1, 1 << 62, size=len(sizes), dtype=np.uint64)
    local = (rng.random(len(sizes)) < 0.3).astype(np.uint8)
    d_off = torch.from_numpy(off).to("cuda")
    succ, _ = ctx.gen_lists_segmented(d_off, torch.from_numpy(seeds).to("cuda"), torch.from_numpy(local).to("cuda"),
                                      weighted=False)
    return d_off, succ, None


def tune_class(ctx, c, reps, kind="list", seed=12345):
    import torch
    n = max(1 << c, MIN_TUNED_N[kind])
    root_dtype = torch.uint64
    if kind == "list":
        succ, w = ctx.gen_list(n, seed + c, weighted=True)
        run = lambda root, rank, kw: ctx.rank_list(succ, w, root, rank, **kw)  # noqa: E731
    elif kind == "tree":
        parent = ctx.gen_tree(n, seed + c)
        run = lambda root, rank, kw: ctx.root_tree(parent, root, rank, **kw)  # noqa: E731
    else:
        off, succ, w = segmented_batch(ctx, n, seed + c)
        n = succ.numel()
        root_dtype = torch.uint32
        run = lambda root, rank, kw: ctx.rank_lists_segmented(succ, off, w, root, rank, **kw)  # noqa: E731
    root = torch.empty(n, dtype=root_dtype, device="cuda")
    rank = torch.empty(n, dtype=torch.int64, device="cuda")
    ref_root = torch.empty_like(root)
    ref_rank = torch.empty_like(rank)
    # start from the context's current row for this size (coordinate descent
    # finds a local optimum: from a fixed start it can miss a joint optimum
    # of spacing and workgroups that the current row already holds)
    cur = ctx.tuning_get(n, kind=kind)
    best = {k: cur[k] for k in START} if cur["spacing_top"] else dict(START)
    best_t = measure(run, ref_root, ref_rank, best, reps)
    for _ in range(2):
        for knob, values in KNOBS.items():
            for v in values:
                if v == best[knob]:
                    continue
                cand = dict(best, **{knob: v})
                if cand["spacing_top"] > n or cand["spacing_deep"] > n:
                    continue
                t = measure(run, root, rank, cand, reps)
                if not (torch.equal(root, ref_root) and torch.equal(rank, ref_rank)):
                    raise RuntimeError(f"output differs for {cand} at n=2^{c}")
                if t < best_t:
                    best, best_t = cand, t
    return best, best_t


def kept_rows(path, kind, classes):
    """Data lines already in the table file that this run does not replace:
    {kind: [(class, line)]} for the list, tree and segmented rows."""
    keep = {k: [] for k in PREFIX}
    if not os.path.exists(path):
        return keep
    with open(path) as f:
        for line in f:
            words = line.split("#", 1)[0].split()
            if not words:
                continue
            k = {"tree": "tree", "seg": "segmented"}.get(words[0], "list")
            c = int(words[1] if k != "list" else words[0])
            if k == kind and c in classes:
                continue
            keep[k].append((c, line.rstrip("\n")))
    return keep


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--kind", choices=["list", "tree", "segmented"], default="list")
    ap.add_argument("--classes", default="")
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--out", default=os.path.join(HERE, "tuning_table.txt"))
    args = ap.parse_args()
    classes = args.classes or {"list": "10,12,14,16,18,20,22,24,26,27", "tree": "10,12,14,16,18,20,22,24,26",
                               "segmented": "24,26,28,30,31"}[args.kind]
    sys.path.insert(0, os.path.dirname(HERE))
    import kascade_amd

    ctx = kascade_amd.Context(0)
    rows = []
    unit = "vertices" if args.kind == "tree" else "elements"
    for c in [int(x) for x in classes.split(",")]:
        best, t = tune_class(ctx, c, args.reps, args.kind)
        rows.append((c, best, t))
        n = max(1 << c, MIN_TUNED_N[args.kind])
        print(f"{args.kind} 2^{c}: {best} {t * 1e3:.3f} ms ({n / t:.3e} {unit}/s)", flush=True)
    keep = kept_rows(args.out, args.kind, {c for c, _, _ in rows})
    keep[args.kind].extend((c, f"{PREFIX[args.kind]}{c} {b['spacing_top']} {b['spacing_deep']} "
                               f"{b['chase_blocks_per_cu']} {b['chase_chunk']}   # {t * 1e3:.3f}") for c, b, t in rows)
    with open(args.out, "w") as f:
        f.write("# kascade_amd tuning table (written by kascade_amd/autotune.py on the GPU)\n")
        f.write("# [tree|seg] log2n spacing_top spacing_deep chase_blocks_per_cu chase_chunk   # median ms\n")
        for k in PREFIX:
            for _, line in sorted(keep[k]):
                f.write(line + "\n")
    print(f"wrote {args.out}", flush=True)


if __name__ == "__main__":
    main()
