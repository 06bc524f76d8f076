"""Per-size-class autotuner for the list-ranking kernels (runs on the GPU).

    python -m kascade_amd.autotune [--classes 10,12,...,27] [--out kascade_amd/tuning_table.txt]

For every size class c it generates a seeded random weighted list of 2^c
elements (at least 12288) on the device and picks, by coordinate descent over the knobs

    ruler spacing at the top level, ruler spacing below it,
    resident chase workgroups per CU, rulers per work-counter atomic,

the configuration with the lowest median time over `reps` calls.  Every
candidate's output must equal the first configuration's output bit for bit
(outputs are unique, so any difference is a bug and aborts the sweep).  The
table is written as text; contexts created afterwards load it from next to
the library.
"""
import argparse
import os
import statistics
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
KNOBS = {
    "spacing_top": [8, 16, 32, 64, 128, 256, 512],
    "spacing_deep": [8, 16, 32],
    "chase_blocks_per_cu": [1, 2, 4, 8],
    "chase_chunk": [64, 256, 1024],
}
START = {"spacing_top": 32, "spacing_deep": 16, "chase_blocks_per_cu": 4, "chase_chunk": 256}


def measure(ctx, succ, w, root, rank, cfg, reps):
    import torch
    kw = dict(ruler_spacing=cfg["spacing_top"], ruler_spacing_deep=cfg["spacing_deep"],
              chase_blocks_per_cu=cfg["chase_blocks_per_cu"], chase_chunk=cfg["chase_chunk"])
    ctx.rank_list(succ, w, root, rank, **kw)
    torch.cuda.synchronize()
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        ctx.rank_list(succ, w, root, rank, **kw)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    return statistics.median(times)


# Lists of at most 8192 elements go to the one-workgroup solver, so the small
# classes are timed where the sparse ruling set starts.
MIN_TUNED_N = 12288


def tune_class(ctx, c, reps, seed=12345):
    import torch
    n = max(1 << c, MIN_TUNED_N)
    succ, w = ctx.gen_list(n, seed + c, weighted=True)
    root = torch.empty(n, dtype=torch.uint64, device="cuda")
    rank = torch.empty(n, dtype=torch.int64, device="cuda")
    ref_root = torch.empty_like(root)
    ref_rank = torch.empty_like(rank)
    best = dict(START)
    best_t = measure(ctx, succ, w, ref_root, ref_rank, best, reps)
    for _ in range(2):
        for knob, values in KNOBS.items():
            for v in values:
                if v == best[knob]:
                    continue
                cand = dict(best, **{knob: v})
                if cand["spacing_top"] > n or cand["spacing_deep"] > n:
                    continue
                t = measure(ctx, succ, w, root, rank, cand, reps)
                if not (torch.equal(root, ref_root) and torch.equal(rank, ref_rank)):
                    raise RuntimeError(f"output differs for {cand} at n=2^{c}")
                if t < best_t:
                    best, best_t = cand, t
    return best, best_t


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--classes", default="10,12,14,16,18,20,22,24,26,27")
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--out", default=os.path.join(HERE, "tuning_table.txt"))
    args = ap.parse_args()
    sys.path.insert(0, os.path.dirname(HERE))
    import kascade_amd

    ctx = kascade_amd.Context(0)
    rows = []
    for c in [int(x) for x in args.classes.split(",")]:
        best, t = tune_class(ctx, c, args.reps)
        rows.append((c, best, t))
        print(f"2^{c}: {best} {t * 1e3:.3f} ms ({max(1 << c, MIN_TUNED_N) / t:.3e} elements/s)", flush=True)
    with open(args.out, "w") as f:
        f.write("# kascade_amd tuning table (written by kascade_amd/autotune.py on the GPU)\n")
        f.write("# log2n spacing_top spacing_deep chase_blocks_per_cu chase_chunk   # median ms\n")
        for c, b, t in rows:
            f.write(f"{c} {b['spacing_top']} {b['spacing_deep']} {b['chase_blocks_per_cu']} "
                    f"{b['chase_chunk']}   # {t * 1e3:.3f}\n")
    print(f"wrote {args.out}", flush=True)


if __name__ == "__main__":
    main()
