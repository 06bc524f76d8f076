"""Continuous-batching server: every job's result is identical alone and in the
mix (bit-exact), in both coalesced and one-job-per-launch modes, with device
or host arrays; budget admission and per-job error isolation."""
import time

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


def make_jobs(count, seed, lo=6, hi=16):
    rng = np.random.default_rng(seed)
    jobs = []
    for k in range(count):
        n = int(2 ** rng.uniform(lo, hi))
        if rng.random() < 0.6:
            succ, w = o.gen_list(n, 1000 + k, True)
            jobs.append(("list", succ, w))
        else:
            jobs.append(("tree", o.gen_tree(n, 2000 + k), None))
    return jobs


def alone(ctx, kind, succ, w):
    if kind == "list":
        r, k = ctx.rank_list(dev(succ), dev(w))
    else:
        r, k = ctx.root_tree(dev(succ))
    return host(r), host(k)


@pytest.mark.parametrize("mode", ["batched", "batched_sorted_tours", "one_job_per_launch"])
def test_server_mix_matches_alone(ctx, mode):
    import torch
    import kascade_amd
    say("server", mode)
    jobs = make_jobs(120, 5)
    sv = kascade_amd.Server(ctx, one_job_per_launch=(mode == "one_job_per_launch"), num_slots=3,
                            tour_build="sorted" if mode == "batched_sorted_tours" else "auto")
    # every job's arrays first, then the submissions in one burst: jobs that
    # arrive while earlier batches run are coalesced
    arrays = []
    for kind, succ, w in jobs:
        s = dev(succ)
        ww = dev(w) if w is not None else None
        ro = torch.empty(len(succ), dtype=torch.uint64, device="cuda")
        ra = torch.empty(len(succ), dtype=torch.int64, device="cuda")
        arrays.append((s, ww, ro, ra))
    torch.cuda.synchronize()
    tickets = []
    for (kind, _, _), (s, ww, ro, ra) in zip(jobs, arrays):
        t, _, _ = sv.submit(kind, s, ww, ro, ra)
        tickets.append((t, s, ww, ro, ra))
    for t, *_ in tickets:
        sv.wait(t)
    st = sv.stats()
    assert st["jobs_completed"] == len(jobs)
    if mode != "one_job_per_launch":
        assert st["batches_launched"] < len(jobs)
    for (kind, succ, w), (t, s, ww, ro, ra) in zip(jobs, tickets):
        er, ek = alone(ctx, kind, succ, w)
        assert (host(ro) == er).all() and (host(ra) == ek).all()
        if kind == "list":
            cr, ck = o.sequential_walk(succ, w)
        else:
            cr, ck = o.euler_tour_rooting(succ)
        assert (er == cr).all() and (ek == ck).all()
    sv.close()


def test_server_host_arrays_and_budget(ctx):
    import kascade_amd
    say("server host/budget")
    jobs = make_jobs(40, 8, 4, 13)
    # a small budget forces admission to wait for memory to free up
    sv = kascade_amd.Server(ctx, mem_budget_bytes=8 << 20, num_slots=2)
    outs = []
    for kind, succ, w in jobs:
        t, ro, ra = sv.submit(kind, succ.copy(), w.copy() if w is not None else None)
        outs.append((t, ro, ra))
    for (kind, succ, w), (t, ro, ra) in zip(jobs, outs):
        sv.wait(t)
        cr, ck = (o.sequential_walk(succ, w) if kind == "list" else o.euler_tour_rooting(succ))
        assert (ro == cr).all() and (ra == ck).all()
    with pytest.raises(kascade_amd.KascadeError) as e:
        sv.submit("list", np.zeros(1 << 20, np.uint64))
    assert e.value.status == kascade_amd.KC_EBUDGET
    sv.close()


def test_server_error_isolation(ctx):
    import kascade_amd
    say("server errors")
    sv = kascade_amd.Server(ctx, num_slots=1, max_batch_jobs=64)
    good = make_jobs(10, 3, 4, 10)
    subs = []
    for i, (kind, succ, w) in enumerate(good):
        subs.append(sv.submit(kind, succ.copy(), w.copy() if w is not None else None))
        if i == 4:
            bad = sv.submit("list", np.array([1, 9, 2], np.uint64))
    for (kind, succ, w), (t, ro, ra) in zip(good, subs):
        sv.wait(t)
        cr, ck = (o.sequential_walk(succ, w) if kind == "list" else o.euler_tour_rooting(succ))
        assert (ro == cr).all() and (ra == ck).all()
    with pytest.raises(kascade_amd.KascadeError) as e:
        sv.wait(bad[0])
    assert e.value.status == kascade_amd.KC_EINVAL
    sv.close()


def test_server_cyclic_tree_isolated(ctx):
    """A rootless 'tree' inside a coalesced batch fails alone with KC_ECYCLE;
    the trees and lists batched with it are still exact."""
    import kascade_amd
    say("server cyclic tree")
    sv = kascade_amd.Server(ctx, num_slots=1, max_batch_jobs=64)
    good = make_jobs(12, 11, 6, 12)
    subs = []
    for i, (kind, succ, w) in enumerate(good):
        subs.append(sv.submit(kind, succ.copy(), w.copy() if w is not None else None))
        if i == 5:
            cyc = np.roll(np.arange(700, dtype=np.uint64), 1)
            bad = sv.submit("tree", cyc)
    for (kind, succ, w), (t, ro, ra) in zip(good, subs):
        sv.wait(t)
        cr, ck = (o.sequential_walk(succ, w) if kind == "list" else o.euler_tour_rooting(succ))
        assert (ro == cr).all() and (ra == ck).all()
    with pytest.raises(kascade_amd.KascadeError) as e:
        sv.wait(bad[0])
    assert e.value.status == kascade_amd.KC_ECYCLE
    sv.close()


def test_scratch_reuse_across_contexts(ctx):
    """A steady workload runs from one scratch chunk: after the first calls
    the pool stops growing; contexts and servers that come and go reuse the
    device cache, and every result stays exact."""
    import torch
    import kascade_amd
    say("scratch reuse")
    jobs = make_jobs(30, 21, 8, 17)
    held = []
    for rep in range(3):
        c = kascade_amd.Context(0)
        for _ in range(3):
            for kind, succ, w in jobs[:6]:
                r, k = (c.rank_list(dev(succ), dev(w)) if kind == "list" else c.root_tree(dev(succ)))
                cr, ck = (o.sequential_walk(succ, w) if kind == "list" else o.euler_tour_rooting(succ))
                assert (host(r) == cr).all() and (host(k) == ck).all()
            held.append(c.scratch_bytes())
        c.close()
        sv = kascade_amd.Server(ctx, num_slots=2)
        subs = [sv.submit(kind, dev(succ), dev(w) if w is not None else None,
                          torch.empty(len(succ), dtype=torch.uint64, device="cuda"),
                          torch.empty(len(succ), dtype=torch.int64, device="cuda")) for kind, succ, w in jobs]
        for (kind, succ, w), (t, ro, ra) in zip(jobs, subs):
            sv.wait(t)
            cr, ck = (o.sequential_walk(succ, w) if kind == "list" else o.euler_tour_rooting(succ))
            assert (host(ro) == cr).all() and (host(ra) == ck).all()
        sv.close()
    # within one context the pool is steady after its first pass over the jobs
    assert held[1] == held[2] and held[4] == held[5] and held[7] == held[8], held
    kascade_amd.release_cached_scratch(0)


def test_server_concurrent_submitters(ctx):
    """Four host threads submit and wait at once (the server is shared, as a
    serving front end would share it); every result matches the oracle."""
    import threading
    import torch
    import kascade_amd
    say("server concurrent submitters")
    jobs = make_jobs(80, 21)
    # device arrays are made here: the submitting threads only submit and wait
    arrays = []
    for kind, succ, w in jobs:
        arrays.append((dev(succ), dev(w) if w is not None else None,
                       torch.empty(len(succ), dtype=torch.uint64, device="cuda"),
                       torch.empty(len(succ), dtype=torch.int64, device="cuda")))
    torch.cuda.synchronize()
    sv = kascade_amd.Server(ctx, num_slots=3)
    errors = []

    def submitter(part):
        try:
            mine = []
            for k in range(part, len(jobs), 4):
                s, ww, ro, ra = arrays[k]
                mine.append(sv.submit(jobs[k][0], s, ww, ro, ra)[0])
            for t in mine:
                sv.wait(t)
        except Exception as e:  # reported on the main thread
            errors.append(e)

    threads = [threading.Thread(target=submitter, args=(p,)) for p in range(4)]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=120)
    assert not errors and not any(th.is_alive() for th in threads)
    assert sv.stats()["jobs_completed"] == len(jobs)
    for (kind, succ, w), (_, _, ro, ra) in zip(jobs, arrays):
        cr, ck = o.sequential_walk(succ, w) if kind == "list" else o.euler_tour_rooting(succ)
        assert (host(ro) == cr).all() and (host(ra) == ck).all()
    sv.close()


@pytest.mark.parametrize("weights", ["small", "distance_overflow", "beyond_int32"])
def test_server_narrow_batches(ctx, weights):
    """Batches of 2^22 records and more pack 8-byte records (ids, weights and
    claim distances in 32 bits).  A weight outside int32 in any list job, or
    claim distances that outgrow int32, must send the batch to 16-byte
    records with every result unchanged."""
    import torch
    import kascade_amd
    say("server narrow batches", weights)
    rng = np.random.default_rng(17)
    jobs = []
    for k in range(14):
        if k % 3 == 2:
            n = int(2 ** rng.uniform(18, 19.5))
            jobs.append(("tree", o.gen_tree(n, 3000 + k), None))
            continue
        n = int(2 ** rng.uniform(19, 20.5))
        succ, w = o.gen_list(n, 4000 + k, True)
        if weights == "distance_overflow":
            w = np.where(w != 0, np.int64(1 << 29), np.int64(0))
        elif weights == "beyond_int32" and k == 4:
            w = np.where(w != 0, np.int64(1 << 33), np.int64(0))
        jobs.append(("list", succ, w))
    sv = kascade_amd.Server(ctx, num_slots=1)
    arrays = []
    for kind, succ, w in jobs:
        s = dev(succ)
        ww = dev(w) if w is not None else None
        ro = torch.empty(len(succ), dtype=torch.uint64, device="cuda")
        ra = torch.empty(len(succ), dtype=torch.int64, device="cuda")
        arrays.append((s, ww, ro, ra))
    torch.cuda.synchronize()
    tickets = [sv.submit(kind, s, ww, ro, ra)[0] for (kind, _, _), (s, ww, ro, ra) in zip(jobs, arrays)]
    for t in tickets:
        sv.wait(t)
    st = sv.stats()
    assert st["jobs_completed"] == len(jobs)
    assert st["batches_launched"] < len(jobs)
    for (kind, succ, w), (s, ww, ro, ra) in zip(jobs, arrays):
        if kind == "list":
            cr, ck = o.sequential_walk(succ, w)
        else:
            cr, ck = o.euler_tour_rooting(succ)
        assert (host(ro) == cr).all() and (host(ra) == ck).all(), kind
    sv.close()
