"""The harness executable on the GPU: the reference runner's loop (generate,
reference run, iterations, verify, statistics) and its JSON report shape."""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as o
from test_gpu_rank import say

pytestmark = pytest.mark.gpu
RUNNER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "kascade_amd", "kascade_runner")


def run(tmp_path, *args):
    out = tmp_path / "report.json"
    r = subprocess.run([RUNNER, *args, "--output", str(out)], capture_output=True, text=True, timeout=300)
    say(r.stderr[-2000:])
    return r, (json.loads(out.read_text()) if out.exists() else None)


def get(d, path):
    for k in path:
        d = d[k]
    return d


def test_runner_list_srs_report(tmp_path):
    say("runner list srs")
    n = 1 << 20
    r, rep = run(tmp_path, "--algorithm", "SparseRulingSet", "--input", "list", "--n", str(n), "--seed", "9",
                 "--weighted", "--iterations", "3", "--statistics", "extensive")
    assert r.returncode == 0
    assert "Root array matches reference." in r.stderr and "Rank array matches reference." in r.stderr
    assert list(rep) == ["config", "stats", "timer", "counters"]
    assert rep["config"]["algorithm"] == "SparseRulingSet" and rep["config"]["iterations"] == 3
    assert rep["config"]["verification_reference"] == "PointerDoubling"
    assert len(rep["timer"]) == 3 and len(rep["counters"]) == 3
    for t in rep["timer"]:
        assert get(t, ["root", "ranking", "max"])[0] > 0          # eval/Config.jl total_time
        assert get(t, ["root", "ranking", "chase_ruler", "max"])[0] > 0
        assert get(t, ["root", "ranking", "ruler_propagation", "max"])[0] > 0
        assert "base_case" in t["root"]["ranking"]
    assert "reference_run" in rep["timer"][0]["root"] and "list_check" in rep["timer"][0]["root"]
    # statistics equal the oracle's on the generator's closed form
    succ, w = o.gen_list(n, 9, True)
    er, ek = o.list_closed_form(n, 9, w)
    want = o.compute_stats(er, ek, "extensive")
    got = rep["stats"]["input_stats"]
    assert got["basic_stats"] == want["basic_stats"]
    e, we = got["extensive_stats"], want["extensive_stats"]
    for k in ("num_nontrivial_trees", "max_size", "avg_size", "nontrivial_avg_size", "nontrivial_avg_rank"):
        assert e[k] == we[k], k
    assert e["per_tree_stats"] == [{k: t[k] for k in ("size", "max_rank", "rank_sum")} for t in we["per_tree_stats"]]


def test_runner_tree_and_forest(tmp_path):
    say("runner tree / forest")
    r, rep = run(tmp_path, "--algorithm", "EulerTour", "--input", "tree", "--n", "300000", "--seed", "4",
                 "--statistics", "basic")
    assert r.returncode == 0 and "Rank array matches reference." in r.stderr
    er, ed = o.tree_closed_form(300000, 4)
    assert rep["stats"]["input_stats"]["basic_stats"] == o.compute_stats(er, ed, "basic")["basic_stats"]
    assert "tour_build" in rep["timer"][0]["root"]["ranking"]
    # pointer doubling under test: the reference run is the sparse ruling set
    r, rep = run(tmp_path, "--algorithm", "PointerDoubling", "--input", "forest", "--n", "200000", "--lists", "300",
                 "--weighted", "--statistics", "extensive-reduced-output")
    assert r.returncode == 0 and rep["config"]["verification_reference"] == "SparseRulingSet"
    assert rep["stats"]["input_stats"]["extensive_stats"]["num_nontrivial_trees"] == 300
    assert "per_tree_stats" in rep["stats"]["input_stats"]["extensive_stats"]
    # pointer doubling on a tree is checked against the Euler tour
    r, rep = run(tmp_path, "--algorithm", "PointerDoubling", "--input", "tree", "--n", "50000")
    assert r.returncode == 0 and rep["config"]["verification_reference"] == "EulerTour"


def test_runner_verify_level_zero_skips_reference(tmp_path):
    say("runner verify 0")
    r, rep = run(tmp_path, "--n", "100000", "--verify-level", "0", "--iterations", "2")
    assert r.returncode == 0 and "matches reference" not in r.stderr
    assert "reference_run" not in rep["timer"][0]["root"] and rep["stats"]["input_stats"] == {}


def test_runner_output_extension_and_errors(tmp_path):
    say("runner output")
    out = tmp_path / "result"
    r = subprocess.run([RUNNER, "--n", "5000", "--output", str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and (tmp_path / "result.json").exists()
    # a tree is not a list: the sparse ruling set rejects it
    r = subprocess.run([RUNNER, "--input", "tree", "--n", "5000", "--algorithm", "SparseRulingSet",
                        "--validate-list"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "[error]" in r.stderr


def test_runner_graph_inputs(tmp_path):
    """--input gnm|rmat with "bfs" processing (a BFS forest, ranked by the
    Euler tour) and "eulertour" processing (the forest's tour, ranked as a
    list), verified against the independent reference run."""
    say("runner graph inputs")
    r, rep = run(tmp_path, "--algorithm", "EulerTour", "--input", "rmat", "--n", "65536", "--m", "400000",
                 "--input-processing", "bfs", "--statistics", "basic")
    assert r.returncode == 0 and "Rank array matches reference." in r.stderr
    assert rep["config"]["input"]["input_processing"] == "bfs"
    src, dst = o.gen_graph("rmat", 65536, 400000, 1)
    xadj, adj = o.csr(src, dst, 65536)
    parent = o.bfs_forest(xadj, adj)
    wts = (parent != np.arange(65536, dtype=np.uint64)).astype(np.int64)
    er, ed = o.sequential_walk(parent, wts)
    assert rep["stats"]["input_stats"]["basic_stats"] == o.compute_stats(er, ed, "basic")["basic_stats"]
    r, rep = run(tmp_path, "--algorithm", "SparseRulingSet", "--input", "gnm", "--n", "100000", "--m", "150000",
                 "--input-processing", "eulertour")
    assert r.returncode == 0 and "Rank array matches reference." in r.stderr
    # the reference's tour: 2 (n - roots) + roots elements
    src, dst = o.gen_graph("gnm", 100000, 150000, 1)
    xadj, adj = o.csr(src, dst, 100000)
    parent = o.bfs_forest(xadj, adj)
    roots = int(np.count_nonzero(parent == np.arange(100000, dtype=np.uint64)))
    assert rep["stats"]["input_distribution"]["total_size"] == 2 * 100000 - roots
    # the submission suites' rgg2d input with eulertour processing
    r, rep = run(tmp_path, "--algorithm", "SparseRulingSet", "--input", "rgg2d", "--n", "65536", "--m", "262144",
                 "--input-processing", "eulertour", "--statistics", "basic")
    assert r.returncode == 0 and "Rank array matches reference." in r.stderr
    src, dst = o.gen_rgg2d(65536, 262144, 1)
    xadj, adj = o.csr(src, dst, 65536)
    parent = o.bfs_forest(xadj, adj)
    roots = int(np.count_nonzero(parent == np.arange(65536, dtype=np.uint64)))
    assert rep["stats"]["input_distribution"]["total_size"] == 2 * 65536 - roots
