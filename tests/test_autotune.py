"""The autotuner's table merge (host logic, no GPU): a run replaces only the
rows of its own kind and classes."""
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def test_kept_rows(tmp_path):
    from kascade_amd.autotune import kept_rows
    f = tmp_path / "t.txt"
    f.write_text("# header\n10 8 8 2 64   # 0.1\n20 16 8 2 64\n27 128 8 2 256\n"
                 "tree 10 8 32 2 64\ntree 20 8 8 1 64   # 0.4\n\nseg 27 256 16 2 64\nseg 30 512 8 1 64\n")
    k = kept_rows(str(f), "list", {10, 27})
    assert [c for c, _ in k["list"]] == [20] and [c for c, _ in k["tree"]] == [10, 20]
    assert [c for c, _ in k["segmented"]] == [27, 30]
    k = kept_rows(str(f), "tree", {20})
    assert [c for c, _ in k["list"]] == [10, 20, 27] and k["tree"] == [(10, "tree 10 8 32 2 64")]
    k = kept_rows(str(f), "segmented", {30})
    assert k["segmented"] == [(27, "seg 27 256 16 2 64")] and len(k["list"]) == 3 and len(k["tree"]) == 2
    assert kept_rows(str(tmp_path / "absent.txt"), "list", {10}) == {"list": [], "tree": [], "segmented": []}
