"""The harness executable's command line (no GPU: it exits while parsing)."""
import os
import subprocess

RUNNER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "kascade_amd", "kascade_runner")


def run(*args):
    return subprocess.run([RUNNER, *args], capture_output=True, text=True, timeout=60)


def test_help_lists_the_reference_options():
    r = run("--help")
    assert r.returncode == 0
    for opt in ("--algorithm", "--iterations", "--verify-level", "--verify-continue-on-mismatch", "--statistics",
                "--output"):
        assert opt in r.stdout


def test_bad_arguments_exit_with_an_error():
    for args in (["--algorithm", "Nope"], ["--statistics", "all"], ["--input", "grid"], ["--input", "gnm", "--input-processing", "none"], ["--n", "0"],
                 ["--bogus"], ["--n", "12x"], ["--input", "forest", "--n", "5", "--lists", "9"]):
        r = run(*args)
        assert r.returncode == 2 and "[error]" in r.stderr, args
