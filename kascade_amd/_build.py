"""Builds libkascade_amd.so in-tree with hipcc for gfx950.

Usage: python -m kascade_amd._build   (or kascade_amd._build.build())
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_obj")
LIB = os.path.join(HERE, "libkascade_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("KASCADE_AMD_ARCH", "gfx950")

CFLAGS = [
    "-O3", "-fPIC", "-std=c++20", f"--offload-arch={ARCH}",
    # no packed-FP32 instructions in any device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
    "-I", os.path.join(HERE, "..", "include"),
]

SOURCES = ["api.cpp", "rank.hip", "gen.hip", "util.hip", "euler.hip", "segmented.hip", "batch.hip",
           "server.cpp", "probe.hip", "listutil.hip", "graph.hip", "noalloc.hip", "contract.hip"]


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(OBJ, src + ".o")
    lang = ["-x", "hip"]
    cmd = [HIPCC, *lang, *CFLAGS, "-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    _build_runner()
    if verbose:
        print(f"built {LIB}")
    return LIB


RUNNER = os.path.join(HERE, "kascade_runner")


def _git_tag():
    try:
        sha = subprocess.run(["git", "-C", HERE, "rev-parse", "--short=8", "HEAD"], capture_output=True,
                             text=True, check=True).stdout.strip()
        dirty = subprocess.run(["git", "-C", HERE, "status", "--porcelain"], capture_output=True,
                               text=True).stdout.strip()
        return sha + ("-dirty" if dirty else "")
    except (OSError, subprocess.CalledProcessError):
        return "unknown"


def _build_runner():
    """The benchmark harness executable (host code over the C ABI)."""
    tmp = RUNNER + ".tmp"
    cmd = [HIPCC, "-O2", "-std=c++20", "-Wall", "-I", os.path.join(HERE, "..", "include"),
           f'-DKC_GIT_TAG="{_git_tag()}"', os.path.join(CSRC, "runner.cpp"), "-o", tmp,
           "-L", HERE, "-lkascade_amd", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"runner build failed\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, RUNNER)


if __name__ == "__main__":
    try:
        build()
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
