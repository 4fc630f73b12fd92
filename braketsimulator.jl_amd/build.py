"""Build libbsim.so (HIP, gfx950) in-tree.

    python braketsimulator.jl_amd/build.py [--jobs N]

Every ``csrc/*.hip`` file is compiled by hipcc for gfx950 only and linked
into ``braketsimulator.jl_amd/lib/libbsim.so``.  Packed-FP32 instructions
are disabled for the device code.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB_DIR = PKG / "lib"
OBJ_DIR = PKG / "_build"
LIB = LIB_DIR / "libbsim.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", str(ROOT / "include"),
    "-I", str(CSRC),
    "-Wall",
    "-Wno-unused-result",
    "-Wno-unused-function",
    "-Wno-unused-value",
]


# Files with the fused tile kernel: its op loop dispatches on uniform
# (per-op) values with the 16-amplitude sub-cube live across every case.
# Structurizing those uniform branches makes the compiler keep two copies of
# the sub-cube and copy it back on every op; leaving uniform regions as plain
# scalar branches keeps it in place (about 10% faster passes).
TILE_FILES = {"fused.hip", "batch.hip", "adj_tile.hip"}
TILE_FLAGS = ["-mllvm", "-structurizecfg-skip-uniform-regions=true"]


def _compile(src: Path) -> Path:
    obj = OBJ_DIR / (src.stem + ".o")
    extra = TILE_FLAGS if src.name in TILE_FILES else []
    cmd = [HIPCC, *FLAGS, *extra, "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{r.stderr}")
    if r.stderr.strip():
        # keep warnings visible but short
        lines = [l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip()]
        if lines:
            sys.stderr.write("\n".join(lines[:40]) + "\n")
    return obj


def build(jobs: int | None = None, verbose: bool = True) -> Path:
    OBJ_DIR.mkdir(exist_ok=True)
    LIB_DIR.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip"))
    if not srcs:
        raise RuntimeError("no HIP sources found")
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources")
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    a = ap.parse_args()
    build(a.jobs)
