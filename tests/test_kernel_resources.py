"""Resource guard for the fused tile kernels (CPU: compiles the device code
to assembly with the library's own flags).  The sub-cube of a tile sweep
must stay in VGPRs: when the compiler puts it in scratch memory instead (a
private segment), the streaming passes run several times slower with the
same results, which no parity test notices."""
from __future__ import annotations

import os
import re
import shutil
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "braketsimulator.jl_amd"


def _metadata(src: str) -> dict[str, dict[str, int]]:
    sys.path.insert(0, str(PKG))
    import build as B  # noqa: E402

    hipcc = B.HIPCC if os.path.exists(B.HIPCC) else shutil.which("hipcc")
    if not hipcc:
        pytest.skip("hipcc not available")
    out = ROOT / "tests" / "native" / "_build" / (Path(src).stem + ".s")
    out.parent.mkdir(parents=True, exist_ok=True)
    cmd = [hipcc, *B.FLAGS, *B.TILE_FLAGS, "--cuda-device-only", "-S", str(PKG / "csrc" / src), "-o", str(out)]
    subprocess.run(cmd, check=True, capture_output=True)
    meta, cur = {}, None
    for line in out.read_text().splitlines():
        m = re.match(r"\s+\.name:\s+(\S+)", line)
        if m:
            cur = m.group(1)
            meta[cur] = {}
            continue
        m = re.match(r"\s+\.(private_segment_fixed_size|vgpr_count|vgpr_spill_count):\s+(\d+)", line)
        if m and cur:
            meta[cur][m.group(1)] = int(m.group(2))
    return meta


def test_tile_kernels_keep_subcube_in_registers():
    meta = _metadata("fused.hip")
    tile = {k: v for k, v in meta.items() if "k_tile" in k}
    assert tile, "no tile kernels found in the assembly"
    for name, m in tile.items():
        assert m.get("private_segment_fixed_size", 0) == 0, f"{name} uses scratch memory: {m}"
        assert m.get("vgpr_spill_count", 0) == 0, f"{name} spills VGPRs: {m}"
