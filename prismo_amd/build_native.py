"""Compile libpmfdtd.so (gfx950) in-tree with hipcc.

Every translation unit is compiled with packed-FP32 instruction selection
disabled; the reference-arithmetic kernels (compat.hip) are additionally
compiled with -ffp-contract=off so no multiply-add is fused.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
LIB = HERE / "libpmfdtd.so"
OBJDIR = HERE / "_build"

ARCH = os.environ.get("PM_OFFLOAD_ARCH", "gfx950")
COMMON = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-Wall",
    "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{HERE.parent / 'include'}",
]
# translation unit -> extra flags
UNITS = {
    "api.hip": [],
    "yee.hip": [],
    "compat.hip": ["-ffp-contract=off"],
    "aux.hip": [],
    "dft_mfma.hip": [],
    "batch2d.hip": [],
    "mlp.hip": [],
    "geom.hip": ["-ffp-contract=off"],
    "fused_f32.hip": [],
    "fused_f64.hip": [],
    "tm2d.hip": [],
}


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = True, jobs: int = 4) -> Path:
    OBJDIR.mkdir(exist_ok=True)
    cc = hipcc()
    procs = []
    objs = []
    for unit, extra in UNITS.items():
        obj = OBJDIR / (unit + ".o")
        objs.append(obj)
        cmd = [cc, *COMMON, *extra, "-c", str(CSRC / unit), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((unit, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        if len([p for _, p in procs if p.poll() is None]) >= jobs:
            procs[0][1].wait()
    failed = []
    for unit, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        if p.returncode != 0:
            failed.append(unit)
            print(text, file=sys.stderr)
        elif verbose and text.strip():
            print(text)
    if failed:
        raise RuntimeError(f"hipcc failed for {failed}")
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
