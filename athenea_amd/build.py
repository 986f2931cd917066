"""Builds the rasterizer's HIP library, libathx.so, in this folder.

Every HIP source under csrc/ is compiled for gfx950 by hipcc and linked into
one shared library whose exported symbols are the C ABI of include/athx.h.
The flags that matter for results:

  -ffp-contract=off   a*b+c in the source stays a multiply and an add, so the
                      projection's keys and the blend's decisions are the same
                      bits a host restatement computes
  -fhip-fp32-correctly-rounded-divide-sqrt   IEEE division and square root
  -Xclang -target-feature -Xclang -packed-fp32-ops   no packed FP32 math
"""
from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
INCLUDE = HERE.parent / "include"
LIB = HERE / "libathx.so"
BUILD = HERE / "_build"

SOURCES = ["project.hip", "scan.hip", "sort.hip", "blend.hip", "project_backward.hip", "ingest.hip", "reference.hip",
           "sky.hip", "ply_header.hip", "capi.hip"]

HIPFLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    "--offload-arch=gfx950",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-fvisibility=hidden",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the rasterizer needs ROCm's hipcc to build")


def _compile(src: Path, obj: Path) -> None:
    cmd = [hipcc(), *HIPFLAGS, "-I", str(INCLUDE), "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src.name}:\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")


def build(verbose: bool = False, jobs: int | None = None) -> Path:
    """Compiles every source and links libathx.so; returns its path."""
    BUILD.mkdir(exist_ok=True)
    objs = []
    jobs = jobs or min(len(SOURCES), os.cpu_count() or 4, 8)
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as pool:
        futures = []
        for name in SOURCES:
            obj = BUILD / (Path(name).stem + ".o")
            objs.append(obj)
            futures.append(pool.submit(_compile, CSRC / name, obj))
        for f in futures:
            f.result()
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), "-shared", "-fPIC", "--offload-arch=gfx950", *map(str, objs), "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
