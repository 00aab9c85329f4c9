"""Build the HIP extension libneuralpom_amd.so for gfx950 (in-tree).

    python -m neuralpom_amd.build            # full rebuild
    python -m neuralpom_amd.build --incremental

Every .hip file under csrc/ is compiled to an object with hipcc (in parallel),
then linked into neuralpom_amd/libneuralpom_amd.so.  The adjoint gathers are
instantiated one operator per compilation unit so the build parallelises.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libneuralpom_amd.so"
INCLUDE = PKG.parent / "include"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
CFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # packed fp32 instructions are disabled for correctness on this target
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}", "-Wno-unused-command-line-argument",
]


def _sources():
    return sorted(CSRC.glob("*.hip"))


def _headers():
    return sorted(CSRC.glob("*.h")) + sorted(INCLUDE.glob("*.h"))


def _compile(src: Path, incremental: bool) -> Path:
    obj = BUILD / (src.stem + ".o")
    if incremental and obj.exists():
        newest = max(p.stat().st_mtime for p in [src] + _headers())
        if obj.stat().st_mtime >= newest:
            return obj
    cmd = [HIPCC, *CFLAGS, "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        err = "\n".join(l for l in res.stderr.splitlines() if "packed-fp32" not in l)
        raise RuntimeError(f"hipcc failed on {src.name}:\n{err[-6000:]}")
    return obj


def build(incremental: bool = False, jobs: int | None = None) -> Path:
    BUILD.mkdir(exist_ok=True)
    srcs = _sources()
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, incremental), srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stderr[-4000:]}")
    os.replace(tmp, LIB)
    return LIB


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--incremental", action="store_true")
    ap.add_argument("-j", "--jobs", type=int, default=None)
    a = ap.parse_args(argv)
    print(build(a.incremental, a.jobs))


if __name__ == "__main__":
    sys.exit(main())
