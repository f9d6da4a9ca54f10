"""Builds the package's native libraries in-tree.

* ``lib/libgraphalgo.so`` -- the gfx950 HIP library behind ``include/graphalgo.h``
  (every ``csrc/*.hip`` file, compiled by hipcc for ``--offload-arch=gfx950``).
* ``lib/libgagraph.so``   -- the host-only graph substrate (generators, Matrix
  Market reader), compiled by g++ without ``-march=native`` so it produces the
  same graphs on every x86-64 host.

Run ``python -m graphalgo_amd.build`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "lib"
OBJ = ROOT / "_build" / "obj"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    "-Wall", "-Wno-unused-function",
    f"-I{ROOT / 'include'}", f"-I{CSRC}",
]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:4])} ...")
    # surface real warnings but not hipcc's notice about the host pass
    for line in (proc.stderr or "").splitlines():
        if "warning" in line and "packed-fp32-ops" not in line and "feature" not in line:
            sys.stderr.write(line + "\n")


def build_host() -> Path:
    LIB.mkdir(parents=True, exist_ok=True)
    out = LIB / "libgagraph.so"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
          "-Wall", "-o", str(out) + ".tmp", str(CSRC / "host" / "graphgen.cpp"), "-lz"])
    os.replace(str(out) + ".tmp", out)
    return out


def build_hip(jobs: int = 8) -> Path:
    LIB.mkdir(parents=True, exist_ok=True)
    OBJ.mkdir(parents=True, exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip"))
    objs = [OBJ / (s.stem + ".o") for s in srcs]

    def compile_one(pair):
        src, obj = pair
        _run([HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(obj)])
        return obj

    with cf.ThreadPoolExecutor(max_workers=max(1, min(jobs, len(srcs)))) as ex:
        list(ex.map(compile_one, zip(srcs, objs)))
    out = LIB / "libgraphalgo.so"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", str(out) + ".tmp", *map(str, objs)])
    os.replace(str(out) + ".tmp", out)
    return out


def build_all() -> None:
    if not Path(HIPCC).exists() and shutil.which("hipcc") is None:
        raise RuntimeError("hipcc not found: the gfx950 library cannot be built")
    build_host()
    build_hip()


if __name__ == "__main__":
    build_all()
    print("built", LIB / "libgraphalgo.so", LIB / "libgagraph.so")
