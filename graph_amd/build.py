"""Builds the native libraries of graph_amd in-tree.

- ``graph_amd/libcosmosgl.so``: the HIP kernels and the C ABI of
  ``include/cosmosgl.h``, compiled for gfx950 only.
- ``graph_amd/libcgdata.so``: host-only synthetic input generators.
- ``graph_amd/napi/cosmosgl.node``: the Node-API addon over the C ABI.

Run ``python -m graph_amd.build`` (or ``__graft_entry__.build()``).  Objects go
to ``graph_amd/_build/``; every call recompiles.
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
BUILD = PKG / "_build"
ARCH = "gfx950"

HIP_SOURCES = ["api.hip", "k_pointwise.hip", "k_manybody.hip", "k_link.hip", "k_collision.hip"]

HIPCC_FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"--offload-arch={ARCH}",
    # no packed-fp32 instructions in any kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # no SLP vectors: with packed fp32 off they are split again, and in the
    # force kernel their splats of the query point spill past its 64 VGPRs
    "-fno-slp-vectorize",
    "-Wall",
    "-Wno-unused-function",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build graph_amd")


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")


def build_cosmosgl() -> Path:
    BUILD.mkdir(exist_ok=True)
    hipcc = _hipcc()
    objs = []
    jobs = []
    for src in HIP_SOURCES:
        obj = BUILD / (Path(src).stem + ".o")
        objs.append(obj)
        jobs.append([hipcc, *HIPCC_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
    with cf.ThreadPoolExecutor(max_workers=min(len(jobs), 8)) as ex:
        list(ex.map(_run, jobs))
    out = PKG / "libcosmosgl.so"
    tmp = PKG / "libcosmosgl.so.tmp"
    _run([hipcc, "-shared", "-fPIC", f"--offload-arch={ARCH}", *map(str, objs), "-o", str(tmp)])
    os.replace(tmp, out)
    return out


def build_datagen() -> Path:
    out = PKG / "libcgdata.so"
    tmp = PKG / "libcgdata.so.tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", str(CSRC / "datagen.cpp"), "-o", str(tmp)])
    os.replace(tmp, out)
    return out


def build_napi() -> Path | None:
    """Node-API addon (graph_amd/napi/cosmosgl.node) when the Node headers are present."""
    if not Path("/usr/include/node/node_api.h").exists():
        return None
    _run(["make", "-s", "-B", "-C", str(PKG / "napi")])
    return PKG / "napi" / "cosmosgl.node"


def build_all() -> None:
    build_datagen()
    build_cosmosgl()
    build_napi()


if __name__ == "__main__":
    build_all()
    print("built", PKG / "libcosmosgl.so", PKG / "libcgdata.so")
