"""Builds libsgpar_mi.so (HIP kernels for gfx950 + the C ABI) in-tree.

    python -m ipdps_2021_graph_coarsening_amd.build

hipcc cross-compiles for gfx950 without a GPU present.  Every HIP source is
compiled with packed-FP32 instruction selection turned off.  The sources are
compiled to objects in parallel (one hipcc per file, in a temporary
directory) and then linked into the shared library.
"""
from __future__ import annotations

import glob
import os
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libsgpar_mi.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = ["-x", "hip", f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument"]
# the host pass does not know the packed-fp32 feature and says so; keep only
# real diagnostics
NOISE = ("packed-fp32-ops", "not a recognized feature", "warning generated", "warnings generated")


def sources() -> list[str]:
    return sorted(glob.glob(os.path.join(CSRC, "*.hip"))) + sorted(glob.glob(os.path.join(CSRC, "*.cpp")))


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    msgs = [l for l in (r.stderr or "").splitlines() if l.strip() and not any(k in l for k in NOISE)]
    if r.returncode != 0:
        sys.stderr.write("\n".join(msgs) + "\n")
        raise RuntimeError(f"hipcc failed with status {r.returncode}: {' '.join(cmd[-3:])}")
    if verbose and msgs:
        sys.stderr.write("\n".join(msgs) + "\n")


def build(verbose: bool = False, jobs: int | None = None) -> str:
    """Compile every source into LIB; returns its path."""
    inc = ["-I", os.path.join(os.path.dirname(PKG), "include")]
    tmpdir = tempfile.mkdtemp(prefix="sgpar_mi_build_")
    try:
        srcs = sources()
        objs = [os.path.join(tmpdir, os.path.basename(s) + ".o") for s in srcs]
        jobs = jobs or min(len(srcs), max(1, min(8, os.cpu_count() or 1)))
        with ThreadPoolExecutor(jobs) as ex:
            list(ex.map(lambda so: _run([HIPCC, *FLAGS, *inc, "-c", so[0], "-o", so[1]], verbose),
                        zip(srcs, objs)))
        tmp = LIB + ".tmp"
        _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp], verbose)
        os.replace(tmp, LIB)
    finally:
        shutil.rmtree(tmpdir, ignore_errors=True)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
