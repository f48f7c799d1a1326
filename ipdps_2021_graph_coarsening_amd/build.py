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
    """Runs a compiler; raises with its diagnostics on failure."""
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    msgs = [l for l in (r.stderr or "").splitlines() if l.strip() and not any(k in l for k in NOISE)]
    if r.returncode != 0:
        sys.stderr.write("\n".join(msgs) + "\n")
        raise RuntimeError(f"{os.path.basename(cmd[0])} failed with status {r.returncode}: {' '.join(cmd[-3:])}")
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


DROPIN = os.path.join(PKG, "dropin")
DROPIN_LIB = os.path.join(PKG, "libsgpar_dropin.so")
DROPIN_EXE = os.path.join(DROPIN, "sgpar_dropin")
# the same compiled with -DFM, as the reference's *_fm binaries are, and with
# -DHEC_V3 (the hec3_* binaries); -DMIS, -DGOSH, -DHEC_V2, -DGOSH_V2, -DHASHMAP, -DPUREMATCH, -DMTMETIS and
# -DHEC_SERIAL likewise
DROPIN_FM_LIB = os.path.join(PKG, "libsgpar_dropin_fm.so")
DROPIN_FM_EXE = os.path.join(DROPIN, "sgpar_dropin_fm")
DROPIN_HEC3_LIB = os.path.join(PKG, "libsgpar_dropin_hec3.so")
DROPIN_HEC3_EXE = os.path.join(DROPIN, "sgpar_dropin_hec3")
DROPIN_HEC3_FM_LIB = os.path.join(PKG, "libsgpar_dropin_hec3_fm.so")
DROPIN_HEC3_FM_EXE = os.path.join(DROPIN, "sgpar_dropin_hec3_fm")
DROPIN_MIS_LIB = os.path.join(PKG, "libsgpar_dropin_mis.so")
DROPIN_MIS_EXE = os.path.join(DROPIN, "sgpar_dropin_mis")
DROPIN_GOSH_LIB = os.path.join(PKG, "libsgpar_dropin_gosh.so")
DROPIN_GOSH_EXE = os.path.join(DROPIN, "sgpar_dropin_gosh")
DROPIN_HASHMAP_LIB = os.path.join(PKG, "libsgpar_dropin_hashmap.so")
DROPIN_HASHMAP_EXE = os.path.join(DROPIN, "sgpar_dropin_hashmap")
DROPIN_MATCH_LIB = os.path.join(PKG, "libsgpar_dropin_match.so")
DROPIN_MATCH_EXE = os.path.join(DROPIN, "sgpar_dropin_match")
DROPIN_MATCH_FM_LIB = os.path.join(PKG, "libsgpar_dropin_match_fm.so")
DROPIN_MATCH_FM_EXE = os.path.join(DROPIN, "sgpar_dropin_match_fm")
DROPIN_MTMETIS_LIB = os.path.join(PKG, "libsgpar_dropin_mtmetis.so")
DROPIN_MTMETIS_EXE = os.path.join(DROPIN, "sgpar_dropin_mtmetis")
DROPIN_HEC2_LIB = os.path.join(PKG, "libsgpar_dropin_hec2.so")
DROPIN_HEC2_EXE = os.path.join(DROPIN, "sgpar_dropin_hec2")
DROPIN_HEC2_FM_LIB = os.path.join(PKG, "libsgpar_dropin_hec2_fm.so")
DROPIN_HEC2_FM_EXE = os.path.join(DROPIN, "sgpar_dropin_hec2_fm")
DROPIN_GOSH2_LIB = os.path.join(PKG, "libsgpar_dropin_gosh2.so")
DROPIN_GOSH2_EXE = os.path.join(DROPIN, "sgpar_dropin_gosh2")
DROPIN_SERIAL_LIB = os.path.join(PKG, "libsgpar_dropin_serial.so")
DROPIN_SERIAL_EXE = os.path.join(DROPIN, "sgpar_dropin_serial")
REF_CMG = os.environ.get("SGPAR_REF_CMG", "/root/reference/cmg")


def build_dropin(verbose: bool = False, ref_cmg: str = REF_CMG) -> bool:
    """The reference-API shim (libsgpar_dropin.so over libsgpar_mi.so) and
    its command-line driver, compiled against the reference's own
    cmg/sgpar.h and cmg/ExperimentLoggerUtil.cpp where they lie.  Skipped
    (returns False) when that directory is absent; the built files are then
    the ones a previous build left."""
    if not os.path.isfile(os.path.join(ref_cmg, "sgpar.h")):
        return False
    cxx = os.environ.get("CXX", "g++")
    common = ["-O2", "-std=c++17", "-DEXPERIMENT", "-I", ref_cmg, "-I", os.path.join(os.path.dirname(PKG), "include"),
              "-w"]
    for lib, exe, extra in ((DROPIN_LIB, DROPIN_EXE, []), (DROPIN_FM_LIB, DROPIN_FM_EXE, ["-DFM"]),
                            (DROPIN_HEC3_LIB, DROPIN_HEC3_EXE, ["-DHEC_V3"]),
                            (DROPIN_HEC3_FM_LIB, DROPIN_HEC3_FM_EXE, ["-DHEC_V3", "-DFM"]),
                            (DROPIN_MIS_LIB, DROPIN_MIS_EXE, ["-DMIS"]),
                            (DROPIN_GOSH_LIB, DROPIN_GOSH_EXE, ["-DGOSH"]),
                            (DROPIN_HASHMAP_LIB, DROPIN_HASHMAP_EXE, ["-DHASHMAP"]),
                            (DROPIN_MATCH_LIB, DROPIN_MATCH_EXE, ["-DPUREMATCH"]),
                            (DROPIN_MATCH_FM_LIB, DROPIN_MATCH_FM_EXE, ["-DPUREMATCH", "-DFM"]),
                            (DROPIN_MTMETIS_LIB, DROPIN_MTMETIS_EXE, ["-DMTMETIS"]),
                            (DROPIN_HEC2_LIB, DROPIN_HEC2_EXE, ["-DHEC_V2"]),
                            (DROPIN_HEC2_FM_LIB, DROPIN_HEC2_FM_EXE, ["-DHEC_V2", "-DFM"]),
                            (DROPIN_GOSH2_LIB, DROPIN_GOSH2_EXE, ["-DGOSH_V2"]),
                            (DROPIN_SERIAL_LIB, DROPIN_SERIAL_EXE, ["-DHEC_SERIAL"])):
        tmp = lib + ".tmp"
        # -Bsymbolic: calls inside the shim bind to its own definitions, not
        # to copies a caller's translation unit of sgpar.h may define
        # the version script exports the API only; the serial helpers the
        # reference's header defines in every translation unit stay local
        # and are dropped
        _run([cxx, *common, *extra, "-fPIC", "-shared", "-ffunction-sections",
              os.path.join(DROPIN, "sgpar_dropin.cpp"), "-o", tmp, "-L", PKG, "-l:libsgpar_mi.so",
              "-Wl,-rpath,$ORIGIN", "-Wl,-Bsymbolic", "-Wl,--gc-sections",
              f"-Wl,--version-script={os.path.join(DROPIN, 'sgpar_dropin.map')}"], verbose)
        os.replace(tmp, lib)
        tmp = exe + ".tmp"
        _run([cxx, *common, *extra, os.path.join(DROPIN, "sgpar_dropin_main.cpp"), "-o", tmp, "-L", PKG,
              f"-l:{os.path.basename(lib)}", "-Wl,-rpath,$ORIGIN/.."], verbose)
        os.replace(tmp, exe)
    return True


if __name__ == "__main__":
    print(build(verbose=True))
    print("drop-in built" if build_dropin(verbose=True) else "drop-in skipped (no reference headers)")
