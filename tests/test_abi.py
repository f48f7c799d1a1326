"""The C ABI of libsgpar_mi.so, checked without a GPU.

* every entry point declared in include/sgpar_mi.h is exported by the built
  library, and the Python binding binds exactly that set;
* the header compiles as C and links against the library with each entry
  point bound through a pointer of its declared type (tests/abi_link_check.c);
* mi_config / mi_stats have the same layout in C and in the binding;
* entry points that need no device reject null handles with MI_STATUS_ARG
  and a missing device with MI_STATUS_HIP;
* the package refuses to run when the library is missing (no CPU fallback).
"""
import ctypes as C
import os
import re
import shutil
import subprocess

import pytest

import ipdps_2021_graph_coarsening_amd as mi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "sgpar_mi.h")


def declared_functions() -> list[str]:
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mi_[a-z0-9_]+)\s*\(", text)))


def exported_symbols() -> set[str]:
    out = subprocess.run(["nm", "-D", "--defined-only", mi.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


@pytest.fixture(scope="module")
def built():
    if not os.path.exists(mi.LIB_PATH):
        pytest.fail(f"{mi.LIB_PATH} is not built (python -m ipdps_2021_graph_coarsening_amd.build)")
    return mi.LIB_PATH


def test_header_declares_the_path_entry_points():
    fns = declared_functions()
    for name in ("mi_partition_graph", "mi_hierarchy_build", "mi_fiedler", "mi_power_iter",
                 "mi_compute_partition", "mi_hec_level", "mi_build_coarse"):
        assert name in fns
    assert len(fns) >= 25


def test_library_exports_every_declared_symbol(built):
    exp = exported_symbols()
    missing = [f for f in declared_functions() if f not in exp]
    assert not missing, f"declared but not exported: {missing}"
    extra = sorted(s for s in exp if s.startswith("mi_") and s not in declared_functions())
    assert not extra, f"exported but not declared: {extra}"


def test_binding_covers_the_header(built):
    assert sorted(mi.EXPORTED) == declared_functions()
    L = mi.lib()
    for name in mi.EXPORTED:
        assert getattr(L, name) is not None


def test_header_compiles_as_c_and_links(built, tmp_path):
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    exe = tmp_path / "abi_link_check"
    src = os.path.join(ROOT, "tests", "abi_link_check.c")
    libdir = os.path.dirname(mi.LIB_PATH)
    subprocess.run([cc, "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), src,
                    "-o", str(exe), "-L", libdir, "-l:libsgpar_mi.so", f"-Wl,-rpath,{libdir}"],
                   check=True, capture_output=True, text=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    fields = dict(kv.split("=") for kv in r.stdout.split())
    assert int(fields["sizeof(mi_config)"]) == C.sizeof(mi.Config)
    assert int(fields["sizeof(mi_stats)"]) == C.sizeof(mi.Stats)
    if not os.path.exists("/dev/kfd"):
        # no GPU here: creating a context must fail cleanly with MI_STATUS_HIP
        assert int(fields["ctx_create"]) == 2


def test_config_layout_mirrors_reference_config_t():
    # config_t, sgpar.h:186-193: four ints, a double, an int
    names = [f[0] for f in mi.Config._fields_]
    assert names == ["coarsening_alg", "refine_alg", "local_search_alg", "num_iter", "tol", "num_partitions"]
    assert mi.Config.tol.offset == 16


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(mi, "LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(mi, "_lib", None)
    with pytest.raises(RuntimeError, match="missing"):
        mi.lib()
    with pytest.raises(RuntimeError):
        mi.Context(0)


DROPIN_API = {
    "sgpar::sgp_partition_graph(unsigned int*, unsigned int*, sgpar::configuration*, int, sgpar::sgp_graph_t, "
    "ExperimentLoggerUtil&, sgpar::sgp_pcg32_random_t*)",
    "sgpar::sgp_load_graph(sgpar::sgp_graph_t*, char*)",
    "sgpar::sgp_free_graph(sgpar::sgp_graph_t*)",
    "sgpar::sgp_load_partition(unsigned int*, unsigned int, char*)",
    "sgpar::sgp_use_partition(unsigned int*, sgpar::sgp_graph_t, sgpar::sgp_graph_t*, sgpar::sgp_graph_t*)",
    "sgpar::sgp_load_config(char const*, sgpar::configuration*)",
    "sgpar::compute_partition_edit_distance(unsigned int const*, unsigned int const*, unsigned int, unsigned int*)",
    "sgpar_mi_fix_clock",
}


def test_dropin_exports_the_reference_api():
    """libsgpar_dropin.so exports the functions of the reference's API
    (cmg/sgpar.h:2230-2244 with EXPERIMENT; demangled names carry the
    reference's own types) and the clock pin, nothing else of the reference
    header's serial code; its driver links it."""
    from ipdps_2021_graph_coarsening_amd import build as mibuild
    if not os.path.exists(mibuild.DROPIN_LIB):
        pytest.skip("the drop-in was not built (it needs the reference's cmg headers at build time)")
    out = subprocess.run(["nm", "-DC", "--defined-only", mibuild.DROPIN_LIB], capture_output=True, text=True,
                         check=True).stdout
    names = {line.split(" ", 2)[2] for line in out.splitlines() if line.count(" ") >= 2}
    ours = {x for x in names if x.startswith("sgpar") or x.startswith("sgp")}
    assert ours == DROPIN_API
    ldd = subprocess.run(["ldd", mibuild.DROPIN_EXE], capture_output=True, text=True, check=True).stdout
    assert "libsgpar_dropin.so" in ldd and "libsgpar_mi.so" in ldd and "not found" not in ldd
