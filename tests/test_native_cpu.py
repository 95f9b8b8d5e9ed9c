"""CPU-side checks of the native libraries: every symbol declared in
include/*.h is exported by the library that implements it, and the host
generators produce well-formed, deterministic inputs."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import cc_mpi_amd as cm
from cc_mpi_amd import _native as nat

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# header -> library that must export every function it declares
HEADER_LIBS = {
    "cc.h": "libccmi355x.so",
    "cc_hip.h": "libccmi355x.so",
    "boundary_sort.h": "libccmi355x.so",
    "graph.h": "libcchost.so",
    "graphgen.h": "libcchost.so",
    "graph_dist.h": "libccmpi.so",
    "exchange_plan.h": "libccmpi.so",
    "cc_mpi.h": "libccmpi.so",
}


def test_every_header_is_covered():
    """Each header that declares functions has a library that must export them."""
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if h.endswith(".h") and declared_functions(h):
            assert h in HEADER_LIBS, f"include/{h} declares functions but no library is checked"


def declared_functions(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    text = re.sub(r"#[^\n]*", "", text)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\([^;{)]*(?:\([^)]*\)[^;{)]*)*\)\s*;", text)
    return sorted(set(n for n in names if n not in ("if", "while", "for", "return", "sizeof")))


def exported(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", os.path.join(nat.LIBDIR, lib)],
                         capture_output=True, text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


@pytest.mark.parametrize("header,lib", sorted(HEADER_LIBS.items()))
def test_header_symbols_exported(header, lib):
    names = declared_functions(header)
    assert names, f"no declarations parsed from {header}"
    missing = [n for n in names if n not in exported(lib)]
    assert not missing, f"{lib} lacks {missing} declared in include/{header}"


def test_libraries_load_without_gpu():
    lib = nat.gpu_lib()
    assert lib.cc_hip_build_info().startswith(b"libccmi355x")
    nat.host_lib()
    # the offload bundle of this library (built here) targets gfx950
    blob = open(os.path.join(nat.LIBDIR, "libccmi355x.so"), "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob


def test_no_device_reports_error_not_fallback():
    lib = nat.gpu_lib()
    if lib.cc_hip_device_count() > 0:
        pytest.skip("a GPU is visible")
    G = cm.rmat(8, 4, seed=1)
    with pytest.raises(RuntimeError, match="no HIP device"):
        cm.compute_connected_components_pthreads(G)


def test_noalloc_solve_never_initialises(tmp_path):
    """The allocation-free solve refuses to run before cc_hip_init (which
    sets up its scratch array) instead of initialising the library, which
    would allocate; checked in a fresh process that never initialises."""
    code = (
        "import ctypes, sys\n"
        f"lib = ctypes.CDLL({os.path.join(nat.LIBDIR, 'libccmi355x.so')!r})\n"
        "lib.cc_hip_last_error.restype = ctypes.c_char_p\n"
        "f = lib.cc_hip_components_device_noalloc\n"
        "f.argtypes = [ctypes.c_void_p] * 2 + [ctypes.c_uint32, ctypes.c_uint64]"
        " + [ctypes.c_void_p] * 2\n"
        "rc = f(None, None, 10, 0, None, None)\n"
        "print(rc, lib.cc_hip_last_error().decode())\n")
    r = subprocess.run(["python3", "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.startswith("2 ") and "cc_hip_init" in r.stdout


def well_formed(G):
    rp, ci = G.row_ptr, G.col_idx
    assert rp[0] == 0 and (np.diff(rp.astype(np.int64)) >= 0).all() and rp[-1] == ci.size
    src = np.repeat(np.arange(G.n, dtype=np.uint32), np.diff(rp).astype(np.int64))
    assert (ci < G.n).all() and (src != ci).all()
    key = (src.astype(np.uint64) << np.uint64(32)) | ci
    assert (np.diff(key.astype(np.uint64)) > 0).all()  # sorted rows, no duplicates
    rev = (ci.astype(np.uint64) << np.uint64(32)) | src
    assert np.array_equal(np.sort(rev), key)  # symmetric


@pytest.mark.parametrize("make", [
    lambda: cm.erdos_renyi(5000, 8.0, seed=1),
    lambda: cm.rmat(12, 16, seed=2),
    lambda: cm.grid(50, 70, 0.7, seed=3),
    lambda: cm.chung_lu(5000, 20000, 2.1, seed=4),
])
def test_generators_well_formed_and_deterministic(make):
    a, b = make(), make()
    well_formed(a)
    assert np.array_equal(a.row_ptr, b.row_ptr) and np.array_equal(a.col_idx, b.col_idx)


def test_grid_structure():
    G = cm.grid(4, 5, 1.0, seed=0, permute=False)
    assert G.m == 2 * (4 * 4 + 3 * 5)
    G = cm.grid(4, 5, 0.0, seed=0, permute=False)
    assert G.m == 0


def test_rmat_sizes():
    G = cm.rmat(10, 16, seed=1002)
    assert G.n == 1024 and G.m <= 2 * 16 * 1024


def test_plan_defaults():
    """The default plan: cache on; one-pass hook; Rem's unions on sparse
    graphs (splicing every step and 4 workgroups per CU from 2^24 vertices,
    8 below) and on dense graphs below 2^23 vertices (each vertex's own
    entry spliced, 8 per CU);
    dense ones of 2^23+ vertices the lockstep hook from the CSR with path
    halving and one root per thread.  Linked vertices skip neighbour 1 (with one-hop links) on
    that hook (48 workgroups per CU) and on sparse graphs of degree 6 and
    more, not on the lattice or R-MAT 22.  The hook from the CSR runs
    low-first (ids below n/16 first), with the owed unions queued per wave
    below 2^31 entries."""
    lib = nat.gpu_lib()
    def plan(n, m):
        v = [ctypes.c_int() for _ in range(4)]
        assert lib.cc_hip_components_plan(n, m, *[ctypes.byref(x) for x in v]) == 0
        return tuple(x.value for x in v)

    def get(n, m):
        return tuple(lib.cc_hip_components_plan_get(n, m, k) for k in (
            b"rem", b"rem_splice", b"halve", b"skip_linked", b"compress_ilp", b"csr_hook"))

    def low(n, m):
        return tuple(lib.cc_hip_components_plan_get(n, m, k) for k in (b"low_shift", b"split_walks"))
    # R-MAT 22, Chung-Lu 65.6M, ER 4M degree 8, ER 4M degree 3, 8192^2
    # grid, R-MAT 24, R-MAT 27
    assert plan(4194304, 128313032) == (1, 1, 1, 8) and get(4194304, 128313032) == (1, 2, 0, 0, 1, 0)
    assert plan(65601024, 3600152054) == (1, 1, 1, 48) and get(65601024, 3600152054) == (0, 0, 1, 2, 4, 1)
    assert low(65601024, 3600152054) == (4, 0) and low(4194304, 128313032) == (0, 0)
    assert plan(4194304, 33554406) == (1, 1, 1, 8) and get(4194304, 33554406) == (1, 0, 0, 2, 4, 0)
    assert plan(4194304, 12582900) == (1, 1, 1, 8) and get(4194304, 12582900) == (1, 0, 0, 0, 1, 0)
    assert plan(67108864, 187873588) == (1, 1, 1, 4) and get(67108864, 187873588) == (1, 1, 0, 0, 1, 0)
    assert plan(16777216, 520769534) == (1, 1, 1, 48) and get(16777216, 520769534) == (0, 0, 1, 2, 4, 1)
    assert low(16777216, 520769534) == (4, 1) and low(67108864, 187873588) == (0, 0)
    assert plan(134217728, 4223308732) == (1, 1, 1, 48) and get(134217728, 4223308732) == (0, 0, 1, 2, 4, 1)
    assert low(134217728, 4223308732) == (4, 0)
    # R-MAT 23 and Chung-Lu 16.4M: dense from 2^23 vertices, 16 workgroups
    # per CU below 2^24
    assert plan(8388608, 256000000) == (1, 1, 1, 16) and get(8388608, 256000000) == (0, 0, 1, 2, 4, 1)
    assert plan(16400256, 865053518) == (1, 1, 1, 16) and low(16400256, 865053518) == (4, 1)
    assert lib.cc_hip_components_plan_get(1, 1, b"hook_mode") == 4
    assert lib.cc_hip_components_plan_get(1, 1, b"nonsense") == -1
