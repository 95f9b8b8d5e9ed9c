"""Host-only MPI checks (run here, no GPU): the exchange-plan build and the
legacy label exchange of include/exchange_plan.h, under mpiexec with 1-4
ranks on localhost, including a rank that has no ghosts (which hangs the
reference's plan build)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "cc_mpi_amd", "lib")
MPIEXEC = "/opt/conda/bin/mpiexec"
MPI_INC = "/opt/conda/include"
MPI_LIB = "/opt/conda/lib/libmpi.so"

pytestmark = pytest.mark.skipif(not (os.path.exists(MPIEXEC) and os.path.exists(MPI_LIB)),
                                reason="MPICH not available")


@pytest.fixture(scope="module")
def plan_check(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("mpi") / "plan_check")
    cmd = ["gcc", "-O2", "-std=gnu11", f"-I{ROOT}/include", f"-I{MPI_INC}",
           os.path.join(ROOT, "tests", "mpi", "plan_check.c"), "-o", out, f"-L{LIB}", "-lccmpi",
           "-lccmi355x", "-Wl,-rpath-link,/usr/lib/x86_64-linux-gnu", MPI_LIB,
           f"-Wl,-rpath,{LIB}", "-Wl,-rpath,/opt/conda/lib"]
    subprocess.run(cmd, check=True)
    return out


@pytest.mark.parametrize("ranks,n", [(1, 100), (2, 1000), (3, 1001), (4, 4096)])
def test_exchange_plan(plan_check, ranks, n):
    env = dict(os.environ)
    env.pop("HIP_VISIBLE_DEVICES", None)
    r = subprocess.run([MPIEXEC, "-launcher", "fork", "-n", str(ranks), plan_check, str(n)],
                       capture_output=True, text=True, timeout=120, env=env, cwd="/tmp")
    assert r.returncode == 0, r.stdout + r.stderr
    assert "plan_check OK" in r.stdout


# ---- delta exchange: the oracle pinned against the reference's own build ----

import sys  # noqa: E402

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import delta_harness as dh  # noqa: E402

needs_ref_mpi = pytest.mark.skipif(
    not os.path.exists(os.path.join(dh.REF, "libccref_mpi.so")),
    reason="oracle/_ref/libccref_mpi.so not built (reference sources absent)")


@pytest.fixture(scope="module")
def delta_ref(tmp_path_factory):
    return dh.build_harness(str(tmp_path_factory.mktemp("dref") / "delta_check_ref"), "reference")


@needs_ref_mpi
@pytest.mark.parametrize("case", dh.CASES, ids=lambda c: "P%d-n%d" % c[:2])
def test_reference_delta_exchange_equals_oracle(delta_ref, case):
    """The reference's exchangeplan_build + exchangeplan_exchange_delta
    (src/exchange_plan.c, compiled from its sources) over several rounds of
    lowered labels: every round's prev_sent, send buffer, receive buffer and
    ghost_labels equal oracle_delta_pack / oracle_delta_unpack (checked inside
    the harness), and the digests equal the committed golden file."""
    lines = dh.run_harness(delta_ref, *case)
    assert lines == dh.golden()[case]


@needs_ref_mpi
@pytest.mark.parametrize("mistake", ["pack_global_index", "unpack_dropped_base"])
def test_planted_oracle_mistake_is_caught(tmp_path, mistake):
    """A plausible mistake planted in a copy of the oracle makes the
    reference comparison fail: the pack emitting the position in
    send_to_flat instead of the position in the destination's segment, or
    the unpack dropping the source's base in need_from_gidx_flat."""
    src = open(os.path.join(ROOT, "oracle", "cc_oracle.c")).read()
    if mistake == "pack_global_index":
        good = "sendbuf[total + (uint64_t)wrote] = ((uint64_t)(uint32_t)i << 32) | lbl;"
        bad = "sendbuf[total + (uint64_t)wrote] = ((uint64_t)(uint32_t)idx << 32) | lbl;"
    else:
        good = "ghost_labels[need_from_gidx_flat[need_base[k] + (int)pos]] = lbl;"
        bad = "ghost_labels[need_from_gidx_flat[(int)pos]] = lbl;"
    assert src.count(good) == 1
    mutated = tmp_path / "cc_oracle_bad.c"
    mutated.write_text(src.replace(good, bad))
    olib = str(tmp_path / "libccoracle_bad.so")
    subprocess.run(["gcc", "-O2", "-std=gnu11", "-fPIC", "-pthread", "-shared", "-o", olib,
                    str(mutated), "-lpthread"], check=True)
    exe = dh.build_harness(str(tmp_path / "delta_check_bad"), "reference", oracle_lib=olib)
    r, _ = dh.run_harness(exe, 3, 5001, 4, 2, check=False)
    assert r.returncode != 0
    assert "differ" in r.stderr
