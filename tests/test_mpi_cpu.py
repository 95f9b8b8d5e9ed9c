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
