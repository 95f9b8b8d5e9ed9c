"""Build and run tests/mpi/delta_check.c (the delta exchange checked against
the oracle under mpiexec) against one of two builds of include/exchange_plan.h:
the reference's own src/exchange_plan.c compiled from its sources
(oracle/_ref/libccref_mpi.so, in the build container only), or this
project's libccmpi.so (GPU pack)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "cc_mpi_amd", "lib")
ORACLE = os.path.join(ROOT, "oracle")
REF = os.path.join(ORACLE, "_ref")
MPIEXEC = "/opt/conda/bin/mpiexec"
MPI_INC = "/opt/conda/include"
MPI_LIB = "/opt/conda/lib/libmpi.so"

# (ranks, vertices, rounds, seed)
CASES = [(2, 1000, 5, 1), (3, 5001, 6, 2), (4, 200000, 6, 3), (4, 64, 4, 4)]


def build_harness(out: str, backend: str, oracle_lib: str | None = None) -> str:
    """gcc the harness against `backend` ("reference" or "gpu")."""
    src = os.path.join(ROOT, "tests", "mpi", "delta_check.c")
    olib = oracle_lib or os.path.join(ORACLE, "libccoracle.so")
    odir = os.path.dirname(olib)
    cmd = ["gcc", "-O2", "-std=gnu11", f"-I{ROOT}/include", f"-I{MPI_INC}", src, "-o", out]
    if backend == "reference":
        cmd += [os.path.join(REF, "libccref_mpi.so"), f"-Wl,-rpath,{REF}"]
    else:
        cmd += [f"-L{LIB}", "-lccmpi", "-lccmi355x", "-lcchost", f"-Wl,-rpath,{LIB}"]
    cmd += [olib, f"-Wl,-rpath,{odir}", "-Wl,-rpath-link,/usr/lib/x86_64-linux-gnu", MPI_LIB,
            "-Wl,-rpath,/opt/conda/lib"]
    subprocess.run(cmd, check=True)
    return out


def run_harness(exe: str, ranks: int, n: int, rounds: int, seed: int, check=True):
    """Run under mpiexec; returns the digest lines sorted by rank (each rank
    prints its own lines, interleaved on stdout)."""
    env = dict(os.environ)
    r = subprocess.run([MPIEXEC, "-launcher", "fork", "-n", str(ranks), exe, str(n), str(rounds),
                        str(seed)], capture_output=True, text=True, timeout=180, env=env,
                       cwd="/tmp")
    if check:
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        assert "delta_check OK" in r.stdout
    lines = [ln for ln in r.stdout.splitlines() if ln.startswith(("plan ", "round "))]

    def key(ln):
        f = dict(kv.split("=", 1) for kv in ln.split()[1:])
        return (int(f["rank"]), ln.startswith("round "), int(f.get("round", -1)))
    return sorted(lines, key=key) if check else (r, lines)


def golden() -> dict:
    """case tuple -> digest lines, from tests/golden/delta_exchange_digests.txt."""
    out, cur = {}, None
    with open(os.path.join(ROOT, "tests", "golden", "delta_exchange_digests.txt")) as f:
        for ln in f.read().splitlines():
            if ln.startswith("case "):
                kv = dict(x.split("=") for x in ln.split()[1:])
                cur = (int(kv["ranks"]), int(kv["n"]), int(kv["rounds"]), int(kv["seed"]))
                out[cur] = []
            elif ln:
                out[cur].append(ln)
    return out
