"""The distributed driver (include/cc_mpi.h) end to end on the GPU: mpiexec
runs cc_mpi_run with 1-4 ranks sharing the one GPU; the gathered
labels_global must equal the oracle's minimum-id labels bit for bit, for
both partition kinds and several exchange intervals."""
import os
import subprocess

import numpy as np
import pytest

import cc_mpi_amd as cm
import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUN = os.path.join(ROOT, "cc_mpi_amd", "lib", "cc_mpi_run")
MPIEXEC = "/opt/conda/bin/mpiexec"

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(MPIEXEC), reason="MPICH not available")]

CASES = [
    (1, ["--gen", "rmat", "--scale", "14", "--ef", "16", "--seed", "21"], 1, False),
    (2, ["--gen", "rmat", "--scale", "14", "--ef", "16", "--seed", "21"], 1, False),
    (4, ["--gen", "rmat", "--scale", "15", "--ef", "8", "--seed", "22"], 2, False),
    (3, ["--gen", "rmat", "--scale", "13", "--ef", "16", "--seed", "23"], 4, True),
    (4, ["--gen", "er", "--n", "30000", "--deg", "1.5", "--seed", "24"], 1, True),
    (3, ["--gen", "grid", "--rows", "120", "--cols", "90", "--keep", "0.6", "--seed", "25"], 2,
     False),
    (4, ["--gen", "grid", "--rows", "64", "--cols", "64", "--keep", "0.9", "--seed", "26"], 1,
     True),
    # larger shares: many split tiles per rank, a giant local component
    # (the boundary build's bitmap path), many small components
    (4, ["--gen", "rmat", "--scale", "18", "--ef", "16", "--seed", "27"], 2, True),
    (2, ["--gen", "er", "--n", "200000", "--deg", "0.8", "--seed", "28"], 3, False),
]


def graph_for(args):
    a = dict(zip(args[::2], args[1::2]))
    seed = int(a["--seed"])
    if a["--gen"] == "rmat":
        return cm.rmat(int(a["--scale"]), int(a["--ef"]), seed=seed)
    if a["--gen"] == "er":
        n = int(a["--n"])
        return cm.erdos_renyi(n, float(a["--deg"]), seed=seed)
    return cm.grid(int(a["--rows"]), int(a["--cols"]), float(a["--keep"]), seed=seed)


@pytest.mark.parametrize("ranks,args,exch,vertex_blocks", CASES)
def test_distributed_driver(tmp_path, ranks, args, exch, vertex_blocks):
    out = str(tmp_path / "labels.bin")
    cmd = [MPIEXEC, "-launcher", "fork", "-n", str(ranks), RUN, *args, "-e", str(exch),
           "--out", out]
    if vertex_blocks:
        cmd.append("--vertex-blocks")
    print("running:", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=240, cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    G = graph_for(args)
    got = np.fromfile(out, dtype=np.uint32)
    want = oracle.cc_min_label(G.row_ptr, G.col_idx)
    assert got.size == G.n
    bad = np.flatnonzero(got != want)
    assert bad.size == 0, f"{bad.size} labels differ, first {bad[:5]}\n{r.stdout[-2000:]}"
    roots = int(np.count_nonzero(want == np.arange(G.n)))
    assert f"Number of connected components: {roots}" in r.stdout


@pytest.mark.parametrize("ranks,flags", [(1, 0), (2, 3), (4, 4)])
def test_driver_reads_mat_file(tmp_path, ranks, flags):
    """The reference's CPU-runnable case end to end through its file path:
    ER with 65,536 vertices and average degree 8 stored as a sparse logical
    .mat (plain, compressed Problem struct, big-endian), read on rank 0 by
    load_dist_csr_from_file, distributed and solved."""
    G = cm.erdos_renyi(65536, 8.0, seed=1001)
    mat = str(tmp_path / "er64k.mat")
    cm.save_mat(mat, G, "A", flags)
    out = str(tmp_path / "labels.bin")
    cmd = [MPIEXEC, "-launcher", "fork", "-n", str(ranks), RUN, "--input", mat, "--out", out]
    print("running:", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=240, cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    got = np.fromfile(out, dtype=np.uint32)
    want = oracle.cc_min_label(G.row_ptr, G.col_idx)
    np.testing.assert_array_equal(got, want)
    roots = int(np.count_nonzero(want == np.arange(G.n)))
    assert f"Number of connected components: {roots}" in r.stdout


def test_driver_rejects_unreadable_input(tmp_path):
    bad = tmp_path / "bad.mat"
    bad.write_bytes(b"\0" * 256)
    cmd = [MPIEXEC, "-launcher", "fork", "-n", "2", RUN, "--input", str(bad)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120, cwd=str(tmp_path))
    assert r.returncode != 0
    assert "cannot load" in r.stderr


import sys  # noqa: E402

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import delta_harness as dh  # noqa: E402


@pytest.fixture(scope="module")
def delta_gpu(tmp_path_factory):
    return dh.build_harness(str(tmp_path_factory.mktemp("dgpu") / "delta_check_gpu"), "gpu")


@pytest.mark.parametrize("case", dh.CASES, ids=lambda c: "P%d-n%d" % c[:2])
def test_host_array_delta_exchange(delta_gpu, case):
    """The reference-prototype exchangeplan_exchange_delta (host arrays; pack
    on the GPU into pinned memory, device state cached per plan after the
    first call) with 2-4 ranks sharing the GPU, over several rounds: every
    round equals the oracle's pack, transport and unpack (checked inside the
    harness), and the plan and per-round digests equal the ones the
    reference's own build produced (tests/golden/delta_exchange_digests.txt)."""
    print("running delta_check", case, flush=True)
    lines = dh.run_harness(delta_gpu, *case)
    assert lines == dh.golden()[case]
