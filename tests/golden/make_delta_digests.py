"""Writes delta_exchange_digests.txt: the observable state of the
reference's own exchangeplan_build + exchangeplan_exchange_delta
(src/exchange_plan.c, compiled from its sources into
oracle/_ref/libccref_mpi.so by `make -C oracle ref`) after every round of
tests/mpi/delta_check.c, for each case in CASES.  The harness checks each
round against the oracle while it runs; this file keeps the per-rank
digests (FNV-1a of the plan arrays, prev_sent, ghost_labels and the
send/receive buffers) so that builds without the reference can be compared
with it line for line.

Run in the build container:  python tests/golden/make_delta_digests.py
"""
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))
from delta_harness import CASES, build_harness, run_harness  # noqa: E402


def main() -> int:
    ref = os.path.join(ROOT, "oracle", "_ref", "libccref_mpi.so")
    if not os.path.exists(ref):
        print("oracle/_ref/libccref_mpi.so missing: run make -C oracle ref", file=sys.stderr)
        return 1
    with tempfile.TemporaryDirectory() as tmp:
        exe = build_harness(os.path.join(tmp, "delta_check_ref"), backend="reference")
        out = []
        for case in CASES:
            lines = run_harness(exe, *case)
            out.append(f"case ranks={case[0]} n={case[1]} rounds={case[2]} seed={case[3]}")
            out.extend(lines)
    path = os.path.join(HERE, "delta_exchange_digests.txt")
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")
    print(f"wrote {path} ({len(out)} lines)")
    return 0


if __name__ == "__main__":
    sys.exit(main())
