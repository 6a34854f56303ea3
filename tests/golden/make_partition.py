"""Writes tests/golden/partition.json: every rank's Nurbs fields after the
reference's rank partition (mesh/nurbs/nurbs_init.c:9-199, set_mpi_part) for
a few meshes and rank counts, from oracle/_ref/partdump (built by
`make -C oracle part` from the reference sources, run under MPICH's mpiexec
in a container that has both).

    make -C oracle part && python tests/golden/make_partition.py
"""
import json
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
EXE = os.path.join(ROOT, "oracle", "_ref", "partdump")
MPIEXEC = os.environ.get("MPIEXEC", "/opt/conda/bin/mpiexec")

CASES = [((6, 5, 7), 4), ((6, 5, 7), 8), ((4, 4, 4), 2), ((9, 3, 5), 6), ((5, 6, 4), 3), ((8, 8, 8), 8)]


def main():
    if not os.path.exists(EXE):
        raise SystemExit("oracle/_ref/partdump missing: make -C oracle part")
    out = []
    for nelem, nproc in CASES:
        with tempfile.TemporaryDirectory() as d:
            subprocess.run([MPIEXEC, "-n", str(nproc), EXE] + [str(n) for n in nelem] + [d], check=True,
                           stdout=subprocess.DEVNULL, timeout=120)
            ranks = [json.load(open(os.path.join(d, "rank_%d.json" % r))) for r in range(nproc)]
        out.append({"nelem": list(nelem), "nproc": nproc, "ranks": ranks})
    with open(os.path.join(HERE, "partition.json"), "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print("wrote", os.path.join(HERE, "partition.json"), file=sys.stderr)


if __name__ == "__main__":
    main()
