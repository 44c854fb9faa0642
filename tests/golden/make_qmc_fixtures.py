"""Copy the reference's committed Sobol golden vectors into tests/golden/qmc.

The reference holds them as data files (tests/qmc/q0_*.txt, q2_*.txt, 1024
points each, produced by its q0_dump.f90/q2_dump.f90 drivers).  They are data,
not source; this keeps the full 1024 rows of the 3-dimension files and the
first 256 rows of the 12-dimension files for a subset of seeds, to keep the
fixtures small.

    python tests/golden/make_qmc_fixtures.py [/root/reference/tests/qmc]
"""
import os
import sys

SRC = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/tests/qmc"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "qmc")

FULL = ["q0_raw.txt", "q0_scr_12345.txt", "q0_scr_777.txt", "q0_scr_2024.txt"]
HEAD = ["q2_raw.txt", "q2_scr_12345.txt", "q2_scr_777.txt", "q2_scr_999983.txt",
        "q2_dif_12345.txt", "q2_dif_777.txt", "q2_dif_999983.txt"]

os.makedirs(DST, exist_ok=True)
for name, nrows in [(n, None) for n in FULL] + [(n, 256) for n in HEAD]:
    with open(os.path.join(SRC, name)) as f:
        rows = f.read().splitlines()
    if nrows is not None:
        rows = rows[:nrows]
    with open(os.path.join(DST, name), "w") as f:
        f.write("\n".join(rows) + "\n")
    print(name, len(rows))
