"""The pose kernel's cell-neighbourhood arithmetic (flexaidds_amd/csrc/
cell_slots.h) against the reference's enumeration of the 27 slots
(LIB/Vcontacts.cpp:1738-1741), on the host: every pair of cells of grids
with 3..14 cells a side (row wraps included), random pairs up to 60 a side,
and the repeat counts of coarse grids."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "flexaidds_amd", "csrc")


def test_cell_slots_equal_the_enumeration(tmp_path):
    exe = str(tmp_path / "cell_slots_check")
    subprocess.run(["g++", "-O2", "-std=c++17", f"-I{CSRC}", "-o", exe, os.path.join(HERE, "cell_slots_check.cpp")],
                   check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert " 0 mismatches" in r.stdout
