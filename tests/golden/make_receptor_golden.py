"""Writes tests/golden/receptor_guanine.npz: the reference's receptor-file
chain (oracle/_ref/libfa_ref_reader.so, built from LIB/ by oracle/ref/Makefile)
run on the reference's RNA target guanine_test_INI.pdb with its default
filters (waters removed, heterogroups kept).  Holds the atoms as read
(coordinates, radius, type, PDB serial) and calc_center's outputs; the GPU
tests dock ligands against this receptor without the reference checkout.

Run from the repository root after `make -C oracle/ref`:
    python tests/golden/make_receptor_golden.py [reference checkout]
"""
from __future__ import annotations

import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

import oracle_lib as ol  # noqa: E402


def main():
    ref_dir = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    lib = ol.load_ref_reader()
    if lib is None:
        raise SystemExit("oracle/_ref/libfa_ref_reader.so is not built")
    with tempfile.TemporaryDirectory() as tmp:
        r = ol.ref_read_receptor(lib, os.path.join(ref_dir, "guanine_test_INI.pdb"),
                                 os.path.join(ref_dir, "WRK", "AMINO.def"),
                                 os.path.join(ref_dir, "WRK", "NUCLEOTIDES.def"),
                                 os.path.join(tmp, "mod.pdb"))
    out = os.path.join(HERE, "receptor_guanine.npz")
    np.savez_compressed(out, xyz=r["xyz"], radius=r["radius"], type=r["type"], number=r["number"],
                        center=r["center"], is_protein=np.int32(r["is_protein"]))
    print(f"{out}: {len(r['xyz'])} atoms")


if __name__ == "__main__":
    main()
