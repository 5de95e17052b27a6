"""MOL2 test inputs (the reference ships none): random tree ligands with
the features the reader has rules for.  Test infrastructure."""
from __future__ import annotations

import numpy as np

from flexaidds_amd import ligand

SYBYL = [k.decode() for k in ligand.SYBYL_TYPE] + ["C.xx", "Du", "Se", "K"]


def write_mol2(path, rng, n_atoms, extra_bonds=3, long_name=False, long_line=False, seventh_bond=False,
               bad_bond=False, short_records=False, origin=None):
    """A MOL2 file of a random tree ligand (1.5 A bonds) with some
    ring-closing bonds; ids shuffled and spaced, records in id order of the
    shuffle; centred near `origin` (default: a random point within 20 A)."""
    xyz = [np.zeros(3)]
    parent = [-1]
    for i in range(1, n_atoms):
        p = int(rng.integers(0, i))
        v = rng.normal(size=3)
        xyz.append(xyz[p] + 1.5 * v / np.linalg.norm(v))
        parent.append(p)
    xyz = np.array(xyz)
    xyz = xyz - xyz.mean(axis=0) + (rng.uniform(-20, 20, size=3) if origin is None else np.asarray(origin))
    ids = rng.permutation(n_atoms) * 3 + 5               # non-contiguous ids
    bonds = [(i, parent[i]) for i in range(1, n_atoms)]
    for _ in range(extra_bonds):
        a, b = rng.choice(n_atoms, 2, replace=False)
        bonds.append((int(a), int(b)))
    if seventh_bond:                                      # atom 0 bonded to seven atoms
        for b in range(1, min(8, n_atoms)):
            if (b, 0) not in bonds and (0, b) not in bonds:
                bonds.append((0, b))
    rng.shuffle(bonds)
    lines = ["# generated", "@<TRIPOS>MOLECULE", "LIGXYZ test", f"{n_atoms} {len(bonds)} 1", "SMALL", "",
             "@<TRIPOS>ATOM"]
    for k in range(n_atoms):
        t = SYBYL[int(rng.integers(0, len(SYBYL)))]
        name = f"A{k}"
        if long_name and k == 3:
            name = "ABCDEFGHIJ"                              # %7s takes seven, the rest breaks the record
        x, y, z = xyz[k]
        rec = f"{ids[k]:7d} {name:<6s} {x:10.4f} {y:10.4f} {z:10.4f} {t:<6s}"
        if not (short_records and k % 4 == 1):
            rec += f" {1:4d} LIG {rng.normal() * 0.3:8.4f}"
        if long_line and k == 5:
            rec += " " * 300 + "0.0"                          # split by the 256-byte buffer
        lines.append(rec)
    lines.append("@<TRIPOS>BOND")
    for q, (a, b) in enumerate(bonds):
        lines.append(f"{q + 1:6d} {ids[a]:5d} {ids[b]:5d} {'ar' if q % 5 == 0 else '1'}")
    if bad_bond:
        lines.append(f"{len(bonds) + 1:6d} {ids[0]:5d} 99999 1")
    lines.append("@<TRIPOS>SUBSTRUCTURE")
    lines.append("     1 LIG         1")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")




ELEMENTS = [k.decode() for k in ligand.ELEMENT_TYPE] + ["Se", "B", "Si"]


def write_sdf(path, rng, n_atoms, extra_bonds=3, origin=None, tail_symbol=False, bad_bond=False, seventh_bond=False,
              short_coord=False):
    """An MDL V2000 file of a random tree ligand: name, program and comment
    lines, the counts line, the atom block (coordinates in 10-character
    columns, the symbol at column 31), the bond block, M  END."""
    xyz = [np.zeros(3)]
    parent = [-1]
    for i in range(1, n_atoms):
        p = int(rng.integers(0, i))
        v = rng.normal(size=3)
        xyz.append(xyz[p] + 1.5 * v / np.linalg.norm(v))
        parent.append(p)
    xyz = np.array(xyz)
    xyz = xyz - xyz.mean(axis=0) + (rng.uniform(-20, 20, size=3) if origin is None else np.asarray(origin))
    bonds = [(i, parent[i]) for i in range(1, n_atoms)]
    for _ in range(extra_bonds):
        a, b = rng.choice(n_atoms, 2, replace=False)
        bonds.append((int(a), int(b)))
    if seventh_bond:
        for b in range(1, min(8, n_atoms)):
            if (b, 0) not in bonds and (0, b) not in bonds:
                bonds.append((0, b))
    rng.shuffle(bonds)
    if bad_bond:
        bonds.append((0, n_atoms + 5))
    lines = ["LIGSDF generated", "  graft  3D", "", f"{n_atoms:3d}{len(bonds):3d}  0  0  0  0  0  0  0  0999 V2000"]
    for k in range(n_atoms):
        e = ELEMENTS[int(rng.integers(0, len(ELEMENTS)))]
        x, y, z = xyz[k]
        line = f"{x:10.4f}{y:10.4f}{z:10.4f} {e:<3s} 0  0  0  0  0  0  0  0  0  0  0  0"
        if tail_symbol and k % 3 == 0:
            line = f"{x:10.4f}{y:10.4f}{z:10.4f} {e}"          # the line ends right after the symbol
        if short_coord and k == 1:
            line = f"{x:10.4f}       abc{z:10.4f} {e:<3s} 0  0"  # y does not convert
        lines.append(line)
    for a, b in bonds:
        lines.append(f"{a + 1:3d}{b + 1:3d}  1  0  0  0  0")
    lines += ["M  END", "$$$$"]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
