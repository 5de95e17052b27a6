"""Receptor PDB reader (flexaidds_amd/receptor.py) against the reference's own
receptor-file chain compiled from its sources (oracle/_ref/libfa_ref_reader.so:
rna_structure, modify_pdb, read_pdb, calc_center, residue_conect,
assign_types, assign_radii_types).  Skipped where the reference checkout (its
type-definition files) or that build is absent.

Inputs: the reference's RNA file guanine_test_INI.pdb, and generated protein
and RNA files that exercise the chain's rules: atom re-ordering and
renumbering, hydrogens and unknown names dropped, alternate locations,
insertion codes, a residue with a missing atom, OXT, chains, a non-natural
ATOM residue turned HETATM, waters with B-factors, ions, a HETATM ligand with
type columns and CONECT records, a short line after a long one (the type
column read from the line buffer's leftover bytes), reversed nucleotide names.
"""
import os

import numpy as np
import pytest

import oracle_lib as ol
from flexaidds_amd import receptor as rc

REF_DIR = os.environ.get("FA_REFERENCE_DIR", "/root/reference")
AMINO = os.path.join(REF_DIR, "WRK", "AMINO.def")
NUCL = os.path.join(REF_DIR, "WRK", "NUCLEOTIDES.def")
GUANINE = os.path.join(REF_DIR, "guanine_test_INI.pdb")

lib = ol.load_ref_reader()
pytestmark = pytest.mark.skipif(lib is None or not os.path.exists(AMINO),
                                reason="reference checkout or oracle/_ref reader build absent")

FLAGS = [
    dict(),
    dict(exclude_het=1, keep_ions=1),
    dict(exclude_het=1, keep_ions=0),
    dict(remove_water=1, keep_structural_waters=1, structural_water_bfactor_max=20.0),
    dict(remove_water=0),
]


def _def_residues(path):
    """Residue name -> atom names in definition order."""
    out, cur = {}, None
    with open(path) as f:
        for ln in f:
            if ln.startswith(("RESIDU", "NUCLEO")):
                cur = ln[7:10]
                out[cur] = []
            elif ln.startswith("ATMTYP") and cur is not None:
                out[cur].append(ln[12:16])
    return out


def _line(rec, serial, name, alt, resn, chain, resnum, ins, xyz, b=20.0, elem=" C"):
    x, y, z = xyz
    return (f"{rec:<6}{serial:5d} {name:4}{alt:1}{resn:>3} {chain:1}{resnum:4d}{ins:1}   "
            f"{x:8.3f}{y:8.3f}{z:8.3f}{1.0:6.2f}{b:6.2f}          {elem:>2}  \n")


def make_protein_pdb(rng, n_res=36, lead_nucleic=False):
    res_def = _def_residues(AMINO)
    names = sorted(res_def)
    out = ["REMARK generated protein\n"]
    serial = 0

    def xyz():
        return rng.uniform(-15.0, 15.0, 3)

    if lead_nucleic:   # a nucleotide as the first ATOM record stays ATOM
        for a in _def_residues(NUCL)["  A"][:6]:
            serial += 1
            out.append(_line("ATOM", serial, a, " ", "  A", "N", 1, " ", xyz()))

    prev_num = 0
    for k in range(n_res):
        chain = "A" if k < n_res // 2 else "B"
        resn = names[int(rng.integers(len(names)))]
        atoms = list(res_def[resn])
        if k % 7 == 3 and len(atoms) > 5:
            atoms.pop(int(rng.integers(5, len(atoms))))          # a missing side-chain atom
        if k % 5 == 1:
            atoms[1:1] = [" H  ", "1HB ", " HA "]                 # dropped by the re-ordering
        if k % 3 == 0:
            atoms = [atoms[i] for i in rng.permutation(len(atoms))]
        if k == 5:
            atoms.append(atoms[0])                                  # a repeated name: file order on ties
        num, ins = (prev_num, "A") if k == 10 else (prev_num + 1, " ")
        if k == n_res // 2:
            num = 1
        prev_num = num
        for a in atoms:
            if k % 6 == 2 and a == " CB ":
                for alt in "AB":
                    serial += 1
                    out.append(_line("ATOM", serial, a, alt, resn, chain, num, ins, xyz()))
                continue
            serial += 1
            out.append(_line("ATOM", serial, a, " ", resn, chain, num, ins, xyz()))
        if k == n_res // 2 - 1:
            serial += 1
            out.append(_line("ATOM", serial, " OXT", " ", resn, chain, num, ins, xyz()))
            out.append("TER\n")
    # a non-natural residue in ATOM records
    for a in (" N  ", " CA ", " C  ", " O  ", " SE "):
        serial += 1
        out.append(_line("ATOM", serial, a, " ", "MSE", "B", prev_num + 1, " ", xyz()))
    out.append("TER\n")
    # waters, ions, a ligand with type columns and bonds
    for b in (10.0, 35.0, 15.0):
        serial += 1
        out.append(_line("HETATM", serial, " O  ", " ", "HOH", "W", serial, " ", xyz(), b=b, elem=" O"))
    out.append(_line("HETATM", 0, " O  ", " ", "HOH", "W", 99, " ", xyz(), b=5.0).replace("    0", "*****", 1))
    # ion residue names as the reference's tables hold them ("ZN ") and as
    # PDB files right-justify them (" ZN", matched by neither table)
    for ion in ("ZN ", " ZN", "MG ", "NA ", "CA "):
        serial += 1
        ln = _line("HETATM", serial, ion.strip().ljust(4), " ", "XXX", "I", serial, " ", xyz(), elem=ion.strip())
        out.append(ln[:17] + ion + ln[20:])
    first = serial + 1
    for i, t in enumerate((12, 3, 3, 14, 2, 36, 7)):
        serial += 1
        out.append(_line("HETATM", serial, f" C{i}".ljust(4), " ", "LIG", "L", 900, " ", xyz(), elem=f"{t:2d}"))
    serial += 1
    short = _line("HETATM", serial, " C7 ", " ", "LIG", "L", 900, " ", xyz())[:66] + "\n"
    out.append(short)                                               # no type column of its own
    serial += 1
    out.append(_line("HETATM", serial, " C8 ", "B", "LIG", "L", 900, " ", xyz(), elem="3"))  # altloc B
    for i in range(first, serial - 1):
        out.append(f"CONECT{i:5d}{i + 1:5d}\n")
    out.append(f"CONECT{first:5d}{first + 2:5d}{first + 3:5d}{first + 4:5d}\n")
    out.append(f"CONECT{first + 1:5d}{first:5d}{first:5d}\n")        # a repeated partner
    out.append(f"CONECT{5:5d}{first:5d}\n")                           # serial from before the renumbering
    out.append("END\n")
    return "".join(out).encode()


def make_rna_pdb(rng, n_res=30):
    res_def = _def_residues(NUCL)
    names = sorted(res_def)
    amino = _def_residues(AMINO)
    out = ["REMARK generated RNA\n"]
    serial = 0
    for k in range(n_res):
        resn = names[int(rng.integers(len(names)))]
        atoms = list(res_def[resn])
        shown = resn[::-1] if k % 4 == 1 else resn                 # "G  " as well as "  G"
        if k % 9 == 4:
            atoms = atoms[:-2]
        for a in atoms:
            serial += 1
            out.append(_line("ATOM", serial, a, " ", shown, "R", k + 1, " ", rng.uniform(-20, 20, 3)))
    for a in amino["LYS"]:                                          # a protein residue in an RNA target
        serial += 1
        out.append(_line("ATOM", serial, a, " ", "LYS", "P", 1, " ", rng.uniform(-20, 20, 3)))
    serial += 1
    out.append(_line("HETATM", serial, " O  ", " ", "HOH", "W", 1, " ", rng.uniform(-20, 20, 3), b=5.0))
    serial += 1
    out.append(_line("HETATM", serial, "MG  ", " ", "MG", "I", 2, " ", rng.uniform(-20, 20, 3), elem="MG"))
    out.append("END\n")
    return "".join(out).encode()


def _compare(tmp_path, data, flags, tag):
    path = str(tmp_path / f"{tag}.pdb")
    with open(path, "wb") as f:
        f.write(data)
    ref = ol.ref_read_receptor(lib, path, AMINO, NUCL, str(tmp_path / f"{tag}_mod.pdb"), **flags)
    got = rc.read_receptor(data, AMINO, NUCL, 40, **flags)
    assert got.is_protein == ref["is_protein"]
    assert got.n_atoms == len(ref["xyz"]), (got.n_atoms, len(ref["xyz"]))
    for f in ("xyz", "radius", "type", "number", "ofres", "nbond"):
        a, b = getattr(got, f), ref[f]
        bad = np.nonzero((a != b).reshape(len(b), -1).any(axis=1))[0]
        assert bad.size == 0, (f, bad[:8], a[bad[:3]], b[bad[:3]])
    mine = np.concatenate([got.ori, got.globalmin, got.globalmax, [got.maxdst, got.maxwidth]]).astype(np.float32)
    assert np.array_equal(mine, ref["center"])
    return got


def test_reference_rna_file(tmp_path):
    """The reference's own target file: 1,097 atoms in 52 nucleotides whose
    left-justified atom names match no definition, so every atom keeps the
    dummy type 39 and its radius 2.0, as the reference reads it."""
    with open(GUANINE, "rb") as f:
        data = f.read()
    got = _compare(tmp_path, data, {}, "guanine")
    assert got.n_atoms == 1097 and not got.is_protein
    assert set(got.type.tolist()) == {39} and set(got.radius.tolist()) == {2.0}


@pytest.mark.parametrize("seed", [1, 2, 3])
@pytest.mark.parametrize("fi", range(len(FLAGS)))
def test_generated_protein(tmp_path, seed, fi):
    data = make_protein_pdb(np.random.default_rng(seed), lead_nucleic=seed == 3)
    got = _compare(tmp_path, data, FLAGS[fi], f"prot{seed}_{fi}")
    assert got.is_protein
    # the rules the file exercises took effect
    assert "MSE" in got.res_name and all(n.strip()[:1] != "H" or n == " HA " for n in got.name[:50])
    types = set(got.type.tolist())
    assert 2 in types and 11 in types and 13 in types
    het = got.res_type[got.ofres] == 1
    assert set(np.unique(got.radius[(got.type == 2) & ~het]).tolist()) <= {np.float32(1.72), np.float32(1.61)}
    assert np.all(got.radius[het] == 0.0)   # HETATM residues keep radius 0
    if FLAGS[fi].get("keep_ions", 1) or not FLAGS[fi].get("exclude_het"):
        assert 35 in types and 36 in types     # ZN, CA by residue name


@pytest.mark.parametrize("seed", [4, 5])
@pytest.mark.parametrize("fi", [0, 1])
def test_generated_rna(tmp_path, seed, fi):
    data = make_rna_pdb(np.random.default_rng(seed))
    got = _compare(tmp_path, data, FLAGS[fi], f"rna{seed}_{fi}")
    assert not got.is_protein
    assert (got.type != 39).sum() > got.n_atoms // 2


def test_receptor_arrays_feed_the_contact_function():
    """ReceptorFile.receptor(): the arrays the system upload takes."""
    data = make_rna_pdb(np.random.default_rng(7))
    got = rc.read_receptor(data, AMINO, NUCL, 40)
    r = got.receptor()
    assert r.xyz.dtype == np.float32 and r.xyz.shape == (got.n_atoms, 3)
    assert r.type.dtype == np.int32 and r.radius.dtype == np.float32
    assert np.allclose(r.center, got.ori)
