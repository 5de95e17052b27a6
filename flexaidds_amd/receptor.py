"""Receptor PDB files into the receptor arrays the contact function takes.

For a target file the reference runs (LIB/top.cpp:582-646, then :856):

  rna_structure      more than half of the ATOM records nucleic -> RNA target
                     (LIB/rna_structure.cpp:20-52)
  modify_pdb         'A'/blank alternate locations only; protein residues
                     re-ordered by atom name and renumbered (atoms not in the
                     name list, hydrogens among them, are dropped); non-natural
                     ATOM residues become HETATM; HETATM and water filters
                     (LIB/modify_pdb.cpp:29-232)
  read_pdb           ATOM/HETATM -> atoms and residues (read_coor), CONECT ->
                     bonds (read_conect) (LIB/read_pdb.cpp:10-86,
                     LIB/read_coor.cpp:8-174, LIB/read_conect.cpp:6-68)
  calc_center        centre of geometry, bounds, widest extent
                     (LIB/calc_center.cpp:8-62)
  residue_conect     protein only: bonds from the type-definition file
                     (LIB/residue_conect.cpp:8-176)
  assign_types       types from the definition file (AMINO.def or
                     NUCLEOTIDES.def), water -> 1, metal ions by residue name
                     (LIB/assign_types.cpp:7-140)
  assign_radii_types radii by type for non-HETATM residues
                     (LIB/assign_radii_types.cpp:7-151)

The reference works on C strings read with fgets into fixed buffers; the
parts of that which decide results are kept: lines longer than a buffer are
split into several records, a field that does not convert leaves the previous
value, and read_coor's type column (atoi at column 77) reads what the line
buffer still holds from a longer previous line when the current line is
shorter.  HETATM residues keep radius 0, as the reference leaves them; type
40 (solvent) keeps radius 0 as well.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

_libc = ctypes.CDLL(None)
_strtof = _libc.strtof
_strtof.restype = ctypes.c_float
_strtof.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_void_p)]

MAX_ATOM_NUMBER = 100000   # LIB/flexaid.h:64: PDB serials index num_atm[]
_WS = b" \t\n\v\f\r"

NUCLEIC = (b"  A", b"  U", b"  C", b"  G")
NUCLEIC_REV = (b"A  ", b"U  ", b"C  ", b"G  ")
AMINO = (b"GLY", b"ALA", b"VAL", b"LEU", b"ILE", b"MET", b"ASN", b"PRO", b"CYS", b"SER",
         b"THR", b"GLN", b"ASP", b"GLU", b"LYS", b"ARG", b"HIS", b"PHE", b"TRP", b"TYR")
# protein atoms in the order modify_pdb writes them (LIB/modify_pdb.cpp:19-26)
PROTEIN_ATOM_ORDER = (b" N  ", b" CA ", b" C  ", b" O  ", b" CB ", b" CG ", b" SG ", b" OG ",
                      b" CG1", b" OG1", b" CG2", b" CD ", b" SD ", b" OD1", b" CD1", b" ND1",
                      b" OD2", b" ND2", b" CD2", b" CE ", b" NE ", b" CE1", b" NE1", b" OE1",
                      b" OE2", b" NE2", b" CE2", b" CE3", b" CZ ", b" NZ ", b" CZ1", b" CZ2",
                      b" CZ3", b" CH ", b" CH1", b" CH2", b" OH ", b" NH1", b" NH2", b" OXT")
_ATOM_RANK = {nm: k for k, nm in enumerate(PROTEIN_ATOM_ORDER)}
ION_RESNAMES = (b"MG ", b"ZN ", b"CA ", b"NA ", b"K  ", b"FE ", b"FE2", b"FE3", b"CU ", b"CU1",
                b"CU2", b"MN ", b"CO ", b"NI ", b"CL ", b"BR ", b"IOD", b"LI ", b"CD ", b"HG ",
                b"PB ")
# SYBYL types of metal-ion residues (LIB/assign_types.cpp:112-138)
ION_TYPES = ((b"MG ", 28), (b"CU ", 30), (b"CU1", 30), (b"CU2", 30), (b"MN ", 31), (b"HG ", 32),
             (b"CD ", 33), (b"NI ", 34), (b"ZN ", 35), (b"CA ", 36), (b"FE ", 37), (b"FE2", 37),
             (b"FE3", 37))
# radius by type for receptor residues (LIB/assign_radii_types.cpp:19-146);
# type 2 is 1.72 below three bonds, else 1.61; type 40 and unknown types keep
# their radius
_TYPE_RADIUS = {1: 1.88, 3: 1.88, 4: 1.76, 5: 1.88, 6: 1.64, 7: 1.64, 8: 1.64, 9: 1.64,
                10: 1.64, 11: 1.64, 12: 1.64, 13: 1.42, 14: 1.46, 15: 1.46, 16: 1.46,
                17: 1.782, 18: 1.782, 19: 1.782, 20: 1.782, 21: 1.782, 22: 1.871, 23: 1.560,
                24: 1.735, 25: 1.978, 26: 2.094, 27: 1.9, 28: 0.72, 29: 1.18, 30: 1.18,
                31: 0.73, 32: 1.02, 33: 0.95, 34: 0.69, 35: 0.74, 36: 1.00, 37: 0.61,
                38: 0.65, 39: 2.0}


# ------------------------------------------------------------ C string helpers
def _fgets(data: bytes, size: int) -> list:
    """The successive strings fgets(buf, size, f) returns on the file's bytes."""
    out, i, n, m = [], 0, len(data), size - 1
    while i < n:
        j = data.find(b"\n", i, i + m)
        j = min(i + m, n) if j < 0 else j + 1
        out.append(data[i:j])
        i = j
    return out


def _cstr(b: bytes) -> bytes:
    k = b.find(b"\0")
    return b if k < 0 else b[:k]


def _scan_int(b: bytes):
    """sscanf("%d"): the value, or None when nothing converts."""
    s = _cstr(b).lstrip(_WS)
    k = 1 if s[:1] in (b"+", b"-") else 0
    d = k
    while d < len(s) and 48 <= s[d] <= 57:
        d += 1
    return int(s[:d]) if d > k else None


def _atoi(b: bytes) -> int:
    v = _scan_int(b)
    return 0 if v is None else v


def _scan_float(b: bytes):
    """sscanf("%f"): the value rounded once to float, or None."""
    buf = ctypes.create_string_buffer(_cstr(b))
    end = ctypes.c_void_p()
    v = _strtof(buf, ctypes.byref(end))
    return None if end.value == ctypes.addressof(buf) else np.float32(v)


# ------------------------------------------------------------ the chain
def rna_structure(data: bytes) -> bool:
    """LIB/rna_structure.cpp:20-52 (80-byte line buffer)."""
    n_atoms = n_na = 0
    for ln in _fgets(data, 80):
        if ln[:6] == b"ATOM  ":
            res = _cstr(ln[17:20])
            if res in AMINO:
                pass
            elif res in NUCLEIC or res in NUCLEIC_REV:
                n_na += 1
            n_atoms += 1
    if n_atoms == 0:
        return False
    return bool(np.float32(n_na) / np.float32(n_atoms) > np.float32(0.5))


def modify_pdb(data: bytes, exclude_het=0, remove_water=1, is_protein=1, keep_ions=0,
               keep_structural_waters=0, structural_water_bfactor_max=20.0) -> bytes:
    """LIB/modify_pdb.cpp:29-232: the rewritten file's bytes."""
    out = []
    lines = []            # buffered protein residue (at most 50 lines)
    wrote = [0]
    prev_resnum, resnum = -1, -1
    prev_insert = b"-"

    def flush():
        # rewrite_residue2 / get_NextLine: smallest rank first, first line on
        # ties; names outside the list are never written
        left = [(_ATOM_RANK.get(_cstr(ln[12:16]), len(PROTEIN_ATOM_ORDER)), i, ln)
                for i, ln in enumerate(lines)]
        for rank, _, ln in sorted(left):
            if rank >= len(PROTEIN_ATOM_ORDER):
                break
            wrote[0] += 1
            out.append(ln[:6] + (b"%5d" % wrote[0])[:5] + ln[11:99])
        lines.clear()

    def natural_nucleic(res):
        return res in NUCLEIC or res in NUCLEIC_REV

    def place(ln, res, check_protein_for_nucleic=True):
        if is_protein and res in AMINO:
            if len(lines) < 50:
                lines.append(ln[:99])
        elif natural_nucleic(res) and (not is_protein or not check_protein_for_nucleic):
            out.append(ln)
        else:
            out.append(b"HETATM" + ln[6:])

    for ln in _fgets(data, 100):
        if ln[:6] == b"ATOM  ":
            res = _cstr(ln[17:20])
            v = _scan_int(ln[22:26])
            if v is not None:
                resnum = v
            insert = ln[26:27]
            if ln[16:17] not in (b" ", b"A"):
                continue
            if resnum == prev_resnum and insert == prev_insert:
                place(ln, res)
            elif prev_resnum != -1:
                if is_protein and lines:
                    flush()
                place(ln, res)
            else:
                # the first ATOM record: nucleic residues pass for proteins too
                place(ln, res, check_protein_for_nucleic=False)
            prev_resnum, prev_insert = resnum, insert
        else:
            if ln[:6] == b"HETATM":
                if exclude_het:
                    if not (keep_ions and _cstr(ln[17:20]) in ION_RESNAMES):
                        continue
                elif ln[17:20] == b"HOH" and remove_water:
                    if keep_structural_waters:
                        bf = _scan_float(ln[60:66])
                        if (0.0 if bf is None else float(bf)) > structural_water_bfactor_max:
                            continue
                    else:
                        continue
            if is_protein and lines:
                flush()
            out.append(ln)
    if is_protein and lines:
        flush()
    return b"".join(out)


@dataclass
class ReceptorFile:
    """Atoms 0..n-1 in the reference's internal order (its atoms 1..n)."""
    xyz: np.ndarray        # (n,3) float32
    radius: np.ndarray     # (n,) float32
    type: np.ndarray       # (n,) int32, 1-based
    number: np.ndarray     # (n,) int32, PDB serial as read
    name: list             # atom names (4 characters)
    ofres: np.ndarray      # (n,) int32, 0-based residue index
    nbond: np.ndarray      # (n,) int32, bond[0]
    res_name: list
    res_type: np.ndarray   # 0 = ATOM, 1 = HETATM
    res_first: np.ndarray  # first atom (0-based)
    res_last: np.ndarray   # last atom (0-based)
    is_protein: bool
    ori: np.ndarray        # (3,) float32, centre of geometry
    globalmin: np.ndarray  # (3,) float32
    globalmax: np.ndarray  # (3,) float32
    maxdst: float
    maxwidth: float

    @property
    def n_atoms(self) -> int:
        return int(self.xyz.shape[0])

    def receptor(self, center=None):
        """The contact function's receptor arrays; `center` is the cavity
        centre poses are drawn around (default: the centre of geometry)."""
        from .synth import Receptor
        c = np.asarray(self.ori if center is None else center, dtype=np.float64)
        return Receptor(xyz=self.xyz.copy(), radius=self.radius.copy(),
                        type=self.type.copy(), center=c)


class _Atoms:
    def __init__(self):
        # index 0 is the reference's unused atom 0 (a bad CONECT writes there)
        self.name, self.number, self.coor = [b""], [0], [np.zeros(3, np.float32)]
        self.type, self.ofres, self.bond = [0], [0], [[0] * 7]


def read_pdb(data: bytes, ntypes: int):
    """LIB/read_pdb.cpp:10-86 with read_coor and read_conect; then the last
    residue's end (LIB/top.cpp:635).  Returns (atoms, residues)."""
    A = _Atoms()
    R = {"name": [b"   "], "type": [0], "fatm": [0], "latm": [0]}
    num_atm = {}
    res_numold = b"-999"
    buf = bytearray(85)   # the line buffer keeps bytes of longer earlier lines
    for ln in _fgets(data, 85):
        buf[:len(ln)] = ln
        buf[len(ln)] = 0
        line = bytes(buf)
        rec = _cstr(line[:6])
        if rec in (b"ATOM  ", b"HETATM"):
            if line[16:17] not in (b" ", b"A"):
                continue
            a = len(A.name)
            A.name.append(_cstr(line[12:16]))
            t = _atoi(line[76:])
            A.type.append(t if 1 <= t <= ntypes else ntypes - 1)
            num = _scan_int(line[6:11])
            num = 4 if num is None else num   # the counter the reference leaves there
            if not 0 <= num < MAX_ATOM_NUMBER:
                raise ValueError(f"PDB atom serial {num} outside 0..{MAX_ATOM_NUMBER - 1}")
            A.number.append(num)
            num_atm[num] = a
            xyz = np.zeros(3, np.float32)
            for j in range(3):
                v = _scan_float(line[30 + 8 * j:38 + 8 * j])
                if v is not None:
                    xyz[j] = v
            A.coor.append(xyz)
            A.bond.append([0] * 7)
            res_new = _cstr(line[17:20])
            res_num = _cstr(line[22:26])
            if res_new != R["name"][-1] or res_num != res_numold:
                R["latm"][-1] = a - 1
                R["name"].append(res_new)
                R["type"].append(0 if rec == b"ATOM  " else 1)
                R["fatm"].append(a)
                R["latm"].append(a)
                res_numold = res_num
            A.ofres.append(len(R["name"]) - 1)
        elif rec == b"CONECT":
            _read_conect(line, A, num_atm)
    R["latm"][-1] = len(A.name) - 1
    return A, R


def _read_conect(line: bytes, A: _Atoms, num_atm: dict):
    """LIB/read_conect.cpp:6-68: up to four partners; bond[i+1] is written at
    the partner's column, bond[0] counts the new ones."""
    s = _cstr(line)
    ncont = max(len(s) - 12, 0) // 5
    atm = _scan_int(s[6:11])
    if atm is None:
        return
    tmp = [0, 0, 0, 0]
    for i in range(min(ncont, 4)):
        v = _scan_int(s[11 + 5 * i:16 + 5 * i])
        if v is not None:
            tmp[i] = v
    me = A.bond[num_atm.get(atm, 0)]
    for i in range(4):
        p = num_atm.get(tmp[i], 0) if tmp[i] != 0 else 0
        if p == 0:
            break
        if p not in me[1:1 + min(me[0], 6)]:
            me[i + 1] = p
            me[0] += 1


def calc_center(A: _Atoms, R: dict):
    """LIB/calc_center.cpp:8-62 in single precision."""
    n = len(A.name) - 1
    gmin = A.coor[1].copy()
    gmax = A.coor[1].copy()
    ori = np.zeros(3, np.float32)
    for j in range(3):
        acc = np.float32(0.0)
        tot = 0
        for k in range(1, len(R["name"])):
            for i in range(R["fatm"][k], R["latm"][k] + 1):
                c = A.coor[i][j]
                acc = np.float32(acc + c)
                if c < gmin[j]:
                    gmin[j] = c
                if c > gmax[j]:
                    gmax[j] = c
                tot += 1
        ori[j] = np.float32(acc / np.float32(tot))
    maxdst = np.float32(0.0)
    for i in range(1, n + 1):
        dx = np.float32(ori[0] - A.coor[i][0])
        dy = np.float32(ori[1] - A.coor[i][1])
        dz = np.float32(ori[2] - A.coor[i][2])
        r2 = np.float32(np.float32(np.float32(dx * dx) + np.float32(dy * dy)) + np.float32(dz * dz))
        d = np.float32(np.sqrt(r2))
        if d >= maxdst:
            maxdst = d
    width = np.float32(0.0)
    for j in range(3):
        w = np.float32(gmax[j] - gmin[j])
        if w > width:
            width = w
    return ori, gmin, gmax, float(maxdst), float(width)


def residue_conect(A: _Atoms, R: dict, deffile: bytes):
    """LIB/residue_conect.cpp:8-176: bonds of complete residues from the
    definition file's CONECT rows, plus the peptide links to the previous
    residue's C and the next residue's N."""
    conect = [[0] * 5 for _ in range(26)]
    natm, rnam, flag = 0, b"", 0
    nres = len(R["name"]) - 1
    for ln in _fgets(deffile, 81):
        field = _cstr(ln[:6])
        if field in (b"RESIDU", b"NUCLEO"):
            flag, natm = 0, 0
            rnam = _cstr(ln[7:10])
        if field == b"CONECT":
            natm += 1
            if natm > 24:
                raise ValueError("definition file: more than 24 CONECT rows in a residue")
            row = conect[natm]
            row[0] = 0
            for j in range(4):
                v = _scan_int(ln[10 + 3 * j:12 + 3 * j])
                atn = 0 if v is None else v
                if atn == 0:
                    break
                row[j + 1] = atn
                row[0] += 1
        if field == b"RESEND":
            flag = 1
        if flag:
            for i in range(1, nres + 1):
                if R["name"][i] != rnam:
                    continue
                f, last = R["fatm"][i], R["latm"][i]
                if last - f + 1 != natm:
                    continue   # "misses atoms": left as read
                for l in range(1, natm + 1):
                    b = A.bond[f + l - 1]
                    b[0] = conect[l][0]
                    for j in range(1, b[0] + 1):
                        b[j] = conect[l][j] + f - 1
                if i != 1 and R["type"][i - 1] == 0:
                    for j in range(R["fatm"][i - 1], R["latm"][i - 1] + 1):
                        if A.name[j] == b" C  ":
                            _push_bond(A.bond[f], j)
                            break
                if i != nres and R["type"][i + 1] == 0:
                    for j in range(R["fatm"][i + 1], R["latm"][i + 1] + 1):
                        if A.name[j] == b" N  ":
                            _push_bond(A.bond[f + 2], j)
                            break
            flag = 0


def _push_bond(b: list, j: int):
    if b[0] + 1 > 6:
        raise ValueError("an atom with more than six bonds")
    b[b[0] + 1] = j
    b[0] += 1


def assign_types(A: _Atoms, R: dict, deffile: bytes, ntypes: int, is_protein: bool):
    """LIB/assign_types.cpp:7-140.  Later definition lines win; RNA targets
    also match the reversed residue name."""
    table = {}
    rnam = rnamrev = b""
    typ = 0
    for ln in _fgets(deffile, 81):
        field = _cstr(ln[:6])
        if field in (b"RESIDU", b"NUCLEO"):
            rnam = _cstr(ln[7:10])
            rnamrev = rnam[::-1]
        if field == b"ATMTYP":
            v = _scan_int(ln[10:12])
            if v is not None:
                typ = v
            anam = _cstr(ln[12:16])
            if typ > ntypes:
                typ = 6   # "set to neutral"; the variable keeps it
            table[(rnam, anam)] = typ
            if not is_protein:
                table[(rnamrev, anam)] = typ
    n = len(A.name) - 1
    for i in range(1, n + 1):
        t = table.get((R["name"][A.ofres[i]], A.name[i]))
        if t is not None:
            A.type[i] = t
    for i in range(1, n + 1):
        if R["name"][A.ofres[i]] == b"HOH":
            A.type[i] = 1
    for k in range(1, len(R["name"])):
        if R["type"][k] != 1:
            continue
        nm = R["name"][k]
        sy = next((s for r, s in ION_TYPES if nm[:3] == r), 0)
        if 0 < sy <= ntypes:
            for i in range(R["fatm"][k], R["latm"][k] + 1):
                A.type[i] = sy


def assign_radii_types(A: _Atoms, R: dict) -> np.ndarray:
    """LIB/assign_radii_types.cpp:7-151: radii of non-HETATM residues."""
    n = len(A.name) - 1
    rad = np.zeros(n + 1, np.float32)
    for k in range(1, len(R["name"])):
        if R["type"][k] == 1:
            continue
        for j in range(R["fatm"][k], R["latm"][k] + 1):
            t = A.type[j]
            if t == 2:
                rad[j] = np.float32(1.72 if 3 - A.bond[j][0] > 0 else 1.61)
            elif t in _TYPE_RADIUS:
                rad[j] = np.float32(_TYPE_RADIUS[t])
    return rad[1:]


def read_receptor(pdb: str | bytes, amino_def: str | bytes, nucleotides_def: str | bytes,
                  ntypes: int = 40, *, exclude_het=0, remove_water=1, keep_ions=0,
                  keep_structural_waters=0, structural_water_bfactor_max=20.0) -> ReceptorFile:
    """A target PDB file (path or bytes) and the two type-definition files
    (AMINO.def, NUCLEOTIDES.def; paths or bytes) -> ReceptorFile.  ntypes is
    the energy matrix's type count."""
    data = _load(pdb)
    is_protein = not rna_structure(data)
    deffile = _load(amino_def if is_protein else nucleotides_def)
    mod = modify_pdb(data, exclude_het, remove_water, int(is_protein), keep_ions,
                     keep_structural_waters, structural_water_bfactor_max)
    A, R = read_pdb(mod, ntypes)
    n = len(A.name) - 1
    if n == 0:
        raise ValueError("receptor file: no ATOM or HETATM records")
    ori, gmin, gmax, maxdst, maxwidth = calc_center(A, R)
    if is_protein:
        residue_conect(A, R, deffile)
    assign_types(A, R, deffile, ntypes, is_protein)
    radius = assign_radii_types(A, R)
    return ReceptorFile(
        xyz=np.array(A.coor[1:], np.float32).reshape(n, 3), radius=radius,
        type=np.array(A.type[1:], np.int32), number=np.array(A.number[1:], np.int32),
        name=[x.decode("latin-1") for x in A.name[1:]],
        ofres=np.array(A.ofres[1:], np.int32) - 1, nbond=np.array([b[0] for b in A.bond[1:]], np.int32),
        res_name=[x.decode("latin-1") for x in R["name"][1:]],
        res_type=np.array(R["type"][1:], np.int32),
        res_first=np.array(R["fatm"][1:], np.int32) - 1, res_last=np.array(R["latm"][1:], np.int32) - 1,
        is_protein=is_protein, ori=ori, globalmin=gmin, globalmax=gmax, maxdst=maxdst, maxwidth=maxwidth)


def _load(src) -> bytes:
    if isinstance(src, (bytes, bytearray)):
        return bytes(src)
    with open(src, "rb") as f:
        return f.read()
