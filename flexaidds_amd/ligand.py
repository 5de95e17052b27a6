"""Ligand MOL2 (and SDF) files into the ligand descriptors the contact function takes.

For a run with a MOL2 ligand the reference takes these steps (direct mode,
LIB/top.cpp:845-897 and :982-998):

  read_mol2_ligand   TRIPOS sections; ATOM records -> coordinates, SYBYL
                     type -> FlexAID type and radius; BOND records -> bond
                     lists (six per atom at most); the reconstruction tree
                     by breadth-first search from the first atom (rec[0..2]:
                     parent, grandparent, great-grandparent, 0 = the ori
                     frame); buildic; the bonded matrix within bloops
                     (bondedlist + update_bonded) (LIB/Mol2Reader.cpp:104-391,
                     LIB/buildic.cpp:6-71, LIB/bondedlist.cpp:8-61,
                     LIB/update_bonded.cpp:5-94)
  assign_radii_types leaves the ligand's radii (ligand residues are skipped,
                     LIB/assign_radii_types.cpp:17)
  GPA block          ori = the ligand's centroid, the global-positioning
                     atoms = its first three atoms, buildic_point of the
                     first (LIB/top.cpp:864-897, LIB/buildic_point.cpp:6-31)
  cleft grid         detect_cleft over the receptor's atoms followed by the
                     ligand's, generate_grid with row 0 = the first GPA atom,
                     calc_cleftic in the ori frame (LIB/top.cpp:904-916,
                     LIB/generate_grid.cpp:24-28)
  ic_bounds + add2_optimiz_vec
                     the translation gene (typ -1 on gpa[0], grid rows
                     1..n-1) and three rotation genes (ang of gpa[1], dih of
                     gpa[1], dih of gpa[2], -180..180), after a second
                     buildic in the final ori frame (LIB/ic_bounds.cpp:6-35,
                     LIB/add2_optimiz_vec.cpp:108-176)
  create_rebuild_list
                     the rebuild order (buildlist with bnum 0: the GPA atoms,
                     then every atom once its tree ancestors are placed,
                     LIB/create_rebuild_list.cpp:8-50, LIB/buildlist.cpp)

A MOL2 ligand has no flexible dihedrals in this path (fdih stays 0); its
genes are the translation and the rotation.  The file is read as the
reference reads it: fgets into a 256-byte buffer (longer lines split), the
ATOM and BOND records parsed with sscanf's rules (a field that does not
convert ends the record, a %7s name longer than seven characters runs into
the next field).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import cleft
from .receptor import _fgets

F = np.float32

_libc = ctypes.CDLL(None)
_strtof = _libc.strtof
_strtof.restype = ctypes.c_float
_strtof.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_void_p)]
_strtol = _libc.strtol
_strtol.restype = ctypes.c_long
_strtol.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_void_p), ctypes.c_int]

BLOOPS = 2            # LIB/top.cpp:296
MAX_BONDS = 6         # bond[0] < 6 before a bond is added (LIB/Mol2Reader.cpp:309-310)
LINE_BUF = 256        # char buf[256] (LIB/Mol2Reader.cpp:135)

# SYBYL atom type -> FlexAID type (LIB/Mol2Reader.cpp:19-69); anything else 39
SYBYL_TYPE = {
    b"C.3": 1, b"C.2": 2, b"C.1": 2, b"C.ar": 3, b"C.cat": 2,
    b"N.3": 4, b"N.2": 5, b"N.1": 5, b"N.ar": 6, b"N.am": 7, b"N.pl3": 8, b"N.4": 9,
    b"O.3": 10, b"O.2": 11, b"O.co2": 12, b"O.spc": 10, b"O.t3p": 10,
    b"S.3": 16, b"S.2": 17, b"S.O": 18, b"S.O2": 19,
    b"P.3": 20,
    b"F": 13, b"Cl": 14, b"Br": 15, b"I": 21,
    b"H": 22, b"H.spc": 22, b"H.t3p": 22,
    b"Fe": 30, b"Zn": 31, b"Ca": 32, b"Mg": 33,
}
DUMMY_TYPE = 39
# radius by the type's first character, Cl and Br first (LIB/Mol2Reader.cpp:72-89)
_RADIUS_2 = {b"Cl": F(1.75), b"Br": F(1.85)}
_RADIUS_1 = {ord("C"): F(1.70), ord("N"): F(1.55), ord("O"): F(1.52), ord("S"): F(1.80), ord("P"): F(1.80),
             ord("F"): F(1.47), ord("H"): F(1.20), ord("I"): F(1.98)}
_WS = b" \t\n\v\f\r"


def sybyl_radius(t: bytes) -> np.float32:
    if t in _RADIUS_2:
        return _RADIUS_2[t]
    return _RADIUS_1.get(t[0] if t else 0, F(1.70))


# ------------------------------------------------------------ sscanf
class _Scan:
    """sscanf over one C string for the conversions the reader uses."""

    def __init__(self, s: bytes):
        k = s.find(b"\0")
        self.s = s if k < 0 else s[:k]
        self.i = 0

    def _ws(self):
        while self.i < len(self.s) and self.s[self.i] in _WS:
            self.i += 1
        return self.i < len(self.s)

    def d(self):
        if not self._ws():
            return None
        buf = ctypes.create_string_buffer(self.s[self.i:])
        end = ctypes.c_void_p()
        v = _strtol(buf, ctypes.byref(end), 10)
        used = end.value - ctypes.addressof(buf)
        if used == 0:
            return None
        self.i += used
        return int(ctypes.c_int(v).value)

    def f(self):
        if not self._ws():
            return None
        buf = ctypes.create_string_buffer(self.s[self.i:])
        end = ctypes.c_void_p()
        v = _strtof(buf, ctypes.byref(end))
        used = end.value - ctypes.addressof(buf)
        if used == 0:
            return None
        self.i += used
        return F(v)

    def s_(self, width):
        if not self._ws():
            return None
        j = self.i
        while j < len(self.s) and j - self.i < width and self.s[j] not in _WS:
            j += 1
        tok = self.s[self.i:j]
        self.i = j
        return tok


def _scan_atom(line: bytes):
    """sscanf(buf, "%d %7s %f %f %f %15s %d %7s %f"): (fields, count)."""
    sc = _Scan(line)
    out = []
    for kind, w in (("d", 0), ("s", 7), ("f", 0), ("f", 0), ("f", 0), ("s", 15), ("d", 0), ("s", 7), ("f", 0)):
        v = sc.d() if kind == "d" else sc.f() if kind == "f" else sc.s_(w)
        if v is None:
            break
        out.append(v)
    return out


def _scan_bond(line: bytes):
    """sscanf(buf, "%*d %d %d %7s"): (origin, target, type token) fields read."""
    sc = _Scan(line)
    if sc.d() is None:
        return []
    out = []
    for kind in ("d", "d", "s"):
        v = sc.d() if kind == "d" else sc.s_(7)
        if v is None:
            break
        out.append(v)
    return out


# ------------------------------------------------------------ the reader
@dataclass
class Mol2Ligand:
    """A ligand as read_mol2_ligand reads it (atoms in file order)."""
    name: bytes
    atom_names: list
    sybyl: list
    xyz: np.ndarray          # (n, 3) float32
    type: np.ndarray         # (n,) int32
    radius: np.ndarray       # (n,) float32
    bonds: list = field(default_factory=list)  # per atom, bonded atom indices in record order (<= 6)

    @property
    def n_atoms(self) -> int:
        return int(self.xyz.shape[0])


def read_mol2(src) -> Mol2Ligand:
    """LIB/Mol2Reader.cpp:104-311: the first molecule's ATOM and BOND
    records (a later @<TRIPOS>ATOM section appends to the same atom list, as
    in the reference)."""
    data = src if isinstance(src, (bytes, bytearray)) else open(src, "rb").read()
    name = b"LIG"
    atoms, bonds = [], []
    sec, mol_line = None, 0
    for raw in _fgets(bytes(data), LINE_BUF):
        line = raw.rstrip(b"\r\n")
        tag = None
        if b"@<TRIPOS>MOLECULE" in line:
            tag = "MOLECULE"
        elif b"@<TRIPOS>ATOM" in line:
            tag = "ATOM"
        elif b"@<TRIPOS>BOND" in line:
            tag = "BOND"
        elif line[:1] == b"@":
            tag = "OTHER"
        if tag is not None:
            sec, mol_line = tag, 0
            continue
        if sec == "MOLECULE":
            mol_line += 1
            if mol_line == 1:
                tok = _Scan(line).s_(63)
                if tok is not None:
                    name = tok
        elif sec == "ATOM":
            v = _scan_atom(line)
            if len(v) >= 6:
                atoms.append(v)
        elif sec == "BOND":
            v = _scan_bond(line)
            if len(v) >= 2:
                bonds.append((v[0], v[1]))
    if not atoms:
        raise ValueError("no atoms found in the MOL2 file")
    n = len(atoms)
    xyz = np.array([[a[2], a[3], a[4]] for a in atoms], F)
    sybyl = [a[5] for a in atoms]
    typ = np.array([SYBYL_TYPE.get(t, DUMMY_TYPE) for t in sybyl], np.int32)
    rad = np.array([sybyl_radius(t) for t in sybyl], F)
    id_map = {}
    for k, a in enumerate(atoms):
        id_map[a[0]] = k           # a repeated id maps to its last atom, as std::map assignment does
    blist = [[] for _ in range(n)]
    for o, t in bonds:
        if o not in id_map or t not in id_map:
            continue
        io, it = id_map[o], id_map[t]
        if len(blist[io]) < MAX_BONDS:
            blist[io].append(it)
        if len(blist[it]) < MAX_BONDS:
            blist[it].append(io)
    return Mol2Ligand(name=name[:3], atom_names=[a[1] for a in atoms], sybyl=sybyl, xyz=xyz, type=typ,
                      radius=rad, bonds=blist)


# element -> FlexAID type and radius of the SDF reader (LIB/SdfReader.cpp:26-60)
ELEMENT_TYPE = {b"C": 1, b"N": 4, b"O": 10, b"S": 16, b"P": 20, b"F": 13, b"Cl": 14, b"Br": 15, b"I": 21,
                b"H": 22, b"Fe": 30, b"Zn": 31, b"Ca": 32, b"Mg": 33}
ELEMENT_RADIUS = {b"C": F(1.70), b"N": F(1.55), b"O": F(1.52), b"S": F(1.80), b"P": F(1.80), b"F": F(1.47),
                  b"Cl": F(1.75), b"Br": F(1.85), b"I": F(1.98), b"H": F(1.20)}


def _atoi3(b: bytes) -> int:
    sc = _Scan(b[:3])
    v = sc.d()
    return 0 if v is None else v


def read_sdf(src) -> Mol2Ligand:
    """LIB/SdfReader.cpp:63-263: the first molecule of an MDL V2000 file --
    the name (line 1), the counts (atoms and bonds in the first two
    three-character columns of line 4), one atom per line (x y z read with
    sscanf, the symbol from columns 31-33 without blanks; the line keeps
    its newline there, which a symbol ending at column 33 takes in), one bond
    per line (three-character columns).  Types and radii come from the
    element; bonds naming an atom outside 1..n are skipped, six per atom at
    most."""
    data = src if isinstance(src, (bytes, bytearray)) else open(src, "rb").read()
    lines = _fgets(bytes(data), LINE_BUF)
    if len(lines) < 4:
        raise ValueError("SDF file ends before its counts line")
    name = b"LIG"
    first = lines[0][:-1] if lines[0].endswith(b"\n") else lines[0]
    tok = _Scan(first).s_(63)
    if tok is not None:
        name = tok
    natoms, nbonds = _atoi3(lines[3][0:3]), _atoi3(lines[3][3:6])
    if natoms <= 0 or natoms > 9999 or nbonds < 0 or nbonds > 9999:
        raise ValueError(f"invalid SDF counts ({natoms} atoms, {nbonds} bonds)")
    if len(lines) < 4 + natoms + nbonds:
        raise ValueError("SDF file ends inside its atom or bond block")
    xyz = np.zeros((natoms, 3), F)
    elems = []
    for i in range(natoms):
        buf = lines[4 + i]
        sc = _Scan(buf)
        for j in range(3):               # a field that does not convert keeps 0
            v = sc.f()
            if v is None:
                break
            xyz[i, j] = v
        k = buf.find(b"\0")
        blen = len(buf) if k < 0 else k
        sym = bytes(c for c in buf[31:min(34, blen)] if c not in (32, 0))
        elems.append(sym[:3])
    typ = np.array([ELEMENT_TYPE.get(e, DUMMY_TYPE) for e in elems], np.int32)
    rad = np.array([ELEMENT_RADIUS.get(e, F(1.70)) for e in elems], F)
    blist = [[] for _ in range(natoms)]
    for b in range(nbonds):
        buf = lines[4 + natoms + b]
        i1, i2 = _atoi3(buf[0:3]), _atoi3(buf[3:6])
        if i1 < 1 or i1 > natoms or i2 < 1 or i2 > natoms:
            continue
        if len(blist[i1 - 1]) < MAX_BONDS:
            blist[i1 - 1].append(i2 - 1)
        if len(blist[i2 - 1]) < MAX_BONDS:
            blist[i2 - 1].append(i1 - 1)
    names = [(e.decode(errors="replace").ljust(2) + str(i % 100))[:4].encode() for i, e in enumerate(elems)]
    return Mol2Ligand(name=name[:3], atom_names=names, sybyl=elems, xyz=xyz, type=typ, radius=rad, bonds=blist)


# ------------------------------------------------------------ set-up
def reconstruction_tree(m: Mol2Ligand) -> np.ndarray:
    """rec[0..2] of every atom (LIB/Mol2Reader.cpp:313-355): breadth-first
    search from atom 0 over the bond lists in order; parent, grandparent and
    great-grandparent, -1 for the ori frame."""
    n = m.n_atoms
    parent = [-1] * n
    grand = [-1] * n
    great = [-1] * n
    seen = [False] * n
    seen[0] = True
    q = [0]
    h = 0
    while h < len(q):
        cur = q[h]
        h += 1
        for nb in m.bonds[cur]:
            if not seen[nb]:
                seen[nb] = True
                parent[nb] = cur
                grand[nb] = parent[cur]
                great[nb] = grand[cur]
                q.append(nb)
    return np.array([parent, grand, great], np.int32).T.copy()


def buildic(xyz: np.ndarray, rec3: np.ndarray, ori) -> np.ndarray:
    """LIB/buildic.cpp:6-71: (dis, ang, dih) of every atom from its
    reconstruction atoms, the ori frame points (ori + x, ori, ori + y)
    standing in for missing ones; float arithmetic."""
    o = np.asarray(ori, F)
    frame = (np.array([F(F(1.0) + o[0]), F(F(0.0) + o[1]), F(F(0.0) + o[2])], F),
             np.array([F(F(0.0) + o[0]), F(F(0.0) + o[1]), F(F(0.0) + o[2])], F),
             np.array([F(F(0.0) + o[0]), F(F(1.0) + o[1]), F(F(0.0) + o[2])], F))
    out = np.zeros((len(xyz), 3), F)
    for i in range(len(xyz)):
        ci = xyz[i].astype(F)
        pts = [xyz[rec3[i, k]].astype(F) if rec3[i, k] >= 0 else frame[k] for k in range(3)]
        out[i, 0] = np.sqrt(cleft._sq3(ci, pts[0]))
        out[i, 1] = cleft._bndang(ci, pts[0], pts[1])
        out[i, 2] = cleft._dihedral(ci, pts[0], pts[1], pts[2])
    return out


def bonded_matrix(bonds: list, n: int, nloops: int = BLOOPS) -> np.ndarray:
    """bondedlist + update_bonded (LIB/bondedlist.cpp:8-61,
    LIB/update_bonded.cpp:58-94): bonds away (0 itself) within nloops, -1
    beyond."""
    B = np.full((n, n), -1, np.int32)
    for a in range(n):
        blist, nbr = [a], [0]
        oldn = 0
        for lvl in range(1, nloops + 1):
            new = []
            for j in range(oldn, len(blist)):
                for nb in bonds[blist[j]]:
                    if nb not in blist:       # only the list before this level is searched
                        new.append(nb)
                        nbr.append(lvl)
            oldn = len(blist)
            blist.extend(new)
        for b, d in zip(blist, nbr):
            B[a, b] = d
    return B


def rebuild_order(rec3: np.ndarray, gpa) -> np.ndarray:
    """buildlist for a ligand with bnum 0 (LIB/buildlist.cpp): every atom,
    the GPA atoms moved to the front and taken first, then repeated passes
    that take an atom once none of its reconstruction atoms is still
    waiting.  Works on 1-based atom numbers as the reference does, so that
    0 (the ori frame) names no atom."""
    n = rec3.shape[0]
    rec1 = np.where(rec3 >= 0, rec3 + 1, 0)
    lst = list(range(1, n + 1)) + [0]          # list[num] is read past the end by the shift
    g1 = [g + 1 for g in gpa]
    for j in range(3):
        i = 0
        while i < n and lst[i] != g1[j]:
            i += 1
        lst[i] = lst[j]
        lst[j] = g1[j]
    out = []
    num = n
    for _ in range(3):
        out.append(lst[0])
        for j in range(num):
            lst[j] = lst[j + 1] if j + 1 < len(lst) else 0
        num -= 1
    while num > 0:
        i = 0
        while i < num:
            waiting = any(rec1[lst[i] - 1, j] in lst[:num] for j in range(3))
            if not waiting:
                out.append(lst[i])
                for j in range(i, num):
                    lst[j] = lst[j + 1]
                num -= 1
            i += 1
    return np.array([a - 1 for a in out], np.int32)


@dataclass
class LigandSetup:
    """What the direct-mode set-up produces for a MOL2 ligand."""
    mol: Mol2Ligand
    ori: np.ndarray          # (3,) float32, the ligand's centroid
    gpa: tuple               # the three global-positioning atoms
    rec: np.ndarray          # (n, 3) int32, -1 = ori frame
    ic: np.ndarray           # (n, 3) float32 dis/ang/dih in the final ori frame
    bonded: np.ndarray       # (n, n) int32 bonds away within bloops, -1 beyond
    mov: np.ndarray          # (n,) rebuild order
    gene_typ: np.ndarray
    gene_atm: np.ndarray
    gene_min: np.ndarray
    gene_max: np.ndarray
    grid0: np.ndarray        # (6,) the cleft grid's row 0: gpa[0]'s coordinates, then dis/ang/dih


def centroid(m: Mol2Ligand) -> np.ndarray:
    """FA->ori in direct mode: FA->ori[j] += coor[j] over the ligand's atoms
    in float, then /= n_lig (LIB/top.cpp:868-880)."""
    ori = np.zeros(3, F)
    for i in range(m.n_atoms):
        for j in range(3):
            ori[j] = F(ori[j] + m.xyz[i, j])
    return (ori / F(m.n_atoms)).astype(F)


def setup(m: Mol2Ligand, n_grid: int) -> LigandSetup:
    """The direct-mode set-up after read_mol2_ligand, for a cleft grid of
    n_grid rows (row 0 the ligand's own position)."""
    n = m.n_atoms
    ori = centroid(m)
    gpa = (0, 1 if n > 1 else 0, 2 if n > 2 else 0)
    rec3 = reconstruction_tree(m)
    ic = buildic(m.xyz, rec3, ori)
    grid0 = np.concatenate([m.xyz[gpa[0]].astype(F), ic[gpa[0]]]).astype(F)
    typ = np.array([-1, 1, 2, 2], np.int32)
    atm = np.array([gpa[0], gpa[1], gpa[1], gpa[2]], np.int32)
    lo = np.array([1.0, -180.0, -180.0, -180.0])
    hi = np.array([float(n_grid - 1), 180.0, 180.0, 180.0])
    return LigandSetup(mol=m, ori=ori, gpa=gpa, rec=rec3, ic=ic, bonded=bonded_matrix(m.bonds, n),
                       mov=rebuild_order(rec3, gpa), gene_typ=typ, gene_atm=atm, gene_min=lo, gene_max=hi,
                       grid0=grid0)


def descriptor(st: LigandSetup, grid_ic: np.ndarray, globalmin, globalmax, maxwidth):
    """The ligand descriptor arrays (a `synth.Ligand`) for the device: the
    set-up plus the cleft grid's (dis, ang, dih) rows 1.. (row 0 is the
    ligand's own position) and the receptor's bounds from calc_center."""
    from .synth import Ligand
    n = st.mol.n_atoms
    rec = np.full((n, 4), -1, np.int32)
    rec[:, :3] = st.rec
    gic = np.vstack([st.grid0[3:][None, :], np.asarray(grid_ic, F)]).astype(F)
    parent = st.rec[:, 0].copy()
    return Ligand(rec=rec, dis=st.ic[:, 0].copy(), ang=st.ic[:, 1].copy(), dih=st.ic[:, 2].copy(),
                  shift=np.zeros(n, F), radius=st.mol.radius.copy(), type=st.mol.type.copy(),
                  mov=st.mov.copy(), bonded=(st.bonded >= 0).astype(np.uint8), gene_typ=st.gene_typ.copy(),
                  gene_atm=st.gene_atm.copy(), gene_min=st.gene_min.copy(), gene_max=st.gene_max.copy(),
                  grid_ic=gic, grid_xyz=np.zeros((len(gic), 3)), ori=st.ori.copy(),
                  globalmin=np.asarray(globalmin, F).copy(), globalmax=np.asarray(globalmax, F).copy(),
                  maxwidth=float(maxwidth), parent=parent)


def direct_system(rec_xyz, rec_radius, rec_type, center, lig_file, params, spacer=0.375, cleft_params=None):
    """A receptor (atoms in the reference's order, e.g. from
    `receptor.read_receptor`, with calc_center's 11 values: ori, globalmin,
    globalmax, maxdst, maxwidth) and a MOL2 or SDF ligand (by extension:
    .sdf/.mol are SDF) set up as the reference's direct mode does
    (LIB/top.cpp:845-998): the ligand's centroid as ori, the automatic
    binding-site grid over the receptor's atoms followed by the ligand's,
    the translation and rotation genes.  An SDF ligand gets the MOL2 path's
    reconstruction tree and bonded matrix: the reference's SDF reader
    builds neither (LIB/SdfReader.cpp:205-260), which leaves it nothing to
    rebuild a pose from.  Returns a `synth.System` ready for
    `Context.upload_system`, and the cleft grid."""
    from .synth import Receptor, System
    is_sdf = isinstance(lig_file, str) and lig_file.lower().endswith((".sdf", ".mol"))
    m = read_sdf(lig_file) if is_sdf else read_mol2(lig_file)
    ori = centroid(m)
    allxyz = np.vstack([np.asarray(rec_xyz, F), m.xyz]).astype(F)
    allrad = np.concatenate([np.asarray(rec_radius, F), m.radius]).astype(F)
    cg = cleft.cleft_grid(allxyz, allrad, ori, spacer, cleft_params)
    st = setup(m, 1 + len(cg.points))
    c = np.asarray(center, F)
    lig = descriptor(st, cg.ic, c[3:6], c[6:9], float(c[10]))
    rec = Receptor(xyz=np.asarray(rec_xyz, F).copy(), radius=np.asarray(rec_radius, F).copy(),
                   type=np.asarray(rec_type, np.int32).copy(), center=np.asarray(ori, np.float64))
    return System(receptor=rec, ligands=[lig], params=params, genes=[]), cg
