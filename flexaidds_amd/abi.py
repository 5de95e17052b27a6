"""ctypes mirror of include/fa_types.h and descriptor builders.

The descriptors borrow numpy buffers; `Descriptors` keeps them alive for as
long as the ctypes structs are in use.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

FA_MAX_LIG_ATOMS = 128
FA_MAX_GENES = 64

FA_MAX_FLEX_RESIDUES = 32
FA_MAX_FLEX_ATOMS = 128
FA_MAX_SCORABLE = 128

FA_OK, FA_HULL_FAIL, FA_CLASH, FA_CAPACITY, FA_GRID_RANGE, FA_BAD_LIGAND, FA_UNDEFINED_FACE = 0, 1, 2, 3, 4, 5, 6
FA_GENE_GRID, FA_GENE_DIS, FA_GENE_ANG, FA_GENE_DIH, FA_GENE_ROT = -1, 0, 1, 2, 4
FA_SHAPE_CELL_LISTS = 1

_f32p = C.POINTER(C.c_float)
_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)
_u8p = C.POINTER(C.c_uint8)


class FlexResidue(C.Structure):
    _fields_ = [("first", C.c_int32), ("n_atoms", C.c_int32), ("n_rot", C.c_int32), ("scorable", _u8p),
                ("rot_xyz", _f32p), ("bonded", _u8p)]


class ReceptorDesc(C.Structure):
    _fields_ = [("n_atoms", C.c_int32), ("xyz", _f32p), ("radius", _f32p), ("type", _i32p),
                ("n_flex", C.c_int32), ("flex", C.POINTER(FlexResidue))]


class Constraint(C.Structure):
    _fields_ = [("atom", C.c_int32), ("partner", C.c_int32), ("partner_is_ligand", C.c_int32), ("type", C.c_int32),
                ("bond_len", C.c_float), ("max_dist", C.c_float)]


class LigandDesc(C.Structure):
    _fields_ = [
        ("n_atoms", C.c_int32), ("n_genes", C.c_int32), ("n_grid", C.c_int32),
        ("rec", _i32p), ("dis", _f32p), ("ang", _f32p), ("dih", _f32p), ("shift", _f32p),
        ("radius", _f32p), ("type", _i32p), ("mov", _i32p), ("bonded", _u8p),
        ("gene_typ", _i32p), ("gene_atm", _i32p), ("gene_min", _f64p), ("gene_max", _f64p),
        ("grid_ic", _f32p),
        ("ori", C.c_float * 3), ("globalmin", C.c_float * 3), ("globalmax", C.c_float * 3),
        ("maxwidth", C.c_float), ("n_cons", C.c_int32), ("cons", C.POINTER(Constraint)),
    ]


class Params(C.Structure):
    _fields_ = [("permeability", C.c_float), ("dee_clash", C.c_float), ("intramolecular", C.c_int32),
                ("solventterm", C.c_float), ("ntypes", C.c_int32), ("emat", _f32p),
                ("normalize_area", C.c_int32), ("dens_off", _i32p), ("dens_xy", _f32p)]


class CFResult(C.Structure):
    _fields_ = [("com", C.c_double), ("wal", C.c_double), ("sas", C.c_double), ("con", C.c_double),
                ("totsas", C.c_double), ("cf", C.c_double), ("app", C.c_double),
                ("rclash", C.c_int32), ("status", C.c_int32)]


class ContactRec(C.Structure):
    _fields_ = [("from_", C.c_int32), ("atom", C.c_int32), ("area", C.c_double), ("dist", C.c_double)]


class Thermo(C.Structure):
    _fields_ = [("temperature", C.c_double), ("log_Z", C.c_double), ("free_energy", C.c_double),
                ("mean_energy", C.c_double), ("mean_energy_sq", C.c_double), ("heat_capacity", C.c_double),
                ("entropy", C.c_double), ("std_energy", C.c_double)]


CF_DTYPE = np.dtype([("com", "f8"), ("wal", "f8"), ("sas", "f8"), ("con", "f8"), ("totsas", "f8"),
                     ("cf", "f8"), ("app", "f8"), ("rclash", "i4"), ("status", "i4")])
REC_DTYPE = np.dtype([("from", "i4"), ("atom", "i4"), ("area", "f8"), ("dist", "f8")])
assert CF_DTYPE.itemsize == C.sizeof(CFResult)
assert REC_DTYPE.itemsize == C.sizeof(ContactRec)


def _ptr(a: np.ndarray, t):
    return a.ctypes.data_as(t)


class Descriptors:
    """ctypes descriptors for one receptor + ligands + params (buffers pinned)."""

    def __init__(self, receptor, ligands, params):
        self._keep = []
        self.receptor = self._rec(receptor)
        self.ligands = [self._lig(l) for l in ligands]
        self.params = self._prm(params)

    def _arr(self, a, dt):
        a = np.ascontiguousarray(a, dtype=dt)
        self._keep.append(a)
        return a

    def _rec(self, r):
        d = ReceptorDesc()
        d.n_atoms = r.n_atoms
        d.xyz = _ptr(self._arr(r.xyz.reshape(-1), np.float32), _f32p)
        d.radius = _ptr(self._arr(r.radius, np.float32), _f32p)
        d.type = _ptr(self._arr(r.type, np.int32), _i32p)
        flex = getattr(r, "flex", None) or []
        d.n_flex = len(flex)
        if flex:
            arr = (FlexResidue * len(flex))()
            for k, f in enumerate(flex):
                arr[k].first, arr[k].n_atoms, arr[k].n_rot = int(f.first), int(f.n_atoms), int(f.n_rot)
                arr[k].scorable = _ptr(self._arr(f.scorable, np.uint8), _u8p)
                arr[k].rot_xyz = _ptr(self._arr(f.rot_xyz.reshape(-1), np.float32), _f32p)
                arr[k].bonded = _ptr(self._arr(f.bonded.reshape(-1), np.uint8), _u8p)
            self._keep.append(arr)
            d.flex = C.cast(arr, C.POINTER(FlexResidue))
        return d

    def _lig(self, l):
        d = LigandDesc()
        d.n_atoms, d.n_genes, d.n_grid = l.n_atoms, l.n_genes, l.n_grid
        d.rec = _ptr(self._arr(l.rec.reshape(-1), np.int32), _i32p)
        for name in ("dis", "ang", "dih", "shift", "radius"):
            setattr(d, name, _ptr(self._arr(getattr(l, name), np.float32), _f32p))
        d.type = _ptr(self._arr(l.type, np.int32), _i32p)
        d.mov = _ptr(self._arr(l.mov, np.int32), _i32p)
        d.bonded = _ptr(self._arr(l.bonded.reshape(-1), np.uint8), _u8p)
        d.gene_typ = _ptr(self._arr(l.gene_typ, np.int32), _i32p)
        d.gene_atm = _ptr(self._arr(l.gene_atm, np.int32), _i32p)
        d.gene_min = _ptr(self._arr(l.gene_min, np.float64), _f64p)
        d.gene_max = _ptr(self._arr(l.gene_max, np.float64), _f64p)
        d.grid_ic = _ptr(self._arr(l.grid_ic.reshape(-1), np.float32), _f32p)
        for k in range(3):
            d.ori[k] = float(l.ori[k])
            d.globalmin[k] = float(l.globalmin[k])
            d.globalmax[k] = float(l.globalmax[k])
        d.maxwidth = float(l.maxwidth)
        cons = getattr(l, "cons", None) or []
        d.n_cons = len(cons)
        if cons:
            arr = (Constraint * len(cons))(*[Constraint(int(a), int(p), int(pl), int(t), float(b), float(m))
                                             for a, p, pl, t, b, m in cons])
            self._keep.append(arr)
            d.cons = C.cast(arr, C.POINTER(Constraint))
        return d

    def _prm(self, p):
        d = Params()
        d.permeability, d.dee_clash = p.permeability, p.dee_clash
        d.intramolecular, d.solventterm, d.ntypes = p.intramolecular, p.solventterm, p.ntypes
        d.emat = _ptr(self._arr(p.emat, np.float32), _f32p)
        d.normalize_area = int(getattr(p, "normalize_area", 0))
        off = getattr(p, "dens_off", None)
        if off is not None:
            d.dens_off = _ptr(self._arr(off, np.int32), _i32p)
            d.dens_xy = _ptr(self._arr(np.asarray(p.dens_xy, np.float32).reshape(-1), np.float32), _f32p)
        return d
