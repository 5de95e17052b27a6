"""Seeded synthetic docking systems for the population-evaluation path.

The reference builds its inputs from PDB/ligand files through a setup stage
(read_lig, buildic, buildlist, assign_shift, bondedlist, calc_cleftic,
add2_optimiz_vec).  This module produces the same *outputs* of that stage
directly, for synthetic receptors and ligands, so that the hot path can be
driven without file I/O:

* receptor atoms uniform in a spherical shell at protein density
  (~1 atom / 11 Å^3), 1-based types, radii U(1.5, 2.0);
* a ligand as a random tree (degree <= 4, 1.5 Å bonds) described by
  internal coordinates with `rec[4]` rebuild references as read by
  LIB/read_lig.cpp:225-396, a rebuild order (`mov`, LIB/buildlist.cpp),
  dihedral shift chains (LIB/assign_shift.cpp:80-188) and a bloops=2 bonded
  matrix (LIB/bondedlist.cpp:8-61, LIB/update_bonded.cpp);
* a gene map in the order add2_optimiz_vec emits it (grid index, then
  ang/dih/dih of the global-positioning atoms, then one dihedral per flexible
  bond, LIB/add2_optimiz_vec.cpp:89-197);
* a cleft grid with internal coordinates relative to the FA->ori frame
  (LIB/calc_cleftic.cpp:5-34).

Everything is drawn from numpy's PCG64 with fixed seeds.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

PI_REF = 3.141592  # LIB/flexaid.h:86

#: per-config seeds
CONFIG_SEED = {0: 1000, 1: 1001, 2: 1002, 3: 1003, 4: 1004}

DENSITY_VOL = 11.0  # Å^3 per atom
CAVITY_RADIUS = (20.0 ** 3 - 3.0 * 1500 * DENSITY_VOL / (4.0 * math.pi)) ** (1.0 / 3.0)


@dataclass
class FlexResidue:
    """A flexible side chain: receptor atoms [first, first + n_atoms) with
    their conformations (0 = the receptor's own coordinates, then the
    rotamers), scorable (non-backbone) atoms and bloops=2 bonded matrix."""
    first: int
    rot_xyz: np.ndarray   # (n_rot, n_atoms, 3) float32
    scorable: np.ndarray  # (n_atoms,) uint8
    bonded: np.ndarray    # (n_atoms, n_atoms) uint8
    chi: np.ndarray = None  # (n_rot, n_chi) degrees, for reference

    @property
    def n_atoms(self) -> int:
        return int(self.rot_xyz.shape[1])

    @property
    def n_rot(self) -> int:
        return int(self.rot_xyz.shape[0])


@dataclass
class Receptor:
    xyz: np.ndarray      # (n,3) float32
    radius: np.ndarray   # (n,) float32
    type: np.ndarray     # (n,) int32, 1-based
    center: np.ndarray   # (3,) float64, cavity centre
    flex: list = field(default_factory=list)  # FlexResidue, increasing `first`

    @property
    def n_atoms(self) -> int:
        return int(self.xyz.shape[0])

    @property
    def n_flex(self) -> int:
        return len(self.flex)


@dataclass
class Ligand:
    rec: np.ndarray       # (n,4) int32, -1 = frame point / none
    dis: np.ndarray       # (n,) float32
    ang: np.ndarray
    dih: np.ndarray
    shift: np.ndarray
    radius: np.ndarray
    type: np.ndarray      # int32, 1-based
    mov: np.ndarray       # (n,) int32 rebuild order
    bonded: np.ndarray    # (n,n) uint8
    gene_typ: np.ndarray  # (g,) int32
    gene_atm: np.ndarray  # (g,) int32
    gene_min: np.ndarray  # (g,) float64
    gene_max: np.ndarray
    grid_ic: np.ndarray   # (n_grid,3) float32
    grid_xyz: np.ndarray  # (n_grid,3) float64 (for reference only)
    ori: np.ndarray       # (3,) float32
    globalmin: np.ndarray  # (3,) float32
    globalmax: np.ndarray
    maxwidth: float
    parent: np.ndarray    # (n,) int32, tree parent (-1 for root)
    flex_bonds: list = field(default_factory=list)
    # distance constraints: (atom, partner, partner_is_ligand, type, bond_len, max_dist)
    cons: list = field(default_factory=list)

    @property
    def n_atoms(self) -> int:
        return int(self.rec.shape[0])

    @property
    def n_genes(self) -> int:
        return int(self.gene_typ.shape[0])

    @property
    def n_grid(self) -> int:
        return int(self.grid_ic.shape[0])


@dataclass
class Params:
    ntypes: int = 40
    emat: np.ndarray = None          # (ntypes*ntypes,) float32
    permeability: float = 1.0
    dee_clash: float = 0.5
    intramolecular: int = 1
    solventterm: float = 0.0
    normalize_area: int = 0
    dens_off: np.ndarray = None      # (ntypes*ntypes + 1,) int32, density-function pairs (optional)
    dens_xy: np.ndarray = None       # (n_nodes, 2) float32


# ---------------------------------------------------------------------------
# geometry helpers (float64; the device and oracle use their own arithmetic)
# ---------------------------------------------------------------------------

def _build_point(x1, x2, x3, dis, ang, dih):
    """Place an atom from three references with the convention of
    LIB/buildcc.cpp:52-129 (rotation by -ang about the plane normal, then by
    dih about the x1->x2 axis)."""
    x1 = np.asarray(x1, np.float64)
    x2 = np.asarray(x2, np.float64)
    x3 = np.asarray(x3, np.float64)
    a = x1[1] * (x2[2] - x3[2]) + x2[1] * (x3[2] - x1[2]) + x3[1] * (x1[2] - x2[2])
    b = x1[2] * (x2[0] - x3[0]) + x2[2] * (x3[0] - x1[0]) + x3[2] * (x1[0] - x2[0])
    c = x1[0] * (x2[1] - x3[1]) + x2[0] * (x3[1] - x1[1]) + x3[0] * (x1[1] - x2[1])
    n = np.array([a, b, c]) / math.sqrt(a * a + b * b + c * c)
    u = x2 - x1
    d = 1.0 / np.linalg.norm(u)
    v = u * d * dis

    def rot(axis, ct, st, w):
        cx, cy, cz = axis
        op = 1.0 - ct
        aa, bb, cc = cx * cx, cy * cy, cz * cz
        return np.array([
            (cx * cz * op - cy * st) * w[2] + ((1 - aa) * ct + aa) * w[0] + (cx * cy * op + cz * st) * w[1],
            (cy * cx * op - cz * st) * w[0] + ((1 - bb) * ct + bb) * w[1] + (cy * cz * op + cx * st) * w[2],
            (cz * cy * op - cx * st) * w[1] + ((1 - cc) * ct + cc) * w[2] + (cz * cx * op + cy * st) * w[0],
        ])

    th = ang * PI_REF / 180.0
    k = rot(n, math.cos(th), -math.sin(th), v)
    ph = dih * PI_REF / 180.0
    return rot(u * d, math.cos(ph), math.sin(ph), k) + x1


def _angle_deg(p, q, r):
    u = np.asarray(p, np.float64) - q
    v = np.asarray(r, np.float64) - q
    c = float(np.dot(u, v) / (np.linalg.norm(u) * np.linalg.norm(v)))
    return math.degrees(math.acos(max(-1.0, min(1.0, c))))


def _dihedral_deg(p1, p2, p3, p4):
    b0 = np.asarray(p1, np.float64) - p2
    b1 = np.asarray(p3, np.float64) - p2
    b2 = np.asarray(p4, np.float64) - p3
    b1n = b1 / np.linalg.norm(b1)
    v = b0 - np.dot(b0, b1n) * b1n
    w = b2 - np.dot(b2, b1n) * b1n
    x = np.dot(v, w)
    y = np.dot(np.cross(b1n, v), w)
    return math.degrees(math.atan2(y, x))


def frame_points(ori):
    o = np.asarray(ori, np.float64)
    return o + np.array([1.0, 0.0, 0.0]), o.copy(), o + np.array([0.0, 1.0, 0.0])


def point_ic(p, ori):
    """(dis, ang, dih) of a point relative to the FA->ori frame such that
    _build_point(frame..., dis, ang, dih) reproduces it (LIB/calc_cleftic.cpp)."""
    ca, cb, cc = frame_points(ori)
    dis = float(np.linalg.norm(np.asarray(p) - ca))
    ang = _angle_deg(p, ca, cb)
    dih = _dihedral_deg(p, ca, cb, cc)
    for cand in (dih, -dih):
        q = _build_point(ca, cb, cc, dis, ang, cand)
        if np.linalg.norm(q - p) < 1e-3:
            return dis, ang, cand
    raise RuntimeError("cleft-grid internal coordinates do not round-trip")


def build_cartesian(lig: Ligand, dis=None, ang=None, dih=None) -> np.ndarray:
    """float64 Cartesian coordinates from internal coordinates (mov order)."""
    dis = lig.dis if dis is None else dis
    ang = lig.ang if ang is None else ang
    dih = lig.dih if dih is None else dih
    n = lig.n_atoms
    xyz = np.zeros((n, 3))
    fr = frame_points(lig.ori)
    for a in lig.mov:
        refs = []
        for i in range(3):
            r = lig.rec[a, i]
            refs.append(xyz[r] if r >= 0 else fr[i])
        xyz[a] = _build_point(refs[0], refs[1], refs[2], float(dis[a]), float(ang[a]), float(dih[a]))
    return xyz


def apply_genes(lig: Ligand, icv: np.ndarray):
    """Decode one gene vector into IC arrays (LIB/ic2cf.cpp:53-105)."""
    dis = lig.dis.astype(np.float32).copy()
    ang = lig.ang.astype(np.float32).copy()
    dih = lig.dih.astype(np.float32).copy()
    for g in range(lig.n_genes):
        t, a, v = int(lig.gene_typ[g]), int(lig.gene_atm[g]), float(icv[g])
        v = min(max(v, float(lig.gene_min[g])), float(lig.gene_max[g]))
        if t == -1:
            k = int(v)
            dis[a], ang[a], dih[a] = lig.grid_ic[k]
        elif t == 0:
            dis[a] = np.float32(v)
        elif t == 1:
            ang[a] = np.float32(v)
        elif t == 2:
            dih[a] = np.float32(v)
            j, cat = a, int(lig.rec[a, 3])
            if cat >= 0:
                while cat != a:
                    dih[cat] = np.float32(dih[j] + lig.shift[cat])
                    j, cat = cat, int(lig.rec[cat, 3])
    return dis, ang, dih


# ---------------------------------------------------------------------------
# receptor
# ---------------------------------------------------------------------------

def make_receptor(rng, n_atoms, center, r_in=CAVITY_RADIUS, r_out=None, ntypes=40,
                  solid=False) -> Receptor:
    """Uniform-in-volume atoms in the shell r_in..r_out around `center` (or a
    solid ball of radius r_out when `solid`)."""
    if r_out is None:
        r_out = (r_in ** 3 + 3.0 * n_atoms * DENSITY_VOL / (4.0 * math.pi)) ** (1.0 / 3.0)
    lo = 0.0 if solid else r_in ** 3
    u = rng.random(n_atoms)
    r = np.cbrt(lo + u * (r_out ** 3 - lo))
    v = rng.normal(size=(n_atoms, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    xyz = (np.asarray(center)[None, :] + v * r[:, None]).astype(np.float32)
    typ = rng.integers(1, ntypes + 1, size=n_atoms).astype(np.int32)
    rad = rng.uniform(1.5, 2.0, size=n_atoms).astype(np.float32)
    return Receptor(xyz=xyz, radius=rad, type=typ, center=np.asarray(center, np.float64))


# ---------------------------------------------------------------------------
# flexible side chains
# ---------------------------------------------------------------------------

#: a methionine-like side chain: backbone N CA C O (not scored), then CB CG
#: SD CE built on chi1 (N-CA-CB-CG), chi2 (CA-CB-CG-SD), chi3 (CB-CG-SD-CE)
FLEX_ATOMS = ("N", "CA", "C", "O", "CB", "CG", "SD", "CE")
FLEX_BONDS = ((0, 1), (1, 2), (2, 3), (1, 4), (4, 5), (5, 6), (6, 7))
_SC_GEOM = ((1.52, 114.0), (1.81, 112.7), (1.79, 100.9))  # bond (A), angle (deg) of CG, SD, CE
CANONICAL_CHI = (-65.0, 180.0, 65.0)


def _place(a, b, c, bond, angle, torsion):
    """Atom d bonded to c with |cd| = bond, angle b-c-d, dihedral a-b-c-d (degrees)."""
    bc = c - b
    bc /= np.linalg.norm(bc)
    n = np.cross(b - a, bc)
    n /= np.linalg.norm(n)
    m = np.cross(n, bc)
    th, ph = math.radians(angle), math.radians(torsion)
    d2 = np.array([-bond * math.cos(th), bond * math.sin(th) * math.cos(ph), bond * math.sin(th) * math.sin(ph)])
    return c + d2[0] * bc + d2[1] * m + d2[2] * n


def _flex_conformation(bb, chi):
    n, ca, _, _, cb = bb
    xyz = list(bb)
    cg = _place(n, ca, cb, _SC_GEOM[0][0], _SC_GEOM[0][1], chi[0])
    sd = _place(ca, cb, cg, _SC_GEOM[1][0], _SC_GEOM[1][1], chi[1])
    ce = _place(cb, cg, sd, _SC_GEOM[2][0], _SC_GEOM[2][1], chi[2])
    return np.array(xyz + [cg, sd, ce])


def _flex_bonded(n_atoms, bonds, bloops=2):
    adj = [[] for _ in range(n_atoms)]
    for a, b in bonds:
        adj[a].append(b)
        adj[b].append(a)
    out = np.zeros((n_atoms, n_atoms), np.uint8)
    for s0 in range(n_atoms):
        seen, frontier = {s0}, [s0]
        for _ in range(bloops):
            frontier = [b for a in frontier for b in adj[a] if b not in seen]
            seen.update(frontier)
        for b in seen:
            out[s0, b] = 1
    return out


def make_flex_receptor(rng, n_atoms, center, n_flex=5, n_rot=10, r_in=CAVITY_RADIUS, ntypes=40,
                       clear=4.2, min_sep_deg=50.0) -> Receptor:
    """A shell receptor of `n_atoms` atoms in all, `n_flex` of whose residues
    are flexible side chains lining the cavity (FLEX_ATOMS, 3 chi angles),
    each with `n_rot` conformations: its own (chi near a canonical
    combination) and n_rot - 1 other canonical combinations with a few
    degrees of noise.  Shell atoms closer than `clear` A to a side-chain atom
    of any conformation are left out (build_rotamers keeps only rotamers free
    of clashes with rigid atoms, LIB/build_rotamers.cpp:214-228), and the
    shell is refilled to `n_atoms`.  Residues sit between shell atoms in calc
    order, as residues of a protein do."""
    center = np.asarray(center, np.float64)
    dirs = []
    while len(dirs) < n_flex:
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        if all(math.degrees(math.acos(max(-1.0, min(1.0, float(v @ w))))) >= min_sep_deg for w in dirs):
            dirs.append(v)
    combos = [(a, b, c) for a in CANONICAL_CHI for b in CANONICAL_CHI for c in CANONICAL_CHI]
    residues = []
    for u in dirs:
        t1 = np.cross(u, [0.0, 0.0, 1.0] if abs(u[2]) < 0.9 else [1.0, 0.0, 0.0])
        t1 /= np.linalg.norm(t1)
        t2 = np.cross(u, t1)
        ca = center + u * (r_in + 0.5)
        nn = ca + 1.46 * (t1 + 0.33 * u) / np.linalg.norm(t1 + 0.33 * u)
        cc = ca + 1.52 * (-0.5 * t1 + 0.87 * t2 + 0.33 * u) / np.linalg.norm(-0.5 * t1 + 0.87 * t2 + 0.33 * u)
        oo = cc + 1.23 * (t2 + 0.5 * u) / np.linalg.norm(t2 + 0.5 * u)
        cb = ca - 1.53 * u
        pick = rng.choice(len(combos), size=n_rot, replace=False)
        chi = np.array([combos[k] for k in pick]) + rng.normal(0.0, 5.0, size=(n_rot, 3))
        conf = np.stack([_flex_conformation([nn, ca, cc, oo, cb], c) for c in chi])
        residues.append((conf, chi))
    n_res_atoms = len(FLEX_ATOMS)
    sc = np.concatenate([conf[:, 4:, :].reshape(-1, 3) for conf, _ in residues])  # scorable, every conformation
    n_shell = n_atoms - n_flex * n_res_atoms
    r_out = (r_in ** 3 + 3.0 * n_atoms * DENSITY_VOL / (4.0 * math.pi)) ** (1.0 / 3.0)
    shell = np.zeros((0, 3))
    while len(shell) < n_shell:
        k = 2 * (n_shell - len(shell)) + 16
        u = rng.random(k)
        r = np.cbrt(r_in ** 3 + u * (r_out ** 3 - r_in ** 3))
        v = rng.normal(size=(k, 3))
        v /= np.linalg.norm(v, axis=1, keepdims=True)
        p = center[None, :] + v * r[:, None]
        d = np.linalg.norm(p[:, None, :] - sc[None, :, :], axis=2).min(axis=1)
        shell = np.vstack([shell, p[d >= clear]])[:n_shell]
    # calc order: shell atoms with the residues spread between them
    cut = np.linspace(0, n_shell, n_flex + 2).round().astype(int)[1:-1]
    xyz, flex, pos = [], [], 0
    for k, (conf, chi) in enumerate(residues):
        xyz.append(shell[pos:cut[k]])
        first = sum(len(x) for x in xyz)
        xyz.append(conf[0])
        flex.append(FlexResidue(first=first, rot_xyz=conf.astype(np.float32),
                                scorable=np.array([0, 0, 0, 0, 1, 1, 1, 1], np.uint8),
                                bonded=_flex_bonded(n_res_atoms, FLEX_BONDS), chi=chi))
        pos = cut[k]
    xyz.append(shell[pos:])
    xyz = np.vstack(xyz).astype(np.float32)
    for f in flex:  # conformation 0 is the receptor's own coordinates (float32 exactly)
        f.rot_xyz[0] = xyz[f.first:f.first + f.n_atoms]
    typ = rng.integers(1, ntypes + 1, size=n_atoms).astype(np.int32)
    rad = rng.uniform(1.5, 2.0, size=n_atoms).astype(np.float32)
    return Receptor(xyz=xyz, radius=rad, type=typ, center=center, flex=flex)


def add_rotamer_genes(lig: "Ligand", rec: Receptor):
    """Append one rotamer gene (typ 4, index gene over 0..n_rot-1) per
    flexible side chain, after the ligand's own genes, as add2_optimiz_vec
    does for "SC" (LIB/add2_optimiz_vec.cpp:28-62)."""
    k = np.arange(rec.n_flex, dtype=np.int32)
    lig.gene_typ = np.concatenate([lig.gene_typ, np.full(rec.n_flex, 4, np.int32)])
    lig.gene_atm = np.concatenate([lig.gene_atm, k])
    lig.gene_min = np.concatenate([lig.gene_min, np.zeros(rec.n_flex)])
    lig.gene_max = np.concatenate([lig.gene_max, np.array([f.n_rot - 1 for f in rec.flex], np.float64)])


# ---------------------------------------------------------------------------
# ligand
# ---------------------------------------------------------------------------

def _random_tree(rng, n):
    parent = [-1, 0, 1]
    deg = [1, 2, 1]
    for i in range(3, n):
        # drug-like trees: mostly chains with occasional branches, attach to
        # one of the last few atoms; degree <= 3 except for rare 4
        while True:
            j = int(rng.integers(max(0, i - 6), i))
            cap = 4 if rng.random() < 0.1 else 3
            if deg[j] < cap:
                break
        parent.append(j)
        deg[j] += 1
        deg.append(1)
    return parent[:n]


def make_ligand(rng, n_atoms, n_tors, ori, anchor, grid_points, ntypes=40,
                max_tries=200) -> Ligand:
    """Random-tree ligand whose reference pose sits at `anchor`.

    `grid_points` (m,3) become cleft-grid entries 1..m; entry 0 is `anchor`.
    """
    assert n_atoms >= 3
    ori = np.asarray(ori, np.float32)
    pts = np.vstack([np.asarray(anchor, np.float64)[None, :], np.asarray(grid_points, np.float64)])
    grid_ic = np.array([point_ic(p, ori) for p in pts], np.float32)
    for _ in range(max_tries):
        lig = _try_ligand(rng, n_atoms, n_tors, ori, pts, grid_ic, ntypes)
        if lig is not None:
            return lig
    raise RuntimeError("could not generate a clash-free ligand")


def _try_ligand(rng, n, n_tors, ori, pts, grid_ic, ntypes):
    parent = _random_tree(rng, n)
    children = [[] for _ in range(n)]
    for i in range(1, n):
        children[parent[i]].append(i)
    depth = [0] * n
    for i in range(1, n):
        depth[i] = depth[parent[i]] + 1
    g0, g1, g2 = 0, 1, 2

    rec = -np.ones((n, 4), np.int32)
    rec[g1, 0] = g0
    rec[g2, 0], rec[g2, 1] = g1, g0
    for x in range(3, n):
        p = parent[x]
        if p == g0:
            rec[x, :3] = (g0, g1, g2)
        elif p == g1:
            rec[x, :3] = (g1, g0, g2)
        else:
            pp = parent[p]
            ppp = parent[pp] if pp >= 0 else -1
            if ppp < 0:  # pp is the root: use gpa2/gpa1 as third reference
                ppp = g2 if p != g2 else g0
            rec[x, :3] = (p, pp, ppp)

    adj0 = [[] for _ in range(n)]
    for i in range(1, n):
        adj0[i].append(parent[i])
        adj0[parent[i]].append(i)

    def within3(x):
        """Atoms at most three bonds from x (x included)."""
        frontier, seen = [x], {x}
        for _ in range(3):
            nxt = []
            for a in frontier:
                for b in adj0[a]:
                    if b not in seen:
                        seen.add(b)
                        nxt.append(b)
            frontier = nxt
        return seen

    near3 = [within3(x) for x in range(n)]

    def depth_sep_ok(x, o):
        """True when atoms x and o are more than three bonds apart."""
        return o not in near3[x]

    dis = np.full(n, 1.5, np.float32)
    ang = np.zeros(n, np.float32)
    dih = np.zeros(n, np.float32)
    # sibling groups share an axis: space their dihedrals 120 degrees apart
    ang[g2] = rng.uniform(105.0, 125.0)
    # reference pose: anchor + a random orientation
    dis[g0], ang[g0], dih[g0] = grid_ic[0]
    ang[g1] = math.degrees(math.acos(rng.uniform(-1.0, 1.0)))
    dih[g1] = rng.uniform(-180.0, 180.0)
    dih[g2] = rng.uniform(-180.0, 180.0)

    groups = {}
    for x in range(3, n):
        groups.setdefault((int(rec[x, 0]), int(rec[x, 1]), int(rec[x, 2])), []).append(x)
    for sib in groups.values():
        for x in sib:
            ang[x] = rng.uniform(105.0, 125.0)

    # rebuild order: global-positioning atoms, then breadth-first
    rest = sorted(range(3, n), key=lambda i: (depth[i], i))
    mov = np.array([g0, g1, g2] + rest, np.int32)

    # greedy placement: each sibling group (atoms sharing rec[0..2]) gets the
    # dihedral base, among a few tries, that keeps it furthest from atoms
    # already built; siblings sit 120 degrees apart
    xyz = np.zeros((n, 3))
    built = np.zeros(n, bool)
    fr = frame_points(ori)

    def place(a):
        refs = [xyz[r] if r >= 0 else fr[i] for i, r in enumerate(rec[a, :3])]
        return _build_point(refs[0], refs[1], refs[2], float(dis[a]), float(ang[a]), float(dih[a]))

    done_groups = set()
    for a in mov:
        key = (int(rec[a, 0]), int(rec[a, 1]), int(rec[a, 2]))
        if a >= 3 and key not in done_groups:
            done_groups.add(key)
            sib = groups[key]
            if key[0] == g1:      # gpa2 is a sibling here: take 120/240 from it
                tries = [120.0]
            elif key[0] == g0:
                tries = [60.0]
            else:
                tries = list(rng.uniform(-180.0, 180.0, size=24))
            best, best_d = None, -1.0
            others = np.nonzero(built)[0]
            for base in tries:
                for k, x in enumerate(sib):
                    dih[x] = ((base + 120.0 * k + 180.0) % 360.0) - 180.0
                pos = np.array([place(x) for x in sib])
                if len(others):
                    dd = np.linalg.norm(pos[:, None, :] - xyz[others][None, :, :], axis=2)
                    far = np.array([[depth_sep_ok(x, o) for o in others] for x in sib])
                    md = dd[far].min() if np.any(far) else 99.0
                else:
                    md = 99.0
                if md > best_d:
                    best, best_d = base, md
            for k, x in enumerate(sib):
                dih[x] = ((best + 120.0 * k + 180.0) % 360.0) - 180.0
        xyz[a] = place(a)
        built[a] = True

    # flexible bonds (p -> c) with c non-terminal and its children built on (c, p, .)
    eligible = []
    for c in range(n):
        p = parent[c]
        if p < 0 or c == g1 or not children[c]:
            continue
        kids = [x for x in children[c] if rec[x, 0] == c and rec[x, 1] == p]
        if len(kids) == len(children[c]):
            eligible.append((p, c))
    nt = min(n_tors, len(eligible))
    flex = [eligible[i] for i in sorted(rng.choice(len(eligible), size=nt, replace=False))] if nt else []
    shift = np.zeros(n, np.float32)
    gene_typ = [-1, 1, 2, 2]
    gene_atm = [g0, g1, g1, g2]
    pos = {int(a): k for k, a in enumerate(mov)}
    for (p, c) in flex:
        kids = sorted(children[c], key=lambda x: pos[x])
        if len(kids) > 1:  # shift chain, LIB/assign_shift.cpp:160-180
            for j in range(len(kids)):
                k = (j + 1) % len(kids)
                rec[kids[j], 3] = kids[k]
                shift[kids[k]] = np.float32(dih[kids[k]] - dih[kids[j]])
        gene_typ.append(2)
        gene_atm.append(kids[0])

    # bonded within bloops = 2 (self, 1-2, 1-3)
    adj = [[] for _ in range(n)]
    for i in range(1, n):
        adj[i].append(parent[i])
        adj[parent[i]].append(i)
    bonded = np.zeros((n, n), np.uint8)
    for s in range(n):
        bonded[s, s] = 1
        for a in adj[s]:
            bonded[s, a] = 1
            for b in adj[a]:
                bonded[s, b] = 1

    ng = len(gene_typ)
    gmin = np.array([0.0] + [-180.0] * (ng - 1))
    gmax = np.array([float(len(pts) - 1)] + [180.0] * (ng - 1))
    lig = Ligand(rec=rec, dis=dis, ang=ang, dih=dih, shift=shift,
                 radius=rng.uniform(1.5, 2.0, size=n).astype(np.float32),
                 type=rng.integers(1, ntypes + 1, size=n).astype(np.int32),
                 mov=mov, bonded=bonded,
                 gene_typ=np.array(gene_typ, np.int32), gene_atm=np.array(gene_atm, np.int32),
                 gene_min=gmin, gene_max=gmax, grid_ic=grid_ic, grid_xyz=pts,
                 ori=ori, globalmin=np.zeros(3, np.float32), globalmax=np.zeros(3, np.float32),
                 maxwidth=0.0, parent=np.array(parent, np.int32), flex_bonds=flex)
    xyz = build_cartesian(lig)
    # reject self-overlapping reference poses: non-bonded pairs closer than
    # 2.2 Å (three bonds apart) or 2.5 Å (further apart)
    d = np.linalg.norm(xyz[:, None, :] - xyz[None, :, :], axis=2)
    sep = np.full((n, n), 99, np.int32)
    for s in range(n):
        sep[s, s] = 0
        frontier, k = [s], 0
        while frontier and k < 3:
            k += 1
            nxt = []
            for a in frontier:
                for b in adj[a]:
                    if sep[s, b] > k:
                        sep[s, b] = k
                        nxt.append(b)
            frontier = nxt
    if np.any(d[sep == 3] < 2.2) or np.any(d[sep > 3] < 2.5):
        return None
    return lig


def finalize_frame(lig: Ligand, rec: Receptor):
    """FA->ori / globalmin / globalmax / maxwidth as calc_center computes them
    over receptor + ligand reference pose (LIB/calc_center.cpp:8-62)."""
    lxyz = build_cartesian(lig).astype(np.float32)
    allxyz = np.vstack([rec.xyz, lxyz])
    lig.globalmin = allxyz.min(axis=0).astype(np.float32)
    lig.globalmax = allxyz.max(axis=0).astype(np.float32)
    lig.maxwidth = float(np.max(lig.globalmax - lig.globalmin))


def make_emat(rng, ntypes=40) -> np.ndarray:
    m = rng.normal(size=(ntypes, ntypes))
    m = np.triu(m) + np.triu(m, 1).T
    return m.astype(np.float32).reshape(-1)


def make_density_emat(rng, ntypes=40, frac=0.5, max_nodes=6):
    """Density-function pairs for a fraction of the type pairs (symmetric):
    2..max_nodes nodes with x ascending in [0, 1.2) (relative contact area)
    and y ~ N(0, 1).  Returns (dens_off, dens_xy) in the fa_params layout;
    pairs left out stay weight-1 pairs of the value matrix."""
    lists = {}
    for i in range(ntypes):
        for j in range(i, ntypes):
            if rng.uniform() < frac:
                n = int(rng.integers(2, max_nodes + 1))
                x = np.sort(rng.uniform(0.0, 1.2, n))
                y = rng.normal(size=n)
                lists[(i, j)] = np.stack([x, y], axis=1).astype(np.float32)
    off = np.zeros(ntypes * ntypes + 1, np.int32)
    xy = []
    for k in range(ntypes * ntypes):
        i, j = divmod(k, ntypes)
        nodes = lists.get((min(i, j), max(i, j)))
        if nodes is not None:
            xy.append(nodes)
        off[k + 1] = off[k] + (0 if nodes is None else len(nodes))
    dens_xy = np.concatenate(xy) if xy else np.zeros((0, 2), np.float32)
    return off, dens_xy


# ---------------------------------------------------------------------------
# poses
# ---------------------------------------------------------------------------

def random_genes(rng, lig: Ligand, n_poses, grid_lo=1, grid_hi=None) -> np.ndarray:
    """Gene vectors (ic values): grid index, uniform rotation, uniform
    torsions, and (rotamer genes, typ 4) a uniform rotamer index."""
    rot = np.nonzero(lig.gene_typ == 4)[0]
    nl = lig.n_genes - len(rot)          # the ligand's own genes come first
    g = np.zeros((n_poses, lig.n_genes), np.float64)
    hi = lig.n_grid - 1 if grid_hi is None else grid_hi
    g[:, 0] = rng.integers(grid_lo, hi + 1, size=n_poses).astype(np.float64)
    g[:, 1] = np.degrees(np.arccos(rng.uniform(-1.0, 1.0, size=n_poses)))
    g[:, 2:nl] = rng.uniform(-180.0, 180.0, size=(n_poses, nl - 2))
    for c in rot:
        g[:, c] = rng.integers(0, int(lig.gene_max[c]) + 1, size=n_poses).astype(np.float64)
    return g


@dataclass
class System:
    receptor: Receptor
    ligands: list
    params: Params
    genes: list          # per ligand: (n_poses, n_genes) float64
    config: int = -1


def ga_limits(lig: Ligand, angle_step=2.0):
    """Gene limits (min, max, del, map) for the GA: the cleft-grid index gene
    in steps of one point (an index gene: map = 1), angles and dihedrals in
    `angle_step` degrees over the ligand's gene range."""
    out = []
    for g in range(lig.n_genes):
        lo, hi = float(lig.gene_min[g]), float(lig.gene_max[g])
        if int(lig.gene_typ[g]) in (-1, 4):  # grid-index and rotamer genes (map_opt_par = 1)
            out.append((lo, hi, 1.0, 1))
        else:
            out.append((lo, hi, angle_step, 0))
    return out


def _ligand_job(job):
    """One ligand of a multi-ligand system from its own seeded stream
    (picklable, so ligands can be built in worker processes): job =
    (seed, k, kind, rec_xyz, center, ori, extra)."""
    seed, k, kind, rec_xyz, center, ori, extra = job
    rng = np.random.default_rng([int(seed), int(k)])
    rec = Receptor(xyz=rec_xyz, radius=np.zeros(len(rec_xyz), np.float32), type=np.zeros(len(rec_xyz), np.int32),
                   center=center)
    if kind == "screen":            # config 2: 12-80 atoms, 0-15 torsions, a 10 A box
        na = int(rng.integers(12, 81))
        nt = int(rng.integers(0, 16))
        pts = _box_points(rng, center, 256)
        anchor = center
    elif kind == "stream":          # config 3: log-normal sizes
        na, nt = config3_ligand_size(rng)
        pts = _box_points(rng, center, 256)
        anchor = center
    else:                           # config 4: 20-60 atoms over the globule's surface
        R, n_grid = extra
        na = int(rng.integers(20, 61))
        nt = int(round(na / 5.0))
        u = rng.random(n_grid)
        r = np.cbrt((R - 4.0) ** 3 + u * ((R + 4.0) ** 3 - (R - 4.0) ** 3))
        v = rng.normal(size=(n_grid, 3))
        v /= np.linalg.norm(v, axis=1, keepdims=True)
        pts = center[None, :] + v * r[:, None]
        anchor = center + v[0] * (R + 4.0)
    lig = make_ligand(rng, na, nt, ori, anchor, pts)
    finalize_frame(lig, rec)
    return lig, rng


def _build_ligands(jobs, pop, workers=None):
    """Ligands of `jobs` and `pop` random gene vectors each, in worker
    processes when there are many (results do not depend on the worker
    count: every ligand has its own stream)."""
    import os
    import sys
    if workers is None:
        workers = min(16, os.cpu_count() or 1) if len(jobs) >= 24 else 1
    # spawned workers re-import the main module by path; a main read from
    # stdin or `-c` has none, and its workers would die at start-up
    main_file = getattr(sys.modules.get("__main__"), "__file__", None)
    if main_file is not None and not os.path.isfile(main_file):
        workers = 1
    if workers > 1:
        import concurrent.futures as cf
        import multiprocessing as mp
        with cf.ProcessPoolExecutor(max_workers=workers, mp_context=mp.get_context("spawn")) as ex:
            built = list(ex.map(_ligand_job, jobs, chunksize=max(1, len(jobs) // (4 * workers))))
    else:
        built = [_ligand_job(j) for j in jobs]
    ligs = [b[0] for b in built]
    genes = [random_genes(b[1], b[0], pop) for b in built]
    return ligs, genes


def _box_points(rng, center, n, edge=10.0):
    return np.asarray(center)[None, :] + rng.uniform(-edge / 2, edge / 2, size=(n, 3))


def make_config0(n_poses=1000, seed=None, n_rec=1500, n_lig=30, n_tors=6) -> System:
    """Plain CF scoring: 1,500-atom shell receptor, 30-atom/6-torsion ligand,
    `n_poses` poses with translations uniform in a 10 Å box."""
    rng = np.random.default_rng(CONFIG_SEED[0] if seed is None else seed)
    center = np.array([22.3, 41.7, -13.9])
    rec = make_receptor(rng, n_rec, center, r_in=CAVITY_RADIUS, r_out=20.0)
    ori = rec.xyz.astype(np.float64).mean(axis=0).astype(np.float32)
    pts = _box_points(rng, center, n_poses)
    lig = make_ligand(rng, n_lig, n_tors, ori, center, pts)
    finalize_frame(lig, rec)
    params = Params(ntypes=40, emat=make_emat(rng))
    genes = np.zeros((n_poses, lig.n_genes))
    genes[:, 0] = np.arange(1, n_poses + 1)
    genes[:, 1] = np.degrees(np.arccos(rng.uniform(-1.0, 1.0, size=n_poses)))
    genes[:, 2:] = rng.uniform(-180.0, 180.0, size=(n_poses, lig.n_genes - 2))
    return System(receptor=rec, ligands=[lig], params=params, genes=[genes], config=0)


def make_config1(pop=2000, seed=None, n_flex=5, n_rot=10) -> System:
    """Plain full-GA dock shape: 4,000-atom receptor whose cavity is lined
    by `n_flex` flexible side chains of 3 chi angles (`n_rot` conformations
    each; n_flex = 0 gives the rigid shell), 45-atom/10-torsion ligand, one
    population of `pop` chromosomes; 14 ligand genes + one rotamer gene per
    side chain."""
    rng = np.random.default_rng(CONFIG_SEED[1] if seed is None else seed)
    center = np.array([-8.6, 17.2, 33.4])
    if n_flex:
        rec = make_flex_receptor(rng, 4000, center, n_flex=n_flex, n_rot=n_rot)
    else:
        rec = make_receptor(rng, 4000, center, r_in=CAVITY_RADIUS)
    ori = rec.xyz.astype(np.float64).mean(axis=0).astype(np.float32)
    pts = _box_points(rng, center, 1000)
    lig = make_ligand(rng, 45, 10, ori, center, pts)
    finalize_frame(lig, rec)
    if n_flex:
        add_rotamer_genes(lig, rec)
    params = Params(ntypes=40, emat=make_emat(rng))
    genes = random_genes(rng, lig, pop)
    return System(receptor=rec, ligands=[lig], params=params, genes=[genes], config=1)


def make_config2(n_ligands=256, pop=512, seed=None) -> System:
    """Batched uneven screen: one 6,000-atom receptor, ligands with 12-80
    heavy atoms and 0-15 torsions, `pop` chromosomes each."""
    base = CONFIG_SEED[2] if seed is None else seed
    rng = np.random.default_rng(base)
    center = np.array([5.1, -12.4, 9.8])
    rec = make_receptor(rng, 6000, center, r_in=CAVITY_RADIUS)
    ori = rec.xyz.astype(np.float64).mean(axis=0).astype(np.float32)
    params = Params(ntypes=40, emat=make_emat(rng))
    ligs, genes = _build_ligands([(base, k, "screen", rec.xyz, center, ori, None) for k in range(n_ligands)], pop)
    return System(receptor=rec, ligands=ligs, params=params, genes=genes, config=2)


def config3_ligand_size(rng):
    """Heavy atoms log-normal(mu = ln 35, sigma = 0.4) clipped to 10..100;
    torsions about atoms / 5."""
    na = int(np.clip(round(float(rng.lognormal(math.log(35.0), 0.4))), 10, 100))
    return na, int(round(na / 5.0))


def make_config3(n_receptors=8, ligands_per_receptor=16, seed=None, n_poses=8) -> list:
    """Served stream: `n_receptors` resident receptors of 2,000..10,000 atoms
    (cleft shells as in config 0), each with a resident library of
    `ligands_per_receptor` ligands sized as config3_ligand_size draws.  A
    request names a receptor and one of its ligands."""
    base = CONFIG_SEED[3] if seed is None else seed
    rng = np.random.default_rng(base)
    sizes = np.linspace(2000, 10000, n_receptors).round().astype(int)
    recs, jobs = [], []
    for q, n_rec in enumerate(sizes):
        center = rng.uniform(-30.0, 30.0, size=3)
        rec = make_receptor(rng, int(n_rec), center, r_in=CAVITY_RADIUS)
        ori = rec.xyz.astype(np.float64).mean(axis=0).astype(np.float32)
        params = Params(ntypes=40, emat=make_emat(rng))
        recs.append((rec, params))
        jobs += [(base, q * 1000 + k, "stream", rec.xyz, center, ori, None) for k in range(ligands_per_receptor)]
    ligs, genes = _build_ligands(jobs, n_poses)
    systems = []
    for q, (rec, params) in enumerate(recs):
        sl = slice(q * ligands_per_receptor, (q + 1) * ligands_per_receptor)
        systems.append(System(receptor=rec, ligands=ligs[sl], params=params, genes=genes[sl], config=3))
    return systems


def make_config4(n_ligands=64, pop=8, n_rec=20000, n_grid=2000, seed=None):
    """Large blind dock: a solid globule of `n_rec` atoms at protein density
    (radius ~37 Å for 20,000), `n_ligands` competing ligands of 20-60 heavy
    atoms (torsions about atoms / 5) whose translation grids cover the whole
    surface (points uniform in a shell from 4 Å inside to 4 Å outside the
    globule's radius), and concentrations log-uniform over 1 nM - 1 mM.
    Returns (system, concentrations in M)."""
    base = CONFIG_SEED[4] if seed is None else seed
    rng = np.random.default_rng(base)
    center = np.array([0.0, 0.0, 0.0])
    R = (3.0 * n_rec * DENSITY_VOL / (4.0 * math.pi)) ** (1.0 / 3.0)
    rec = make_receptor(rng, n_rec, center, r_in=0.0, r_out=R, solid=True)
    ori = rec.xyz.astype(np.float64).mean(axis=0).astype(np.float32)
    params = Params(ntypes=40, emat=make_emat(rng))
    ligs, genes = _build_ligands([(base, k, "blind", rec.xyz, center, ori, (R, n_grid)) for k in range(n_ligands)],
                                 pop)
    conc = np.exp(rng.uniform(math.log(1e-9), math.log(1e-3), n_ligands))
    return System(receptor=rec, ligands=ligs, params=params, genes=genes, config=4), conc


def poisson_requests(rng, systems, rate_hz, n_requests, generations=100):
    """Request stream: Poisson arrivals at `rate_hz`; receptor uniform, ligand
    uniform in its library.  Returns dicts sorted by arrival time (s)."""
    t = np.cumsum(rng.exponential(1.0 / rate_hz, size=n_requests))
    out = []
    for i in range(n_requests):
        r = int(rng.integers(0, len(systems)))
        l = int(rng.integers(0, len(systems[r].ligands)))
        out.append({"t": float(t[i]), "receptor": r, "ligand": l, "generations": generations,
                    "seed": int(rng.integers(1, 2**62))})
    return out


def add_constraints(system: System, lig_index=0, n_receptor=2, n_ligand=1, bond_len=3.2, max_dist=4.6, seed=0):
    """Covalent (type 1) distance constraints on one ligand: `n_receptor`
    pairs between random ligand atoms and the receptor atoms nearest the
    cleft centre, and the `n_ligand` closest non-bonded ligand pairs of the
    reference pose (often in contact); contact distances fall inside the
    bell [2 bond_len - max_dist, max_dist] or, outside it, on the even-power
    tails the reference's Gaussian has."""
    rng = np.random.default_rng(seed)
    lig = system.ligands[lig_index]
    rec = system.receptor
    d = np.linalg.norm(rec.xyz.astype(np.float64) - rec.center[None, :], axis=1)
    near = np.argsort(d)[:max(8, n_receptor)]
    atoms = rng.choice(lig.n_atoms, size=n_receptor, replace=False)
    cons = [(int(atoms[q]), int(near[q]), 0, 1, float(bond_len), float(max_dist)) for q in range(n_receptor)]
    xyz = build_cartesian(lig).astype(np.float64)
    dd = np.linalg.norm(xyz[:, None, :] - xyz[None, :, :], axis=2)
    pairs = sorted((dd[a, b], a, b) for a in range(lig.n_atoms) for b in range(a + 1, lig.n_atoms)
                   if not lig.bonded[a, b])
    for _, a, b in pairs[:n_ligand]:
        cons.append((int(a), int(b), 1, 1, float(bond_len), float(max_dist)))
    lig.cons = cons
    return cons


def make_engulf_system(n_poses=200, seed=1010) -> System:
    """Config-0 shape whose ligand alternates small (1.5 A) and large (2.0 A)
    atoms along its bonds (by depth parity in the tree).  A bonded neighbour
    1.5 A away with a radius 0.5 A larger puts the radical plane beyond the
    atom's centre (an engulfing plane, D > 0, LIB/Vcontacts.cpp:470-478), and
    bonded pairs add no clash wall, so small atoms with two or three large
    neighbours reach calc_areas' two- and three-plane engulfing branches
    (LIB/Vcontacts.cpp:963-1042)."""
    s = make_config0(n_poses=n_poses, seed=seed)
    lig = s.ligands[0]
    depth = np.zeros(lig.n_atoms, np.int32)
    for i in range(1, lig.n_atoms):
        depth[i] = depth[lig.parent[i]] + 1
    lig.radius = np.where(depth % 2 == 0, 1.5, 2.0).astype(np.float32)
    return s


def fibonacci_sphere(n):
    """n unit vectors spread evenly over the sphere (golden-angle spiral)."""
    k = np.arange(n) + 0.5
    z = 1.0 - 2.0 * k / n
    r = np.sqrt(1.0 - z * z)
    phi = k * math.pi * (3.0 - math.sqrt(5.0))
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)


def make_shell_system(n_shell=80, dist=5.0, n_poses=64, n_far=63, seed=1011) -> System:
    """A three-atom ligand (bonds of 0.6 A, so every pair is within two
    bonds) whose cleft grid holds the centre of a shell of `n_shell`
    receptor atoms (radius 1.8 A) `dist` A from it (every odd grid point)
    and points 30 A away from any atom (the even ones).  At the centre each
    ligand atom has about n_shell contacts, none within the clash distance,
    so its polyhedron has about 3 n_shell - 6 edges: the hull walk reaches its
    200-edge limit (LIB/Vcontacts.cpp:378-407) for n_shell >= ~70, and more
    than 128 contacts exceed the device's per-atom capacity.  Far away the
    ligand touches nothing and scores its solvent term only."""
    rng = np.random.default_rng(seed)
    center = np.array([3.0, -2.0, 5.0])
    # a little jitter keeps four planes from meeting in one vertex (an exact
    # degeneracy where the hull walk finds no plane ahead)
    u = fibonacci_sphere(n_shell) + rng.normal(0.0, 0.03, size=(n_shell, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    xyz = (center[None, :] + dist * u).astype(np.float32)
    typ = rng.integers(1, 41, size=n_shell).astype(np.int32)
    rec = Receptor(xyz=xyz, radius=np.full(n_shell, 1.8, np.float32), type=typ, center=center)
    ori = (center + np.array([7.0, 5.0, -6.0])).astype(np.float32)  # any frame origin off the anchor
    far = center + np.array([30.0, 0.0, 0.0])
    pts = np.vstack([center if k % 2 else far for k in range(1, n_far + 1)])
    lig = make_ligand(rng, 3, 0, ori, center, pts)
    lig.dis[1:] = np.float32(0.6)
    lig.radius = np.full(lig.n_atoms, 1.6, np.float32)
    finalize_frame(lig, rec)
    params = Params(ntypes=40, emat=make_emat(rng))
    genes = random_genes(rng, lig, n_poses)
    genes[:, 0] = np.arange(n_poses) % (n_far + 1)
    return System(receptor=rec, ligands=[lig], params=params, genes=[genes], config=-1)
