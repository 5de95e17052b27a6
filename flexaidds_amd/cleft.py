"""Binding-cleft grid of a receptor, as the reference builds it when no grid
is given (LIB/top.cpp:904-916): gap-sphere cleft detection
(LIB/CleftDetector.cpp: detect_cleft), the grid inside the cleft's spheres
(LIB/generate_grid.cpp:11-107) and the grid points' internal coordinates in
the ori frame (LIB/calc_cleftic.cpp:5-34).  Single-precision arithmetic as
the reference's float code does it, so the grid is the same point for point.

detect_cleft: for every pair of atoms i < j (atoms 1 .. n-1 of the
reference's 1-based array, i.e. all but the last receptor atom; atoms at the
origin with radius 0 are skipped) with 1 <= d^2 <= max_pair_dist^2, a probe at
the midpoint with radius min(d/2, probe_radius_max) shrinks by
probe_shrink_step until no other atom k overlaps it (d_k^2 < (r_k + r)^2);
probes that keep at least probe_radius_min survive.  Probes closer than
cluster_cutoff are joined by union-find in pair order (i, then j); the first
label (ascending root index) of the largest cluster with at least
min_cluster_size members wins, and its probes are listed last one first.
The probe order is that of the reference's serial build (its OpenMP build
appends per-thread lists in scheduling order, so its grid order varies from
run to run).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

_libm = ctypes.CDLL("libm.so.6")
_acosf = _libm.acosf
_acosf.restype = ctypes.c_float
_acosf.argtypes = [ctypes.c_float]
_atan2f = _libm.atan2f
_atan2f.restype = ctypes.c_float
_atan2f.argtypes = [ctypes.c_float, ctypes.c_float]

F = np.float32
PI = 3.141592          # LIB/flexaid.h:86


@dataclass
class CleftParams:
    """LIB/CleftDetector.h default_cleft_params."""
    max_pair_dist: float = 12.0
    probe_radius_max: float = 5.0
    probe_radius_min: float = 1.5
    probe_shrink_step: float = 0.1
    cluster_cutoff: float = 4.0
    min_cluster_size: int = 10


@dataclass
class CleftGrid:
    spheres: np.ndarray   # (s,4) float32: x, y, z, radius in the order generate_grid visits them
    points: np.ndarray    # (g,3) float32: grid points 1 .. num_grd-1
    ic: np.ndarray        # (g,3) float32: dis, ang, dih of each point (calc_cleftic)


def _sq3(a, b):
    """((a0-b0)^2 + (a1-b1)^2) + (a2-b2)^2 in float, broadcasting."""
    d0 = (a[..., 0] - b[..., 0]).astype(F)
    d1 = (a[..., 1] - b[..., 1]).astype(F)
    d2 = (a[..., 2] - b[..., 2]).astype(F)
    return ((d0 * d0 + d1 * d1).astype(F) + d2 * d2).astype(F)


def _pairs_within(xyz, cut2, lo_ok=None):
    """All pairs i < j with _sq3 <= cut2 (and >= lo_ok if given), in (i, j)
    order, found through a cell list then tested in float."""
    n = len(xyz)
    if n < 2:
        return np.zeros(0, np.int64), np.zeros(0, np.int64)
    cell = float(np.sqrt(cut2)) + 1e-3
    key = np.floor((xyz.astype(np.float64) - xyz.min(axis=0)) / cell).astype(np.int64)
    dims = key.max(axis=0) + 1
    lin = (key[:, 0] * dims[1] + key[:, 1]) * dims[2] + key[:, 2]
    order = np.argsort(lin, kind="stable")
    starts = np.searchsorted(lin[order], np.arange(dims.prod() + 1))
    I, J = [], []
    for dx in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dz in (-1, 0, 1):
                nb = key + np.array([dx, dy, dz])
                ok = np.all((nb >= 0) & (nb < dims), axis=1)
                nl = (nb[:, 0] * dims[1] + nb[:, 1]) * dims[2] + nb[:, 2]
                idx = np.nonzero(ok)[0]
                s, e = starts[nl[idx]], starts[nl[idx] + 1]
                cnt = e - s
                rep_i = np.repeat(idx, cnt)
                pos = np.repeat(s - np.cumsum(cnt) + cnt, cnt) + np.arange(cnt.sum())
                rep_j = order[pos]
                keep = rep_j > rep_i
                I.append(rep_i[keep])
                J.append(rep_j[keep])
    I = np.concatenate(I)
    J = np.concatenate(J)
    d2 = _sq3(xyz[I], xyz[J])
    keep = d2 <= F(cut2)
    if lo_ok is not None:
        keep &= d2 >= F(lo_ok)
    I, J = I[keep], J[keep]
    o = np.lexsort((J, I))
    return I[o], J[o]


def _probes(xyz, rad, p: CleftParams):
    """generate_probes (serial branch): probe centres and radii in pair order."""
    xyz = xyz.astype(F)
    rad = rad.astype(F)
    mp = F(p.max_pair_dist)
    I, J = _pairs_within(xyz, F(mp * mp), lo_ok=1.0)
    if I.size == 0:
        return np.zeros((0, 3), F), np.zeros(0, F)
    d = np.sqrt(_sq3(xyz[I], xyz[J])).astype(F)
    cen = (F(0.5) * (xyz[I] + xyz[J]).astype(F)).astype(F)
    r0 = np.minimum((F(0.5) * d).astype(F), F(p.probe_radius_max)).astype(F)
    # the shrink sequence r_t = r_{t-1} - step while r_t >= r_min
    rmin, step = F(p.probe_radius_min), F(p.probe_shrink_step)
    seq = [r0]
    while True:
        nxt = (seq[-1] - step).astype(F)
        if not np.any(seq[-1] >= rmin):
            break
        seq.append(nxt)
    R = np.stack(seq, axis=1)                       # (P, T)
    T = (R >= rmin).sum(axis=1)                     # steps available per probe
    # clash(t) = any_k d_k^2 < (r_k + r_t)^2 is monotone in t: binary search
    # for the first t without a clash, vectorised over chunks of probes
    keep_r = np.full(len(I), -1.0, F)
    chunk = 1024
    for s in range(0, len(I), chunk):
        e = min(s + chunk, len(I))
        dk2 = _sq3(cen[s:e, None, :], xyz[None, :, :])          # (c, n)
        excl = np.zeros_like(dk2, dtype=bool)
        rows = np.arange(e - s)
        excl[rows, I[s:e]] = True
        excl[rows, J[s:e]] = True
        lo = np.zeros(e - s, np.int64)
        hi = T[s:e].copy()                                      # answer in [lo, hi]; hi = none
        while True:
            act = lo < hi
            if not act.any():
                break
            mid = (lo + hi) // 2
            rt = R[np.arange(s, e), np.minimum(mid, R.shape[1] - 1)]
            ov = (rad[None, :] + rt[:, None]).astype(F)
            clash = ((dk2 < (ov * ov).astype(F)) & ~excl).any(axis=1)
            hi = np.where(act & ~clash, mid, hi)
            lo = np.where(act & clash, mid + 1, lo)
        ok = lo < T[s:e]
        keep_r[s:e] = np.where(ok, R[np.arange(s, e), np.minimum(lo, R.shape[1] - 1)], F(-1.0))
    sel = keep_r >= rmin
    return cen[sel], keep_r[sel]


def _find(parent, x):
    while parent[x] != x:
        parent[x] = parent[parent[x]]
        x = parent[x]
    return x


def _cluster(cen, cutoff):
    """cluster_probes: union-find with path halving and union by rank, pairs
    in (i, j) order; returns each probe's root after the final pass."""
    n = len(cen)
    parent = list(range(n))
    rank = [0] * n
    cut2 = F(F(cutoff) * F(cutoff))
    I, J = _pairs_within(cen, cut2)
    strict = _sq3(cen[I], cen[J]) < cut2
    for a, b in zip(I[strict].tolist(), J[strict].tolist()):
        a = _find(parent, a)
        b = _find(parent, b)
        if a == b:
            continue
        if rank[a] < rank[b]:
            a, b = b, a
        parent[b] = a
        if rank[a] == rank[b]:
            rank[a] += 1
    for i in range(n):
        parent[i] = _find(parent, i)
    return np.array(parent, np.int64)


def detect_cleft(xyz, radius, params: CleftParams | None = None) -> np.ndarray:
    """Spheres (x, y, z, r) of the largest cleft, in the order of the
    reference's linked list (last probe first).  xyz/radius are the receptor
    atoms in the reference's order."""
    p = params or CleftParams()
    xyz = np.asarray(xyz, F)
    radius = np.asarray(radius, F)
    n = len(xyz)
    # the reference scans its atoms[0 .. atm_cnt_real-1]: atom 0 is the unused
    # zero atom, so receptor atoms 1 .. n-1 (not the last one)
    use = np.arange(max(n - 1, 0))
    zero = np.all(xyz[use] == 0.0, axis=1) & (radius[use] == 0.0)
    use = use[~zero]
    cen, r = _probes(xyz[use], radius[use], p)
    if len(cen) == 0:
        raise ValueError("cleft detection found no cavities")
    labels = _cluster(cen, p.cluster_cutoff)
    vals, counts = np.unique(labels, return_counts=True)    # ascending label, as std::map
    best, best_n = -1, 0
    for v, c in zip(vals.tolist(), counts.tolist()):
        if c > best_n and c >= p.min_cluster_size:
            best, best_n = v, c
    if best < 0:
        for v, c in zip(vals.tolist(), counts.tolist()):
            if c > best_n:
                best, best_n = v, c
    idx = np.nonzero(labels == best)[0][::-1]
    return np.concatenate([cen[idx], r[idx, None]], axis=1).astype(F)


def _seq(lo, hi, step):
    out, c = [], F(lo)
    while c < hi:
        out.append(c)
        c = F(c + step)
    return np.array(out, F)


def _roundf_int(v):
    v = np.asarray(v, np.float64)              # float values are exact in double
    t = np.trunc(v)
    return (t + np.where(np.abs(v - t) >= 0.5, np.sign(v), 0.0)).astype(np.int64)


def generate_grid(spheres, spacer=0.375) -> np.ndarray:
    """LIB/generate_grid.cpp:11-107: points 1 .. num_grd-1."""
    sp = F(spacer)
    inv = 1.0 / float(sp)
    coarse = F(F(inv) - F(int(inv))) > 0.001
    seen = set()
    pts = []
    for x, y, z, r in np.asarray(spheres, F):
        c = (x, y, z)
        if coarse:
            mn = [F(F(int(F(F(c[k] - r) / sp))) * sp) for k in range(3)]
            mx = [F(F(int(F(F(c[k] + r) / sp)) + 1.0) * sp) for k in range(3)]
        else:
            mn = [F(int(F(F(c[k] - r) - sp))) for k in range(3)]
            mx = [F(int(float(F(F(c[k] + r) + sp)) + 1.0)) for k in range(3)]
        xs, ys, zs = (_seq(mn[k], mx[k], sp) for k in range(3))
        if not (len(xs) and len(ys) and len(zs)):
            continue
        Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")          # z outer, x inner
        P = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
        d2 = _sq3(np.array([x, y, z], F)[None, :], P)
        P = P[d2 < F(r * r)]
        keys = np.stack([_roundf_int(F(1000.0) * P[:, k]) for k in range(3)], axis=1)
        for q, kk in zip(P, map(tuple, keys.tolist())):
            if kk not in seen:
                seen.add(kk)
                pts.append(q)
    return np.array(pts, F).reshape(-1, 3)


def _bndang(a, b, c):
    cosa = absu = absv = F(0.0)
    for i in range(3):
        u = F(a[i] - b[i])
        v = F(c[i] - b[i])
        cosa = F(cosa + F(u * v))
        absu = F(absu + F(u * u))
        absv = F(absv + F(v * v))
    cosa = F(cosa / np.sqrt(F(absu * absv)))
    return F(float(_acosf(float(cosa))) * 180.0 / PI)


def _dihedral(a1, a2, a3, a4):
    r1 = (a2 - a1).astype(F)
    r2 = (a3 - a2).astype(F)
    r3 = (a4 - a3).astype(F)

    def cross(u, v):
        return np.array([F(F(u[1] * v[2]) - F(u[2] * v[1])), F(F(u[2] * v[0]) - F(u[0] * v[2])),
                         F(F(u[0] * v[1]) - F(u[1] * v[0]))], F)

    def dot(u, v):
        return F(F(F(u[0] * v[0]) + F(u[1] * v[1])) + F(u[2] * v[2]))

    n1, n2 = cross(r1, r2), cross(r2, r3)
    psin = F(dot(n1, r3) * np.sqrt(dot(r2, r2)))
    pcos = dot(n1, n2)
    return F(F(57.2958) * F(_atan2f(float(psin), float(pcos))))


def calc_cleftic(points, ori) -> np.ndarray:
    """LIB/calc_cleftic.cpp:5-34: (dis, ang, dih) of each point from the
    frame ca = ori + x, cb = ori, cc = ori + y."""
    o = np.asarray(ori, F)
    ca = np.array([F(o[0] + F(1.0)), o[1], o[2]], F)
    cb = o.copy()
    cc = np.array([o[0], F(o[1] + F(1.0)), o[2]], F)
    out = np.zeros((len(points), 3), F)
    for i, ci in enumerate(np.asarray(points, F)):
        out[i, 0] = np.sqrt(_sq3(ci, ca))
        out[i, 1] = _bndang(ci, ca, cb)
        out[i, 2] = _dihedral(ci, ca, cb, cc)
    return out


def cleft_grid(xyz, radius, ori, spacer=0.375, params: CleftParams | None = None) -> CleftGrid:
    """The reference's automatic binding-site grid for a receptor (atoms in
    the reference's order, e.g. ReceptorFile.xyz/.radius) and the ori frame
    (e.g. ReceptorFile.ori)."""
    sph = detect_cleft(xyz, radius, params)
    pts = generate_grid(sph, spacer)
    return CleftGrid(spheres=sph, points=pts, ic=calc_cleftic(pts, ori))
