"""Energy-matrix files (read_emat, LIB/read_emat.cpp:8-149) into the
fa_params layout.

The file holds one line per unordered type pair (i <= j, row-major over the
upper triangle), so n lines give ntypes = (sqrt(1 + 8 n) - 1) / 2.  Text
after an '=' is the pair's data: one number is a weight-1 pair (a value per
unit contact area); an even count is a density function given as x y x y ...
nodes (relative contact area -> energy).  Values are stored in single
precision, as the reference's float fields hold them.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


@dataclass
class EnergyMatrix:
    ntypes: int
    emat: np.ndarray          # (ntypes*ntypes,) float32, weight-1 values (0 for density pairs)
    dens_off: np.ndarray      # (ntypes*ntypes + 1,) int32
    dens_xy: np.ndarray       # (n_nodes, 2) float32

    @property
    def has_density(self) -> bool:
        return int(self.dens_off[-1]) > 0


def parse_emat(text: str) -> EnergyMatrix:
    # only newline-terminated lines count, as the reference's line loop keeps them
    lines = [ln for ln in text.splitlines(keepends=True) if ln.endswith("\n")]
    n = len(lines)
    z = (math.sqrt(1.0 + 8.0 * n) - 1.0) / 2.0
    if abs(z - int(z)) > 0.001:
        raise ValueError(f"energy matrix: {n} lines is not a triangular count")
    nt = int(z + 0.001)
    if nt < 1:
        raise ValueError("energy matrix: no pairs")
    emat = np.zeros(nt * nt, np.float32)
    nodes: dict[int, np.ndarray] = {}
    it = iter(lines)
    for i in range(nt):
        for j in range(i, nt):
            line = next(it)
            if "=" in line:
                line = line[line.index("=") + 1:]
            vals = line.split()
            if len(vals) == 1:
                emat[i * nt + j] = emat[j * nt + i] = np.float32(float(vals[0]))
            elif len(vals) % 2 == 0 and vals:
                xy = np.array([float(v) for v in vals], np.float64).astype(np.float32).reshape(-1, 2)
                nodes[i * nt + j] = nodes[j * nt + i] = xy
            else:
                raise ValueError(f"energy matrix: invalid number of values for pair {i + 1}-{j + 1}")
    off = np.zeros(nt * nt + 1, np.int32)
    chunks = []
    for k in range(nt * nt):
        xy = nodes.get(k)
        off[k + 1] = off[k] + (0 if xy is None else len(xy))
        if xy is not None:
            chunks.append(xy)
    dens = np.concatenate(chunks) if chunks else np.zeros((0, 2), np.float32)
    return EnergyMatrix(nt, emat, off, dens)


def read_emat(path: str) -> EnergyMatrix:
    with open(path, "r") as f:
        return parse_emat(f.read())


def write_emat(m: EnergyMatrix) -> str:
    """The file text of an energy matrix (upper triangle, one pair per line)."""
    out = []
    nt = m.ntypes
    for i in range(nt):
        for j in range(i, nt):
            k = i * nt + j
            b, e = int(m.dens_off[k]), int(m.dens_off[k + 1])
            if e > b:
                body = " ".join(f"{float(x)!r} {float(y)!r}" for x, y in m.dens_xy[b:e])
            else:
                body = repr(float(m.emat[k]))
            out.append(f"{i + 1:>3}-{j + 1:<3} = {body}\n")
    return "".join(out)


def apply(params, m: EnergyMatrix):
    """Set a synth.Params' energy terms from an energy matrix."""
    params.ntypes = m.ntypes
    params.emat = m.emat
    params.dens_off = m.dens_off if m.has_density else None
    params.dens_xy = m.dens_xy if m.has_density else None
    return params
