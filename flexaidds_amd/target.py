"""Target-level state over served docking: every finished session of a
receptor feeds one grand canonical partition function.

Mirrors the reference's TargetServer session API (LIB/TargetServer.h:39-52,
:81-102; LIB/TargetServer.cpp create_session / register_result /
binding_probability / selectivity_ratio / rank_ligands /
completed_sessions) and the GrandPartitionFunction it owns
(LIB/GrandPartitionFunction.h, .cpp:18-205): the same method names,
argument meanings and errors (ValueError where the reference throws
std::invalid_argument).

For a site that is empty or holds one of N ligands,
    Xi = 1 + sum_i z_i Z_i,   z_i = c_i / 1 M,
so ln Xi = LSE(0, ln c_1 + ln Z_1, ...), p_empty = 1/Xi and
p_bound(i) = z_i Z_i / Xi (LIB/GrandPartitionFunction.cpp:96-137).

The registry (ligand name -> ln Z_i, ln c_i, ln z_i Z_i) is host state, as
in the reference.  ln Xi, p_empty and every p_bound are computed on the
device by fa_grand_device over a device copy of the registry, recomputed
only after a change (the reference caches ln Xi the same way,
:113-118).  There is no CPU fallback for them.

A served job's ln Z is the ensemble of every pose the job scored
(fa_job_result.thermo, the same fold as the batched GA's per-job ln Z);
register_result takes the reference's 1 M default concentration.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

KB_KCAL = 0.001987206          # LIB/statmech.h:26


@dataclass
class DockingSession:
    """LIB/TargetServer.h:39-52 (conformer populations omitted: one receptor
    model per context)."""
    session_id: int
    ligand_name: str
    log_Z: float = 0.0
    completed: bool = False
    n_poses: int = 0
    best_center: tuple = (0.0, 0.0, 0.0)
    best_energy: float = 0.0


@dataclass
class LigandRank:
    """GrandPartitionFunction::LigandRank (LIB/GrandPartitionFunction.h)."""
    name: str
    log_Z: float
    dG: float
    p_bound: float


@dataclass
class _Entry:
    log_Z: float
    log_c: float
    log_zZ: float


def _log_c(concentration_M: float) -> float:
    # LIB/GrandPartitionFunction.cpp:21-26
    if not concentration_M > 0.0:
        raise ValueError("Concentration must be positive")
    if concentration_M > 1e3:
        raise ValueError("Concentration > 1000 M -- did you pass uM or nM without conversion to M?")
    return math.log(concentration_M)


class GrandPartitionFunction:
    """Grand canonical partition function over the ligands of one site,
    evaluated on `ctx`'s device."""

    def __init__(self, ctx, temperature_K: float = 300.0):
        if not temperature_K > 0.0:
            raise ValueError("Temperature must be positive")
        self.ctx = ctx
        self.T = float(temperature_K)
        self.beta = 1.0 / (KB_KCAL * self.T)
        self._lig: dict[str, _Entry] = {}
        self._cache = None                      # (names, log_xi, p_empty, p_bound) or None

    # registration (LIB/GrandPartitionFunction.cpp:18-94)
    def add_ligand(self, name: str, log_Z: float, concentration_M: float = 1.0):
        lc = _log_c(concentration_M)
        if name in self._lig:
            raise ValueError(f"Ligand '{name}' already registered")
        self._lig[name] = _Entry(float(log_Z), lc, lc + float(log_Z))
        self._cache = None

    def add_or_overwrite(self, name: str, log_Z: float, concentration_M: float = 1.0):
        lc = _log_c(concentration_M)
        self._lig[name] = _Entry(float(log_Z), lc, lc + float(log_Z))
        self._cache = None

    def overwrite_ligand(self, name: str, new_log_Z: float):
        e = self._get(name)
        e.log_Z = float(new_log_Z)
        e.log_zZ = e.log_c + e.log_Z
        self._cache = None

    def merge_ligand(self, name: str, new_log_Z: float):
        """ln zZ <- LSE(ln zZ, ln c + new ln Z): two independent ensembles of
        one ligand at its registered concentration."""
        e = self._get(name)
        a, b = e.log_zZ, e.log_c + float(new_log_Z)
        m = max(a, b)
        e.log_zZ = m + math.log(math.exp(a - m) + math.exp(b - m))
        e.log_Z = e.log_zZ - e.log_c
        self._cache = None

    def remove_ligand(self, name: str):
        if self._lig.pop(name, None) is None:
            raise ValueError(f"Ligand '{name}' not found")
        self._cache = None

    # device evaluation
    def _get(self, name: str) -> _Entry:
        e = self._lig.get(name)
        if e is None:
            raise ValueError(f"Ligand '{name}' not found")
        return e

    def _evaluate(self):
        if self._cache is not None:
            return self._cache
        import torch
        names = list(self._lig)
        n = len(names)
        dev = torch.device("cuda", self.ctx.device)
        lz = torch.tensor([self._lig[k].log_Z for k in names] or [0.0], dtype=torch.float64, device=dev)
        lc = torch.tensor([self._lig[k].log_c for k in names] or [0.0], dtype=torch.float64, device=dev)
        out = torch.empty(n + 2, dtype=torch.float64, device=dev)
        with torch.cuda.device(dev):
            s = torch.cuda.current_stream()
            self.ctx.grand_device(lz.data_ptr(), 1, lc.data_ptr(), n, out.data_ptr(), s.cuda_stream)
            s.synchronize()
        o = out.cpu().numpy()
        self._cache = (names, float(o[0]), float(o[1]), dict(zip(names, o[2:].tolist())))
        return self._cache

    # queries (LIB/GrandPartitionFunction.cpp:119-205)
    def log_Xi(self) -> float:
        return self._evaluate()[1]

    def empty_probability(self) -> float:
        return self._evaluate()[2]

    def binding_probability(self, name: str) -> float:
        self._get(name)
        return self._evaluate()[3][name]

    def F_bound(self, name: str) -> float:
        return -(1.0 / self.beta) * self._get(name).log_Z

    def delta_G_bind(self, name: str, F_ref: float = 0.0) -> float:
        return self.F_bound(name) - F_ref

    def log_selectivity(self, a: str, b: str) -> float:
        ea, eb = self._get(a), self._get(b)
        return ea.log_zZ - eb.log_zZ

    def selectivity(self, a: str, b: str) -> float:
        d = self.log_selectivity(a, b)
        if d > 700.0:
            return math.inf
        if d < -700.0:
            return 0.0
        return math.exp(d)

    def log_intrinsic_selectivity(self, a: str, b: str) -> float:
        ea, eb = self._get(a), self._get(b)
        return ea.log_Z - eb.log_Z

    def rank(self) -> list[LigandRank]:
        """Every ligand by dG = -kT ln Z ascending (most favourable first)."""
        _, _, _, pb = self._evaluate()
        kT = 1.0 / self.beta
        r = [LigandRank(k, e.log_Z, -kT * e.log_Z, pb[k]) for k, e in self._lig.items()]
        r.sort(key=lambda x: x.dG)
        return r

    def num_ligands(self) -> int:
        return len(self._lig)

    def has_ligand(self, name: str) -> bool:
        return name in self._lig

    def temperature(self) -> float:
        return self.T

    def all_log_Z(self):
        return [(k, e.log_Z) for k, e in self._lig.items()]

    def all_log_zZ(self):
        return [(k, e.log_zZ) for k, e in self._lig.items()]


class TargetServer:
    """One receptor's sessions and their grand partition function.

    `ctx` is the receptor's context (its device evaluates ln Xi)."""

    def __init__(self, ctx, temperature_K: float = 300.0):
        self.ctx = ctx
        self.T = float(temperature_K)
        self._xi = GrandPartitionFunction(ctx, temperature_K)
        self._next_id = 0
        self._completed = 0

    def temperature(self) -> float:
        return self.T

    def create_session(self, ligand_name: str) -> DockingSession:
        s = DockingSession(self._next_id, ligand_name)
        self._next_id += 1
        return s

    def register_result(self, session: DockingSession):
        """An incomplete session is ignored; a completed one enters Xi with
        insert-or-overwrite at 1 M (LIB/TargetServer.cpp register_result)."""
        if not session.completed:
            return
        self._xi.add_or_overwrite(session.ligand_name, session.log_Z)
        self._completed += 1

    def grand_partition(self) -> GrandPartitionFunction:
        return self._xi

    def binding_probability(self, ligand_name: str) -> float:
        return self._xi.binding_probability(ligand_name)

    def selectivity_ratio(self, a: str, b: str) -> float:
        return self._xi.selectivity(a, b)

    def rank_ligands(self) -> list[LigandRank]:
        return self._xi.rank()

    def completed_sessions(self) -> int:
        return self._completed


def session_from_job(server: TargetServer, name: str, result: dict, ligand=None) -> DockingSession:
    """A completed session from a finished pool job (fa_job_result as
    hip.Pool.collect returns it): ln Z of every pose the job scored, the
    best pose's energy, and its cleft-grid point as the binding centre when
    the ligand's first gene is the grid index."""
    s = server.create_session(name)
    s.log_Z = float(result["thermo"]["log_Z"])
    s.n_poses = int(result["n_poses"])
    s.best_energy = float(result["best"]["cf"])
    if ligand is not None and len(result["best_ic"]) and int(ligand.gene_typ[0]) == -1:
        k = int(result["best_ic"][0])
        if 0 <= k < len(ligand.grid_xyz):
            s.best_center = tuple(float(v) for v in np.asarray(ligand.grid_xyz[k]))
    s.completed = True
    return s
