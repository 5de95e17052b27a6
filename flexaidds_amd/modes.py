"""Binding modes of docked poses and their thermodynamics.

Mirrors the reference's BindingPopulation / BindingMode
(LIB/BindingMode.h, LIB/BindingMode.cpp:18-55, :114-147, :224-335) over the
modes that the device's CF clustering (fa_cluster_cf, the reference's
default CLUSTA CF) finds: a mode's poses are its head and members, each
pose's energy is the energy it was clustered on (Pose::CF = app_evalue,
LIB/BindingMode.cpp:613-625).

  BindingMode.thermodynamics()  StatMechEngine over the mode's poses, one
                                sample each (rebuild_engine, :239-251)
  BindingMode.boltzmann_weights()  the engine's normalised weights (:330-335)
  BindingPopulation.modes       ordered by free energy (Entropize, :37-45)
  global_ensemble()             every pose of every mode with its mode's
                                Boltzmann weight as multiplicity (:114-147);
                                StatMechEngine::add_sample takes an int, so
                                the weight is truncated: only poses whose
                                weight is exactly 1 (a mode's only pose, or
                                one whose companions underflow) count
  delta_G(a, b), deltaG_matrix()   free-energy differences (:99-102, :207-222)
  super_cluster_ensemble()      the dominant energy basin (:153-180): the
                                FastOPTICS super-cluster of every pose energy
                                (fa_super_cluster, minPts = max(4, n / 20)),
                                one sample per member

The vibrational and co-translational corrections the reference adds to a
mode's free energy are 0 unless their (default-off) models are enabled;
they are not part of this path.  The statistics run on the device
(fa_thermo_host, fa_boltzmann_device); there is no CPU fallback.
"""
from __future__ import annotations

import numpy as np

from . import hip


class BindingMode:
    def __init__(self, population: "BindingPopulation", record: dict, poses: np.ndarray):
        self.population = population
        self.top = int(record["top"])
        self.freq = int(record["freq"])
        self.tcf = float(record["tcf"])
        self.acf = float(record["acf"])
        self.poses = poses                               # pose indices, increasing
        self.energies = population.energies[poses]
        self._thermo = None

    def size(self) -> int:
        return len(self.poses)

    def thermodynamics(self) -> dict:
        if self._thermo is None:
            self._thermo = self.population.ctx.thermo(self.energies, self.population.temperature)
        return self._thermo

    def free_energy(self) -> float:
        return self.thermodynamics()["free_energy"]

    def boltzmann_weights(self) -> np.ndarray:
        """n_i exp(-beta E_i) / sum_j n_j exp(-beta E_j) over the mode's poses."""
        import torch
        ctx = self.population.ctx
        dev = torch.device("cuda", ctx.device)
        e = torch.tensor(self.energies, dtype=torch.float64, device=dev)
        w = torch.empty_like(e)
        with torch.cuda.device(dev):
            s = torch.cuda.current_stream()
            ctx.boltzmann_device(e.data_ptr(), 1, len(self.energies), self.population.temperature, w.data_ptr(),
                                 want_log_z=True, stream=s.cuda_stream)
            s.synchronize()
        raw = w.cpu().numpy()
        tot = 0.0
        for x in raw:                                    # the reference's sequential sum (statmech.cpp:192-194)
            tot += x
        return raw / tot if tot > 0.0 else np.zeros_like(raw)


class BindingPopulation:
    """Binding modes of `ligand`'s poses on `ctx` (clustered on the device)."""

    def __init__(self, ctx: hip.Context, ligand: int, genes, energies, temperature: float = 300.0,
                 rmsd: float = 2.0, max_results: int = 10):
        self.ctx = ctx
        self.temperature = float(temperature)
        self.energies = np.ascontiguousarray(energies, np.float64)
        self.clusters = ctx.cluster_cf(ligand, genes, self.energies, temperature=temperature, rmsd=rmsd,
                                       max_results=max_results)
        member = self.clusters["member"]
        modes = []
        for rec in self.clusters["modes"]:
            if rec["freq"] == 0:                         # a record the reference never filled
                continue
            modes.append(BindingMode(self, rec, np.flatnonzero(member == rec["top"])))
        # add_BindingMode's partition function: exp(-app / T) per pose (Pose::Pose, :625)
        self.partition_function = float(sum(np.exp(-m.energies / self.temperature).sum() for m in modes))
        modes.sort(key=lambda m: m.free_energy())        # Entropize
        self.modes = modes

    def size(self) -> int:
        return len(self.modes)

    def delta_G(self, a: int, b: int) -> float:
        return self.modes[a].free_energy() - self.modes[b].free_energy()

    def deltaG_matrix(self) -> np.ndarray:
        n = len(self.modes)
        m = np.zeros((n, n))
        for i in range(n):
            for j in range(i + 1, n):
                m[i, j] = self.delta_G(i, j)
                m[j, i] = -m[i, j]
        return m

    def global_ensemble(self) -> dict | None:
        """Thermodynamics of every pose with multiplicity int(weight); None
        when no pose keeps a multiplicity (the reference's engine then has
        only zero counts)."""
        e, c = [], []
        for m in self.modes:
            e.append(m.energies)
            c.append(m.boltzmann_weights().astype(np.int32))     # add_sample(double -> int)
        if not e:
            return None
        e, c = np.concatenate(e), np.concatenate(c)
        if not c.any():
            return None
        return self.ctx.thermo(e, self.temperature, count=c)

    def super_cluster_ensemble(self) -> dict | None:
        """Thermodynamics of the poses in the super-cluster of all pose
        energies (modes in free-energy order, each mode's poses in order),
        one sample each in the flood fill's order; five poses or fewer fall
        back to global_ensemble(), as in the reference.  fa_super_cluster
        takes at most FA_SUPER_CLUSTER_MAX (8192) poses and raises above."""
        if not self.modes:
            return self.global_ensemble()
        e = np.concatenate([m.energies for m in self.modes])
        if len(e) <= 4:
            return self.global_ensemble()
        members = self.ctx.super_cluster(e, max(4, len(e) // 20))["members"]
        if len(members) == 0:
            return self.global_ensemble()
        return self.ctx.thermo(e[members], self.temperature)
