"""Served docking: a request loop over resident receptors.

Each receptor is uploaded once into its own context (resident in HBM with
its ligand library) and gets a docking pool (fa_pool_*) whose slot count
follows from an HBM budget.  The loop (fa_serve_run, in the library)
admits requests as their arrival time passes and steps every pool that has
work as soon as that pool's previous step has ended (each step enqueues one
generation of all of the pool's resident jobs on the pool's stream and
returns); finished jobs are collected then.  The pools' launches overlap on
the device, and jobs that arrive while others run join their pool at its
next generation.

Each receptor also has a TargetServer (flexaidds_amd.target): every
finished job is registered there as a completed docking session under its
request's ligand name, so the receptor's grand partition function, binding
probabilities, selectivities and ranking follow the served stream
(LIB/TargetServer.h:81-102).

There is no CPU fallback: the pools run through the HIP library.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import hip, synth, target


@dataclass
class Served:
    results: dict = field(default_factory=dict)      # request index -> result dict
    latency_s: dict = field(default_factory=dict)    # request index -> completion - arrival
    finished: list = field(default_factory=list)      # request indices in completion order
    wall_s: float = 0.0
    steps: int = 0
    poses: int = 0

    def summary(self) -> dict:
        lat = np.array(sorted(self.latency_s.values()))
        return {"requests": len(lat), "wall_s": round(self.wall_s, 3), "steps": self.steps, "poses": self.poses,
                "poses_per_s": round(self.poses / self.wall_s, 1) if self.wall_s > 0 else 0.0,
                "latency_p50_s": round(float(np.percentile(lat, 50)), 4) if len(lat) else None,
                "latency_p99_s": round(float(np.percentile(lat, 99)), 4) if len(lat) else None,
                "latency_max_s": round(float(lat.max()), 4) if len(lat) else None}


class Server:
    """Resident receptors with one docking pool each."""

    def __init__(self, systems, population=1024, offspring=None, max_slots=64, hbm_budget_per_pool=4 << 30,
                 temperature=300.0, fitness="LINEAR", device=0, max_running=0):
        self.systems = systems
        self.ctx, self.pools, self.limits, self.targets = [], [], [], []
        for s in systems:
            c = hip.Context(device)
            c.upload_system(s)
            self.ctx.append(c)
            self.pools.append(hip.Pool(c, population=population, offspring=offspring, max_slots=max_slots,
                                       hbm_budget=hbm_budget_per_pool, temperature=temperature, fitness=fitness,
                                       max_running=max_running))
            self.limits.append([synth.ga_limits(l) for l in s.ligands])
            self.targets.append(target.TargetServer(c, temperature))

    def close(self):
        for p in self.pools:
            p.close()
        for c in self.ctx:
            c.close()

    def run(self, requests, realtime=True) -> Served:
        """Serve `requests` (dicts with t, receptor, ligand, generations, seed
        and optionally name, the ligand's name for the receptor's
        TargetServer, by default "L<ligand>") in arrival order.  With
        realtime, a request is admitted once the wall clock passes its
        arrival time; otherwise all arrive at once.

        The loop is the library's (fa_serve_run, C++): each pool is stepped
        as soon as its own previous step has ended (a non-blocking query),
        so one pool's slow launch does not hold the others back and the
        pools' launches keep overlapping on the device.  Finished jobs are
        then registered with their receptor's TargetServer in completion
        order."""
        reqs = [{"t": r["t"], "pool": r["receptor"], "ligand": r["ligand"], "generations": r["generations"],
                 "seed": r["seed"], "limits": self.limits[r["receptor"]][r["ligand"]]} for r in requests]
        done, st = hip.serve_run(self.pools, reqs, realtime=realtime)
        out = Served(wall_s=st["wall_s"], steps=st["steps"], poses=st["poses"])
        for i, lat, res in done:
            r = requests[i]
            p = r["receptor"]
            out.results[i] = res
            out.latency_s[i] = lat
            out.finished.append(i)
            lig = self.systems[p].ligands[r["ligand"]]
            self.targets[p].register_result(target.session_from_job(
                self.targets[p], r.get("name", f"L{r['ligand']}"), res, lig))
        return out
