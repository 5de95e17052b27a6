"""Served docking: a request loop over resident receptors.

Each receptor is uploaded once into its own context (resident in HBM with
its ligand library) and gets a docking pool (fa_pool_*) whose slot count
follows from an HBM budget.  The loop admits requests as their arrival time
passes and steps every pool that has work as soon as that pool's previous
step has ended (each step enqueues one generation of all of the pool's
resident jobs on the pool's stream and returns); finished jobs are collected
then.  The pools' launches overlap on the device, and jobs that arrive while
others run join their pool at its next generation.

Each receptor also has a TargetServer (flexaidds_amd.target): every
finished job is registered there as a completed docking session under its
request's ligand name, so the receptor's grand partition function, binding
probabilities, selectivities and ranking follow the served stream
(LIB/TargetServer.h:81-102).

There is no CPU fallback: the pools run through the HIP library.
"""
from __future__ import annotations

import time
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

    def run(self, requests, realtime=True, on_step=None) -> Served:
        """Serve `requests` (dicts with t, receptor, ligand, generations, seed
        and optionally name, the ligand's name for the receptor's
        TargetServer, by default "L<ligand>") in arrival order.  With realtime, a request is admitted once the wall
        clock passes its arrival time; otherwise all arrive at once.

        Each pool is stepped as soon as its own previous step has ended (a
        non-blocking query), so one pool's slow launch does not hold the
        others back and the pools' launches keep overlapping on the device."""
        out = Served()
        order = sorted(range(len(requests)), key=lambda i: requests[i]["t"])
        nxt = 0
        owner = {}                                   # (pool, job id) -> request index
        P = len(self.pools)
        busy = [0] * P                               # resident + queued jobs per pool
        inflight = [False] * P                       # a step enqueued and not yet collected
        t0 = time.perf_counter()
        while nxt < len(order) or any(busy) or any(inflight):
            now = time.perf_counter() - t0
            while nxt < len(order) and (not realtime or requests[order[nxt]]["t"] <= now):
                i = order[nxt]
                r = requests[i]
                p = r["receptor"]
                jid = self.pools[p].submit(r["ligand"], self.limits[p][r["ligand"]], r["generations"], r["seed"])
                owner[(p, jid)] = i
                busy[p] += 1
                nxt += 1
            progressed = False
            for p in range(P):
                if inflight[p]:
                    if self.pools[p].busy():
                        continue
                    for res in self.pools[p].collect():
                        i = owner.pop((p, res["id"]))
                        done = time.perf_counter() - t0
                        out.results[i] = res
                        out.finished.append(i)
                        out.latency_s[i] = done - (requests[i]["t"] if realtime else 0.0)
                        out.poses += res["n_poses"]
                        r = requests[i]
                        lig = self.systems[p].ligands[r["ligand"]]
                        self.targets[p].register_result(target.session_from_job(
                            self.targets[p], r.get("name", f"L{r['ligand']}"), res, lig))
                        busy[p] -= 1
                    inflight[p] = False
                    progressed = True
                if busy[p] > 0:
                    self.pools[p].step()
                    inflight[p] = True
                    out.steps += 1
                    progressed = True
            if progressed:
                if on_step:
                    on_step(out)
            elif not any(inflight) and nxt < len(order):
                time.sleep(max(0.0, requests[order[nxt]]["t"] - (time.perf_counter() - t0)))
            else:
                time.sleep(0.0002)
        out.wall_s = time.perf_counter() - t0
        return out
