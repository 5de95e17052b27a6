#!/usr/bin/env python3
"""Population-evaluation benchmark on one MI355X.

A step is one pass of the hot path over one batch: one GA generation's
population (2,000 chromosomes of the 45-atom / 10-torsion ligand against the
4,000-atom receptor) decoded, rebuilt and scored with the Voronoi contact
function, inputs already resident in HBM.  `value` is poses/s of the plain
path (one ligand population per launch).  Beside it:

  ga_dock   the full plain dock of that shape on the device GA (2,000 x 500
            generations, one generation captured as a HIP graph)
  batched   the batched screen: many uneven ligands (12-80 atoms, 0-15
            torsions, 512 chromosomes each) evolved together, every job's
            children in one pose launch per generation
  served    the served stream: 8 resident receptors (2k-10k atoms), Poisson
            requests (1,024 x 100 generations each) admitted into per-receptor
            docking pools as they arrive; p50/p99 latency, aggregate poses/s,
            and the plain path on the same mix (one request at a time, one
            synchronous launch per generation)
  grand     the blind dock of 64 competing ligands on a 20,000-atom globule
            (32,768 chromosomes each, one batched GA) and the grand partition
            function over them; a few generations timed, 300 projected
  real_inputs  the plain launch and a device-GA dock on a receptor read from
            a PDB file (the reference's RNA target) with the automatic
            binding-site grid built from it
  binding_modes  the reference's default clustering of docked poses into
            binding modes on the device, over 200,000 poses, and the
            FastOPTICS super-cluster of 8,192 pose energies (ga_dock also
            reports the modes and super-cluster of its final population and
            the same dock with duplicate rejection, whose snapshot of every
            generation is sorted, thinned and clustered as the reference's
            run ends)

The roofline numbers come from the kernel's own operation counters (the
algorithmic FLOP model, see include/fa_types.h fa_opcount) and HIP-event
timing of each launch on the launch stream.  The CPU baseline is the
reference's own contact function (its sources compiled with its flags, one
process per host core) on a bounded sample of the same workload, with the
C++ restatement (OpenMP over poses) on the same cores beside it.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Voronoi-CF poses/s on 1 MI355X, batched vs plain; p50/p99 dock latency"
# the compute ceiling ("mfma" in the report's two-way hbm/mfma vocabulary):
# MI355X's FP64 peak, the same 78.6 TFLOP/s for its vector and matrix units;
# the kernel issues FP64 vector instructions (no MFMA: nothing is a dense
# contraction)
FP64_PEAK_TFLOPS = 78.6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _host_cores():
    """Cores this run may use: the OpenMP share of the box when set, else all."""
    n = os.cpu_count() or 1
    try:
        return max(1, min(n, int(os.environ.get("OMP_NUM_THREADS", n))))
    except ValueError:
        return n


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _ref_worker(job):
    """One core's share of the CPU baseline: the reference's own contact
    function (compiled from its sources with its flags, -O3 -ffast-math,
    oracle/_ref/libfa_ref_fast.so) over chunks of the poses until the time
    budget is spent.  The reference keeps static state in Vcontacts, so each
    core runs it in a process of its own."""
    system, genes, budget_s, chunk = job
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    from flexaidds_amd import abi
    lib = oracle_lib.load_ref(fast=True)
    desc = abi.Descriptors(system.receptor, system.ligands, system.params)
    done, k, t0 = 0, 0, time.perf_counter()
    while time.perf_counter() - t0 < budget_s:
        g = genes[(k * chunk) % len(genes):][:chunk]
        oracle_lib.ref_eval(lib, desc, 0, g, cap=1)
        done += len(g)
        k += 1
    return done, time.perf_counter() - t0


def cpu_baseline(system, genes, budget_s=12.0):
    """The reference's CPU path on the host cores (one process per core, the
    reference's own build), and beside it the C++ restatement (the oracle,
    OpenMP over poses) on the same cores; poses/s of the same workload."""
    import concurrent.futures as cf
    import multiprocessing as mp
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    from flexaidds_amd import abi
    cores = _host_cores()
    out = {"unit": "poses/s", "cores": cores, "cpu_model": _cpu_model(), "nproc": os.cpu_count()}
    # restatement: OpenMP over poses on every core
    lib = oracle_lib.load_oracle()
    desc = abi.Descriptors(system.receptor, system.ligands, system.params)
    done, k, t0 = 0, 0, time.perf_counter()
    while time.perf_counter() - t0 < budget_s / 2:
        g = genes[(k * 2000) % len(genes):][:2000]
        oracle_lib.oracle_eval(lib, desc, 0, g, n_threads=cores)
        done += len(g)
        k += 1
    port = done / (time.perf_counter() - t0)
    out["restatement"] = {"value": round(port, 1), "threads": cores, "poses": done,
                          "kind": "oracle C++ FP64 restatement, OpenMP over poses"}
    if not os.path.exists(oracle_lib.REF_FAST_SO):
        out.update({"value": round(port, 1), "kind": "port",
                    "sample": f"{done} poses, oracle restatement on {cores} threads (reference build absent)"})
        return out
    ctx = mp.get_context("spawn")
    jobs = [(system, genes[(i * 977) % len(genes):], budget_s, 250) for i in range(cores)]
    t0 = time.perf_counter()
    with cf.ProcessPoolExecutor(max_workers=cores, mp_context=ctx) as ex:
        res = list(ex.map(_ref_worker, jobs))
    wall = time.perf_counter() - t0
    poses = sum(r[0] for r in res)
    per_core = [r[0] / r[1] for r in res]
    ref = sum(per_core)
    out.update({"value": round(ref, 1), "kind": "reference",
                "single_core": round(float(np.median(per_core)), 1),
                "restatement_vs_reference": round(port / ref, 3),
                "sample": f"{poses} poses of the same workload in {cores} processes x {budget_s:.0f} s "
                          f"({wall:.1f} s wall incl. start-up): the reference's Vcontacts/vcfunction/ic2cf/buildcc "
                          f"compiled from its sources with -O3 -ffast-math, one process per core"})
    return out


def population_sweep():
    """Launch time against poses per launch (plain config-1 workload): one
    pose alone shows the single-pose latency, large launches the throughput
    once every resident wave has several poses."""
    import torch
    from flexaidds_amd import abi, hip, synth
    system = synth.make_config1(pop=16384)
    lig = system.ligands[0]
    genes = synth.random_genes(np.random.default_rng(7), lig, 16384)
    ctx = hip.Context(max_pop=16384)
    ctx.upload_system(system)
    d_genes = torch.from_numpy(genes).cuda()
    d_out = torch.zeros(16384 * abi.CF_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
    st = torch.cuda.current_stream()
    rows = []
    for n in (1, 16, 256, 1024, 2048, 4096, 8192, 16384):
        ctx.eval_device(n, d_genes.data_ptr(), lig.n_genes, 0, d_out.data_ptr(), st.cuda_stream)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(3):
            ctx.eval_device(n, d_genes.data_ptr(), lig.n_genes, 0, d_out.data_ptr(), st.cuda_stream)
        e1.record(st)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / 3
        rows.append({"poses": n, "ms": round(ms, 4), "poses_per_s": round(n / ms * 1e3, 1)})
        log(json.dumps(rows[-1]))
    print(json.dumps({"sweep": rows, "launch_waves": ctx.launch_waves()}), flush=True)
    ctx.close()


def measured_traffic(pop):
    """HBM bytes per launch of the pose kernel for this workload, from the
    committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary (calibrated on a copy
    kernel), or None when no summary covers it."""
    path = os.path.join(ROOT, "profiles", "cf_pose_kernel_traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    want = f"plain config-1 generation with flexible side chains, {pop} poses"
    return rec.get("bytes_per_launch") if rec.get("workload") == want else None


def _unscored(ctx):
    """Poses a context's launches did not score since its last reset."""
    t = ctx.totals()
    return {"capacity": t["capacity"], "grid_range": t["grid_range"], "bad_ligand": t["bad_ligand"],
            "undefined_face": t["undefined_face"]}


def config0_leg(args):
    """The reference's CPU-runnable case: plain CF scoring of 1,000 poses
    (1,500-atom receptor, 30-atom / 6-torsion ligand) in one launch, then
    Z and F at 300 K in FP64 on the device over the scored poses' CF."""
    import torch
    from flexaidds_amd import abi, hip, synth
    s = synth.make_config0(n_poses=1000)
    G = s.ligands[0].n_genes
    ctx = hip.Context(max_pop=1000)
    ctx.upload_system(s)
    ctx.reset_totals()
    d_genes = torch.from_numpy(np.ascontiguousarray(s.genes[0])).cuda()
    d_out = torch.zeros(1000 * abi.CF_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
    st = torch.cuda.current_stream()
    for _ in range(3):
        ctx.eval_device(1000, d_genes.data_ptr(), G, 0, d_out.data_ptr(), st.cuda_stream)
    n = max(args.steps, 5)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(st)
    for _ in range(n):
        ctx.eval_device(1000, d_genes.data_ptr(), G, 0, d_out.data_ptr(), st.cuda_stream)
    e1.record(st)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / n
    res = np.frombuffer(d_out.cpu().numpy().tobytes(), abi.CF_DTYPE)
    ok = res["status"] == 0
    th = ctx.thermo(res["cf"][ok], 300.0)
    out = {"poses": 1000, "launches": n, "ms_per_launch": round(ms, 4), "poses_per_s": round(1000 / ms * 1e3, 1),
           "scored_fraction": round(float(ok.mean()), 4), "log_Z_300K": round(th["log_Z"], 6),
           "free_energy_300K": round(th["free_energy"], 6), "unscored": _unscored(ctx)}
    ctx.close()
    return out


def ga_dock_leg(ctx, system, args):
    """The plain full dock: one ligand, population = offspring = --pop, the
    given generations as replays of one captured generation."""
    from flexaidds_amd import hip, synth
    lig = system.ligands[0]
    log(f"GA dock: {args.pop} x {args.ga_generations} generations")
    ctx.reset_totals()
    t0 = time.perf_counter()
    ga = hip.GA(ctx, synth.ga_limits(lig), population=args.pop, offspring=args.pop, seed=2024, temperature=300.0)
    ms = ga.run(args.ga_generations, graph=True)
    th, n = ga.thermo()
    _, ic, res = ga.population()
    wall = time.perf_counter() - t0
    ga.close()
    poses = args.pop * args.ga_generations
    w, f = ctx.launch_shape()
    out = {"population": args.pop, "generations": args.ga_generations, "poses": int(n),
           "launch_shape": {"waves": w, "flags": f},
           "generations_ms": round(ms, 2), "ms_per_generation": round(ms / args.ga_generations, 4),
           "poses_per_s": round(poses / (ms / 1e3), 1), "dock_wall_s": round(wall, 3),
           "best_cf": round(float(res["cf"][0]), 6), "free_energy_300K": round(th["free_energy"], 6),
           "unscored": _unscored(ctx)}
    # binding modes of the final population (survivors in energy order)
    from flexaidds_amd import modes
    pop = modes.BindingPopulation(ctx, 0, ic[:args.pop, :lig.n_genes], res["app"][:args.pop], temperature=300.0,
                                  rmsd=2.0, max_results=10)
    out["binding_modes"] = {"modes": pop.size(), "clustered": pop.clusters["n_clustered"],
                            "device_ms": round(pop.clusters["device_ms"], 4),
                            "best_mode_free_energy_300K": round(pop.modes[0].free_energy(), 6) if pop.modes else None,
                            "best_mode_size": pop.modes[0].size() if pop.modes else 0}
    # the dominant energy basin of those poses (FastOPTICS super-cluster)
    if pop.modes:
        e_all = np.concatenate([m.energies for m in pop.modes])
        if 4 < len(e_all) <= hip.SUPER_CLUSTER_MAX:
            sc = ctx.super_cluster(e_all, max(4, len(e_all) // 20))
            sce = pop.super_cluster_ensemble()
            out["binding_modes"]["super_cluster"] = {"poses": int(len(e_all)), "members": int(len(sc["members"])),
                                                     "device_ms": round(sc["device_ms"], 4),
                                                     "free_energy_300K": round(sce["free_energy"], 6)}
    # the same dock with the reference's default duplicate rejection
    if args.dedup_generations > 0:
        ga = hip.GA(ctx, synth.ga_limits(lig), population=args.pop, offspring=args.pop, seed=2024, temperature=300.0,
                    dedup_rounds=8)
        ga.snapshot_enable(args.dedup_generations)
        dms = ga.run(args.dedup_generations, graph=True)
        st = ga.dedup_stats()[0]
        _, _, dres = ga.population()
        # the end of the reference's run: the snapshot sorted and thinned,
        # its ensemble, its binding modes
        t0 = time.perf_counter()
        snap = ga.snapshot()
        snap_ms = (time.perf_counter() - t0) * 1e3
        sth = ga.snapshot_thermo()
        ga.close()
        spop = modes.BindingPopulation(ctx, 0, snap["ic"], snap["app"], temperature=300.0, rmsd=2.0, max_results=10)
        out["dedup"] = {"generations": args.dedup_generations, "rounds": 8,
                        "ms_per_generation": round(dms / args.dedup_generations, 4),
                        "duplicates_drawn": int(st[0]), "duplicates_kept_at_cap": int(st[1]),
                        "best_cf": round(float(dres["cf"][0]), 6),
                        "snapshot": {"records": snap["n_raw"], "kept": int(len(snap["evalue"])),
                                     "sort_and_thin_ms": round(snap_ms, 3),
                                     "free_energy_300K": round(sth["free_energy"], 6),
                                     "modes": spop.size(), "clustered": spop.clusters["n_clustered"],
                                     "cluster_device_ms": round(spop.clusters["device_ms"], 4)},
                        "note": "ms_per_generation includes recording the population into the snapshot"}
    return out


def modes_leg(args):
    """Clustering throughput: --modes-poses jittered poses of the config-1
    ligand (a GA history's size) in energy order, T = 300 K, 10 modes."""
    from flexaidds_amd import hip, synth
    log(f"binding modes: {args.modes_poses} poses")
    system = synth.make_config1(pop=8)
    lig = system.ligands[0]
    rng = np.random.default_rng(77)
    seeds = synth.random_genes(rng, lig, max(1, args.modes_poses // 400))
    genes = np.repeat(seeds, 400, axis=0)[:args.modes_poses]
    ang = (np.asarray(lig.gene_typ) != -1) & (np.asarray(lig.gene_typ) != 4)
    genes[:, ang] += rng.normal(0.0, 5.0, size=(len(genes), int(ang.sum())))
    genes = np.clip(genes, lig.gene_min, lig.gene_max)
    e = np.sort(rng.normal(-20.0, 8.0, len(genes)))
    with hip.Context() as ctx:
        ctx.upload_system(system)
        ctx.cluster_cf(0, genes[:1000], e[:1000], temperature=300.0)          # warm
        r = [ctx.cluster_cf(0, genes, e, temperature=300.0, rmsd=2.0, max_results=10) for _ in range(3)]
        # the FastOPTICS super-cluster at its largest size
        n_sc = hip.SUPER_CLUSTER_MAX
        e_sc = rng.normal(-20.0, 8.0, n_sc)
        ctx.super_cluster(e_sc[:100], 4)                                      # warm
        sc = [ctx.super_cluster(e_sc, max(4, n_sc // 20)) for _ in range(3)]
    ms = float(np.median([x["device_ms"] for x in r]))
    passes = len(r[0]["modes"])
    # coordinate bytes one head pass reads (component-major float32)
    pass_bytes = len(genes) * 3 * lig.n_atoms * 4
    return {"poses": len(genes), "ligand_atoms": lig.n_atoms, "modes": passes, "clustered": r[0]["n_clustered"],
            "device_ms": round(ms, 4), "poses_per_s": round(len(genes) / (ms / 1e3), 1),
            "coordinate_bytes_per_pass": pass_bytes,
            "note": "device_ms covers the coordinate build and every head pass",
            "super_cluster": {"points": n_sc, "min_pts": max(4, n_sc // 20), "members": int(len(sc[0]["members"])),
                              "device_ms": round(float(np.median([x["device_ms"] for x in sc])), 4)}}


def _tune_generation(ctx, system, pop, seed=5):
    """Autotune the launch shape on one generation-sized launch of every
    ligand of `system` (random genes, one launch for all of them), as the
    batched GA scores its children."""
    import torch
    from flexaidds_amd import abi, autotune, synth
    rng = np.random.default_rng(seed)
    gmax = max(l.n_genes for l in system.ligands)
    rows, lig = [], []
    for li, l in enumerate(system.ligands):
        g = np.zeros((pop, gmax))
        g[:, :l.n_genes] = synth.random_genes(rng, l, pop)
        rows.append(g)
        lig.append(np.full(pop, li, np.int32))
    d_g = torch.from_numpy(np.concatenate(rows)).cuda()
    d_l = torch.from_numpy(np.concatenate(lig)).cuda()
    n = d_l.numel()
    d_o = torch.zeros(n * abi.CF_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
    sh = torch.cuda.current_stream().cuda_stream
    return autotune.tune(ctx, lambda: ctx.eval_device(n, d_g.data_ptr(), gmax, d_l.data_ptr(), d_o.data_ptr(), sh),
                         reps=2)


def batched_leg(args):
    """The batched screen: every ligand a job of one batched GA."""
    import torch  # noqa: F401  (same HIP runtime as the library)
    from flexaidds_amd import hip, synth
    log(f"batched screen: {args.batched_ligands} ligands x {args.batched_pop}")
    bs = synth.make_config2(n_ligands=args.batched_ligands, pop=8)
    bctx = hip.Context()
    bctx.upload_system(bs)
    shape = _tune_generation(bctx, bs, args.batched_pop)   # kept on the context: the GA's graph uses it
    log(f"batched launch shape: {shape['waves']} waves, flags {shape['flags']}")
    bctx.reset_totals()
    J = len(bs.ligands)
    ga = hip.GA(bctx, [synth.ga_limits(l) for l in bs.ligands], ligands=list(range(J)), population=args.batched_pop,
                offspring=args.batched_pop, seed=77, temperature=300.0)
    ga.run(1, graph=True)                                     # capture + warm
    ms = ga.run(args.batched_generations, graph=True)
    ths, n = ga.thermo_jobs()
    conc = np.exp(np.random.default_rng(3).uniform(np.log(1e-9), np.log(1e-3), J))
    gpf = ga.grand(conc)
    ga.close()
    uns = _unscored(bctx)
    bctx.close()
    poses = J * args.batched_pop * args.batched_generations
    return {"ligands": J, "population": args.batched_pop, "generations": args.batched_generations,
            "poses_per_generation": J * args.batched_pop, "ms_per_generation": round(ms / args.batched_generations, 3),
            "poses_per_s": round(poses / (ms / 1e3), 1), "ln_xi_300K": round(gpf["log_xi"], 6),
            "p_empty": gpf["p_empty"], "unscored": uns, "launch_shape": shape}


def grand_leg(args):
    """Config-4 shape: a blind dock of competing ligands on a 20,000-atom
    globule as one batched GA, then the grand partition function over the
    ligands' ensembles at their concentrations.  --grand-generations are
    timed (a few by default, the full dock at 300); the 300-generation time
    is also projected from them."""
    import torch  # noqa: F401  (same HIP runtime as the library)
    from flexaidds_amd import hip, synth
    log(f"grand-canonical blind dock: {args.grand_ligands} ligands x {args.grand_pop}")
    gs, conc = synth.make_config4(n_ligands=args.grand_ligands, pop=8)
    gctx = hip.Context()
    gctx.upload_system(gs)
    gctx.reset_totals()
    J = len(gs.ligands)
    ga = hip.GA(gctx, [synth.ga_limits(l) for l in gs.ligands], ligands=list(range(J)), population=args.grand_pop,
                offspring=args.grand_pop, seed=404, temperature=300.0)
    ga.run(1, graph=True)                                     # capture + warm
    if args.grand_warm > 0:                                   # let the population move off clashing poses
        ga.run(args.grand_warm, graph=True)
    ms, done = 0.0, 0
    while done < args.grand_generations:                      # chunks, with a progress line each
        k = min(20, args.grand_generations - done)
        ms += ga.run(k, graph=True)
        done += k
        if args.grand_generations > 20:
            log(f"grand: {done}/{args.grand_generations} generations, {ms / done:.1f} ms each")
    gpf = ga.grand(conc)
    _, _, res = ga.population()
    kids = res[J * args.grand_pop:]
    ga.close()
    uns = _unscored(gctx)
    gctx.close()
    poses = J * args.grand_pop * args.grand_generations
    per_gen = ms / args.grand_generations
    return {"ligands": J, "receptor_atoms": gs.receptor.n_atoms, "population": args.grand_pop,
            "generations": args.grand_generations, "poses_per_generation": J * args.grand_pop,
            "ms_per_generation": round(per_gen, 3), "poses_per_s": round(poses / (ms / 1e3), 1),
            "generations_s": round(ms / 1e3, 2), "warm_generations": 1 + args.grand_warm,
            "clash_fraction_last_generation": round(float((kids["status"] == 2).mean()), 4),
            "ln_xi_300K": round(gpf["log_xi"], 6), "p_empty": gpf["p_empty"],
            "max_occupancy": round(float(np.max(gpf["p_bound"])), 6), "unscored": uns,
            "note": "a few generations after warm-up; the per-generation cost rises while selection moves "
                    "the population off early-exit clash poses, so this rate is not a projection of the "
                    "full 300-generation dock (measured separately: --grand-generations 300)"}


def real_leg(args):
    """The plain launch on real inputs: the reference's RNA target as its own
    reader chain reads it (tests/golden/receptor_guanine.npz, 1,097 atoms),
    its automatic binding-site grid (gap-sphere cleft, 9,043 points at
    0.375 A) as the translation genes, a 45-atom / 10-torsion ligand,
    population --pop per launch; then a device-GA dock of --real-generations."""
    import torch
    from flexaidds_amd import abi, cleft, hip, synth
    g = np.load(os.path.join(ROOT, "tests", "golden", "receptor_guanine.npz"))
    ori = g["center"][:3].astype(np.float32)
    cg = cleft.cleft_grid(g["xyz"], g["radius"], ori)
    anchor = cg.points.astype(np.float64).mean(axis=0)
    rng = np.random.default_rng(3003)
    rec = synth.Receptor(xyz=g["xyz"].copy(), radius=g["radius"].copy(), type=g["type"].copy(), center=anchor)
    lig = synth.make_ligand(rng, 45, 10, ori, anchor, cg.points)
    lig.grid_ic[1:] = cg.ic
    synth.finalize_frame(lig, rec)
    system = synth.System(receptor=rec, ligands=[lig], params=synth.Params(ntypes=40, emat=synth.make_emat(rng)),
                          genes=[])
    log(f"real inputs: {rec.n_atoms}-atom RNA target, {len(cg.points)} cleft points")
    G = lig.n_genes
    n = args.real_steps + 2
    pops = np.stack([synth.random_genes(rng, lig, args.pop) for _ in range(n)])
    ctx = hip.Context(max_pop=args.pop)
    ctx.upload_system(system)
    ctx.reset_totals()
    d_genes = torch.from_numpy(pops).cuda()
    d_out = torch.zeros((n, args.pop * abi.CF_DTYPE.itemsize), dtype=torch.uint8, device="cuda")
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream
    for k in range(2):
        ctx.eval_device(args.pop, d_genes[k].data_ptr(), G, 0, d_out[k].data_ptr(), sh)
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0.record(stream)
    for k in range(2, n):
        ctx.eval_device(args.pop, d_genes[k].data_ptr(), G, 0, d_out[k].data_ptr(), sh)
    t1.record(stream)
    torch.cuda.synchronize()
    ms = t0.elapsed_time(t1)
    res = np.frombuffer(d_out[2:].cpu().numpy().tobytes(), abi.CF_DTYPE)
    ga = hip.GA(ctx, synth.ga_limits(lig), population=args.pop, offspring=args.pop, seed=77, temperature=300.0)
    ga_ms = ga.run(args.real_generations, graph=True)
    _, _, best = ga.population()
    ga.close()
    caps = ctx.totals()["capacity"]
    ctx.close()
    return {"receptor_atoms": rec.n_atoms, "cleft_points": int(len(cg.points)), "cleft_spheres": int(len(cg.spheres)),
            "ligand_atoms": lig.n_atoms, "population": args.pop, "launches": args.real_steps,
            "ms_per_launch": round(ms / args.real_steps, 4),
            "poses_per_s": round(args.real_steps * args.pop / (ms / 1e3), 1),
            "scored_fraction": round(float((res["status"] == 0).mean()), 4), "capacity_overflows": caps,
            "ga_generations": args.real_generations,
            "ga_ms_per_generation": round(ga_ms / max(args.real_generations, 1), 4),
            "ga_best_cf": round(float(best["cf"][0]), 6)}


def served_leg(args):
    """The served stream, and the plain path on the same kind of requests.

    Scenarios on the same resident receptors: the specified stream (Poisson
    at --served-rate, every resident job advanced each step), the same stream
    with at most --served-cap running jobs per receptor, and a stream at
    --served-low-rate, below what the device scores."""
    from flexaidds_amd import hip, serve, synth
    log(f"served stream: {args.served_requests} requests at {args.served_rate}/s")
    systems = synth.make_config3(n_receptors=8, ligands_per_receptor=args.served_library)
    srv = serve.Server(systems, population=args.served_pop, max_slots=64, hbm_budget_per_pool=2 << 30)
    # warm every pool's kernels with one short job, untimed
    srv.run([{"t": 0.0, "receptor": r, "ligand": 0, "generations": 1, "seed": 1} for r in range(len(systems))],
            realtime=False)

    def scenario(rate, n, cap, seed):
        for p in srv.pools:
            p.set_max_running(cap)
        reqs = synth.poisson_requests(np.random.default_rng(seed), systems, rate, n,
                                      generations=args.served_generations)
        got = srv.run(reqs, realtime=True).summary()
        got.update({"rate_hz": rate, "max_running_per_receptor": cap,
                    "offered_poses_per_s": round(rate * args.served_pop * (args.served_generations + 1), 1)})
        return got, reqs

    main_s, reqs = scenario(args.served_rate, args.served_requests, 0, 11)
    scen = [main_s]
    if args.served_cap > 0:
        scen.append(scenario(args.served_rate, args.served_requests, args.served_cap, 11)[0])
    if args.served_low_rate > 0:
        scen.append(scenario(args.served_low_rate, args.served_low_requests, 0, 12)[0])
    # plain path: one request at a time, one synchronous launch per generation
    plain_ms, plain_poses = 0.0, 0
    for r in reqs[:2]:
        c = srv.ctx[r["receptor"]]
        ga = hip.GA(c, srv.limits[r["receptor"]][r["ligand"]], ligand=r["ligand"], population=args.served_pop,
                    offspring=args.served_pop, seed=r["seed"])
        for _ in range(10):
            plain_ms += ga.run(1, graph=False)
            plain_poses += args.served_pop
        ga.close()
    # each receptor's grand partition function over every session it served
    # (warm-up jobs included; a re-docked ligand keeps its last ln Z)
    targets = []
    for q, ts in enumerate(srv.targets):
        g = ts.grand_partition()
        top = ts.rank_ligands()[0]
        targets.append({"receptor": q, "sessions": ts.completed_sessions(), "ligands": g.num_ligands(),
                        "ln_xi_300K": round(g.log_Xi(), 6), "p_empty": g.empty_probability(),
                        "top": top.name, "top_dG": round(top.dG, 4), "top_p_bound": round(top.p_bound, 6)})
    srv.close()
    plain = plain_poses / (plain_ms / 1e3)
    out = dict(main_s)
    out.update({"receptors": len(systems), "population": args.served_pop, "generations": args.served_generations,
                "plain_poses_per_s": round(plain, 1), "vs_plain": round(main_s["poses_per_s"] / plain, 3),
                "scenarios": scen, "targets": targets,
                "note": "at the specified rate the stream offers more poses/s than one GPU scores, so every "
                        "request waits for the backlog and latency grows with the stream's length (no steady "
                        "state exists); the low-rate scenario is below capacity"})
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--pop", type=int, default=2000)
    ap.add_argument("--ga-generations", type=int, default=500, help="plain GA dock length (0: skip)")
    ap.add_argument("--dedup-generations", type=int, default=100,
                    help="the GA dock again with duplicate rejection (0: skip)")
    ap.add_argument("--modes-poses", type=int, default=200000, help="binding-mode clustering poses (0: skip)")
    ap.add_argument("--config0", type=int, default=1, help="time the config-0 case (1,000 poses, Z/F at 300 K; 0: skip)")
    ap.add_argument("--batched-ligands", type=int, default=256, help="batched screen ligands (0: skip)")
    ap.add_argument("--batched-pop", type=int, default=512)
    ap.add_argument("--batched-generations", type=int, default=200)
    ap.add_argument("--served-requests", type=int, default=60, help="served stream requests (0: skip)")
    ap.add_argument("--served-rate", type=float, default=50.0)
    ap.add_argument("--served-library", type=int, default=6, help="resident ligands per receptor")
    ap.add_argument("--served-generations", type=int, default=100)
    ap.add_argument("--served-pop", type=int, default=1024)
    ap.add_argument("--served-cap", type=int, default=2, help="running jobs per receptor in the capped scenario")
    ap.add_argument("--served-low-rate", type=float, default=8.0, help="arrival rate below capacity (0: skip)")
    ap.add_argument("--served-low-requests", type=int, default=96)
    ap.add_argument("--grand-ligands", type=int, default=64, help="config-4 blind dock ligands (0: skip)")
    ap.add_argument("--grand-pop", type=int, default=32768)
    ap.add_argument("--grand-generations", type=int, default=3)
    ap.add_argument("--grand-warm", type=int, default=2, help="untimed generations before the timed ones")
    ap.add_argument("--real-steps", type=int, default=20, help="launches on the RNA target's cleft (0: skip)")
    ap.add_argument("--real-generations", type=int, default=200)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default launch shape (no autotuning launches; for profiler runs)")
    ap.add_argument("--sweep", action="store_true",
                    help="only time one launch per population size (1 .. 16,384 poses of the plain workload) and exit")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device 5 times (calibrates rocprofv3 byte counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "poses are independent and would shard by population across GPUs, which is not run here.",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from flexaidds_amd import abi, hip, synth

    torch.cuda.set_device(0)
    if args.copy_probe > 0:
        n = args.copy_probe * (1 << 20)
        src = torch.ones(n, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": n, "copies": 5}), flush=True)
        return
    if args.sweep:
        population_sweep()
        return
    log("building config-1 workload (five flexible side chains)")
    system = synth.make_config1(pop=args.pop)
    lig = system.ligands[0]
    G = lig.n_genes
    rng = np.random.default_rng(2024)
    n_pop = args.steps + args.warmup
    pops = np.stack([synth.random_genes(rng, lig, args.pop) for _ in range(n_pop)])
    ctx = hip.Context(max_pop=args.pop)
    ctx.upload_system(system)
    ctx.reset_totals()
    d_genes = torch.from_numpy(pops).cuda()
    out_bytes = args.pop * abi.CF_DTYPE.itemsize
    d_out = torch.zeros((n_pop, out_bytes), dtype=torch.uint8, device="cuda")
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream

    def launch(k):
        ctx.eval_device(args.pop, d_genes[k].data_ptr(), G, 0, d_out[k].data_ptr(), sh)

    log("warmup")
    for k in range(args.warmup):
        launch(k)
    torch.cuda.synchronize()
    from flexaidds_amd import autotune
    if args.no_tune:
        w0, f0 = ctx.launch_shape()
        plain_shape = {"waves": w0, "flags": f0, "ms": None, "table": []}
    else:
        plain_shape = autotune.tune(ctx, lambda: launch(0), reps=3)   # left set for the timed launches
    log(f"launch shape: {plain_shape['waves']} waves, flags {plain_shape['flags']} ({plain_shape['ms']} ms)")
    ctx.reset_totals()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    t_begin = torch.cuda.Event(enable_timing=True)
    t_end = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t_begin.record(stream)
    for i in range(args.steps):
        starts[i].record(stream)
        launch(args.warmup + i)
        ends[i].record(stream)
    t_end.record(stream)
    torch.cuda.synchronize()
    total_ms = t_begin.elapsed_time(t_end)
    launch_ms = np.array([s.elapsed_time(e) for s, e in zip(starts, ends)])
    log(f"timed {args.steps} steps: {total_ms:.2f} ms")

    # algorithmic FLOPs of the timed launches (deterministic: replay untimed)
    flop_total = 0
    status = np.zeros(5, np.int64)
    for i in range(args.steps):
        launch(args.warmup + i)
        flop_total += hip.flops(ctx.opcount())
        res = np.frombuffer(d_out[args.warmup + i].cpu().numpy().tobytes(), abi.CF_DTYPE)
        status += np.bincount(res["status"], minlength=5)
    plain_unscored = _unscored(ctx)
    poses = args.steps * args.pop
    value = poses / (total_ms / 1e3)
    scored = int(status[0])
    achieved = flop_total / args.steps / (launch_ms.mean() / 1e3) / 1e12

    # thermodynamics of the last population (FP64 GPU reduction, in place)
    cf_off = abi.CF_DTYPE.fields["cf"][1]
    th = ctx.thermo_device(d_out[args.warmup + args.steps - 1].data_ptr() + cf_off, abi.CF_DTYPE.itemsize // 8,
                           args.pop, 300.0, stream=sh)

    c0 = config0_leg(args) if args.config0 else None
    ga_dock = ga_dock_leg(ctx, system, args) if args.ga_generations > 0 else None
    batched = batched_leg(args) if args.batched_ligands > 0 else None
    served = served_leg(args) if args.served_requests > 0 else None
    grand = grand_leg(args) if args.grand_ligands > 0 else None
    real = real_leg(args) if args.real_steps > 0 else None
    modes_l = modes_leg(args) if args.modes_poses > 0 else None

    ctx.close()
    log("CPU baseline")
    cpu = None if args.no_cpu else cpu_baseline(system, pops.reshape(-1, G))
    line = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "poses/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(total_ms / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "plain GA-generation evaluation, config 1: 4,000-atom receptor shell whose cavity "
                               "is lined by 5 flexible side chains (3 chi angles, 10 conformations each), "
                               "45-atom / 10-torsion ligand, 14 ligand genes + 5 rotamer genes, population "
                               "2,000 per launch",
                   "receptor_atoms": system.receptor.n_atoms, "flexible_side_chains": system.receptor.n_flex,
                   "ligand_atoms": lig.n_atoms, "genes": G, "population": args.pop},
        "scored_poses_per_s": round(value * scored / max(1, int(status.sum())), 1),
        "clash_fraction": round(float(status[2]) / max(1, int(status.sum())), 4),
        "roofline": {"bound": "mfma", "achieved": round(achieved, 5), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / FP64_PEAK_TFLOPS, 7), "traffic": measured_traffic(args.pop),
                     "peak_kind": "FP64 vector (VALU) peak; the kernel issues FP64 VALU instructions and no MFMA "
                                  "(\"mfma\" names the compute-bound side of the hbm/mfma choice)",
                     "flop_per_launch": int(flop_total / args.steps),
                     "launch_ms_mean": round(float(launch_ms.mean()), 4)},
        "cpu_baseline": cpu,
        "latency_ms": {"generation_p50": round(float(np.percentile(launch_ms, 50)), 4),
                       "generation_p99": round(float(np.percentile(launch_ms, 99)), 4)},
        "config0": c0,
        "ga_dock": ga_dock,
        "batched": batched,
        "served": served,
        "grand": grand,
        "real_inputs": real,
        "binding_modes": modes_l,
        "status_counts": {"ok": int(status[0]), "hull_fail": int(status[1]), "clash": int(status[2]),
                          "capacity": int(status[3]), "grid_range": int(status[4])},
        "unscored": plain_unscored,
        "launch_shape": plain_shape,
        "free_energy_300K_last_population": round(th["free_energy"], 6),
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
