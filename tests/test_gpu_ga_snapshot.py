"""The GA's snapshot on the device against the oracle (GPU tests).

fa_ga_snapshot_enable records each job's population after every
generation (the reference's save_snapshot after reproduce,
LIB/gaboom.cpp:494); fa_ga_snapshot sorts a job's records by energy and
thins them with remove_dups (LIB/gaboom.cpp:517-528, :2728-2746).  The
records are rebuilt on the host from the population after each generation
(one generation per run), and the oracle's sort and thinning of them must
give the device's kept records bit for bit: genes, cf and app.  Then the
end-of-run ensemble over the kept energies (LIB/gaboom.cpp:535-540) and
the binding modes of the kept records (the reference clusters the
snapshot, LIB/top.cpp:1350-1353) run on them.
"""
import numpy as np
import pytest

import oracle_lib as ol
from flexaidds_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    from flexaidds_amd import hip as h
    return h


@pytest.fixture(scope="module")
def ctx(hip):
    c = hip.Context()
    c.upload_system(synth.make_config0(n_poses=8))
    yield c
    c.close()


def _records(ga, generations, graph):
    ic, cf, app = [], [], []
    pop = ga.params.population
    for _ in range(generations):
        ga.run(1, graph=graph)
        _, gic, res = ga.population()
        rows = slice(0, pop)
        ic.append(gic[rows, :ga.job_genes[0]])
        cf.append(res["cf"][rows])
        app.append(res["app"][rows])
    return np.concatenate(ic), np.concatenate(cf), np.concatenate(app)


@pytest.mark.parametrize("graph", [True, False])
def test_snapshot_equals_oracle(hip, ctx, oracle, graph):
    s = synth.make_config0(n_poses=8)
    lim = synth.ga_limits(s.ligands[0])
    ga = hip.GA(ctx, lim, population=300, offspring=300, seed=77, temperature=300.0)
    try:
        ga.snapshot_enable(30)
        print(f"snapshot graph={graph}", flush=True)
        ic, cf, app = _records(ga, 24, graph)
        snap = ga.snapshot()
        assert snap["n_raw"] == 24 * 300
        keep = ol.snapshot(oracle, ic, cf)
        assert len(keep) == len(snap["evalue"]) and 1 < len(keep) < snap["n_raw"]
        assert np.array_equal(snap["ic"].view(np.uint64), ic[keep].view(np.uint64))
        assert np.array_equal(snap["evalue"].view(np.uint64), cf[keep].view(np.uint64))
        assert np.array_equal(snap["app"].view(np.uint64), app[keep].view(np.uint64))
        th = ga.snapshot_thermo()
        ref = ctx.thermo(cf[keep], 300.0)
        assert th["log_Z"] == ref["log_Z"] and th["free_energy"] == ref["free_energy"]
    finally:
        ga.close()


def test_snapshot_limit_and_clusters(hip, ctx, oracle):
    # recording stops after max_generations; the kept records cluster into
    # binding modes in the reference's order (energies ascending)
    from flexaidds_amd import modes
    s = synth.make_config0(n_poses=8)
    ga = hip.GA(ctx, synth.ga_limits(s.ligands[0]), population=200, offspring=200, seed=5, temperature=300.0)
    assert ga.snapshot()["n_raw"] == 0
    ga.snapshot_enable(5)
    ga.run(8)
    snap = ga.snapshot()
    assert snap["n_raw"] == 5 * 200
    assert np.all(np.diff(snap["evalue"]) >= 0)
    pop = modes.BindingPopulation(ctx, 0, snap["ic"], snap["app"], temperature=300.0, rmsd=2.0, max_results=10)
    assert pop.size() >= 1 and pop.clusters["n_clustered"] <= len(snap["app"])
    ga.snapshot_enable(0)
    ga.run(1)
    assert ga.snapshot()["n_raw"] == 0
    with pytest.raises(hip.FAError):
        ga.snapshot(job=1)


def test_snapshot_super_cluster_prefilter(hip, ctx, oracle):
    # use_super_cluster: the end-of-run ensemble over the FastOPTICS
    # super-cluster of the kept energies when it leaves records out
    s = synth.make_config0(n_poses=8)
    ga = hip.GA(ctx, synth.ga_limits(s.ligands[0]), population=200, offspring=200, seed=9, temperature=300.0)
    try:
        ga.snapshot_enable(12)
        ga.run(12)
        e = ga.snapshot()["evalue"]
        assert 4 < len(e) <= 8192
        members = ol.super_cluster(oracle, e, max(4, len(e) // 20))
        want = e[members] if 0 < len(members) < len(e) else e
        got = ga.snapshot_thermo(super_cluster=True)
        ref = ctx.thermo(want, 300.0)
        assert got["log_Z"] == ref["log_Z"] and got["free_energy"] == ref["free_energy"]
        print(f"super-cluster kept {len(members)} of {len(e)}", flush=True)
    finally:
        ga.close()
