"""Continuous-batching docking pool (GPU tests).

A job's draws depend only on its own seed and generation, so its result must
be bit-identical alone in a pool, in a mix of jobs that arrive and finish at
other steps, and as the single-job GA (fa_ga) with the same seed.  Admission
is bounded by the pool's slots and every submitted job finishes once."""
import numpy as np
import pytest

from flexaidds_amd import synth

pytestmark = pytest.mark.gpu

POP, NEW = 128, 128


@pytest.fixture(scope="module")
def hip():
    from flexaidds_amd import hip as h
    return h


@pytest.fixture(scope="module")
def system():
    return synth.make_config2(n_ligands=5, pop=8)


@pytest.fixture(scope="module")
def ctx(hip, system):
    c = hip.Context()
    c.upload_system(system)
    yield c
    c.close()


@pytest.fixture(scope="module")
def limits(system):
    return [synth.ga_limits(l) for l in system.ligands]


def _drain(pool, max_steps=1000):
    out = []
    for _ in range(max_steps):
        running, queued = pool.step()
        out += pool.collect()
        if running == 0 and queued == 0:
            return out
    raise AssertionError("pool did not drain")


def _same(a, b):
    assert a["best"] == b["best"]
    assert np.array_equal(a["best_genes"], b["best_genes"])
    assert np.array_equal(a["best_ic"], b["best_ic"])
    assert a["thermo"] == b["thermo"]
    assert a["n_poses"] == b["n_poses"]


@pytest.mark.parametrize("fitness,rep", [("LINEAR", "BOOM"), ("PSHARE", "BOOM"), ("LINEAR", "STEADY")])
def test_job_equals_single_ga(hip, ctx, limits, fitness, rep):
    gens, seed, lig = 6, 1234, 3
    new = NEW // 2 if rep == "STEADY" else NEW
    ga = hip.GA(ctx, limits[lig], ligand=lig, population=POP, offspring=new, seed=seed, fitness=fitness,
                reproduction=rep)
    ga.run(gens, graph=True)
    g, ic, res = ga.population()
    th, n = ga.thermo()
    ga.close()
    pool = hip.Pool(ctx, population=POP, offspring=new, max_slots=4, fitness=fitness, reproduction=rep)
    jid = pool.submit(lig, limits[lig], gens, seed)
    out = _drain(pool)
    pool.close()
    assert len(out) == 1 and out[0]["id"] == jid and out[0]["generations"] == gens
    r = out[0]
    best = int(np.argmin(res["cf"][:POP]))
    assert r["best"]["cf"] == res["cf"][best]
    G = len(limits[lig])
    assert np.array_equal(r["best_genes"], g[best, :G])
    assert np.array_equal(r["best_ic"], ic[best, :G])
    assert r["thermo"] == th
    assert r["n_poses"] == n == POP + gens * new


def test_result_same_alone_and_in_mix(hip, ctx, limits):
    target = dict(ligand=1, generations=8, seed=99)
    alone = hip.Pool(ctx, population=POP, offspring=NEW, max_slots=8)
    alone.submit(target["ligand"], limits[1], target["generations"], target["seed"])
    ref = _drain(alone)
    alone.close()
    assert len(ref) == 1
    # a mix: jobs of every ligand arriving at different steps, with other
    # lengths, the target arriving third, more jobs than slots
    mix = hip.Pool(ctx, population=POP, offspring=NEW, max_slots=3)
    arrivals = {0: [(0, 5, 1), (4, 3, 2)], 2: [(1, 8, 99)], 3: [(2, 2, 3), (3, 9, 4)], 5: [(0, 1, 5)]}
    ids, out = {}, []
    for step in range(400):
        for lig, gens, seed in arrivals.get(step, []):
            ids[mix.submit(lig, limits[lig], gens, seed)] = (lig, gens, seed)
        running, queued = mix.step()
        assert running <= mix.slots
        out += mix.collect()
        if step > 5 and running == 0 and queued == 0:
            break
    mix.close()
    assert sorted(r["id"] for r in out) == sorted(ids)
    got = [r for r in out if ids[r["id"]] == (1, 8, 99)]
    assert len(got) == 1
    _same(got[0], ref[0])
    for r in out:
        lig, gens, seed = ids[r["id"]]
        assert r["ligand"] == lig and r["generations"] == gens
        assert r["n_poses"] == POP + gens * NEW


def test_admission_bounded_by_slots(hip, ctx, limits):
    pool = hip.Pool(ctx, population=POP, offspring=NEW, max_slots=2)
    assert pool.slots == 2
    ids = [pool.submit(k % 5, limits[k % 5], 3, 10 + k) for k in range(5)]
    seen = []
    for _ in range(100):
        running, queued = pool.step()
        assert running <= 2 and running + queued <= 5
        seen += pool.collect()
        if running == 0 and queued == 0:
            break
    pool.close()
    assert sorted(r["id"] for r in seen) == ids


def test_running_cap(hip, ctx, limits):
    # at most max_running jobs resident; the rest wait in arrival order and
    # every result is still the one the job gets alone
    pool = hip.Pool(ctx, population=POP, offspring=NEW, max_slots=4, max_running=2)
    ids = [pool.submit(k % 5, limits[k % 5], 2 + k % 3, 40 + k) for k in range(5)]
    seen = []
    for _ in range(100):
        running, queued = pool.step()
        assert running <= 2
        seen += pool.collect()
        if running == 0 and queued == 0:
            break
    pool.close()
    assert sorted(r["id"] for r in seen) == ids
    for r in seen:
        k = ids.index(r["id"])
        alone = hip.Pool(ctx, population=POP, offspring=NEW, max_slots=1)
        alone.submit(k % 5, limits[k % 5], 2 + k % 3, 40 + k)
        ref = _drain(alone)
        alone.close()
        _same(r, ref[0])


def test_hbm_budget_sets_slots(hip, ctx):
    probe = hip.Pool(ctx, population=POP, offspring=NEW, max_slots=1)
    per = probe.slot_bytes
    probe.close()
    p = hip.Pool(ctx, population=POP, offspring=NEW, max_slots=100, hbm_budget=3 * per + per // 2)
    assert p.slots == 3
    p.close()
    with pytest.raises(hip.FAError):
        hip.Pool(ctx, population=POP, offspring=NEW, max_slots=4, hbm_budget=per // 2)


def test_zero_generation_job(hip, ctx, limits):
    pool = hip.Pool(ctx, population=POP, offspring=NEW, max_slots=2)
    jid = pool.submit(0, limits[0], 0, 5)
    pool.step()
    out = pool.collect()
    assert [r["id"] for r in out] == [jid]
    assert out[0]["n_poses"] == POP and out[0]["generations"] == 0
    assert pool.step() == (0, 0)
    pool.close()


def test_server_results_match_pools_alone(hip):
    # two resident receptors stepped back to back (their pools overlap on the
    # device), requests arriving together: each result equals the same job
    # alone in a fresh pool of its receptor
    from flexaidds_amd import serve
    systems = synth.make_config3(n_receptors=2, ligands_per_receptor=2, n_poses=2)
    reqs = [{"t": 0.0, "receptor": k % 2, "ligand": (k // 2) % 2, "generations": 3 + k % 3, "seed": 500 + k}
            for k in range(6)]
    srv = serve.Server(systems, population=64, max_slots=2, hbm_budget_per_pool=1 << 28)
    got = srv.run(reqs, realtime=False)
    assert sorted(got.results) == list(range(len(reqs)))
    for i, r in enumerate(reqs):
        alone = hip.Pool(srv.ctx[r["receptor"]], population=64, max_slots=1)
        alone.submit(r["ligand"], srv.limits[r["receptor"]][r["ligand"]], r["generations"], r["seed"])
        ref = _drain(alone)
        alone.close()
        _same(got.results[i], ref[0])
    s = got.summary()
    assert s["requests"] == len(reqs) and s["poses"] == sum(64 + r["generations"] * 64 for r in reqs)
    srv.close()


def test_busy_query(hip, ctx, limits):
    # the non-blocking query reports a step in flight until it ends, never
    # waits, and stepping after it went idle gives the usual result
    import time
    pool = hip.Pool(ctx, population=POP, offspring=NEW, max_slots=2)
    assert not pool.busy()
    jid = pool.submit(1, limits[1], 2, 77)
    out = []
    t0 = time.time()
    for _ in range(20):
        running, queued = pool.step()
        while pool.busy():
            assert time.time() - t0 < 60
            time.sleep(0.0005)
        out += pool.collect()
        if running == 0 and queued == 0:
            break
    assert not pool.busy()
    pool.close()
    alone = hip.Pool(ctx, population=POP, offspring=NEW, max_slots=1)
    alone.submit(1, limits[1], 2, 77)
    ref = _drain(alone)
    alone.close()
    assert [r["id"] for r in out] == [jid]
    _same(out[0], ref[0])
