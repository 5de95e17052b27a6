"""Device GA with duplicate rejection against the GA oracle (GPU tests).

The oracle draws candidates one at a time as the reference's reproduce()
and populate loops do (LIB/gaboom.cpp:971-1018, :1966-1978) and keeps those
whose signature is new to the job; the device draws batches of candidates
and keeps the new ones in draw order.  The initial population (as a
multiset: the device ranks it by energy), every generation's children row
for row, and the rejected / kept-duplicate counts must be identical, on a
gene space narrow enough that candidates repeat often, on one that runs out
of new chromosomes (the batch cap then keeps duplicates), through graph
replays across a growth of the signature table, and per job in a batched GA.
Independently of the signature, no kept row repeats an earlier kept row's
rounded genes (int32(ic + 0.5) per gene) while the cap is not hit.
"""
import numpy as np
import pytest

import oracle_lib as ol
from flexaidds_amd import synth

pytestmark = pytest.mark.gpu

SEED = 314


@pytest.fixture(scope="module")
def hip():
    from flexaidds_amd import hip as h
    return h


@pytest.fixture(scope="module")
def system():
    return synth.make_config0(n_poses=8)


@pytest.fixture(scope="module")
def ctx(hip, system):
    c = hip.Context()
    c.upload_system(system)
    yield c
    c.close()


@pytest.fixture(scope="module")
def orc(oracle):
    return ol.ga_bind(oracle)


def narrow(lig, wide_genes, bins=3):
    """The ligand's limits with the grid gene over 4 points, `wide_genes`
    angle genes over `bins` values and every other gene fixed."""
    out = []
    for g, (lo, hi, d, m) in enumerate(synth.ga_limits(lig)):
        if m:
            out.append((lo, lo + 3.0, 1.0, 1))
        elif g <= wide_genes:
            out.append((lo, lo + 2.0 * bins, 2.0, 0))
        else:
            out.append((lo, lo + 2.0, 2.0, 0))
    return out


def rounded(ic):
    return {tuple(int(v) for v in np.trunc(r + 0.5)) for r in ic}


def check_run(hip, ctx, orc, limits, pop, new, R, gens, graph=False, expect_fill=False):
    G = len(limits)
    ga = hip.GA(ctx, limits, population=pop, offspring=new, seed=SEED, dedup_rounds=R)
    g0, ic0, _ = ga.population()
    og, oic, ost = ol.ga_init_dedup(orc, SEED, pop, limits, R)
    assert sorted(map(tuple, g0[:pop, :G])) == sorted(map(tuple, og))
    stats = ost.copy()
    assert np.array_equal(ga.dedup_stats()[0], stats)
    seen = {ol.genes_sig(orc, r) for r in oic}
    kept_rounded = rounded(oic)
    distinct_checked = 0
    for gen in range(1, gens + 1):
        g, _, _ = ga.population()
        f, _ = ga.fitness()
        ga.run(1, graph=graph)
        g1, ic1, _ = ga.population()
        child, cic, st = ol.ga_breed_dedup(orc, SEED, gen, g[:pop, :G], f, new, limits, R, seen)
        assert np.array_equal(g1[pop:, :G], child), gen
        assert np.array_equal(ic1[pop:, :G], cic), gen
        stats += st
        assert np.array_equal(ga.dedup_stats()[0], stats), gen
        if int(st[1]) == 0:
            r = rounded(cic)
            assert len(r) == new and not (r & kept_rounded), gen
            distinct_checked += 1
        kept_rounded |= rounded(cic)
        seen |= {ol.genes_sig(orc, r) for r in cic}
    if expect_fill:
        assert int(stats[1]) > 0
    assert int(stats[0]) > 0
    ga.close()
    return stats, distinct_checked


def test_frequent_repeats(hip, ctx, system, orc):
    # 4 x 3^7 = 8,748 vectors: parents converge, so children repeat them often
    stats, checked = check_run(hip, ctx, orc, narrow(system.ligands[0], 7), 96, 96, 8, 4)
    assert checked >= 1


def test_exhausted_space_keeps_duplicates_at_cap(hip, ctx, system, orc):
    # 4 x 3^3 = 108 vectors: the second generation cannot find 48 new ones
    check_run(hip, ctx, orc, narrow(system.ligands[0], 3), 64, 48, 4, 3, expect_fill=True)


def test_graph_replay_across_table_growth(hip, ctx, system):
    # the signature table is sized for 16 generations at creation (4,096
    # entries here); 40 generations regrow it and re-capture the generation:
    # replays equal eager
    lim = narrow(system.ligands[0], 6)
    a = hip.GA(ctx, lim, population=64, offspring=64, seed=SEED + 1, dedup_rounds=6)
    b = hip.GA(ctx, lim, population=64, offspring=64, seed=SEED + 1, dedup_rounds=6)
    for _ in range(40):
        a.run(1, graph=False)
    b.run(20, graph=True)
    b.run(20, graph=True)
    pa, pb = a.population(), b.population()
    for x, y in zip(pa, pb):
        assert np.array_equal(x, y) if x.dtype.names is None else x.tobytes() == y.tobytes()
    assert np.array_equal(a.dedup_stats(), b.dedup_stats()) and int(a.dedup_stats()[0, 0]) > 0
    a.close()
    b.close()


def test_batched_jobs_follow_their_own_streams(hip, ctx, system, orc):
    # two jobs of one ligand with different gene spaces: job k draws from
    # streams (k << 32) | t and keeps its own signature set
    lims = [narrow(system.ligands[0], 5), narrow(system.ligands[0], 3)]
    pop, new, R = 64, 48, 4
    ga = hip.GA(ctx, lims, ligands=[0, 0], population=pop, offspring=new, seed=SEED + 2, dedup_rounds=R)
    g0, _, _ = ga.population()
    seen, stats = [], []
    for k in range(2):
        og, oic, ost = ol.ga_init_dedup(orc, SEED + 2, pop, lims[k], R, job=k)
        assert sorted(map(tuple, g0[k * pop:(k + 1) * pop, :len(lims[k])])) == sorted(map(tuple, og))
        seen.append({ol.genes_sig(orc, r) for r in oic})
        stats.append(ost.copy())
    for gen in (1, 2):
        g, _, _ = ga.population()
        f, _ = ga.fitness()
        ga.run(1, graph=False)
        g1, _, _ = ga.population()
        for k in range(2):
            G = len(lims[k])
            child, cic, st = ol.ga_breed_dedup(orc, SEED + 2, gen, g[k * pop:(k + 1) * pop, :G],
                                               f[k * pop:(k + 1) * pop], new, lims[k], R, seen[k], job=k)
            rows = slice(2 * pop + k * new, 2 * pop + (k + 1) * new)
            assert np.array_equal(g1[rows, :G], child), (gen, k)
            seen[k] |= {ol.genes_sig(orc, r) for r in cic}
            stats[k] += st
    assert np.array_equal(ga.dedup_stats(), np.stack(stats))
    ga.close()


def test_pool_jobs_equal_the_ga(hip, ctx, system):
    # a pool with duplicate rejection: each job equals the GA with the same
    # seed, also when it reuses the slot of a finished job (its set starts
    # empty); a job longer than the pool's sizing is refused
    lim = narrow(system.ligands[0], 5)
    pop, new, R, gens = 64, 48, 4, 5
    seeds = (SEED + 7, SEED + 8, SEED + 9)
    refs = []
    for sd in seeds:
        ga = hip.GA(ctx, lim, population=pop, offspring=new, seed=sd, dedup_rounds=R)
        ga.run(gens, graph=False)
        g, ic, res = ga.population()
        th, n = ga.thermo()
        assert int(ga.dedup_stats()[0, 0]) > 0
        ga.close()
        b = int(np.argmin(res["cf"][:pop]))
        refs.append((g[b], ic[b], res["cf"][b], th, n))
    pool = hip.Pool(ctx, population=pop, offspring=new, max_slots=1, dedup_rounds=R, dedup_max_generations=gens)
    ids = [pool.submit(0, lim, gens, sd) for sd in seeds]
    with pytest.raises(hip.FAError):
        pool.submit(0, lim, gens + 1, 1)
    out = {}
    for _ in range(100):
        running, queued = pool.step()
        for r in pool.collect():
            out[r["id"]] = r
        if running == 0 and queued == 0:
            break
    pool.close()
    assert sorted(out) == sorted(ids)
    G = len(lim)
    for jid, (gb, icb, cf, th, n) in zip(ids, refs):
        r = out[jid]
        assert r["best"]["cf"] == cf
        assert np.array_equal(r["best_genes"], gb[:G]) and np.array_equal(r["best_ic"], icb[:G])
        assert r["thermo"] == th and r["n_poses"] == n
