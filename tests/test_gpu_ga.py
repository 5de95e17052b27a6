"""Device GA against the GA oracle (GPU tests).

The operators use counter-based draws, so one generation on the GPU must
reproduce the oracle's children bit for bit from the same parents; the
survivors must be the oracle's stable ranking of parents + children by the
GPU energies; the energies themselves match the contact-function oracle to
its tolerance.  A generation captured as a HIP graph and replayed must give
exactly the eager generation's population after every replay."""
import numpy as np
import pytest

import oracle_lib as ol
from flexaidds_amd import abi, synth

pytestmark = pytest.mark.gpu

POP, NEW, SEED = 192, 160, 77


def tol(cf):
    return np.maximum(1e-4 * np.abs(cf), 1e-3)


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
def limits(system):
    return synth.ga_limits(system.ligands[0])


@pytest.fixture(scope="module")
def orc(oracle):
    return ol.ga_bind(oracle)


def _cf_oracle(oracle, system, ic):
    desc = abi.Descriptors(system.receptor, system.ligands, system.params)
    with ol.gpu_mirror(oracle):
        return ol.oracle_eval(oracle, desc, 0, ic)


def test_initial_population(hip, ctx, system, limits, orc, oracle):
    ga = hip.GA(ctx, limits, population=POP, offspring=NEW, seed=SEED)
    genes, ic, res = ga.population()
    og, oic = ol.ga_init(orc, SEED, POP, limits)
    # same chromosomes, ranked by energy
    key = {tuple(r): k for k, r in enumerate(og)}
    idx = np.array([key[tuple(r)] for r in genes[:POP]])
    assert sorted(idx) == list(range(POP))
    assert np.array_equal(ic[:POP], oic[idx])
    ref = _cf_oracle(oracle, system, oic)
    assert np.array_equal(res["status"][:POP], ref["status"][idx])
    assert np.all(np.abs(res["cf"][:POP] - ref["cf"][idx]) <= tol(ref["cf"][idx]))
    assert np.all(np.diff(res["cf"][:POP]) >= 0)
    assert ga.generation() == 0
    ga.close()


@pytest.mark.parametrize("pop,new", [(POP, NEW), (5000, 5000)])
def test_generation_matches_oracle(hip, ctx, system, limits, orc, oracle, pop, new):
    # (5000, 5000): the roulette narrows 5,000 rows in three rounds and the
    # ranking compares 10,000 candidates in 40 chunks
    ga = hip.GA(ctx, limits, population=pop, offspring=new, seed=SEED)
    g0, ic0, r0 = ga.population()
    ga.run(1, graph=False)
    g1, ic1, r1 = ga.population()
    assert ga.generation() == 1
    child, cic = ol.ga_breed(orc, SEED, 1, g0[:pop], new, limits)
    assert np.array_equal(g1[pop:], child)
    assert np.array_equal(ic1[pop:], cic)
    ref = _cf_oracle(oracle, system, cic)
    assert np.array_equal(r1["status"][pop:], ref["status"])
    assert np.all(np.abs(r1["cf"][pop:] - ref["cf"]) <= tol(ref["cf"]))
    # survivors: stable ranking of parents + children by the GPU energies
    energy = np.concatenate([r0["cf"][:pop], r1["cf"][pop:]])
    pool = np.concatenate([g0[:pop], child])
    order = ol.ga_survivors(orc, energy, pop)
    assert np.array_equal(g1[:pop], pool[order])
    assert r1["cf"][0] <= r0["cf"][0]
    ga.close()


def test_graph_replay_equals_eager(hip, ctx, limits):
    a = hip.GA(ctx, limits, population=POP, offspring=NEW, seed=SEED + 1)
    b = hip.GA(ctx, limits, population=POP, offspring=NEW, seed=SEED + 1)
    for gen in range(1, 5):
        a.run(1, graph=False)
        b.run(1, graph=True)
        ga_, ia, ra = a.population()
        gb, ib, rb = b.population()
        assert np.array_equal(ga_, gb), gen
        assert np.array_equal(ia, ib), gen
        assert np.array_equal(ra.tobytes(), rb.tobytes()), gen
    assert a.generation() == b.generation() == 4
    a.close()
    b.close()


def test_ensemble_over_all_evaluated_poses(hip, ctx, limits, oracle):
    ga = hip.GA(ctx, limits, population=POP, offspring=NEW, seed=SEED + 2, temperature=300.0)
    energies = [ga.population()[2]["cf"][:POP]]
    best = [energies[0][0]]
    for _ in range(3):
        ga.run(1, graph=True)
        _, _, r = ga.population()
        energies.append(r["cf"][POP:])
        best.append(r["cf"][0])
    th, n = ga.thermo()
    e = np.concatenate(energies)
    assert n == POP + 3 * NEW == len(e)
    ref = ol_thermo(oracle, e, 300.0)
    assert abs(th["free_energy"] - ref["free_energy"]) <= 1e-6 * abs(ref["free_energy"])
    assert abs(th["mean_energy"] - ref["mean_energy"]) <= 1e-6 * max(1.0, abs(ref["mean_energy"]))
    assert all(b2 <= b1 for b1, b2 in zip(best, best[1:]))
    ga.close()


def ol_thermo(oracle, energies, T):
    import ctypes as C
    e = np.ascontiguousarray(energies, np.float64)
    out = abi.Thermo()
    oracle.fa_oracle_thermo(e.ctypes.data, None, len(e), T, C.byref(out))
    return {f: getattr(out, f) for f, _ in abi.Thermo._fields_}


@pytest.mark.parametrize("model", ["PSHARE", "SMFREE"])
def test_fitness_models(hip, ctx, limits, orc, model):
    # niche radius from the limits, fitness of the survivors against the
    # oracle's (to rounding: the device's exp/pow/sqrt differ by ulps), PSHARE
    # survivors ordered by fitness, and children bred from the device fitness
    # exactly as the oracle breeds them
    code = {"PSHARE": 1, "SMFREE": 2}[model]
    T = 300.0
    ga = hip.GA(ctx, limits, population=POP, offspring=NEW, seed=SEED + 3, fitness=model, temperature=T,
                entropy_weight=0.5)
    f0, sig = ga.fitness()
    assert abs(sig[0] - ol.sig_share(orc, limits)) <= 1e-12 * sig[0]
    for gen in range(3):
        g, ic, r = ga.population()
        f, _ = ga.fitness()
        E = r["cf"][:POP]
        order = np.argsort(E, kind="stable")                     # survivors by increasing energy
        ref = ol.ga_fitness(orc, code, ic[order], E[order], T=T, sig=sig[0], alpha=1.0, ew=0.5)
        # rows of equal energy (duplicate children) may take their ranks in
        # either order: compare those as a multiset, the others row by row
        Es = E[order]
        tied = np.zeros(POP, bool)
        tied[1:] |= Es[1:] == Es[:-1]
        tied[:-1] |= Es[1:] == Es[:-1]
        assert (~tied).sum() > POP // 2
        assert np.allclose(f[order][~tied], ref[~tied], rtol=1e-12, atol=0), gen
        assert np.allclose(np.sort(f), np.sort(ref), rtol=1e-12, atol=0), gen
        if model == "PSHARE":
            assert np.all(np.diff(f) <= 0), gen                  # ordered by decreasing fitness
        else:
            assert np.all(np.diff(E) >= 0), gen                  # SMFREE keeps the energy order
        ga.run(1, graph=(gen % 2 == 1))
        g1, _, _ = ga.population()
        child, cic = ol.ga_breed_fit(orc, SEED + 3, gen + 1, g[:POP], f, NEW, limits)
        assert np.array_equal(g1[POP:], child), gen
    ga.close()


def test_steady_reproduction(hip, ctx, limits, orc):
    # STEADY (LIB/gaboom.cpp:1022-1027): the parents by energy, the worst
    # offspring-many replaced by the children whatever their energies, the
    # population then in energy order (ties by candidate index); children
    # bred exactly as under BOOM
    pop, new = 192, 64
    ga = hip.GA(ctx, limits, population=pop, offspring=new, seed=SEED + 5, reproduction="STEADY")
    for gen in (1, 2, 3):
        g0, _, r0 = ga.population()
        f, _ = ga.fitness()
        ga.run(1, graph=(gen == 2))
        g1, _, r1 = ga.population()
        child, _ = ol.ga_breed_fit(orc, SEED + 5, gen, g0[:pop], f, new, limits)
        assert np.array_equal(g1[pop:], child), gen
        pe = r0["cf"][:pop]
        keep = np.argsort(pe, kind="stable")[:pop - new]
        cand_e = np.concatenate([pe[keep], r1["cf"][pop:]])
        cand_i = np.concatenate([keep, pop + np.arange(new)])
        order = np.lexsort((cand_i, cand_e))
        rows = np.concatenate([g0[:pop], child])[cand_i[order]]
        assert np.array_equal(g1[:pop], rows), gen
        assert np.all(np.diff(r1["cf"][:pop]) >= 0), gen
    ga.close()
    with pytest.raises(hip.FAError):
        hip.GA(ctx, limits, population=64, offspring=65, reproduction="STEADY")
