"""Batched device GA (several ligands docked together) and the grand
partition function on the device (GPU tests).

Each job of a batched GA draws from its own streams, so job k's children
must equal the oracle's breed for job k from job k's parents, bit for bit;
its survivors are the oracle's stable ranking within the job; job 0 is the
single-ligand GA exactly (a ligand's result is the same alone and in the
mix).  The grand partition function over the jobs' ensembles matches the
oracle's ln Xi and occupancies."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as ol
from flexaidds_amd import abi, synth

pytestmark = pytest.mark.gpu

POP, NEW, SEED = 96, 80, 31
LIGS = [2, 0, 3, 1]


def tol(cf):
    return np.maximum(1e-4 * np.abs(cf), 1e-3)


@pytest.fixture(scope="module")
def hip():
    from flexaidds_amd import hip as h
    return h


@pytest.fixture(scope="module")
def system():
    return synth.make_config2(n_ligands=4, pop=8)


@pytest.fixture(scope="module")
def ctx(hip, system):
    c = hip.Context()
    c.upload_system(system)
    yield c
    c.close()


@pytest.fixture(scope="module")
def limits(system):
    return [synth.ga_limits(system.ligands[k]) for k in LIGS]


@pytest.fixture(scope="module")
def orc(oracle):
    return ol.ga_bind(oracle)


def _cf_oracle(oracle, system, lig, ic):
    desc = abi.Descriptors(system.receptor, system.ligands, system.params)
    with ol.gpu_mirror(oracle):
        return ol.oracle_eval(oracle, desc, lig, np.ascontiguousarray(ic))


def _batch(hip, ctx, limits, seed=SEED, **kw):
    return hip.GA(ctx, limits, ligands=LIGS, population=POP, offspring=NEW, seed=seed, **kw)


def test_shape_and_padding(hip, ctx, limits):
    ga = _batch(hip, ctx, limits)
    assert ga.jobs == len(LIGS)
    assert ga.genes == max(len(l) for l in limits)
    for gen in range(2):
        g, ic, res = ga.population()
        assert g.shape == (len(LIGS) * (POP + NEW), ga.genes)
        for k, lim in enumerate(limits):
            rows, crow = ga.job_rows(k)
            assert np.all(np.diff(res["cf"][rows]) >= 0)
            for r in (rows, crow) if gen else (rows,):
                assert np.all(g[r, len(lim):] == 0) and np.all(ic[r, len(lim):] == 0.0), (gen, k)
        ga.run(1, graph=False)
    ga.close()


def test_initial_population_per_job(hip, ctx, system, limits, orc, oracle):
    ga = _batch(hip, ctx, limits)
    genes, ic, res = ga.population()
    for k, (lig, lim) in enumerate(zip(LIGS, limits)):
        G = len(lim)
        rows, _ = ga.job_rows(k)
        og, oic = ol.ga_init(orc, SEED, POP, lim, job=k)
        key = {tuple(r): i for i, r in enumerate(og)}
        idx = np.array([key[tuple(r)] for r in genes[rows, :G]])
        assert sorted(idx) == list(range(POP)), k
        assert np.array_equal(ic[rows, :G], oic[idx]), k
        ref = _cf_oracle(oracle, system, lig, oic)
        assert np.array_equal(res["status"][rows], ref["status"][idx]), k
        assert np.all(np.abs(res["cf"][rows] - ref["cf"][idx]) <= tol(ref["cf"][idx])), k
    ga.close()


def test_generation_per_job_matches_oracle(hip, ctx, system, limits, orc, oracle):
    ga = _batch(hip, ctx, limits)
    g0, ic0, r0 = ga.population()
    ga.run(1, graph=False)
    g1, ic1, r1 = ga.population()
    for k, (lig, lim) in enumerate(zip(LIGS, limits)):
        G = len(lim)
        prow, crow = ga.job_rows(k)
        child, cic = ol.ga_breed(orc, SEED, 1, g0[prow, :G], NEW, lim, job=k)
        assert np.array_equal(g1[crow, :G], child), k
        assert np.array_equal(ic1[crow, :G], cic), k
        ref = _cf_oracle(oracle, system, lig, cic)
        assert np.array_equal(r1["status"][crow], ref["status"]), k
        assert np.all(np.abs(r1["cf"][crow] - ref["cf"]) <= tol(ref["cf"])), k
        energy = np.concatenate([r0["cf"][prow], r1["cf"][crow]])
        pool = np.concatenate([g0[prow], g1[crow]])
        order = ol.ga_survivors(orc, energy, POP)
        assert np.array_equal(g1[prow], pool[order]), k
    ga.close()


def test_job_zero_is_the_single_ligand_ga(hip, ctx, limits):
    # the same ligand alone and first in a mix: identical populations and
    # energies after every generation
    batch = _batch(hip, ctx, limits, seed=SEED + 5)
    alone = hip.GA(ctx, limits[0], ligand=LIGS[0], population=POP, offspring=NEW, seed=SEED + 5)
    G = len(limits[0])
    for gen in range(1, 4):
        batch.run(1, graph=True)
        alone.run(1, graph=True)
        gb, ib, rb = batch.population()
        ga_, ia, ra = alone.population()
        prow, crow = batch.job_rows(0)
        assert np.array_equal(gb[prow, :G], ga_[:POP]), gen
        assert np.array_equal(gb[crow, :G], ga_[POP:]), gen
        assert np.array_equal(rb[prow].tobytes(), ra[:POP].tobytes()), gen
        assert np.array_equal(rb[crow].tobytes(), ra[POP:].tobytes()), gen
    tb, _ = batch.thermo_jobs()
    ta, _ = alone.thermo()
    assert tb[0]["log_Z"] == ta["log_Z"] and tb[0]["mean_energy"] == ta["mean_energy"]
    batch.close()
    alone.close()


def test_batched_graph_replay_equals_eager(hip, ctx, limits):
    a = _batch(hip, ctx, limits, seed=SEED + 7)
    b = _batch(hip, ctx, limits, seed=SEED + 7)
    for gen in range(1, 4):
        a.run(1, graph=False)
        b.run(1, graph=True)
        pa, pb = a.population(), b.population()
        for x, y in zip(pa, pb):
            assert np.array_equal(np.asarray(x).tobytes(), np.asarray(y).tobytes()), gen
    a.close()
    b.close()


def test_job_ensembles_and_grand_partition(hip, ctx, limits, oracle):
    T = 300.0
    ga = _batch(hip, ctx, limits, seed=SEED + 9, temperature=T)
    energies = [[] for _ in LIGS]
    _, _, r = ga.population()
    for k in range(len(LIGS)):
        energies[k].append(r["cf"][ga.job_rows(k)[0]])
    for _ in range(2):
        ga.run(1, graph=True)
        _, _, r = ga.population()
        for k in range(len(LIGS)):
            energies[k].append(r["cf"][ga.job_rows(k)[1]])
    th, n = ga.thermo_jobs()
    assert n == len(LIGS) * (POP + 2 * NEW)
    log_z = []
    for k in range(len(LIGS)):
        e = np.ascontiguousarray(np.concatenate(energies[k]))
        ref = abi.Thermo()
        oracle.fa_oracle_thermo(e.ctypes.data, None, len(e), T, C.byref(ref))
        assert abs(th[k]["log_Z"] - ref.log_Z) <= 1e-9 * max(1.0, abs(ref.log_Z)), k
        assert abs(th[k]["free_energy"] - ref.free_energy) <= 1e-6 * abs(ref.free_energy), k
        log_z.append(th[k]["log_Z"])
    conc = np.array([1e-6, 1e-3, 5e-9, 2e-5])
    got = ga.grand(conc)
    lz = np.ascontiguousarray(log_z, np.float64)
    pb = np.zeros(len(LIGS))
    pe = C.c_double()
    lx = oracle.fa_oracle_log_xi(lz.ctypes.data, conc.ctypes.data, len(LIGS), pb.ctypes.data, C.byref(pe))
    assert abs(got["log_xi"] - lx) <= 1e-12 * max(1.0, abs(lx))
    assert np.allclose(got["p_bound"], pb, rtol=1e-12, atol=1e-300)
    assert abs(got["p_empty"] - pe.value) <= 1e-12 * max(pe.value, 1e-300)
    assert abs(got["p_empty"] + got["p_bound"].sum() - 1.0) <= 1e-12
    ga.close()


def test_grand_device_extremes(hip, ctx, oracle):
    # wide ln Z spreads (one ligand dominating, all negligible, a single
    # ligand, none): the max shift keeps every term finite
    import torch
    cases = [np.array([4000.0, 3990.0, -50.0]), np.array([-900.0, -800.0]), np.array([12.5]),
             np.linspace(-30.0, 30.0, 1000), np.zeros(0)]
    for lz in cases:
        n = len(lz)
        lc = np.log(np.linspace(1e-9, 1e-3, max(n, 1)))[:n]
        d_lz = torch.from_numpy(np.ascontiguousarray(lz)).cuda()
        d_lc = torch.from_numpy(np.ascontiguousarray(lc)).cuda()
        d_out = torch.zeros(n + 2, dtype=torch.float64, device="cuda")
        s = torch.cuda.current_stream()
        ctx.grand_device(d_lz.data_ptr() if n else d_out.data_ptr(), 1,
                         d_lc.data_ptr() if n else d_out.data_ptr(), n, d_out.data_ptr(), s.cuda_stream)
        s.synchronize()
        out = d_out.cpu().numpy()
        pb = np.zeros(max(n, 1))
        pe = C.c_double()
        conc = np.ascontiguousarray(np.exp(lc)) if n else np.ones(1)
        lx = oracle.fa_oracle_log_xi(np.ascontiguousarray(lz).ctypes.data, conc.ctypes.data, n, pb.ctypes.data,
                                     C.byref(pe))
        assert abs(out[0] - lx) <= 1e-12 * max(1.0, abs(lx)), n
        assert abs(out[1] - pe.value) <= 1e-12 * max(pe.value, 1e-300), n
        assert np.allclose(out[2:], pb[:n], rtol=1e-11, atol=1e-300), n
