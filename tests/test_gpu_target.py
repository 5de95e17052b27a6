"""Grand partition function of a receptor on the device, and over served
docking sessions (GPU tests).

The device queries (ln Xi, p_empty, p_bound, ranking) take the reference's
own test values (tests/test_grand_partition.cpp: empty site, one ligand,
two competing ligands, equal ligands, ranking, removal, extreme ln Z,
concentrations), then a served stream: every finished job of a receptor is
registered under its request's name (re-docked names overwrite), and the
receptor's ln Xi, p_empty and p_bound must equal the oracle's
(fa_oracle_log_xi) over the registry the stream leaves -- each name's ln Z
from its last finished job -- to 1e-12 relative.
"""
import ctypes as C
import math

import numpy as np
import pytest

from flexaidds_amd import synth, target

pytestmark = pytest.mark.gpu
KT_300 = 0.001987206 * 300.0


@pytest.fixture(scope="module")
def ctx():
    from flexaidds_amd import hip
    c = hip.Context()
    yield c
    c.close()


def test_reference_values(ctx):
    g = target.GrandPartitionFunction(ctx)
    assert g.log_Xi() == 0.0 and g.empty_probability() == 1.0          # EmptyXi
    g.add_ligand("ligandA", 10.0)                                      # SingleLigand
    lx = math.log(1.0 + math.exp(10.0))
    assert abs(g.log_Xi() - lx) < 1e-10
    assert abs(g.binding_probability("ligandA") - math.exp(10.0) / (1.0 + math.exp(10.0))) < 1e-8
    assert abs(g.empty_probability() - 1.0 / (1.0 + math.exp(10.0))) < 1e-8
    h = target.GrandPartitionFunction(ctx)                             # TwoLigandsCompetitive
    h.add_ligand("A", 10.0)
    h.add_ligand("B", 8.0)
    xi = 1.0 + math.exp(10.0) + math.exp(8.0)
    assert abs(math.exp(h.log_Xi()) - xi) < xi * 1e-10
    assert abs(h.binding_probability("A") + h.binding_probability("B") + h.empty_probability() - 1.0) < 1e-10
    e = target.GrandPartitionFunction(ctx)                             # EqualLigands
    for k in "XYZ":
        e.add_ligand(k, 5.0)
    p = [e.binding_probability(k) for k in "XYZ"]
    assert max(p) - min(p) < 1e-12
    r = target.GrandPartitionFunction(ctx)                             # Ranking
    r.add_ligand("weak", 2.0)
    r.add_ligand("strong", 20.0)
    r.add_ligand("medium", 10.0)
    ranks = r.rank()
    assert [x.name for x in ranks] == ["strong", "medium", "weak"]
    assert all(abs(x.dG + KT_300 * x.log_Z) < 1e-10 for x in ranks)
    assert abs(sum(x.p_bound for x in ranks) + r.empty_probability() - 1.0) < 1e-10
    before = h.log_Xi()                                                # RemoveLigand
    h.remove_ligand("A")
    assert h.log_Xi() < before and abs(h.log_Xi() - math.log(1.0 + math.exp(8.0))) < 1e-12
    x = target.GrandPartitionFunction(ctx)                             # ExtremeLogZ
    x.add_ligand("huge", 500.0)
    x.add_ligand("tiny", -500.0)
    assert abs(x.log_Xi() - 500.0) < 1.0 and abs(x.binding_probability("huge") - 1.0) < 1e-6
    assert x.binding_probability("tiny") < 1e-6 and math.isfinite(x.empty_probability())
    c = target.GrandPartitionFunction(ctx)                             # ConcentrationAffectsProbability
    c.add_ligand("A", 10.0, 1.0)
    c.add_ligand("B", 10.0, 0.01)
    assert c.binding_probability("A") > c.binding_probability("B")
    assert abs(c.binding_probability("A") / c.binding_probability("B") - 100.0) < 1e-9


def test_cache_follows_changes(ctx):
    # every change re-evaluates ln Xi (a stale cache would keep the old value)
    g = target.GrandPartitionFunction(ctx)
    g.add_ligand("A", 4.0)
    a = g.log_Xi()
    g.overwrite_ligand("A", 6.0)
    b = g.log_Xi()
    g.merge_ligand("A", 6.0)
    c = g.log_Xi()
    g.add_or_overwrite("B", 1.0, 0.5)
    d = g.log_Xi()
    assert abs(a - math.log1p(math.exp(4.0))) < 1e-12 and abs(b - math.log1p(math.exp(6.0))) < 1e-12
    assert abs(c - math.log1p(2.0 * math.exp(6.0))) < 1e-12
    assert abs(d - math.log(1.0 + 2.0 * math.exp(6.0) + 0.5 * math.exp(1.0))) < 1e-12


def test_served_sessions_feed_xi(oracle):
    from flexaidds_amd import serve
    systems = synth.make_config3(n_receptors=2, ligands_per_receptor=3, n_poses=2)
    names = ["aspirin", "ibuprofen", "caffeine"]
    reqs = []
    for k in range(10):
        q, lig = k % 2, (k // 2) % 3
        reqs.append({"t": 0.0, "receptor": q, "ligand": lig, "generations": 2 + k % 4, "seed": 900 + k,
                     "name": names[lig]})
    srv = serve.Server(systems, population=64, max_slots=2, hbm_budget_per_pool=1 << 28)
    got = srv.run(reqs, realtime=False)
    try:
        assert sorted(got.finished) == list(range(len(reqs)))
        for q in range(2):
            reg = {}
            for i in got.finished:                                   # completion order: last one wins
                if reqs[i]["receptor"] == q:
                    reg[reqs[i]["name"]] = got.results[i]["thermo"]["log_Z"]
            ts = srv.targets[q]
            assert ts.completed_sessions() == sum(r["receptor"] == q for r in reqs)
            assert dict(ts.grand_partition().all_log_Z()) == reg
            lz = np.array(list(reg.values()))
            conc = np.ones(len(lz))
            pb = np.zeros(len(lz))
            pe = C.c_double()
            lx = oracle.fa_oracle_log_xi(lz.ctypes.data, conc.ctypes.data, len(lz), pb.ctypes.data, C.byref(pe))
            g = ts.grand_partition()
            assert abs(g.log_Xi() - lx) <= 1e-12 * max(1.0, abs(lx))
            assert abs(g.empty_probability() - pe.value) <= 1e-12 * max(pe.value, 1e-300)
            for k, name in enumerate(reg):
                assert abs(ts.binding_probability(name) - pb[k]) <= 1e-12 * max(pb[k], 1e-300)
            ranks = ts.rank_ligands()
            assert [x.name for x in ranks] == sorted(reg, key=lambda n: -reg[n])
    finally:
        srv.close()
