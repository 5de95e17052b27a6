"""Target registry semantics on the host: registration, overwrite, merge,
removal, errors, selectivities and free energies.

The values are the reference's own test cases for these queries
(tests/test_grand_partition.cpp, tests/test_target_server.cpp); everything
here is host arithmetic on the registry, so no device is needed.  ln Xi,
p_empty, p_bound and the ranking are device queries and are tested in
test_gpu_target.py.
"""
import math

import numpy as np
import pytest

from flexaidds_amd import target

KT_300 = 0.001987206 * 300.0


def gpf():
    return target.GrandPartitionFunction(ctx=None, temperature_K=300.0)


def test_invalid_temperature():
    for t in (0.0, -100.0):
        with pytest.raises(ValueError):
            target.GrandPartitionFunction(None, t)


def test_f_bound_and_delta_g():
    g = gpf()
    g.add_ligand("A", 10.0)
    assert abs(g.F_bound("A") - (-KT_300 * 10.0)) < 1e-10
    assert abs(g.delta_G_bind("A", 0.0) - g.F_bound("A")) < 1e-10
    assert abs(g.delta_G_bind("A", 2.0) - (g.F_bound("A") - 2.0)) < 1e-10


def test_overwrite_keeps_concentration_merge_is_lse():
    g = gpf()
    g.add_ligand("A", 5.0, 0.25)
    g.overwrite_ligand("A", 8.0)
    assert abs(g.F_bound("A") + KT_300 * 8.0) < 1e-10
    assert dict(g.all_log_zZ())["A"] == math.log(0.25) + 8.0
    h = gpf()
    h.add_ligand("A", 5.0)
    h.merge_ligand("A", 5.0)                     # Z_merged = 2 e^5
    assert abs(h.F_bound("A") + KT_300 * (5.0 + math.log(2.0))) < 1e-10
    # merging at c = 0.25: ln zZ = LSE(ln c + 5, ln c + 3), ln Z = that - ln c
    k = gpf()
    k.add_ligand("B", 5.0, 0.25)
    k.merge_ligand("B", 3.0)
    want = 5.0 + math.log1p(math.exp(-2.0))
    assert abs(dict(k.all_log_Z())["B"] - want) < 1e-12


def test_add_or_overwrite_and_removal():
    g = gpf()
    g.add_or_overwrite("A", 5.0)
    g.add_or_overwrite("A", 7.0, 0.5)           # replaces ln Z and the concentration
    assert g.num_ligands() == 1
    assert dict(g.all_log_Z())["A"] == 7.0 and dict(g.all_log_zZ())["A"] == math.log(0.5) + 7.0
    g.add_ligand("B", 5.0)
    g.remove_ligand("A")
    assert not g.has_ligand("A") and g.has_ligand("B") and g.num_ligands() == 1


def test_errors():
    g = gpf()
    g.add_ligand("A", 5.0)
    with pytest.raises(ValueError):
        g.add_ligand("A", 10.0)
    for c in (0.0, -1.0, 1e6):
        with pytest.raises(ValueError):
            g.add_ligand("bad", 5.0, c)
        with pytest.raises(ValueError):
            g.add_or_overwrite("bad", 5.0, c)
    for f in (lambda: g.binding_probability("X"), lambda: g.F_bound("X"), lambda: g.overwrite_ligand("X", 5.0),
              lambda: g.merge_ligand("X", 5.0), lambda: g.remove_ligand("X"), lambda: g.selectivity("A", "X")):
        with pytest.raises(ValueError):
            f()


def test_selectivities():
    g = gpf()
    g.add_ligand("A", 10.0, 1.0)
    g.add_ligand("B", 10.0, 0.01)
    assert abs(g.selectivity("A", "B") - 100.0) < 1e-9
    assert abs(g.log_intrinsic_selectivity("A", "B")) < 1e-12
    assert abs(g.log_selectivity("A", "B") - math.log(100.0)) < 1e-10
    h = gpf()
    h.add_ligand("A", 500.0)
    h.add_ligand("B", 3.0)
    assert abs(h.log_selectivity("A", "B") - 497.0) < 1e-10
    assert math.isfinite(h.selectivity("A", "B")) and h.selectivity("B", "A") < 1e-200
    x = gpf()
    x.add_ligand("X", 800.0)
    x.add_ligand("Y", 0.0)
    assert math.isinf(x.selectivity("X", "Y")) and x.selectivity("Y", "X") == 0.0


def test_sessions():
    s = target.TargetServer(ctx=None, temperature_K=310.0)
    assert s.temperature() == 310.0 and s.completed_sessions() == 0
    a, b = s.create_session("aspirin"), s.create_session("ibuprofen")
    assert (a.session_id, b.session_id) == (0, 1) and not a.completed and a.ligand_name == "aspirin"
    s.register_result(a)                         # incomplete: ignored
    assert s.completed_sessions() == 0 and not s.grand_partition().has_ligand("aspirin")
    a.log_Z, a.completed = 12.0, True
    s.register_result(a)
    b.log_Z, b.completed = 3.0, True
    b.ligand_name = "aspirin"                    # re-dock: overwrites
    s.register_result(b)
    assert s.completed_sessions() == 2 and s.grand_partition().num_ligands() == 1
    assert dict(s.grand_partition().all_log_Z())["aspirin"] == 3.0
    assert abs(s.selectivity_ratio("aspirin", "aspirin") - 1.0) < 1e-15


def test_session_from_job():
    class Lig:
        gene_typ = np.array([-1, 1, 2, 2])
        grid_xyz = np.array([[0.0, 0, 0], [1.5, -2.0, 3.25], [4.0, 4, 4]])
    res = {"thermo": {"log_Z": 41.5}, "n_poses": 640, "best": {"cf": -17.25}, "best_ic": np.array([1.0, 10, 20, 30])}
    s = target.TargetServer(None)
    x = target.session_from_job(s, "lig1", res, Lig())
    assert x.completed and x.log_Z == 41.5 and x.n_poses == 640 and x.best_energy == -17.25
    assert x.best_center == (1.5, -2.0, 3.25)
