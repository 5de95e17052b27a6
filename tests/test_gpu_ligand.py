"""A MOL2 ligand docked on the device the way the reference's direct mode
sets it up: the reference's RNA target as its own reader chain reads it
(tests/golden/receptor_guanine.npz), a generated 12-atom MOL2 ligand at the
most open point of the target's cleft, ori = the ligand's centroid, the
automatic binding-site grid over the receptor's atoms followed by the
ligand's, the translation and rotation genes (flexaidds_amd/ligand.py,
pinned against the reference's reader chain by tests/test_ligand.py).
Random poses of this set-up nearly all clash, so the poses compared are a
device-GA dock's: its final population, device against the oracle."""
import os

import numpy as np
import pytest

import oracle_lib as ol
from flexaidds_amd import abi, cleft, ligand, synth
from mol2_gen import write_mol2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    from flexaidds_amd import hip as h
    return h


def test_mol2_direct_mode_dock_against_oracle(hip, oracle, tmp_path):
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "receptor_guanine.npz"))
    cg0 = cleft.cleft_grid(g["xyz"], g["radius"], g["center"][:3].astype(np.float32))
    far = np.array([np.min(np.sum((g["xyz"] - p) ** 2, axis=1)) for p in cg0.points])
    site = cg0.points[int(np.argmax(far))].astype(np.float64)
    rng = np.random.default_rng(77)
    path = str(tmp_path / "lig.mol2")
    write_mol2(path, rng, 12, extra_bonds=2, origin=site)
    params = synth.Params(ntypes=40, emat=synth.make_emat(rng))
    s, cg = ligand.direct_system(g["xyz"], g["radius"], g["type"], g["center"], path, params)
    lig = s.ligands[0]
    assert lig.n_genes == 4 and len(cg.points) > 1000
    with hip.Context() as c:
        c.upload_system(s)
        ga = hip.GA(c, synth.ga_limits(lig), population=2000, seed=5, temperature=300.0)
        ga.run(1, graph=False)
        _, _, r0 = ga.population()
        ga.run(40, graph=True)
        _, ic, r1 = ga.population()
        ga.close()
        genes = np.ascontiguousarray(ic[:2000, :lig.n_genes])
        out = c.eval(0, genes)
    assert float(r1["cf"][:2000].min()) <= float(r0["cf"][:2000].min())
    d = abi.Descriptors(s.receptor, s.ligands, s.params)
    with ol.gpu_mirror(oracle):
        ref = ol.oracle_eval(oracle, d, 0, genes)
    assert np.array_equal(out["status"], ref["status"])
    ok = ref["status"] == 0
    assert ok.sum() >= 50, np.unique(ref["status"], return_counts=True)
    tol = np.maximum(1e-4 * np.abs(ref["cf"][ok]), 1e-3)
    assert np.all(np.abs(out["cf"][ok] - ref["cf"][ok]) <= tol)
    assert np.abs(out["cf"] - ref["cf"])[ok].max() < 1e-9
