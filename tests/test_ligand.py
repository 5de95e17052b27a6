"""MOL2 ligand reader and direct-mode set-up against the reference's own.

The reference's reader chain (LIB/Mol2Reader.cpp, buildic, buildic_point,
bondedlist, update_bonded, add2_optimiz_vec, ic_bounds, buildlist,
create_rebuild_list) compiles from its sources into
oracle/_ref/libfa_ref_reader.so (oracle/ref/Makefile), and fa_ref_read_mol2
runs it in LIB/top.cpp's order.  Generated MOL2 files -- random trees with
ring-closing bonds, shuffled and non-contiguous atom ids, SYBYL types from
the reader's table and unknown ones, records with and without the optional
columns, a name longer than %7s takes, lines longer than the 256-byte read
buffer, a seventh bond on one atom, a bond naming an unknown id -- are read
by both; coordinates, types, radii, the reconstruction tree, the internal
coordinates, the bonded matrix, the rebuild order, the genes, ori and the
cleft grid's row 0 must agree bit for bit.  The reference ships no MOL2
file, so the inputs are generated.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle_lib as ol
from flexaidds_amd import abi, ligand, synth
from mol2_gen import write_mol2

ref = ol.load_ref_reader()
needs_ref = pytest.mark.skipif(ref is None, reason="reference reader build (oracle/_ref) absent")

CASES = [
    dict(n_atoms=12, seed=1),
    dict(n_atoms=30, seed=2, extra_bonds=6),
    dict(n_atoms=45, seed=3, short_records=True),
    dict(n_atoms=20, seed=4, long_name=True),
    dict(n_atoms=25, seed=5, long_line=True),
    dict(n_atoms=18, seed=6, seventh_bond=True, bad_bond=True),
    dict(n_atoms=3, seed=7, extra_bonds=0),
    dict(n_atoms=80, seed=8, extra_bonds=10),
]


@needs_ref
@pytest.mark.parametrize("case", CASES, ids=[f"n{c['n_atoms']}s{c['seed']}" for c in CASES])
def test_mol2_setup_equals_reference(tmp_path, case):
    case = dict(case)
    rng = np.random.default_rng(case.pop("seed"))
    path = str(tmp_path / "lig.mol2")
    write_mol2(path, rng, **case)
    n_grid = 500
    r = ol.ref_read_mol2(ref, path, n_grid)
    m = ligand.read_mol2(path)
    st = ligand.setup(m, n_grid)
    assert m.n_atoms == len(r["xyz"])
    assert np.array_equal(m.xyz, r["xyz"])
    assert np.array_equal(m.type, r["type"])
    assert np.array_equal(m.radius, r["radius"])
    assert np.array_equal(st.rec, r["rec"])
    assert np.array_equal(st.ic.view(np.uint32), r["ic"].view(np.uint32))
    assert np.array_equal(st.bonded, r["bonded"])
    assert np.array_equal(st.mov, r["mov"])
    assert np.array_equal(st.gene_typ, r["gene_typ"]) and np.array_equal(st.gene_atm, r["gene_atm"])
    assert np.array_equal(st.gene_min, r["gene_min"]) and np.array_equal(st.gene_max, r["gene_max"])
    assert np.array_equal(st.ori, r["ori"])
    assert np.array_equal(st.grid0, r["grid0"])


@needs_ref
def test_planted_mistakes_fail(tmp_path):
    """Each rule the comparison rests on, broken on purpose in the
    restatement, makes the reference disagree (a check that the test can
    see it)."""
    rng = np.random.default_rng(11)
    path = str(tmp_path / "lig.mol2")
    write_mol2(path, rng, 40, extra_bonds=6, seventh_bond=True)
    r = ol.ref_read_mol2(ref, path, 300)
    m = ligand.read_mol2(path)
    st = ligand.setup(m, 300)
    # breadth-first tree taken depth first
    rec_df = st.rec.copy()
    rec_df[:, 1:] = -1
    assert not np.array_equal(rec_df, r["rec"])
    # bonds beyond six per atom kept
    m2 = ligand.read_mol2(path)
    assert max(len(b) for b in m2.bonds) == ligand.MAX_BONDS
    # bonded within three loops instead of two
    assert not np.array_equal(ligand.bonded_matrix(m.bonds, m.n_atoms, 3), r["bonded"])
    # centroid in double instead of float accumulation
    ori64 = m.xyz.astype(np.float64).mean(axis=0).astype(np.float32)
    ic64 = ligand.buildic(m.xyz, st.rec, ori64)
    assert not np.array_equal(ic64.view(np.uint32), r["ic"].view(np.uint32)) or \
        np.array_equal(ori64, r["ori"])
    # rebuild order without moving the GPA atoms to the front
    assert not np.array_equal(ligand.rebuild_order(st.rec, (5, 6, 7)), r["mov"])


def test_mol2_ligand_builds_its_own_pose(tmp_path):
    """The descriptor a MOL2 ligand yields rebuilds the file's coordinates:
    genes equal to the set-up's own values (grid row 0, the GPA atoms'
    angles) give back the input pose through the oracle's buildcc (up to
    the +1e-10 the reference adds per reconstruction step and float
    rounding)."""
    rng = np.random.default_rng(21)
    path = str(tmp_path / "lig.mol2")
    write_mol2(path, rng, 35, extra_bonds=4)
    m = ligand.read_mol2(path)
    st = ligand.setup(m, 10)
    grid_ic = np.tile(st.grid0[3:], (9, 1))
    lig = ligand.descriptor(st, grid_ic, m.xyz.min(0) - 5, m.xyz.max(0) + 5, 50.0)
    genes = np.array([[0.0, st.ic[st.gpa[1], 1], st.ic[st.gpa[1], 2], st.ic[st.gpa[2], 2]]])
    lig.gene_min[0] = 0.0
    oracle = ol.load_oracle()
    rec = synth.Receptor(xyz=np.zeros((4, 3), np.float32) + 60, radius=np.full(4, 1.8, np.float32),
                         type=np.ones(4, np.int32), center=np.zeros(3))
    d = abi.Descriptors(rec, [lig], synth.Params(ntypes=40, emat=np.zeros((40, 40), np.float32)))
    _, xyz, _ = ol.oracle_pose(oracle, d, 0, genes[0])
    assert np.abs(xyz - m.xyz).max() < 2e-3


SDF_CASES = [
    dict(n_atoms=15, seed=31),
    dict(n_atoms=40, seed=32, extra_bonds=5, tail_symbol=True),
    dict(n_atoms=22, seed=33, bad_bond=True, seventh_bond=True),
    dict(n_atoms=30, seed=34, short_coord=True),
]


@needs_ref
@pytest.mark.parametrize("case", SDF_CASES, ids=[f"n{c['n_atoms']}s{c['seed']}" for c in SDF_CASES])
def test_sdf_reader_equals_reference(tmp_path, case):
    """Atoms (coordinates, element types and radii, including a symbol that
    runs into the line's newline and a coordinate that does not convert)
    and bond lists as the reference's SDF reader leaves them."""
    from mol2_gen import write_sdf
    case = dict(case)
    rng = np.random.default_rng(case.pop("seed"))
    path = str(tmp_path / "lig.sdf")
    write_sdf(path, rng, **case)
    r = ol.ref_read_sdf(ref, path)
    m = ligand.read_sdf(path)
    assert m.n_atoms == len(r["xyz"])
    assert np.array_equal(m.xyz, r["xyz"])
    assert np.array_equal(m.type, r["type"])
    assert np.array_equal(m.radius, r["radius"])
    assert [list(b) for b in m.bonds] == [[int(x) for x in b] for b in r["bonds"]]
    if case.get("tail_symbol"):  # the symbol runs into the newline: no element matches, dummy type
        assert np.all(r["type"][::3] == ligand.DUMMY_TYPE) and np.all(m.type[::3] == ligand.DUMMY_TYPE)


def test_sdf_ligand_builds_its_own_pose(tmp_path):
    """An SDF ligand set up with the MOL2 path's tree rebuilds its own pose."""
    from mol2_gen import write_sdf
    rng = np.random.default_rng(41)
    path = str(tmp_path / "lig.sdf")
    write_sdf(path, rng, 26, extra_bonds=3)
    m = ligand.read_sdf(path)
    st = ligand.setup(m, 10)
    lig = ligand.descriptor(st, np.tile(st.grid0[3:], (9, 1)), m.xyz.min(0) - 5, m.xyz.max(0) + 5, 50.0)
    lig.gene_min[0] = 0.0
    genes = np.array([0.0, st.ic[st.gpa[1], 1], st.ic[st.gpa[1], 2], st.ic[st.gpa[2], 2]])
    rec = synth.Receptor(xyz=np.zeros((4, 3), np.float32) + 60, radius=np.full(4, 1.8, np.float32),
                         type=np.ones(4, np.int32), center=np.zeros(3))
    d = abi.Descriptors(rec, [lig], synth.Params(ntypes=40, emat=np.zeros((40, 40), np.float32)))
    _, xyz, _ = ol.oracle_pose(ol.load_oracle(), d, 0, genes)
    assert np.abs(xyz - m.xyz).max() < 2e-3
