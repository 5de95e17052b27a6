"""Oracle against the reference compiled from its own sources (oracle/_ref,
built by oracle/ref/Makefile where the reference checkout is present; the
test is skipped elsewhere).  The strict IEEE build must agree bit for bit;
the reference's own -ffast-math build is reported, since the contact
function's discontinuities make it differ on a few poses."""
import numpy as np
import pytest

import oracle_lib as ol
from flexaidds_amd import abi, synth

ref = ol.load_ref()
pytestmark = pytest.mark.skipif(ref is None, reason="oracle/_ref not built (no reference checkout)")


def _tol(cf):
    return np.maximum(1e-4 * np.abs(cf), 1e-3)


@pytest.fixture(scope="module")
def sys0():
    return synth.make_config0(n_poses=1000)


def test_strict_reference_bit_exact_config0(oracle, sys0):
    desc = abi.Descriptors(sys0.receptor, sys0.ligands, sys0.params)
    genes = sys0.genes[0][:400]
    r_out, r_xyz, r_recs = ol.ref_eval(ref, desc, 0, genes)
    with ol.trig_mode(oracle, 0):
        out = ol.oracle_eval(oracle, desc, 0, genes)
        assert np.array_equal(out["status"], r_out["status"])
        for f in ("com", "wal", "sas", "totsas", "cf", "rclash"):
            assert np.array_equal(out[f], r_out[f]), f
        for p in range(0, len(genes), 7):
            o, xyz, orec = ol.oracle_pose(oracle, desc, 0, genes[p])
            assert np.array_equal(xyz, r_xyz[p])
            a = sorted((int(r["from"]), int(r["atom"]), float(r["area"])) for r in orec)
            b = sorted((int(r["from"]), int(r["atom"]), float(r["area"])) for r in r_recs[p])
            assert a == b, p


def test_strict_reference_bit_exact_uneven_ligands(oracle):
    s = synth.make_config2(n_ligands=6, pop=48)
    desc = abi.Descriptors(s.receptor, s.ligands, s.params)
    for li in range(6):
        r_out, _, _ = ol.ref_eval(ref, desc, li, s.genes[li])
        with ol.trig_mode(oracle, 0):
            out = ol.oracle_eval(oracle, desc, li, s.genes[li])
        assert np.array_equal(out["status"], r_out["status"]), li
        assert np.array_equal(out["cf"], r_out["cf"]), li


def test_fast_math_reference_sensitivity(oracle, sys0):
    fast = ol.load_ref(fast=True)
    if fast is None:
        pytest.skip("fast-math reference not built")
    desc = abi.Descriptors(sys0.receptor, sys0.ligands, sys0.params)
    genes = sys0.genes[0][:400]
    f_out, _, _ = ol.ref_eval(fast, desc, 0, genes)
    with ol.trig_mode(oracle, 0):
        out = ol.oracle_eval(oracle, desc, 0, genes)
    ok = (out["status"] == 0) & (f_out["status"] == 0)
    off = np.abs(out["cf"] - f_out["cf"])[ok] > _tol(out["cf"][ok])
    print(f"fast-math reference outside the CF tolerance on {off.sum()} of {ok.sum()} scored poses")
    assert off.sum() <= 0.2 * ok.sum()


@pytest.mark.parametrize("normalize", [0, 1])
def test_strict_reference_bit_exact_density_matrix(oracle, sys0, normalize):
    # half of the type pairs as density functions (energy_matrix.weight = 0),
    # the rest weight 1, with and without normalize_area
    import dataclasses
    off, xy = synth.make_density_emat(np.random.default_rng(17), sys0.params.ntypes, frac=0.5)
    prm = dataclasses.replace(sys0.params, dens_off=off, dens_xy=xy, normalize_area=normalize)
    desc = abi.Descriptors(sys0.receptor, sys0.ligands, prm)
    genes = sys0.genes[0][:300]
    r_out, _, _ = ol.ref_eval(ref, desc, 0, genes)
    with ol.trig_mode(oracle, 0):
        out = ol.oracle_eval(oracle, desc, 0, genes)
    assert np.array_equal(out["status"], r_out["status"])
    for f in ("com", "sas", "cf"):
        assert np.array_equal(out[f], r_out[f]), f
    # the density terms changed the scores (the test is not vacuous)
    plain = ol.oracle_eval(oracle, abi.Descriptors(sys0.receptor, sys0.ligands, sys0.params), 0, genes)
    ok = out["status"] == 0
    assert np.mean(out["com"][ok] != plain["com"][ok]) > 0.9


def test_strict_reference_bit_exact_constraints(oracle):
    # covalent constraints (ligand-receptor and ligand-ligand): the con term,
    # the shortened clash distance and the skipped contact energy
    s = synth.make_config0(n_poses=400)
    cons = synth.add_constraints(s, n_receptor=3, n_ligand=2)
    start = 1000.0 * (3 + 2 * 2)                  # KDIST per constrained ligand atom
    desc = abi.Descriptors(s.receptor, s.ligands, s.params)
    genes = s.genes[0][:300]
    r_out, _, _ = ref_eval_cons(desc, genes)
    with ol.trig_mode(oracle, 0):
        out = ol.oracle_eval(oracle, desc, 0, genes)
    assert np.array_equal(out["status"], r_out["status"])
    for f in ("com", "wal", "sas", "con", "cf", "rclash"):
        assert np.array_equal(out[f], r_out[f]), f
    ok = out["status"] == 0
    assert np.all(out["con"][ok] <= start)
    # many poses bring a constrained pair into contact (con below the start)
    assert np.sum(out["con"][ok] < start - 1e-6) >= 20


def ref_eval_cons(desc, genes):
    return ol.ref_eval(ref, desc, 0, genes)


def test_strict_reference_bit_exact_rna_receptor_file(oracle):
    """The reference's RNA target as its reader chain reads it, with a
    generated ligand in two of its grooves."""
    from file_systems import rna_receptor_system
    for groove in (0, 1):
        s = rna_receptor_system(21 + groove, groove=groove, n_poses=400)
        desc = abi.Descriptors(s.receptor, s.ligands, s.params)
        r_out, _, _ = ol.ref_eval(ref, desc, 0, s.genes[0])
        with ol.trig_mode(oracle, 0):
            out = ol.oracle_eval(oracle, desc, 0, s.genes[0])
        assert (r_out["status"] == 0).sum() >= 10
        assert np.array_equal(out["status"], r_out["status"])
        for f in ("com", "wal", "sas", "totsas", "cf", "rclash"):
            assert np.array_equal(out[f], r_out[f]), f


def test_strict_reference_bit_exact_rna_cleft_grid(oracle):
    """The RNA target with the reference's automatic binding-site grid."""
    from file_systems import rna_cleft_system
    s = rna_cleft_system(52, n_poses=300)
    desc = abi.Descriptors(s.receptor, s.ligands, s.params)
    r_out, _, _ = ol.ref_eval(ref, desc, 0, s.genes[0])
    with ol.trig_mode(oracle, 0):
        out = ol.oracle_eval(oracle, desc, 0, s.genes[0])
    assert (r_out["status"] == 0).sum() >= 20
    assert np.array_equal(out["status"], r_out["status"])
    for f in ("com", "wal", "sas", "totsas", "cf", "rclash"):
        assert np.array_equal(out[f], r_out[f]), f
