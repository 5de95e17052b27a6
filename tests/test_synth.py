"""Synthetic system generator: determinism and geometric invariants of the
internal-coordinate encoding (bond lengths, bond angles and dihedrals that the
rebuild must reproduce, LIB/buildcc.cpp)."""
import numpy as np

from flexaidds_amd import synth


def test_deterministic():
    a = synth.make_config0(n_poses=16)
    b = synth.make_config0(n_poses=16)
    assert np.array_equal(a.receptor.xyz, b.receptor.xyz)
    assert np.array_equal(a.genes[0], b.genes[0])
    assert np.array_equal(a.ligands[0].grid_ic, b.ligands[0].grid_ic)


def test_receptor_shell_density():
    s = synth.make_config0(n_poses=4)
    r = np.linalg.norm(s.receptor.xyz - s.receptor.center, axis=1)
    assert r.min() >= synth.CAVITY_RADIUS - 1e-3 and r.max() <= 20.0 + 1e-3
    assert s.receptor.type.min() >= 1 and s.receptor.type.max() <= 40
    assert 1.5 <= s.receptor.radius.min() and s.receptor.radius.max() <= 2.0


def test_torsions_keep_bond_geometry():
    s = synth.make_config0(n_poses=32)
    lig = s.ligands[0]
    for g in s.genes[0][:8]:
        d, a, h = synth.apply_genes(lig, g)
        x = synth.build_cartesian(lig, d, a, h)
        for i in range(1, lig.n_atoms):
            p = lig.parent[i]
            assert abs(np.linalg.norm(x[i] - x[p]) - 1.5) < 1e-3


def test_grid_gene_places_anchor():
    s = synth.make_config0(n_poses=8)
    lig = s.ligands[0]
    for k in range(1, 8):
        g = s.genes[0][k - 1].copy()
        d, a, h = synth.apply_genes(lig, g)
        x = synth.build_cartesian(lig, d, a, h)
        assert np.linalg.norm(x[0] - lig.grid_xyz[int(g[0])]) < 1e-3


def test_flexible_bond_rotates_subtree_only():
    s = synth.make_config0(n_poses=2)
    lig = s.ligands[0]
    g0 = s.genes[0][0].copy()
    g1 = g0.copy()
    g1[4] += 90.0
    x0 = synth.build_cartesian(lig, *synth.apply_genes(lig, g0))
    x1 = synth.build_cartesian(lig, *synth.apply_genes(lig, g1))
    moved = np.linalg.norm(x0 - x1, axis=1) > 1e-4
    assert moved.any() and not moved.all()
    # distances within the rigid remainder are unchanged
    keep = ~moved
    d0 = np.linalg.norm(x0[keep][:, None] - x0[keep][None], axis=2)
    d1 = np.linalg.norm(x1[keep][:, None] - x1[keep][None], axis=2)
    assert np.allclose(d0, d1, atol=1e-4)
