"""Systems around receptors read from files (test helper).

rna_receptor_system: the reference's RNA target as its own reader chain reads
it (tests/golden/receptor_guanine.npz: 1,097 atoms along a 147 A helix, type
39, radius 2.0), with a generated ligand docked in one of its grooves.  The
receptor is real geometry: bonded neighbours 1.2-1.6 A apart, so the spheres
overlap far more than in the generated receptors."""
from __future__ import annotations

import os

import numpy as np

from flexaidds_amd import synth

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "receptor_guanine.npz")
# groove points 4-5 A from the nearest atom with the most atoms within 9 A
GROOVES = ((-0.15, 6.85, 45.0), (-5.15, -4.15, 103.0), (-2.15, -3.15, 87.0))


def rna_receptor_system(seed, groove=0, n_atoms=30, n_tors=6, n_poses=600, spread=2.0):
    g = np.load(GOLDEN)
    rng = np.random.default_rng(seed)
    center = np.asarray(GROOVES[groove], np.float64)
    rec = synth.Receptor(xyz=g["xyz"].copy(), radius=g["radius"].copy(), type=g["type"].copy(), center=center)
    ori = g["center"][:3].astype(np.float32)          # calc_center's centre of geometry
    pts = center[None, :] + rng.uniform(-spread, spread, size=(300, 3))
    lig = synth.make_ligand(rng, n_atoms, n_tors, ori, center, pts)
    synth.finalize_frame(lig, rec)
    params = synth.Params(ntypes=40, emat=synth.make_emat(rng))
    return synth.System(receptor=rec, ligands=[lig], params=params,
                        genes=[synth.random_genes(rng, lig, n_poses)])


def rna_cleft_system(seed, n_atoms=30, n_tors=6, n_poses=1000):
    """The same RNA target with the reference's automatic binding-site grid
    (gap-sphere cleft, grid at 0.375 A, internal coordinates as calc_cleftic
    computes them) as the ligand's translation genes."""
    from flexaidds_amd import cleft
    g = np.load(GOLDEN)
    rng = np.random.default_rng(seed)
    ori = g["center"][:3].astype(np.float32)
    cg = cleft.cleft_grid(g["xyz"], g["radius"], ori)
    anchor = cg.points.astype(np.float64).mean(axis=0)
    rec = synth.Receptor(xyz=g["xyz"].copy(), radius=g["radius"].copy(), type=g["type"].copy(), center=anchor)
    lig = synth.make_ligand(rng, n_atoms, n_tors, ori, anchor, cg.points)
    lig.grid_ic[1:] = cg.ic                         # entries 1..m exactly as the reference's
    synth.finalize_frame(lig, rec)
    params = synth.Params(ntypes=40, emat=synth.make_emat(rng))
    return synth.System(receptor=rec, ligands=[lig], params=params,
                        genes=[synth.random_genes(rng, lig, n_poses)])
