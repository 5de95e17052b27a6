"""Writes the golden fixtures of the contact-function path from the reference
built from source (oracle/_ref/libfa_ref.so, strict IEEE build of
LIB/{ic2cf,buildcc,Vcontacts,vcfunction}.cpp plus oracle/ref/ref_driver.cpp).

Run from the repository root after `make -C oracle/ref`:
    python tests/golden/make_golden.py

Each fixture holds the seeded inputs (config, ligand index, genes), a digest
of the generated system (to detect generator drift) and the reference's
outputs: per-pose cfstr terms and status, ligand coordinates, and the contact
records of the first poses.
"""
from __future__ import annotations

import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import oracle_lib as ol  # noqa: E402
from flexaidds_amd import abi, synth  # noqa: E402


def system_digest(system, lig_index) -> str:
    h = hashlib.sha256()
    r = system.receptor
    for a in (r.xyz, r.radius, r.type):
        h.update(np.ascontiguousarray(a).tobytes())
    for f in getattr(r, "flex", None) or []:  # flexible side chains (absent: rigid receptor)
        h.update(np.int64(f.first).tobytes())
        for a in (f.rot_xyz, f.scorable, f.bonded):
            h.update(np.ascontiguousarray(a).tobytes())
    lig = system.ligands[lig_index]
    for name in ("rec", "dis", "ang", "dih", "shift", "radius", "type", "mov", "bonded", "gene_typ", "gene_atm",
                 "gene_min", "gene_max", "grid_ic"):
        h.update(np.ascontiguousarray(getattr(lig, name)).tobytes())
    h.update(np.ascontiguousarray(system.params.emat).tobytes())
    return h.hexdigest()


def make(system, lig_index, genes, n_rec_poses, path):
    ref = ol.load_ref()
    if ref is None:
        raise SystemExit("oracle/_ref/libfa_ref.so is not built (make -C oracle/ref)")
    desc = abi.Descriptors(system.receptor, system.ligands, system.params)
    out, xyz, recs = ol.ref_eval(ref, desc, lig_index, genes)
    rec_rows = [r for p in range(n_rec_poses) for r in recs[p]]
    rec_pose = np.array([p for p in range(n_rec_poses) for _ in recs[p]], np.int32)
    rec_arr = np.array(rec_rows, dtype=abi.REC_DTYPE) if rec_rows else np.zeros(0, abi.REC_DTYPE)
    np.savez_compressed(path, config=system.config, lig_index=lig_index, digest=system_digest(system, lig_index),
                        genes=genes, out=out, xyz=xyz, rec=rec_arr, rec_pose=rec_pose)
    print(path, "poses", len(genes), "statuses", np.bincount(out["status"], minlength=5), "records", len(rec_arr))


def main():
    s0 = synth.make_config0(n_poses=1000)
    make(s0, 0, np.ascontiguousarray(s0.genes[0][:96]), 8, os.path.join(HERE, "ref_config0.npz"))
    s2 = synth.make_config2(n_ligands=4, pop=64)
    for li in range(4):
        make(s2, li, np.ascontiguousarray(s2.genes[li][:32]), 4, os.path.join(HERE, f"ref_config2_lig{li}.npz"))
    s1 = synth.make_config1(pop=200)  # five flexible side chains, rotamer genes
    make(s1, 0, np.ascontiguousarray(s1.genes[0][:96]), 12, os.path.join(HERE, "ref_config1_flex.npz"))


if __name__ == "__main__":
    main()
