"""Binding-cleft grid (flexaidds_amd/cleft.py) against the reference's own
detect_cleft + generate_grid + calc_cleftic compiled from its sources
(oracle/_ref/libfa_ref_reader.so, serial build).  Spheres, grid points and
their internal coordinates must be equal bit for bit.  Skipped where that
build is absent."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as ol
from file_systems import GOLDEN
from flexaidds_amd import cleft, synth

lib = ol.load_ref_reader()
pytestmark = pytest.mark.skipif(lib is None, reason="oracle/_ref reader build absent")


def _ref(xyz, rad, ori, spacer):
    mp, ms = 400000, 100000
    pts = np.zeros((mp, 3), np.float32)
    ic = np.zeros((mp, 3), np.float32)
    sph = np.zeros((ms, 4), np.float32)
    ns = np.zeros(1, np.int32)
    xyz = np.ascontiguousarray(xyz, np.float32)
    rad = np.ascontiguousarray(rad, np.float32)
    ori = np.ascontiguousarray(ori, np.float32)
    f = lib.fa_ref_cleft_grid
    f.restype = C.c_int
    n = f(xyz.ctypes.data_as(C.c_void_p), rad.ctypes.data_as(C.c_void_p), len(xyz),
          ori.ctypes.data_as(C.c_void_p), C.c_float(spacer), mp, pts.ctypes.data_as(C.c_void_p),
          ic.ctypes.data_as(C.c_void_p), ms, sph.ctypes.data_as(C.c_void_p), ns.ctypes.data_as(C.c_void_p))
    assert n >= 0, n
    return pts[:n], ic[:n], sph[:ns[0]]


def _check(xyz, rad, ori, spacer=0.375):
    pts, ic, sph = _ref(xyz, rad, ori, spacer)
    g = cleft.cleft_grid(xyz, rad, ori, spacer)
    assert np.array_equal(g.spheres, sph)
    assert np.array_equal(g.points, pts)
    assert np.array_equal(g.ic, ic)
    # every point lies inside a sphere of the cleft, and no point repeats
    d2 = ((g.points[:, None, :].astype(np.float64) - g.spheres[None, :, :3]) ** 2).sum(-1)
    assert np.all((d2 < g.spheres[None, :, 3].astype(np.float64) ** 2 + 1e-3).any(axis=1))
    assert len(np.unique(np.round(g.points * 1000).astype(np.int64), axis=0)) == len(g.points)
    return g


def test_reference_rna_receptor():
    """The reference's RNA target as its reader chain reads it: 161 spheres,
    9,043 grid points."""
    gz = np.load(GOLDEN)
    g = _check(gz["xyz"], gz["radius"], gz["center"][:3])
    assert len(g.spheres) == 161 and len(g.points) == 9043


@pytest.mark.parametrize("spacer", [0.375, 0.5])
def test_generated_receptor_with_cavity(spacer):
    """A 900-atom shell around a cavity, radii 1.5-2.0 and a few at radius 0
    (as HETATM atoms read), plus a last atom in the middle of the cavity,
    which the reference's scan leaves out; spacer 0.5 takes generate_grid's
    integer-ratio branch."""
    rng = np.random.default_rng(77)
    c = np.array([1.0, -2.0, 0.5])
    rec = synth.make_receptor(rng, 900, c, r_in=9.0, r_out=15.0)
    xyz = np.vstack([rec.xyz, c[None, :].astype(np.float32)])
    rad = np.append(rec.radius, np.float32(1.8))
    rad[:-1:97] = 0.0
    ori = xyz.astype(np.float64).mean(axis=0).astype(np.float32)
    g = _check(xyz, rad, ori, spacer)
    assert len(g.points) > 100


def test_too_few_atoms():
    with pytest.raises(ValueError):
        cleft.detect_cleft(np.zeros((1, 3), np.float32), np.ones(1, np.float32))
