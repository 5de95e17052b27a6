"""Why a pose's atoms are evaluated in the reference's serial order.

The reference computes the contact polyhedra of a pose's scorable atoms one
after another in box order (LIB/Vcontacts.cpp:74-139), and each atom inherits
state from the earlier ones: the done/stub rule drops earlier atoms that
recorded no contact with it from its contact list (LIB/Vcontacts.cpp:1750-1836)
and a hull seed starts its walk at a vertex an earlier atom found
(LIB/Vcontacts.cpp:499-590, :595-629).  Evaluating every atom independently
(one wave per atom) would need both to be neutral.  They are not, within the
CF tolerance of max(1e-4 |CF|, 1e-3) kcal/mol:

* without seeds the unseeded start (closest plane, closest line, nearest
  cut) is sometimes not a polyhedron vertex, and the walk from it builds a
  different polyhedron (hundreds of kcal/mol on a few poses);
* even from a true vertex -- the oracle's speculative form casts rays from
  the atom's centre to find one, and hands every pose whose reference start
  it cannot vouch for to the serial form -- the areas differ beyond rounding:
  order_faces adds the point opposite the face's first arc point
  (LIB/Vcontacts.cpp:747-770) and calc_areas fans each face from its first
  point (LIB/Vcontacts.cpp:909-930), and which point is first follows the
  walk's order, i.e. the start vertex and the plane list.

So the areas are a function of the walk order, which only the serial chain
fixes.  These tests pin that finding on the CPU oracle (mode 0 is the
reference semantics, equal to the reference build bit for bit elsewhere).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import oracle_lib as ol
from flexaidds_amd import abi, synth


def _eval(lib, desc, genes, bits):
    lib.fa_oracle_set_mode_bits.argtypes = [C.c_int]
    lib.fa_oracle_eval_fallback.argtypes = [C.c_void_p] * 6 + [C.c_int, C.c_void_p]
    lib.fa_oracle_set_mode_bits(bits)
    try:
        g = np.ascontiguousarray(genes, np.float64)
        out = np.zeros(len(g), abi.CF_DTYPE)
        fb = np.zeros(len(g), np.int32)
        lib.fa_oracle_eval_fallback(ol._addr(desc.receptor), ol._addr(desc.ligands[0]), ol._addr(desc.params),
                                    len(g), g.ctypes.data, out.ctypes.data, 0, fb.ctypes.data)
    finally:
        lib.fa_oracle_set_mode_bits(0)
    return out, fb


@pytest.fixture(scope="module")
def sys0():
    return synth.make_config0(n_poses=1000)


def _outside(a, b):
    ok = (a["status"] == 0) & (b["status"] == 0)
    d = np.abs(a["cf"] - b["cf"])
    tol = np.maximum(1e-4 * np.abs(a["cf"]), 1e-3)
    return ok, d, (d > tol) & ok


def test_independent_atoms_leave_the_tolerance(oracle, sys0):
    """No done/stub rule and no seeds: same statuses, but poses whose CF
    moves by far more than the tolerance."""
    desc = abi.Descriptors(sys0.receptor, sys0.ligands, sys0.params)
    ref, _ = _eval(oracle, desc, sys0.genes[0], 0)
    ind, _ = _eval(oracle, desc, sys0.genes[0], 3)
    assert np.array_equal(ref["status"], ind["status"])
    ok, d, bad = _outside(ref, ind)
    assert bad.sum() >= 5 and d[ok].max() > 10.0, (int(bad.sum()), float(d[ok].max()))


def test_speculative_atoms_still_depend_on_walk_order(oracle, sys0):
    """Starts that are polyhedron vertices and serial evaluation of every
    pose the checks cannot vouch for remove the large errors, not the
    order-dependent area terms: contact pairs agree, CF does not within the
    tolerance on a few percent of the poses."""
    desc = abi.Descriptors(sys0.receptor, sys0.ligands, sys0.params)
    ref, _ = _eval(oracle, desc, sys0.genes[0], 0)
    spec, fb = _eval(oracle, desc, sys0.genes[0], 8)
    assert np.array_equal(ref["status"], spec["status"])
    ok, d, bad = _outside(ref, spec)
    scored = ref["status"] == 0
    assert 0 < fb[scored].sum() < 0.05 * scored.sum()  # poses handed to the serial form
    assert d[ok].max() < 0.1                            # no wrong polyhedron left
    assert bad.sum() > 0                                # but the areas follow the walk order
    # the same contact pairs on a pose that leaves the tolerance
    p = int(np.where(bad)[0][0])
    oracle.fa_oracle_set_mode_bits.argtypes = [C.c_int]
    _, _, ra = ol.oracle_pose(oracle, desc, 0, sys0.genes[0][p])
    oracle.fa_oracle_set_mode_bits(8)
    try:
        _, _, rb = ol.oracle_pose(oracle, desc, 0, sys0.genes[0][p])
    finally:
        oracle.fa_oracle_set_mode_bits(0)
    pairs = lambda r: {(int(x["from"]), int(x["atom"])) for x in r}
    assert pairs(ra) == pairs(rb)
    area = lambda r: {(int(x["from"]), int(x["atom"])): float(x["area"]) for x in r}
    aa, ab = area(ra), area(rb)
    assert max(abs(aa[k] - ab[k]) for k in aa) > 1e-4   # far above rounding


def test_speculative_form_serial_fallback_is_the_reference(oracle, sys0):
    """A pose the speculative form hands over is evaluated in the reference
    semantics: bit for bit the mode-0 result."""
    desc = abi.Descriptors(sys0.receptor, sys0.ligands, sys0.params)
    ref, _ = _eval(oracle, desc, sys0.genes[0], 0)
    spec, fb = _eval(oracle, desc, sys0.genes[0], 8)
    m = fb == 1
    assert m.any()
    assert np.array_equal(ref[m].tobytes(), spec[m].tobytes())
