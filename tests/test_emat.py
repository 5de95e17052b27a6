"""Energy-matrix files and the pair term (get_yval + contribution).

The density function's value is fixed by its definition (LIB/vcfunction.cpp:
576-621): 0 left of the first node, the last node's y right of the last
node, linear interpolation in between; a weight-1 pair scales the contact
area (or area / sphere with normalize_area).  The file reader follows
LIB/read_emat.cpp: one line per unordered pair over the upper triangle."""
import ctypes as C
import os

import numpy as np
import pytest

from flexaidds_amd import abi, emat, synth

REF_EMAT = "/root/reference/WRK/MC_st0r5.2_6.dat"


@pytest.fixture(scope="module")
def pair_term(oracle):
    oracle.fa_oracle_pair_term.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double]
    oracle.fa_oracle_pair_term.restype = C.c_double
    return oracle.fa_oracle_pair_term


def _params(nt, emat_v, off=None, xy=None, normalize=0):
    p = synth.Params(ntypes=nt, emat=np.asarray(emat_v, np.float32), dens_off=off, dens_xy=xy,
                     normalize_area=normalize)
    d = abi.Descriptors(synth.make_config0(n_poses=2).receptor, [], p)
    return d


def _interp(nodes, rel):
    x = nodes[:, 0].astype(np.float64)
    y = nodes[:, 1].astype(np.float64)
    if rel < x[0]:
        return 0.0
    if rel >= x[-1]:
        return float(y[-1])
    return float(np.interp(rel, x, y))


def test_density_function_closed_form(pair_term):
    rng = np.random.default_rng(4)
    nt = 3
    off, xy = synth.make_density_emat(rng, nt, frac=1.0, max_nodes=5)
    d = _params(nt, np.zeros(nt * nt), off, xy)
    for k in range(nt * nt):
        nodes = xy[off[k]:off[k + 1]]
        for rel in np.concatenate([rng.uniform(-0.2, 1.5, 300), nodes[:, 0].astype(np.float64)]):
            got = pair_term(C.byref(d.params), k, rel * 10.0, 10.0)
            want = _interp(nodes, rel * 10.0 / 10.0)
            assert abs(got - want) <= 1e-6 * max(1.0, abs(want)), (k, rel)


def test_density_exact_at_nodes_and_outside(pair_term):
    nodes = np.array([[0.2, 1.5], [0.5, -2.0], [0.9, 4.0]], np.float32)
    off = np.array([0, 3], np.int32)
    d = _params(1, [7.0], off, nodes)
    surf = 1.0
    assert pair_term(C.byref(d.params), 0, 0.1999, surf) == 0.0          # left of the first node
    assert pair_term(C.byref(d.params), 0, float(np.float32(0.2)), surf) == 1.5
    # an interior node belongs to the segment on its left (the walk advances
    # only while rel > next x), so its value is that segment's end point
    at = pair_term(C.byref(d.params), 0, float(np.float32(0.5)), surf)
    assert at != -2.0 and abs(at + 2.0) < 1e-6
    assert pair_term(C.byref(d.params), 0, 3.0, surf) == 4.0             # right of the last node
    mid = pair_term(C.byref(d.params), 0, 0.35, surf)
    assert abs(mid - (1.5 + (0.35 - np.float32(0.2)) / np.float32(0.3) * -3.5)) < 1e-6


def test_weight_one_and_normalized(pair_term):
    d = _params(2, [1.0, -3.0, -3.0, 0.5])
    assert pair_term(C.byref(d.params), 1, 2.5, 40.0) == -7.5
    dn = _params(2, [1.0, -3.0, -3.0, 0.5], normalize=1)
    assert pair_term(C.byref(dn.params), 1, 2.5, 40.0) == -3.0 * 2.5 / 40.0
    # an empty node range is a weight-1 pair
    off = np.array([0, 0, 2, 2, 2], np.int32)
    xy = np.array([[0.0, 1.0], [1.0, 2.0]], np.float32)
    dm = _params(2, [1.0, -3.0, -3.0, 0.5], off, xy)
    assert pair_term(C.byref(dm.params), 0, 2.0, 4.0) == 2.0
    assert pair_term(C.byref(dm.params), 1, 2.0, 4.0) == 1.5


def test_file_round_trip():
    rng = np.random.default_rng(9)
    nt = 6
    off, xy = synth.make_density_emat(rng, nt, frac=0.4)
    values = synth.make_emat(rng, nt)
    values[np.diff(off) > 0] = 0.0
    m = emat.EnergyMatrix(nt, values, off, xy)
    back = emat.parse_emat(emat.write_emat(m))
    assert back.ntypes == nt
    assert np.array_equal(back.emat, m.emat)
    assert np.array_equal(back.dens_off, m.dens_off)
    assert np.array_equal(back.dens_xy, m.dens_xy)


def test_file_errors():
    with pytest.raises(ValueError):
        emat.parse_emat("1-1 = 1\n1-2 = 2\n")                  # 2 lines: not triangular
    with pytest.raises(ValueError):
        emat.parse_emat("1-1 = 1 2 3\n")                       # odd node count
    m = emat.parse_emat("1-1 = 1\n1-2 = 0 1 1 2\n2-2 = -1\nno newline")
    assert m.ntypes == 2 and m.has_density
    assert list(m.dens_off) == [0, 0, 2, 4, 4]


@pytest.mark.skipif(not os.path.exists(REF_EMAT), reason="reference checkout absent")
def test_reference_matrix_file():
    m = emat.read_emat(REF_EMAT)
    assert m.ntypes == 40 and not m.has_density
    M = m.emat.reshape(40, 40)
    assert np.array_equal(M, M.T)
    # spot values as written in the file
    with open(REF_EMAT) as f:
        lines = f.readlines()
    assert M[0, 1] == np.float32(float(lines[1].split("=")[1]))
    assert M[0, 9] == np.float32(float(lines[9].split("=")[1]))
