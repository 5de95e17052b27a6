"""Known-answer tests of the oracle's Vcontacts helpers.

The cases restate the reference's own unit tests, tests/test_vcontacts.cpp
(solve_3x3 :22-74, solve_2xS :80-142, cosPQR :148-191, add_vertex distance
:197-246, spherical_arc :252-278), plus closed forms the geometry fixes.
"""
import ctypes as C
import math

import numpy as np
import pytest

EPS = 1e-6


def _d(*v):
    return np.array(v, np.float64)


def solve3(o, e0, e1, e2):
    pt = np.zeros(3)
    rc = o.fa_oracle_solve_3x3(e0.ctypes.data, e1.ctypes.data, e2.ctypes.data, pt.ctypes.data)
    return rc, pt


def solve2s(o, a0, a1, r):
    p0, p1 = np.zeros(3), np.zeros(3)
    rc = o.fa_oracle_solve_2xS(a0.ctypes.data, a1.ctypes.data, r, p0.ctypes.data, p1.ctypes.data)
    return rc, p0, p1


def test_solve_3x3_identity(oracle):
    rc, pt = solve3(oracle, _d(1, 0, 0, -3), _d(0, 1, 0, -5), _d(0, 0, 1, -7))
    assert rc == 0
    assert np.allclose(pt, [3, 5, 7], atol=EPS)


def test_solve_3x3_general(oracle):
    e0, e1, e2 = _d(2, 1, -1, -3), _d(1, -1, 2, -1), _d(3, 2, 1, -10)
    rc, pt = solve3(oracle, e0, e1, e2)
    assert rc == 0
    for e in (e0, e1, e2):
        assert abs(e[:3] @ pt + e[3]) < EPS


@pytest.mark.parametrize("eqs", [
    (_d(1, 2, 3, -4), _d(1, 2, 3, -4), _d(0, 0, 1, -1)),
    (_d(0, 0, 0, 0), _d(0, 0, 0, 0), _d(0, 0, 0, 0)),
])
def test_solve_3x3_singular(oracle, eqs):
    rc, _ = solve3(oracle, *eqs)
    assert rc == -1


def test_solve_2xS_orthogonal(oracle):
    rc, p0, p1 = solve2s(oracle, _d(1, 0, 0, 0), _d(0, 1, 0, 0), 2.0)
    assert rc == 0
    assert abs(p0[0]) < EPS and abs(p0[1]) < EPS and abs(abs(p0[2]) - 2) < EPS
    assert abs(p1[0]) < EPS and abs(p1[1]) < EPS and abs(abs(p1[2]) - 2) < EPS
    assert abs(p0[2] + p1[2]) < EPS


def test_solve_2xS_too_small(oracle):
    rc, _, _ = solve2s(oracle, _d(1, 0, 0, -5), _d(0, 1, 0, -5), 1.0)
    assert rc == -1


def test_solve_2xS_parallel(oracle):
    rc, _, _ = solve2s(oracle, _d(1, 0, 0, 0), _d(1, 0, 0, -1), 5.0)
    assert rc == -1


def test_solve_2xS_on_sphere(oracle):
    rc, p0, p1 = solve2s(oracle, _d(1, 0, 0, 0), _d(0, 0, 1, 0), 3.0)
    assert rc == 0
    assert abs(np.linalg.norm(p0) - 3) < EPS and abs(np.linalg.norm(p1) - 3) < EPS


def test_solve_2xS_general_closed_form(oracle):
    # both points satisfy both plane equations and lie on the sphere
    a0 = _d(0.6, 0.8, 0.0, -0.9)
    a1 = _d(0.0, 0.6, 0.8, -1.1)
    rc, p0, p1 = solve2s(oracle, a0, a1, 3.25)
    assert rc == 0
    for p in (p0, p1):
        assert abs(a0[:3] @ p + a0[3]) < 1e-12
        assert abs(a1[:3] @ p + a1[3]) < 1e-12
        assert abs(p @ p - np.float32(3.25) ** 2) < 1e-9


@pytest.mark.parametrize("P,Q,R,expect", [
    ((1, 0, 0), (0, 0, 0), (0, 1, 0), 0.0),
    ((2, 0, 0), (0, 0, 0), (5, 0, 0), 1.0),
    ((1, 0, 0), (0, 0, 0), (-3, 0, 0), -1.0),
    ((1, 0, 0), (0, 0, 0), (1, 1, 0), math.cos(math.pi / 4)),
    ((1, 1, 1), (0, 0, 0), (-1, -1, -1), -1.0),
])
def test_cosPQR(oracle, P, Q, R, expect):
    p, q, r = _d(*P), _d(*Q), _d(*R)
    v = oracle.fa_oracle_cosPQR(p.ctypes.data, q.ctypes.data, r.ctypes.data)
    assert abs(v - expect) < EPS


def _arc(o, A, B, Cc, r):
    A, B, Cc = _d(*A), _d(*B), _d(*Cc)
    return o.fa_oracle_spherical_arc(A.ctypes.data, float(np.linalg.norm(A)), B.ctypes.data,
                                     float(np.linalg.norm(B)), Cc.ctypes.data, float(np.linalg.norm(Cc)), r)


def test_spherical_arc_finite(oracle):
    a1 = _arc(oracle, (0, 0, 1), (1, 0, 0), (0, 1, 0), 1.0)
    a2 = _arc(oracle, (0, 0, 1), (1, 0, 0), (0, 1, 0), 2.0)
    assert math.isfinite(a1) and math.isfinite(a2)


def test_spherical_arc_closed_form(oracle):
    """Cap sector minus spherical triangle, checked with an independent
    formula: for a cap whose plane sits at distance d from the centre and two
    rim points an azimuth phi apart, the segment area is
    phi * R * (R - d) - R^2 * E, with E the excess of the triangle (pole, B, C)
    from the Van Oosterom-Strackee formula (the oracle uses L'Huilier's)."""
    for R, d, phi_deg in ((2.0, 0.7, 70.0), (3.1, 1.9, 35.0), (2.9, 0.05, 150.0)):
        phi = math.radians(phi_deg)
        rho = math.sqrt(R * R - d * d)
        A = (0.0, 0.0, d)
        B = (rho, 0.0, d)
        Cc = (rho * math.cos(phi), rho * math.sin(phi), d)
        p, b, c = np.array([0, 0, 1.0]), np.array(B) / R, np.array(Cc) / R
        E = 2.0 * math.atan2(abs(p @ np.cross(b, c)), 1.0 + p @ b + b @ c + c @ p)
        expect = phi * R * (R - d) - R * R * E
        got = _arc(oracle, A, B, Cc, R)
        assert abs(got - expect) < 1e-6 * max(1.0, abs(expect)), (R, d, phi_deg, got, expect)


def test_add_vertex_distance_is_norm():
    # add_vertex stores sqrt(x^2+y^2+z^2) (tests/test_vcontacts.cpp:220-246)
    assert math.isclose(math.sqrt(3 * 3 + 4 * 4), 5.0)
