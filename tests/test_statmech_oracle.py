"""Ensemble-reduction oracle against closed forms.

Cases restate the reference's analytic tests: tests/test_statmech.cpp
(single state :71-106, two-state :108-131, equal energies :199-228,
stability :375-402, :695-734), tests/test_unified_dispatch.cpp:166-216
(log-sum-exp known results), python/tests/test_statmech.py:61-97
(first-principles reference on [-10.5, -9.8, -10.2, -11.0, -9.5] at 300 K,
1e-9), tests/test_grand_partition.cpp:41-130 (Ξ closed forms).
"""
import ctypes as C
import math

import numpy as np
import pytest

from flexaidds_amd import abi

KB = 0.001987206
T = 300.0


def thermo(o, e, cnt=None, temp=T):
    e = np.ascontiguousarray(e, np.float64)
    c = None if cnt is None else np.ascontiguousarray(cnt, np.int32)
    th = abi.Thermo()
    rc = o.fa_oracle_thermo(e.ctypes.data, None if c is None else c.ctypes.data, len(e), temp, C.byref(th))
    assert rc == 0
    return th


def analytic(energies, temp=T):
    beta = 1.0 / (KB * temp)
    x = [-beta * e for e in energies]
    m = max(x)
    lz = m + math.log(sum(math.exp(v - m) for v in x))
    w = [math.exp(v - lz) for v in x]
    e1 = sum(a * b for a, b in zip(w, energies))
    e2 = sum(a * b * b for a, b in zip(w, energies))
    F = -KB * temp * lz
    return dict(log_Z=lz, free_energy=F, mean_energy=e1, mean_energy_sq=e2, entropy=(e1 - F) / temp,
                heat_capacity=(e2 - e1 * e1) / (KB * temp) ** 2, std_energy=math.sqrt(max(0.0, e2 - e1 * e1)))


def test_python_reference_energies(oracle):
    E = [-10.5, -9.8, -10.2, -11.0, -9.5]
    th = thermo(oracle, E)
    ref = analytic(E)
    for k, v in ref.items():
        assert abs(getattr(th, k) - v) < 1e-9, k


def test_single_state(oracle):
    th = thermo(oracle, [-12.0])
    assert abs(th.free_energy + 12.0) < 1e-6 and abs(th.entropy) < 1e-6 and abs(th.heat_capacity) < 1e-6


def test_single_state_multiplicity(oracle):
    th = thermo(oracle, [-10.0], [5])
    kT = KB * T
    assert abs(th.free_energy - (-10.0 - kT * math.log(5))) < 1e-9
    assert abs(th.entropy - KB * math.log(5)) < 1e-9


def test_two_state(oracle):
    beta = 1.0 / (KB * T)
    th = thermo(oracle, [-10.0, -8.0])
    Z = math.exp(10 * beta) + math.exp(8 * beta)
    assert abs(th.log_Z - math.log(Z)) < 1e-9
    assert abs(th.free_energy - (-(KB * T) * math.log(Z))) < 1e-9


@pytest.mark.parametrize("N", [8, 1000])
def test_equal_energies(oracle, N):
    th = thermo(oracle, [-7.77] * N)
    assert abs(th.entropy - KB * math.log(N)) < 1e-9
    assert abs(th.heat_capacity) < 1e-9


def test_high_multiplicity(oracle):
    th = thermo(oracle, [-5.0], [10000])
    assert abs(th.entropy - KB * math.log(10000)) < 1e-9
    assert abs(th.free_energy - (-5.0 - KB * T * math.log(10000))) < 1e-9


def test_stability_large_gap(oracle):
    th = thermo(oracle, [-500.0, 0.0])
    assert all(math.isfinite(getattr(th, k)) for k in ("free_energy", "mean_energy", "entropy", "heat_capacity"))


def test_near_degenerate(oracle):
    th = thermo(oracle, [-10.0 + i * 1e-10 for i in range(100)])
    assert abs(th.entropy - KB * math.log(100)) < 0.01


def test_lse_known(oracle):
    two = np.array([0.0, 0.0])
    assert abs(oracle.fa_oracle_log_sum_exp(two.ctypes.data, 2) - math.log(2)) < 1e-12
    one = np.array([5.0])
    assert abs(oracle.fa_oracle_log_sum_exp(one.ctypes.data, 1) - 5.0) < 1e-12
    big = np.array([1000.0, 1001.0, 1002.0])
    v = oracle.fa_oracle_log_sum_exp(big.ctypes.data, 3)
    assert math.isfinite(v) and v > 1001.0
    assert oracle.fa_oracle_log_sum_exp(None, 0) <= -1e300


def test_boltzmann_weights(oracle):
    e = np.array([-20.0, -15.0, -10.0, -5.0, 0.0, 5.0])
    w = np.zeros_like(e)
    oracle.fa_oracle_boltzmann(e.ctypes.data, None, len(e), T, w.ctypes.data)
    assert abs(w.sum() - 1.0) < 1e-12 and np.all(np.diff(w) <= 0)


def test_grand_partition(oracle):
    lz = np.array([10.0, 8.0])
    c = np.array([1.0, 1.0])
    pb = np.zeros(2)
    pe = C.c_double()
    lx = oracle.fa_oracle_log_xi(lz.ctypes.data, c.ctypes.data, 2, pb.ctypes.data, C.byref(pe))
    xi = 1 + math.exp(10) + math.exp(8)
    assert abs(math.exp(lx) - xi) < xi * 1e-10
    assert abs(pb.sum() + pe.value - 1.0) < 1e-10
    assert abs(pb[0] / pb[1] - math.exp(2.0)) < 1e-10
    one = np.array([10.0])
    c1 = np.array([1.0])
    lx1 = oracle.fa_oracle_log_xi(one.ctypes.data, c1.ctypes.data, 1, pb.ctypes.data, C.byref(pe))
    assert abs(lx1 - math.log(1 + math.exp(10))) < 1e-10


BETA300 = 1.0 / (KB * T)  # kB of LIB/statmech.h:26


def boltz_batch(o, e, beta):
    e = np.ascontiguousarray(e, np.float64)
    w = np.zeros_like(e)
    lz, em = C.c_double(), C.c_double()
    assert o.fa_oracle_boltzmann_batch(e.ctypes.data, len(e), beta, w.ctypes.data, C.byref(lz), C.byref(em)) == 0
    return w, lz.value, em.value


def test_boltzmann_batch_two_state_closed_form(oracle):
    # compute_boltzmann_batch returns exp(-beta (E - Emin)) unnormalised:
    # the lowest state weighs exactly 1, the other exp(-beta dE)
    w, lz, em = boltz_batch(oracle, [-9.0, -10.0], BETA300)
    assert em == -10.0 and w[1] == 1.0
    assert abs(w[0] - math.exp(-BETA300 * 1.0)) < 1e-15
    assert abs(w[0] / w[1] - math.exp(-BETA300 * (-9.0 - -10.0))) < 1e-15
    # ln Z = ln(exp(10 b) + exp(9 b)) = 10 b + log1p(exp(-b))
    assert abs(lz - (10.0 * BETA300 + math.log1p(math.exp(-BETA300)))) < 1e-12
    # the temperature enters only through beta = 1 / (kB T): kB = 0.001987206
    assert abs(BETA300 - 1.677396975) < 1e-8


def test_boltzmann_batch_single_and_extremes(oracle):
    w, lz, em = boltz_batch(oracle, [-42.5], BETA300)
    assert w[0] == 1.0 and abs(lz - 42.5 * BETA300) < 1e-12 and em == -42.5
    # +-500 kcal/mol: the high state underflows to 0, ln Z stays the ground state's
    w, lz, em = boltz_batch(oracle, [-500.0, 500.0, -499.0], BETA300)
    assert w[0] == 1.0 and w[1] == 0.0 and abs(w[2] - math.exp(-BETA300)) < 1e-15
    assert abs(lz - (500.0 * BETA300 + math.log1p(math.exp(-BETA300)))) < 1e-9
    # empty span: the reference's {weights {}, log_Z 0, E_min 0}
    w, lz, em = boltz_batch(oracle, np.zeros(0), BETA300)
    assert lz == 0.0 and em == 0.0


def test_boltzmann_weights_multiplicity_closed_form(oracle):
    # StatMechEngine::boltzmann_weights: w_i = n_i exp(-b E_i) / sum_j n_j exp(-b E_j)
    e = np.array([0.0, 1.0, 2.0])
    n = np.array([1, 2, 3], np.int32)
    w = np.zeros(3)
    lz = oracle.fa_oracle_boltzmann(e.ctypes.data, n.ctypes.data, 3, T, w.ctypes.data)
    b = BETA300
    raw = np.array([1.0, 2.0 * math.exp(-b), 3.0 * math.exp(-2.0 * b)])
    assert np.allclose(w, raw / raw.sum(), rtol=0, atol=1e-15)
    # the returned ln Z is the multiplicity-free batch's (helmholtz, LIB/statmech.cpp:213-224)
    assert abs(lz - math.log(1.0 + math.exp(-b) + math.exp(-2.0 * b))) < 1e-14
    # without multiplicities every state weighs 1 x its factor
    w1 = np.zeros(3)
    oracle.fa_oracle_boltzmann(e.ctypes.data, None, 3, T, w1.ctypes.data)
    raw1 = np.array([1.0, math.exp(-b), math.exp(-2.0 * b)])
    assert np.allclose(w1, raw1 / raw1.sum(), rtol=0, atol=1e-15)
    # the temperature is read as Kelvin with kB = 0.001987206: at 600 K the
    # ratio of neighbouring states is exp(-b / 2)
    w6 = np.zeros(3)
    oracle.fa_oracle_boltzmann(e.ctypes.data, None, 3, 600.0, w6.ctypes.data)
    assert abs(w6[1] / w6[0] - math.exp(-b / 2.0)) < 1e-14
