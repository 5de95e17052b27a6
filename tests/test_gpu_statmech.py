"""Device Boltzmann weights and log-sum-exp against the oracle's restatement
of compute_boltzmann_batch / log_sum_exp_dispatch
(LIB/UnifiedHardwareDispatch.cpp:730-746, :768-843), and the ensemble
reduction's treatment of unscored poses (energy +inf, weight 0)."""
import ctypes as C
import math

import numpy as np
import pytest

from flexaidds_amd import abi

pytestmark = pytest.mark.gpu

KB = 0.001987206


@pytest.fixture(scope="module")
def ctx():
    from flexaidds_amd import hip
    c = hip.Context(0)
    yield c
    c.close()


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, np.float64)).cuda()


def _oracle_batch(oracle, e, T):
    e = np.ascontiguousarray(e, np.float64)
    w = np.zeros_like(e)
    lz, em = C.c_double(), C.c_double()
    oracle.fa_oracle_boltzmann_batch(e.ctypes.data, len(e), 1.0 / (KB * T), w.ctypes.data, C.byref(lz), C.byref(em))
    return w, lz.value


@pytest.mark.parametrize("case", ["random", "spread", "single", "two_state"])
@pytest.mark.parametrize("T", [300.0, 600.0])
def test_boltzmann_device_against_oracle(ctx, oracle, case, T):
    import torch
    rng = np.random.default_rng(17)
    e = {"random": rng.normal(-30.0, 8.0, 5000), "spread": np.array([-500.0, 500.0, -499.0, 0.0, 123.4]),
         "single": np.array([-42.5]), "two_state": np.array([-9.0, -10.0])}[case]
    d = _dev(e)
    w = torch.zeros_like(d)
    lz = ctx.boltzmann_device(d.data_ptr(), 1, len(e), T, w.data_ptr())
    ref_w, ref_lz = _oracle_batch(oracle, e, T)
    got = w.cpu().numpy()
    assert np.all(np.abs(got - ref_w) <= 1e-12 * np.maximum(ref_w, 1e-300) + 1e-300)
    assert abs(lz - ref_lz) <= 1e-12 * abs(ref_lz) + 1e-12
    # unnormalised: the ground state weighs exactly 1
    assert got[np.argmin(e)] == 1.0
    if case == "two_state":
        b = 1.0 / (KB * T)
        assert abs(got[0] - math.exp(-b)) < 1e-15
        assert abs(lz - (10.0 * b + math.log1p(math.exp(-b)))) < 1e-12


def test_boltzmann_device_strided_result_column(ctx, oracle):
    """The cf column of an fa_cf_result array (stride 8 doubles) read in place."""
    import torch
    rng = np.random.default_rng(5)
    n = 777
    res = np.zeros(n, abi.CF_DTYPE)
    res["cf"] = rng.normal(-20.0, 5.0, n)
    res["com"] = 1e9  # neighbouring fields must not be read
    d = torch.from_numpy(np.frombuffer(res.tobytes(), np.uint8).copy()).cuda()
    w = torch.zeros(n, dtype=torch.float64, device="cuda")
    off = abi.CF_DTYPE.fields["cf"][1]
    lz = ctx.boltzmann_device(d.data_ptr() + off, abi.CF_DTYPE.itemsize // 8, n, 300.0, w.data_ptr())
    ref_w, ref_lz = _oracle_batch(oracle, res["cf"], 300.0)
    assert np.allclose(w.cpu().numpy(), ref_w, rtol=1e-12, atol=0)
    assert abs(lz - ref_lz) <= 1e-12 * abs(ref_lz)


@pytest.mark.parametrize("case", ["random", "ln2", "big", "neg_inf", "pos_inf", "strided"])
def test_log_sum_exp_device_against_oracle(ctx, oracle, case):
    rng = np.random.default_rng(9)
    x = {"random": rng.normal(0.0, 30.0, 10001), "ln2": np.zeros(2), "big": np.array([1000.0, 1001.0, 1002.0]),
         "neg_inf": np.array([-np.inf, 3.0, -np.inf]), "pos_inf": np.array([1.0, np.inf, 2.0]),
         "strided": rng.normal(0.0, 5.0, 3 * 999)}[case]
    stride = 3 if case == "strided" else 1
    n = len(x) // stride
    d = _dev(x)
    got = ctx.log_sum_exp_device(d.data_ptr(), stride, n)
    xs = np.ascontiguousarray(x[::stride][:n])
    ref = oracle.fa_oracle_log_sum_exp(xs.ctypes.data, n)
    if math.isinf(ref):
        assert got == ref
    else:
        assert abs(got - ref) <= 1e-12 * max(1.0, abs(ref))
    if case == "ln2":
        assert abs(got - math.log(2.0)) < 1e-15


def test_log_sum_exp_device_empty(ctx):
    assert ctx.log_sum_exp_device(0, 1, 0) == -math.inf


def test_thermo_skips_unscored_poses(ctx, oracle):
    """An energy of +inf (a pose the evaluator could not score) has weight 0:
    the ensemble equals the ensemble without it."""
    rng = np.random.default_rng(3)
    e = rng.normal(-15.0, 4.0, 300)
    e_inf = e.copy()
    e_inf[::7] = np.inf
    th = ctx.thermo(e_inf, 300.0)
    fin = np.ascontiguousarray(e_inf[np.isfinite(e_inf)])
    o = abi.Thermo()
    oracle.fa_oracle_thermo(fin.ctypes.data, None, len(fin), 300.0, C.byref(o))
    for f in ("log_Z", "free_energy", "mean_energy", "entropy", "heat_capacity"):
        assert math.isfinite(th[f]), f
        assert abs(th[f] - getattr(o, f)) <= 1e-9 * abs(getattr(o, f)) + 1e-12, f
