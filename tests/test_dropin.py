"""The C++ drop-in layer (include/hip_eval.h) run the way a reference-side
caller uses it: tests/dropin/dropin.cpp calls hip_eval::init / upload_system
/ batch_evaluate / batch_evaluate_full (LIB/hip_eval.h:41-166),
voronoi_cf_hip::batch_eval with a status column and gene limits
(LIB/VoronoiCFBatch.h:184-247: clamp, 'n' rows skipped and left zero, cf and
apparent CF per chromosome) and rocm_dispatch::compute / log_sum_exp /
boltzmann_batch (LIB/statmech.cpp:73-173, LIB/UnifiedHardwareDispatch.cpp:
730-843), and is loaded here with ctypes.  The GPU tests compare every output
with the oracle and with the C ABI path; the CPU tests check the layer's
argument errors, which fire before any device work."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle_lib as ol
from flexaidds_amd import abi, synth

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "dropin", "_build", "libfa_dropin.so")
KB = 0.001987206


class DropinOutputs(C.Structure):
    _fields_ = [("com", C.c_void_p), ("wal", C.c_void_p), ("sas", C.c_void_p), ("full", C.c_void_p),
                ("vcf", C.c_void_p), ("vapp", C.c_void_p), ("wall_ms", C.c_double), ("thermo", abi.Thermo),
                ("lse", C.c_double), ("weights", C.c_void_p), ("log_z", C.c_double), ("e_min", C.c_double),
                ("available", C.c_int)]


@pytest.fixture(scope="module")
def dropin():
    if not os.path.exists(SO):
        pytest.fail(f"{SO} is not built (python tests/dropin/build.py)")
    # one HIP runtime per process: torch's, loaded first with the library (as
    # flexaidds_amd.hip does), before this object pulls the library in
    from flexaidds_amd import hip  # noqa: F401
    lib = C.CDLL(SO)
    lib.dropin_session.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                   C.c_char_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_double,
                                   C.POINTER(DropinOutputs), C.c_char_p, C.c_int]
    lib.dropin_session.restype = C.c_int
    lib.dropin_error_case.argtypes = [C.c_int, C.c_char_p, C.c_int]
    lib.dropin_error_case.restype = C.c_int
    return lib


@pytest.mark.parametrize("which,kind,text", [
    (0, "invalid_argument", "ligand index out of range"),
    (1, "invalid_argument", "bad gene shape"),
    (2, "invalid_argument", "gene_lim size differs"),
    (3, "invalid_argument", "only one value per type pair"),
    (4, "runtime_error", "init() has not been called"),
])
def test_dropin_argument_errors(dropin, which, kind, text):
    buf = C.create_string_buffer(512)
    assert dropin.dropin_error_case(which, buf, 512) == -1
    msg = buf.value.decode()
    assert msg.startswith(kind + ":") and text in msg, msg


def _session(dropin, system, lig, genes, status, glim, energies, counts, T=300.0):
    d = abi.Descriptors(system.receptor, system.ligands, system.params)
    pop = len(genes)
    genes = np.ascontiguousarray(genes, np.float64)
    energies = np.ascontiguousarray(energies, np.float64)
    arr = {k: np.zeros(pop) for k in ("com", "wal", "sas", "vapp")}
    full = np.zeros(pop, abi.CF_DTYPE)
    vcf = np.zeros(pop, abi.CF_DTYPE)
    w = np.zeros(len(energies))
    o = DropinOutputs(arr["com"].ctypes.data, arr["wal"].ctypes.data, arr["sas"].ctypes.data, full.ctypes.data,
                      vcf.ctypes.data, arr["vapp"].ctypes.data, 0.0, abi.Thermo(), 0.0, w.ctypes.data, 0.0, 0.0, 0)
    ligs = (abi.LigandDesc * len(d.ligands))(*d.ligands)
    cnt = None if counts is None else np.ascontiguousarray(counts, np.int32)
    gl = None if glim is None else np.ascontiguousarray(glim, np.float64)
    buf = C.create_string_buffer(512)
    rc = dropin.dropin_session(ol._addr(d.receptor), ol._addr(d.params), C.addressof(ligs), len(d.ligands), lig, pop,
                               genes.ctypes.data, status, None if gl is None else gl.ctypes.data, energies.ctypes.data,
                               None if cnt is None else cnt.ctypes.data, len(energies), T, C.byref(o), buf, 512)
    assert rc == 0, buf.value.decode()
    return arr, full, vcf, w, o, d


def _check_scored(out, ref, tol_rel=1e-9):
    assert np.array_equal(out["status"], ref["status"])
    ok = ref["status"] == 0
    assert np.all(np.abs(out["cf"] - ref["cf"])[ok] <= tol_rel * np.maximum(1.0, np.abs(ref["cf"][ok])))
    cl = ref["status"] == 2
    assert np.all(np.abs(out["cf"] - ref["cf"])[cl] <= 1e-9 * np.abs(ref["cf"][cl]))
    return int(ok.sum())


@pytest.mark.gpu
def test_dropin_session_flexible_side_chains(dropin, oracle):
    """Config 1 (five flexible side chains, rotamer genes): every C++ entry
    against the oracle and against the C ABI path on the same inputs."""
    from flexaidds_amd import hip
    s = synth.make_config1(pop=256)
    lig = s.ligands[0]
    genes = np.ascontiguousarray(s.genes[0])
    pop = len(genes)
    status = b"".join(b"n" if i % 7 == 3 else b"a" for i in range(pop))
    lo, hi = lig.gene_min.astype(np.float64), lig.gene_max.astype(np.float64)
    span = hi - lo
    glim = np.stack([lo + 0.1 * span, hi - 0.1 * span], axis=1)  # narrower than the ligand's: clamps bite
    rng = np.random.default_rng(3)
    energies = rng.normal(-25.0, 6.0, 3000)
    counts = rng.integers(1, 5, 3000)
    arr, full, vcf, w, o, d = _session(dropin, s, 0, genes, status, glim, energies, counts)

    with ol.gpu_mirror(oracle):
        ref = ol.oracle_eval(oracle, d, 0, genes)
        clamped = np.clip(genes, glim[:, 0], glim[:, 1])
        ref_c = ol.oracle_eval(oracle, d, 0, clamped)
    assert _check_scored(full, ref) >= 50
    # the three-array form carries the same terms
    for f in ("com", "wal", "sas"):
        assert np.array_equal(arr[f], full[f]), f
    # the C ABI path gives the same bits
    with hip.Context() as c:
        c.upload_system(s)
        capi = c.eval(0, genes)
        capi_c = c.eval(0, clamped)
    assert capi.tobytes() == full.tobytes()
    # voronoi_cf_hip::batch_eval: 'n' rows zero, the others scored on clamped genes
    skip = np.frombuffer(status, np.uint8) == ord("n")
    assert np.all(vcf[skip]["cf"] == 0.0) and np.all(vcf[skip]["status"] == 0) and np.all(arr["vapp"][skip] == 0.0)
    assert vcf[~skip].tobytes() == capi_c[~skip].tobytes()
    assert np.array_equal(arr["vapp"][~skip], capi_c["app"][~skip])
    assert _check_scored(vcf[~skip], ref_c[~skip]) >= 40
    assert not np.array_equal(clamped, genes)
    assert o.wall_ms > 0.0 and o.available == 1
    # rocm_dispatch against the oracle's statmech restatement
    th = abi.Thermo()
    cnt = np.ascontiguousarray(counts, np.int32)
    oracle.fa_oracle_thermo(np.ascontiguousarray(energies).ctypes.data, cnt.ctypes.data, len(energies), 300.0,
                            C.byref(th))
    for f, _ in abi.Thermo._fields_:
        assert abs(getattr(o.thermo, f) - getattr(th, f)) <= 1e-9 * abs(getattr(th, f)) + 1e-12, f
    e = np.ascontiguousarray(energies)
    ref_lse = oracle.fa_oracle_log_sum_exp(e.ctypes.data, len(e))
    assert abs(o.lse - ref_lse) <= 1e-12 * abs(ref_lse)
    ref_w = np.zeros_like(e)
    lz, em = C.c_double(), C.c_double()
    oracle.fa_oracle_boltzmann_batch(e.ctypes.data, len(e), 1.0 / (KB * 300.0), ref_w.ctypes.data, C.byref(lz),
                                     C.byref(em))
    assert np.all(np.abs(w - ref_w) <= 1e-12 * ref_w + 1e-300)
    assert abs(o.log_z - lz.value) <= 1e-12 * abs(lz.value)
    assert o.e_min == em.value == e.min()


@pytest.mark.gpu
def test_dropin_session_second_ligand(dropin, oracle):
    """A multi-ligand upload scored through the C++ layer at ligand index 1,
    without gene limits or a status column."""
    s = synth.make_config2(n_ligands=3, pop=96)
    genes = np.ascontiguousarray(s.genes[1])
    _arr, full, vcf, _w, _o, d = _session(dropin, s, 1, genes, None, None, np.array([-3.0, -1.0]), None)
    with ol.gpu_mirror(oracle):
        ref = ol.oracle_eval(oracle, d, 1, genes)
    assert _check_scored(full, ref) >= 10
    assert vcf.tobytes() == full.tobytes()
