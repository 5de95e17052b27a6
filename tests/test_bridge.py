"""The reference-side bridge (include/bridge/flexaids_rocm_bridge.h) on the
reference's own structures, as ref_driver.cpp builds them from descriptors
the way LIB/top.cpp sets a run up (oracle/_ref/libfa_bridge.so).

CPU: descriptors → reference structures (FA_Global, atoms with rotamer
copies, residues with fatm/latm and bonded matrices, optres, map_par, mov,
constraints, energy matrix with density lists) → system_from_reference →
descriptors must give back every field, for rigid and flexible receptors,
several ligands, covalent constraints and density-function matrices.

GPU: flexaids_rocm::batch_eval with voronoi_cf::batch_eval's signature
(LIB/VoronoiCFBatch.h:184-247) over chromosomes, against the oracle: 'n'
rows untouched, cfstr fields per chromosome, clash and hull-failure records
as ic2cf returns them (wal = penalty, metal_coord = 1, LIB/ic2cf.cpp:128-131),
app_evalue = get_apparent_cf_evalue; and the flexaids:: dispatch entries."""
import ctypes as C
import dataclasses
import os

import numpy as np
import pytest

import oracle_lib as ol
from flexaidds_amd import abi, synth

SO = os.path.join(ol.ROOT, "oracle", "_ref", "libfa_bridge.so")
KB = 0.001987206


@pytest.fixture(scope="module")
def bridge():
    if not os.path.exists(SO):
        pytest.skip("oracle/_ref/libfa_bridge.so is not built (needs the reference sources: make -C oracle/ref)")
    # one HIP runtime per process: torch's, loaded first with the library (as
    # flexaidds_amd.hip does), before this object pulls the library in
    from flexaidds_amd import hip  # noqa: F401
    lib = C.CDLL(SO)
    lib.fa_bridge_compare.argtypes = [C.c_void_p] * 6 + [C.c_char_p, C.c_int]
    lib.fa_bridge_compare.restype = C.c_int
    lib.fa_bridge_batch_eval.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_char_p,
                                         C.c_void_p, C.c_void_p, C.c_void_p, C.c_char_p, C.c_int]
    lib.fa_bridge_batch_eval.restype = C.c_int
    lib.fa_bridge_dispatch.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.c_void_p, C.c_char_p, C.c_int]
    lib.fa_bridge_dispatch.restype = C.c_int
    return lib


def _compare(bridge, da, db, li=0):
    buf = C.create_string_buffer(256)
    rc = bridge.fa_bridge_compare(ol._addr(da.receptor), ol._addr(da.ligands[li]), ol._addr(da.params),
                                  ol._addr(db.receptor), ol._addr(db.ligands[li]), ol._addr(db.params), buf, 256)
    return rc, buf.value.decode()


_SYSTEMS = {}


def _system(name):
    if name not in _SYSTEMS:
        if name == "config0":
            s = synth.make_config0(n_poses=4)
        elif name == "config1_flex":
            s = synth.make_config1(pop=4)
        elif name == "config2":
            s = synth.make_config2(n_ligands=3, pop=4)
        elif name == "constraints":
            s = synth.make_config0(n_poses=4)
            synth.add_constraints(s, n_receptor=3, n_ligand=2)
        else:  # flexible side chains with a density-function matrix and normalised areas
            s = synth.make_config1(pop=4, n_flex=3, n_rot=4)
            off, xy = synth.make_density_emat(np.random.default_rng(5), s.params.ntypes, frac=0.5)
            s.params = dataclasses.replace(s.params, dens_off=off, dens_xy=xy, normalize_area=1)
        _SYSTEMS[name] = s
    return _SYSTEMS[name]


@pytest.mark.parametrize("name,li", [("config0", 0), ("config1_flex", 0), ("config2", 0), ("config2", 1),
                                     ("config2", 2), ("constraints", 0), ("flex_density", 0)])
def test_bridge_roundtrip(bridge, name, li):
    s = _system(name)
    d = abi.Descriptors(s.receptor, s.ligands, s.params)
    rc, msg = _compare(bridge, d, d, li)
    assert rc == 0, msg


@pytest.mark.parametrize("change,field", [("rot_xyz", "flex.rot_xyz [7]"), ("scorable", "flex.scorable [2]"),
                                          ("gene_atm", "ligand.gene_atm"), ("cons", "ligand.cons [1]")])
def test_bridge_compare_reports_a_changed_field(bridge, change, field):
    """The comparison itself: structures built from one system, compared
    with a copy that differs in one field, name that field."""
    s = synth.make_config1(pop=4, n_flex=2, n_rot=3)
    synth.add_constraints(s, n_receptor=2, n_ligand=1)
    rec, lig = s.receptor, s.ligands[0]
    if change in ("rot_xyz", "scorable"):
        f = rec.flex[1]
        a = getattr(f, change).copy()
        if change == "rot_xyz":
            a.reshape(-1)[7] += 0.25
        else:
            a.reshape(-1)[2] ^= 1
        rec = dataclasses.replace(rec, flex=[rec.flex[0], dataclasses.replace(f, **{change: a})])
    elif change == "gene_atm":
        ga = lig.gene_atm.copy()
        k = int(np.flatnonzero(lig.gene_typ == abi.FA_GENE_ROT)[0])
        ga[k] = 1 - ga[k]                                  # the other side chain
        lig = dataclasses.replace(lig, gene_atm=ga)
    else:
        cons = list(lig.cons)
        a, p, pl, t, b, m = cons[1]
        cons[1] = (a, p, pl, t, b + 0.5, m)
        lig = dataclasses.replace(lig, cons=cons)
    da = abi.Descriptors(s.receptor, s.ligands, s.params)
    db = abi.Descriptors(rec, [lig], s.params)
    rc, msg = _compare(bridge, da, db)
    assert rc == 1 and msg.startswith(field), msg


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["config1_flex", "constraints"])
def test_bridge_batch_eval_exact_signature(bridge, oracle, name):
    if name == "config1_flex":
        s = synth.make_config1(pop=192)
    else:  # the constrained system of the strict-reference constraint test
        s = synth.make_config0(n_poses=400)
        synth.add_constraints(s, n_receptor=3, n_ligand=2)
    d = abi.Descriptors(s.receptor, s.ligands, s.params)
    genes = np.ascontiguousarray(s.genes[0][:300], np.float64)
    n = len(genes)
    status = b"".join(b"n" if i % 5 == 2 else b"o" for i in range(n))
    cf = np.zeros((n, 11))
    app = np.zeros(n)
    wall = C.c_double()
    buf = C.create_string_buffer(512)
    rc = bridge.fa_bridge_batch_eval(ol._addr(d.receptor), ol._addr(d.ligands[0]), ol._addr(d.params), n,
                                     genes.ctypes.data, status, cf.ctypes.data, app.ctypes.data, C.byref(wall), buf,
                                     512)
    assert rc == 0, buf.value.decode()
    with ol.gpu_mirror(oracle):
        ref = ol.oracle_eval(oracle, d, 0, genes)
    skip = np.frombuffer(status, np.uint8) == ord("n")
    assert np.all(cf[skip] == 0.0) and np.all(app[skip] == 0.0)
    com, con, wal, sas, elec, gist, hb, gd, metal, totsas, rclash = cf.T
    keep = ~skip
    st = ref["status"]
    ok = keep & (st == 0)
    assert ok.sum() >= 30
    tol = 1e-9 * np.maximum(1.0, np.abs(ref["cf"]))
    total = com + con + wal + sas + elec + hb + gd + metal       # get_cf_evalue
    assert np.all(np.abs(total - ref["cf"])[keep] <= tol[keep])
    assert np.all(np.abs(app - ref["app"])[keep] <= tol[keep])   # get_apparent_cf_evalue
    for name_, v in (("com", com), ("wal", wal), ("sas", sas), ("con", con), ("totsas", totsas)):
        assert np.all(np.abs(v - ref[name_])[ok] <= 1e-8 * np.maximum(1.0, np.abs(ref[name_][ok]))), name_
    assert np.array_equal(rclash[ok], ref["rclash"][ok])
    assert np.all(elec[keep] == 0.0) and np.all(gist[keep] == 0.0) and np.all(hb[keep] == 0.0)
    # ic2cf's clash record: the wall penalty alone, metal_coord 1
    cl = keep & (st == 2)
    assert cl.sum() >= 10
    assert np.all(metal[cl] == 1.0) and np.all(com[cl] == 0.0) and np.all(sas[cl] == 0.0)
    assert np.all(metal[ok] == 0.0)
    assert wall.value > 0.0
    if name == "constraints":
        assert np.sum(con[ok] < 1000.0 * 7 - 1e-6) >= 5          # constrained pairs in contact


@pytest.mark.gpu
def test_bridge_dispatch(bridge, oracle):
    rng = np.random.default_rng(12)
    e = np.ascontiguousarray(rng.normal(-40.0, 9.0, 4000))
    w = np.zeros_like(e)
    out3 = np.zeros(3)
    backend, tb = C.c_int(), C.c_int()
    buf = C.create_string_buffer(256)
    beta = 1.0 / (KB * 310.0)
    assert bridge.fa_bridge_dispatch(e.ctypes.data, len(e), beta, w.ctypes.data, out3.ctypes.data, C.byref(backend),
                                     C.byref(tb), buf, 256) == 0, buf.value.decode()
    ref_w = np.zeros_like(e)
    lz, em = C.c_double(), C.c_double()
    oracle.fa_oracle_boltzmann_batch(e.ctypes.data, len(e), beta, ref_w.ctypes.data, C.byref(lz), C.byref(em))
    assert np.all(np.abs(w - ref_w) <= 1e-12 * ref_w + 1e-300)
    assert abs(out3[0] - lz.value) <= 1e-12 * abs(lz.value)
    assert out3[1] == em.value
    ref_lse = oracle.fa_oracle_log_sum_exp(e.ctypes.data, len(e))
    assert abs(out3[2] - ref_lse) <= 1e-12 * abs(ref_lse)
    assert backend.value == 6 and tb.value == 6                   # HardwareBackend::ROCM, Backend::ROCM
