"""HIP library against the CPU oracle (GPU tests).

Tolerances (per pose, scored poses): |CF_gpu - CF_oracle| <= max(1e-4 |CF|,
1e-3 kcal/mol); F at 300 K within 1e-6 relative on identical energies;
contact-pair sets identical.  The pose build's cos/sin are correctly rounded
on the GPU; the oracle's trig mode 1 computes the same bits, so against it
coordinates are bit-exact and contact sets identical.  Against the host-libm
oracle (mode 0, the reference binary's own bits) a pose may only differ where
the oracle itself differs between the two trig modes (a documented
discontinuity of the contact function, not a kernel error).
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle_lib as ol
from flexaidds_amd import abi, synth

pytestmark = pytest.mark.gpu


def tol(cf):
    return np.maximum(1e-4 * np.abs(cf), 1e-3)


@pytest.fixture(scope="module")
def hip():
    from flexaidds_amd import hip as h
    return h


@pytest.fixture(scope="module")
def sys0():
    return synth.make_config0(n_poses=1000)


@pytest.fixture(scope="module")
def ctx0(hip, sys0):
    c = hip.Context()
    c.upload_system(sys0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def desc0(sys0):
    return abi.Descriptors(sys0.receptor, sys0.ligands, sys0.params)


@pytest.fixture(scope="module")
def oracle_cr(oracle, desc0, sys0):
    with ol.gpu_mirror(oracle):
        out, ops = ol.oracle_eval_counted(oracle, desc0, 0, sys0.genes[0])
    return out, ops


@pytest.fixture(scope="module")
def oracle_libm(oracle, desc0, sys0):
    return ol.oracle_eval(oracle, desc0, 0, sys0.genes[0])


def test_device_is_gfx950(hip):
    info = hip.device_info(0)
    print("device", info, flush=True)
    assert info["gcn_arch"] == 950
    assert info["wavefront_size"] == 64


def test_config0_against_oracle(ctx0, sys0, oracle_cr, oracle_libm):
    g = ctx0.eval(0, sys0.genes[0])
    ref, _ = oracle_cr
    diff = np.nonzero(g["status"] != ref["status"])[0]
    assert diff.size == 0, [(int(i), int(g["status"][i]), int(ref["status"][i])) for i in diff[:20]]
    assert ctx0.counters()["capacity"] == 0
    ok = ref["status"] == 0
    assert ok.sum() > 500
    d = np.abs(g["cf"] - ref["cf"])
    assert np.all(d[ok] <= tol(ref["cf"][ok])), np.nonzero(d[ok] > tol(ref["cf"][ok]))
    assert d[ok].max() < 1e-9                     # same arithmetic: rounding-level only
    for f in ("com", "wal", "sas", "totsas"):
        assert np.abs(g[f] - ref[f])[ok].max() < 1e-8, f
    assert np.array_equal(g["rclash"], ref["rclash"])
    cl = ref["status"] == 2
    assert np.all(np.abs(g["cf"] - ref["cf"])[cl] <= 1e-9 * np.abs(ref["cf"][cl]))
    # against the host-libm oracle: only trig-ulp discontinuities may differ
    lm = oracle_libm
    okl = ok & (lm["status"] == 0)
    bad = okl & (np.abs(g["cf"] - lm["cf"]) > tol(lm["cf"]))
    trig_ties = okl & (np.abs(ref["cf"] - lm["cf"]) > tol(lm["cf"]))
    assert np.all(bad <= trig_ties), np.nonzero(bad & ~trig_ties)
    assert bad.sum() <= max(2, okl.sum() // 200)


@pytest.mark.parametrize("normalize", [0, 1])
def test_density_matrix_against_oracle(hip, sys0, oracle, normalize):
    # density-function pairs (weight 0) for half the type pairs, with and
    # without normalize_area: same tolerance as the weight-1 matrix
    import dataclasses
    off, xy = synth.make_density_emat(np.random.default_rng(17), sys0.params.ntypes, frac=0.5)
    prm = dataclasses.replace(sys0.params, dens_off=off, dens_xy=xy, normalize_area=normalize)
    s = dataclasses.replace(sys0, params=prm)
    genes = sys0.genes[0][:400]
    with hip.Context() as c:
        c.upload_system(s)
        g = c.eval(0, genes)
    desc = abi.Descriptors(s.receptor, s.ligands, prm)
    with ol.gpu_mirror(oracle):
        ref = ol.oracle_eval(oracle, desc, 0, genes)
    assert np.array_equal(g["status"], ref["status"])
    ok = ref["status"] == 0
    assert ok.sum() > 200
    d = np.abs(g["cf"] - ref["cf"])
    assert np.all(d[ok] <= tol(ref["cf"][ok]))
    for f in ("com", "sas"):
        assert np.abs(g[f] - ref[f])[ok].max() < 1e-8, f


def test_constraints_against_oracle(hip, oracle):
    # covalent constraints: con term, shortened clash distance, skipped
    # contact energy; the reference's Gaussian explodes outside its window,
    # so the tolerance is relative for those poses
    s = synth.make_config0(n_poses=600)
    synth.add_constraints(s, n_receptor=3, n_ligand=2)
    genes = s.genes[0]
    with hip.Context() as c:
        c.upload_system(s)
        g = c.eval(0, genes)
    desc = abi.Descriptors(s.receptor, s.ligands, s.params)
    with ol.gpu_mirror(oracle):
        ref = ol.oracle_eval(oracle, desc, 0, genes)
    assert np.array_equal(g["status"], ref["status"])
    ok = ref["status"] == 0
    assert np.sum(ref["con"][ok] < 7000.0 - 1e-6) >= 40
    for f in ("cf", "con", "com", "wal"):
        d = np.abs(g[f] - ref[f])[ok]
        assert np.all(d <= np.maximum(1e-4 * np.abs(ref[f][ok]), 1e-3)), f
    assert np.array_equal(g["rclash"], ref["rclash"])


def test_pose_build_bit_exact(ctx0, sys0, oracle, desc0):
    g = sys0.genes[0][:200]
    _, xyz, _ = ctx0.eval_debug(0, g)
    with ol.gpu_mirror(oracle):
        for p in range(len(g)):
            ref = np.zeros((sys0.ligands[0].n_atoms, 3), np.float32)
            gp = np.ascontiguousarray(g[p])
            oracle.fa_oracle_build_pose(ol._addr(desc0.ligands[0]), gp.ctypes.data, ref.ctypes.data)
            assert np.array_equal(ref, xyz[p]), p


def test_contact_records(ctx0, sys0, oracle, desc0):
    g = sys0.genes[0][:120]
    out, _, recs = ctx0.eval_debug(0, g)
    n_checked = 0
    with ol.gpu_mirror(oracle):
        for p in range(len(g)):
            o, _, orec = ol.oracle_pose(oracle, desc0, 0, g[p])
            assert out[p]["status"] == o["status"]
            if o["status"] != 0:
                continue
            a = {(int(r["from"]), int(r["atom"])): r for r in recs[p]}
            b = {(int(r["from"]), int(r["atom"])): r for r in orec}
            assert a.keys() == b.keys(), p
            for k in a:
                assert abs(a[k]["area"] - b[k]["area"]) <= 1e-9 * max(1.0, abs(b[k]["area"])), (p, k)
                assert a[k]["dist"] == b[k]["dist"]
            # record order per atom: reverse contact order, as the ca_rec lists
            for frm in np.unique(orec["from"]):
                ga = [int(x) for x in recs[p]["atom"][recs[p]["from"] == frm]]
                oa = [int(x) for x in orec["atom"][orec["from"] == frm]]
                assert ga == oa, (p, int(frm))
            n_checked += 1
    assert n_checked > 50


def test_opcount_matches_oracle(ctx0, sys0, oracle_cr):
    ctx0.eval(0, sys0.genes[0])
    got, want = ctx0.opcount(), oracle_cr[1]
    assert got == want, {k: (got[k], want[k]) for k in want if got[k] != want[k]}


def test_repeatable(ctx0, sys0):
    a = ctx0.eval(0, sys0.genes[0][:300])
    b = ctx0.eval(0, sys0.genes[0][:300])
    assert a.tobytes() == b.tobytes()


def test_gene_clamping(ctx0, sys0, oracle, desc0):
    g = sys0.genes[0][:40].copy()
    g[:, 0] = np.where(np.arange(40) % 2, 5000.0, -3.0)   # grid index beyond both limits
    g[:, 1:] += np.where(np.arange(40)[:, None] % 3 == 0, 400.0, 0.0)
    out = ctx0.eval(0, g)
    with ol.gpu_mirror(oracle):
        ref = ol.oracle_eval(oracle, desc0, 0, g)
    assert np.array_equal(out["status"], ref["status"])
    ok = ref["status"] == 0
    assert np.all(np.abs(out["cf"] - ref["cf"])[ok] <= tol(ref["cf"][ok]))


def test_empty_and_invalid(ctx0, sys0, hip):
    out = ctx0.eval(0, np.zeros((0, sys0.ligands[0].n_genes)))
    assert out.shape == (0,)
    g = np.ascontiguousarray(sys0.genes[0][:2])
    res = np.zeros(2, abi.CF_DTYPE)
    assert hip._lib.fa_eval_batch(ctx0._ctx, 5, 2, g.ctypes.data, res.ctypes.data) != 0   # no such ligand
    assert hip._lib.fa_eval_batch(ctx0._ctx, 0, -1, g.ctypes.data, res.ctypes.data) != 0  # negative size
    bad = synth.make_config0(n_poses=4)
    bad.ligands[0].mov = bad.ligands[0].mov[::-1].copy()  # rebuild order violates dependencies
    c = hip.Context()
    with pytest.raises(hip.FAError):
        c.upload_system(bad)
    assert hip._lib.fa_eval_batch(c._ctx, 0, 2, g.ctypes.data, res.ctypes.data) != 0      # nothing uploaded
    c.close()


def test_slow_grid_path(hip, oracle):
    """Poses that place ligand atoms outside the receptor's bounding box widen
    the cell grid per pose (LIB/Vcontacts.cpp:1626-1641)."""
    s = synth.make_config0(n_poses=64)
    lig = s.ligands[0]
    center = s.receptor.center
    far = [center + np.array([dx, 0.0, 0.0]) for dx in np.linspace(26.0, 34.0, 16)]
    ic = np.array([synth.point_ic(p, lig.ori) for p in far], np.float32)
    lig.grid_ic = np.vstack([lig.grid_ic, ic])
    lig.gene_max = lig.gene_max.copy()
    lig.gene_max[0] = lig.n_grid - 1
    g = s.genes[0].copy()
    g[:, 0] = lig.n_grid - 16 + (np.arange(64) % 16)
    d = abi.Descriptors(s.receptor, s.ligands, s.params)
    c = hip.Context()
    c.upload_system(s)
    out = c.eval(0, g)
    n_slow = c.counters()["slow_grid"]
    assert n_slow > 0
    # per-pose cycle counts flag exactly the widened-grid poses
    c.set_phase_profiling(True)
    again = c.eval(0, g)
    cyc, slow = c.pose_cycles(len(g))
    c.set_phase_profiling(False)
    assert np.array_equal(again["cf"], out["cf"])
    assert slow.sum() == n_slow and np.all(cyc > 0)
    with ol.gpu_mirror(oracle):
        ref = ol.oracle_eval(oracle, d, 0, g)
    c.close()
    assert np.array_equal(out["status"], ref["status"])
    ok = ref["status"] == 0
    assert ok.sum() > 5
    assert np.all(np.abs(out["cf"] - ref["cf"])[ok] <= tol(ref["cf"][ok]))


def test_multi_ligand_launch(hip, oracle):
    """Uneven ligands in one launch: each pose equals its single-ligand result."""
    import torch
    s = synth.make_config2(n_ligands=4, pop=48)
    c = hip.Context()
    c.upload_system(s)
    G = max(l.n_genes for l in s.ligands)
    P = sum(len(g) for g in s.genes)
    genes = np.zeros((P, G))
    lig_of = np.zeros(P, np.int32)
    k = 0
    for li, gl in enumerate(s.genes):
        genes[k:k + len(gl), :gl.shape[1]] = gl
        lig_of[k:k + len(gl)] = li
        k += len(gl)
    perm = np.random.default_rng(7).permutation(P)   # interleave ligands
    genes, lig_of = genes[perm], lig_of[perm]
    dg = torch.from_numpy(genes).cuda()
    dl = torch.from_numpy(lig_of).cuda()
    dout = torch.zeros(P * abi.CF_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
    c.eval_device(P, dg.data_ptr(), G, dl.data_ptr(), dout.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    mixed = np.frombuffer(dout.cpu().numpy().tobytes(), abi.CF_DTYPE)
    d = abi.Descriptors(s.receptor, s.ligands, s.params)
    for li, gl in enumerate(s.genes):
        sel = np.nonzero(lig_of == li)[0]
        alone = c.eval(li, genes[sel, :gl.shape[1]])
        assert alone.tobytes() == mixed[sel].tobytes(), li
        with ol.gpu_mirror(oracle):
            ref = ol.oracle_eval(oracle, d, li, genes[sel, :gl.shape[1]])
        assert np.array_equal(alone["status"], ref["status"])
        ok = ref["status"] == 0
        assert np.all(np.abs(alone["cf"] - ref["cf"])[ok] <= tol(ref["cf"][ok]))
    c.close()


def test_thermo_against_oracle(ctx0, oracle_cr, oracle):
    ref, _ = oracle_cr
    e = np.ascontiguousarray(ref["cf"][ref["status"] == 0])
    th = ctx0.thermo(e, 300.0)
    o = abi.Thermo()
    oracle.fa_oracle_thermo(e.ctypes.data, None, len(e), 300.0, C.byref(o))
    for f in ("log_Z", "free_energy", "mean_energy", "entropy"):
        assert abs(th[f] - getattr(o, f)) <= 1e-6 * abs(getattr(o, f)) + 1e-12, f
    # closed forms: equal energies give S = k ln N, multiplicity adds k ln n
    kb = 0.001987206
    th = ctx0.thermo(np.full(1000, -7.77), 300.0)
    assert abs(th["entropy"] - kb * math.log(1000)) < 1e-9
    th = ctx0.thermo(np.array([-5.0]), 300.0, count=np.array([10000]))
    assert abs(th["free_energy"] - (-5.0 - kb * 300.0 * math.log(10000))) < 1e-9
    th = ctx0.thermo(np.array([-10.5, -9.8, -10.2, -11.0, -9.5]), 300.0)
    o2 = abi.Thermo()
    ee = np.array([-10.5, -9.8, -10.2, -11.0, -9.5])
    oracle.fa_oracle_thermo(ee.ctypes.data, None, 5, 300.0, C.byref(o2))
    assert abs(th["free_energy"] - o2.free_energy) < 1e-9


def test_thermo_over_results_on_device(ctx0, sys0, hip, oracle_cr, oracle):
    """The reduction reads the cf column of the result array in place."""
    import torch
    g = torch.from_numpy(np.ascontiguousarray(sys0.genes[0])).cuda()
    P = g.shape[0]
    out = torch.zeros(P * abi.CF_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    ctx0.eval_device(P, g.data_ptr(), g.shape[1], 0, out.data_ptr(), s)
    cf_off = abi.CF_DTYPE.fields["cf"][1]
    th = ctx0.thermo_device(out.data_ptr() + cf_off, abi.CF_DTYPE.itemsize // 8, P, 300.0, stream=s)
    ref, _ = oracle_cr
    e = np.ascontiguousarray(ref["cf"])
    o = abi.Thermo()
    oracle.fa_oracle_thermo(e.ctypes.data, None, len(e), 300.0, C.byref(o))
    assert abs(th["free_energy"] - o.free_energy) <= 1e-6 * abs(o.free_energy)


def _bit_level_parity(out, ref, min_ok):
    """Statuses equal, scored poses within the CF tolerance and in practice
    at rounding level (the same arithmetic), clash walls equal."""
    diff = np.nonzero(out["status"] != ref["status"])[0]
    assert diff.size == 0, [(int(i), int(out["status"][i]), int(ref["status"][i])) for i in diff[:20]]
    ok = ref["status"] == 0
    assert ok.sum() >= min_ok
    d = np.abs(out["cf"] - ref["cf"])
    assert np.all(d[ok] <= tol(ref["cf"][ok]))
    assert np.all(d[ok] < 1e-9)
    cl = ref["status"] == 2
    assert np.all(np.abs(out["cf"] - ref["cf"])[cl] <= 1e-9 * np.abs(ref["cf"][cl]))


def test_config1_against_oracle(hip, oracle):
    """The benchmark's own workload (4,000-atom receptor, 45-atom /
    10-torsion ligand) on 400 of its poses, with the operation counters."""
    s = synth.make_config1(pop=400)
    d = abi.Descriptors(s.receptor, s.ligands, s.params)
    c = hip.Context()
    c.upload_system(s)
    out = c.eval(0, s.genes[0])
    ops = c.opcount()
    caps = c.counters()["capacity"]
    c.close()
    with ol.gpu_mirror(oracle):
        ref, rops = ol.oracle_eval_counted(oracle, d, 0, s.genes[0])
    assert caps == 0
    _bit_level_parity(out, ref, min_ok=100)
    assert ops == rops


def test_config3_receptor_sizes_against_oracle(hip, oracle):
    """Served-stream receptors of 2,000, 6,000 and 10,000 atoms, each with
    ligands of the stream's size distribution, 48 poses per ligand."""
    systems = synth.make_config3(n_receptors=3, ligands_per_receptor=3, n_poses=48)
    for s in systems:
        d = abi.Descriptors(s.receptor, s.ligands, s.params)
        c = hip.Context()
        c.upload_system(s)
        n_ok = 0
        for li, gl in enumerate(s.genes):
            out = c.eval(li, gl)
            assert c.counters()["capacity"] == 0
            with ol.gpu_mirror(oracle):
                ref = ol.oracle_eval(oracle, d, li, gl)
            _bit_level_parity(out, ref, min_ok=0)
            n_ok += int((ref["status"] == 0).sum())
        c.close()
        print("receptor", s.receptor.n_atoms, "scored poses", n_ok, flush=True)
        assert n_ok > 0


def test_config4_globule_against_oracle(hip, oracle):
    """The blind-dock receptor: a 20,000-atom globule (its cell lists far
    beyond one CU's LDS), ligands of 20-60 atoms placed over the whole
    surface, 48 poses for each of three ligands."""
    s, conc = synth.make_config4(n_ligands=3, pop=48)
    assert s.receptor.n_atoms == 20000 and conc.shape == (3,)
    d = abi.Descriptors(s.receptor, s.ligands, s.params)
    c = hip.Context()
    c.upload_system(s)
    n_ok = 0
    for li, gl in enumerate(s.genes):
        out = c.eval(li, gl)
        assert c.counters()["capacity"] == 0
        with ol.gpu_mirror(oracle):
            ref = ol.oracle_eval(oracle, d, li, gl)
        _bit_level_parity(out, ref, min_ok=0)
        n_ok += int((ref["status"] == 0).sum())
    c.close()
    print("scored poses", n_ok, flush=True)
    assert n_ok > 0


def test_coarse_grid_receptor_against_oracle(hip, oracle):
    """A receptor of 60 atoms in a 6 Å ball: fewer than three cells a side, so
    a cell can be named by several of an atom's 27 slots (the coarse paths
    of both the precomputed and the widened grid)."""
    rng = np.random.default_rng(2718)
    center = np.array([3.0, -2.0, 5.0])
    rec = synth.make_receptor(rng, 60, center, r_in=0.0, r_out=6.0, solid=True)
    ori = rec.xyz.astype(np.float64).mean(axis=0).astype(np.float32)
    pts = center[None, :] + rng.uniform(-9.0, 9.0, size=(160, 3))
    lig = synth.make_ligand(rng, 20, 3, ori, center + np.array([7.0, 0.0, 0.0]), pts)
    synth.finalize_frame(lig, rec)
    params = synth.Params(ntypes=40, emat=synth.make_emat(rng))
    genes = synth.random_genes(rng, lig, 160)
    s = synth.System(receptor=rec, ligands=[lig], params=params, genes=[genes], config=0)
    d = abi.Descriptors(s.receptor, s.ligands, s.params)
    with hip.Context() as c:
        c.upload_system(s)
        out = c.eval(0, genes)
        caps = c.counters()["capacity"]
    with ol.gpu_mirror(oracle):
        ref = ol.oracle_eval(oracle, d, 0, genes)
    assert caps == 0
    _bit_level_parity(out, ref, min_ok=5)


def test_largest_ligand_against_oracle(hip, oracle):
    """A ligand at the atom limit (128 atoms, 20 torsions): ranks and atoms
    in both lane slots everywhere in the pose."""
    rng = np.random.default_rng(1618)
    center = np.array([-4.0, 6.0, 1.0])
    rec = synth.make_receptor(rng, 3000, center, r_in=24.0)   # a cavity the ligand fits in
    ori = rec.xyz.astype(np.float64).mean(axis=0).astype(np.float32)
    pts = center[None, :] + rng.uniform(-4.0, 4.0, size=(300, 3))
    lig = synth.make_ligand(rng, 128, 20, ori, center, pts)
    synth.finalize_frame(lig, rec)
    params = synth.Params(ntypes=40, emat=synth.make_emat(rng))
    s = synth.System(receptor=rec, ligands=[lig], params=params, genes=[synth.random_genes(rng, lig, 300)])
    assert s.ligands[0].n_atoms == 128
    d = abi.Descriptors(s.receptor, s.ligands, s.params)
    with hip.Context() as c:
        c.upload_system(s)
        out = c.eval(0, s.genes[0])
        caps = c.counters()["capacity"]
    with ol.gpu_mirror(oracle):
        ref = ol.oracle_eval(oracle, d, 0, s.genes[0])
    assert caps == 0
    _bit_level_parity(out, ref, min_ok=5)


def test_rna_receptor_file_against_oracle(hip, oracle):
    """The reference's RNA target as its own reader chain reads it (real
    geometry, bonded receptor atoms 1.2-1.6 A apart), a generated ligand in
    three of its grooves, 800 poses each."""
    from file_systems import rna_receptor_system
    n_ok = 0
    for groove in range(3):
        s = rna_receptor_system(41 + groove, groove=groove, n_poses=800, spread=3.0)
        d = abi.Descriptors(s.receptor, s.ligands, s.params)
        with hip.Context() as c:
            c.upload_system(s)
            out = c.eval(0, s.genes[0])
            caps = c.counters()["capacity"]
        with ol.gpu_mirror(oracle):
            ref = ol.oracle_eval(oracle, d, 0, s.genes[0])
        assert caps == 0
        _bit_level_parity(out, ref, min_ok=5)
        n_ok += int((ref["status"] == 0).sum())
    assert n_ok >= 60


def test_rna_cleft_grid_against_oracle(hip, oracle):
    """The RNA target with the reference's automatic binding-site grid
    (9,043 cleft points, internal coordinates as calc_cleftic computes them)
    as the translation genes: 1,000 poses over the whole cleft."""
    from file_systems import rna_cleft_system
    s = rna_cleft_system(51)
    d = abi.Descriptors(s.receptor, s.ligands, s.params)
    with hip.Context() as c:
        c.upload_system(s)
        out = c.eval(0, s.genes[0])
        caps = c.counters()["capacity"]
    with ol.gpu_mirror(oracle):
        ref = ol.oracle_eval(oracle, d, 0, s.genes[0])
    assert caps == 0
    _bit_level_parity(out, ref, min_ok=100)
