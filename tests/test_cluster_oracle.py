"""Binding-mode clustering: the oracle's restatement of cluster()
(LIB/cluster.cpp, the reference's default CLUSTA CF) against the reference's
own cluster() built from its sources, and against what the algorithm fixes.

Pose sets are generated: scoring poses of the config-0 ligand, each a few
degrees away from one of several seed poses, so that they fall into modes of
various sizes; the reference scores them (its energies and its rebuilt
coordinates), they are put in energy order, and both clusterings run on the
same chromosomes.  The reference's output is its .cad file: heads, sizes,
TCF and ACF printed to six decimals, and the RMSD between heads.  The strict
build (no AVX2: scalar distance sums) is compared with the oracle's scalar
mode, the reference's x86 release flags (-O3 -ffast-math -mavx2 -mfma) with
its AVX2 mode.
"""
import numpy as np
import pytest

import oracle_lib as ol
from flexaidds_amd import abi, synth


@pytest.fixture(scope="module")
def system():
    return synth.make_config0(n_poses=8)


def _pose_set(ref, system, n_seeds, per_seed, spread, seed):
    """Scoring poses around n_seeds seed poses, in increasing energy."""
    rng = np.random.default_rng(seed)
    lig = system.ligands[0]
    desc = abi.Descriptors(system.receptor, system.ligands, system.params)
    lo, hi = np.asarray(lig.gene_min, float), np.asarray(lig.gene_max, float)
    angle = np.asarray(lig.gene_typ) != -1
    rows = []
    tries = 0
    while len(rows) < n_seeds and tries < 200:
        tries += 1
        g = synth.random_genes(rng, lig, 1)[0]
        out, _, _ = ol.ref_eval(ref, desc, 0, g[None, :])
        if out["status"][0] != 0:
            continue
        kids = np.repeat(g[None, :], per_seed, axis=0)
        kids[:, angle] += rng.normal(0.0, spread, size=(per_seed, int(angle.sum())))
        kids = np.clip(kids, lo, hi)
        rows.append(kids)
    genes = np.concatenate(rows)
    out, xyz, _ = ol.ref_eval(ref, desc, 0, genes)
    ok = out["status"] == 0
    genes, out, xyz = genes[ok], out[ok], xyz[ok]
    order = np.argsort(out["cf"], kind="stable")
    return desc, genes[order], out[order], xyz[order]


def _same_modes(ora, refc, acf_tol=2e-6, rmsd_rtol=0.0):
    # six printed decimals: 5e-7 from the print, the rest from the sums
    assert len(ora["modes"]) == len(refc["modes"])
    for a, b in zip(ora["modes"], refc["modes"]):
        assert a["top"] == b["top"] and a["freq"] == b["freq"]
        for k in ("tcf", "acf"):
            assert abs(a[k] - b[k]) <= acf_tol + 1e-9 * abs(b[k]), (k, a, b)
    assert np.allclose(ora["top_rmsd"], refc["top_rmsd"], atol=1e-6, rtol=rmsd_rtol)


@pytest.mark.parametrize("fast", [False, True], ids=["strict", "release-flags"])
@pytest.mark.parametrize("T,max_results", [(0, 10), (300, 10), (300, 40), (0, 3)])
def test_equals_reference_cluster(tmp_path, oracle, system, fast, T, max_results):
    ref = ol.load_ref(fast=fast)
    if ref is None:
        pytest.skip("reference build (oracle/_ref) absent")
    desc, genes, out, xyz = _pose_set(ref, system, 8, 12, 6.0, 5)
    assert len(genes) >= 40
    refc = ol.ref_cluster(ref, desc, 0, genes, out["app"], out["cf"], T, 2.0, max_results, tmp_path)
    ora = ol.oracle_cluster(oracle, xyz, out["app"], T=T, rmsd=2.0, max_results=max_results,
                            mode=1 if fast else 0)
    # -ffast-math lets the compiler reassociate the reference's float sum of
    # squared distances between heads (calc_rmsd_chrom): a few ulps
    _same_modes(ora, refc, rmsd_rtol=4e-6 if fast else 0.0)
    if T > 0 and max_results == 40:
        # fewer modes than records: the never-filled zero records take part
        # in the ordering, as in the reference; any of them printed is all zeros
        assert len(ora["modes"]) < 40
        assert all(m["acf"] == 0 and m["top"] == 0 for m in ora["modes"] if m["freq"] == 0)


def test_modes_by_definition(oracle, system):
    """Brute force over the oracle's output: heads are the first unclustered
    pose at their turn, every member lies within the RMSD of its head and
    after it, no pose left for a later head was within the RMSD of an earlier
    head, and sizes add up."""
    ref = ol.load_ref()
    if ref is None:
        pytest.skip("reference build (oracle/_ref) absent")
    _, genes, out, xyz = _pose_set(ref, system, 6, 15, 8.0, 9)
    n = len(genes)
    res = ol.oracle_cluster(oracle, xyz, out["app"], T=0, rmsd=1.5, max_results=n, mode=0)
    member = res["member"]
    assert np.all(member >= 0)                       # max_results = n: everything clustered

    def rmsd(a, b):
        return float(np.sqrt(((xyz[a].astype(np.float64) - xyz[b]) ** 2).sum() / xyz.shape[1]))

    heads = [m["top"] for m in res["modes"]]
    assert heads == sorted(heads) and heads[0] == 0
    for k, h in enumerate(heads):
        assert member[h] == h
        mem = np.flatnonzero(member == h)
        assert len(mem) == res["modes"][k]["freq"] and mem.min() == h
        for i in mem:
            assert rmsd(i, h) <= 1.5 + 1e-5
        for i in np.flatnonzero(member > h):          # clustered later: not within reach of h
            if i > h:
                assert rmsd(i, h) > 1.5 - 1e-5
    assert sum(m["freq"] for m in res["modes"]) == n
    # T = 0: TCF is the head's energy, ACF the sum over the mode
    for m in res["modes"]:
        assert m["tcf"] == out["app"][m["top"]]
        assert abs(m["acf"] - out["app"][member == m["top"]].sum()) <= 1e-9 * max(1.0, abs(m["acf"]))


def test_planted_mistakes_fail(tmp_path, oracle, system):
    """The checks see a dropped sign in the member term, the scalar sum
    where AVX2 order is used, and a sort by TCF instead of ACF."""
    ref = ol.load_ref()
    if ref is None:
        pytest.skip("reference build (oracle/_ref) absent")
    desc, genes, out, xyz = _pose_set(ref, system, 8, 12, 6.0, 5)
    refc = ol.ref_cluster(ref, desc, 0, genes, out["app"], out["cf"], 300, 2.0, 10, tmp_path)
    ora = ol.oracle_cluster(oracle, xyz, out["app"], T=300, rmsd=2.0, max_results=10, mode=0)
    # a member term with -T P ln P (the head's sign) moves ACF
    T = 300.0
    e = out["app"]
    z = np.exp(-e / T).sum()
    member = ora["member"]
    wrong = []
    for m in ora["modes"]:
        if m["freq"] < 2:
            continue
        mem = np.flatnonzero(member == m["top"])
        p = np.exp(-e[mem] / T) / z
        wrong.append(m["tcf"] + sum(p[1:] * e[mem][1:] - T * p[1:] * np.log(p[1:])))
        assert abs(wrong[-1] - m["acf"]) > 1e-5
    assert wrong
    # ordering by TCF instead of ACF changes the record order
    by_tcf = sorted(refc["modes"], key=lambda m: m["tcf"])
    assert [m["top"] for m in by_tcf] != [m["top"] for m in refc["modes"]]


def test_zero_records_sorted_in_like_the_reference(tmp_path, oracle, system):
    # energies shifted up: most modes get a positive ACF, so the records the
    # reference never filled (ACF 0) sort in front of them and are printed
    ref = ol.load_ref()
    if ref is None:
        pytest.skip("reference build (oracle/_ref) absent")
    desc, genes, out, xyz = _pose_set(ref, system, 8, 12, 6.0, 5)
    app = out["app"] + 400.0
    refc = ol.ref_cluster(ref, desc, 0, genes, app, out["cf"] + 400.0, 300, 2.0, 40, tmp_path)
    ora = ol.oracle_cluster(oracle, xyz, app, T=300, rmsd=2.0, max_results=40, mode=0)
    _same_modes(ora, refc)
    zeros = [m for m in ora["modes"] if m["freq"] == 0]
    assert zeros and all(m["acf"] == 0 and m["tcf"] == 0 and m["top"] == 0 for m in zeros)


@pytest.mark.parametrize("case", range(8))
def test_super_cluster_equals_reference(oracle, case):
    # the oracle's FastOPTICS super-cluster against the reference's own
    # (fast_optics.cpp built from its sources) on one-dimensional energy
    # sets: spread, clustered, with repeated values (ties in the reach and
    # in the core distances), the minPts the binding population uses
    ref = ol.load_ref()
    if ref is None:
        pytest.skip("reference build (oracle/_ref) absent")
    rng = np.random.default_rng(100 + case)
    n = [5, 12, 60, 200, 333, 400, 90, 150][case]
    if case % 3 == 0:
        e = rng.normal(-20.0, 6.0, n)
    elif case % 3 == 1:
        e = np.concatenate([rng.normal(c, 0.5, n // 3) for c in (-40.0, -25.0, -5.0)] + [rng.normal(0, 30, n - 3 * (n // 3))])
    else:
        e = np.round(rng.normal(-15.0, 4.0, n), 1)        # many equal energies
    mp = max(4, n // 20)
    got = ol.super_cluster(oracle, e, mp)
    want = ol.super_cluster(ref, e, mp, ref=True)
    assert np.array_equal(got, want)
    assert len(got) >= 1 and len(set(got.tolist())) == len(got)


def test_super_cluster_by_definition(oracle):
    # two far groups: the flood fill never crosses the gap, every member is
    # within 0.8 core distances of a member visited before it, and no
    # outsider is within reach of any member
    rng = np.random.default_rng(7)
    e = np.concatenate([rng.normal(-50.0, 0.3, 40), rng.normal(10.0, 0.3, 25)])
    mp = max(4, len(e) // 20)
    got = ol.super_cluster(oracle, e, mp)
    inside = set(got.tolist())
    assert inside <= set(range(40)) or inside <= set(range(40, 65))
    d = np.abs(e[:, None] - e[None, :])
    cd = np.array([np.sort(np.delete(d[v], v))[mp - 2] for v in range(len(e))])
    for q, v in enumerate(got[1:], 1):
        assert np.any(d[got[:q], v] <= 0.8 * cd[v])
    for v in set(range(len(e))) - inside:
        assert not np.any(d[got, v] <= 0.8 * cd[v])


@pytest.mark.parametrize("mp", [1, 2, 3, 7, 149, 150, 151, 1000])
def test_super_cluster_min_pts_edges(oracle, mp):
    # zero core distances (minPts <= 1), minPts near and above n (infinite
    # core distances: every point opens its own run), against the reference
    ref = ol.load_ref()
    if ref is None:
        pytest.skip("reference build (oracle/_ref) absent")
    rng = np.random.default_rng(mp)
    e = np.round(rng.normal(-12.0, 5.0, 150), 2)
    assert np.array_equal(ol.super_cluster(oracle, e, mp), ol.super_cluster(ref, e, mp, ref=True))


def test_super_cluster_tiny_and_huge_energies(oracle):
    # differences whose square underflows or overflows: the oracle's
    # sqrt(d * d) against the reference's own
    ref = ol.load_ref()
    if ref is None:
        pytest.skip("reference build (oracle/_ref) absent")
    rng = np.random.default_rng(44)
    tiny = rng.uniform(1.0, 2.0, 40) * 1e-300
    huge = rng.uniform(1.0, 2.0, 10) * 1e200
    e = np.concatenate([rng.normal(-20.0, 5.0, 150), tiny, -tiny[:10], huge, [0.0, 0.0]])
    e = e[rng.permutation(len(e))]
    for mp in (max(4, len(e) // 20), 2, 30):
        assert np.array_equal(ol.super_cluster(oracle, e, mp), ol.super_cluster(ref, e, mp, ref=True)), mp
