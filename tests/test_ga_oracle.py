"""The GA oracle's operators against what their definitions fix: genetoic's
bin arithmetic, crossover as a bit swap, the per-bit mutation rate, roulette
selection in proportion to fitness, and the uniformity of the counter-based
draws (include/fa_rng.h)."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as ol
from flexaidds_amd import synth

MAXR = 2147483647


class Lim(C.Structure):
    _fields_ = [("min", C.c_double), ("max", C.c_double), ("del_", C.c_double), ("map", C.c_int32),
                ("pad", C.c_int32)]


@pytest.fixture(scope="module")
def ga(oracle):
    lib = oracle
    lib.fa_oracle_genetoic.argtypes = [C.c_void_p, C.c_int32]
    lib.fa_oracle_genetoic.restype = C.c_double
    lib.fa_oracle_ga_init.argtypes = [C.c_uint64, C.c_uint64, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                      C.c_void_p]
    lib.fa_oracle_ga_breed.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                       C.c_double, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.fa_oracle_rng_u01.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64]
    lib.fa_oracle_rng_u01.restype = C.c_double
    lib.fa_oracle_roulette.argtypes = [C.c_void_p, C.c_int, C.c_double]
    lib.fa_oracle_roulette.restype = C.c_int
    lib.fa_oracle_crossover.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_uint64, C.c_uint64,
                                        C.c_uint64]
    return lib


def test_genetoic_closed_form(ga):
    # 64 bins of 1/64 add up exactly: i = ceil(64 u) - 1 for u > 0, else 0
    lim = Lim(-10.0, 54.0, 1.0, 0, 0)
    rng = np.random.default_rng(3)
    genes = np.concatenate([[0, 1, 2, MAXR, MAXR - 1, -5, 1 << 25, (1 << 25) + 1],
                            rng.integers(-2**31, 2**31 - 1, 2000)]).astype(np.int64)
    for g in genes:
        u = g / 2.0**31
        i = max(int(np.ceil(64 * u)) - 1, 0) if u > 0 else 0
        assert ga.fa_oracle_genetoic(C.byref(lim), int(g)) == -10.0 + i, g


def test_genetoic_index_gene_covers_every_point(ga):
    lim = Lim(0.0, 99.0, 1.0, 1, 0)   # 100 grid points, map gene: 100 bins
    vals = {ga.fa_oracle_genetoic(C.byref(lim), int(g)) for g in np.linspace(1, MAXR, 20000).astype(np.int64)}
    assert vals == {float(v) for v in range(100)}


def test_crossover_is_a_bit_swap(ga):
    rng = np.random.default_rng(11)
    n_cross = 0
    for t in range(300):
        G = int(rng.integers(1, 20))
        j0 = rng.integers(-2**31, 2**31 - 1, G).astype(np.int32)
        m0 = rng.integers(-2**31, 2**31 - 1, G).astype(np.int32)
        j1, m1 = j0.copy(), m0.copy()
        ga.fa_oracle_crossover(j1.ctypes.data, m1.ctypes.data, G, int(t % 2), 7, 3, t)
        assert np.array_equal(j1 ^ m1, j0 ^ m0)          # every bit position keeps its pair
        assert np.array_equal(j1 & m1, j0 & m0)
        n_cross += int(not np.array_equal(j1, j0))
    assert n_cross > 200


def test_mutation_flips_bits_at_the_rate(ga):
    G, pop, n_new, rate = 12, 4, 4000, 0.03
    lim = (Lim * G)(*[Lim(-180.0, 180.0, 2.0, 0, 0) for _ in range(G)])
    parents = np.full((pop, G), 0x1234567, np.int32)     # identical parents: crossover is a no-op
    child = np.zeros((n_new, G), np.int32)
    ic = np.zeros((n_new, G))
    ga.fa_oracle_ga_breed(5, 0, 1, pop, n_new, G, 0, 0.8, rate, lim, parents.ctypes.data, child.ctypes.data,
                          ic.ctypes.data)
    flips = np.unpackbits((child ^ parents[0]).view(np.uint8)).sum()
    n = n_new * G * 32
    assert abs(flips - rate * n) < 5 * np.sqrt(n * rate * (1 - rate))


def test_roulette_in_proportion_to_fitness(ga):
    pop = 50
    fit = np.array([pop - i for i in range(pop)], np.float64)     # LINEAR fitness
    N = 127500                                                     # = 100 * sum(fit)
    counts = np.zeros(pop, int)
    for k in range(N):
        counts[ga.fa_oracle_roulette(fit.ctypes.data, pop, (k + 0.5) / N)] += 1
    assert np.all(np.abs(counts - N * fit / fit.sum()) <= 1)


def test_draws_are_uniform(ga):
    u = np.array([ga.fa_oracle_rng_u01(1, 2, s, k) for s in range(200) for k in range(200)])
    assert np.all((u >= 0) & (u < 1))
    assert abs(u.mean() - 0.5) < 0.005 and abs(u.var() - 1 / 12) < 0.002
    # neighbouring counters are uncorrelated
    assert abs(np.corrcoef(u[:-1], u[1:])[0, 1]) < 0.02


def test_init_population_uses_the_limits(ga):
    s = synth.make_config0(n_poses=4)
    lig = s.ligands[0]
    lims = synth.ga_limits(lig)
    G = len(lims)
    lim = (Lim * G)(*[Lim(a, b, d, m, 0) for a, b, d, m in lims])
    genes = np.zeros((500, G), np.int32)
    ic = np.zeros((500, G))
    ga.fa_oracle_ga_init(9, 0, 500, G, lim, genes.ctypes.data, ic.ctypes.data)
    assert np.all(genes >= 1)
    lo = np.array([l[0] for l in lims])
    hi = np.array([l[1] for l in lims])
    step = np.array([l[2] for l in lims])
    assert np.all(ic >= lo) and np.all(ic <= hi + step)
    assert np.allclose((ic - lo) / step, np.round((ic - lo) / step))


def test_jobs_draw_from_disjoint_streams(ga):
    # job k's item t draws from stream (k << 32) | t: job 0 is the plain GA,
    # other jobs are independent of it
    G, pop = 6, 300
    lim = (Lim * G)(*[Lim(-180.0, 180.0, 2.0, 0, 0) for _ in range(G)])
    rows = []
    for job in (0, 1, 2):
        genes = np.zeros((pop, G), np.int32)
        ic = np.zeros((pop, G))
        ga.fa_oracle_ga_init(4, job, pop, G, lim, genes.ctypes.data, ic.ctypes.data)
        rows.append(genes)
    assert not np.array_equal(rows[0], rows[1]) and not np.array_equal(rows[1], rows[2])
    # a gene's high bits agree between jobs no more often than chance
    same = np.mean((rows[0] >> 24) == (rows[1] >> 24))
    assert same < 0.02


@pytest.fixture(scope="module")
def gao(oracle):
    return ol.ga_bind(oracle)


def test_sig_share_closed_form(gao):
    # every gene spans R: sqrt(mean R^2) = R
    for G, R, peaks, scale in [(4, 360.0, 5.0, 10.0), (10, 2.0, 3.0, 1.0), (1, 7.0, 5.0, 10.0)]:
        lims = [(-R / 2, R / 2, 1.0, 0)] * G
        assert abs(ol.sig_share(gao, lims, peaks, scale) - R / (2.0 * peaks ** (1.0 / G)) / scale) < 1e-12


def test_pshare_niche_counts(gao):
    pop, G = 40, 5
    rng = np.random.default_rng(2)
    # all rows identical: every niche count is pop
    same = np.tile(rng.normal(size=G), (pop, 1))
    fit = ol.ga_fitness(gao, 1, same, np.arange(pop, dtype=float), sig=1.0)
    assert np.allclose(fit, (pop - np.arange(pop)) / pop, rtol=0, atol=1e-15)
    # rows far apart: niche counts 1, LINEAR fitness
    far = np.arange(pop)[:, None] * 100.0 + np.zeros((1, G))
    fit = ol.ga_fitness(gao, 1, far, np.arange(pop, dtype=float), sig=1.0)
    assert np.array_equal(fit, (pop - np.arange(pop)).astype(float))
    # two rows at rmsp r < sig share 1 - (r/sig)^alpha with each other
    two = np.zeros((2, G))
    two[1] = 0.3                                    # rmsp = 0.3
    for alpha in (1.0, 2.0, 0.5):
        fit = ol.ga_fitness(gao, 1, two, np.array([0.0, 1.0]), sig=1.2, alpha=alpha)
        share = 1.0 + (1.0 - (0.3 / 1.2) ** alpha)
        assert np.allclose(fit, np.array([2.0, 1.0]) / share, rtol=1e-14)


def test_smfree_limits(gao):
    pop, G = 30, 3
    far = np.arange(pop)[:, None] * 100.0 + np.zeros((1, G))
    E = np.sort(np.random.default_rng(8).normal(scale=0.5, size=pop))
    # entropy weight 0: rank only (niche counts 1)
    fit = ol.ga_fitness(gao, 2, far, E, T=300.0, sig=1.0, ew=0.0)
    assert np.allclose(fit, (pop - np.arange(pop)).astype(float), rtol=1e-15)
    # entropy weight 1: pop * Boltzmann factor relative to the best pose
    fit = ol.ga_fitness(gao, 2, far, E, T=300.0, sig=1.0, ew=1.0)
    beta = 1.0 / (0.001987206 * 300.0)
    assert np.allclose(fit, pop * np.exp(-beta * (E - E.min())), rtol=1e-12)


def test_breed_fit_linear_equals_breed(gao):
    G, pop = 6, 50
    lims = [(-180.0, 180.0, 2.0, 0)] * G
    parents, _ = ol.ga_init(gao, 3, pop, lims)
    a, ai = ol.ga_breed(gao, 3, 2, parents, 40, lims)
    b, bi = ol.ga_breed_fit(gao, 3, 2, parents, (pop - np.arange(pop)).astype(float), 40, lims)
    assert np.array_equal(a, b) and np.array_equal(ai, bi)


def test_crossover_swaps_one_contiguous_segment(ga):
    # with john all zeros and mary all ones, john's set bits after crossover
    # are exactly the swapped positions; reading genes in order and each
    # gene's bits from 31 down to 0, they must form a single run (two-point
    # crossover): a mask built in the wrong direction or an off-by-one at a
    # cut leaves a gap or a stray bit
    rng = np.random.default_rng(21)
    runs = 0
    for t in range(400):
        G = int(rng.integers(1, 12))
        john = np.zeros(G, np.int32)
        mary = np.full(G, -1, np.int32)
        ga.fa_oracle_crossover(john.ctypes.data, mary.ctypes.data, G, int(t % 2), 13, 2, t)
        bits = np.unpackbits(john.astype('>u4').view(np.uint8))        # gene order, bit 31 first
        assert np.array_equal(bits ^ 1, np.unpackbits(mary.astype('>u4').view(np.uint8)))
        on = np.nonzero(bits)[0]
        if on.size:
            assert on[-1] - on[0] + 1 == on.size, (t, G, bits)
            runs += 1
    assert runs > 300


# ------------------------------------------------------------ duplicate rejection
def _rounded(ic):
    # the duplicate relation the reference's signature encodes: every gene's
    # internal coordinate as int32(ic + 0.5) (truncation toward zero)
    return tuple(int(np.trunc(v + 0.5)) for v in ic)


def _first_new(cands_ic, rows, seen_tuples):
    # the reference's loop over candidates in draw order, on rounded tuples
    # rather than signatures (equal barring 64-bit collisions)
    seen = set(seen_tuples)
    kept = []
    for a, ic in enumerate(cands_ic):
        t = _rounded(ic)
        if t in seen:
            continue
        seen.add(t)
        kept.append(a)
        if len(kept) == rows:
            break
    return kept


def test_genes_sig_closed_forms(gao):
    C9 = 0x9E3779B9
    M = (1 << 64) - 1
    # one gene: the int's hash plus the constant; rounding is int32(v + 0.5),
    # truncating toward zero (-3.7 -> -3, not -4), and the int's hash
    # sign-extends (-3 -> 2^64 - 3, not 2^32 - 3)
    assert ol.genes_sig(gao, np.array([3.7])) == 4 + C9
    assert ol.genes_sig(gao, np.array([-3.7])) == ((1 << 64) - 3 + C9) & M
    assert ol.genes_sig(gao, np.array([-0.6])) == C9                     # int32(-0.1) = 0
    # two genes: the second term mixes in the first's hash shifted both ways
    h1 = 4 + C9
    assert ol.genes_sig(gao, np.array([3.7, 1.2])) == h1 ^ ((1 + C9 + (h1 << 6) + (h1 >> 2)) & M)
    # order matters; equal rounded vectors collide
    assert ol.genes_sig(gao, np.array([1.2, 3.7])) != ol.genes_sig(gao, np.array([3.7, 1.2]))
    assert ol.genes_sig(gao, np.array([3.6, 1.4, -6.8])) == ol.genes_sig(gao, np.array([3.9, 0.5, -7.4]))


def test_init_dedup_keeps_first_new_draws(gao):
    # candidate a of the initial population is draw stream a; with rejection
    # the kept rows are the first `pop` candidates whose rounded genes are new
    G, pop, R = 4, 40, 6
    lims = [(0.0, 3.0, 1.0, 0)] * G                      # 3 bins a gene: 81 vectors, frequent repeats
    cands, cic = ol.ga_init(gao, 5, pop * R, lims)       # every candidate the R batches can draw
    genes, ic, st = ol.ga_init_dedup(gao, 5, pop, lims, R)
    kept = _first_new(cic, pop, [])
    assert len(kept) == pop and kept[-1] >= pop          # some draws were rejected
    assert np.array_equal(genes, cands[kept]) and np.array_equal(ic, cic[kept])
    assert int(st[0]) == kept[-1] + 1 - pop and int(st[1]) == 0
    # no repeats among the rows; an empty rejection list reproduces the plain draws
    assert len({_rounded(r) for r in ic}) == pop
    wide = [(-180.0, 180.0, 1.0, 0)] * 6
    g0, i0 = ol.ga_init(gao, 5, pop, wide)
    g1, i1, s1 = ol.ga_init_dedup(gao, 5, pop, wide, 1)
    assert np.array_equal(g0, g1) and np.array_equal(i0, i1) and int(s1[0]) == 0


def test_init_dedup_cap_fills_from_last_batch(gao):
    # 2 genes of 2 bins: 4 distinct vectors for 10 rows; after R batches the
    # rest are the last batch's rejected candidates in draw order
    G, pop, R = 2, 10, 3
    lims = [(0.0, 2.0, 1.0, 0)] * G
    cands, cic = ol.ga_init(gao, 8, pop * R, lims)
    genes, ic, st = ol.ga_init_dedup(gao, 8, pop, lims, R)
    kept = _first_new(cic, pop, [])
    assert len(kept) == 4
    last = [a for a in range(pop * (R - 1), pop * R) if a not in kept][:pop - 4]
    assert np.array_equal(genes, cands[kept + last])
    assert int(st[1]) == pop - 4 and int(st[0]) == pop * R - 4


def test_breed_dedup_against_seen(gao):
    # children: candidate a is child a & 1 of pair a >> 1; signatures accepted
    # earlier (the parents here) and earlier candidates are rejected
    G, pop, n_new, R = 4, 30, 25, 8
    lims = [(0.0, 4.0, 1.0, 0)] * G
    parents, pic = ol.ga_init(gao, 2, pop, lims)
    fit = (pop - np.arange(pop)).astype(float)
    P = (n_new + 1) // 2
    cands, cic = ol.ga_breed_fit(gao, 2, 3, parents, fit, 2 * P * R, lims, mut=0.05)
    seen = {ol.genes_sig(gao, r) for r in pic}
    child, ic, st = ol.ga_breed_dedup(gao, 2, 3, parents, fit, n_new, lims, R, seen, mut=0.05)
    kept = _first_new(cic, n_new, [_rounded(r) for r in pic])
    assert len(kept) == n_new and int(st[1]) == 0
    assert np.array_equal(child, cands[kept])
    assert not {_rounded(r) for r in ic} & {_rounded(r) for r in pic}
    assert int(st[0]) == kept[-1] + 1 - n_new > 0


def _snapshot_brute(ic, e):
    # the definition: records by increasing energy (ties in record order),
    # then each kept unless every gene is within 0.1 of the last kept one
    order = sorted(range(len(e)), key=lambda i: (e[i], i))
    kept = []
    for i in order:
        if kept and all(abs(ic[i][g] - ic[kept[-1]][g]) < 0.1 for g in range(len(ic[i]))):
            continue
        kept.append(i)
    return kept


def test_snapshot_oracle_by_definition(oracle):
    rng = np.random.default_rng(3)
    base = rng.uniform(-10, 10, (40, 6))
    rows, e = [], []
    for k in range(300):
        b = base[rng.integers(40)]
        kind = k % 4
        if kind == 0:
            r = b.copy()                                     # an exact repeat (a survivor recorded again)
        elif kind == 1:
            r = b + rng.uniform(-0.09, 0.09, 6)              # within 0.1 of it
        elif kind == 2:
            r = b.copy()
            r[rng.integers(6)] += rng.choice([-1, 1]) * rng.uniform(0.1, 0.3)   # one gene apart
        else:
            r = rng.uniform(-10, 10, 6)
        rows.append(r)
        e.append(round(float(np.sum(r)) * 0.3, 1 if k % 5 == 0 else 9))   # some equal energies
    ic, e = np.array(rows), np.array(e)
    got = ol.snapshot(oracle, ic, e).tolist()
    assert got == _snapshot_brute(ic, e)
    assert len(got) < len(e)
    assert np.all(np.diff(e[got]) >= 0)


def test_snapshot_oracle_chain(oracle):
    # not transitive: b is within 0.1 of a and dropped, c is within 0.1 of b
    # but 0.16 from a (the last kept), so c is kept
    ic = np.array([[0.0, 0.0], [0.08, 0.0], [0.16, 0.0], [0.17, 0.0]])
    e = np.array([-3.0, -2.0, -1.0, 0.0])
    assert ol.snapshot(oracle, ic, e).tolist() == [0, 2]
    assert ol.snapshot(oracle, ic[:1], e[:1]).tolist() == [0]
    assert ol.snapshot(oracle, ic[:0], e[:0]).tolist() == []
