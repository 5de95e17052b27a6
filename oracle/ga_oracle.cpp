/*
 * ga_oracle.cpp — CPU restatement of the device GA's operators.
 *
 * TEST INFRASTRUCTURE ONLY (tests/, bench.py's CPU baseline); the shipped
 * library never links it.  The operators follow the reference line by line,
 * with its RandomDouble()/dice() draws taken from include/fa_rng.h in the
 * layout documented there, so a generation can be replayed exactly:
 *
 *   genetoic          LIB/gaboom.cpp:2887-2906
 *   set_bins          LIB/gaboom.cpp:2291-2306
 *   roullete_wheel    LIB/gaboom.cpp:1158-1182 (sequential scan, as written)
 *   crossover         LIB/gaboom.cpp:2517-2618
 *   mutate            LIB/gaboom.cpp:2620-2641
 *   reproduce (BOOM)  LIB/gaboom.cpp:886-973 (fa_oracle_ga_breed_fit)
 *   duplicate rejection  LIB/gaboom.cpp:971-1018 (children) and :1966-1978
 *                     (initial population) with hash_genes, :1040-1045:
 *                     candidates drawn one at a time, kept when their
 *                     signature is new to the job (fa_oracle_ga_*_dedup)
 *   LINEAR fitness    LIB/gaboom.cpp:1627-1635
 *   PSHARE fitness    LIB/gaboom.cpp:1637-1664 (calc_rmsp, :2874-2885)
 *   SMFREE fitness    LIB/gaboom.cpp:1666-1747 (boltzmann_weights,
 *                     LIB/statmech.cpp:178-201; boltzmann_scalar,
 *                     LIB/UnifiedHardwareDispatch.cpp:608-624)
 *   sig_share         LIB/gaboom.cpp:300-309
 *   snapshot          LIB/gaboom.cpp:517-528: the records sorted by evalue
 *                     (a stable sort: records of equal energy keep record
 *                     order, where the reference's QuickSort, :2674-2724,
 *                     leaves an order of its own) and remove_dups
 *                     (:2728-2746) (fa_oracle_snapshot)
 */
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <numeric>
#include <unordered_set>
#include <vector>

#include "fa_rng.h"
#include "flexaidds_amd.h"

namespace {

double genetoic(const fa_gene_lim& g, double bin, int32_t gene)
{
    int i = 0;
    double tot = bin;
    while (tot < gene / ((double)FA_MAX_RANDOM_VALUE + 1.0)) {
        tot += bin;
        i++;
    }
    return g.min + g.del * (double)i;
}

double nbins(const fa_gene_lim& g)
{
    double nbin = (g.max - g.min) / g.del;
    if (nbin - (int)nbin > 0.0) nbin += 1.0;
    if (g.map) nbin += 1.0;
    return nbin;
}

int roulette(const double* fit, int n, double u)
{
    double tot = 0.0;
    for (int i = 0; i < n; i++) tot += fit[i];
    if (tot <= 0.0) return static_cast<int>(u * n) % n;
    const double r = u * tot;
    int i = 0;
    tot = 0.0;
    while (tot <= r && i < n) {
        tot += fit[i];
        i++;
    }
    i--;
    if (i < 0) i = 0;
    return i;
}

void crossover(int32_t* john, int32_t* mary, int num_genes, int intragenes, uint64_t seed, uint64_t gen,
               uint64_t stream)
{
    auto rd = [&](int k) { return fa_rng_u01(seed, gen, stream, k); };
    int gen_a = (int)(rd(FA_DRAW_GEN_A) * (double)num_genes);
    int gen_b = (int)(rd(FA_DRAW_GEN_B) * (double)num_genes);
    if (gen_a >= num_genes) gen_a = num_genes - 1;
    if (gen_b >= num_genes) gen_b = num_genes - 1;
    int pnt_a, pnt_b, aux_gen, aux_pnt;
    if (intragenes) {
        pnt_a = (int)(rd(FA_DRAW_PNT_A) * 32.0);
        pnt_b = (int)(rd(FA_DRAW_PNT_B) * 32.0);
        if (gen_a > gen_b) {
            aux_gen = gen_a;
            aux_pnt = pnt_a;
            gen_a = gen_b;
            pnt_a = pnt_b;
            gen_b = aux_gen;
            pnt_b = aux_pnt;
        }
        if (gen_a == gen_b && pnt_a < pnt_b) {
            aux_pnt = pnt_b;
            pnt_b = pnt_a;
            pnt_a = aux_pnt;
        }
    } else {
        if (gen_a > gen_b) {
            aux_gen = gen_a;
            gen_a = gen_b;
            gen_b = aux_gen;
        }
        if (gen_a != gen_b) {
            pnt_a = (rd(FA_DRAW_PNT_A) < 0.5) ? 32 : 0;
            pnt_b = (rd(FA_DRAW_PNT_B) < 0.5) ? 32 : 0;
            if ((gen_b - gen_a) == 1 && pnt_a == 0 && pnt_b == 32) pnt_b = pnt_a;
        } else {
            pnt_a = 32;
            pnt_b = 0;
        }
    }
    for (int j = gen_a; j <= gen_b; j++) {
        unsigned int optr = 1u;
        aux_pnt = (j == gen_a) ? pnt_a : 32;
        for (int i = 0; i < aux_pnt; i++) optr |= (optr << 1);
        const unsigned int uj = (unsigned int)john[j], um = (unsigned int)mary[j];
        john[j] = (int32_t)((uj & ~optr) | (um & optr));
        mary[j] = (int32_t)((uj & optr) | (um & ~optr));
    }
    if (pnt_b > 0) {
        unsigned int optr = 1u;
        for (int i = 0; i < pnt_b - 1; i++) optr |= (optr << 1);
        const unsigned int uj = (unsigned int)john[gen_b], um = (unsigned int)mary[gen_b];
        john[gen_b] = (int32_t)((uj & ~optr) | (um & optr));
        mary[gen_b] = (int32_t)((uj & optr) | (um & ~optr));
    }
}

void mutate(int32_t* john, int num_genes, double mut_rate, uint64_t seed, uint64_t gen, uint64_t stream, int child)
{
    for (int j = 0; j < num_genes; j++) {
        unsigned int optr = 0u, test = 1u;
        for (int i = 0; i < 32; i++) {
            if (fa_rng_u01(seed, gen, stream, FA_DRAW_MUT + ((uint64_t)child * num_genes + j) * 32 + i) < mut_rate)
                optr |= test;
            test <<= 1;
        }
        john[j] ^= (int32_t)optr;
    }
}

}  // namespace

extern "C" {

double fa_oracle_genetoic(const fa_gene_lim* g, int32_t gene) { return genetoic(*g, 1.0 / nbins(*g), gene); }

/* initial population of job `job` (draw streams (job << 32) | c): genes
 * [pop * G] and their internal coordinates */
void fa_oracle_ga_init(uint64_t seed, uint64_t job, int pop, int G, const fa_gene_lim* lim, int32_t* genes,
                       double* ic)
{
    for (int c = 0; c < pop; ++c)
        for (int j = 0; j < G; ++j) {
            const int32_t g = fa_rng_gene(seed, 0, (job << 32) | (uint64_t)c, (uint64_t)j);
            genes[(std::size_t)c * G + j] = g;
            ic[(std::size_t)c * G + j] = genetoic(lim[j], 1.0 / nbins(lim[j]), g);
        }
}

/* one generation's children of job `job` from its parents with roulette
 * fitness fit[pop]; pair t draws from stream (job << 32) | t */
void fa_oracle_ga_breed_fit(uint64_t seed, uint64_t job, uint64_t gen, int pop, int n_new, int G, int intragenes,
                            double cross_rate, double mut_rate, const fa_gene_lim* lim, const double* fit_in,
                            const int32_t* parents, int32_t* child_genes, double* child_ic)
{
    std::vector<double> fit(fit_in, fit_in + pop);
    std::vector<int32_t> john(G), mary(G);
    for (int t = 0; 2 * t < n_new; ++t) {
        const uint64_t st = (job << 32) | (uint64_t)t;
        const int p1 = roulette(fit.data(), pop, fa_rng_u01(seed, gen, st, FA_DRAW_P1));
        const int p2 = roulette(fit.data(), pop, fa_rng_u01(seed, gen, st, FA_DRAW_P2));
        for (int j = 0; j < G; ++j) {
            john[j] = parents[(std::size_t)p1 * G + j];
            mary[j] = parents[(std::size_t)p2 * G + j];
        }
        if (fa_rng_u01(seed, gen, st, FA_DRAW_CROSS) < cross_rate)
            crossover(john.data(), mary.data(), G, intragenes, seed, gen, st);
        mutate(john.data(), G, mut_rate, seed, gen, st, 0);
        mutate(mary.data(), G, mut_rate, seed, gen, st, 1);
        for (int k = 0; k < 2 && 2 * t + k < n_new; ++k) {
            const int32_t* src = k ? mary.data() : john.data();
            for (int j = 0; j < G; ++j) {
                child_genes[(std::size_t)(2 * t + k) * G + j] = src[j];
                child_ic[(std::size_t)(2 * t + k) * G + j] = genetoic(lim[j], 1.0 / nbins(lim[j]), src[j]);
            }
        }
    }
}

/* hash_genes (LIB/gaboom.cpp:1040-1045) over internal coordinates: the
 * boost-style combine of std::hash<int32_t>(int32(ic + 0.5)), the hash of an
 * int being its value as a size_t */
uint64_t fa_oracle_genes_sig(const double* ic, int G)
{
    std::size_t h = 0;
    for (int i = 0; i < G; ++i)
        h ^= std::hash<int32_t>{}(static_cast<int32_t>(ic[i] + 0.5)) + 0x9e3779b9 + (h << 6) + (h >> 2);
    return (uint64_t)h;
}

}  // extern "C"

namespace {

// The reference's loop: candidates in draw order, each kept when its
// signature is not yet in `seen` (then added), until `rows` are kept.  The
// device draws at most `rounds` batches of `batch` candidates; when they do
// not fill the rows, the rest take the last batch's rejected candidates in
// draw order.  draw(a, genes) makes candidate a.  stats += {rejected, filled}.
template <class Draw>
void dedup_fill(int rows, int batch, int rounds, int G, const fa_gene_lim* lim, const uint64_t* seen_in, int64_t n_seen,
                Draw draw, int32_t* out_genes, double* out_ic, uint64_t* stats)
{
    std::unordered_set<uint64_t> seen(seen_in, seen_in + n_seen);
    std::vector<int32_t> g((std::size_t)batch * rounds * G);
    std::vector<double> ic(g.size());
    std::vector<char> kept((std::size_t)batch * rounds, 0);
    int n = 0;
    long long a = 0;
    const long long total = (long long)batch * rounds;
    for (; a < total && n < rows; ++a) {
        int32_t* ga = &g[(std::size_t)a * G];
        double* ia = &ic[(std::size_t)a * G];
        draw(a, ga);
        for (int j = 0; j < G; ++j) ia[j] = genetoic(lim[j], 1.0 / nbins(lim[j]), ga[j]);
        const uint64_t sig = fa_oracle_genes_sig(ia, G);
        if (seen.count(sig)) continue;
        seen.insert(sig);
        kept[a] = 1;
        std::copy(ga, ga + G, out_genes + (std::size_t)n * G);
        std::copy(ia, ia + G, out_ic + (std::size_t)n * G);
        ++n;
    }
    const int kept_n = n;
    if (stats) stats[0] += (uint64_t)(a - kept_n);
    for (long long b = total - batch; n < rows && b < total; ++b) {   // the last batch's rejected candidates
        if (kept[b]) continue;
        std::copy(&g[(std::size_t)b * G], &g[(std::size_t)b * G] + G, out_genes + (std::size_t)n * G);
        std::copy(&ic[(std::size_t)b * G], &ic[(std::size_t)b * G] + G, out_ic + (std::size_t)n * G);
        ++n;
    }
    if (stats) stats[1] += (uint64_t)(n - kept_n);
}

}  // namespace

extern "C" {

/* initial population with duplicate rejection: candidate a draws from stream
 * (job << 32) | a; batches of pop candidates */
void fa_oracle_ga_init_dedup(uint64_t seed, uint64_t job, int pop, int G, const fa_gene_lim* lim, int rounds,
                             const uint64_t* seen, int64_t n_seen, int32_t* genes, double* ic, uint64_t* stats)
{
    auto draw = [&](long long a, int32_t* out) {
        for (int j = 0; j < G; ++j) out[j] = fa_rng_gene(seed, 0, (job << 32) | (uint64_t)a, (uint64_t)j);
    };
    dedup_fill(pop, pop, rounds, G, lim, seen, n_seen, draw, genes, ic, stats);
}

/* children with duplicate rejection: candidate a is child a & 1 of pair a >> 1
 * (stream (job << 32) | pair), as fa_oracle_ga_breed_fit makes it; batches of
 * 2 * ceil(n_new / 2) candidates */
void fa_oracle_ga_breed_dedup(uint64_t seed, uint64_t job, uint64_t gen, int pop, int n_new, int G, int intragenes,
                              double cross_rate, double mut_rate, const fa_gene_lim* lim, const double* fit_in,
                              const int32_t* parents, int rounds, const uint64_t* seen, int64_t n_seen,
                              int32_t* child_genes, double* child_ic, uint64_t* stats)
{
    std::vector<double> fit(fit_in, fit_in + pop);
    std::vector<int32_t> john(G), mary(G);
    long long pair = -1;
    auto draw = [&](long long a, int32_t* out) {
        if ((a >> 1) != pair) {
            pair = a >> 1;
            const uint64_t st = (job << 32) | (uint64_t)pair;
            const int p1 = roulette(fit.data(), pop, fa_rng_u01(seed, gen, st, FA_DRAW_P1));
            const int p2 = roulette(fit.data(), pop, fa_rng_u01(seed, gen, st, FA_DRAW_P2));
            for (int j = 0; j < G; ++j) {
                john[j] = parents[(std::size_t)p1 * G + j];
                mary[j] = parents[(std::size_t)p2 * G + j];
            }
            if (fa_rng_u01(seed, gen, st, FA_DRAW_CROSS) < cross_rate)
                crossover(john.data(), mary.data(), G, intragenes, seed, gen, st);
            mutate(john.data(), G, mut_rate, seed, gen, st, 0);
            mutate(mary.data(), G, mut_rate, seed, gen, st, 1);
        }
        const int32_t* src = (a & 1) ? mary.data() : john.data();
        std::copy(src, src + G, out);
    };
    dedup_fill(n_new, 2 * ((n_new + 1) / 2), rounds, G, lim, seen, n_seen, draw, child_genes, child_ic, stats);
}

/* as fa_oracle_ga_breed_fit with LINEAR fitness (parents ranked best first) */
void fa_oracle_ga_breed(uint64_t seed, uint64_t job, uint64_t gen, int pop, int n_new, int G, int intragenes,
                        double cross_rate, double mut_rate, const fa_gene_lim* lim, const int32_t* parents,
                        int32_t* child_genes, double* child_ic)
{
    std::vector<double> fit(pop);
    for (int i = 0; i < pop; ++i) fit[i] = (double)(pop - i);
    fa_oracle_ga_breed_fit(seed, job, gen, pop, n_new, G, intragenes, cross_rate, mut_rate, lim, fit.data(), parents,
                           child_genes, child_ic);
}

/* niche radius from the gene limits */
double fa_oracle_sig_share(const fa_gene_lim* lim, int G, double peaks, double scale)
{
    double sig = 0.0;
    for (int i = 0; i < G; i++) sig += (lim[i].max - lim[i].min) * (lim[i].max - lim[i].min);
    sig = std::sqrt(sig / (double)G) / (2.0 * std::pow(peaks, (1.0 / (double)G)));
    return sig / scale;
}

/* fitness of `pop` survivors ordered by increasing energy: model 0 LINEAR,
 * 1 PSHARE, 2 SMFREE (temperature T in K) */
void fa_oracle_ga_fitness(int model, int pop, int G, const double* ic, const double* energy, double T, double sig,
                          double alpha, double ew, double* fit)
{
    auto rmsp = [&](int a, int b) {
        double ss = 0.0;
        for (int j = 0; j < G; ++j) {
            const double d = ic[(std::size_t)a * G + j] - ic[(std::size_t)b * G + j];
            ss += d * d;
        }
        return std::sqrt(ss / (double)G);
    };
    auto niche = [&](int i) {
        double pshare = 0.0;
        for (int j = 0; j < pop; j++) {
            const double r = rmsp(i, j);
            if (r <= sig) pshare += (1.0 - std::pow((r / sig), alpha));
        }
        return pshare;
    };
    if (model == 0) {
        for (int i = 0; i < pop; i++) fit[i] = (double)(pop - i);
        return;
    }
    if (model == 1) {
        for (int i = 0; i < pop; i++) fit[i] = (double)(pop - i) / niche(i);
        return;
    }
    const double kB = 0.001987206;  /* kcal/(mol K), LIB/statmech.h:26 */
    const double beta = 1.0 / (kB * T);
    const double emin = *std::min_element(energy, energy + pop);
    std::vector<double> w(pop);
    for (int i = 0; i < pop; ++i) w[i] = std::exp(-beta * (energy[i] - emin));
    double Z = 0.0;
    for (int i = 0; i < pop; ++i) Z += w[i];
    for (int i = 0; i < pop; ++i) w[i] = (Z > 0.0) ? w[i] / Z : 0.0;
    double max_bw = 0.0;
    for (int i = 0; i < pop; ++i)
        if (w[i] > max_bw) max_bw = w[i];
    if (max_bw <= 0.0) max_bw = 1.0;
    for (int i = 0; i < pop; ++i) {
        const double rank = (double)(pop - i) / (double)pop;
        const double blended = (1.0 - ew) * rank + ew * (w[i] / max_bw);
        fit[i] = blended * (double)pop / niche(i);
    }
}

/* the best `pop` of `n` rows by energy, ties by index (stable order) */
void fa_oracle_ga_survivors(int n, int pop, const double* energy, int32_t* order)
{
    std::vector<int32_t> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int32_t a, int32_t b) { return energy[a] < energy[b]; });
    for (int i = 0; i < pop && i < n; ++i) order[i] = idx[i];
}

/* a single draw, for distribution tests */
double fa_oracle_rng_u01(uint64_t seed, uint64_t gen, uint64_t stream, uint64_t k)
{
    return fa_rng_u01(seed, gen, stream, k);
}

/* one roulette pick over `fit` (distribution tests) */
int fa_oracle_roulette(const double* fit, int n, double u) { return roulette(fit, n, u); }

/* crossover of two chromosomes with the draws of (seed, gen, stream) */
void fa_oracle_crossover(int32_t* john, int32_t* mary, int G, int intragenes, uint64_t seed, uint64_t gen,
                         uint64_t stream)
{
    crossover(john, mary, G, intragenes, seed, gen, stream);
}

/* Kept records of a snapshot (rows of G internal coordinates, energies
 * evalue) in order: order[0..*n_kept) receives record indices. */
int fa_oracle_snapshot(const double* ic, const double* evalue, int n, int G, int* order, int* n_kept)
{
    std::vector<int> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return evalue[a] < evalue[b]; });
    int m = 0;
    for (int j = 0; j < n; ++j) {
        if (m > 0) {
            const double* last = ic + (size_t)order[m - 1] * G;
            const double* cur = ic + (size_t)idx[j] * G;
            int close = 0;
            for (int g = 0; g < G; ++g) close += std::fabs(cur[g] - last[g]) < 0.1;
            if (close == G) continue;
        }
        order[m++] = idx[j];
    }
    *n_kept = m;
    return 0;
}

}  // extern "C"
