// ga.hip — one genetic-algorithm generation on the device, graph-capturable.
//
// The reference's generation (GA(), LIB/gaboom.cpp:346-500, with BOOM
// reproduction, :886-1036, and LINEAR fitness, :1625-1633) becomes a fixed
// sequence of launches on one stream:
//
//   tick        generation counter + 1 (the counter seeds the draws)
//   breed       one thread per offspring pair: roulette parents, crossover,
//               mutation, genes -> internal coordinates
//   poses       the contact-function kernel over the offspring
//   ensemble    fold the offspring energies into the FP64 ensemble state
//   rank        parents + offspring ranked by energy (ties by index), the
//               best `population` gathered in order
//   copy back   survivors into the population rows
//   fitness     LINEAR, PSHARE or SMFREE fitness of the survivors
//               (LIB/gaboom.cpp:1627-1747) and its running sums for the
//               next roulette; PSHARE then orders the survivors by fitness
//               as GA() does (LIB/gaboom.cpp:498-500)
//
// A GA holds J independent jobs (ligands) that advance together: rows
// [0, J*pop) are the parents of job 0, 1, ... and rows [J*pop, J*(pop+new))
// their children, so one pose launch scores every job's children (the
// batched form; J = 1 is the plain GA).  Each job keeps its own FP64
// ensemble, which gives its ln Z and, over jobs, the grand partition
// function.  Nothing is allocated or memset inside, so the sequence is
// captured once as a HIP graph and replayed.  Random draws are counter-based (fa_rng.h): a
// generation is a pure function of the population, the seed and the
// generation number.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "fa_ctx.h"
#include "fa_rng.h"
#include "lse.h"

namespace fa {

struct GaDev {
    int jobs, pop, n_new, n_all, Gmax;
    int intragenes;
    double cross, mut, beta;
    uint64_t seed;
    const int* G;          // [jobs] genes of each job
    const double* lmin;    // [jobs * Gmax]
    const double* ldel;
    const double* lbin;
    int32_t* gi;           // [jobs * n_all * Gmax] genes: parents, then children
    double* gic;           // internal coordinates, same rows
    fa_cf_result* res;     // [jobs * n_all]
    int32_t* gi2;          // [jobs * pop * Gmax] survivors being gathered
    double* gic2;
    fa_cf_result* res2;
    int model;             // FA_FIT_LINEAR / FA_FIT_PSHARE / FA_FIT_SMFREE
    double alpha, ew;      // sharing exponent, SMFREE entropy weight
    const double* sig;     // [jobs] sharing radius of each job
    double* fit;           // [jobs * pop] fitness of the parent rows
    double* fit2;          // [jobs * pop] fitness being gathered
    double* share;         // [jobs * pop] niche counts
    double* cum;           // [jobs * pop] running sums of the fitness (roulette)
    LseState* ens;         // [jobs] ensemble of every evaluated pose
    unsigned long long* n_eval;
    unsigned long long* gen;
};

// row of parent r / child i of job k
__device__ __forceinline__ size_t parent_row(const GaDev& g, int k, int r) { return (size_t)k * g.pop + r; }
__device__ __forceinline__ size_t child_row(const GaDev& g, int k, int i)
{
    return (size_t)g.jobs * g.pop + (size_t)k * g.n_new + i;
}
// draw stream of item t of job k (job 0 uses t itself)
__device__ __forceinline__ uint64_t job_stream(int k, int t) { return ((uint64_t)k << 32) | (uint32_t)t; }

// LIB/gaboom.cpp:2887-2906 with RandomDouble(gene) = gene / (2^31)
__host__ __device__ inline double genetoic(double mn, double del, double bin, int32_t gene)
{
    const double u = gene / ((double)FA_MAX_RANDOM_VALUE + 1.0);
    int i = 0;
    double tot = bin;
    while (tot < u) {
        tot += bin;
        ++i;
    }
    return mn + del * (double)i;
}

// LIB/gaboom.cpp:1158-1182: first index whose running fitness exceeds r
__device__ __forceinline__ int roulette(const double* cum, int n, double u)
{
    const double tot = cum[n - 1];
    if (tot <= 0.0) return (int)(u * n) % n;
    const double r = u * tot;
    int lo = 0, hi = n - 1;  // smallest i with cum[i] > r, n - 1 if none
    if (!(cum[hi] > r)) return n - 1;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (cum[mid] > r)
            hi = mid;
        else
            lo = mid + 1;
    }
    return lo;
}

__device__ __forceinline__ unsigned low_mask(int n)
{
    // the reference's mask loop: optr = 1; repeat n times optr |= optr << 1
    unsigned optr = 1u;
    for (int i = 0; i < n; ++i) optr |= (optr << 1);
    return optr;
}

// LIB/gaboom.cpp:2517-2618
__device__ void crossover(int32_t* john, int32_t* mary, int G, int intragenes, uint64_t seed, uint64_t gen,
                          uint64_t stream)
{
    int gen_a = (int)(fa_rng_u01(seed, gen, stream, FA_DRAW_GEN_A) * (double)G);
    int gen_b = (int)(fa_rng_u01(seed, gen, stream, FA_DRAW_GEN_B) * (double)G);
    if (gen_a >= G) gen_a = G - 1;
    if (gen_b >= G) gen_b = G - 1;
    int pnt_a, pnt_b;
    if (intragenes) {
        pnt_a = (int)(fa_rng_u01(seed, gen, stream, FA_DRAW_PNT_A) * 32.0);
        pnt_b = (int)(fa_rng_u01(seed, gen, stream, FA_DRAW_PNT_B) * 32.0);
        if (gen_a > gen_b) {
            const int ag = gen_a, ap = pnt_a;
            gen_a = gen_b;
            pnt_a = pnt_b;
            gen_b = ag;
            pnt_b = ap;
        }
        if (gen_a == gen_b && pnt_a < pnt_b) {
            const int ap = pnt_b;
            pnt_b = pnt_a;
            pnt_a = ap;
        }
    } else {
        if (gen_a > gen_b) {
            const int ag = gen_a;
            gen_a = gen_b;
            gen_b = ag;
        }
        if (gen_a != gen_b) {
            pnt_a = (fa_rng_u01(seed, gen, stream, FA_DRAW_PNT_A) < 0.5) ? 32 : 0;
            pnt_b = (fa_rng_u01(seed, gen, stream, FA_DRAW_PNT_B) < 0.5) ? 32 : 0;
            if ((gen_b - gen_a) == 1 && pnt_a == 0 && pnt_b == 32) pnt_b = pnt_a;
        } else {
            pnt_a = 32;
            pnt_b = 0;
        }
    }
    for (int j = gen_a; j <= gen_b; ++j) {
        const unsigned optr = low_mask((j == gen_a) ? pnt_a : 32);
        const unsigned uj = (unsigned)john[j], um = (unsigned)mary[j];
        john[j] = (int32_t)((uj & ~optr) | (um & optr));
        mary[j] = (int32_t)((uj & optr) | (um & ~optr));
    }
    if (pnt_b > 0) {
        const unsigned optr = low_mask(pnt_b - 1);
        const unsigned uj = (unsigned)john[gen_b], um = (unsigned)mary[gen_b];
        john[gen_b] = (int32_t)((uj & ~optr) | (um & optr));
        mary[gen_b] = (int32_t)((uj & optr) | (um & ~optr));
    }
}

// LIB/gaboom.cpp:2620-2641
__device__ void mutate(int32_t* genes, int G, double rate, uint64_t seed, uint64_t gen, uint64_t stream, int child)
{
    for (int j = 0; j < G; ++j) {
        unsigned optr = 0u;
        for (int i = 0; i < 32; ++i)
            if (fa_rng_u01(seed, gen, stream, FA_DRAW_MUT + ((uint64_t)child * G + j) * 32 + i) < rate) optr |= 1u << i;
        genes[j] ^= (int32_t)optr;
    }
}

__global__ void ga_tick_kernel(unsigned long long* gen)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *gen += 1ull;
}

// initial population, generation 0 (LIB/gaboom.cpp:1871-1902)
__global__ void ga_init_kernel(GaDev g)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)g.jobs * g.pop * g.Gmax) return;
    const int k = (int)(t / ((long long)g.pop * g.Gmax));
    const int c = (int)((t / g.Gmax) % g.pop), j = (int)(t % g.Gmax);
    const size_t at = parent_row(g, k, c) * g.Gmax + j;
    if (j >= g.G[k]) {
        g.gi[at] = 0;
        g.gic[at] = 0.0;
        return;
    }
    const int32_t gene = fa_rng_gene(g.seed, 0ull, job_stream(k, c), (uint64_t)j);
    g.gi[at] = gene;
    g.gic[at] = genetoic(g.lmin[k * g.Gmax + j], g.ldel[k * g.Gmax + j], g.lbin[k * g.Gmax + j], gene);
}

// one thread per offspring pair (LIB/gaboom.cpp:920-973)
__global__ void ga_breed_kernel(GaDev g)
{
    const int pairs = (g.n_new + 1) / 2;
    const long long tt = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (tt >= (long long)g.jobs * pairs) return;
    const int k = (int)(tt / pairs), t = (int)(tt % pairs);
    const uint64_t gen = *g.gen;
    const uint64_t st = job_stream(k, t);
    const int G = g.G[k], Gm = g.Gmax;
    const double* lmin = g.lmin + (size_t)k * Gm;
    const double* ldel = g.ldel + (size_t)k * Gm;
    const double* lbin = g.lbin + (size_t)k * Gm;
    const double* cum = g.cum + (size_t)k * g.pop;
    const int p1 = roulette(cum, g.pop, fa_rng_u01(g.seed, gen, st, FA_DRAW_P1));
    const int p2 = roulette(cum, g.pop, fa_rng_u01(g.seed, gen, st, FA_DRAW_P2));
    int32_t john[FA_MAX_GENES], mary[FA_MAX_GENES];
    for (int j = 0; j < G; ++j) {
        john[j] = g.gi[parent_row(g, k, p1) * Gm + j];
        mary[j] = g.gi[parent_row(g, k, p2) * Gm + j];
    }
    if (fa_rng_u01(g.seed, gen, st, FA_DRAW_CROSS) < g.cross) crossover(john, mary, G, g.intragenes, g.seed, gen, st);
    mutate(john, G, g.mut, g.seed, gen, st, 0);
    mutate(mary, G, g.mut, g.seed, gen, st, 1);
    // genes past the job's own count stay 0 (rows are Gmax wide)
    const size_t c1 = child_row(g, k, 2 * t);
    for (int j = 0; j < Gm; ++j) {
        g.gi[c1 * Gm + j] = (j < G) ? john[j] : 0;
        g.gic[c1 * Gm + j] = (j < G) ? genetoic(lmin[j], ldel[j], lbin[j], john[j]) : 0.0;
    }
    if (2 * t + 1 < g.n_new) {
        for (int j = 0; j < Gm; ++j) {
            g.gi[(c1 + 1) * Gm + j] = (j < G) ? mary[j] : 0;
            g.gic[(c1 + 1) * Gm + j] = (j < G) ? genetoic(lmin[j], ldel[j], lbin[j], mary[j]) : 0.0;
        }
    }
}

// energies of each job's parents (children = 0) or children (children = 1)
// into that job's ensemble state; one block per job
__global__ __launch_bounds__(256) void ga_ensemble_kernel(GaDev g, int children)
{
    __shared__ LseState sh[256];
    const int k = blockIdx.x;
    const int n = children ? g.n_new : g.pop;
    LseState st{0.0, 0.0, 0.0, 0.0};
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const double e = g.res[children ? child_row(g, k, i) : parent_row(g, k, i)].cf;
        lse_push(st, -g.beta * e, e);
    }
    sh[threadIdx.x] = st;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) lse_merge(sh[threadIdx.x], sh[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        LseState a = g.ens[k];
        lse_merge(a, sh[0]);
        g.ens[k] = a;
        atomicAdd(g.n_eval, (unsigned long long)n);
    }
}

// total order on energies (NaN last)
__device__ __forceinline__ uint64_t order_key(double e)
{
    const uint64_t u = (uint64_t)__double_as_longlong(e);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}

// rank of each of a job's first n candidates (its parents, then its
// children) by energy, ties by candidate index; the best `pop` gathered in
// order (QuickSort by evalue, LIB/gaboom.cpp:2674-2726).  Job = blockIdx.y.
__device__ __forceinline__ size_t cand_row(const GaDev& g, int k, int r)
{
    return r < g.pop ? parent_row(g, k, r) : child_row(g, k, r - g.pop);
}

__global__ __launch_bounds__(256) void ga_rank_kernel(GaDev g, int n)
{
    __shared__ uint64_t keys[1024];
    const int k = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t ki = (i < n) ? order_key(g.res[cand_row(g, k, i)].cf) : 0ull;
    int rank = 0;
    for (int base = 0; base < n; base += 1024) {
        __syncthreads();
        for (int q = threadIdx.x; q < 1024 && base + q < n; q += blockDim.x)
            keys[q] = order_key(g.res[cand_row(g, k, base + q)].cf);
        __syncthreads();
        const int lim = min(1024, n - base);
        if (i < n)
            for (int q = 0; q < lim; ++q) {
                const uint64_t kj = keys[q];
                rank += (kj < ki) || (kj == ki && base + q < i);
            }
    }
    if (i < n && rank < g.pop) {
        const size_t src = cand_row(g, k, i), dst = parent_row(g, k, rank);
        for (int j = 0; j < g.Gmax; ++j) {
            g.gi2[dst * g.Gmax + j] = g.gi[src * g.Gmax + j];
            g.gic2[dst * g.Gmax + j] = g.gic[src * g.Gmax + j];
        }
        g.res2[dst] = g.res[src];
    }
}

__global__ void ga_copy_back_kernel(GaDev g)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long rows = (long long)g.jobs * g.pop;
    if (t < rows * g.Gmax) {
        g.gi[t] = g.gi2[t];
        g.gic[t] = g.gic2[t];
    }
    if (t < rows) g.res[t] = g.res2[t];
}

// LINEAR fitness pop - rank and its inclusive prefix sums, one block per
// job (the values are integers, so the sums are exact in any order)
__global__ __launch_bounds__(256) void ga_linear_kernel(GaDev g)
{
    __shared__ double part[256];
    const int k = blockIdx.x;
    double carry = 0.0;
    for (int base = 0; base < g.pop; base += 256) {
        const int r = base + threadIdx.x;
        const double f = (r < g.pop) ? (double)(g.pop - r) : 0.0;
        part[threadIdx.x] = f;
        __syncthreads();
        for (int off = 1; off < 256; off <<= 1) {
            const double v = (threadIdx.x >= (unsigned)off) ? part[threadIdx.x - off] : 0.0;
            __syncthreads();
            part[threadIdx.x] += v;
            __syncthreads();
        }
        if (r < g.pop) {
            g.fit[(size_t)k * g.pop + r] = f;
            g.cum[(size_t)k * g.pop + r] = carry + part[threadIdx.x];
        }
        carry += part[255];
        __syncthreads();
    }
}

// Niche counts (LIB/gaboom.cpp:1649-1660): for survivor i of job k,
// sum over survivors j of 1 - (rmsp / sig)^alpha where rmsp <= sig, with
// rmsp = sqrt(sum of squared internal-coordinate differences / genes)
// (calc_rmsp, LIB/gaboom.cpp:2874-2885).  Survivors j stream through LDS
// in tiles; each thread keeps its own row's sums in the reference's order.
constexpr int SHARE_TILE = 32;

__global__ __launch_bounds__(256) void ga_share_kernel(GaDev g)
{
    __shared__ double tile[SHARE_TILE * FA_MAX_GENES];
    const int k = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int G = g.G[k], Gm = g.Gmax;
    const double sig = g.sig[k];
    const double* own = g.gic + parent_row(g, k, min(i, g.pop - 1)) * Gm;
    double pshare = 0.0;
    for (int base = 0; base < g.pop; base += SHARE_TILE) {
        const int rows = min(SHARE_TILE, g.pop - base);
        __syncthreads();
        for (int q = threadIdx.x; q < rows * G; q += blockDim.x) {
            const int r = q / G, j = q % G;
            tile[r * FA_MAX_GENES + j] = g.gic[parent_row(g, k, base + r) * Gm + j];
        }
        __syncthreads();
        if (i < g.pop)
            for (int r = 0; r < rows; ++r) {
                double ss = 0.0;
                for (int j = 0; j < G; ++j) {
                    const double d = own[j] - tile[r * FA_MAX_GENES + j];
                    ss += d * d;
                }
                const double rmsp = sqrt(ss / (double)G);
                if (rmsp <= sig) pshare += 1.0 - pow(rmsp / sig, g.alpha);
            }
    }
    if (i < g.pop) g.share[(size_t)k * g.pop + i] = pshare;
}

__device__ double block_sum(double v, double* sh)
{
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

__device__ double block_max(double v, double* sh)
{
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// PSHARE / SMFREE fitness of job k's survivors (rows by increasing energy),
// one block per job (LIB/gaboom.cpp:1637-1747).  SMFREE's Boltzmann weights
// follow StatMechEngine::boltzmann_weights over the survivors' energies
// (LIB/statmech.cpp:178-201 with compute_boltzmann_batch's exp(-beta (E -
// Emin)), LIB/UnifiedHardwareDispatch.cpp:608-624), normalised by their
// maximum.
__global__ __launch_bounds__(256) void ga_shared_fitness_kernel(GaDev g)
{
    __shared__ double sh[256];
    const int k = blockIdx.x;
    const int n = g.pop;
    double Z = 1.0, maxw = 1.0, emin = 0.0;
    const bool smfree = g.model == FA_FIT_SMFREE;
    if (smfree) {
        double m = INFINITY;
        for (int i = threadIdx.x; i < n; i += blockDim.x) m = fmin(m, g.res[parent_row(g, k, i)].cf);
        emin = -block_max(-m, sh);
        double z = 0.0;
        for (int i = threadIdx.x; i < n; i += blockDim.x) z += exp(-g.beta * (g.res[parent_row(g, k, i)].cf - emin));
        Z = block_sum(z, sh);
        double mw = 0.0;
        for (int i = threadIdx.x; i < n; i += blockDim.x)
            mw = fmax(mw, exp(-g.beta * (g.res[parent_row(g, k, i)].cf - emin)) / Z);
        maxw = block_max(mw, sh);
        if (maxw <= 0.0) maxw = 1.0;
    }
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const double ps = g.share[(size_t)k * n + i];
        double f;
        if (smfree) {
            const double w = (Z > 0.0) ? exp(-g.beta * (g.res[parent_row(g, k, i)].cf - emin)) / Z : 0.0;
            const double rank = (double)(n - i) / (double)n;
            const double blended = (1.0 - g.ew) * rank + g.ew * (w / maxw);
            f = blended * (double)n / ps;
        } else {
            f = (double)(n - i) / ps;
        }
        g.fit[(size_t)k * n + i] = f;
    }
}

// position of each of job k's survivors when ordered by decreasing fitness
// (QuickSort(chrom, 0, num_chrom - 1, false) in GA(), LIB/gaboom.cpp:498-500;
// ties keep row order), gathered into the second buffers
__global__ __launch_bounds__(256) void ga_fitness_order_kernel(GaDev g)
{
    __shared__ uint64_t keys[1024];
    const int k = blockIdx.y;
    const int n = g.pop;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t ki = (i < n) ? order_key(-g.fit[(size_t)k * n + i]) : 0ull;
    int rank = 0;
    for (int base = 0; base < n; base += 1024) {
        __syncthreads();
        for (int q = threadIdx.x; q < 1024 && base + q < n; q += blockDim.x)
            keys[q] = order_key(-g.fit[(size_t)k * n + base + q]);
        __syncthreads();
        const int lim = min(1024, n - base);
        if (i < n)
            for (int q = 0; q < lim; ++q) {
                const uint64_t kj = keys[q];
                rank += (kj < ki) || (kj == ki && base + q < i);
            }
    }
    if (i < n) {
        const size_t src = parent_row(g, k, i), dst = parent_row(g, k, rank);
        for (int j = 0; j < g.Gmax; ++j) {
            g.gi2[dst * g.Gmax + j] = g.gi[src * g.Gmax + j];
            g.gic2[dst * g.Gmax + j] = g.gic[src * g.Gmax + j];
        }
        g.res2[dst] = g.res[src];
        g.fit2[(size_t)k * n + rank] = g.fit[(size_t)k * n + i];
    }
}

__global__ void ga_fit_back_kernel(GaDev g)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < (long long)g.jobs * g.pop) g.fit[t] = g.fit2[t];
}

// running sums of a job's fitness in row order, one thread per job: the
// sequential order of roullete_wheel's loop (LIB/gaboom.cpp:1158-1182), so
// a draw picks the row the reference's scan would
__global__ void ga_cum_kernel(GaDev g)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= g.jobs) return;
    const double* f = g.fit + (size_t)k * g.pop;
    double* c = g.cum + (size_t)k * g.pop;
    double tot = 0.0;
    for (int i = 0; i < g.pop; ++i) {
        tot += f[i];
        c[i] = tot;
    }
}

// each job's ensemble thermodynamics, out[k] for job k
__global__ void ga_thermo_kernel(GaDev g, double T, fa_thermo* out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < g.jobs) out[k] = thermo_of(g.ens[k], T);
}

}  // namespace fa

struct fa_ga {
    fa_ctx* ctx = nullptr;
    fa_ga_params p{};
    int jobs = 0, Gmax = 0;
    fa::GaDev d{};
    DevBuf gi, gic, res, gi2, gic2, res2, cum, lim, ngenes, ens, cnt, lig, thermo, grand, fit, fit2, share, sig;
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

int ga_blocks(long long n) { return (int)std::max(1LL, (n + 255) / 256); }

// each job's survivors among its first n candidates into its parent rows,
// then the (shared) rank fitness
int ga_select(fa_ga* ga, int n)
{
    fa::GaDev& d = ga->d;
    hipLaunchKernelGGL(fa::ga_rank_kernel, dim3(ga_blocks(n), d.jobs), dim3(256), 0, ga->stream, d, n);
    hipLaunchKernelGGL(fa::ga_copy_back_kernel, dim3(ga_blocks((long long)d.jobs * d.pop * d.Gmax)), dim3(256), 0,
                       ga->stream, d);
    if (d.model == FA_FIT_LINEAR) {
        hipLaunchKernelGGL(fa::ga_linear_kernel, dim3(d.jobs), dim3(256), 0, ga->stream, d);
    } else {
        hipLaunchKernelGGL(fa::ga_share_kernel, dim3(ga_blocks(d.pop), d.jobs), dim3(256), 0, ga->stream, d);
        hipLaunchKernelGGL(fa::ga_shared_fitness_kernel, dim3(d.jobs), dim3(256), 0, ga->stream, d);
        if (d.model == FA_FIT_PSHARE) {
            hipLaunchKernelGGL(fa::ga_fitness_order_kernel, dim3(ga_blocks(d.pop), d.jobs), dim3(256), 0, ga->stream, d);
            hipLaunchKernelGGL(fa::ga_copy_back_kernel, dim3(ga_blocks((long long)d.jobs * d.pop * d.Gmax)), dim3(256),
                               0, ga->stream, d);
            hipLaunchKernelGGL(fa::ga_fit_back_kernel, dim3(ga_blocks((long long)d.jobs * d.pop)), dim3(256), 0,
                               ga->stream, d);
        }
        hipLaunchKernelGGL(fa::ga_cum_kernel, dim3(ga_blocks(d.jobs)), dim3(256), 0, ga->stream, d);
    }
    FA_HIP(hipGetLastError());
    return FA_SUCCESS;
}

// one generation of every job, all on ga->stream (capturable)
int ga_generation(fa_ga* ga)
{
    fa::GaDev& d = ga->d;
    const size_t rows_parents = (size_t)d.jobs * d.pop;
    hipLaunchKernelGGL(fa::ga_tick_kernel, dim3(1), dim3(64), 0, ga->stream, d.gen);
    hipLaunchKernelGGL(fa::ga_breed_kernel, dim3(ga_blocks((long long)d.jobs * ((d.n_new + 1) / 2))), dim3(256), 0,
                       ga->stream, d);
    FA_HIP(hipGetLastError());
    const int* lig = static_cast<const int*>(ga->lig.p) + rows_parents;
    int rc = fa_launch_poses(ga->ctx, d.jobs * d.n_new, d.gic + rows_parents * d.Gmax, d.Gmax, lig,
                             d.res + rows_parents, nullptr, 0, nullptr, nullptr, 0, ga->stream);
    if (rc != FA_SUCCESS) return rc;
    hipLaunchKernelGGL(fa::ga_ensemble_kernel, dim3(d.jobs), dim3(256), 0, ga->stream, d, 1);
    return ga_select(ga, d.n_all);
}

int ga_build(fa_ctx* c, const fa_ga_params* p, int n_jobs, const int* ligands, const fa_gene_lim* lim, fa_ga** out)
{
    if (!c || !p || !lim || !out || !ligands) return fa_fail(FA_ERR_INVALID, "null argument");
    *out = nullptr;
    if (!c->have_system) return fa_fail(FA_ERR_NOSYSTEM, "no system uploaded");
    if (n_jobs < 1) return fa_fail(FA_ERR_INVALID, "at least one job");
    if (p->population < 2 || p->offspring < 1) return fa_fail(FA_ERR_INVALID, "population >= 2 and offspring >= 1");
    if (p->temperature <= 0.0) return fa_fail(FA_ERR_INVALID, "temperature must be positive");
    if (p->fitness < FA_FIT_LINEAR || p->fitness > FA_FIT_SMFREE) return fa_fail(FA_ERR_INVALID, "unknown fitness model");
    if (p->fitness != FA_FIT_LINEAR) {
        if (!(p->share_alpha > 0.0)) return fa_fail(FA_ERR_INVALID, "sharing alpha must be positive");
        if (!(p->sig_share > 0.0) && !(p->share_peaks > 0.0 && p->share_scale > 0.0))
            return fa_fail(FA_ERR_INVALID, "sharing needs sig_share > 0 or peaks and scale > 0");
        if (!(p->entropy_weight >= 0.0 && p->entropy_weight <= 1.0))
            return fa_fail(FA_ERR_INVALID, "entropy weight must lie in [0, 1]");
    }
    const long long pop = p->population, n_new = p->offspring;
    if ((long long)n_jobs * (pop + n_new) > (1LL << 30) || pop + n_new > (1LL << 30))
        return fa_fail(FA_ERR_LIMIT, "jobs * (population + offspring) above 2^30");
    int Gmax = 0;
    std::vector<int> G(n_jobs);
    for (int k = 0; k < n_jobs; ++k) {
        if (ligands[k] < 0 || ligands[k] >= c->n_ligands) return fa_fail(FA_ERR_INVALID, "ligand index out of range");
        G[k] = c->lig_genes[ligands[k]];
        Gmax = std::max(Gmax, G[k]);
    }
    if (Gmax < 1 || Gmax > FA_MAX_GENES) return fa_fail(FA_ERR_LIMIT, "between 1 and FA_MAX_GENES genes per ligand");
    // set_bins, LIB/gaboom.cpp:2291-2306, per job; padded to Gmax.  Niche
    // radius per job: sqrt(mean (max - min)^2) / (2 peaks^(1/genes)) / scale
    // (LIB/gaboom.cpp:300-309) unless given.
    std::vector<double> host(3 * (size_t)n_jobs * Gmax, 0.0), sig(n_jobs, 1.0);
    const size_t plane = (size_t)n_jobs * Gmax;
    for (int k = 0, off = 0; k < n_jobs; off += G[k], ++k) {
        if (p->sig_share > 0.0) {
            sig[k] = p->sig_share;
        } else if (p->fitness != FA_FIT_LINEAR) {
            double ss = 0.0;
            for (int j = 0; j < G[k]; ++j) ss += (lim[off + j].max - lim[off + j].min) * (lim[off + j].max - lim[off + j].min);
            sig[k] = std::sqrt(ss / (double)G[k]) / (2.0 * std::pow(p->share_peaks, 1.0 / (double)G[k]));
            sig[k] /= p->share_scale;
            if (!(sig[k] > 0.0)) return fa_fail(FA_ERR_INVALID, "gene limits give a zero niche radius");
        }
        for (int j = 0; j < G[k]; ++j) {
            const fa_gene_lim& L = lim[off + j];
            if (!(L.del > 0.0) || !(L.max >= L.min)) return fa_fail(FA_ERR_INVALID, "gene limits need del > 0 and max >= min");
            double nbin = (L.max - L.min) / L.del;
            if (nbin - (int)nbin > 0.0) nbin += 1.0;
            if (L.map) nbin += 1.0;
            if (!(nbin >= 1.0) || nbin > 1.0e6) return fa_fail(FA_ERR_LIMIT, "between 1 and 1e6 bins per gene");
            host[(size_t)k * Gmax + j] = L.min;
            host[plane + (size_t)k * Gmax + j] = L.del;
            host[2 * plane + (size_t)k * Gmax + j] = 1.0 / nbin;
        }
    }
    FA_HIP(hipSetDevice(c->device));
    fa_ga* ga = new fa_ga();
    ga->ctx = c;
    ga->p = *p;
    ga->jobs = n_jobs;
    ga->Gmax = Gmax;
    const size_t J = (size_t)n_jobs, n_all = (size_t)(pop + n_new);
    auto fail_free = [&](int code) {
        fa_ga_destroy(ga);
        return code;
    };
    int rc;
    if ((rc = fa_ensure(ga->gi, sizeof(int32_t) * J * n_all * Gmax)) ||
        (rc = fa_ensure(ga->gic, sizeof(double) * J * n_all * Gmax)) ||
        (rc = fa_ensure(ga->res, sizeof(fa_cf_result) * J * n_all)) ||
        (rc = fa_ensure(ga->gi2, sizeof(int32_t) * J * pop * Gmax)) ||
        (rc = fa_ensure(ga->gic2, sizeof(double) * J * pop * Gmax)) ||
        (rc = fa_ensure(ga->res2, sizeof(fa_cf_result) * J * pop)) || (rc = fa_ensure(ga->cum, sizeof(double) * J * pop)) ||
        (rc = fa_ensure(ga->fit, sizeof(double) * J * pop)) || (rc = fa_ensure(ga->fit2, sizeof(double) * J * pop)) ||
        (rc = fa_ensure(ga->share, sizeof(double) * J * pop)) || (rc = fa_ensure(ga->sig, sizeof(double) * J)) ||
        (rc = fa_ensure(ga->lim, sizeof(double) * host.size())) || (rc = fa_ensure(ga->ngenes, sizeof(int) * J)) ||
        (rc = fa_ensure(ga->ens, sizeof(fa::LseState) * J)) || (rc = fa_ensure(ga->cnt, 2 * sizeof(unsigned long long))) ||
        (rc = fa_ensure(ga->lig, sizeof(int) * J * n_all)) || (rc = fa_ensure(ga->thermo, sizeof(fa_thermo) * J)) ||
        (rc = fa_ensure(ga->grand, sizeof(double) * (2 * J + 2))))
        return fail_free(rc);
    // ligand of every row: parents of job 0, 1, ..., then children of job 0, 1, ...
    std::vector<int> lig(J * n_all);
    for (size_t k = 0; k < J; ++k) {
        std::fill(lig.begin() + k * pop, lig.begin() + (k + 1) * pop, ligands[k]);
        std::fill(lig.begin() + J * pop + k * n_new, lig.begin() + J * pop + (k + 1) * n_new, ligands[k]);
    }
    hipError_t e = hipMemcpy(ga->lim.p, host.data(), sizeof(double) * host.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(ga->lig.p, lig.data(), sizeof(int) * lig.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(ga->ngenes.p, G.data(), sizeof(int) * J, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(ga->sig.p, sig.data(), sizeof(double) * J, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(ga->ens.p, 0, sizeof(fa::LseState) * J);
    if (e == hipSuccess) e = hipMemset(ga->cnt.p, 0, 2 * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ga->stream, hipStreamNonBlocking);
    if (e != hipSuccess) return fail_free(fa_fail(FA_ERR_HIP, std::string("GA setup: ") + hipGetErrorString(e)));
    fa::GaDev& d = ga->d;
    d.jobs = n_jobs;
    d.pop = (int)pop;
    d.n_new = (int)n_new;
    d.n_all = (int)n_all;
    d.Gmax = Gmax;
    d.intragenes = p->intragenes ? 1 : 0;
    d.cross = p->cross_rate;
    d.mut = p->mut_rate;
    d.beta = 1.0 / (fa::KB_KCAL * p->temperature);
    d.seed = p->seed;
    d.G = static_cast<const int*>(ga->ngenes.p);
    const double* lp = static_cast<const double*>(ga->lim.p);
    d.lmin = lp;
    d.ldel = lp + plane;
    d.lbin = lp + 2 * plane;
    d.gi = static_cast<int32_t*>(ga->gi.p);
    d.gic = static_cast<double*>(ga->gic.p);
    d.res = static_cast<fa_cf_result*>(ga->res.p);
    d.gi2 = static_cast<int32_t*>(ga->gi2.p);
    d.gic2 = static_cast<double*>(ga->gic2.p);
    d.res2 = static_cast<fa_cf_result*>(ga->res2.p);
    d.cum = static_cast<double*>(ga->cum.p);
    d.model = p->fitness;
    d.alpha = p->share_alpha;
    d.ew = p->entropy_weight;
    d.sig = static_cast<const double*>(ga->sig.p);
    d.fit = static_cast<double*>(ga->fit.p);
    d.fit2 = static_cast<double*>(ga->fit2.p);
    d.share = static_cast<double*>(ga->share.p);
    d.ens = static_cast<fa::LseState*>(ga->ens.p);
    d.n_eval = static_cast<unsigned long long*>(ga->cnt.p);
    d.gen = d.n_eval + 1;
    // initial population of every job: draw, evaluate, fold, rank
    hipLaunchKernelGGL(fa::ga_init_kernel, dim3(ga_blocks((long long)J * pop * Gmax)), dim3(256), 0, ga->stream, d);
    if ((e = hipGetLastError()) != hipSuccess)
        return fail_free(fa_fail(FA_ERR_HIP, std::string("GA init: ") + hipGetErrorString(e)));
    rc = fa_launch_poses(c, (int)(J * pop), d.gic, Gmax, static_cast<const int*>(ga->lig.p), d.res, nullptr, 0, nullptr,
                         nullptr, 0, ga->stream);
    if (rc == FA_SUCCESS) {
        hipLaunchKernelGGL(fa::ga_ensemble_kernel, dim3(n_jobs), dim3(256), 0, ga->stream, d, 0);
        rc = ga_select(ga, (int)pop);
    }
    e = hipStreamSynchronize(ga->stream);
    if (rc != FA_SUCCESS) return fail_free(rc);
    if (e != hipSuccess) return fail_free(fa_fail(FA_ERR_HIP, std::string("GA init: ") + hipGetErrorString(e)));
    *out = ga;
    return FA_SUCCESS;
}

}  // namespace

extern "C" {

int fa_ga_create(fa_ctx* c, const fa_ga_params* p, const fa_gene_lim* lim, fa_ga** out)
{
    if (!p) return fa_fail(FA_ERR_INVALID, "null argument");
    return ga_build(c, p, 1, &p->ligand, lim, out);
}

int fa_ga_create_batch(fa_ctx* c, const fa_ga_params* p, int n_jobs, const int* ligands, const fa_gene_lim* lim,
                       fa_ga** out)
{
    return ga_build(c, p, n_jobs, ligands, lim, out);
}

int fa_ga_shape(fa_ga* ga, int* n_jobs, int* max_genes)
{
    if (!ga) return fa_fail(FA_ERR_INVALID, "null GA");
    if (n_jobs) *n_jobs = ga->jobs;
    if (max_genes) *max_genes = ga->Gmax;
    return FA_SUCCESS;
}

int fa_ga_run(fa_ga* ga, int generations, int use_graph, float* ms)
{
    if (!ga || generations < 0) return fa_fail(FA_ERR_INVALID, "bad GA run");
    FA_HIP(hipSetDevice(ga->ctx->device));
    if (use_graph && !ga->exec) {
        FA_HIP(hipStreamBeginCapture(ga->stream, hipStreamCaptureModeThreadLocal));
        const int rc = ga_generation(ga);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(ga->stream, &g);
        if (rc != FA_SUCCESS) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        FA_HIP(e);
        ga->graph = g;
        FA_HIP(hipGraphInstantiate(&ga->exec, ga->graph, nullptr, nullptr, 0));
    }
    hipEvent_t e0, e1;
    FA_HIP(hipEventCreate(&e0));
    FA_HIP(hipEventCreate(&e1));
    FA_HIP(hipEventRecord(e0, ga->stream));
    int rc = FA_SUCCESS;
    for (int g = 0; g < generations && rc == FA_SUCCESS; ++g) {
        if (use_graph) {
            const hipError_t e = hipGraphLaunch(ga->exec, ga->stream);
            if (e != hipSuccess) rc = fa_fail(FA_ERR_HIP, std::string("hipGraphLaunch: ") + hipGetErrorString(e));
        } else {
            rc = ga_generation(ga);
        }
    }
    FA_HIP(hipEventRecord(e1, ga->stream));
    FA_HIP(hipEventSynchronize(e1));
    float t = 0.0f;
    FA_HIP(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (ms) *ms = t;
    return rc;
}

int fa_ga_population(fa_ga* ga, int32_t* genes, double* ic, fa_cf_result* results)
{
    if (!ga) return fa_fail(FA_ERR_INVALID, "null GA");
    FA_HIP(hipSetDevice(ga->ctx->device));
    FA_HIP(hipStreamSynchronize(ga->stream));
    const size_t rows = (size_t)ga->jobs * ga->d.n_all;
    if (genes) FA_HIP(hipMemcpy(genes, ga->gi.p, sizeof(int32_t) * rows * ga->Gmax, hipMemcpyDeviceToHost));
    if (ic) FA_HIP(hipMemcpy(ic, ga->gic.p, sizeof(double) * rows * ga->Gmax, hipMemcpyDeviceToHost));
    if (results) FA_HIP(hipMemcpy(results, ga->res.p, sizeof(fa_cf_result) * rows, hipMemcpyDeviceToHost));
    return FA_SUCCESS;
}

int fa_ga_thermo(fa_ga* ga, fa_thermo* out, uint64_t* n_poses)
{
    if (!ga || !out) return fa_fail(FA_ERR_INVALID, "null argument");
    FA_HIP(hipSetDevice(ga->ctx->device));
    hipLaunchKernelGGL(fa::ga_thermo_kernel, dim3(ga_blocks(ga->jobs)), dim3(256), 0, ga->stream, ga->d,
                       ga->p.temperature, static_cast<fa_thermo*>(ga->thermo.p));
    FA_HIP(hipGetLastError());
    FA_HIP(hipStreamSynchronize(ga->stream));
    FA_HIP(hipMemcpy(out, ga->thermo.p, sizeof(fa_thermo) * ga->jobs, hipMemcpyDeviceToHost));
    if (n_poses) FA_HIP(hipMemcpy(n_poses, ga->d.n_eval, sizeof(uint64_t), hipMemcpyDeviceToHost));
    return FA_SUCCESS;
}

int fa_ga_grand(fa_ga* ga, const double* conc_M, double* log_xi, double* p_bound, double* p_empty)
{
    if (!ga || !conc_M) return fa_fail(FA_ERR_INVALID, "null argument");
    const int J = ga->jobs;
    std::vector<double> lc(J);
    for (int k = 0; k < J; ++k) {
        if (!(conc_M[k] > 0.0)) return fa_fail(FA_ERR_INVALID, "concentrations must be positive");
        lc[k] = std::log(conc_M[k]);
    }
    FA_HIP(hipSetDevice(ga->ctx->device));
    double* g = static_cast<double*>(ga->grand.p);  // [log c (J) | log Xi, p_empty, p_bound (J)]
    FA_HIP(hipMemcpyAsync(g, lc.data(), sizeof(double) * J, hipMemcpyHostToDevice, ga->stream));
    fa_thermo* th = static_cast<fa_thermo*>(ga->thermo.p);
    hipLaunchKernelGGL(fa::ga_thermo_kernel, dim3(ga_blocks(J)), dim3(256), 0, ga->stream, ga->d, ga->p.temperature, th);
    FA_HIP(hipGetLastError());
    FA_HIP(fa::launch_grand(&th->log_Z, (int)(sizeof(fa_thermo) / sizeof(double)), g, J, g + J, ga->stream));
    std::vector<double> r(J + 2);
    FA_HIP(hipMemcpyAsync(r.data(), g + J, sizeof(double) * (J + 2), hipMemcpyDeviceToHost, ga->stream));
    FA_HIP(hipStreamSynchronize(ga->stream));
    if (log_xi) *log_xi = r[0];
    if (p_empty) *p_empty = r[1];
    if (p_bound) std::copy(r.begin() + 2, r.end(), p_bound);
    return FA_SUCCESS;
}

int fa_ga_generation(fa_ga* ga, uint64_t* generation)
{
    if (!ga || !generation) return fa_fail(FA_ERR_INVALID, "null argument");
    FA_HIP(hipSetDevice(ga->ctx->device));
    FA_HIP(hipStreamSynchronize(ga->stream));
    FA_HIP(hipMemcpy(generation, ga->d.gen, sizeof(uint64_t), hipMemcpyDeviceToHost));
    return FA_SUCCESS;
}

int fa_ga_fitness(fa_ga* ga, double* fitness, double* sig_share)
{
    if (!ga) return fa_fail(FA_ERR_INVALID, "null GA");
    FA_HIP(hipSetDevice(ga->ctx->device));
    FA_HIP(hipStreamSynchronize(ga->stream));
    const size_t rows = (size_t)ga->jobs * ga->d.pop;
    if (fitness) FA_HIP(hipMemcpy(fitness, ga->fit.p, sizeof(double) * rows, hipMemcpyDeviceToHost));
    if (sig_share) FA_HIP(hipMemcpy(sig_share, ga->sig.p, sizeof(double) * ga->jobs, hipMemcpyDeviceToHost));
    return FA_SUCCESS;
}

int fa_ga_destroy(fa_ga* ga)
{
    if (!ga) return FA_SUCCESS;
    (void)hipSetDevice(ga->ctx->device);
    if (ga->stream) (void)hipStreamSynchronize(ga->stream);
    if (ga->exec) (void)hipGraphExecDestroy(ga->exec);
    if (ga->graph) (void)hipGraphDestroy(ga->graph);
    for (DevBuf* b : {&ga->gi, &ga->gic, &ga->res, &ga->gi2, &ga->gic2, &ga->res2, &ga->cum, &ga->lim, &ga->ngenes,
                      &ga->ens, &ga->cnt, &ga->lig, &ga->thermo, &ga->grand, &ga->fit, &ga->fit2, &ga->share,
                      &ga->sig})
        if (b->p) (void)hipFree(b->p);
    if (ga->stream) (void)hipStreamDestroy(ga->stream);
    delete ga;
    return FA_SUCCESS;
}

}  // extern "C"
