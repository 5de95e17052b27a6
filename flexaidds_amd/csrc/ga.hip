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
//   fitness     rank fitness and its prefix sums for the next roulette
//
// Nothing is allocated or memset inside, so the sequence is captured once as
// a HIP graph and replayed.  Random draws are counter-based (fa_rng.h): a
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
    int pop, n_new, n_all, G;
    int intragenes;
    double cross, mut, beta;
    uint64_t seed;
    const double* lmin;    // [G]
    const double* ldel;    // [G]
    const double* lbin;    // [G]
    int32_t* gi;           // [n_all * G] genes
    double* gic;           // [n_all * G] internal coordinates
    fa_cf_result* res;     // [n_all]
    int32_t* gi2;          // [pop * G] survivors being gathered
    double* gic2;
    fa_cf_result* res2;
    double* cum;           // [pop] inclusive prefix sums of the fitness
    LseState* ens;         // ensemble of every evaluated pose
    unsigned long long* n_eval;
    unsigned long long* gen;
};

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
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.pop * g.G) return;
    const int c = t / g.G, j = t % g.G;
    const int32_t gene = fa_rng_gene(g.seed, 0ull, (uint64_t)c, (uint64_t)j);
    g.gi[t] = gene;
    g.gic[t] = genetoic(g.lmin[j], g.ldel[j], g.lbin[j], gene);
}

// one thread per offspring pair (LIB/gaboom.cpp:920-973)
__global__ void ga_breed_kernel(GaDev g)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int pairs = (g.n_new + 1) / 2;
    if (t >= pairs) return;
    const uint64_t gen = *g.gen;
    const int G = g.G;
    const int p1 = roulette(g.cum, g.pop, fa_rng_u01(g.seed, gen, t, FA_DRAW_P1));
    const int p2 = roulette(g.cum, g.pop, fa_rng_u01(g.seed, gen, t, FA_DRAW_P2));
    int32_t john[FA_MAX_GENES], mary[FA_MAX_GENES];
    for (int j = 0; j < G; ++j) {
        john[j] = g.gi[(size_t)p1 * G + j];
        mary[j] = g.gi[(size_t)p2 * G + j];
    }
    if (fa_rng_u01(g.seed, gen, t, FA_DRAW_CROSS) < g.cross) crossover(john, mary, G, g.intragenes, g.seed, gen, t);
    mutate(john, G, g.mut, g.seed, gen, t, 0);
    mutate(mary, G, g.mut, g.seed, gen, t, 1);
    const int c1 = g.pop + 2 * t;
    for (int j = 0; j < G; ++j) {
        g.gi[(size_t)c1 * G + j] = john[j];
        g.gic[(size_t)c1 * G + j] = genetoic(g.lmin[j], g.ldel[j], g.lbin[j], john[j]);
    }
    if (2 * t + 1 < g.n_new) {
        for (int j = 0; j < G; ++j) {
            g.gi[(size_t)(c1 + 1) * G + j] = mary[j];
            g.gic[(size_t)(c1 + 1) * G + j] = genetoic(g.lmin[j], g.ldel[j], g.lbin[j], mary[j]);
        }
    }
}

// energies of rows [off, off + n) into the ensemble state (one block)
__global__ __launch_bounds__(256) void ga_ensemble_kernel(GaDev g, int off, int n)
{
    __shared__ LseState sh[256];
    LseState st{0.0, 0.0, 0.0, 0.0};
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        const double e = g.res[off + k].cf;
        lse_push(st, -g.beta * e, e);
    }
    sh[threadIdx.x] = st;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) lse_merge(sh[threadIdx.x], sh[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        LseState a = *g.ens;
        lse_merge(a, sh[0]);
        *g.ens = a;
        *g.n_eval += (unsigned long long)n;
    }
}

// total order on energies (NaN last)
__device__ __forceinline__ uint64_t order_key(double e)
{
    const uint64_t u = (uint64_t)__double_as_longlong(e);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}

// rank of each of rows [0, n) by energy, ties by index; the best `pop`
// rows are gathered in order (QuickSort by evalue, LIB/gaboom.cpp:2674-2726)
__global__ __launch_bounds__(256) void ga_rank_kernel(GaDev g, int n)
{
    __shared__ uint64_t keys[1024];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t ki = (i < n) ? order_key(g.res[i].cf) : 0ull;
    int rank = 0;
    for (int base = 0; base < n; base += 1024) {
        __syncthreads();
        for (int k = threadIdx.x; k < 1024 && base + k < n; k += blockDim.x) keys[k] = order_key(g.res[base + k].cf);
        __syncthreads();
        const int lim = min(1024, n - base);
        if (i < n)
            for (int k = 0; k < lim; ++k) {
                const uint64_t kj = keys[k];
                rank += (kj < ki) || (kj == ki && base + k < i);
            }
    }
    if (i < n && rank < g.pop) {
        for (int j = 0; j < g.G; ++j) {
            g.gi2[(size_t)rank * g.G + j] = g.gi[(size_t)i * g.G + j];
            g.gic2[(size_t)rank * g.G + j] = g.gic[(size_t)i * g.G + j];
        }
        g.res2[rank] = g.res[i];
    }
}

__global__ void ga_copy_back_kernel(GaDev g)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < g.pop * g.G) {
        g.gi[t] = g.gi2[t];
        g.gic[t] = g.gic2[t];
    }
    if (t < g.pop) g.res[t] = g.res2[t];
}

// LINEAR fitness pop - rank and its inclusive prefix sums (one block; the
// values are integers, so the sums are exact in any order)
__global__ __launch_bounds__(256) void ga_fitness_kernel(GaDev g)
{
    __shared__ double part[256];
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
        if (r < g.pop) g.cum[r] = carry + part[threadIdx.x];
        carry += part[255];
        __syncthreads();
    }
}

__global__ void ga_thermo_kernel(GaDev g, double T, fa_thermo* out)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *out = thermo_of(*g.ens, T);
}

}  // namespace fa

struct fa_ga {
    fa_ctx* ctx = nullptr;
    fa_ga_params p{};
    int G = 0;
    fa::GaDev d{};
    DevBuf gi, gic, res, gi2, gic2, res2, cum, lim, ens, cnt, lig, thermo;
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

int ga_blocks(long long n) { return (int)std::max(1LL, (n + 255) / 256); }

// survivors of rows [0, n) into [0, pop), then their fitness
int ga_select(fa_ga* ga, int n)
{
    fa::GaDev& d = ga->d;
    hipLaunchKernelGGL(fa::ga_rank_kernel, dim3(ga_blocks(n)), dim3(256), 0, ga->stream, d, n);
    hipLaunchKernelGGL(fa::ga_copy_back_kernel, dim3(ga_blocks((long long)d.pop * d.G)), dim3(256), 0, ga->stream, d);
    hipLaunchKernelGGL(fa::ga_fitness_kernel, dim3(1), dim3(256), 0, ga->stream, d);
    FA_HIP(hipGetLastError());
    return FA_SUCCESS;
}

// one generation, all on ga->stream (capturable)
int ga_generation(fa_ga* ga)
{
    fa::GaDev& d = ga->d;
    hipLaunchKernelGGL(fa::ga_tick_kernel, dim3(1), dim3(64), 0, ga->stream, d.gen);
    hipLaunchKernelGGL(fa::ga_breed_kernel, dim3(ga_blocks((d.n_new + 1) / 2)), dim3(256), 0, ga->stream, d);
    FA_HIP(hipGetLastError());
    int rc = fa_launch_poses(ga->ctx, d.n_new, d.gic + (size_t)d.pop * d.G, d.G, static_cast<const int*>(ga->lig.p),
                             d.res + d.pop, nullptr, 0, nullptr, nullptr, 0, ga->stream);
    if (rc != FA_SUCCESS) return rc;
    hipLaunchKernelGGL(fa::ga_ensemble_kernel, dim3(1), dim3(256), 0, ga->stream, d, d.pop, d.n_new);
    return ga_select(ga, d.n_all);
}

}  // namespace

extern "C" {

int fa_ga_create(fa_ctx* c, const fa_ga_params* p, const fa_gene_lim* lim, fa_ga** out)
{
    if (!c || !p || !lim || !out) return fa_fail(FA_ERR_INVALID, "null argument");
    *out = nullptr;
    if (!c->have_system) return fa_fail(FA_ERR_NOSYSTEM, "no system uploaded");
    if (p->ligand < 0 || p->ligand >= c->n_ligands) return fa_fail(FA_ERR_INVALID, "ligand index out of range");
    if (p->population < 2 || p->offspring < 1) return fa_fail(FA_ERR_INVALID, "population >= 2 and offspring >= 1");
    if (p->temperature <= 0.0) return fa_fail(FA_ERR_INVALID, "temperature must be positive");
    const int G = c->lig_genes[p->ligand];
    std::vector<double> host(3 * (size_t)G);
    for (int j = 0; j < G; ++j) {
        // set_bins, LIB/gaboom.cpp:2291-2306
        if (!(lim[j].del > 0.0) || !(lim[j].max >= lim[j].min))
            return fa_fail(FA_ERR_INVALID, "gene limits need del > 0 and max >= min");
        double nbin = (lim[j].max - lim[j].min) / lim[j].del;
        if (nbin - (int)nbin > 0.0) nbin += 1.0;
        if (lim[j].map) nbin += 1.0;
        if (!(nbin >= 1.0) || nbin > 1.0e6) return fa_fail(FA_ERR_LIMIT, "between 1 and 1e6 bins per gene");
        host[j] = lim[j].min;
        host[G + j] = lim[j].del;
        host[2 * G + j] = 1.0 / nbin;
    }
    FA_HIP(hipSetDevice(c->device));
    fa_ga* ga = new fa_ga();
    ga->ctx = c;
    ga->p = *p;
    ga->G = G;
    const int pop = p->population, n_new = p->offspring, n_all = pop + n_new;
    int rc;
    auto fail_free = [&](int code) {
        fa_ga_destroy(ga);
        return code;
    };
    if ((rc = fa_ensure(ga->gi, sizeof(int32_t) * (size_t)n_all * G)) ||
        (rc = fa_ensure(ga->gic, sizeof(double) * (size_t)n_all * G)) ||
        (rc = fa_ensure(ga->res, sizeof(fa_cf_result) * (size_t)n_all)) ||
        (rc = fa_ensure(ga->gi2, sizeof(int32_t) * (size_t)pop * G)) ||
        (rc = fa_ensure(ga->gic2, sizeof(double) * (size_t)pop * G)) ||
        (rc = fa_ensure(ga->res2, sizeof(fa_cf_result) * (size_t)pop)) ||
        (rc = fa_ensure(ga->cum, sizeof(double) * (size_t)pop)) || (rc = fa_ensure(ga->lim, sizeof(double) * host.size())) ||
        (rc = fa_ensure(ga->ens, sizeof(fa::LseState))) || (rc = fa_ensure(ga->cnt, 2 * sizeof(unsigned long long))) ||
        (rc = fa_ensure(ga->lig, sizeof(int) * (size_t)n_new)) || (rc = fa_ensure(ga->thermo, sizeof(fa_thermo))))
        return fail_free(rc);
    std::vector<int> lig(n_new, p->ligand);
    hipError_t e = hipMemcpy(ga->lim.p, host.data(), sizeof(double) * host.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(ga->lig.p, lig.data(), sizeof(int) * lig.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(ga->ens.p, 0, sizeof(fa::LseState));
    if (e == hipSuccess) e = hipMemset(ga->cnt.p, 0, 2 * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ga->stream, hipStreamNonBlocking);
    if (e != hipSuccess) return fail_free(fa_fail(FA_ERR_HIP, std::string("GA setup: ") + hipGetErrorString(e)));
    fa::GaDev& d = ga->d;
    d.pop = pop;
    d.n_new = n_new;
    d.n_all = n_all;
    d.G = G;
    d.intragenes = p->intragenes ? 1 : 0;
    d.cross = p->cross_rate;
    d.mut = p->mut_rate;
    d.beta = 1.0 / (fa::KB_KCAL * p->temperature);
    d.seed = p->seed;
    const double* lp = static_cast<const double*>(ga->lim.p);
    d.lmin = lp;
    d.ldel = lp + G;
    d.lbin = lp + 2 * G;
    d.gi = static_cast<int32_t*>(ga->gi.p);
    d.gic = static_cast<double*>(ga->gic.p);
    d.res = static_cast<fa_cf_result*>(ga->res.p);
    d.gi2 = static_cast<int32_t*>(ga->gi2.p);
    d.gic2 = static_cast<double*>(ga->gic2.p);
    d.res2 = static_cast<fa_cf_result*>(ga->res2.p);
    d.cum = static_cast<double*>(ga->cum.p);
    d.ens = static_cast<fa::LseState*>(ga->ens.p);
    d.n_eval = static_cast<unsigned long long*>(ga->cnt.p);
    d.gen = d.n_eval + 1;
    // initial population: draw, evaluate, fold, rank
    hipLaunchKernelGGL(fa::ga_init_kernel, dim3(ga_blocks((long long)pop * G)), dim3(256), 0, ga->stream, d);
    if ((e = hipGetLastError()) != hipSuccess)
        return fail_free(fa_fail(FA_ERR_HIP, std::string("GA init: ") + hipGetErrorString(e)));
    std::vector<int> lig_pop(pop, p->ligand);
    DevBuf lig0;
    if ((rc = fa_ensure(lig0, sizeof(int) * (size_t)pop))) return fail_free(rc);
    e = hipMemcpy(lig0.p, lig_pop.data(), sizeof(int) * lig_pop.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(lig0.p);
        return fail_free(fa_fail(FA_ERR_HIP, std::string("GA init: ") + hipGetErrorString(e)));
    }
    rc = fa_launch_poses(c, pop, d.gic, G, static_cast<const int*>(lig0.p), d.res, nullptr, 0, nullptr, nullptr, 0,
                         ga->stream);
    if (rc == FA_SUCCESS) {
        hipLaunchKernelGGL(fa::ga_ensemble_kernel, dim3(1), dim3(256), 0, ga->stream, d, 0, pop);
        rc = ga_select(ga, pop);
    }
    e = hipStreamSynchronize(ga->stream);
    (void)hipFree(lig0.p);
    if (rc != FA_SUCCESS) return fail_free(rc);
    if (e != hipSuccess) return fail_free(fa_fail(FA_ERR_HIP, std::string("GA init: ") + hipGetErrorString(e)));
    *out = ga;
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
    const size_t rows = (size_t)ga->d.n_all;
    if (genes) FA_HIP(hipMemcpy(genes, ga->gi.p, sizeof(int32_t) * rows * ga->G, hipMemcpyDeviceToHost));
    if (ic) FA_HIP(hipMemcpy(ic, ga->gic.p, sizeof(double) * rows * ga->G, hipMemcpyDeviceToHost));
    if (results) FA_HIP(hipMemcpy(results, ga->res.p, sizeof(fa_cf_result) * rows, hipMemcpyDeviceToHost));
    return FA_SUCCESS;
}

int fa_ga_thermo(fa_ga* ga, fa_thermo* out, uint64_t* n_poses)
{
    if (!ga || !out) return fa_fail(FA_ERR_INVALID, "null argument");
    FA_HIP(hipSetDevice(ga->ctx->device));
    hipLaunchKernelGGL(fa::ga_thermo_kernel, dim3(1), dim3(64), 0, ga->stream, ga->d, ga->p.temperature,
                       static_cast<fa_thermo*>(ga->thermo.p));
    FA_HIP(hipGetLastError());
    FA_HIP(hipStreamSynchronize(ga->stream));
    FA_HIP(hipMemcpy(out, ga->thermo.p, sizeof(fa_thermo), hipMemcpyDeviceToHost));
    if (n_poses) FA_HIP(hipMemcpy(n_poses, ga->d.n_eval, sizeof(uint64_t), hipMemcpyDeviceToHost));
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

int fa_ga_destroy(fa_ga* ga)
{
    if (!ga) return FA_SUCCESS;
    (void)hipSetDevice(ga->ctx->device);
    if (ga->stream) (void)hipStreamSynchronize(ga->stream);
    if (ga->exec) (void)hipGraphExecDestroy(ga->exec);
    if (ga->graph) (void)hipGraphDestroy(ga->graph);
    for (DevBuf* b : {&ga->gi, &ga->gic, &ga->res, &ga->gi2, &ga->gic2, &ga->res2, &ga->cum, &ga->lim, &ga->ens,
                      &ga->cnt, &ga->lig, &ga->thermo})
        if (b->p) (void)hipFree(b->p);
    if (ga->stream) (void)hipStreamDestroy(ga->stream);
    delete ga;
    return FA_SUCCESS;
}

}  // extern "C"
