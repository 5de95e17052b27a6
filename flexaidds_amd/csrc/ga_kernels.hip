// ga_kernels.hip — one genetic-algorithm generation on the device.
//
// The reference's generation (GA(), LIB/gaboom.cpp:346-500, with BOOM
// reproduction, :886-1036, and the fitness models, :1627-1747) becomes a
// fixed sequence of launches on one stream:
//
//   tick        generation counter(s) + 1 (the counter seeds the draws)
//   breed       one wave per offspring pair: roulette parents, crossover;
//               then one quad per child gene: mutation, gene -> internal
//               coordinate
//   poses       the contact-function kernel over the offspring (caller)
//   ensemble    fold the offspring energies into each job's FP64 ensemble
//   rank        parents + offspring ranked by energy (ties by index), the
//               best `population` gathered in order
//   copy back   survivors into the population rows
//   fitness     LINEAR, PSHARE or SMFREE fitness of the survivors and its
//               running sums for the next roulette; PSHARE then orders the
//               survivors by fitness as GA() does (LIB/gaboom.cpp:498-500)
//
// Random draws are counter-based (fa_rng.h): a generation is a pure function
// of the population, the seed and the generation number, so a job's result
// does not depend on which other jobs share its launches.
#include "ga_kernels.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "bin_index.h"
#include "fa_rng.h"

namespace fa {

__device__ __forceinline__ int slot_of(const GaDev& g, int a) { return g.act ? g.act[a] : a; }
__device__ __forceinline__ size_t parent_row(const GaDev& g, int k, int r) { return (size_t)k * g.pop + r; }
__device__ __forceinline__ size_t child_row(const GaDev& g, int a, int i)
{
    return (size_t)g.jobs * g.pop + (size_t)a * g.n_new + i;
}
__device__ __forceinline__ size_t stage_row(const GaDev& g, int a, int c) { return g.stage + (size_t)a * g.pop + c; }
__device__ __forceinline__ uint64_t stream_of(const GaDev& g, int k, int t)
{
    return g.job_streams ? (((uint64_t)k << 32) | (uint32_t)t) : (uint64_t)(uint32_t)t;
}
__device__ __forceinline__ uint64_t seed_of(const GaDev& g, int k) { return g.seeds ? g.seeds[k] : g.seed; }
__device__ __forceinline__ uint64_t gen_of(const GaDev& g, int k) { return g.gens ? g.gens[k] : *g.gen; }

// LIB/gaboom.cpp:2887-2906 with RandomDouble(gene) = gene / (2^31); the bin
// count of its accumulation loop without the loop (bin_index.h)
__host__ __device__ inline double genetoic(double mn, double del, double bin, int32_t gene)
{
    const double u = gene / ((double)FA_MAX_RANDOM_VALUE + 1.0);
    return mn + del * (double)fa_bin_index(u, bin);
}

__device__ __forceinline__ unsigned low_mask(int n)
{
    // the reference's mask loop: optr = 1; repeat n times optr |= optr << 1
    unsigned optr = 1u;
    for (int i = 0; i < n; ++i) optr |= (optr << 1);
    return optr;
}

// Roulette pick (roullete_wheel, LIB/gaboom.cpp:1158-1182, scans the running
// fitness sums)
// by a whole wave (lane l of 64): the smallest i with cum[i] > u * total
// over non-decreasing running sums (the last row when none, u * rows when
// the total is not positive), found by narrowing the range 64 blocks at a
// time: two rounds of loads for up to 4,096 rows
__device__ __forceinline__ int roulette_wave(const double* cum, int n, double u, int l)
{
    const double tot = cum[n - 1];
    if (tot <= 0.0) return (int)(u * n) % n;
    const double r = u * tot;
    if (!(cum[n - 1] > r)) return n - 1;
    int lo = 0, hi = n - 1;  // answer in [lo, hi], cum[hi] > r
    while (lo < hi) {
        const int S = (hi - lo + 64) / 64;  // block length
        const int idx = min(lo + (l + 1) * S - 1, hi);
        const unsigned long long m = __builtin_amdgcn_ballot_w64(cum[idx] > r);
        const int b = __ffsll((long long)m) - 1;  // first block whose last row passes
        const int nhi = min(lo + (b + 1) * S - 1, hi);
        lo = lo + b * S;
        hi = (S == 1) ? lo : nhi;
    }
    return lo;
}

// crossover points of one pair (LIB/gaboom.cpp:2517-2618): genes
// [gen_a, gen_b] swap, gen_a from bit pnt_a, then gen_b's low pnt_b - 1 bits
// swap back when pnt_b > 0
__device__ __forceinline__ void cross_points(int G, int intragenes, uint64_t seed, uint64_t gen, uint64_t stream,
                                             int& gen_a, int& gen_b, int& pnt_a, int& pnt_b)
{
    gen_a = (int)(fa_rng_u01(seed, gen, stream, FA_DRAW_GEN_A) * (double)G);
    gen_b = (int)(fa_rng_u01(seed, gen, stream, FA_DRAW_GEN_B) * (double)G);
    if (gen_a >= G) gen_a = G - 1;
    if (gen_b >= G) gen_b = G - 1;
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
}

// LIB/gaboom.cpp:2620-2641
__global__ void ga_tick_kernel(GaDev g)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (g.gens) {
        if (a < g.n_act) g.gens[slot_of(g, a)] += 1ull;
    } else if (a == 0) {
        *g.gen += 1ull;
    }
}

// initial population of each job into its staged rows, generation 0
// (LIB/gaboom.cpp:1871-1902); genes past a job's own count are 0.  In
// candidate mode (duplicate rejection) the batch's draws go to the job's
// candidate rows instead, from draw stream t0 + c.
__global__ void ga_init_kernel(GaDev g)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)g.n_act * g.pop * g.Gmax) return;
    const int a = (int)(t / ((long long)g.pop * g.Gmax));
    const int c = (int)((t / g.Gmax) % g.pop), j = (int)(t % g.Gmax);
    const int k = slot_of(g, a);
    if (g.cand_mode && g.need[k] == 0) return;
    const size_t row = g.cand_mode ? g.cand + (size_t)a * g.n_cand + c : stage_row(g, a, c);
    const size_t at = row * g.Gmax + j;
    if (j >= g.G[k]) {
        g.gi[at] = 0;
        g.gic[at] = 0.0;
        return;
    }
    const size_t l = (size_t)k * g.Gmax + j;
    const int32_t gene = fa_rng_gene(seed_of(g, k), 0ull, stream_of(g, k, g.t0 + c), (uint64_t)j);
    g.gi[at] = gene;
    g.gic[at] = genetoic(g.lmin[l], g.ldel[l], g.lbin[l], gene);
}

// Reproduction (LIB/gaboom.cpp:920-973) in two launches: one wave per
// offspring pair picks the parents by roulette and crosses them over into the
// children rows (lane j: gene j); then one quad per child gene applies that
// gene's 32 mutation draws and converts it to an internal coordinate.  The
// draws are the ones a single thread per pair would make (fa_rng.h layout),
// so the split changes nothing but the parallelism.  In candidate mode
// (duplicate rejection) pair t of the batch draws from stream t0 + t and
// both of its children go to the job's candidate rows 2t, 2t + 1.
__device__ __forceinline__ size_t out_row(const GaDev& g, int a, int i)
{
    return g.cand_mode ? g.cand + (size_t)a * g.n_cand + i : child_row(g, a, i);
}

__global__ __launch_bounds__(256) void ga_pair_kernel(GaDev g)
{
    // one wave per offspring pair, lane j holding gene j of both children
    const int pairs = (g.n_new + 1) / 2;
    const long long tt = (long long)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    const int l = threadIdx.x & 63;
    if (tt >= (long long)g.n_act * pairs) return;  // whole waves
    const int a = (int)(tt / pairs), t = (int)(tt % pairs);
    const int k = slot_of(g, a);
    if (g.cand_mode && g.need[k] == 0) return;     // whole waves: one job per wave
    const uint64_t gen = gen_of(g, k), seed = seed_of(g, k);
    const uint64_t st = stream_of(g, k, g.t0 + t);
    const int G = g.G[k], Gm = g.Gmax;
    const double* cum = g.cum + (size_t)k * g.pop;
    const int p1 = roulette_wave(cum, g.pop, fa_rng_u01(seed, gen, st, FA_DRAW_P1), l);
    const int p2 = roulette_wave(cum, g.pop, fa_rng_u01(seed, gen, st, FA_DRAW_P2), l);
    unsigned jo = 0u, ma = 0u;
    if (l < G) {
        jo = (unsigned)g.gi[parent_row(g, k, p1) * Gm + l];
        ma = (unsigned)g.gi[parent_row(g, k, p2) * Gm + l];
    }
    if (fa_rng_u01(seed, gen, st, FA_DRAW_CROSS) < g.cross) {
        int ga, gb, pa, pb;
        cross_points(G, g.intragenes, seed, gen, st, ga, gb, pa, pb);
        // the reference's crossover on gene l: the span swap, then the end
        // gene's second swap
        if (l >= ga && l <= gb) {
            const unsigned optr = low_mask((l == ga) ? pa : 32);
            const unsigned nj = (jo & ~optr) | (ma & optr), nm = (jo & optr) | (ma & ~optr);
            jo = nj;
            ma = nm;
        }
        if (pb > 0 && l == gb) {
            const unsigned optr = low_mask(pb - 1);
            const unsigned nj = (jo & ~optr) | (ma & optr), nm = (jo & optr) | (ma & ~optr);
            jo = nj;
            ma = nm;
        }
    }
    const size_t c1 = out_row(g, a, 2 * t);
    if (l < G) {
        g.gi[c1 * Gm + l] = (int32_t)jo;
        if (g.cand_mode || 2 * t + 1 < g.n_new) g.gi[(c1 + 1) * Gm + l] = (int32_t)ma;
    }
}

// per child gene: mutation (LIB/gaboom.cpp:2620-2641) and genetoic; genes
// past the job's own count are 0 (rows are Gmax wide).  Four lanes per gene
// (one quad), eight of the gene's 32 bit draws each; the quad ORs its flip
// masks and its first lane writes the gene.  Grid y is the job.
__global__ __launch_bounds__(256) void ga_mutate_kernel(GaDev g)
{
    const int a = blockIdx.y;
    const int tq = blockIdx.x * blockDim.x + threadIdx.x;
    const int gi = tq >> 2, part = tq & 3;  // child gene of the job, quarter of its bits
    const int n_out = g.cand_mode ? 2 * ((g.n_new + 1) / 2) : g.n_new;
    const bool live = gi < n_out * g.Gmax;
    const int i = gi / g.Gmax, j = gi - i * g.Gmax;
    const int k = slot_of(g, a);
    if (g.cand_mode && g.need[k] == 0) return;  // the whole block: grid y is the job
    const int G = g.G[k];
    unsigned optr = 0u;
    if (live && j < G) {
        const uint64_t gen = gen_of(g, k), seed = seed_of(g, k);
        const uint64_t st = stream_of(g, k, g.t0 + (i >> 1));
        const uint64_t base = FA_DRAW_MUT + ((uint64_t)(i & 1) * G + j) * 32 + 8 * part;
        for (int b = 0; b < 8; ++b)
            if (fa_rng_u01(seed, gen, st, base + b) < g.mut) optr |= 1u << (8 * part + b);
    }
    optr |= (unsigned)__shfl_xor((int)optr, 1);
    optr |= (unsigned)__shfl_xor((int)optr, 2);
    if (!live || part != 0) return;
    const size_t at = out_row(g, a, i) * g.Gmax + j;
    if (j >= G) {
        g.gi[at] = 0;
        g.gic[at] = 0.0;
        return;
    }
    const int32_t gene = g.gi[at] ^ (int32_t)optr;
    const size_t lk = (size_t)k * g.Gmax + j;
    g.gi[at] = gene;
    g.gic[at] = genetoic(g.lmin[lk], g.ldel[lk], g.lbin[lk], gene);
}

// energies of each job's children (children = 1) or staged initial
// population (children = 0) into its ensemble state; one block per job
__global__ __launch_bounds__(256) void ga_ensemble_kernel(GaDev g, int children)
{
    __shared__ LseState sh[256];
    const int a = blockIdx.x;
    const int k = slot_of(g, a);
    const int n = children ? g.n_new : g.pop;
    LseState st{0.0, 0.0, 0.0, 0.0};
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const double e = g.res[children ? child_row(g, a, i) : stage_row(g, a, i)].cf;
        lse_push(st, -g.beta * e, e);
    }
    sh[threadIdx.x] = st;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) lse_merge(sh[threadIdx.x], sh[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        LseState acc = g.ens[k];
        lse_merge(acc, sh[0]);
        g.ens[k] = acc;
        atomicAdd(g.n_eval, (unsigned long long)n);
    }
}

// total order on energies (NaN last)
__device__ __forceinline__ uint64_t order_key(double e)
{
    const uint64_t u = (uint64_t)__double_as_longlong(e);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}

// candidate r of the job at position a: its parents then its children, or
// its staged initial population
__device__ __forceinline__ size_t cand_row(const GaDev& g, int a, int k, int r, int initial)
{
    if (initial) return stage_row(g, a, r);
    return r < g.pop ? parent_row(g, k, r) : child_row(g, a, r - g.pop);
}

// What a ranking launch ranks: every candidate (parents then children), the
// staged initial population, or the parents alone (STEADY's first step)
enum RankMode { RANK_ALL = 0, RANK_INITIAL = 1, RANK_PARENTS = 2 };

__device__ __forceinline__ int rank_n(const GaDev& g, int mode) { return mode == RANK_ALL ? g.n_all : g.pop; }

// the key a candidate is ranked by: its energy's order key; under STEADY
// the parents beyond the best pop - offspring (by their rank among the
// parents) rank after every other candidate, so they never survive
__device__ __forceinline__ uint64_t rank_key(const GaDev& g, int a, int k, int r, int mode)
{
    const uint64_t e = order_key(g.res[cand_row(g, a, k, r, mode == RANK_INITIAL)].cf);
    if (mode == RANK_ALL && g.steady && r < g.pop && g.prank[(size_t)a * g.pop + r] >= g.pop - g.n_new) return ~0ull;
    return e;
}

// Rank of each of a job's n candidates by energy, ties by candidate index
// (a total order, so ranks are a permutation); the best `pop` are gathered
// in order (QuickSort by evalue, LIB/gaboom.cpp:2674-2726).  Above
// RANK_COUNT_MAX candidates, ranking by sorted tiles: every tile of
// RANK_TILE candidates is sorted in LDS (bitonic, by (energy key, index)); a
// candidate's rank is the number of candidates before it in each sorted
// tile, found by binary search with the tile staged in LDS:
// n (n / RANK_TILE) log2(RANK_TILE) steps instead of n^2 comparisons, exact,
// no atomics.  Smaller sets are counted directly (below), which spreads
// over more blocks.
constexpr int RANK_TILE = 1024;
constexpr int RANK_SORT_THREADS = RANK_TILE / 2;

__global__ __launch_bounds__(RANK_SORT_THREADS) void ga_rank_sort_kernel(GaDev g, int mode)
{
    __shared__ unsigned long long sk[RANK_TILE];
    __shared__ int si[RANK_TILE];
    const int a = blockIdx.y;
    const int k = slot_of(g, a);
    const int n = rank_n(g, mode);
    const int base = blockIdx.x * RANK_TILE;
    const int m = min(RANK_TILE, n - base);
    for (int q = threadIdx.x; q < RANK_TILE; q += RANK_SORT_THREADS) {
        if (q < m) {
            sk[q] = rank_key(g, a, k, base + q, mode);
            si[q] = base + q;
        } else {  // padding sorts last
            sk[q] = ~0ull;
            si[q] = 0x7fffffff;
        }
    }
    __syncthreads();
    const int q = threadIdx.x;  // one compare-exchange per thread per stage
    for (int size = 2; size <= RANK_TILE; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            const int lo = 2 * stride * (q / stride) + (q % stride), hi = lo + stride;
            const bool up = (lo & size) == 0;
            const unsigned long long ka = sk[lo], kb = sk[hi];
            const int ia = si[lo], ib = si[hi];
            const bool gt = (ka > kb) || (ka == kb && ia > ib);
            if (gt == up) {
                sk[lo] = kb;
                sk[hi] = ka;
                si[lo] = ib;
                si[hi] = ia;
            }
            __syncthreads();
        }
    }
    const size_t out = (size_t)a * g.n_all + base;
    for (int p = threadIdx.x; p < m; p += RANK_SORT_THREADS) {
        g.skey[out + p] = sk[p];
        g.sidx[out + p] = si[p];
    }
}

// Small rankings (n <= RANK_COUNT_MAX, where the tiles would leave the
// device idle): every candidate counts the candidates before it, in chunks
// of RANK_CHUNK keys staged in LDS, one block per (candidates, chunk) pair,
// integer atomics (order-free) into zeroed ranks.
constexpr int RANK_CHUNK = 256;
constexpr int RANK_COUNT_MAX = 8192;

__device__ __forceinline__ int* rank_out(const GaDev& g, int a, int mode)
{
    return mode == RANK_PARENTS ? g.prank + (size_t)a * g.pop : g.rank + (size_t)a * g.n_all;
}

__global__ void ga_rank_zero_kernel(GaDev g, int mode)
{
    const int n = rank_n(g, mode);
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < (long long)g.n_act * n) rank_out(g, (int)(t / n), mode)[t % n] = 0;
}

__global__ __launch_bounds__(256) void ga_rank_count_kernel(GaDev g, int mode)
{
    __shared__ uint64_t keys[RANK_CHUNK];
    const int a = blockIdx.z;
    const int k = slot_of(g, a);
    const int n = rank_n(g, mode);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int base = blockIdx.y * RANK_CHUNK;
    const int lim = min(RANK_CHUNK, n - base);
    for (int q = threadIdx.x; q < lim; q += blockDim.x) keys[q] = rank_key(g, a, k, base + q, mode);
    __syncthreads();
    if (i >= n) return;
    const uint64_t ki = rank_key(g, a, k, i, mode);
    int cnt = 0;
#pragma unroll 8
    for (int q = 0; q < lim; ++q) {
        const uint64_t kj = keys[q];
        cnt += (kj < ki) || (kj == ki && base + q < i);
    }
    if (cnt) atomicAdd(&rank_out(g, a, mode)[i], cnt);
}

// four candidates per thread (one tile per block) against every sorted tile
__global__ __launch_bounds__(256) void ga_rank_search_kernel(GaDev g, int mode)
{
    __shared__ unsigned long long sk[RANK_TILE];
    __shared__ int si[RANK_TILE];
    const int a = blockIdx.y;
    const int k = slot_of(g, a);
    const int n = rank_n(g, mode);
    const int tiles = (n + RANK_TILE - 1) / RANK_TILE;
    unsigned long long key[RANK_TILE / 256];
    int idx[RANK_TILE / 256], rank[RANK_TILE / 256];
#pragma unroll
    for (int r = 0; r < RANK_TILE / 256; ++r) {
        idx[r] = blockIdx.x * RANK_TILE + r * 256 + (int)threadIdx.x;
        key[r] = idx[r] < n ? rank_key(g, a, k, idx[r], mode) : 0ull;
        rank[r] = 0;
    }
    const size_t row = (size_t)a * g.n_all;
    for (int c = 0; c < tiles; ++c) {
        const int m = min(RANK_TILE, n - c * RANK_TILE);
        __syncthreads();
        for (int q = threadIdx.x; q < m; q += 256) {
            sk[q] = g.skey[row + (size_t)c * RANK_TILE + q];
            si[q] = g.sidx[row + (size_t)c * RANK_TILE + q];
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < RANK_TILE / 256; ++r) {
            if (idx[r] >= n) continue;
            int lo = 0, hi = m;  // first entry not before (key, idx)
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                const bool before = sk[mid] < key[r] || (sk[mid] == key[r] && si[mid] < idx[r]);
                if (before)
                    lo = mid + 1;
                else
                    hi = mid;
            }
            rank[r] += lo;
        }
    }
    int* out = rank_out(g, a, mode);
#pragma unroll
    for (int r = 0; r < RANK_TILE / 256; ++r)
        if (idx[r] < n) out[idx[r]] = rank[r];
}

// one thread per (candidate, gene): a survivor's row moves in consecutive words
__global__ void ga_rank_gather_kernel(GaDev g, int initial)
{
    const int n = initial ? g.pop : g.n_all;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)g.n_act * n * g.Gmax) return;
    const long long c = t / g.Gmax;
    const int j = (int)(t - c * g.Gmax);
    const int a = (int)(c / n), i = (int)(c % n);
    const int k = slot_of(g, a);
    const int rank = g.rank[(size_t)a * g.n_all + i];
    if (rank >= g.pop) return;
    const size_t src = cand_row(g, a, k, i, initial), dst = parent_row(g, k, rank);
    g.gi2[dst * g.Gmax + j] = g.gi[src * g.Gmax + j];
    g.gic2[dst * g.Gmax + j] = g.gic[src * g.Gmax + j];
    if (j == 0) g.res2[dst] = g.res[src];
}

// gathered survivors (and, with fit_too, their fitness) into the parent rows
__global__ void ga_copy_back_kernel(GaDev g, int fit_too)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long per = (long long)g.pop * g.Gmax;
    if (t < (long long)g.n_act * per) {
        const int a = (int)(t / per);
        const size_t at = (size_t)slot_of(g, a) * per + (size_t)(t % per);
        g.gi[at] = g.gi2[at];
        g.gic[at] = g.gic2[at];
    }
    if (t < (long long)g.n_act * g.pop) {
        const int a = (int)(t / g.pop);
        const size_t at = (size_t)slot_of(g, a) * g.pop + (size_t)(t % g.pop);
        g.res[at] = g.res2[at];
        if (fit_too) g.fit[at] = g.fit2[at];
    }
}

// LINEAR fitness pop - rank and its inclusive prefix sums, one block per
// job (the values are integers, so the sums are exact in any order)
__global__ __launch_bounds__(256) void ga_linear_kernel(GaDev g)
{
    __shared__ double part[256];
    const int k = slot_of(g, blockIdx.x);
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

// Niche counts (LIB/gaboom.cpp:1649-1660): for survivor i of a job, the sum
// over its survivors j of 1 - (rmsp / sig)^alpha where rmsp <= sig, with
// rmsp = sqrt(sum of squared internal-coordinate differences / genes)
// (calc_rmsp, LIB/gaboom.cpp:2874-2885).  Survivors j stream through LDS in
// tiles; each thread keeps its own row's sums in the reference's order.
constexpr int SHARE_TILE = 32;

__global__ __launch_bounds__(256) void ga_share_kernel(GaDev g)
{
    __shared__ double tile[SHARE_TILE * FA_MAX_GENES];
    const int k = slot_of(g, blockIdx.y);
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

// PSHARE / SMFREE fitness of a job's survivors (rows by increasing energy),
// one block per job (LIB/gaboom.cpp:1637-1747).  SMFREE's Boltzmann weights
// follow StatMechEngine::boltzmann_weights over the survivors' energies
// (LIB/statmech.cpp:178-201 with compute_boltzmann_batch's exp(-beta (E -
// Emin)), LIB/UnifiedHardwareDispatch.cpp:608-624), normalised by their
// maximum.
__global__ __launch_bounds__(256) void ga_shared_fitness_kernel(GaDev g)
{
    __shared__ double sh[256];
    const int k = slot_of(g, blockIdx.x);
    const int n = g.pop;
    double Z = 1.0, maxw = 1.0, emin = 0.0;
    const bool smfree = g.model == FA_FIT_SMFREE;
    if (smfree) {
        double m = INFINITY;
        for (int i = threadIdx.x; i < n; i += blockDim.x) m = fmin(m, g.res[parent_row(g, k, i)].cf);
        emin = -block_max(-m, sh);
        // unscored poses (energy +inf) have weight 0, also when every
        // survivor is one (no finite minimum)
        auto bw = [&](double e) { return (e < INFINITY) ? exp(-g.beta * (e - emin)) : 0.0; };
        double z = 0.0;
        for (int i = threadIdx.x; i < n; i += blockDim.x) z += bw(g.res[parent_row(g, k, i)].cf);
        Z = block_sum(z, sh);
        double mw = 0.0;
        if (Z > 0.0)
            for (int i = threadIdx.x; i < n; i += blockDim.x) mw = fmax(mw, bw(g.res[parent_row(g, k, i)].cf) / Z);
        maxw = block_max(mw, sh);
        if (maxw <= 0.0) maxw = 1.0;
    }
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const double ps = g.share[(size_t)k * n + i];
        double f;
        if (smfree) {
            const double e = g.res[parent_row(g, k, i)].cf;
            const double w = (Z > 0.0 && e < INFINITY) ? exp(-g.beta * (e - emin)) / Z : 0.0;
            const double rank = (double)(n - i) / (double)n;
            const double blended = (1.0 - g.ew) * rank + g.ew * (w / maxw);
            f = blended * (double)n / ps;
        } else {
            f = (double)(n - i) / ps;
        }
        g.fit[(size_t)k * n + i] = f;
    }
}

// position of each of a job's survivors when ordered by decreasing fitness
// (QuickSort(chrom, 0, num_chrom - 1, false) in GA(), LIB/gaboom.cpp:498-500;
// ties keep row order), gathered into the second buffers
__global__ __launch_bounds__(256) void ga_fitness_order_kernel(GaDev g)
{
    __shared__ uint64_t keys[1024];
    const int k = slot_of(g, blockIdx.y);
    const int n = g.pop;
    const double* f = g.fit + (size_t)k * n;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t ki = (i < n) ? order_key(-f[i]) : 0ull;
    int rank = 0;
    for (int base = 0; base < n; base += 1024) {
        __syncthreads();
        for (int q = threadIdx.x; q < 1024 && base + q < n; q += blockDim.x) keys[q] = order_key(-f[base + q]);
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
        g.fit2[(size_t)k * n + rank] = f[i];
    }
}

// running sums of a job's fitness in row order, one thread per job: the
// sequential order of roullete_wheel's loop (LIB/gaboom.cpp:1158-1182), so
// a draw picks the row the reference's scan would
__global__ void ga_cum_kernel(GaDev g)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= g.n_act) return;
    const int k = slot_of(g, a);
    const double* f = g.fit + (size_t)k * g.pop;
    double* c = g.cum + (size_t)k * g.pop;
    double tot = 0.0;
    for (int i = 0; i < g.pop; ++i) {
        tot += f[i];
        c[i] = tot;
    }
}

// ensemble thermodynamics of slots [0, n)
__global__ void ga_thermo_kernel(GaDev g, int n, double T, fa_thermo* out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) out[k] = thermo_of(g.ens[k], T);
}

// ------------------------------------------------------------ duplicates
// Duplicate rejection (GB->duplicates = 0, the reference's default;
// LIB/gaboom.cpp:971-1018 for children, :1966-1978 for the initial
// population).  The reference draws candidates one at a time and keeps a
// candidate whose signature (hash_genes, :1040-1045) is not yet in the job's
// set, adding it, until the rows are filled; the set lives for the whole run.
// Here a batch of candidates is drawn at once and kept in draw order: a
// candidate is new when its signature is in neither the set nor an earlier
// candidate of the batch; the first `need` new ones fill the rows and enter
// the set, exactly the candidates the one-at-a-time loop keeps.  After
// dd_rounds batches the rows still empty take the last batch's rejected
// candidates in draw order (the reference would keep drawing).

// hash_genes: h ^= hash(int32(ic + 0.5)) + 0x9e3779b9 + (h << 6) + (h >> 2)
// over the genes, std::hash<int32_t> being the value as a size_t
// (sign-extended); the same value, so signatures collide where the
// reference's do
__device__ __forceinline__ unsigned long long genes_sig(const double* ic, int G)
{
    unsigned long long h = 0ull;
    for (int j = 0; j < G; ++j) {
        const int32_t v = (int32_t)(ic[j] + 0.5);
        h ^= (unsigned long long)(long long)v + 0x9e3779b9ull + (h << 6) + (h >> 2);
    }
    return h;
}

// table slot of a signature (splitmix64 finalizer: the signatures of nearby
// angle vectors differ in few bits)
__device__ __forceinline__ size_t sig_slot(unsigned long long x, int log2)
{
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ull;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebull;
    x ^= x >> 31;
    return (size_t)(x & ((1ull << log2) - 1ull));
}

// open addressing, linear probing, key 0 = empty (signature 0 is kept
// beside the table); tables stay at most half full, so probes end early
__device__ bool tab_has(const unsigned long long* t, int log2, unsigned long long key)
{
    const size_t mask = ((size_t)1 << log2) - 1;
    size_t h = sig_slot(key, log2);
    for (size_t p = 0; p <= mask; ++p) {
        const unsigned long long v = t[h];
        if (v == key) return true;
        if (v == 0ull) return false;
        h = (h + 1) & mask;
    }
    return false;
}

__device__ long long tab_insert(unsigned long long* t, int log2, unsigned long long key)
{
    const size_t mask = ((size_t)1 << log2) - 1;
    size_t h = sig_slot(key, log2);
    for (size_t p = 0; p <= mask; ++p) {
        const unsigned long long prev = atomicCAS(&t[h], 0ull, key);
        if (prev == 0ull || prev == key) return (long long)h;
        h = (h + 1) & mask;
    }
    return -1;
}

__device__ long long tab_find(const unsigned long long* t, int log2, unsigned long long key)
{
    const size_t mask = ((size_t)1 << log2) - 1;
    size_t h = sig_slot(key, log2);
    for (size_t p = 0; p <= mask; ++p) {
        const unsigned long long v = t[h];
        if (v == key) return (long long)h;
        if (v == 0ull) return -1;
        h = (h + 1) & mask;
    }
    return -1;
}

__global__ void ga_dd_start_kernel(GaDev g, int initial)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a < g.n_act) g.need[slot_of(g, a)] = initial ? g.pop : g.n_new;
}

// clears each job's batch table (jobs with rows to fill); grid y is the job
__global__ void ga_dd_clear_kernel(GaDev g)
{
    const int a = blockIdx.y, k = slot_of(g, a);
    if (g.need[k] == 0) return;
    const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < ((size_t)1 << g.rtab_log2)) {
        g.rkey[((size_t)k << g.rtab_log2) + j] = 0ull;
        g.rval[((size_t)k << g.rtab_log2) + j] = 0x7fffffff;
    }
    if (j == 0) g.rzero[k] = 0x7fffffff;
}

// signature of every candidate of the batch; whether the set holds it; the
// first candidate of the batch holding each new signature
__global__ __launch_bounds__(256) void ga_dd_sig_kernel(GaDev g, int n_round)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)g.n_act * n_round) return;
    const int a = (int)(t / n_round), i = (int)(t % n_round);
    const int k = slot_of(g, a);
    if (g.need[k] == 0) return;
    const size_t row = g.cand + (size_t)a * g.n_cand + i;
    const unsigned long long sig = genes_sig(g.gic + row * g.Gmax, g.G[k]);
    const size_t c = (size_t)a * g.n_cand + i;
    g.csig[c] = sig;
    const bool dup = sig == 0ull ? g.dzero[k] != 0 : tab_has(g.dset + ((size_t)k << g.dset_log2), g.dset_log2, sig);
    g.cdup[c] = dup ? 1 : 0;
    if (dup) return;
    if (sig == 0ull) {
        atomicMin(&g.rzero[k], i);
    } else {
        const long long h = tab_insert(g.rkey + ((size_t)k << g.rtab_log2), g.rtab_log2, sig);
        if (h >= 0) atomicMin(&g.rval[((size_t)k << g.rtab_log2) + h], i);
    }
}

// exclusive rank of p among the block's 256 threads (four waves), and the count
__device__ __forceinline__ int block_rank(bool p, int& total, int* wsum)
{
    const unsigned long long m = __builtin_amdgcn_ballot_w64(p);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int before = __popcll(m & ((1ull << l) - 1ull));
    if (l == 0) wsum[w] = __popcll(m);
    __syncthreads();
    int off = 0, tot = 0;
    for (int q = 0; q < 4; ++q) {
        if (q < w) off += wsum[q];
        tot += wsum[q];
    }
    __syncthreads();
    total = tot;
    return off + before;
}

__device__ __forceinline__ void copy_row(const GaDev& g, size_t from, size_t to)
{
    for (int j = 0; j < g.Gmax; ++j) {
        g.gi[to * g.Gmax + j] = g.gi[from * g.Gmax + j];
        g.gic[to * g.Gmax + j] = g.gic[from * g.Gmax + j];
    }
}

// one block of 256 per job: the batch's new candidates in draw order fill
// the job's remaining rows and enter its set; on the last batch the rows
// still empty take the batch's rejected candidates in draw order
__global__ __launch_bounds__(256) void ga_dd_accept_kernel(GaDev g, int n_round, int initial, int last)
{
    __shared__ int wsum[4];
    __shared__ int s_last;
    const int a = blockIdx.x, k = slot_of(g, a);
    const int need0 = g.need[k];
    if (need0 <= 0) return;
    const int rows = initial ? g.pop : g.n_new;
    const size_t c0 = (size_t)a * g.n_cand;
    const unsigned long long* rkey = g.rkey + ((size_t)k << g.rtab_log2);
    const int* rval = g.rval + ((size_t)k << g.rtab_log2);
    auto is_new = [&](int i) {
        if (i >= n_round || g.cdup[c0 + i]) return false;
        const unsigned long long sig = g.csig[c0 + i];
        if (sig == 0ull) return g.rzero[k] == i;
        const long long h = tab_find(rkey, g.rtab_log2, sig);
        return h >= 0 && rval[h] == i;
    };
    auto dest = [&](int r) { return initial ? stage_row(g, a, r) : child_row(g, a, r); };
    if (threadIdx.x == 0) s_last = n_round - 1;
    int base = 0;
    for (int s0 = 0; s0 < n_round && base < need0; s0 += 256) {
        const int i = s0 + (int)threadIdx.x;
        const bool acc = is_new(i);
        int total;
        const int r = block_rank(acc, total, wsum);
        if (acc && base + r < need0) {
            copy_row(g, g.cand + c0 + i, dest(rows - need0 + base + r));
            const unsigned long long sig = g.csig[c0 + i];
            if (sig == 0ull)
                g.dzero[k] = 1;
            else
                (void)tab_insert(g.dset + ((size_t)k << g.dset_log2), g.dset_log2, sig);
            if (base + r == need0 - 1) s_last = i;   // the candidate that fills the last row
        }
        base += total;
    }
    __syncthreads();
    const int kept = min(base, need0);
    const int examined = (base >= need0) ? s_last + 1 : n_round;
    int filled = 0;
    if (last && kept < need0) {
        int fb = 0;
        for (int s0 = 0; s0 < n_round && fb < need0 - kept; s0 += 256) {
            const int i = s0 + (int)threadIdx.x;
            const bool rej = i < n_round && !is_new(i);
            int total;
            const int r = block_rank(rej, total, wsum);
            if (rej && fb + r < need0 - kept) copy_row(g, g.cand + c0 + i, dest(rows - need0 + kept + fb + r));
            fb += total;
        }
        filled = min(fb, need0 - kept);
    }
    if (threadIdx.x == 0) {
        g.need[k] = need0 - kept - filled;
        g.dd_stats[2 * (size_t)k] += (unsigned long long)(examined - kept);
        g.dd_stats[2 * (size_t)k + 1] += (unsigned long long)filled;
    }
}

// grid y is the job: its set, signature-0 flag and counts cleared
__global__ void ga_dd_reset_kernel(GaDev g)
{
    const int a = blockIdx.y, k = slot_of(g, a);
    const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < ((size_t)1 << g.dset_log2)) g.dset[((size_t)k << g.dset_log2) + j] = 0ull;
    if (j == 0) {
        g.dzero[k] = 0;
        g.dd_stats[2 * (size_t)k] = 0ull;
        g.dd_stats[2 * (size_t)k + 1] = 0ull;
    }
}

__global__ void ga_dset_rehash_kernel(GaDev g, const unsigned long long* old, int old_log2)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ((size_t)g.jobs << old_log2)) return;
    const unsigned long long key = old[t];
    if (key == 0ull) return;
    const size_t k = t >> old_log2;
    (void)tab_insert(g.dset + (k << g.dset_log2), g.dset_log2, key);
}

namespace {
unsigned blocks_of(long long n) { return (unsigned)std::max(1LL, (n + 255) / 256); }
}  // namespace

void ga_launch_tick(const GaDev& g, hipStream_t s)
{
    if (g.n_act < 1) return;
    hipLaunchKernelGGL(ga_tick_kernel, dim3(blocks_of(g.n_act)), dim3(256), 0, s, g);
}

void ga_launch_init(const GaDev& g, hipStream_t s)
{
    if (g.n_act < 1) return;
    hipLaunchKernelGGL(ga_init_kernel, dim3(blocks_of((long long)g.n_act * g.pop * g.Gmax)), dim3(256), 0, s, g);
}

void ga_launch_breed(const GaDev& g, hipStream_t s)
{
    if (g.n_act < 1) return;
    // four pairs per 256-thread block (a wave each); 64 child genes per
    // block (a quad each), one row of blocks per job
    const long long pairs = (long long)g.n_act * ((g.n_new + 1) / 2);
    const long long n_out = g.cand_mode ? 2ll * ((g.n_new + 1) / 2) : g.n_new;
    hipLaunchKernelGGL(ga_pair_kernel, dim3((unsigned)std::max(1LL, (pairs + 3) / 4)), dim3(256), 0, s, g);
    hipLaunchKernelGGL(ga_mutate_kernel, dim3(blocks_of(4ll * n_out * g.Gmax), g.n_act), dim3(256), 0, s, g);
}

void ga_launch_dedup_fill(const GaDev& g, int initial, hipStream_t s)
{
    if (g.n_act < 1 || g.dd_rounds < 1) return;
    const int P = (g.n_new + 1) / 2;
    const int n_round = initial ? g.pop : 2 * P;
    hipLaunchKernelGGL(ga_dd_start_kernel, dim3(blocks_of(g.n_act)), dim3(256), 0, s, g, initial);
    for (int r = 0; r < g.dd_rounds; ++r) {
        GaDev b = g;
        b.cand_mode = 1;
        b.dd_round = r;
        b.t0 = r * (initial ? g.pop : P);
        hipLaunchKernelGGL(ga_dd_clear_kernel, dim3(blocks_of(1ll << g.rtab_log2), g.n_act), dim3(256), 0, s, b);
        if (initial)
            ga_launch_init(b, s);
        else
            ga_launch_breed(b, s);
        hipLaunchKernelGGL(ga_dd_sig_kernel, dim3(blocks_of((long long)g.n_act * n_round)), dim3(256), 0, s, b,
                           n_round);
        hipLaunchKernelGGL(ga_dd_accept_kernel, dim3(g.n_act), dim3(256), 0, s, b, n_round, initial,
                           r == g.dd_rounds - 1 ? 1 : 0);
    }
}

void ga_launch_dedup_reset(const GaDev& g, hipStream_t s)
{
    if (g.n_act < 1 || g.dd_rounds < 1) return;
    hipLaunchKernelGGL(ga_dd_reset_kernel, dim3(blocks_of(1ll << g.dset_log2), g.n_act), dim3(256), 0, s, g);
}

void ga_launch_dset_rehash(const GaDev& g, const unsigned long long* old_keys, int old_log2, hipStream_t s)
{
    hipLaunchKernelGGL(ga_dset_rehash_kernel, dim3(blocks_of((long long)g.jobs << old_log2)), dim3(256), 0, s, g,
                       old_keys, old_log2);
}

void ga_launch_ensemble(const GaDev& g, int children, hipStream_t s)
{
    if (g.n_act < 1) return;
    hipLaunchKernelGGL(ga_ensemble_kernel, dim3(g.n_act), dim3(256), 0, s, g, children);
}

void ga_launch_select(const GaDev& g, int initial, hipStream_t s)
{
    if (g.n_act < 1) return;
    const int n = initial ? g.pop : g.n_all;
    const unsigned rows = blocks_of((long long)g.n_act * g.pop * g.Gmax);
    const long long cand = (long long)g.n_act * n;
    auto rank = [&](int mode) {
        const int m = mode == RANK_ALL ? g.n_all : g.pop;
        if (m <= RANK_COUNT_MAX) {
            hipLaunchKernelGGL(ga_rank_zero_kernel, dim3(blocks_of((long long)g.n_act * m)), dim3(256), 0, s, g, mode);
            hipLaunchKernelGGL(ga_rank_count_kernel, dim3(blocks_of(m), (m + RANK_CHUNK - 1) / RANK_CHUNK, g.n_act),
                               dim3(256), 0, s, g, mode);
            return;
        }
        const unsigned tiles = (unsigned)((m + RANK_TILE - 1) / RANK_TILE);
        hipLaunchKernelGGL(ga_rank_sort_kernel, dim3(tiles, g.n_act), dim3(RANK_SORT_THREADS), 0, s, g, mode);
        hipLaunchKernelGGL(ga_rank_search_kernel, dim3(tiles, g.n_act), dim3(256), 0, s, g, mode);
    };
    // STEADY (LIB/gaboom.cpp:1022-1027): the parents by energy first; the
    // worst offspring-many of them give way to the children
    if (!initial && g.steady) rank(RANK_PARENTS);
    rank(initial ? RANK_INITIAL : RANK_ALL);
    hipLaunchKernelGGL(ga_rank_gather_kernel, dim3(blocks_of(cand * g.Gmax)), dim3(256), 0, s, g, initial);
    hipLaunchKernelGGL(ga_copy_back_kernel, dim3(rows), dim3(256), 0, s, g, 0);
    if (g.model == FA_FIT_LINEAR) {
        hipLaunchKernelGGL(ga_linear_kernel, dim3(g.n_act), dim3(256), 0, s, g);
        return;
    }
    hipLaunchKernelGGL(ga_share_kernel, dim3(blocks_of(g.pop), g.n_act), dim3(256), 0, s, g);
    hipLaunchKernelGGL(ga_shared_fitness_kernel, dim3(g.n_act), dim3(256), 0, s, g);
    if (g.model == FA_FIT_PSHARE) {
        hipLaunchKernelGGL(ga_fitness_order_kernel, dim3(blocks_of(g.pop), g.n_act), dim3(256), 0, s, g);
        hipLaunchKernelGGL(ga_copy_back_kernel, dim3(rows), dim3(256), 0, s, g, 1);
    }
    hipLaunchKernelGGL(ga_cum_kernel, dim3(blocks_of(g.n_act)), dim3(256), 0, s, g);
}

void ga_launch_thermo(const GaDev& g, int n_slots, double T, fa_thermo* out, hipStream_t s)
{
    if (n_slots < 1) return;
    hipLaunchKernelGGL(ga_thermo_kernel, dim3(blocks_of(n_slots)), dim3(256), 0, s, g, n_slots, T, out);
}

}  // namespace fa
