// ga.hip — the device GA behind fa_ga_* (C ABI).
//
// A GA holds J independent jobs (ligands) that advance together: rows
// [0, J*pop) are the parents of job 0, 1, ... and rows [J*pop, J*(pop+new))
// their children, so one pose launch scores every job's children (the
// batched form; J = 1 is the plain GA).  Each job keeps its own FP64
// ensemble, which gives its ln Z and, over jobs, the grand partition
// function.  One generation (ga_kernels.h) allocates nothing and records no
// memset, so it is captured once as a HIP graph and replayed.
//
// Snapshot (fa_ga_snapshot_enable / fa_ga_snapshot): after each generation
// the population is recorded in HBM (the reference's save_snapshot after
// reproduce, SAVE_CHROM_FRACTION 1.0: LIB/gaboom.cpp:333, :494); at the end
// a job's records are sorted by energy (a radix sort of order-preserving
// keys, stable: records of equal energy keep record order) and thinned by
// remove_dups (LIB/gaboom.cpp:2728-2746: a record is dropped when every gene
// lies within 0.1 of the last record kept), one wave testing 64 records
// (held in registers) against the last kept record at a time.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_select.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "fa_ctx.h"
#include "ga_kernels.h"

struct fa_ga {
    fa_ctx* ctx = nullptr;
    fa_ga_params p{};
    int jobs = 0, Gmax = 0;
    fa::GaDev d{};
    DevBuf rank, gi, gic, res, gi2, gic2, res2, cum, lim, ngenes, ens, cnt, lig, thermo, grand, fit, fit2, share, sig;
    DevBuf dset, dzero, rkey, rval, rzero, csig, cdup, need, ddstats;  // duplicate rejection
    DevBuf skey, sidx;  // survivor ranking: candidates sorted within tiles
    DevBuf prank;       // STEADY: parent ranks among the parents
    DevBuf snap, snapwork;  // snapshot records; sorting and thinning scratch
    int snap_max = 0, snap_gens = 0;  // generations the snapshot holds / has taken
    std::vector<int> Gjob;  // genes of each job
    long long gens_run = 0;   // generations enqueued (bounds the signatures accepted)
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

int ga_blocks(long long n) { return (int)std::max(1LL, (n + 255) / 256); }

// ------------------------------------------------------------ snapshot
// record slot `slot` of every job: its population rows -> records
// [job * cap + slot * pop, +pop)
__global__ void ga_snap_append_kernel(const double* __restrict__ gic, const fa_cf_result* __restrict__ res, int jobs,
                                      int pop, int Gmax, long long cap, int slot, double* __restrict__ sic,
                                      double* __restrict__ se, double* __restrict__ sa)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)jobs * pop * Gmax) return;
    const long long row = i / Gmax;
    const int g = (int)(i % Gmax);
    const long long job = row / pop, r = row % pop;
    const long long dst = job * cap + (long long)slot * pop + r;
    sic[dst * Gmax + g] = gic[row * Gmax + g];
    if (g == 0) {
        se[dst] = res[row].cf;
        sa[dst] = res[row].app;
    }
}

// ascending doubles -> ascending unsigned keys
__global__ void ga_snap_keys_kernel(const double* __restrict__ e, int n, unsigned long long* __restrict__ key,
                                    int* __restrict__ idx)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double v = e[i] == 0.0 ? 0.0 : e[i];  // -0 sorts with +0, as the comparison e_a < e_b has it
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    key[i] = (b >> 63) ? ~b : (b | 0x8000000000000000ull);
    idx[i] = i;
}

// records in sorted order, the job's G genes each
__global__ void ga_snap_gather_kernel(const int* __restrict__ order, int n, int G, int Gmax,
                                      const double* __restrict__ sic, const double* __restrict__ se,
                                      const double* __restrict__ sa, double* __restrict__ oic,
                                      double* __restrict__ oe, double* __restrict__ oa)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)n * G) return;
    const long long k = i / G;
    const int g = (int)(i % G);
    const long long r = order[k];
    oic[i] = sic[r * Gmax + g];
    if (g == 0) {
        oe[k] = se[r];
        oa[k] = sa[r];
    }
}

// remove_dups over sorted records (rows of G <= GM genes): record 0 is
// kept; a record is kept when some gene differs by 0.1 or more from the last
// record kept.  One wave: each lane holds one of 64 consecutive records and
// the last record kept in registers; all test at once, the first that
// differs is kept (its genes move to every lane through readlane) and the
// records after it are tested again.
// The kept records are 0, nxt(0), nxt(nxt(0)), ... with nxt(i) the first
// record after i that differs from it (some gene 0.1 or more apart).  nxt by
// a forward scan per record (at most SNAP_SCAN steps; longer runs of close
// records set *overflow and the one-wave kernel below thins instead); the
// chain from record 0 by pointer doubling: jump[k+1] = jump[k] o jump[k],
// then, from the largest jump down, every marked record marks its jump[k]
// (a record marked during a pass is on the chain as well, so the races of
// a pass only add chain records).
constexpr int SNAP_SCAN = 1024;

__global__ void ga_snap_next_kernel(const double* __restrict__ sic, int n, int G, int* __restrict__ nxt,
                                    int* __restrict__ overflow)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    if (i == n) {
        nxt[n] = n;
        return;
    }
    const double* a = sic + (size_t)i * G;
    int j = i + 1;
    for (int steps = 0; j < n; ++j, ++steps) {
        if (steps == SNAP_SCAN) {
            *overflow = 1;
            break;
        }
        const double* b = sic + (size_t)j * G;
        int flag = 0;
        for (int g = 0; g < G; ++g) flag += fabs(b[g] - a[g]) < 0.1;
        if (flag != G) break;
    }
    nxt[i] = j;
}

__global__ void ga_snap_jump_kernel(const int* __restrict__ from, int n, int* __restrict__ to)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= n) to[i] = from[from[i]];
}

__global__ void ga_snap_mark_kernel(const int* __restrict__ jump, int n, unsigned char* __restrict__ mark)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && mark[i]) {
        const int j = jump[i];
        if (j < n) mark[j] = 1;
    }
}

template <int GM>
__global__ __launch_bounds__(64) void ga_snap_dedup_kernel(const double* __restrict__ sic, int n, int G,
                                                            int* __restrict__ kept, int* __restrict__ n_kept)
{
    const int lane = threadIdx.x;
    if (n <= 0) {
        if (lane == 0) *n_kept = 0;
        return;
    }
    double last[GM], row[GM];
#pragma unroll
    for (int g = 0; g < GM; ++g) last[g] = g < G ? sic[g] : 0.0;
    if (lane == 0) kept[0] = 0;
    int cnt = 1;
    for (int base = 1; base < n; base += 64) {
        const int j = base + lane;
        const bool have = j < n;
#pragma unroll
        for (int g = 0; g < GM; ++g) row[g] = (have && g < G) ? sic[(size_t)j * G + g] : 0.0;
        int start = 0;
        while (true) {
            bool differs = false;
            if (have && lane >= start) {
                int flag = 0;
#pragma unroll
                for (int g = 0; g < GM; ++g)
                    if (g < G) flag += fabs(row[g] - last[g]) < 0.1;
                differs = flag != G;
            }
            const unsigned long long b = __builtin_amdgcn_ballot_w64(differs);
            if (b == 0ull) break;
            const int f = (int)__builtin_ctzll(b);
            if (lane == 0) kept[cnt] = base + f;
            ++cnt;
#pragma unroll
            for (int g = 0; g < GM; ++g) {
                const unsigned long long u = (unsigned long long)__double_as_longlong(row[g]);
                const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, f);
                const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), f);
                last[g] = __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
            }
            start = f + 1;
        }
    }
    if (lane == 0) *n_kept = cnt;
}



int ceil_log2(long long n)
{
    int l = 0;
    while ((1LL << l) < n) ++l;
    return l;
}

// signature set of each job large enough (at most half full) for the
// signatures accepted up to `generations` generations: population + n_new
// per generation; grown by rehashing into a table twice as large or more
int ensure_dset(fa_ga* ga, long long generations)
{
    fa::GaDev& d = ga->d;
    const long long want = 2 * ((long long)d.pop + generations * (long long)d.n_new);
    if (d.dset && (1LL << d.dset_log2) >= want) return FA_SUCCESS;
    const int log2 = std::max(ceil_log2(want), d.dset ? d.dset_log2 + 1 : 0);
    if (log2 > 34) return fa_fail(FA_ERR_LIMIT, "duplicate-rejection table above 2^34 entries per job");
    DevBuf nb;
    int rc = fa_ensure(nb, sizeof(unsigned long long) * ((size_t)ga->jobs << log2));
    if (rc) return rc;
    fa::GaDev nd = d;
    nd.dset = static_cast<unsigned long long*>(nb.p);
    nd.dset_log2 = log2;
    hipError_t e = hipMemsetAsync(nb.p, 0, sizeof(unsigned long long) * ((size_t)ga->jobs << log2), ga->stream);
    if (e == hipSuccess && d.dset) {
        fa::ga_launch_dset_rehash(nd, d.dset, d.dset_log2, ga->stream);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ga->stream);
    if (e != hipSuccess) {
        (void)hipFree(nb.p);
        return fa_fail(FA_ERR_HIP, std::string("duplicate table: ") + hipGetErrorString(e));
    }
    if (ga->dset.p) (void)hipFree(ga->dset.p);
    ga->dset = nb;
    d.dset = nd.dset;
    d.dset_log2 = log2;
    // the captured generation holds the old table's address
    if (ga->exec) (void)hipGraphExecDestroy(ga->exec);
    if (ga->graph) (void)hipGraphDestroy(ga->graph);
    ga->exec = nullptr;
    ga->graph = nullptr;
    return FA_SUCCESS;
}

// one generation of every job, all on ga->stream (capturable)
int ga_generation(fa_ga* ga)
{
    const fa::GaDev& d = ga->d;
    const size_t rows_parents = (size_t)d.jobs * d.pop;
    fa::ga_launch_tick(d, ga->stream);
    if (d.dd_rounds > 0)
        fa::ga_launch_dedup_fill(d, 0, ga->stream);
    else
        fa::ga_launch_breed(d, ga->stream);
    FA_HIP(hipGetLastError());
    const int* lig = static_cast<const int*>(ga->lig.p) + rows_parents;
    int rc = fa_launch_poses(ga->ctx, d.jobs * d.n_new, d.gic + rows_parents * d.Gmax, d.Gmax, lig,
                             d.res + rows_parents, nullptr, 0, nullptr, nullptr, 0, ga->stream);
    if (rc != FA_SUCCESS) return rc;
    fa::ga_launch_ensemble(d, 1, ga->stream);
    fa::ga_launch_select(d, 0, ga->stream);
    FA_HIP(hipGetLastError());
    return FA_SUCCESS;
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
    if (p->reproduction != FA_REP_BOOM && p->reproduction != FA_REP_STEADY)
        return fa_fail(FA_ERR_INVALID, "unknown reproduction model");
    if (p->reproduction == FA_REP_STEADY && p->offspring > p->population)
        return fa_fail(FA_ERR_INVALID, "STEADY reproduction needs offspring <= population");
    const int R = p->dedup_rounds;
    if (R < 0 || R > 1024) return fa_fail(FA_ERR_INVALID, "dedup_rounds in [0, 1024]");
    // candidate rows per job (a batch of initial chromosomes or of children)
    const long long n_cand = R > 0 ? std::max(pop, 2 * ((n_new + 1) / 2)) : 0;
    if ((long long)R * n_cand >= (1LL << 31)) return fa_fail(FA_ERR_LIMIT, "dedup_rounds * batch above 2^31 draws");
    if ((long long)n_jobs * (pop + n_new + n_cand) > (1LL << 30))
        return fa_fail(FA_ERR_LIMIT, "jobs * (population + offspring + candidates) above 2^30");
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
    ++c->live_users;  // released by fa_ga_destroy
    ga->p = *p;
    ga->jobs = n_jobs;
    ga->Gmax = Gmax;
    ga->Gjob = G;
    const size_t J = (size_t)n_jobs, n_all = (size_t)(pop + n_new), rows_all = J * (n_all + (size_t)n_cand);
    auto fail_free = [&](int code) {
        fa_ga_destroy(ga);
        return code;
    };
    int rc;
    if ((rc = fa_ensure(ga->rank, sizeof(int) * J * n_all)) || (rc = fa_ensure(ga->skey, sizeof(unsigned long long) * J * n_all)) ||
        (rc = fa_ensure(ga->sidx, sizeof(int) * J * n_all)) || (rc = fa_ensure(ga->gi, sizeof(int32_t) * rows_all * Gmax)) ||
        (rc = fa_ensure(ga->gic, sizeof(double) * rows_all * Gmax)) ||
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
    if (p->reproduction == FA_REP_STEADY && (rc = fa_ensure(ga->prank, sizeof(int) * J * pop))) return fail_free(rc);
    const int rlog2 = R > 0 ? ceil_log2(2 * n_cand) : 0;
    if (R > 0 &&
        ((rc = fa_ensure(ga->dzero, sizeof(int) * J)) || (rc = fa_ensure(ga->rkey, sizeof(unsigned long long) * (J << rlog2))) ||
         (rc = fa_ensure(ga->rval, sizeof(int) * (J << rlog2))) || (rc = fa_ensure(ga->rzero, sizeof(int) * J)) ||
         (rc = fa_ensure(ga->csig, sizeof(unsigned long long) * J * n_cand)) ||
         (rc = fa_ensure(ga->cdup, sizeof(int) * J * n_cand)) || (rc = fa_ensure(ga->need, sizeof(int) * J)) ||
         (rc = fa_ensure(ga->ddstats, sizeof(unsigned long long) * 2 * J))))
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
    if (e == hipSuccess && R > 0) e = hipMemset(ga->dzero.p, 0, sizeof(int) * J);
    if (e == hipSuccess && R > 0) e = hipMemset(ga->ddstats.p, 0, sizeof(unsigned long long) * 2 * J);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ga->stream, hipStreamNonBlocking);
    if (e != hipSuccess) return fail_free(fa_fail(FA_ERR_HIP, std::string("GA setup: ") + hipGetErrorString(e)));
    fa::GaDev& d = ga->d;
    d = fa::GaDev{};
    d.jobs = n_jobs;
    d.n_act = n_jobs;
    d.act = nullptr;
    d.stage = 0;          // initial populations are drawn straight into the parent rows
    d.job_streams = 1;
    d.seeds = nullptr;
    d.gens = nullptr;
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
    d.rank = static_cast<int*>(ga->rank.p);
    d.skey = static_cast<unsigned long long*>(ga->skey.p);
    d.sidx = static_cast<int*>(ga->sidx.p);
    d.steady = p->reproduction == FA_REP_STEADY ? 1 : 0;
    d.prank = static_cast<int*>(ga->prank.p);
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
    d.dd_rounds = R;
    if (R > 0) {
        d.n_cand = (int)n_cand;
        d.cand = J * n_all;
        d.rtab_log2 = rlog2;
        d.dzero = static_cast<int*>(ga->dzero.p);
        d.rkey = static_cast<unsigned long long*>(ga->rkey.p);
        d.rval = static_cast<int*>(ga->rval.p);
        d.rzero = static_cast<int*>(ga->rzero.p);
        d.csig = static_cast<unsigned long long*>(ga->csig.p);
        d.cdup = static_cast<int*>(ga->cdup.p);
        d.need = static_cast<int*>(ga->need.p);
        d.dd_stats = static_cast<unsigned long long*>(ga->ddstats.p);
        if ((rc = ensure_dset(ga, 16)) != FA_SUCCESS) return fail_free(rc);
    }
    // initial population of every job: draw, evaluate, fold, rank
    if (R > 0)
        fa::ga_launch_dedup_fill(d, 1, ga->stream);
    else
        fa::ga_launch_init(d, ga->stream);
    if ((e = hipGetLastError()) != hipSuccess)
        return fail_free(fa_fail(FA_ERR_HIP, std::string("GA init: ") + hipGetErrorString(e)));
    rc = fa_launch_poses(c, (int)(J * pop), d.gic, Gmax, static_cast<const int*>(ga->lig.p), d.res, nullptr, 0, nullptr,
                         nullptr, 0, ga->stream);
    if (rc == FA_SUCCESS) {
        fa::ga_launch_ensemble(d, 0, ga->stream);
        fa::ga_launch_select(d, 1, ga->stream);
        if ((e = hipGetLastError()) != hipSuccess)
            rc = fa_fail(FA_ERR_HIP, std::string("GA init: ") + hipGetErrorString(e));
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
    if (ga->d.dd_rounds > 0) {
        const int rc = ensure_dset(ga, ga->gens_run + generations);
        if (rc != FA_SUCCESS) return rc;
    }
    ga->gens_run += generations;
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
        if (rc == FA_SUCCESS && ga->snap_gens < ga->snap_max) {
            const fa::GaDev& d = ga->d;
            const long long cap = (long long)ga->snap_max * d.pop;
            double* sic = static_cast<double*>(ga->snap.p);
            double* se = sic + (size_t)ga->jobs * cap * d.Gmax;
            hipLaunchKernelGGL(ga_snap_append_kernel, dim3(ga_blocks((long long)ga->jobs * d.pop * d.Gmax)), dim3(256),
                               0, ga->stream, d.gic, d.res, ga->jobs, d.pop, d.Gmax, cap, ga->snap_gens, sic, se,
                               se + (size_t)ga->jobs * cap);
            const hipError_t e = hipGetLastError();
            if (e != hipSuccess) rc = fa_fail(FA_ERR_HIP, std::string("snapshot: ") + hipGetErrorString(e));
            ++ga->snap_gens;
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
    fa::ga_launch_thermo(ga->d, ga->jobs, ga->p.temperature, static_cast<fa_thermo*>(ga->thermo.p), ga->stream);
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
    fa::ga_launch_thermo(ga->d, J, ga->p.temperature, th, ga->stream);
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

int fa_ga_dedup_stats(fa_ga* ga, uint64_t* stats)
{
    if (!ga || !stats) return fa_fail(FA_ERR_INVALID, "null argument");
    FA_HIP(hipSetDevice(ga->ctx->device));
    FA_HIP(hipStreamSynchronize(ga->stream));
    if (ga->d.dd_rounds < 1) {
        std::fill(stats, stats + 2 * (size_t)ga->jobs, (uint64_t)0);
        return FA_SUCCESS;
    }
    FA_HIP(hipMemcpy(stats, ga->ddstats.p, sizeof(uint64_t) * 2 * ga->jobs, hipMemcpyDeviceToHost));
    return FA_SUCCESS;
}

int fa_ga_snapshot_enable(fa_ga* ga, int max_generations)
{
    if (!ga) return fa_fail(FA_ERR_INVALID, "null GA");
    if (max_generations < 0) return fa_fail(FA_ERR_INVALID, "negative generation count");
    const long long cap = (long long)max_generations * ga->d.pop;
    if (cap >= (1LL << 31)) return fa_fail(FA_ERR_LIMIT, "snapshot above 2^31 records per job");
    FA_HIP(hipSetDevice(ga->ctx->device));
    FA_HIP(hipStreamSynchronize(ga->stream));
    ga->snap_max = 0;
    ga->snap_gens = 0;
    if (cap > 0) {
        const int rc = fa_ensure(ga->snap, sizeof(double) * (size_t)ga->jobs * cap * (ga->d.Gmax + 2));
        if (rc != FA_SUCCESS) return rc;
    }
    ga->snap_max = max_generations;
    return FA_SUCCESS;
}

int fa_ga_snapshot(fa_ga* ga, int job, double* ic, double* evalue, double* app, int32_t* n_raw, int32_t* n_kept)
{
    if (!ga || !n_raw || !n_kept) return fa_fail(FA_ERR_INVALID, "null argument");
    if (job < 0 || job >= ga->jobs) return fa_fail(FA_ERR_INVALID, "job index out of range");
    const long long n = (long long)ga->snap_gens * ga->d.pop;
    *n_raw = (int32_t)n;
    *n_kept = 0;
    if (!ic && !evalue && !app) return FA_SUCCESS;  // the size alone
    if (!ic || !evalue || !app) return fa_fail(FA_ERR_INVALID, "ic, evalue and app together");
    if (n == 0) return FA_SUCCESS;
    FA_HIP(hipSetDevice(ga->ctx->device));
    hipStream_t s = ga->stream;
    const int G = ga->Gjob[job], Gmax = ga->d.Gmax;
    const long long cap = (long long)ga->snap_max * ga->d.pop;
    const double* sic = static_cast<const double*>(ga->snap.p) + (size_t)job * cap * Gmax;
    const double* se = static_cast<const double*>(ga->snap.p) + (size_t)ga->jobs * cap * Gmax + (size_t)job * cap;
    const double* sa = se + (size_t)ga->jobs * cap;
    size_t tmp = 0;
    FA_HIP(rocprim::radix_sort_pairs(nullptr, tmp, (unsigned long long*)nullptr, (unsigned long long*)nullptr,
                                     (int*)nullptr, (int*)nullptr, (unsigned)n, 0, 64, s));
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t bk = al(8 * (size_t)n), bi = al(4 * (size_t)n), bg = al(8 * (size_t)n * G);
    int K = 1;  // jump levels: 2^K > n
    while ((1LL << K) <= n) ++K;
    size_t tsel = 0;
    FA_HIP(rocprim::select(nullptr, tsel, rocprim::counting_iterator<int>(0), (unsigned char*)nullptr, (int*)nullptr,
                           (int*)nullptr, (size_t)n, s));
    tmp = std::max(tmp, tsel);
    const size_t bj = al(4 * (size_t)(n + 1));
    FA_HIP(hipStreamSynchronize(s));  // the scratch may be regrown
    const int rc = fa_ensure(ga->snapwork, 2 * bk + 2 * bi + 2 * bg + 4 * bk + bi + 256 + al(tmp) + (size_t)K * bj +
                                               al((size_t)n));
    if (rc != FA_SUCCESS) return rc;
    char* w = static_cast<char*>(ga->snapwork.p);
    auto take = [&](size_t b) {
        char* q = w;
        w += b;
        return q;
    };
    auto* k0 = reinterpret_cast<unsigned long long*>(take(bk));
    auto* k1 = reinterpret_cast<unsigned long long*>(take(bk));
    int* v0 = reinterpret_cast<int*>(take(bi));
    int* v1 = reinterpret_cast<int*>(take(bi));
    double* tic = reinterpret_cast<double*>(take(bg));
    double* oic = reinterpret_cast<double*>(take(bg));
    double* te = reinterpret_cast<double*>(take(bk));
    double* ta = reinterpret_cast<double*>(take(bk));
    double* oe = reinterpret_cast<double*>(take(bk));
    double* oa = reinterpret_cast<double*>(take(bk));
    int* kept = reinterpret_cast<int*>(take(bi));
    int* nk = reinterpret_cast<int*>(take(256));
    void* t = take(al(tmp));
    int* jump = reinterpret_cast<int*>(take((size_t)K * bj));
    auto* mark = reinterpret_cast<unsigned char*>(take(al((size_t)n)));
    hipLaunchKernelGGL(ga_snap_keys_kernel, dim3(ga_blocks(n)), dim3(256), 0, s, se, (int)n, k0, v0);
    FA_HIP(hipGetLastError());
    FA_HIP(rocprim::radix_sort_pairs(t, tmp, k0, k1, v0, v1, (unsigned)n, 0, 64, s));
    hipLaunchKernelGGL(ga_snap_gather_kernel, dim3(ga_blocks(n * G)), dim3(256), 0, s, v1, (int)n, G, Gmax, sic, se,
                       sa, tic, te, ta);
    // nxt, then the chain from record 0; nk[0] = kept count, nk[1] = overflow
    FA_HIP(hipMemsetAsync(nk, 0, 2 * sizeof(int), s));
    FA_HIP(hipMemsetAsync(mark, 0, (size_t)n, s));
    FA_HIP(hipMemsetAsync(mark, 1, 1, s));
    hipLaunchKernelGGL(ga_snap_next_kernel, dim3(ga_blocks(n + 1)), dim3(256), 0, s, tic, (int)n, G, jump, nk + 1);
    for (int k = 0; k + 1 < K; ++k)
        hipLaunchKernelGGL(ga_snap_jump_kernel, dim3(ga_blocks(n + 1)), dim3(256), 0, s, jump + (size_t)k * (bj / 4),
                           (int)n, jump + (size_t)(k + 1) * (bj / 4));
    for (int k = K - 1; k >= 0; --k)
        hipLaunchKernelGGL(ga_snap_mark_kernel, dim3(ga_blocks(n)), dim3(256), 0, s, jump + (size_t)k * (bj / 4),
                           (int)n, mark);
    FA_HIP(hipGetLastError());
    FA_HIP(rocprim::select(t, tsel, rocprim::counting_iterator<int>(0), mark, kept, nk, (size_t)n, s));
    int ovf = 0;
    FA_HIP(hipMemcpyAsync(&ovf, nk + 1, sizeof(int), hipMemcpyDeviceToHost, s));
    FA_HIP(hipStreamSynchronize(s));
    if (ovf) {  // a run of more than SNAP_SCAN close records: thin in one wave
        if (G <= 16)
            hipLaunchKernelGGL(ga_snap_dedup_kernel<16>, dim3(1), dim3(64), 0, s, tic, (int)n, G, kept, nk);
        else if (G <= 32)
            hipLaunchKernelGGL(ga_snap_dedup_kernel<32>, dim3(1), dim3(64), 0, s, tic, (int)n, G, kept, nk);
        else
            hipLaunchKernelGGL(ga_snap_dedup_kernel<FA_MAX_GENES>, dim3(1), dim3(64), 0, s, tic, (int)n, G, kept, nk);
    }
    FA_HIP(hipGetLastError());
    int m = 0;
    FA_HIP(hipMemcpyAsync(&m, nk, sizeof(int), hipMemcpyDeviceToHost, s));
    FA_HIP(hipStreamSynchronize(s));
    if (m < 1 || m > n) return fa_fail(FA_ERR_INVALID, "snapshot: bad kept count from the device");
    hipLaunchKernelGGL(ga_snap_gather_kernel, dim3(ga_blocks((long long)m * G)), dim3(256), 0, s, kept, m, G, G, tic,
                       te, ta, oic, oe, oa);
    FA_HIP(hipGetLastError());
    FA_HIP(hipMemcpyAsync(ic, oic, sizeof(double) * (size_t)m * G, hipMemcpyDeviceToHost, s));
    FA_HIP(hipMemcpyAsync(evalue, oe, sizeof(double) * (size_t)m, hipMemcpyDeviceToHost, s));
    FA_HIP(hipMemcpyAsync(app, oa, sizeof(double) * (size_t)m, hipMemcpyDeviceToHost, s));
    FA_HIP(hipStreamSynchronize(s));
    *n_kept = m;
    return FA_SUCCESS;
}

int fa_ga_destroy(fa_ga* ga)
{
    if (!ga) return FA_SUCCESS;
    (void)hipSetDevice(ga->ctx->device);
    if (ga->stream) (void)hipStreamSynchronize(ga->stream);
    if (ga->exec) (void)hipGraphExecDestroy(ga->exec);
    if (ga->graph) (void)hipGraphDestroy(ga->graph);
    for (DevBuf* b : {&ga->rank, &ga->gi, &ga->gic, &ga->res, &ga->gi2, &ga->gic2, &ga->res2, &ga->cum, &ga->lim, &ga->ngenes,
                      &ga->ens, &ga->cnt, &ga->lig, &ga->thermo, &ga->grand, &ga->fit, &ga->fit2, &ga->share,
                      &ga->sig, &ga->dset, &ga->dzero, &ga->rkey, &ga->rval, &ga->rzero, &ga->csig, &ga->cdup,
                      &ga->need, &ga->ddstats, &ga->skey, &ga->sidx, &ga->prank, &ga->snap, &ga->snapwork})
        if (b->p) (void)hipFree(b->p);
    if (ga->stream) (void)hipStreamDestroy(ga->stream);
    --ga->ctx->live_users;
    delete ga;
    return FA_SUCCESS;
}

}  // extern "C"
