// pool.hip — continuous-batching docking pool behind fa_pool_* (C ABI).
//
// A pool keeps `slots` docking jobs resident on one context's device.  Each
// step admits queued jobs into free slots (draws their initial population),
// advances every running job by one GA generation, and scores the running
// jobs' children together with the admitted jobs' initial populations in ONE
// pose launch.  Jobs join and leave between steps, so the device always
// works on whatever requests are in flight.  A job's draws depend only on
// its own seed and generation (stream = item, as a single-job GA), so its
// result is the same alone, in any mix, and as fa_ga with the same seed.
//
// Device memory is allocated once at creation: the number of slots follows
// from the HBM budget divided by one slot's state.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <deque>
#include <string>
#include <vector>

#include "fa_ctx.h"
#include "ga_kernels.h"

namespace fa {

// one admitted job's parameters, scattered into the slot arrays on the device
struct PoolAdmit {
    int slot, G, ligand, pad;
    uint64_t seed;
    double sig;
    double lmin[FA_MAX_GENES], ldel[FA_MAX_GENES], lbin[FA_MAX_GENES];
};

// the outcome of one finished job, written on the device
struct PoolOut {
    fa_thermo thermo;
    fa_cf_result best;
    int32_t best_genes[FA_MAX_GENES];
    double best_ic[FA_MAX_GENES];
};

// the slot arrays an admission writes
struct PoolSlots {
    int Gmax;
    size_t lim_plane;          // S * Gmax: offset between the min, del and bin planes
    double* lim;               // [3][S * Gmax]
    int* G;
    uint64_t* seeds;
    double* sig;
    unsigned long long* gens;
    LseState* ens;
    int* lig;
};

__global__ void pool_admit_kernel(PoolSlots t, const PoolAdmit* adm, int n)
{
    const int b = blockIdx.x;
    if (b >= n) return;
    const PoolAdmit& p = adm[b];
    const int k = p.slot;
    for (int j = threadIdx.x; j < t.Gmax; j += blockDim.x) {
        const size_t at = (size_t)k * t.Gmax + j;
        t.lim[at] = (j < p.G) ? p.lmin[j] : 0.0;
        t.lim[t.lim_plane + at] = (j < p.G) ? p.ldel[j] : 0.0;
        t.lim[2 * t.lim_plane + at] = (j < p.G) ? p.lbin[j] : 0.0;
    }
    if (threadIdx.x == 0) {
        t.G[k] = p.G;
        t.seeds[k] = p.seed;
        t.sig[k] = p.sig;
        t.gens[k] = 0ull;
        t.ens[k] = LseState{0.0, 0.0, 0.0, 0.0};
        t.lig[k] = p.ligand;
    }
}

// ligand of every scored row: children of running position a, then the
// staged initial populations of the admitted jobs
__global__ void pool_lig_kernel(const int* run, int n_run, int n_new, const int* adm, int n_adm, int pop,
                                const int* slot_lig, int* lig)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long nc = (long long)n_run * n_new;
    if (t < nc) {
        lig[t] = slot_lig[run[t / n_new]];
    } else if (t < nc + (long long)n_adm * pop) {
        lig[t] = slot_lig[adm[(t - nc) / pop]];
    }
}

// per finished slot: its lowest-energy survivor (first in row order on ties)
// and its ensemble thermodynamics; one block per finished job
__global__ __launch_bounds__(256) void pool_result_kernel(GaDev g, const int* done, int n_done, double T,
                                                          PoolOut* out)
{
    __shared__ double be[256];
    __shared__ int bi[256];
    const int d = blockIdx.x;
    if (d >= n_done) return;
    const int k = done[d];
    double e = INFINITY;
    int r = g.pop;
    for (int i = threadIdx.x; i < g.pop; i += blockDim.x) {
        const double v = g.res[(size_t)k * g.pop + i].cf;
        if (v < e) {  // strict: the first row of an energy wins; NaN never does
            e = v;
            r = i;
        }
    }
    be[threadIdx.x] = e;
    bi[threadIdx.x] = r;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            const double e2 = be[threadIdx.x + w];
            const int i2 = bi[threadIdx.x + w];
            if (e2 < be[threadIdx.x] || (e2 == be[threadIdx.x] && i2 < bi[threadIdx.x])) {
                be[threadIdx.x] = e2;
                bi[threadIdx.x] = i2;
            }
        }
        __syncthreads();
    }
    const int best = bi[0] < g.pop ? bi[0] : 0;  // every energy NaN: the first row
    const size_t row = (size_t)k * g.pop + best;
    for (int j = threadIdx.x; j < FA_MAX_GENES; j += blockDim.x) {
        out[d].best_genes[j] = (j < g.Gmax) ? g.gi[row * g.Gmax + j] : 0;
        out[d].best_ic[j] = (j < g.Gmax) ? g.gic[row * g.Gmax + j] : 0.0;
    }
    if (threadIdx.x == 0) {
        out[d].best = g.res[row];
        out[d].thermo = thermo_of(g.ens[k], T);
    }
}

}  // namespace fa

struct PoolJob {
    int64_t id;
    int ligand, generations;
    uint64_t seed;
    std::vector<fa_gene_lim> lim;
};

struct PoolSlot {
    bool busy = false;
    bool ready = false;  // initial population drawn (takes part in generations)
    int64_t id = -1;
    int ligand = 0, target = 0, done = 0;
};

struct fa_pool {
    fa_ctx* ctx = nullptr;
    fa_pool_params p{};
    int slots = 0, Gmax = 0;
    size_t slot_bytes = 0;
    fa::GaDev d{};
    DevBuf skey, sidx;  // survivor ranking: candidates sorted within tiles
    DevBuf prank;       // STEADY: parent ranks among the parents
    DevBuf dset, dzero, rkey, rval, rzero, csig, cdup, need, ddstats;  // duplicate rejection, per slot
    DevBuf rank, gi, gic, res, gi2, gic2, res2, cum, fit, fit2, share, sig, lim, ngenes, seeds, gens, ens, cnt, slot_lig,
        lig_rows, run, adm, admrec, done, out;
    hipStream_t stream = nullptr;
    // pinned host staging
    fa::PoolAdmit* h_admrec = nullptr;
    int* h_lists = nullptr;          // [run (slots) | adm (slots) | done (slots)]
    fa::PoolOut* h_out = nullptr;    // [slots]
    std::vector<PoolSlot> slot;
    std::deque<PoolJob> queue;
    std::vector<int> pending_done;   // slots whose results are being copied back
    int64_t next_id = 0;
    bool in_flight = false;
};

namespace {

double nbins(const fa_gene_lim& L)
{
    // set_bins, LIB/gaboom.cpp:2291-2306
    double nbin = (L.max - L.min) / L.del;
    if (nbin - (int)nbin > 0.0) nbin += 1.0;
    if (L.map) nbin += 1.0;
    return nbin;
}

int pool_wait(fa_pool* pl)
{
    if (!pl->in_flight) return FA_SUCCESS;
    FA_HIP(hipStreamSynchronize(pl->stream));
    pl->in_flight = false;
    return FA_SUCCESS;
}

}  // namespace

extern "C" {

int fa_pool_create(fa_ctx* c, const fa_pool_params* p, fa_pool** out)
{
    if (!c || !p || !out) return fa_fail(FA_ERR_INVALID, "null argument");
    *out = nullptr;
    if (!c->have_system) return fa_fail(FA_ERR_NOSYSTEM, "no system uploaded");
    if (p->population < 2 || p->offspring < 1) return fa_fail(FA_ERR_INVALID, "population >= 2 and offspring >= 1");
    if (p->temperature <= 0.0) return fa_fail(FA_ERR_INVALID, "temperature must be positive");
    if (p->fitness < FA_FIT_LINEAR || p->fitness > FA_FIT_SMFREE) return fa_fail(FA_ERR_INVALID, "unknown fitness model");
    if (p->reproduction != FA_REP_BOOM && p->reproduction != FA_REP_STEADY)
        return fa_fail(FA_ERR_INVALID, "unknown reproduction model");
    if (p->reproduction == FA_REP_STEADY && p->offspring > p->population)
        return fa_fail(FA_ERR_INVALID, "STEADY reproduction needs offspring <= population");
    if (p->fitness != FA_FIT_LINEAR && !(p->share_alpha > 0.0))
        return fa_fail(FA_ERR_INVALID, "sharing alpha must be positive");
    if (p->fitness != FA_FIT_LINEAR && !(p->sig_share > 0.0) && !(p->share_peaks > 0.0 && p->share_scale > 0.0))
        return fa_fail(FA_ERR_INVALID, "sharing needs sig_share > 0 or peaks and scale > 0");
    if (p->max_slots < 1) return fa_fail(FA_ERR_INVALID, "max_slots >= 1");
    if (p->dedup_rounds < 0 || p->dedup_rounds > 1024) return fa_fail(FA_ERR_INVALID, "dedup_rounds in [0, 1024]");
    if (p->dedup_rounds > 0 && p->dedup_max_generations < 0)
        return fa_fail(FA_ERR_INVALID, "dedup_max_generations must be >= 0");
    if (p->max_running < 0) return fa_fail(FA_ERR_INVALID, "max_running >= 0");
    int Gmax = p->max_genes;
    if (Gmax <= 0)
        for (int g : c->lig_genes) Gmax = std::max(Gmax, g);
    if (Gmax < 1 || Gmax > FA_MAX_GENES) return fa_fail(FA_ERR_LIMIT, "between 1 and FA_MAX_GENES genes per row");
    const size_t pop = (size_t)p->population, nn = (size_t)p->offspring;
    // duplicate rejection: candidate rows, a batch table and a signature set
    // sized for dedup_max_generations per slot
    const int R = p->dedup_rounds;
    const size_t n_cand = R > 0 ? std::max(pop, 2 * ((nn + 1) / 2)) : 0;
    auto log2_at_least = [](long long x) {
        int l = 0;
        while ((1LL << l) < x) ++l;
        return l;
    };
    const int rlog2 = R > 0 ? log2_at_least(2 * (long long)n_cand) : 0;
    const int slog2 = R > 0 ? log2_at_least(2 * ((long long)pop + (long long)p->dedup_max_generations * (long long)nn)) : 0;
    if (slog2 > 34) return fa_fail(FA_ERR_LIMIT, "duplicate-rejection table above 2^34 entries per slot");
    if ((long long)R * (long long)n_cand >= (1LL << 31)) return fa_fail(FA_ERR_LIMIT, "dedup_rounds * batch above 2^31 draws");
    const size_t dd_bytes = R > 0 ? n_cand * ((size_t)Gmax * (sizeof(int32_t) + sizeof(double)) + sizeof(unsigned long long) + sizeof(int)) +
                                        (sizeof(unsigned long long) + sizeof(int)) * ((size_t)1 << rlog2) +
                                        sizeof(unsigned long long) * ((size_t)1 << slog2) + 32
                                  : 0;
    // one slot's device state: its parent rows, a share of the scoring area
    // (children + a staged initial population), gathered survivors, fitness
    // arrays, limits and small per-slot values
    const size_t row = (size_t)Gmax * (sizeof(int32_t) + sizeof(double)) + sizeof(fa_cf_result);
    const size_t slot_bytes = row * (pop + nn + pop) + row * pop + 4 * sizeof(double) * pop +
                              3 * sizeof(double) * Gmax + 2 * sizeof(int) * (nn + pop) +
                              (sizeof(unsigned long long) + sizeof(int)) * (nn + pop) + sizeof(int) * pop + dd_bytes + 256;
    long long slots = std::min<long long>(p->max_slots, (long long)(p->hbm_budget / slot_bytes));
    if (slots < 1) return fa_fail(FA_ERR_LIMIT, "HBM budget below one job's state");
    if ((long long)slots * (long long)(pop + nn + pop + n_cand) > (1LL << 30)) return fa_fail(FA_ERR_LIMIT, "pool rows above 2^30");
    FA_HIP(hipSetDevice(c->device));
    fa_pool* pl = new fa_pool();
    pl->ctx = c;
    ++c->live_users;  // released by fa_pool_destroy
    pl->p = *p;
    pl->slots = (int)slots;
    pl->Gmax = Gmax;
    pl->slot_bytes = slot_bytes;
    pl->slot.resize(slots);
    const size_t S = (size_t)slots;
    const size_t rows = S * pop + S * (nn + pop) + S * n_cand;  // parents, the scoring area, candidates
    auto fail_free = [&](int code) {
        fa_pool_destroy(pl);
        return code;
    };
    int rc;
    if ((rc = fa_ensure(pl->rank, sizeof(int) * S * (pop + nn))) ||
        (rc = fa_ensure(pl->skey, sizeof(unsigned long long) * S * (pop + nn))) ||
        (rc = fa_ensure(pl->sidx, sizeof(int) * S * (pop + nn))) ||
        (p->reproduction == FA_REP_STEADY && (rc = fa_ensure(pl->prank, sizeof(int) * S * pop))) || (rc = fa_ensure(pl->gi, sizeof(int32_t) * rows * Gmax)) || (rc = fa_ensure(pl->gic, sizeof(double) * rows * Gmax)) ||
        (rc = fa_ensure(pl->res, sizeof(fa_cf_result) * rows)) || (rc = fa_ensure(pl->gi2, sizeof(int32_t) * S * pop * Gmax)) ||
        (rc = fa_ensure(pl->gic2, sizeof(double) * S * pop * Gmax)) || (rc = fa_ensure(pl->res2, sizeof(fa_cf_result) * S * pop)) ||
        (rc = fa_ensure(pl->cum, sizeof(double) * S * pop)) || (rc = fa_ensure(pl->fit, sizeof(double) * S * pop)) ||
        (rc = fa_ensure(pl->fit2, sizeof(double) * S * pop)) || (rc = fa_ensure(pl->share, sizeof(double) * S * pop)) ||
        (rc = fa_ensure(pl->sig, sizeof(double) * S)) || (rc = fa_ensure(pl->lim, 3 * sizeof(double) * S * Gmax)) ||
        (rc = fa_ensure(pl->ngenes, sizeof(int) * S)) || (rc = fa_ensure(pl->seeds, sizeof(uint64_t) * S)) ||
        (rc = fa_ensure(pl->gens, sizeof(unsigned long long) * S)) || (rc = fa_ensure(pl->ens, sizeof(fa::LseState) * S)) ||
        (rc = fa_ensure(pl->cnt, 2 * sizeof(unsigned long long))) || (rc = fa_ensure(pl->slot_lig, sizeof(int) * S)) ||
        (rc = fa_ensure(pl->lig_rows, sizeof(int) * S * (nn + pop))) || (rc = fa_ensure(pl->run, sizeof(int) * S)) ||
        (rc = fa_ensure(pl->adm, sizeof(int) * S)) || (rc = fa_ensure(pl->admrec, sizeof(fa::PoolAdmit) * S)) ||
        (rc = fa_ensure(pl->done, sizeof(int) * S)) || (rc = fa_ensure(pl->out, sizeof(fa::PoolOut) * S)))
        return fail_free(rc);
    if (R > 0 &&
        ((rc = fa_ensure(pl->dset, sizeof(unsigned long long) * (S << slog2))) ||
         (rc = fa_ensure(pl->dzero, sizeof(int) * S)) || (rc = fa_ensure(pl->rkey, sizeof(unsigned long long) * (S << rlog2))) ||
         (rc = fa_ensure(pl->rval, sizeof(int) * (S << rlog2))) || (rc = fa_ensure(pl->rzero, sizeof(int) * S)) ||
         (rc = fa_ensure(pl->csig, sizeof(unsigned long long) * S * n_cand)) ||
         (rc = fa_ensure(pl->cdup, sizeof(int) * S * n_cand)) || (rc = fa_ensure(pl->need, sizeof(int) * S)) ||
         (rc = fa_ensure(pl->ddstats, sizeof(unsigned long long) * 2 * S))))
        return fail_free(rc);
    hipError_t e = hipHostMalloc((void**)&pl->h_admrec, sizeof(fa::PoolAdmit) * S, hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void**)&pl->h_lists, sizeof(int) * 3 * S, hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void**)&pl->h_out, sizeof(fa::PoolOut) * S, hipHostMallocDefault);
    if (e == hipSuccess) e = hipMemset(pl->cnt.p, 0, 2 * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMemset(pl->gi.p, 0, sizeof(int32_t) * rows * Gmax);
    if (e == hipSuccess) e = hipMemset(pl->gic.p, 0, sizeof(double) * rows * Gmax);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&pl->stream, hipStreamNonBlocking);
    if (e != hipSuccess) return fail_free(fa_fail(FA_ERR_HIP, std::string("pool setup: ") + hipGetErrorString(e)));
    fa::GaDev& d = pl->d;
    d = fa::GaDev{};
    d.jobs = (int)slots;
    d.pop = (int)pop;
    d.n_new = (int)nn;
    d.n_all = (int)(pop + nn);
    d.Gmax = Gmax;
    d.intragenes = p->intragenes ? 1 : 0;
    d.job_streams = 0;
    d.cross = p->cross_rate;
    d.mut = p->mut_rate;
    d.beta = 1.0 / (fa::KB_KCAL * p->temperature);
    d.seeds = static_cast<const uint64_t*>(pl->seeds.p);
    d.gens = static_cast<unsigned long long*>(pl->gens.p);
    d.G = static_cast<const int*>(pl->ngenes.p);
    const double* lp = static_cast<const double*>(pl->lim.p);
    d.lmin = lp;
    d.ldel = lp + S * Gmax;
    d.lbin = lp + 2 * S * Gmax;
    d.rank = static_cast<int*>(pl->rank.p);
    d.skey = static_cast<unsigned long long*>(pl->skey.p);
    d.sidx = static_cast<int*>(pl->sidx.p);
    d.steady = p->reproduction == FA_REP_STEADY ? 1 : 0;
    d.prank = static_cast<int*>(pl->prank.p);
    d.dd_rounds = R;
    if (R > 0) {
        d.n_cand = (int)n_cand;
        d.cand = S * pop + S * (nn + pop);
        d.dset_log2 = slog2;
        d.rtab_log2 = rlog2;
        d.dset = static_cast<unsigned long long*>(pl->dset.p);
        d.dzero = static_cast<int*>(pl->dzero.p);
        d.rkey = static_cast<unsigned long long*>(pl->rkey.p);
        d.rval = static_cast<int*>(pl->rval.p);
        d.rzero = static_cast<int*>(pl->rzero.p);
        d.csig = static_cast<unsigned long long*>(pl->csig.p);
        d.cdup = static_cast<int*>(pl->cdup.p);
        d.need = static_cast<int*>(pl->need.p);
        d.dd_stats = static_cast<unsigned long long*>(pl->ddstats.p);
    }
    d.gi = static_cast<int32_t*>(pl->gi.p);
    d.gic = static_cast<double*>(pl->gic.p);
    d.res = static_cast<fa_cf_result*>(pl->res.p);
    d.gi2 = static_cast<int32_t*>(pl->gi2.p);
    d.gic2 = static_cast<double*>(pl->gic2.p);
    d.res2 = static_cast<fa_cf_result*>(pl->res2.p);
    d.model = p->fitness;
    d.alpha = p->share_alpha;
    d.ew = p->entropy_weight;
    d.sig = static_cast<const double*>(pl->sig.p);
    d.fit = static_cast<double*>(pl->fit.p);
    d.fit2 = static_cast<double*>(pl->fit2.p);
    d.share = static_cast<double*>(pl->share.p);
    d.cum = static_cast<double*>(pl->cum.p);
    d.ens = static_cast<fa::LseState*>(pl->ens.p);
    d.n_eval = static_cast<unsigned long long*>(pl->cnt.p);
    d.gen = d.n_eval + 1;
    *out = pl;
    return FA_SUCCESS;
}

int fa_pool_shape(fa_pool* pl, int* slots, uint64_t* slot_bytes, int* max_genes)
{
    if (!pl) return fa_fail(FA_ERR_INVALID, "null pool");
    if (slots) *slots = pl->slots;
    if (slot_bytes) *slot_bytes = pl->slot_bytes;
    if (max_genes) *max_genes = pl->Gmax;
    return FA_SUCCESS;
}

int fa_pool_submit(fa_pool* pl, const fa_job* job, int64_t* id)
{
    if (!pl || !job || !job->limits) return fa_fail(FA_ERR_INVALID, "null argument");
    fa_ctx* c = pl->ctx;
    if (job->ligand < 0 || job->ligand >= c->n_ligands) return fa_fail(FA_ERR_INVALID, "ligand index out of range");
    if (job->generations < 0) return fa_fail(FA_ERR_INVALID, "generations must be >= 0");
    if (pl->p.dedup_rounds > 0 && job->generations > pl->p.dedup_max_generations)
        return fa_fail(FA_ERR_LIMIT, "job longer than the pool's dedup_max_generations");
    const int G = c->lig_genes[job->ligand];
    if (G > pl->Gmax) return fa_fail(FA_ERR_LIMIT, "ligand has more genes than the pool's rows");
    for (int j = 0; j < G; ++j) {
        const fa_gene_lim& L = job->limits[j];
        if (!(L.del > 0.0) || !(L.max >= L.min)) return fa_fail(FA_ERR_INVALID, "gene limits need del > 0 and max >= min");
        const double nb = nbins(L);
        if (!(nb >= 1.0) || nb > 1.0e6) return fa_fail(FA_ERR_LIMIT, "between 1 and 1e6 bins per gene");
    }
    PoolJob pj;
    pj.id = pl->next_id++;
    pj.ligand = job->ligand;
    pj.generations = job->generations;
    pj.seed = job->seed;
    pj.lim.assign(job->limits, job->limits + G);
    pl->queue.push_back(std::move(pj));
    if (id) *id = pl->queue.back().id;
    return FA_SUCCESS;
}

int fa_pool_step(fa_pool* pl, int* n_running, int* n_queued)
{
    if (!pl) return fa_fail(FA_ERR_INVALID, "null pool");
    int rc = pool_wait(pl);
    if (rc) return rc;
    if (!pl->pending_done.empty()) return fa_fail(FA_ERR_INVALID, "collect the finished jobs before the next step");
    FA_HIP(hipSetDevice(pl->ctx->device));
    const int S = pl->slots, Gm = pl->Gmax;
    const fa_pool_params& p = pl->p;
    int* run = pl->h_lists;
    int* adm = pl->h_lists + S;
    int* done = pl->h_lists + 2 * S;
    int n_run = 0, n_adm = 0, n_done = 0;
    for (int k = 0; k < S; ++k)
        if (pl->slot[k].busy && pl->slot[k].ready) run[n_run++] = k;
    // admission: queued jobs into free slots, in arrival order, while the
    // resident count stays within max_running
    int resident = 0;
    for (int k = 0; k < S; ++k) resident += pl->slot[k].busy;
    const int cap = (p.max_running > 0) ? std::min(p.max_running, S) : S;
    for (int k = 0; k < S && !pl->queue.empty() && resident < cap; ++k) {
        if (pl->slot[k].busy) continue;
        ++resident;
        PoolJob j = std::move(pl->queue.front());
        pl->queue.pop_front();
        PoolSlot& sl = pl->slot[k];
        sl.busy = true;
        sl.ready = false;
        sl.id = j.id;
        sl.ligand = j.ligand;
        sl.target = j.generations;
        sl.done = 0;
        fa::PoolAdmit& r = pl->h_admrec[n_adm];
        std::memset(&r, 0, sizeof(r));
        const int G = (int)j.lim.size();
        r.slot = k;
        r.G = G;
        r.ligand = j.ligand;
        r.seed = j.seed;
        double ss = 0.0;
        for (int q = 0; q < G; ++q) {
            r.lmin[q] = j.lim[q].min;
            r.ldel[q] = j.lim[q].del;
            r.lbin[q] = 1.0 / nbins(j.lim[q]);
            ss += (j.lim[q].max - j.lim[q].min) * (j.lim[q].max - j.lim[q].min);
        }
        // niche radius, LIB/gaboom.cpp:300-309, unless given
        r.sig = (p.sig_share > 0.0) ? p.sig_share
                                    : std::sqrt(ss / (double)G) / (2.0 * std::pow(p.share_peaks, 1.0 / (double)G)) /
                                          p.share_scale;
        if (p.fitness != FA_FIT_LINEAR && !(r.sig > 0.0)) r.sig = 1.0;
        adm[n_adm++] = k;
    }
    hipStream_t s = pl->stream;
    const size_t pop = (size_t)p.population, nn = (size_t)p.offspring;
    if (n_adm) {
        FA_HIP(hipMemcpyAsync(pl->admrec.p, pl->h_admrec, sizeof(fa::PoolAdmit) * n_adm, hipMemcpyHostToDevice, s));
        FA_HIP(hipMemcpyAsync(pl->adm.p, adm, sizeof(int) * n_adm, hipMemcpyHostToDevice, s));
        fa::PoolSlots t;
        t.Gmax = Gm;
        t.lim_plane = (size_t)S * Gm;
        t.lim = static_cast<double*>(pl->lim.p);
        t.G = static_cast<int*>(pl->ngenes.p);
        t.seeds = static_cast<uint64_t*>(pl->seeds.p);
        t.sig = static_cast<double*>(pl->sig.p);
        t.gens = static_cast<unsigned long long*>(pl->gens.p);
        t.ens = static_cast<fa::LseState*>(pl->ens.p);
        t.lig = static_cast<int*>(pl->slot_lig.p);
        hipLaunchKernelGGL(fa::pool_admit_kernel, dim3(n_adm), dim3(64), 0, s, t,
                           static_cast<const fa::PoolAdmit*>(pl->admrec.p), n_adm);
    }
    if (n_run) FA_HIP(hipMemcpyAsync(pl->run.p, run, sizeof(int) * n_run, hipMemcpyHostToDevice, s));
    fa::GaDev dr = pl->d;   // running jobs: one generation
    dr.n_act = n_run;
    dr.act = static_cast<const int*>(pl->run.p);
    fa::GaDev da = pl->d;   // admitted jobs: initial population, staged after the children
    da.n_act = n_adm;
    da.act = static_cast<const int*>(pl->adm.p);
    da.stage = (size_t)S * pop + (size_t)n_run * nn;
    fa::ga_launch_tick(dr, s);
    if (p.dedup_rounds > 0) {
        fa::ga_launch_dedup_fill(dr, 0, s);
        fa::ga_launch_dedup_reset(da, s);   // a new job in the slot: an empty set
        fa::ga_launch_dedup_fill(da, 1, s);
    } else {
        fa::ga_launch_breed(dr, s);
        fa::ga_launch_init(da, s);
    }
    const long long n_rows = (long long)n_run * nn + (long long)n_adm * pop;
    if (n_rows > 0) {
        const size_t base = (size_t)S * pop;
        hipLaunchKernelGGL(fa::pool_lig_kernel, dim3((unsigned)((n_rows + 255) / 256)), dim3(256), 0, s,
                           static_cast<const int*>(pl->run.p), n_run, (int)nn, static_cast<const int*>(pl->adm.p), n_adm,
                           (int)pop, static_cast<const int*>(pl->slot_lig.p), static_cast<int*>(pl->lig_rows.p));
        FA_HIP(hipGetLastError());
        rc = fa_launch_poses(pl->ctx, (int)n_rows, pl->d.gic + base * Gm, Gm, static_cast<const int*>(pl->lig_rows.p),
                             pl->d.res + base, nullptr, 0, nullptr, nullptr, 0, s);
        if (rc) return rc;
    }
    fa::ga_launch_ensemble(dr, 1, s);
    fa::ga_launch_select(dr, 0, s);
    fa::ga_launch_ensemble(da, 0, s);
    fa::ga_launch_select(da, 1, s);
    FA_HIP(hipGetLastError());
    // bookkeeping: generations done, finished jobs
    for (int q = 0; q < n_run; ++q) pl->slot[run[q]].done += 1;
    for (int q = 0; q < n_adm; ++q) pl->slot[adm[q]].ready = true;
    for (int k = 0; k < S; ++k)
        if (pl->slot[k].busy && pl->slot[k].ready && pl->slot[k].done >= pl->slot[k].target) done[n_done++] = k;
    if (n_done) {
        FA_HIP(hipMemcpyAsync(pl->done.p, done, sizeof(int) * n_done, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(fa::pool_result_kernel, dim3(n_done), dim3(256), 0, s, pl->d,
                           static_cast<const int*>(pl->done.p), n_done, p.temperature,
                           static_cast<fa::PoolOut*>(pl->out.p));
        FA_HIP(hipGetLastError());
        FA_HIP(hipMemcpyAsync(pl->h_out, pl->out.p, sizeof(fa::PoolOut) * n_done, hipMemcpyDeviceToHost, s));
        pl->pending_done.assign(done, done + n_done);
    }
    pl->in_flight = true;
    int busy = 0;
    for (int k = 0; k < S; ++k) busy += pl->slot[k].busy;
    if (n_running) *n_running = busy - n_done;
    if (n_queued) *n_queued = (int)pl->queue.size();
    return FA_SUCCESS;
}

int fa_pool_collect(fa_pool* pl, fa_job_result* out, int max_out, int* n_out)
{
    if (!pl || !n_out || (max_out > 0 && !out)) return fa_fail(FA_ERR_INVALID, "null argument");
    int rc = pool_wait(pl);
    if (rc) return rc;
    const int n = (int)pl->pending_done.size();
    if (n > max_out) return fa_fail(FA_ERR_LIMIT, "result buffer smaller than the finished jobs");
    const fa_pool_params& p = pl->p;
    for (int q = 0; q < n; ++q) {
        const int k = pl->pending_done[q];
        PoolSlot& sl = pl->slot[k];
        const fa::PoolOut& o = pl->h_out[q];
        fa_job_result& r = out[q];
        std::memset(&r, 0, sizeof(r));
        r.id = sl.id;
        r.ligand = sl.ligand;
        r.generations = sl.done;
        r.n_genes = pl->ctx->lig_genes[sl.ligand];
        r.best = o.best;
        for (int j = 0; j < r.n_genes; ++j) {
            r.best_genes[j] = o.best_genes[j];
            r.best_ic[j] = o.best_ic[j];
        }
        r.thermo = o.thermo;
        r.n_poses = (uint64_t)p.population + (uint64_t)sl.done * (uint64_t)p.offspring;
        sl = PoolSlot{};
    }
    pl->pending_done.clear();
    *n_out = n;
    return FA_SUCCESS;
}

int fa_pool_population(fa_pool* pl, int64_t job_id, int32_t* genes, double* ic, fa_cf_result* results)
{
    if (!pl) return fa_fail(FA_ERR_INVALID, "null pool");
    int rc = pool_wait(pl);
    if (rc) return rc;
    int k = -1;
    for (int q = 0; q < pl->slots; ++q)
        if (pl->slot[q].busy && pl->slot[q].ready && pl->slot[q].id == job_id) k = q;
    if (k < 0) return fa_fail(FA_ERR_INVALID, "job is not resident");
    const size_t pop = (size_t)pl->p.population, Gm = (size_t)pl->Gmax;
    FA_HIP(hipSetDevice(pl->ctx->device));
    if (genes) FA_HIP(hipMemcpy(genes, pl->d.gi + (size_t)k * pop * Gm, sizeof(int32_t) * pop * Gm, hipMemcpyDeviceToHost));
    if (ic) FA_HIP(hipMemcpy(ic, pl->d.gic + (size_t)k * pop * Gm, sizeof(double) * pop * Gm, hipMemcpyDeviceToHost));
    if (results)
        FA_HIP(hipMemcpy(results, pl->d.res + (size_t)k * pop, sizeof(fa_cf_result) * pop, hipMemcpyDeviceToHost));
    return FA_SUCCESS;
}

int fa_pool_set_max_running(fa_pool* pl, int max_running)
{
    if (!pl || max_running < 0) return fa_fail(FA_ERR_INVALID, "bad pool or max_running");
    pl->p.max_running = max_running;
    return FA_SUCCESS;
}

int fa_pool_busy(fa_pool* pl, int* busy)
{
    if (!pl || !busy) return fa_fail(FA_ERR_INVALID, "null argument");
    *busy = 0;
    if (!pl->in_flight) return FA_SUCCESS;
    FA_HIP(hipSetDevice(pl->ctx->device));
    const hipError_t e = hipStreamQuery(pl->stream);
    if (e == hipErrorNotReady) {
        *busy = 1;
        return FA_SUCCESS;
    }
    FA_HIP(e);
    return FA_SUCCESS;
}

int fa_pool_destroy(fa_pool* pl)
{
    if (!pl) return FA_SUCCESS;
    (void)hipSetDevice(pl->ctx->device);
    if (pl->stream) (void)hipStreamSynchronize(pl->stream);
    for (DevBuf* b : {&pl->skey, &pl->sidx, &pl->prank, &pl->dset, &pl->dzero, &pl->rkey, &pl->rval, &pl->rzero,
                      &pl->csig, &pl->cdup, &pl->need, &pl->ddstats, &pl->rank, &pl->gi, &pl->gic, &pl->res, &pl->gi2, &pl->gic2, &pl->res2, &pl->cum, &pl->fit, &pl->fit2,
                      &pl->share, &pl->sig, &pl->lim, &pl->ngenes, &pl->seeds, &pl->gens, &pl->ens, &pl->cnt,
                      &pl->slot_lig, &pl->lig_rows, &pl->run, &pl->adm, &pl->admrec, &pl->done, &pl->out})
        if (b->p) (void)hipFree(b->p);
    if (pl->h_admrec) (void)hipHostFree(pl->h_admrec);
    if (pl->h_lists) (void)hipHostFree(pl->h_lists);
    if (pl->h_out) (void)hipHostFree(pl->h_out);
    if (pl->stream) (void)hipStreamDestroy(pl->stream);
    --pl->ctx->live_users;
    delete pl;
    return FA_SUCCESS;
}

}  // extern "C"
