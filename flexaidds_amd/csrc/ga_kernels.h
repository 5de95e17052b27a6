// ga_kernels.h — device GA state and the launches of one generation
// (internal to the library).
//
// Rows of a GA's buffers are laid out for `jobs` slots: parents of slot k at
// rows [k*pop, (k+1)*pop), then the children area [jobs*pop, jobs*pop +
// jobs*n_new) where the children of the a-th job taking part in a launch
// sit at rows jobs*pop + a*n_new + i.  A launch covers `n_act` jobs; `act`
// maps position a to its slot (null: slot = position), so a pool can run a
// generation over whichever slots are busy.  Initial populations are drawn
// into staged rows (stage + a*pop + c) and ranked from there into the slot.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "flexaidds_amd.h"
#include "lse.h"

namespace fa {

struct GaDev {
    int jobs;              // slots the rows are laid out for
    int pop, n_new, n_all, Gmax;
    int n_act;             // jobs taking part in this launch
    const int* act;        // [n_act] slot of each position, or null (slot = position)
    size_t stage;          // first staged row of initial populations
    int intragenes;
    int job_streams;       // 1: item t of slot k draws from stream (k << 32) | t; 0: stream t
    double cross, mut, beta;
    uint64_t seed;                 // used when seeds is null
    const uint64_t* seeds;         // [jobs] per-slot seed, or null
    unsigned long long* gens;      // [jobs] per-slot generation, or null (shared *gen)
    const int* G;                  // [jobs] genes of each slot
    const double* lmin;            // [jobs * Gmax]
    const double* ldel;
    const double* lbin;
    int32_t* gi;                   // genes of every row
    double* gic;                   // internal coordinates, same rows
    fa_cf_result* res;             // energies of every row
    int* rank;                     // [jobs * n_all] candidate ranks (by launch position)
    unsigned long long* skey;      // [jobs * n_all] energy keys, sorted within tiles of RANK_TILE
    int* sidx;                     // [jobs * n_all] their candidate indices
    int steady;                    // 1: STEADY reproduction (the worst offspring-many parents replaced)
    int* prank;                    // [jobs * pop] parent ranks among the parents (steady)
    int32_t* gi2;                  // [jobs * pop * Gmax] survivors being gathered
    double* gic2;
    fa_cf_result* res2;
    int model;                     // fa_fitness_model
    double alpha, ew;              // sharing exponent, SMFREE entropy weight
    const double* sig;             // [jobs] niche radius
    double* fit;                   // [jobs * pop] fitness of the parent rows
    double* fit2;
    double* share;                 // [jobs * pop] niche counts
    double* cum;                   // [jobs * pop] running sums of the fitness
    LseState* ens;                 // [jobs] ensemble of every evaluated pose
    unsigned long long* n_eval;    // poses folded into ensembles
    unsigned long long* gen;       // shared generation counter (gens == null)
    // Duplicate rejection (dd_rounds > 0).  Candidates are drawn in batches
    // into candidate rows; a batch's launches set cand_mode, t0 (first draw
    // stream of the batch) and dd_round.
    int dd_rounds;                 // batches of candidates per generation at most
    int dd_round;                  // batch of this launch
    int cand_mode;                 // 1: init / breed write candidate rows
    int t0;                        // first draw stream of this batch
    int n_cand;                    // candidate rows per job
    size_t cand;                   // first candidate row
    int dset_log2, rtab_log2;      // table sizes (log2 entries per job)
    unsigned long long* dset;      // [jobs << dset_log2] signatures of every accepted chromosome (0: empty)
    int* dzero;                    // [jobs] signature 0 accepted
    unsigned long long* rkey;      // [jobs << rtab_log2] signatures of this batch (0: empty)
    int* rval;                     // [jobs << rtab_log2] first candidate of the batch holding each
    int* rzero;                    // [jobs] first candidate of the batch with signature 0
    unsigned long long* csig;      // [jobs * n_cand] candidate signatures
    int* cdup;                     // [jobs * n_cand] signature accepted before this batch
    int* need;                     // [jobs] rows still to fill
    unsigned long long* dd_stats;  // [jobs * 2] duplicate candidates drawn, duplicates kept at the cap
};

// Launches on `s`; each covers the n_act jobs of `g` and is skipped when
// n_act is 0.  None allocates or clears memory with a memset node, so the
// sequence can be captured in a HIP graph.
void ga_launch_tick(const GaDev& g, hipStream_t s);
void ga_launch_init(const GaDev& g, hipStream_t s);              // staged initial populations
void ga_launch_breed(const GaDev& g, hipStream_t s);             // children of every job
void ga_launch_ensemble(const GaDev& g, int children, hipStream_t s);
void ga_launch_select(const GaDev& g, int initial, hipStream_t s);  // rank, survivors, fitness
void ga_launch_thermo(const GaDev& g, int n_slots, double T, fa_thermo* out, hipStream_t s);
// Initial population (initial = 1) or children with duplicate rejection:
// up to dd_rounds batches of candidates, each kept in draw order when its
// signature is new, until the job's rows are filled.
void ga_launch_dedup_fill(const GaDev& g, int initial, hipStream_t s);
// Empties the signature sets and counts of the n_act jobs of `g` (a job
// newly admitted into a reused slot).
void ga_launch_dedup_reset(const GaDev& g, hipStream_t s);
// Re-inserts the signatures of `old_keys` (jobs << old_log2 entries) into
// g.dset (cleared beforehand).
void ga_launch_dset_rehash(const GaDev& g, const unsigned long long* old_keys, int old_log2, hipStream_t s);

}  // namespace fa
