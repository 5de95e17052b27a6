/*
 * flexaidds_amd.h — C ABI of the MI355X population-evaluation library.
 *
 * Plain pointers, sizes and int status codes; no C++ or torch types.  Every
 * entry point lists the reference interface it replaces.  The C++ drop-ins
 * (hip_eval::, voronoi_cf::batch_eval, UnifiedHardwareDispatch ROCM routing)
 * are thin layers over these calls (include/hip_eval.h).
 *
 * Ownership: contexts are created and destroyed by the caller; host arrays
 * are borrowed for the duration of a call.  Threading: one context serves one
 * thread at a time; distinct contexts may be used concurrently.
 */
#ifndef FLEXAIDDS_AMD_H
#define FLEXAIDDS_AMD_H

#include <stddef.h>
#include <stdint.h>

#include "fa_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define FA_SUCCESS 0
#define FA_ERR_INVALID (-1)   /* bad argument or malformed system */
#define FA_ERR_HIP (-2)       /* HIP runtime error */
#define FA_ERR_NODEVICE (-3)  /* no usable GPU */
#define FA_ERR_NOSYSTEM (-4)  /* no system uploaded */
#define FA_ERR_LIMIT (-5)     /* a size exceeds a device limit */

typedef struct fa_ctx fa_ctx;

/* Device properties (hip_eval::DeviceInfo, LIB/hip_eval.h:48-56). */
typedef struct fa_device_info {
    char name[256];
    int32_t compute_units;
    uint64_t global_mem;
    int32_t gcn_arch;        /* e.g. 950 */
    int32_t wavefront_size;
    int32_t max_threads_per_cu;
    int32_t shared_mem_per_cu;
} fa_device_info;

/* Replaces hip_eval::get_device_info (LIB/hip_eval.h:60). */
int fa_get_device_info(int device_id, fa_device_info* out);

/* Last error message of the calling thread (never NULL). */
const char* fa_last_error(void);

/* Replaces hip_eval::init (LIB/hip_eval.h:74) and cuda_eval_init's context
 * creation (LIB/cuda_eval.cuh:46-56): opens `device_id` and returns a context
 * with buffers for `max_pop` poses (0 = 4096, as FLEXAIDS_HIP_DEFAULT_MAX_POP,
 * LIB/hip_eval.hip:63-65); buffers grow on demand. */
int fa_ctx_create(int device_id, int max_pop, fa_ctx** out);

/* Replaces hip_eval::shutdown (LIB/hip_eval.h:78) / cuda_eval_shutdown
 * (LIB/cuda_eval.cuh:75). */
int fa_ctx_destroy(fa_ctx* ctx);

/* Additive: the system upload the real contact function needs (the reference
 * hip_eval.h has none; cuda_eval_init carries a partial one,
 * LIB/cuda_eval.cuh:46-56).  Receptor + scoring parameters + `n_ligands`
 * ligand topologies are copied to HBM and stay resident; receptor cell lists
 * are built once per distinct grid.  Replaces any previous system.
 * Lifetime rule: a GA (fa_ga_*) or docking pool (fa_pool_*) built on the
 * context keeps device pointers into the resident system (also inside its
 * captured HIP graph), so while one is alive this call fails with
 * FA_ERR_INVALID and changes nothing; destroy them first.  The context itself
 * must outlive them too. */
int fa_upload_system(fa_ctx* ctx, const fa_receptor_desc* receptor, const fa_params* params, int n_ligands,
                     const fa_ligand_desc* ligands);

/* Replaces hip_eval::batch_evaluate (LIB/hip_eval.h:117-127) and
 * cuda_eval_batch (LIB/cuda_eval.cuh:58-73): synchronous, host buffers.
 * `genes` is [n_poses * n_genes] internal-coordinate values of ligand
 * `ligand` (the `to_ic` doubles the GA packs, LIB/gaboom.cpp:1288-1294);
 * `out` receives one fa_cf_result per pose. */
int fa_eval_batch(fa_ctx* ctx, int ligand, int n_poses, const double* genes, fa_cf_result* out);

/* Device-resident form for batched and served callers: all pointers are
 * device pointers, the call only enqueues work on `stream` (a hipStream_t,
 * NULL = default stream).  `pose_lig` (may be NULL = ligand 0) selects the
 * ligand of each pose; rows of `genes` are `gene_stride` doubles apart.
 * Preconditions: gene_stride >= the largest gene count of the uploaded
 * ligands (checked: FA_ERR_INVALID otherwise); a pose_lig entry outside
 * [0, n_ligands) is not read through but returned as status FA_BAD_LIGAND.
 * No reference counterpart (the reference path is synchronous per call). */
int fa_eval_batch_device(fa_ctx* ctx, int n_poses, const double* genes, int gene_stride, const int* pose_lig,
                         fa_cf_result* out, void* stream);

/* Allocation-free form of fa_eval_batch_device, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  No counterpart in the reference.
 * Same arguments and results as fa_eval_batch_device; it calls no hipMalloc
 * or hipFree: its scratch (launch counters, each resident wave's neighbour
 * list and widened-grid cells) is one static device array of
 * fa_noalloc_scratch_bytes() (128 MiB) per GPU, held until the process ends,
 * so it runs as many resident waves as that array holds for the uploaded
 * system (at most the normal count) and fails with FA_ERR_LIMIT when one
 * wave's scratch does not fit.  Calls that overlap on one GPU race on the
 * array (one call at a time per GPU); fa_last_counters and fa_last_opcount
 * do not see these launches (fa_ctx_totals does). */
int fa_eval_batch_device_noalloc(fa_ctx* ctx, int n_poses, const double* genes, int gene_stride, const int* pose_lig,
                                 fa_cf_result* out, void* stream);
size_t fa_noalloc_scratch_bytes(void);

/* Debug/parity form of fa_eval_batch: also returns the rebuilt ligand
 * coordinates ([n_poses * n_atoms * 3], may be NULL) and each pose's contact
 * records (ca_rec lists of LIB/Vcontacts.cpp:1149-1220, calc-order indices;
 * at most `rec_cap` per pose in `recs` [n_poses * rec_cap], true count in
 * `rec_count`). */
int fa_eval_batch_debug(fa_ctx* ctx, int ligand, int n_poses, const double* genes, fa_cf_result* out,
                        float* lig_xyz, fa_contact_rec* recs, int rec_cap, int* rec_count);

/* Replaces statmech::StatMechEngine::compute (LIB/statmech.cpp:73-173) with
 * the log-sum-exp of flexaids::log_sum_exp_dispatch
 * (LIB/UnifiedHardwareDispatch.cpp:800-843), FP64 on the GPU.  `energy` has
 * element i at energy[i * stride] (stride in doubles, so the `cf` column of an
 * fa_cf_result array can be passed directly); `count` (may be NULL) holds the
 * multiplicities.  Device pointers; `out` is a host pointer; synchronous. */
int fa_thermo_device(fa_ctx* ctx, const double* energy, int stride, const int* count, int64_t n, double temperature,
                     fa_thermo* out, void* stream);

/* Same over host arrays (copies them in). */
int fa_thermo_host(fa_ctx* ctx, const double* energy, const int* count, int64_t n, double temperature, fa_thermo* out);

/* Replaces hip_eval::upload_emat (LIB/hip_eval.h:92) for type-pair matrices
 * with one value per pair (weight = 1): replaces the resident matrix. */
int fa_upload_emat(fa_ctx* ctx, const float* emat, int ntypes);

/* Replaces flexaids::compute_boltzmann_batch (LIB/UnifiedHardwareDispatch.cpp:
 * 768-794, :730-747): weights[i] = exp(-beta E_i - max_j(-beta E_j)) on the GPU
 * in FP64 and log_Z = ln sum exp(-beta E_i).  Device pointers (`log_Z` host,
 * may be NULL); synchronous when `log_Z` is requested. */
int fa_boltzmann_device(fa_ctx* ctx, const double* energy, int stride, int64_t n, double temperature,
                        double* weights, double* log_Z, void* stream);

/* Replaces flexaids::log_sum_exp_dispatch (LIB/UnifiedHardwareDispatch.cpp:
 * 800-843): ln sum exp(x_i), FP64 on the GPU; n = 0 gives -inf.  Device
 * pointer `values`, host `out`; synchronous. */
int fa_log_sum_exp_device(fa_ctx* ctx, const double* values, int stride, int64_t n, double* out, void* stream);

/* Device grand partition function: log_Z[i * stride] and log_c[i] (ln of
 * concentration) for n ligands, device pointers; out (device, n + 2
 * doubles) receives [ln Xi, p_empty, p_0 .. p_{n-1}] (GrandPartitionFunction
 * compute_log_Xi_fresh / binding_probability / empty_probability,
 * LIB/GrandPartitionFunction.cpp:112-160).
 * Asynchronous on `stream`. */
int fa_grand_device(fa_ctx* ctx, const double* log_Z, int stride, const double* log_c, int n, double* out,
                    void* stream);

/* Counters of the last launch: poses that needed a per-pose cell grid, poses
 * that hit a device capacity.  Synchronises the device. */
int fa_last_counters(fa_ctx* ctx, uint32_t* slow_grid, uint32_t* capacity);

/* Poses the device did not score (status FA_CAPACITY, FA_GRID_RANGE,
 * FA_BAD_LIGAND) and poses that needed a widened cell grid, summed over every
 * launch of the context since its creation or the last fa_ctx_reset_totals
 * (single evaluations, GA generations, pool steps alike):
 * totals4 = [slow_grid, capacity, grid_range, bad_ligand].  Synchronises the
 * device.  No counterpart in the reference (it has no device limits). */
int fa_ctx_totals(fa_ctx* ctx, uint64_t* totals4);
/* The same counts and more: the first n of [slow_grid, capacity, grid_range,
 * bad_ligand, undefined_face] (n <= FA_TOTALS_N; undefined_face counts the
 * FA_UNDEFINED_FACE poses).  No counterpart in the reference. */
#define FA_TOTALS_N 5
int fa_ctx_totals_n(fa_ctx* ctx, uint64_t* totals, int n);
int fa_ctx_reset_totals(fa_ctx* ctx);

/* Operation counters (fa_opcount) summed over the last launch; they define
 * the algorithmic FLOP count of that launch.  Synchronises the device. */
int fa_last_opcount(fa_ctx* ctx, fa_opcount* out);

/* Diagnostics (no counterpart in the reference; for profiling the kernel):
 * when enabled, launches sum the clock cycles each wave spends in the
 * kernel's phases (pose build, pre-pass, neighbour cull, first vertex, hull
 * walk, arc points, seeds, face ordering, areas, records) into 16 counters
 * read by fa_last_phase_cycles (synchronises the device). */
/* Changing the setting while a GA or pool of the context is alive fails with
 * FA_ERR_INVALID (their captured graphs hold the profiling buffer). */
int fa_set_phase_profiling(fa_ctx* ctx, int enable);
int fa_last_phase_cycles(fa_ctx* ctx, uint64_t* cycles16);
/* Per-pose clock cycles of the last profiled launch (n <= its poses); the
 * top bit marks poses that needed a widened cell grid.  Diagnostics only:
 * profiling grows a buffer, so it stays off under graph capture. */
int fa_last_pose_cycles(fa_ctx* ctx, int n, uint64_t* cycles);

/* ---------------------------------------------------------------------
 * Device GA (replaces the generation loop of GA(), LIB/gaboom.cpp:346-500,
 * with BOOM reproduction, LIB/gaboom.cpp:886-1036, and LINEAR fitness,
 * LIB/gaboom.cpp:1625-1633).  One generation = roulette selection of parent
 * pairs (roullete_wheel, :1158-1182), bitwise crossover at cross_rate
 * (crossover, :2517-2618), per-bit mutation at mut_rate (mutate, :2620-2641),
 * genes to internal coordinates (genetoic, :2887-2906), evaluation of the
 * offspring, survivors = best `population` of parents + offspring by energy
 * (QuickSort by evalue; ties keep index order; or STEADY replacement,
 * fa_ga_params.reproduction), rank fitness, and the
 * ensemble reduction over every evaluated pose at `temperature`.
 * Random draws come from include/fa_rng.h; duplicate rejection is an option
 * (fa_ga_params.dedup_rounds).
 * ------------------------------------------------------------------- */
typedef struct fa_gene_lim {   /* LIB/gaboom.h:46-54 (genlim) */
    double min, max, del;
    int32_t map;               /* 1 for index genes: one more bin (set_bins, LIB/gaboom.cpp:2291-2306) */
    int32_t pad;
} fa_gene_lim;

/* Fitness models (GB->fitness_model, LIB/gaboom.cpp:1627-1747). */
enum fa_fitness_model {
    FA_FIT_LINEAR = 0,         /* num_chrom - rank */
    FA_FIT_PSHARE = 1,         /* (num_chrom - rank) / niche count; survivors then ordered by fitness */
    FA_FIT_SMFREE = 2          /* rank blended with Boltzmann weight at `temperature`, / niche count */
};

typedef struct fa_ga_params {
    int32_t ligand;            /* ligand of the context to dock */
    int32_t population;        /* survivors per generation (num_chrom) */
    int32_t offspring;         /* children per generation (BOOM: pbfrac * num_chrom) */
    int32_t intragenes;        /* crossover points inside genes */
    double cross_rate;         /* e.g. 0.8 */
    double mut_rate;           /* per bit, e.g. 0.03 */
    double temperature;        /* K, ensemble reduction and SMFREE weights */
    uint64_t seed;
    int32_t fitness;           /* fa_fitness_model */
    int32_t dedup_rounds;      /* 0: duplicate chromosomes allowed (GB->duplicates = 1).  R > 0:
                                  rejected as GB->duplicates = 0 (the reference's default) does,
                                  LIB/gaboom.cpp:971-1018, :1966-1978: a child or initial chromosome
                                  whose signature (hash_genes, :1040-1045) was accepted before in the
                                  job is drawn again; at most R batches of candidates per generation,
                                  then the remaining rows keep duplicates (fa_ga_dedup_stats) */
    double share_alpha;        /* sharing_alpha (reference default 1) */
    double share_peaks;        /* sharing_peaks (5) */
    double share_scale;        /* sharing_scale (10) */
    double sig_share;          /* > 0: niche radius for every job; else derived per job from its
                                  limits as LIB/gaboom.cpp:300-309 does */
    double entropy_weight;     /* SMFREE blend w (0.5) */
    int32_t reproduction;      /* fa_reproduction (GB->rep_model) */
    int32_t pad;
} fa_ga_params;

/* Reproduction models (GB->rep_model, LIB/gaboom.cpp:1022-1032). */
enum fa_reproduction {
    FA_REP_BOOM = 0,           /* survivors = the best `population` of parents + offspring */
    FA_REP_STEADY = 1          /* the worst `offspring` parents (by energy) replaced by the
                                  offspring, whatever their energies; needs offspring <= population */
};

typedef struct fa_ga fa_ga;

/* Allocates the GA on the context's device, draws and evaluates the initial
 * population (generate_random_individual, LIB/gaboom.cpp:1871-1902). */
int fa_ga_create(fa_ctx* ctx, const fa_ga_params* params, const fa_gene_lim* limits, fa_ga** out);
/* Batched form: n_jobs ligands (ligands[k], params->ligand ignored) docked
 * together with the same GA parameters; each generation breeds every job and
 * scores all children in one pose launch.  limits holds each job's gene
 * limits back to back (job k's G_k genes follow job k-1's).  Job k draws
 * from streams (k << 32) | item, so job 0 reproduces the single-job GA.
 * Replaces looping GA() over ligands (LIB/gaboom.cpp:346-500). */
int fa_ga_create_batch(fa_ctx* ctx, const fa_ga_params* params, int n_jobs, const int* ligands,
                       const fa_gene_lim* limits, fa_ga** out);
/* Number of jobs and the row stride (largest gene count) of the population. */
int fa_ga_shape(fa_ga* ga, int* n_jobs, int* max_genes);
/* Runs `generations` generations on the GA's stream and waits.  With
 * use_graph != 0 one generation is captured once as a HIP graph and
 * replayed.  *ms (optional) receives the device time of the run. */
int fa_ga_run(fa_ga* ga, int generations, int use_graph, float* ms);
/* Rows of max_genes values: [0, J*population) the survivors of job 0, 1, ...
 * (population rows each, by increasing energy); then [J*population,
 * J*(population + offspring)) the last generation's children of job 0, 1,
 * ... as bred.  Genes past a job's own count are 0.  Any output may be
 * null.  Synchronous. */
int fa_ga_population(fa_ga* ga, int32_t* genes, double* ic, fa_cf_result* results);
/* Ensemble thermodynamics of each job (out[n_jobs]) over every pose it has
 * evaluated (initial population and all children), FP64 on the device;
 * *n_poses counts the poses of all jobs. */
int fa_ga_thermo(fa_ga* ga, fa_thermo* out, uint64_t* n_poses);
/* Grand partition function over the jobs' ensembles at concentrations
 * conc_M[n_jobs] (mol/L): ln Xi = LSE(0, ln c_k + ln Z_k), p_bound[k] =
 * exp(ln c_k + ln Z_k - ln Xi), p_empty = exp(-ln Xi)
 * (GrandPartitionFunction::log_Xi, binding_probability, empty_probability,
 * LIB/GrandPartitionFunction.h:85-91).  Outputs may be null. */
int fa_ga_grand(fa_ga* ga, const double* conc_M, double* log_xi, double* p_bound, double* p_empty);
int fa_ga_generation(fa_ga* ga, uint64_t* generation);
/* Fitness of the parent rows ([n_jobs * population], rows as in
 * fa_ga_population) as the next generation's roulette uses it, and each
 * job's niche radius (sig_share, [n_jobs]); either may be null. */
int fa_ga_fitness(fa_ga* ga, double* fitness, double* sig_share);
/* Duplicate rejection counts per job since creation: stats[2k] candidates
 * drawn whose signature the job had already accepted, stats[2k + 1] rows
 * that kept a duplicate because dedup_rounds batches did not fill them (the
 * reference draws until they are filled).  Zeros when dedup_rounds is 0. */
int fa_ga_dedup_stats(fa_ga* ga, uint64_t* stats);

/* Snapshot of the GA's populations for the end-of-run ensemble and the
 * clustering (the reference's chrom_snapshot, LIB/gaboom.cpp:333, :494,
 * :517-528).  fa_ga_snapshot_enable records each job's population after
 * every later generation of fa_ga_run (save_snapshot after reproduce,
 * SAVE_CHROM_FRACTION 1.0), for up to max_generations generations (the
 * reference sizes its snapshot for max_generations); it clears earlier
 * records (0 turns recording off).  fa_ga_snapshot sorts job `job`'s
 * records by energy (QuickSort by evalue; records of equal energy keep
 * record order: generation, then population rank) and drops every record
 * whose genes all lie within 0.1 of the last record kept (remove_dups,
 * LIB/gaboom.cpp:2728-2746).  n_raw receives the records taken, n_kept the
 * records kept; ic [n_raw * genes of the job], evalue [n_raw] (cf) and app
 * [n_raw] (host) receive the kept records in order.  With ic, evalue and
 * app all null only n_raw is set.  Synchronous. */
int fa_ga_snapshot_enable(fa_ga* ga, int max_generations);
int fa_ga_snapshot(fa_ga* ga, int job, double* ic, double* evalue, double* app, int32_t* n_raw, int32_t* n_kept);
int fa_ga_destroy(fa_ga* ga);

/* ---------------------------------------------------------------------
 * Binding modes: clustering of docked poses (replaces cluster(),
 * LIB/cluster.cpp:9-176, the reference's default CLUSTA CF, called after
 * the GA at LIB/top.cpp:1350-1353).  The poses, in the order given (the
 * reference passes them by increasing energy), are grouped by repeatedly
 * taking the first pose not yet clustered as the head of a new mode and
 * adding every later unclustered pose whose ligand RMSD to it (no
 * superposition) is at most `rmsd`, until every pose is clustered or
 * max_results modes exist.  Each mode records its head (TOP), its size
 * (FRE) and TCF / ACF: energies at temperature 0; otherwise, with
 * P = exp(-E / T) / sum exp(-E / T) (the reference's beta = 1 / T),
 * TCF = P E - T P ln P of the head and ACF = TCF + sum over members of
 * P E + T P ln P; at T > 0 the first max_results records are then ordered
 * by ACF with the reference's quicksort (records never filled take part as
 * zeros, as there).  Energies are what the reference clusters on
 * (app_evalue, fa_cf_result.app).
 * ------------------------------------------------------------------- */
typedef struct fa_cluster_params {
    double temperature;        /* K (FA->temperature, an integer there); 0: plain energies */
    float rmsd;                /* FA->cluster_rmsd, A (reference default 2.0) */
    int32_t max_results;       /* FA->max_results (default 10), 1..64 */
} fa_cluster_params;

typedef struct fa_cluster_mode {
    int32_t top;               /* pose index of the mode's head (Clus_TOP) */
    int32_t freq;              /* poses in the mode (Clus_FRE) */
    double tcf;                /* Clus_TCF */
    double acf;                /* Clus_ACF */
} fa_cluster_mode;

typedef struct fa_cluster_out {
    int32_t n_modes;           /* records written to `modes` */
    int32_t n_clustered;       /* sum of their freq */
    float device_ms;           /* coordinates and clustering kernels */
    int32_t pad;
} fa_cluster_out;

/* Clusters n_poses poses of `ligand` (genes: [n_poses * gene_stride]
 * internal coordinates, energy: [n_poses]; host arrays) on the context's
 * device.  modes: [max_results]; member (optional, [n_poses]): head pose of
 * each pose, -1 when unclustered; rmsd_head (optional, [n_poses]): RMSD to
 * that head; top_rmsd (optional, [max_results^2], row-major n_modes x
 * n_modes): RMSD between the heads of the output modes (the .cad file's
 * "RMSD between clusters").  Synchronous. */
int fa_cluster_cf(fa_ctx* ctx, int ligand, int n_poses, const double* genes, int gene_stride, const double* energy,
                  const fa_cluster_params* params, fa_cluster_mode* modes, int* member, float* rmsd_head,
                  float* top_rmsd, fa_cluster_out* out);

/* ---------------------------------------------------------------------
 * Super-cluster of pose energies: FastOPTICS over the energies as
 * one-dimensional points with SUPER_CLUSTER_ONLY extraction, the filter of
 * BindingPopulation::get_super_cluster_ensemble (LIB/BindingMode.cpp:153-180;
 * FastOPTICS::FastOPTICS, computeOrdering, coreDist and extractSuperCluster,
 * LIB/fast_optics.cpp:8-135).
 */
#define FA_SUPER_CLUSTER_MAX 8192

typedef struct fa_super_cluster_out {
    int32_t n_members;         /* indices written to `members` */
    int32_t seed;              /* the fill's first point: the ordering entry of least (reach, index) */
    float device_ms;           /* ranking, core distances, ordering and fill */
    int32_t pad;
} fa_super_cluster_out;

/* energy: [n] host, finite; min_pts: FastOPTICS minPts (the binding
 * population passes max(4, n / 20)); members: [n] host, receives the
 * super-cluster's point indices in the flood fill's visiting order.
 * n <= FA_SUPER_CLUSTER_MAX (FA_ERR_LIMIT above).  Synchronous. */
int fa_super_cluster(fa_ctx* ctx, const double* energy, int n, int min_pts, int* members, fa_super_cluster_out* out);

/* ---------------------------------------------------------------------
 * Docking pool: continuous batching of docking jobs on one context.
 * Replaces running GA() once per request (LIB/gaboom.cpp:346-500) with a
 * resident pool that advances every in-flight job by one generation per
 * step and scores all of their children, plus the initial populations of
 * jobs admitted at that step, in one pose launch.  Jobs are admitted in
 * arrival order while slots are free; the slot count is fixed at creation
 * from the HBM budget.  A job's result equals fa_ga's for the same seed and
 * parameters, whatever else shares the pool.  One pool per context and
 * thread; pools of different contexts may be stepped back to back so their
 * work overlaps on the device.
 * ------------------------------------------------------------------- */
typedef struct fa_pool_params {
    int32_t population;        /* survivors per generation, every job */
    int32_t offspring;         /* children per generation */
    int32_t intragenes;
    int32_t fitness;           /* fa_fitness_model */
    double cross_rate;
    double mut_rate;
    double temperature;        /* K */
    double share_alpha, share_peaks, share_scale, sig_share, entropy_weight;  /* as fa_ga_params */
    int32_t max_slots;         /* upper bound on resident jobs */
    int32_t max_genes;         /* row width; 0 = the largest gene count of the context's ligands */
    uint64_t hbm_budget;       /* bytes of job state; slots = min(max_slots, budget / slot bytes) */
    int32_t max_running;       /* at most this many resident jobs (0: every slot); a smaller
                                  batch per step finishes each job sooner */
    int32_t reproduction;      /* fa_reproduction, as fa_ga_params */
    int32_t dedup_rounds;      /* as fa_ga_params (0: duplicates allowed) */
    int32_t dedup_max_generations;  /* with dedup_rounds > 0: the longest job a slot's signature
                                       set is sized for; longer jobs are refused at submission */
} fa_pool_params;

typedef struct fa_job {
    int32_t ligand;            /* ligand of the context */
    int32_t generations;       /* generations after the initial population */
    uint64_t seed;
    const fa_gene_lim* limits; /* [genes of the ligand]; copied at submission */
} fa_job;

typedef struct fa_job_result {
    int64_t id;
    int32_t ligand;
    int32_t generations;
    int32_t n_genes;
    int32_t pad;
    fa_cf_result best;                   /* lowest-energy survivor */
    int32_t best_genes[FA_MAX_GENES];
    double best_ic[FA_MAX_GENES];
    fa_thermo thermo;                    /* ensemble of every pose the job evaluated */
    uint64_t n_poses;
} fa_job_result;

typedef struct fa_pool fa_pool;

int fa_pool_create(fa_ctx* ctx, const fa_pool_params* params, fa_pool** out);
/* Resident job slots, device bytes per slot, row width. */
int fa_pool_shape(fa_pool* pool, int* slots, uint64_t* slot_bytes, int* max_genes);
/* Queues a job (host side; no device work). */
int fa_pool_submit(fa_pool* pool, const fa_job* job, int64_t* id);
/* Waits for the previous step, admits queued jobs into free slots and
 * enqueues one generation of every resident job; returns without waiting.
 * *n_running: jobs still resident after this step; *n_queued: waiting. */
int fa_pool_step(fa_pool* pool, int* n_running, int* n_queued);
/* Waits for the last step and returns the jobs it finished (up to max). */
int fa_pool_collect(fa_pool* pool, fa_job_result* out, int max, int* n);
/* Survivors of a resident job (population rows of max_genes values). */
int fa_pool_population(fa_pool* pool, int64_t job_id, int32_t* genes, double* ic, fa_cf_result* results);
/* Changes the cap on resident jobs for later admissions (0: every slot). */
int fa_pool_set_max_running(fa_pool* pool, int max_running);
/* *busy = 1 while the last step still runs on the device, else 0; never
 * waits (lets a host loop step each pool as soon as its own step ends).
 * No counterpart in the reference (its GA loop is synchronous). */
int fa_pool_busy(fa_pool* pool, int* busy);
int fa_pool_destroy(fa_pool* pool);

/* ---------------------------------------------------------------------
 * Served stream: the request loop over resident receptors' pools, in C++.
 * No counterpart in the reference (it docks one ligand per process run);
 * its TargetServer (LIB/TargetServer.h) is what receives the results.
 * Requests are admitted in arrival order once the wall clock passes their
 * arrival time (or at once when realtime is 0) into their receptor's pool;
 * each pool is stepped as soon as its own previous step has ended, so the
 * pools' launches overlap on the device.  out[i] receives the i-th request
 * to finish: its index, its job result and its latency (completion minus
 * arrival, from the loop's start when not realtime).
 * ------------------------------------------------------------------- */
typedef struct fa_request {
    double t;                  /* arrival, seconds after the loop starts */
    int32_t pool;              /* index into the pools array (the receptor) */
    int32_t pad;
    fa_job job;                /* ligand of that pool's context, generations, seed, limits */
} fa_request;

typedef struct fa_served {
    int64_t request;           /* index into the requests array */
    double latency_s;
    fa_job_result result;
} fa_served;

typedef struct fa_serve_stats {
    double wall_s;
    int64_t steps;             /* pool steps taken */
    uint64_t poses;            /* poses the finished jobs evaluated */
} fa_serve_stats;

int fa_serve_run(fa_pool* const* pools, int n_pools, const fa_request* requests, int n_requests, int realtime,
                 fa_served* out, fa_serve_stats* stats);

/* Size of the kernel's per-wave LDS workspace (bytes). */
int fa_kernel_lds_bytes(void);

/* Waves a pose launch of this context runs (one per resident slot on the
 * device, each looping over poses); 0 before a system is uploaded.  Launches
 * of many times this many poses keep every slot busy to the end. */
int fa_ctx_launch_waves(fa_ctx* ctx);

/* Launch shape of the context's pose launches (every entry point that scores
 * poses, the GA and the pool included): `waves` caps the resident waves
 * (0 = one per resident slot), flags select variants of the kernel's
 * neighbour scan.  A wave cap leaves every result unchanged; the scan
 * variants find the same neighbours in the same order, so they differ only in
 * the summation order of the clash pre-pass (the wall of an FA_CLASH pose, in
 * its last bits).
 * flexaidds_amd.autotune picks a shape per workload by timing the
 * candidates.  No counterpart in the reference (its CPU loop and its GPU
 * placeholder have compile-time launch constants, LIB/hip_eval.hip:57-71).
 * fa_ctx_get_launch_shape reports the wave count a launch uses. */
#define FA_SHAPE_CELL_LISTS 1  /* scan the receptor cell lists instead of the per-cell neighbourhood lists */
int fa_ctx_set_launch_shape(fa_ctx* ctx, int waves, int flags);
int fa_ctx_get_launch_shape(fa_ctx* ctx, int* waves, int* flags);

#ifdef __cplusplus
}
#endif

#endif /* FLEXAIDDS_AMD_H */
