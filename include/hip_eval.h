// hip_eval.h — C++ drop-in for the reference's ROCm evaluation API.
//
// namespace hip_eval mirrors LIB/hip_eval.h:41-166 (same names, argument
// meaning, synchronous semantics and exceptions: HIP failures throw
// std::runtime_error as HIP_CHECK does, LIB/hip_eval.hip:77-87).  It adds
// upload_system(), which the real contact function needs and the reference
// header lacks (its kernels compute a placeholder, LIB/hip_eval.hip:241-245).
//
// voronoi_cf_hip::batch_eval mirrors voronoi_cf::batch_eval
// (LIB/VoronoiCFBatch.h:184-247): clamp to gene limits, skip chromosomes with
// status 'n', return one cf record and apparent CF per chromosome plus the
// wall-clock time.
//
// rocm_dispatch:: routes the UnifiedHardwareDispatch entry points the path
// uses (FITNESS_EVAL / PARTITION_FUNC / BOLTZMANN_WEIGHTS, Backend::ROCM = 6,
// LIB/UnifiedHardwareDispatch.h:34-43, .cpp:172-214, :768-843) to the GPU,
// with runtime device detection instead of the compile-time flag (:132-137).
#pragma once

#include <hip/hip_runtime_api.h>  // hipStream_t; host compilers need -D__HIP_PLATFORM_AMD__

#include <cstddef>
#include <span>
#include <string>
#include <vector>

#include "fa_types.h"
#include "flexaidds_amd.h"

namespace hip_eval {

struct DeviceInfo {
    std::string name;
    int compute_units;
    std::size_t global_mem;
    int gcn_arch;
    int wavefront_size;
    int max_threads_per_cu;
    int shared_mem_per_cu;
};

DeviceInfo get_device_info(int device_id = 0);

// Idempotent on the same device; a different device re-initialises.
void init(int device_id = 0, int max_pop = 0);
void shutdown();

// Additive: receptor, scoring parameters and ligand topologies.  Ligand
// `ligand` (default 0) is the one batch_evaluate scores.
void upload_system(const fa_receptor_desc& receptor, const fa_params& params, const fa_ligand_desc* ligands,
                   int n_ligands);

// Type-pair matrix with one value per pair (n_samples == 1, weight = 1).
// Density functions sampled at n_samples > 1 points are rejected.
void upload_emat(const float* emat_sampled, int n_types, int n_samples);

// h_genes: [pop_size x num_genes] internal-coordinate values (to_ic).
// Outputs widened to double.  Synchronous.
void batch_evaluate(const double* h_genes, double* h_com, double* h_wal, double* h_sas, int pop_size, int num_genes,
                    const float* emat_sampled, int n_types, int n_samples);

// Full per-pose results (cfstr terms + status) for callers that need them.
void batch_evaluate_full(const double* h_genes, fa_cf_result* h_out, int pop_size, int num_genes, int ligand = 0);

// Low-level launchers (device pointers): d_genes float [pop x num_genes].
void launch_cf_score_kernel(const float* d_genes, float* d_com, float* d_wal, float* d_sas, int pop_size,
                            int num_genes, hipStream_t stream = hipStreamDefault);
void launch_pack_genes_kernel(const double* d_genes_f64, float* d_genes_f32, int total,
                              hipStream_t stream = hipStreamDefault);

// The context behind the namespace API (for the batched/served layers).
fa_ctx* context();

}  // namespace hip_eval

namespace voronoi_cf_hip {

struct GeneLimit {
    double max;
    double min;
};

struct BatchResult {
    std::vector<fa_cf_result> cf;
    std::vector<double> app_evalue;
    double wall_ms = 0.0;
};

// genes: [n x num_genes]; status (may be empty) — 'n' rows are skipped and
// left zero, as LIB/VoronoiCFBatch.h:221.
BatchResult batch_eval(std::span<const double> genes, int num_genes, std::span<const char> status,
                       std::span<const GeneLimit> gene_lim, int ligand = 0);

}  // namespace voronoi_cf_hip

namespace rocm_dispatch {

bool is_available();  // runtime: a gfx9xx device is present

// statmech::StatMechEngine::compute over (energy, multiplicity)
fa_thermo compute(std::span<const double> energies, std::span<const int> counts, double temperature);

// flexaids::log_sum_exp_dispatch
double log_sum_exp(std::span<const double> values);

// flexaids::compute_boltzmann_batch: unnormalised weights exp(-beta (E - Emin)),
// returns log_Z
double boltzmann_batch(std::span<const double> energies, double beta, std::vector<double>& weights);

// The same with hw::BoltzmannBatchResult's fields (LIB/UnifiedHardwareDispatch.h:
// 124-129, less the telemetry): weights, log_Z = ln sum w - beta E_min, E_min.
// An empty ensemble gives {}, 0, 0 as the reference's early return (:768-772).
struct BoltzmannBatch {
    std::vector<double> weights;
    double log_Z = 0.0;
    double E_min = 0.0;
};
BoltzmannBatch boltzmann_batch(std::span<const double> energies, double beta);

}  // namespace rocm_dispatch
