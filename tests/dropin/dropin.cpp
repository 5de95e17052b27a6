// dropin.cpp — test-only shared object that drives the C++ drop-in layer
// (include/hip_eval.h) the way a reference-side caller does: hip_eval::init /
// upload_system / batch_evaluate (LIB/hip_eval.h:41-166), then
// voronoi_cf_hip::batch_eval with a status column and gene limits
// (LIB/VoronoiCFBatch.h:184-247), then the rocm_dispatch statistical-mechanics
// entries (LIB/UnifiedHardwareDispatch.cpp:768-843, LIB/statmech.cpp:73-173).
//
// Everything runs behind plain extern "C" entries so a test can load the
// object with ctypes.  Exceptions thrown by the drop-in layer come back as
// return code -1 with the exception's type and message in `err`.
#include <cstdio>
#include <cstring>
#include <exception>
#include <span>
#include <stdexcept>
#include <string>
#include <typeinfo>
#include <vector>

#include "hip_eval.h"

namespace {

int fail(const std::exception& e, char* err, int err_len)
{
    if (err && err_len > 0) {
        const char* kind = dynamic_cast<const std::invalid_argument*>(&e) ? "invalid_argument"
                           : dynamic_cast<const std::runtime_error*>(&e)  ? "runtime_error"
                                                                          : "exception";
        std::snprintf(err, (size_t)err_len, "%s: %s", kind, e.what());
    }
    return -1;
}

}  // namespace

extern "C" {

// Outputs of one drop-in session (every pointer is a host array the caller owns).
struct dropin_outputs {
    double* com;          // [pop]  hip_eval::batch_evaluate, three-array form
    double* wal;
    double* sas;
    fa_cf_result* full;   // [pop]  hip_eval::batch_evaluate_full
    fa_cf_result* vcf;    // [pop]  voronoi_cf_hip::batch_eval .cf
    double* vapp;         // [pop]  voronoi_cf_hip::batch_eval .app_evalue
    double wall_ms;       //        voronoi_cf_hip::batch_eval .wall_ms
    fa_thermo thermo;     //        rocm_dispatch::compute
    double lse;           //        rocm_dispatch::log_sum_exp
    double* weights;      // [n_e]  rocm_dispatch::boltzmann_batch
    double log_z;
    double e_min;
    int available;        //        rocm_dispatch::is_available
};

// One session: init on device 0, upload, score ligand `lig` three ways, the
// statistics of `energies` (counts may be NULL), shutdown.  `status` is the
// chromosome status column ('n' rows are skipped), `glim` the gene limits as
// (min, max) pairs per gene (NULL: no clamp).
int dropin_session(const fa_receptor_desc* rec, const fa_params* prm, const fa_ligand_desc* ligs, int n_ligs, int lig,
                   int pop, const double* genes, const char* status, const double* glim, const double* energies,
                   const int* counts, int n_e, double temperature, dropin_outputs* o, char* err, int err_len)
{
    try {
        hip_eval::init(0, pop);
        hip_eval::upload_system(*rec, *prm, ligs, n_ligs);
        const int ng = ligs[lig].n_genes;
        if (lig == 0) hip_eval::batch_evaluate(genes, o->com, o->wal, o->sas, pop, ng, nullptr, 0, 0);
        hip_eval::batch_evaluate_full(genes, o->full, pop, ng, lig);

        std::vector<voronoi_cf_hip::GeneLimit> gl;
        if (glim)
            for (int k = 0; k < ng; ++k) gl.push_back({glim[2 * k + 1], glim[2 * k]});
        std::span<const char> st;
        if (status) st = std::span<const char>(status, (size_t)pop);
        const voronoi_cf_hip::BatchResult br =
            voronoi_cf_hip::batch_eval(std::span<const double>(genes, (size_t)pop * ng), ng, st, gl, lig);
        for (int i = 0; i < pop; ++i) {
            o->vcf[i] = br.cf[(size_t)i];
            o->vapp[i] = br.app_evalue[(size_t)i];
        }
        o->wall_ms = br.wall_ms;

        o->available = rocm_dispatch::is_available() ? 1 : 0;
        const std::span<const double> es(energies, (size_t)n_e);
        o->thermo = rocm_dispatch::compute(es, counts ? std::span<const int>(counts, (size_t)n_e) : std::span<const int>(),
                                           temperature);
        o->lse = rocm_dispatch::log_sum_exp(es);
        const double beta = 1.0 / (0.001987206 * temperature);
        std::vector<double> w;
        const double lz = rocm_dispatch::boltzmann_batch(es, beta, w);
        const rocm_dispatch::BoltzmannBatch bb = rocm_dispatch::boltzmann_batch(es, beta);
        if (bb.log_Z != lz || bb.weights != w) throw std::runtime_error("the two boltzmann_batch forms disagree");
        for (int i = 0; i < n_e; ++i) o->weights[i] = w[(size_t)i];
        o->log_z = bb.log_Z;
        o->e_min = bb.E_min;
        hip_eval::shutdown();
        return 0;
    } catch (const std::exception& e) {
        hip_eval::shutdown();
        return fail(e, err, err_len);
    }
}

// Argument checks of the drop-in layer that fire before any device work:
// case 0 batch_evaluate before upload_system, 1 batch_eval with a gene shape
// that does not divide, 2 batch_eval with gene limits of the wrong length,
// 3 upload_emat with sampled density functions, 4 context() before init.
int dropin_error_case(int which, char* err, int err_len)
{
    try {
        std::vector<double> g(6, 0.0);
        switch (which) {
        case 0: {
            std::vector<fa_cf_result> r(2);
            hip_eval::batch_evaluate_full(g.data(), r.data(), 2, 3, 0);
            break;
        }
        case 1:
            (void)voronoi_cf_hip::batch_eval(std::span<const double>(g.data(), 5), 3, {}, {}, 0);
            break;
        case 2: {
            std::vector<voronoi_cf_hip::GeneLimit> gl(2);
            (void)voronoi_cf_hip::batch_eval(g, 3, {}, gl, 0);
            break;
        }
        case 3: {
            std::vector<float> m(4 * 4 * 8, 0.0f);
            hip_eval::upload_emat(m.data(), 4, 8);
            break;
        }
        case 4:
            (void)hip_eval::context();
            break;
        default:
            return 1;
        }
        return 0;
    } catch (const std::exception& e) {
        return fail(e, err, err_len);
    }
}

int dropin_is_available() { return rocm_dispatch::is_available() ? 1 : 0; }

}  // extern "C"
