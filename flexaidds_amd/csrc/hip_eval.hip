// hip_eval.hip — the C++ drop-in layers over the C ABI (include/hip_eval.h).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "hip_eval.h"

namespace {

void check(int rc, const char* what)
{
    if (rc != FA_SUCCESS) throw std::runtime_error(std::string(what) + ": " + fa_last_error());
}

void hip_check(hipError_t e, const char* what)
{
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// process-wide state of the namespace API (the reference keeps a global g_state,
// LIB/hip_eval.hip:438-474; callers serialise access as GPUContextPool does)
struct State {
    fa_ctx* ctx = nullptr;
    int device = -1;
    int ntypes = 0;
    std::vector<int> genes_per_ligand;
    // scratch for the low-level launchers
    double* d_genes64 = nullptr;
    fa_cf_result* d_out = nullptr;
    size_t cap = 0;
};
State g;
std::mutex g_mu;

void ensure_scratch(size_t n_genes_total, size_t pop)
{
    const size_t need = n_genes_total > pop ? n_genes_total : pop;
    if (need <= g.cap) return;
    if (g.d_genes64) (void)hipFree(g.d_genes64);
    if (g.d_out) (void)hipFree(g.d_out);
    g.d_genes64 = nullptr;
    g.d_out = nullptr;
    hip_check(hipMalloc(&g.d_genes64, sizeof(double) * need), "hipMalloc genes");
    hip_check(hipMalloc(&g.d_out, sizeof(fa_cf_result) * need), "hipMalloc results");
    g.cap = need;
}

__global__ void pack_genes_kernel(const double* __restrict__ in, float* __restrict__ out, int total)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) out[i] = (float)in[i];
}

__global__ void widen_genes_kernel(const float* __restrict__ in, double* __restrict__ out, int total)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) out[i] = (double)in[i];
}

__global__ void split_terms_kernel(const fa_cf_result* __restrict__ r, float* com, float* wal, float* sas, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        com[i] = (float)r[i].com;
        wal[i] = (float)r[i].wal;
        sas[i] = (float)r[i].sas;
    }
}

int blocks_for(int n) { return n <= 0 ? 1 : (n + 255) / 256 > 4096 ? 4096 : (n + 255) / 256; }

}  // namespace

namespace hip_eval {

DeviceInfo get_device_info(int device_id)
{
    fa_device_info d;
    check(fa_get_device_info(device_id, &d), "get_device_info");
    return DeviceInfo{d.name, d.compute_units, (std::size_t)d.global_mem, d.gcn_arch, d.wavefront_size,
                      d.max_threads_per_cu, d.shared_mem_per_cu};
}

void init(int device_id, int max_pop)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (g.ctx && g.device == device_id) return;
    if (g.ctx) {
        fa_ctx_destroy(g.ctx);
        g.ctx = nullptr;
    }
    check(fa_ctx_create(device_id, max_pop, &g.ctx), "init");
    g.device = device_id;
}

void shutdown()
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (g.d_genes64) (void)hipFree(g.d_genes64);
    if (g.d_out) (void)hipFree(g.d_out);
    g.d_genes64 = nullptr;
    g.d_out = nullptr;
    g.cap = 0;
    if (g.ctx) fa_ctx_destroy(g.ctx);
    g.ctx = nullptr;
    g.device = -1;
    g.genes_per_ligand.clear();
}

fa_ctx* context()
{
    if (!g.ctx) throw std::runtime_error("hip_eval: init() has not been called");
    return g.ctx;
}

void upload_system(const fa_receptor_desc& receptor, const fa_params& params, const fa_ligand_desc* ligands,
                   int n_ligands)
{
    check(fa_upload_system(context(), &receptor, &params, n_ligands, ligands), "upload_system");
    g.ntypes = params.ntypes;
    g.genes_per_ligand.clear();
    for (int i = 0; i < n_ligands; ++i) g.genes_per_ligand.push_back(ligands[i].n_genes);
}

void upload_emat(const float* emat_sampled, int n_types, int n_samples)
{
    if (!emat_sampled) throw std::invalid_argument("upload_emat: null matrix");
    if (n_samples != 1)
        throw std::invalid_argument("upload_emat: only one value per type pair (weight = 1) is supported");
    if ((long long)n_types * n_types * n_samples > 128LL * 1024)  // LIB/hip_eval.hip:539-552 limit
        throw std::invalid_argument("upload_emat: matrix larger than 128 x 1024 floats");
    check(fa_upload_emat(context(), emat_sampled, n_types), "upload_emat");
}

void batch_evaluate_full(const double* h_genes, fa_cf_result* h_out, int pop_size, int num_genes, int ligand)
{
    if (pop_size < 0) throw std::invalid_argument("batch_evaluate: negative population");
    if (ligand < 0 || ligand >= (int)g.genes_per_ligand.size())
        throw std::invalid_argument("batch_evaluate: ligand index out of range (upload_system first)");
    if (num_genes != g.genes_per_ligand[ligand])
        throw std::invalid_argument("batch_evaluate: num_genes differs from the uploaded gene map");
    check(fa_eval_batch(context(), ligand, pop_size, h_genes, h_out), "batch_evaluate");
}

void batch_evaluate(const double* h_genes, double* h_com, double* h_wal, double* h_sas, int pop_size, int num_genes,
                    const float* emat_sampled, int n_types, int n_samples)
{
    if (emat_sampled) upload_emat(emat_sampled, n_types, n_samples);
    std::vector<fa_cf_result> r((size_t)(pop_size > 0 ? pop_size : 0));
    batch_evaluate_full(h_genes, r.data(), pop_size, num_genes, 0);
    for (int i = 0; i < pop_size; ++i) {
        if (h_com) h_com[i] = r[i].com;
        if (h_wal) h_wal[i] = r[i].wal;
        if (h_sas) h_sas[i] = r[i].sas;
    }
}

void launch_pack_genes_kernel(const double* d_genes_f64, float* d_genes_f32, int total, hipStream_t stream)
{
    if (total <= 0) return;
    hipLaunchKernelGGL(pack_genes_kernel, dim3(blocks_for(total)), dim3(256), 0, stream, d_genes_f64, d_genes_f32,
                       total);
    hip_check(hipGetLastError(), "launch_pack_genes_kernel");
}

void launch_cf_score_kernel(const float* d_genes, float* d_com, float* d_wal, float* d_sas, int pop_size,
                            int num_genes, hipStream_t stream)
{
    if (pop_size <= 0) return;
    if (g.genes_per_ligand.empty() || num_genes != g.genes_per_ligand[0])
        throw std::invalid_argument("launch_cf_score_kernel: num_genes differs from the uploaded gene map");
    const int total = pop_size * num_genes;
    ensure_scratch((size_t)total, (size_t)pop_size);
    hipLaunchKernelGGL(widen_genes_kernel, dim3(blocks_for(total)), dim3(256), 0, stream, d_genes, g.d_genes64,
                       total);
    check(fa_eval_batch_device(context(), pop_size, g.d_genes64, num_genes, nullptr, g.d_out, stream),
          "launch_cf_score_kernel");
    hipLaunchKernelGGL(split_terms_kernel, dim3(blocks_for(pop_size)), dim3(256), 0, stream, g.d_out, d_com, d_wal,
                       d_sas, pop_size);
    hip_check(hipGetLastError(), "launch_cf_score_kernel");
}

}  // namespace hip_eval

namespace voronoi_cf_hip {

BatchResult batch_eval(std::span<const double> genes, int num_genes, std::span<const char> status,
                       std::span<const GeneLimit> gene_lim, int ligand)
{
    if (num_genes <= 0 || genes.size() % (size_t)num_genes) throw std::invalid_argument("batch_eval: bad gene shape");
    if (!gene_lim.empty() && (int)gene_lim.size() != num_genes)
        throw std::invalid_argument("batch_eval: gene_lim size differs from num_genes");
    const int n = (int)(genes.size() / num_genes);
    BatchResult res;
    res.cf.assign((size_t)n, fa_cf_result{});
    res.app_evalue.assign((size_t)n, 0.0);
    std::vector<int> rows;
    std::vector<double> packed;
    for (int i = 0; i < n; ++i) {
        if (!status.empty() && status[i] == 'n') continue;
        rows.push_back(i);
        for (int k = 0; k < num_genes; ++k) {
            double v = genes[(size_t)i * num_genes + k];
            if (!gene_lim.empty()) {
                if (v > gene_lim[k].max) v = gene_lim[k].max;
                if (v < gene_lim[k].min) v = gene_lim[k].min;
            }
            packed.push_back(v);
        }
    }
    auto t0 = std::chrono::steady_clock::now();
    std::vector<fa_cf_result> out(rows.size());
    if (!rows.empty()) hip_eval::batch_evaluate_full(packed.data(), out.data(), (int)rows.size(), num_genes, ligand);
    auto t1 = std::chrono::steady_clock::now();
    for (size_t r = 0; r < rows.size(); ++r) {
        res.cf[rows[r]] = out[r];
        res.app_evalue[rows[r]] = out[r].app;
    }
    res.wall_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    return res;
}

}  // namespace voronoi_cf_hip

namespace rocm_dispatch {

bool is_available()
{
    fa_device_info d;
    return fa_get_device_info(0, &d) == FA_SUCCESS && d.gcn_arch >= 900 && d.gcn_arch < 1000;
}

namespace {
struct DevArr {
    void* p = nullptr;
    ~DevArr()
    {
        if (p) (void)hipFree(p);
    }
};
}  // namespace

fa_thermo compute(std::span<const double> energies, std::span<const int> counts, double temperature)
{
    if (energies.empty()) throw std::runtime_error("compute: empty ensemble");
    if (!counts.empty() && counts.size() != energies.size())
        throw std::invalid_argument("compute: counts size differs from energies");
    fa_thermo th;
    check(fa_thermo_host(hip_eval::context(), energies.data(), counts.empty() ? nullptr : counts.data(),
                         (int64_t)energies.size(), temperature, &th),
          "compute");
    return th;
}

double log_sum_exp(std::span<const double> values)
{
    if (values.empty()) return -HUGE_VAL;
    DevArr x;
    hip_check(hipMalloc(&x.p, sizeof(double) * values.size()), "hipMalloc");
    hip_check(hipMemcpy(x.p, values.data(), sizeof(double) * values.size(), hipMemcpyHostToDevice), "hipMemcpy");
    double out = 0.0;
    check(fa_log_sum_exp_device(hip_eval::context(), static_cast<double*>(x.p), 1, (int64_t)values.size(), &out,
                                nullptr),
          "log_sum_exp");
    return out;
}

double boltzmann_batch(std::span<const double> energies, double beta, std::vector<double>& weights)
{
    weights.assign(energies.size(), 0.0);
    if (energies.empty()) return 0.0;
    constexpr double kB = 0.001987206;
    DevArr e, w;
    hip_check(hipMalloc(&e.p, sizeof(double) * energies.size()), "hipMalloc");
    hip_check(hipMalloc(&w.p, sizeof(double) * energies.size()), "hipMalloc");
    hip_check(hipMemcpy(e.p, energies.data(), sizeof(double) * energies.size(), hipMemcpyHostToDevice), "hipMemcpy");
    double logz = 0.0;
    check(fa_boltzmann_device(hip_eval::context(), static_cast<double*>(e.p), 1, (int64_t)energies.size(),
                              1.0 / (kB * beta), static_cast<double*>(w.p), &logz, nullptr),
          "boltzmann_batch");
    hip_check(hipMemcpy(weights.data(), w.p, sizeof(double) * energies.size(), hipMemcpyDeviceToHost), "hipMemcpy");
    return logz;
}

BoltzmannBatch boltzmann_batch(std::span<const double> energies, double beta)
{
    BoltzmannBatch r;
    if (energies.empty()) return r;
    r.log_Z = boltzmann_batch(energies, beta, r.weights);
    r.E_min = energies[0];
    for (double e : energies)
        if (e < r.E_min) r.E_min = e;
    return r;
}

}  // namespace rocm_dispatch
