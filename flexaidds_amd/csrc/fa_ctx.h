// fa_ctx.h — the context behind the C ABI (internal to the library).
//
// A context owns one GPU's resident state: the receptor (calc order), one
// receptor cell list per distinct grid, the ligand topologies, the scoring
// parameters, growable pose buffers and the pose kernel's per-wave scratch.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "fa_device.h"
#include "flexaidds_amd.h"

namespace fa {
hipError_t launch_cf(const DevReceptor& R, const DevGrid* grids, const DevLigand* ligs, const DevParams& prm,
                     const DevBatch& B, int n_blocks, hipStream_t stream);
int cf_waves_per_cu();
int smem_bytes();
hipError_t launch_coords(const DevLigand* ligs, int ligand, int n_atoms, const double* genes, int stride, int n,
                         float* xyzT, float* icT, hipStream_t stream);
size_t lse_scratch_bytes();
hipError_t launch_thermo(const double* d_energy, int stride, const int* d_count, long long n, double T,
                         void* d_scratch, fa_thermo* d_out, hipStream_t stream);
hipError_t launch_boltzmann(const double* d_energy, int stride, long long n, double T, void* d_scratch, double* d_w,
                            hipStream_t stream);
hipError_t launch_lse_values(const double* d_x, int stride, long long n, void* d_scratch, hipStream_t stream);
hipError_t launch_zero_u32(unsigned int* p, int n_words, hipStream_t stream);
hipError_t launch_grand(const double* d_log_Z, int stride, const double* d_log_c, int n, double* d_out,
                        hipStream_t stream);
}  // namespace fa

// error state of the calling thread (fa_last_error)
int fa_fail(int code, const std::string& msg);

// static scratch of the allocation-free evaluation: its address on `device`
// (the current device), looked up once (noalloc.hip)
int fa_noalloc_init_device(int device);

#define FA_HIP(expr)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fa_fail(FA_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct DevBuf {
    void* p = nullptr;
    size_t n = 0;
};

// grows a device buffer to at least `bytes` (contents are not kept)
int fa_ensure(DevBuf& b, size_t bytes);

struct fa_ctx {
    int device = 0;
    int n_cu = 0;
    int max_pop = 0;
    // resident system
    std::vector<void*> allocs;
    fa::DevReceptor rec{};
    fa::DevParams prm{};
    fa::DevGrid* d_grids = nullptr;
    fa::DevLigand* d_ligs = nullptr;
    std::vector<fa::DevLigand> h_ligs;
    std::vector<int> lig_genes;
    int n_ligands = 0;
    int max_lig_atoms = 0;
    bool have_system = false;
    // pose-kernel grid: one resident wave per slot, each with its own
    // neighbour-list scratch of near_cap entries
    int n_waves = 0;
    int shape_waves = 0;   // launch shape (fa_ctx_set_launch_shape): wave cap, 0 = every resident slot
    int shape_flags = 0;   //   FA_SHAPE_* flags
    int near_cap = 0;
    // pose buffers
    DevBuf genes, out, counters, recs, reccnt, xyz, scratch, thermo, ops, prof, near_idx, near_xyz, rbox;
    DevBuf totals;  // FA_TOTALS_N x u64 unscored-pose counts since the last reset (not cleared per launch)
    DevBuf cluster; // pose clustering (fa_cluster_cf): coordinates, inputs, per-pose and per-mode arrays
    int max_genes = 0;  // largest gene count of the uploaded ligands
    int n_sc = 0;       // scorable side-chain atoms of the resident receptor
    int prof_poses = 0;  // poses of the last profiled launch
    bool profile = false;
    std::vector<uint32_t> last_counters = std::vector<uint32_t>(4, 0);
    // GAs and pools built on this context: their buffers and captured graphs
    // hold device pointers into the resident system, so while any is alive
    // the system may not be replaced (fa_upload_system) nor the profiling
    // buffers regrown (fa_set_phase_profiling)
    int live_users = 0;

    void free_system()
    {
        for (void* p : allocs) (void)hipFree(p);
        allocs.clear();
        d_grids = nullptr;
        d_ligs = nullptr;
        h_ligs.clear();
        lig_genes.clear();
        have_system = false;
    }
};

// One pose-kernel launch on `stream` over device genes/results (and optional
// debug outputs).  Counters are cleared with a kernel, so the launch can be
// captured in a HIP graph.
int fa_launch_poses(fa_ctx* c, int n_poses, const double* d_genes, int stride, const int* d_lig,
                    fa_cf_result* d_out, fa_contact_rec* d_recs, int rec_cap, int* d_reccnt, float* d_xyz,
                    int xyz_stride, hipStream_t stream);
