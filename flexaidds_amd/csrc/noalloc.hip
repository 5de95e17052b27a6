// noalloc.hip — the pose evaluation for callers that cannot allocate device
// memory (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc).  Its scratch — the launch counters and each resident
// wave's neighbour list and widened-grid cells — lives in one file-scope
// device array of fixed size instead of the context's growable buffers.
// Every call on a GPU shares that GPU's array: calls that overlap on one GPU
// race.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <mutex>
#include <string>

#include "fa_ctx.h"

namespace {

constexpr size_t SCRATCH_BYTES = size_t(128) << 20;  // 128 MiB per GPU, held until the process ends
constexpr size_t HEADER = 256;                        // launch counters (4 x u32), then 256-aligned lists
constexpr int MAX_DEVICES = 64;
constexpr int NEAR_TAIL = 192;                        // the pose kernel's slow-path parking (SLOW_TMP)

static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

// the array's device address on each GPU (a host-side name is no address)
unsigned char* g_ptr[MAX_DEVICES] = {};
std::mutex g_mu;

size_t round256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace

// Looks up the array's address on `device` (current) once; FA_SUCCESS or an
// error code.  Called by fa_ctx_create for the context's device.
int fa_noalloc_init_device(int device)
{
    if (device < 0 || device >= MAX_DEVICES) return fa_fail(FA_ERR_LIMIT, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_ptr[device]) return FA_SUCCESS;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fa_fail(FA_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_ptr[device] = static_cast<unsigned char*>(p);
    return FA_SUCCESS;
}

extern "C" {

size_t fa_noalloc_scratch_bytes(void) { return SCRATCH_BYTES; }

int fa_eval_batch_device_noalloc(fa_ctx* c, int n_poses, const double* genes, int gene_stride, const int* pose_lig,
                                 fa_cf_result* out, void* stream)
{
    if (!c) return fa_fail(FA_ERR_INVALID, "null context");
    if (!c->have_system) return fa_fail(FA_ERR_NOSYSTEM, "no system uploaded");
    if (n_poses < 0 || (n_poses > 0 && (!genes || !out))) return fa_fail(FA_ERR_INVALID, "bad batch");
    if (gene_stride < c->max_genes)
        return fa_fail(FA_ERR_INVALID, "gene_stride below the largest gene count of the uploaded ligands");
    unsigned char* base = (c->device >= 0 && c->device < MAX_DEVICES) ? g_ptr[c->device] : nullptr;
    if (!base) return fa_fail(FA_ERR_INVALID, "scratch address not set up for this device");
    // one resident wave needs its neighbour list and its widened-grid rows
    const size_t r4 = (size_t)((c->rec.n + 3) & ~3);
    const size_t per_wave = (size_t)c->near_cap * (sizeof(int2) + sizeof(float4)) + 2 * r4 * sizeof(int);
    const size_t room = SCRATCH_BYTES - HEADER - 3 * 256;
    const long long fit = (long long)(room / per_wave);
    if (fit < 1)
        return fa_fail(FA_ERR_LIMIT, "one wave's scratch (" + std::to_string(per_wave) + " B for this system) exceeds the " +
                                          std::to_string(SCRATCH_BYTES >> 20) + " MiB static scratch");
    const int cap = c->shape_waves > 0 ? std::min(c->shape_waves, c->n_waves) : c->n_waves;  // launch shape
    const int waves = (int)std::min<long long>(cap, fit);
    fa::DevBatch B{};
    B.n_poses = n_poses;
    B.gene_stride = gene_stride;
    B.genes = genes;
    B.pose_lig = pose_lig;
    B.out = out;
    B.n_ligands = c->n_ligands;
    B.counters = reinterpret_cast<unsigned int*>(base);
    B.totals = static_cast<unsigned long long*>(c->totals.p);  // resident since the context's creation
    B.ops = nullptr;
    B.prof = nullptr;
    size_t off = HEADER;
    B.near_idx = reinterpret_cast<int2*>(base + off);
    off = round256(off + sizeof(int2) * (size_t)waves * c->near_cap);
    B.near_xyz = reinterpret_cast<float4*>(base + off);
    off = round256(off + sizeof(float4) * (size_t)waves * c->near_cap);
    B.rbox = reinterpret_cast<int*>(base + off);
    B.near_cap = c->near_cap;
    B.cell_lists = (c->shape_flags & FA_SHAPE_CELL_LISTS) ? 1 : 0;
    FA_HIP(hipSetDevice(c->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    // the work-queue and overflow counters start at zero (a fill kernel, no memset node)
    FA_HIP(fa::launch_zero_u32(B.counters, 4, s));
    FA_HIP(fa::launch_cf(c->rec, c->d_grids, c->d_ligs, c->prm, B, waves, s));
    return FA_SUCCESS;
}

}  // extern "C"
