// cluster.hip — clustering of docked poses into binding modes on the device
// (fa_cluster_cf, C ABI).
//
// The reference's default post-GA clustering (CLUSTA CF: cluster(),
// LIB/cluster.cpp:9-176) takes the poses in energy order and repeatedly
// makes the first pose not yet clustered the head of a new cluster, adding
// every later unclustered pose within cluster_rmsd of it (RMSD over the
// ligand atoms, no superposition), until every pose is clustered or
// max_results clusters exist.  Each cluster records its head (TOP), size
// (FRE) and two energies (TCF, ACF): plain energies at T = 0, otherwise
// P E - T P ln P of the head and P E + T P ln P per member with
// P = exp(-E / T) / sum_j exp(-E_j / T) (FA->beta = 1 / T, no Boltzmann
// constant).  At T > 0 the first max_results cluster records are then
// ordered by ACF with the reference's quicksort (QuickSort_Clusters,
// :323-359; records never filled take part as zeros).
//
// Device form: the poses' ligand coordinates are built once (the pose
// kernel's build, a thread per pose, tiles of 256 poses); each cluster is one pass over the
// poses (a thread per pose) against its head's coordinates in LDS, which
// also finds the next head (the first pose it leaves unclustered); a small
// kernel folds the pass's member terms in block order and starts the next
// pass.  The squared distance of two poses is summed in the lane order of
// the reference's AVX2 build (flexaids::sum_sq_distances_f,
// LIB/simd_distance.h:693-710: eight fused multiply-add accumulators, the
// same horizontal sum, a fused tail).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "fa_ctx.h"

namespace fa {
namespace {

constexpr int CL_BLOCK = 256;
constexpr int NO_HEAD = 0x7fffffff;

struct ClDev {
    int n, n_atoms, max_results;
    const float* xyzT;        // 3 * n_atoms rows per pose, tiled (coord_at)
    const double* energy;     // [n]
    double T, beta;           // temperature (K) and 1 / T; T = 0: plain energies
    float rmsd;
    double* z;                // partition function sum
    int* member;              // [n] head pose of each pose, -1 unclustered
    float* rmsd_head;         // [n]
    double* part;             // [blocks] per-block sums of a pass
    int* pcount;              // [blocks]
    int* head;                // [max_results + 1]: head of each pass (NO_HEAD: none)
    int* next;                // first pose a pass leaves unclustered (atomicMin)
    int* top;                 // [max(max_results, 1)] records
    int* fre;
    double* tcf;
    double* acf;
    int* n_clusters;
};

// flexaids::sum_sq_distances_f in its AVX2 form
__device__ __forceinline__ float sum_sq_avx2(const ClDev& c, const float* __restrict__ hx, int pose)
{
    const int nf = 3 * c.n_atoms;
    float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    int i = 0;
    for (; i + 7 < nf; i += 8) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const float d = c.xyzT[coord_at(nf, i + q, pose)] - hx[i + q];
            acc[q] = __builtin_fmaf(d, d, acc[q]);
        }
    }
    // hi + lo, then (0 + 2, 1 + 3), then their sum
    const float s0 = acc[0] + acc[4], s1 = acc[1] + acc[5], s2 = acc[2] + acc[6], s3 = acc[3] + acc[7];
    const float t0 = s0 + s2, t1 = s1 + s3;
    float sum = t0 + t1;
    for (; i < nf; ++i) {
        const float d = c.xyzT[coord_at(nf, i, pose)] - hx[i];
        sum = __builtin_fmaf(d, d, sum);
    }
    return sum;
}

__device__ __forceinline__ double prob(const ClDev& c, double e) { return exp(-c.beta * e) / *c.z; }

__global__ __launch_bounds__(CL_BLOCK) void cl_init_kernel(ClDev c)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < c.n) {
        c.member[i] = i == 0 ? 0 : -1;
        c.rmsd_head[i] = 0.f;
    }
    if (i < c.max_results) {
        c.top[i] = 0;
        c.fre[i] = 0;
        c.tcf[i] = 0.0;
        c.acf[i] = 0.0;
    }
    if (i <= c.max_results) c.head[i] = (i == 0 && c.n > 0) ? 0 : NO_HEAD;
    if (i == 0) {
        *c.next = NO_HEAD;
        *c.n_clusters = 0;
        *c.z = 0.0;
    }
}

// sum_j exp(-beta E_j): per-block tree sums, then their fold in block order
// (the reference sums in pose order; equal to rounding)
__global__ __launch_bounds__(CL_BLOCK) void cl_z_kernel(ClDev c)
{
    __shared__ double sh[CL_BLOCK];
    const int i = blockIdx.x * CL_BLOCK + threadIdx.x;
    sh[threadIdx.x] = i < c.n ? exp(-c.beta * c.energy[i]) : 0.0;
    __syncthreads();
    for (int w = CL_BLOCK / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) c.part[blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(CL_BLOCK) void cl_z_final_kernel(ClDev c, int n_blocks)
{
    __shared__ double sh[CL_BLOCK];
    double v = 0.0;
    for (int b = threadIdx.x; b < n_blocks; b += CL_BLOCK) v += c.part[b];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int w = CL_BLOCK / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *c.z = sh[0];
}

// pass r: every unclustered pose after the head within the RMSD joins it
__global__ __launch_bounds__(CL_BLOCK) void cl_assign_kernel(ClDev c, int r)
{
    __shared__ float hx[3 * FA_MAX_LIG_ATOMS];
    __shared__ double sh[CL_BLOCK];
    __shared__ int shn[CL_BLOCK];
    const int j = c.head[r];
    if (j == NO_HEAD) return;  // the whole grid: no cluster left
    const int nf = 3 * c.n_atoms;
    for (int k = threadIdx.x; k < nf; k += CL_BLOCK) hx[k] = c.xyzT[coord_at(nf, k, j)];
    __syncthreads();
    __shared__ int wfirst[CL_BLOCK / 64];
    const int i = blockIdx.x * CL_BLOCK + threadIdx.x;
    double v = 0.0;
    int joined = 0;
    bool left = false;  // unclustered after this pass
    if (i > j && i < c.n && c.member[i] == -1) {
        const float d = sum_sq_avx2(c, hx, i);
        const float rm = sqrtf(d / (float)c.n_atoms);
        if (rm <= c.rmsd) {
            c.member[i] = j;
            c.rmsd_head[i] = rm;
            if (c.T > 0.0) {
                const double p = prob(c, c.energy[i]);
                v = (p * c.energy[i]) + (c.T * p * log(p));
            } else {
                v = c.energy[i];
            }
            joined = 1;
        } else {
            left = true;
        }
    }
    // the first pose of the block left unclustered (poses increase with
    // the thread), one atomic per block for the next head
    const unsigned long long m = __builtin_amdgcn_ballot_w64(left);
    if ((threadIdx.x & 63) == 0) wfirst[threadIdx.x >> 6] = m ? (int)(blockIdx.x * CL_BLOCK + (threadIdx.x & ~63) + __ffsll((long long)m) - 1) : NO_HEAD;
    sh[threadIdx.x] = v;
    shn[threadIdx.x] = joined;
    __syncthreads();
    if (threadIdx.x == 0) {
        int f = NO_HEAD;
        for (int w = 0; w < CL_BLOCK / 64; ++w) f = min(f, wfirst[w]);
        if (f != NO_HEAD) atomicMin(c.next, f);
    }
    for (int w = CL_BLOCK / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            sh[threadIdx.x] += sh[threadIdx.x + w];
            shn[threadIdx.x] += shn[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        c.part[blockIdx.x] = sh[0];
        c.pcount[blockIdx.x] = shn[0];
    }
}

// the pass's record from its block sums (in block order), and the next head
__global__ __launch_bounds__(CL_BLOCK) void cl_finish_kernel(ClDev c, int r, int n_blocks)
{
    __shared__ double sh[CL_BLOCK];
    __shared__ int shn[CL_BLOCK];
    const int j = c.head[r];
    if (j == NO_HEAD) return;
    double v = 0.0;
    int cnt = 0;
    // block b's sum goes to thread b % CL_BLOCK in increasing b, then a tree
    for (int b = threadIdx.x; b < n_blocks; b += CL_BLOCK) {
        v += c.part[b];
        cnt += c.pcount[b];
    }
    sh[threadIdx.x] = v;
    shn[threadIdx.x] = cnt;
    __syncthreads();
    for (int w = CL_BLOCK / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            sh[threadIdx.x] += sh[threadIdx.x + w];
            shn[threadIdx.x] += shn[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    const double e = c.energy[j];
    double lead = e;
    if (c.T > 0.0) {
        const double p = prob(c, e);
        lead = (p * e) - (c.T * p * log(p));
    }
    c.top[r] = j;
    c.fre[r] = 1 + shn[0];
    c.tcf[r] = lead;
    c.acf[r] = lead + sh[0];
    *c.n_clusters = r + 1;
    const int nx = *c.next;
    *c.next = NO_HEAD;
    if (r + 1 < c.max_results && nx != NO_HEAD) {
        c.head[r + 1] = nx;
        c.member[nx] = nx;
        c.rmsd_head[nx] = 0.f;
    }
}

// QuickSort_Clusters (LIB/cluster.cpp:323-359) over the first max_results
// records (one thread: at most a few dozen records); the member array is
// not permuted (the reference's swap of its first max_results entries only
// reaches its .rrd output)
__device__ void swap_rec(const ClDev& c, int a, int b)
{
    const int t = c.top[a], f = c.fre[a];
    const double x = c.tcf[a], y = c.acf[a];
    c.top[a] = c.top[b];
    c.fre[a] = c.fre[b];
    c.tcf[a] = c.tcf[b];
    c.acf[a] = c.acf[b];
    c.top[b] = t;
    c.fre[b] = f;
    c.tcf[b] = x;
    c.acf[b] = y;
}

__global__ void cl_sort_kernel(ClDev c)
{
    if (threadIdx.x != 0 || blockIdx.x != 0 || !(c.T > 0.0)) return;
    // explicit stack of (beg, end) ranges in place of the recursion
    int stk[2 * 64];
    int sp = 0;
    stk[sp++] = 0;
    stk[sp++] = c.max_results - 1;
    while (sp > 0) {
        int end = stk[--sp];
        int beg = stk[--sp];
        while (beg < end) {
            int l = beg, p = beg + (end - beg) / 2, r = end;
            const double piv = c.acf[p];
            for (;;) {
                while (l <= r && c.acf[l] - piv <= 0) ++l;
                while (l <= r && c.acf[r] - piv > 0) --r;
                if (l > r) break;
                swap_rec(c, l, r);
                if (p == r) p = l;
                ++l;
                --r;
            }
            swap_rec(c, p, r);
            --r;
            // the shorter side first (the reference recurses into it)
            if ((r - beg) < (end - l)) {
                if (sp + 2 <= 2 * 64) {
                    stk[sp++] = l;
                    stk[sp++] = end;
                }
                end = r;
            } else {
                if (sp + 2 <= 2 * 64) {
                    stk[sp++] = beg;
                    stk[sp++] = r;
                }
                beg = l;
            }
        }
    }
}

// RMSD between the heads of the output clusters as calc_rmsd_chrom computes
// it (LIB/calc_rmsd_chrom.cpp: sqrdist summed in atom order, sqrt of the
// mean): one thread per pair
__global__ void cl_top_rmsd_kernel(ClDev c, int n_out, float* out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_out * n_out) return;
    const int a = t / n_out, b = t % n_out;
    if (b <= a) {
        out[t] = 0.f;
        return;
    }
    const int pa = c.top[a], pb = c.top[b];
    float s = 0.f;
    const int nf = 3 * c.n_atoms;
    for (int i = 0; i < c.n_atoms; ++i) {
        const float d0 = c.xyzT[coord_at(nf, 3 * i, pa)] - c.xyzT[coord_at(nf, 3 * i, pb)];
        const float d1 = c.xyzT[coord_at(nf, 3 * i + 1, pa)] - c.xyzT[coord_at(nf, 3 * i + 1, pb)];
        const float d2 = c.xyzT[coord_at(nf, 3 * i + 2, pa)] - c.xyzT[coord_at(nf, 3 * i + 2, pb)];
        s += d0 * d0 + d1 * d1 + d2 * d2;
    }
    out[t] = (float)sqrt((double)(s / (float)c.n_atoms));
}

}  // namespace
}  // namespace fa

extern "C" {

int fa_cluster_cf(fa_ctx* ctx, int ligand, int n_poses, const double* genes, int gene_stride, const double* energy,
                  const fa_cluster_params* p, fa_cluster_mode* modes, int* member, float* rmsd_head, float* top_rmsd,
                  fa_cluster_out* out)
{
    if (!ctx || !genes || !energy || !p || !modes || !out) return fa_fail(FA_ERR_INVALID, "null argument");
    if (!ctx->have_system) return fa_fail(FA_ERR_NOSYSTEM, "no system uploaded");
    if (ligand < 0 || ligand >= ctx->n_ligands) return fa_fail(FA_ERR_INVALID, "ligand index out of range");
    if (n_poses < 0 || gene_stride < ctx->lig_genes[ligand]) return fa_fail(FA_ERR_INVALID, "bad pose count or stride");
    if (p->max_results < 1 || p->max_results > 64) return fa_fail(FA_ERR_INVALID, "max_results in [1, 64]");
    if (!(p->temperature >= 0.0) || !(p->rmsd >= 0.0f)) return fa_fail(FA_ERR_INVALID, "negative temperature or rmsd");
    const int L = ctx->h_ligs[ligand].n_atoms, G = ctx->lig_genes[ligand];
    const int R = p->max_results;
    if ((long long)n_poses * 3 * L > (1LL << 31)) return fa_fail(FA_ERR_LIMIT, "poses * atoms above 2^31 coordinates");
    *out = fa_cluster_out{};
    FA_HIP(hipSetDevice(ctx->device));
    const int n = n_poses, ld = std::max(n, 1);
    const int nb = std::max(1, (n + fa::CL_BLOCK - 1) / fa::CL_BLOCK);
    // one allocation: coordinates, genes, energies, per-pose and per-block arrays, records
    const size_t b_xyz = sizeof(float) * fa::coord_floats(3 * L, std::max(n, 1)), b_ic = b_xyz, b_gen = sizeof(double) * G * (size_t)ld,
                 b_e = sizeof(double) * (size_t)ld, b_pose = (sizeof(int) + sizeof(float)) * (size_t)ld,
                 b_blk = (sizeof(double) + sizeof(int)) * (size_t)nb, b_rec = (2 * sizeof(int) + 2 * sizeof(double)) * R +
                                                                         sizeof(int) * (R + 4) + sizeof(double),
                 b_trm = sizeof(float) * R * R;
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t total = al(b_xyz) + al(b_ic) + al(b_gen) + al(b_e) + al(b_pose) + al(b_blk) + al(b_rec) + al(b_trm);
    FA_HIP(hipDeviceSynchronize());  // the buffer may be regrown: no earlier clustering still reads it
    int rc = fa_ensure(ctx->cluster, total);
    if (rc) return rc;
    char* base = static_cast<char*>(ctx->cluster.p);
    size_t off = 0;
    auto take = [&](size_t bytes) {
        char* q = base + off;
        off += al(bytes);
        return q;
    };
    float* xyzT = reinterpret_cast<float*>(take(b_xyz));
    float* icT = reinterpret_cast<float*>(take(b_ic));
    double* d_genes = reinterpret_cast<double*>(take(b_gen));
    double* d_e = reinterpret_cast<double*>(take(b_e));
    char* pose_area = take(b_pose);
    char* blk_area = take(b_blk);
    char* rec_area = take(b_rec);
    float* d_trm = reinterpret_cast<float*>(take(b_trm));
    fa::ClDev c{};
    c.n = n;
    c.n_atoms = L;
    c.max_results = R;
    c.xyzT = xyzT;
    c.energy = d_e;
    c.T = p->temperature;
    c.beta = p->temperature > 0.0 ? 1.0 / p->temperature : 0.0;
    c.rmsd = p->rmsd;
    c.member = reinterpret_cast<int*>(pose_area);
    c.rmsd_head = reinterpret_cast<float*>(pose_area + sizeof(int) * (size_t)ld);
    c.part = reinterpret_cast<double*>(blk_area);
    c.pcount = reinterpret_cast<int*>(blk_area + sizeof(double) * (size_t)nb);
    c.tcf = reinterpret_cast<double*>(rec_area);
    c.acf = c.tcf + R;
    c.z = c.acf + R;
    c.top = reinterpret_cast<int*>(c.z + 1);
    c.fre = c.top + R;
    c.head = c.fre + R;
    c.next = c.head + R + 1;
    c.n_clusters = c.next + 1;
    hipStream_t s = nullptr;
    // host genes to the device (row stride gene_stride -> G)
    std::vector<double> packed((size_t)G * ld, 0.0);
    for (int i = 0; i < n; ++i)
        std::copy(genes + (size_t)i * gene_stride, genes + (size_t)i * gene_stride + G, packed.begin() + (size_t)i * G);
    FA_HIP(hipMemcpyAsync(d_genes, packed.data(), sizeof(double) * packed.size(), hipMemcpyHostToDevice, s));
    if (n > 0) FA_HIP(hipMemcpyAsync(d_e, energy, sizeof(double) * n, hipMemcpyHostToDevice, s));
    hipEvent_t e0, e1;
    FA_HIP(hipEventCreate(&e0));
    FA_HIP(hipEventCreate(&e1));
    FA_HIP(hipEventRecord(e0, s));
    FA_HIP(fa::launch_coords(ctx->d_ligs, ligand, L, d_genes, G, n, xyzT, icT, s));
    const int ni = std::max(nb, (R + fa::CL_BLOCK) / fa::CL_BLOCK);
    hipLaunchKernelGGL(fa::cl_init_kernel, dim3(ni), dim3(fa::CL_BLOCK), 0, s, c);
    if (c.T > 0.0 && n > 0) {
        hipLaunchKernelGGL(fa::cl_z_kernel, dim3(nb), dim3(fa::CL_BLOCK), 0, s, c);
        hipLaunchKernelGGL(fa::cl_z_final_kernel, dim3(1), dim3(fa::CL_BLOCK), 0, s, c, nb);
    }
    for (int r = 0; r < R && n > 0; ++r) {
        hipLaunchKernelGGL(fa::cl_assign_kernel, dim3(nb), dim3(fa::CL_BLOCK), 0, s, c, r);
        hipLaunchKernelGGL(fa::cl_finish_kernel, dim3(1), dim3(fa::CL_BLOCK), 0, s, c, r, nb);
    }
    hipLaunchKernelGGL(fa::cl_sort_kernel, dim3(1), dim3(64), 0, s, c);
    if (n > 0) hipLaunchKernelGGL(fa::cl_top_rmsd_kernel, dim3((R * R + 255) / 256), dim3(256), 0, s, c, R, d_trm);
    FA_HIP(hipGetLastError());
    FA_HIP(hipEventRecord(e1, s));
    FA_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    FA_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    int nc = 0;
    FA_HIP(hipMemcpy(&nc, c.n_clusters, sizeof(int), hipMemcpyDeviceToHost));
    if (c.T > 0.0 && n > 0) {
        double z = 0.0;
        FA_HIP(hipMemcpy(&z, c.z, sizeof(double), hipMemcpyDeviceToHost));
        // the reference stops here ("The Partition Function is NULL")
        if (!(z > 0.0)) return fa_fail(FA_ERR_INVALID, "partition function is zero: every exp(-E / T) underflows");
    }
    std::vector<int> top(R), fre(R);
    std::vector<double> tcf(R), acf(R);
    FA_HIP(hipMemcpy(top.data(), c.top, sizeof(int) * R, hipMemcpyDeviceToHost));
    FA_HIP(hipMemcpy(fre.data(), c.fre, sizeof(int) * R, hipMemcpyDeviceToHost));
    FA_HIP(hipMemcpy(tcf.data(), c.tcf, sizeof(double) * R, hipMemcpyDeviceToHost));
    FA_HIP(hipMemcpy(acf.data(), c.acf, sizeof(double) * R, hipMemcpyDeviceToHost));
    for (int k = 0; k < nc; ++k) modes[k] = fa_cluster_mode{top[k], fre[k], tcf[k], acf[k]};
    if (member && n > 0) FA_HIP(hipMemcpy(member, c.member, sizeof(int) * n, hipMemcpyDeviceToHost));
    if (rmsd_head && n > 0) FA_HIP(hipMemcpy(rmsd_head, c.rmsd_head, sizeof(float) * n, hipMemcpyDeviceToHost));
    if (top_rmsd) {
        std::vector<float> t((size_t)R * R);
        FA_HIP(hipMemcpy(t.data(), d_trm, sizeof(float) * t.size(), hipMemcpyDeviceToHost));
        for (int a = 0; a < nc; ++a)
            for (int b = 0; b < nc; ++b) top_rmsd[(size_t)a * nc + b] = b > a ? t[(size_t)a * R + b] : b < a ? t[(size_t)b * R + a] : 0.f;
    }
    int clustered = 0;
    for (int k = 0; k < nc; ++k) clustered += fre[k];
    out->n_modes = nc;
    out->n_clustered = clustered;
    out->device_ms = ms;
    return FA_SUCCESS;
}

}  // extern "C"
