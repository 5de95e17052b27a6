// statmech.hip — FP64 canonical-ensemble reduction on the GPU.
//
// Computes ln Z, <E>, <E^2> over energies E_i with multiplicities n_i at
// temperature T, the quantities of statmech::StatMechEngine::compute
// (LIB/statmech.cpp:73-173): log w_i = ln n_i - beta E_i, ln Z = LSE(log w),
// <E> = sum E_i exp(log w_i - ln Z).  Each thread folds its elements into an
// online log-sum-exp state (max m, s = sum exp(lw - m), s1 = sum E exp(..),
// s2 = sum E^2 exp(..)); states merge by rescaling to the larger max.  One
// pass over HBM, FP64 throughout.
#include <hip/hip_runtime.h>

#include <math.h>

#include "lse.h"

namespace fa {

__device__ void block_reduce(LseState& st, LseState* sh)
{
    for (int off = 32; off > 0; off >>= 1) {
        LseState o;
        o.m = __shfl_xor(st.m, off);
        o.s = __shfl_xor(st.s, off);
        o.s1 = __shfl_xor(st.s1, off);
        o.s2 = __shfl_xor(st.s2, off);
        lse_merge(st, o);
    }
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    if (l == 0) sh[w] = st;
    __syncthreads();
    if (threadIdx.x == 0) {
        LseState acc = sh[0];
        for (int k = 1; k < (int)(blockDim.x / 64); ++k) lse_merge(acc, sh[k]);
        sh[0] = acc;
    }
    __syncthreads();
    st = sh[0];
}

// pass 1: one state per block
__global__ __launch_bounds__(256) void lse_partial_kernel(const double* __restrict__ energy, int stride,
                                                           const int* __restrict__ count, long long n, double beta,
                                                           LseState* __restrict__ part)
{
    __shared__ LseState sh[4];
    LseState st{0.0, 0.0, 0.0, 0.0};
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const double e = energy[i * stride];
        const double c = count ? (double)count[i] : 1.0;
        lse_push(st, log(c) - beta * e, e);
    }
    block_reduce(st, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = st;
}

// pass 2: fold the block states, write the thermodynamics
__global__ __launch_bounds__(256) void lse_final_kernel(const LseState* __restrict__ part, int nparts, double T,
                                                         fa_thermo* __restrict__ out)
{
    __shared__ LseState sh[4];
    LseState st{0.0, 0.0, 0.0, 0.0};
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) lse_merge(st, part[i]);
    block_reduce(st, sh);
    if (threadIdx.x == 0) *out = thermo_of(st, T);
}

// w_i = exp(-beta E_i - m) with m = max_i(-beta E_i): the unnormalised
// weights of compute_boltzmann_batch (LIB/UnifiedHardwareDispatch.cpp:730-747)
__global__ __launch_bounds__(256) void boltzmann_map_kernel(const double* __restrict__ energy, int stride, long long n,
                                                             double beta, const LseState* __restrict__ total,
                                                             double* __restrict__ w)
{
    const double m = total->m;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        w[i] = exp(-beta * energy[i * stride] - m);
}

// fold the block states into part[0] (single block)
__global__ __launch_bounds__(256) void lse_fold_kernel(LseState* __restrict__ part, int nparts)
{
    __shared__ LseState sh[4];
    LseState st{0.0, 0.0, 0.0, 0.0};
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) lse_merge(st, part[i]);
    block_reduce(st, sh);
    __syncthreads();
    if (threadIdx.x == 0) part[0] = st;
}

// clears n 32-bit words (a kernel instead of a memset node, so that launches
// captured into a HIP graph replay correctly)
__global__ void zero_u32_kernel(unsigned int* __restrict__ p, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
}

hipError_t launch_zero_u32(unsigned int* p, int n_words, hipStream_t stream)
{
    if (n_words <= 0) return hipSuccess;
    const int blocks = (n_words + 255) / 256 > 1024 ? 1024 : (n_words + 255) / 256;
    hipLaunchKernelGGL(zero_u32_kernel, dim3(blocks), dim3(256), 0, stream, p, n_words);
    return hipGetLastError();
}

constexpr int LSE_MAX_PARTS = 1024;

size_t lse_scratch_bytes() { return sizeof(LseState) * LSE_MAX_PARTS; }

hipError_t launch_thermo(const double* d_energy, int stride, const int* d_count, long long n, double T,
                         void* d_scratch, fa_thermo* d_out, hipStream_t stream)
{
    const double beta = 1.0 / (KB_KCAL * T);
    long long blocks = (n + 255) / 256;
    if (blocks > LSE_MAX_PARTS) blocks = LSE_MAX_PARTS;
    if (blocks < 1) blocks = 1;
    LseState* part = static_cast<LseState*>(d_scratch);
    hipLaunchKernelGGL(lse_partial_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, d_energy, stride, d_count, n,
                       beta, part);
    hipLaunchKernelGGL(lse_final_kernel, dim3(1), dim3(256), 0, stream, part, (int)blocks, T, d_out);
    return hipGetLastError();
}

// Boltzmann weights (unnormalised, max-shifted) and ln Z = m + ln sum w.
hipError_t launch_boltzmann(const double* d_energy, int stride, long long n, double T, void* d_scratch, double* d_w,
                            hipStream_t stream)
{
    const double beta = 1.0 / (KB_KCAL * T);
    long long blocks = (n + 255) / 256;
    if (blocks > LSE_MAX_PARTS) blocks = LSE_MAX_PARTS;
    if (blocks < 1) blocks = 1;
    LseState* part = static_cast<LseState*>(d_scratch);
    hipLaunchKernelGGL(lse_partial_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, d_energy, stride,
                       (const int*)nullptr, n, beta, part);
    hipLaunchKernelGGL(lse_fold_kernel, dim3(1), dim3(256), 0, stream, part, (int)blocks);
    long long mb = (n + 255) / 256;
    if (mb > 4096) mb = 4096;
    hipLaunchKernelGGL(boltzmann_map_kernel, dim3((unsigned)mb), dim3(256), 0, stream, d_energy, stride, n, beta,
                       part, d_w);
    return hipGetLastError();
}

// log-sum-exp of raw values x_i (x at stride): part[0] = (max, sum exp(x - max))
hipError_t launch_lse_values(const double* d_x, int stride, long long n, void* d_scratch, hipStream_t stream)
{
    long long blocks = (n + 255) / 256;
    if (blocks > LSE_MAX_PARTS) blocks = LSE_MAX_PARTS;
    if (blocks < 1) blocks = 1;
    LseState* part = static_cast<LseState*>(d_scratch);
    // lw = log(1) - beta * e with beta = -1 gives lw = x
    hipLaunchKernelGGL(lse_partial_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, d_x, stride,
                       (const int*)nullptr, n, -1.0, part);
    hipLaunchKernelGGL(lse_fold_kernel, dim3(1), dim3(256), 0, stream, part, (int)blocks);
    return hipGetLastError();
}

// Grand partition function over n ligands (GrandPartitionFunction::
// compute_log_Xi_fresh and the occupancies, LIB/GrandPartitionFunction.cpp:112-160):
// log zZ_i = ln c_i + ln Z_i, ln Xi = LSE(0, log zZ_i), p_i = exp(log zZ_i -
// ln Xi), p_empty = exp(-ln Xi).  One block; out = [ln Xi, p_empty, p_0..].
__global__ __launch_bounds__(256) void grand_kernel(const double* __restrict__ log_Z, int stride,
                                                    const double* __restrict__ log_c, int n, double* __restrict__ out)
{
    __shared__ double sh[256];
    double m = 0.0;  // the empty state's term is exp(0)
    for (int i = threadIdx.x; i < n; i += blockDim.x) m = fmax(m, log_c[i] + log_Z[(size_t)i * stride]);
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
        __syncthreads();
    }
    m = sh[0];
    __syncthreads();
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += exp(log_c[i] + log_Z[(size_t)i * stride] - m);
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    const double lx = m + log(exp(-m) + sh[0]);
    for (int i = threadIdx.x; i < n; i += blockDim.x) out[2 + i] = exp(log_c[i] + log_Z[(size_t)i * stride] - lx);
    if (threadIdx.x == 0) {
        out[0] = lx;
        out[1] = exp(-lx);
    }
}

hipError_t launch_grand(const double* d_log_Z, int stride, const double* d_log_c, int n, double* d_out,
                        hipStream_t stream)
{
    hipLaunchKernelGGL(grand_kernel, dim3(1), dim3(256), 0, stream, d_log_Z, stride, d_log_c, n, d_out);
    return hipGetLastError();
}

}  // namespace fa
