// optics.hip — the FastOPTICS super-cluster of pose energies on the device
// (fa_super_cluster, C ABI).
//
// BindingPopulation::get_super_cluster_ensemble (LIB/BindingMode.cpp:153-180)
// keeps the poses of the dominant energy basin: FastOPTICS over the poses'
// energies as one-dimensional points (LIB/fast_optics.cpp:8-135) with
// minPts = max(4, n / 20) and SUPER_CLUSTER_ONLY extraction.  That is
//   core distance  coreDist(v): the (minPts-1)-th smallest distance from v to
//                  the other points; 0 when minPts <= 1, infinite when
//                  n < minPts (computeOrdering then pushes nothing)
//   ordering       point 0 first, then repeatedly the unprocessed point j of
//                  least (reach, j) with reach = min over processed u of
//                  max(coreDist(u), |e_u - e_j|) (the priority queue of
//                  computeOrdering); when no processed point reaches any
//                  point, the next run starts at the least unprocessed index
//                  with an infinite reach
//   seed           the ordering entry of least (reach, index)
//   flood fill     breadth first from the seed; the points v that a visited u
//                  adds, in index order, are those not yet visited with
//                  |e_u - e_v| <= 0.8 coreDist(v)
// and the super-cluster is the fill's visiting order.
//
// Device form: the energies are ranked (a thread per point counts the points
// before it, chunks of 1,024 in separate blocks), and a point's core distance is the (minPts-1)-th step of a
// merge of its left and right neighbours in energy order.  The ordering and
// the fill run in one workgroup that holds every point's energy, core
// distance and reach in registers (up to 512 threads, at most 8 waves, so
// that a step's cross-wave part stays short; a thread owns up to 16
// consecutive points): an ordering step is an argmin over (reach, index) — a wave
// butterfly, one barrier, a read of the waves' winners — and a fill step
// takes up to 32 queue entries at once (each point finds the first of them
// that it joins; ballots per entry order the joins, the waves' counts pass
// through one barrier).  The reference's distance sqrt(d * d)
// is |d|, exactly, in binary floating point while d * d neither underflows
// nor overflows; energies outside the range that guarantees it take a
// guarded form that evaluates sqrt(d * d) there.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <string>

#include "fa_ctx.h"

namespace fa {
namespace {

constexpr int SC_MAX = FA_SUPER_CLUSTER_MAX;
constexpr int SC_THREADS = 512;
constexpr int SC_BLOCK = 256;
constexpr int SC_BATCH = 32;  // fill: queue entries taken at a time

__device__ __forceinline__ double sc_dist(double a, double b)
{
    const double d = a - b;
    const double ad = fabs(d);
    if (ad == 0.0 || (ad >= 0x1p-500 && ad <= 0x1p+500)) return ad;
    return sqrt(d * d);
}

constexpr int SC_RANK_CHUNK = 1024;  // points counted by one block of the ranking

// rank[j] += #points i of this block's chunk before j in (energy, index)
// order (rank zeroed first); blocks (j tile, i chunk)
__global__ __launch_bounds__(SC_BLOCK) void sc_rank_kernel(const double* __restrict__ x, int n, int* __restrict__ rank)
{
    __shared__ double tile[SC_BLOCK];
    const int j = blockIdx.x * SC_BLOCK + threadIdx.x;
    const double xj = j < n ? x[j] : 0.0;
    int r = 0;
    const int lo = blockIdx.y * SC_RANK_CHUNK, hi = min(n, lo + SC_RANK_CHUNK);
    for (int base = lo; base < hi; base += SC_BLOCK) {
        const int i = base + threadIdx.x;
        tile[threadIdx.x] = i < n ? x[i] : 0.0;
        __syncthreads();
        const int m = min(SC_BLOCK, hi - base);
        for (int q = 0; q < m; ++q) {
            const double xi = tile[q];
            r += (xi < xj) || (xi == xj && base + q < j);
        }
        __syncthreads();
    }
    if (j < n && r > 0) atomicAdd(&rank[j], r);
}

// sorted[rank[j]] = e_j
__global__ __launch_bounds__(SC_BLOCK) void sc_scatter_kernel(const double* __restrict__ x, int n,
                                                              const int* __restrict__ rank, double* __restrict__ sorted)
{
    const int j = blockIdx.x * SC_BLOCK + threadIdx.x;
    if (j < n) sorted[rank[j]] = x[j];
}

// coreDist of every point: the distances to its neighbours grow away from
// it on either side of the energy order, so the k-th smallest is the k-th
// step of merging the two sides
__global__ __launch_bounds__(SC_BLOCK) void sc_core_kernel(const double* __restrict__ x, const int* __restrict__ rank,
                                                           const double* __restrict__ sorted, int n, int min_pts,
                                                           double* __restrict__ cd)
{
    const int j = blockIdx.x * SC_BLOCK + threadIdx.x;
    if (j >= n) return;
    double c = 0.0;
    if (min_pts <= 1) {
        c = 0.0;
    } else if (n < min_pts) {
        c = INFINITY;
    } else {
        const double xj = x[j];
        const int p = rank[j];
        int l = p - 1, r = p + 1;
        for (int t = 0; t < min_pts - 1; ++t) {
            const double dl = l >= 0 ? sc_dist(xj, sorted[l]) : INFINITY;
            const double dr = r < n ? sc_dist(xj, sorted[r]) : INFINITY;
            if (dl <= dr) {
                c = dl;
                --l;
            } else {
                c = dr;
                ++r;
            }
        }
    }
    cd[j] = c;
}

// ------------------------------------------------------------ wave helpers
// 64-bit DPP move inside 16-lane rows (the butterflies below never read
// outside a row); all lanes must be active
template <int CTRL>
__device__ __forceinline__ uint64_t dpp_u64(uint64_t v)
{
    const int lo = __builtin_amdgcn_update_dpp((int)(uint32_t)v, (int)(uint32_t)v, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(uint32_t)(v >> 32), (int)(uint32_t)(v >> 32), CTRL, 0xF, 0xF, false);
    return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
}

__device__ __forceinline__ uint64_t umin64(uint64_t a, uint64_t b) { return b < a ? b : a; }

__device__ __forceinline__ uint64_t readlane_u64(uint64_t v, int lane)
{
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, lane);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), lane);
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ double readlane_f64(double v, int lane)
{
    return __longlong_as_double((long long)readlane_u64((uint64_t)__double_as_longlong(v), lane));
}

// minimum of v over each 16-lane row, in every lane of the row
__device__ __forceinline__ uint64_t row_min_u64(uint64_t v)
{
    v = umin64(v, dpp_u64<0xB1>(v));   // quad_perm [1,0,3,2]
    v = umin64(v, dpp_u64<0x4E>(v));   // quad_perm [2,3,0,1]
    v = umin64(v, dpp_u64<0x141>(v));  // row_half_mirror
    v = umin64(v, dpp_u64<0x140>(v));  // row_mirror
    return v;
}

__device__ __forceinline__ uint64_t wave_min_u64(uint64_t v)
{
    v = row_min_u64(v);
    return umin64(umin64(readlane_u64(v, 0), readlane_u64(v, 16)), umin64(readlane_u64(v, 32), readlane_u64(v, 48)));
}

__device__ __forceinline__ unsigned long long ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

__device__ __forceinline__ int lanes_below(unsigned long long m)
{
    return __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

// |a - b|: equal to the reference's sqrt(d * d) when the host has checked
// that every nonzero difference of the energies lies in [2^-500, 2^500]
// (EXACT); otherwise the guarded form
template <bool EXACT>
__device__ __forceinline__ double sc_dist_t(double a, double b)
{
    if constexpr (EXACT)
        return fabs(a - b);
    else
        return sc_dist(a, b);
}

struct ScSlot {
    uint64_t key;  // reach bits: non-negative doubles order as their bit patterns
    int idx, pad;
};

// One workgroup of blockDim.x (a multiple of 64, blockDim.x * K >= n)
// threads; thread t owns points t*K .. t*K+K-1, so lane order, then k, is
// point order.  out[0..info[0]) receives the fill's visiting order, info[1]
// the seed.  Reaches are kept as doubles whose bit patterns order like their
// values (all are >= +0 or +inf); a processed or absent point holds -inf,
// which the minimum keeps and whose bit pattern sorts above +inf's.  The
// winner's energy and core distance come from LDS (the core distances'
// area holds the fill's queue afterwards).
template <int K, bool EXACT>
__global__ __launch_bounds__(SC_THREADS) void sc_optics_kernel(const double* __restrict__ x,
                                                               const double* __restrict__ cd, int n,
                                                               int* __restrict__ out, int* __restrict__ info)
{
    __shared__ double xs[SC_MAX];
    __shared__ double cds[SC_MAX];
    int* const queue = reinterpret_cast<int*>(cds);
    __shared__ ScSlot slot[2][SC_THREADS / 64];
    __shared__ int wsum[SC_BATCH][SC_THREADS / 64];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6, nw = blockDim.x >> 6;
    const double taken = -INFINITY;
    double xr[K], cr[K], br[K];
    unsigned absent = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int v = t * K + k;
        xr[k] = v < n ? x[v] : 0.0;
        cr[k] = v < n ? cd[v] : 0.0;
        br[k] = v < n ? INFINITY : taken;
        if (v >= n) absent |= 1u << k;
    }
    for (int v = t; v < n; v += blockDim.x) {
        xs[v] = x[v];
        cds[v] = cd[v];
    }
    if (t == 0) br[0] = taken;  // point 0 opens the ordering (reach infinite)
    double xu = x[0], cu = cd[0];
    // least (reach, index) of the ordering so far: (+inf, 0)
    uint64_t s_key = (uint64_t)__double_as_longlong(INFINITY);
    int s_idx = 0;
    for (int step = 1; step < n; ++step) {
        uint64_t bkey = ~0ull;
        int bk = 0;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            // no operand is a NaN: fmax / fmin are the reference's max and
            // the queue's least entry
            br[k] = fmin(br[k], fmax(cu, sc_dist_t<EXACT>(xu, xr[k])));
            const uint64_t key = (uint64_t)__double_as_longlong(br[k]);
            if (key < bkey) {
                bkey = key;
                bk = k;
            }
        }
        // the wave's least key; its first lane holds the least index
        const uint64_t wmin = wave_min_u64(bkey);
        const int wl = (int)__builtin_ctzll(ballot(bkey == wmin));
        const int par = step & 1;
        if (lane == wl) slot[par][wave] = ScSlot{wmin, t * K + bk, 0};
        __syncthreads();
        // the block's: lanes 0..nw-1 hold the waves' winners, in point order
        const bool live = lane < nw;
        const ScSlot c = slot[par][live ? lane : 0];
        const uint64_t ck = live ? c.key : ~0ull;
        const uint64_t r = row_min_u64(ck);
        const uint64_t gmin = readlane_u64(r, 0);
        const int ww = (int)__builtin_ctzll(ballot(live && ck == gmin));
        const int widx = __builtin_amdgcn_readlane(c.idx, ww);
        xu = xs[widx];
        cu = cds[widx];
        if (gmin < s_key || (gmin == s_key && widx < s_idx)) {
            s_key = gmin;
            s_idx = widx;
        }
#pragma unroll
        for (int k = 0; k < K; ++k)
            if (t * K + k == widx) br[k] = taken;
    }
    // flood fill from the seed, up to SC_BATCH queue entries at a time: the
    // entries [head, head + B) are all in the queue when the batch starts,
    // so the one-at-a-time fill adds, in order, the points whose first
    // qualifying entry is the batch's first, then its second, ..., each
    // group in point order; here each point finds its first qualifying
    // entry, and the groups are counted per wave and laid out in that order
    unsigned vis = absent;
#pragma unroll
    for (int k = 0; k < K; ++k)
        if (t * K + k == s_idx) vis |= 1u << k;
    __syncthreads();  // the queue takes the core distances' area
    if (t == 0) queue[0] = s_idx;
    __syncthreads();
    int head = 0, tail = 1;
    while (head < tail) {
        const int B = min(SC_BATCH, tail - head);
        const double xb = lane < B ? xs[queue[head + lane]] : 0.0;
        int fp[K];
#pragma unroll
        for (int k = 0; k < K; ++k) fp[k] = -1;
        for (int b = 0; b < B; ++b) {
            const double xq = readlane_f64(xb, b);
#pragma unroll
            for (int k = 0; k < K; ++k)
                if (fp[k] < 0 && !(vis >> k & 1u) && sc_dist_t<EXACT>(xq, xr[k]) <= 0.8 * cr[k]) fp[k] = b;
        }
        // the batch entries that add points in this wave
        unsigned lm = 0;
#pragma unroll
        for (int k = 0; k < K; ++k)
            if (fp[k] >= 0) lm |= 1u << fp[k];
        unsigned wm = lm;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) wm |= (unsigned)__shfl_xor((int)wm, o);
        wm = (unsigned)__builtin_amdgcn_readfirstlane((int)wm);
        int rank[K];  // a joining point's place among this wave's points of its entry
        int mycnt = 0;  // lane l: this wave's points added by entry l
        for (unsigned rem = wm; rem != 0u; rem &= rem - 1u) {
            const int f = __builtin_ctz(rem);
            int below = 0, tot = 0;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const unsigned long long bk = ballot(fp[k] == f);
                below += lanes_below(bk);
                tot += __popcll(bk);
            }
            int own = 0;
#pragma unroll
            for (int k = 0; k < K; ++k)
                if (fp[k] == f) rank[k] = below + own++;
            if (lane == f) mycnt = tot;
        }
        if (lane < SC_BATCH) wsum[lane][wave] = mycnt;
        __syncthreads();
        int T = 0, P = 0;
        if (lane < B)
            for (int w = 0; w < nw; ++w) {
                const int c = wsum[lane][w];
                T += c;
                P += w < wave ? c : 0;
            }
        int incl = T;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(incl, o);
            if (lane >= o) incl += y;
        }
        const int base = incl - T + P;  // lane l: where this wave's points of entry l start
        const int total = __builtin_amdgcn_readlane(incl, 63);
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int at = __shfl(base, fp[k] < 0 ? 0 : fp[k]);
            if (fp[k] >= 0) {
                queue[tail + at + rank[k]] = t * K + k;
                vis |= 1u << k;
            }
        }
        tail += total;
        head += B;
        __syncthreads();
    }
    __syncthreads();
    for (int q = t; q < tail; q += blockDim.x) out[q] = queue[q];
    if (t == 0) {
        info[0] = tail;
        info[1] = s_idx;
    }
}

template <int K>
hipError_t launch_optics(bool exact, int threads, const double* x, const double* cd, int n, int* out, int* info,
                         hipStream_t s)
{
    if (exact)
        hipLaunchKernelGGL((sc_optics_kernel<K, true>), dim3(1), dim3(threads), 0, s, x, cd, n, out, info);
    else
        hipLaunchKernelGGL((sc_optics_kernel<K, false>), dim3(1), dim3(threads), 0, s, x, cd, n, out, info);
    return hipGetLastError();
}

}  // namespace
}  // namespace fa

extern "C" {

int fa_super_cluster(fa_ctx* ctx, const double* energy, int n, int min_pts, int* members, fa_super_cluster_out* out)
{
    if (!ctx || !out || (n > 0 && (!energy || !members))) return fa_fail(FA_ERR_INVALID, "null argument");
    if (n < 0) return fa_fail(FA_ERR_INVALID, "negative point count");
    if (n > FA_SUPER_CLUSTER_MAX)
        return fa_fail(FA_ERR_LIMIT, "super cluster over more than " + std::to_string(FA_SUPER_CLUSTER_MAX) + " points");
    // |e_a - e_b| is the reference's sqrt(d * d) while d * d neither
    // underflows nor overflows: certain when every nonzero |e| lies in
    // [2^-440, 2^498] (a nonzero difference is then at least 2^-492, the
    // spacing of doubles at 2^-440, and below 2^499)
    bool exact = true;
    for (int i = 0; i < n; ++i) {
        if (!std::isfinite(energy[i])) return fa_fail(FA_ERR_INVALID, "non-finite energy");
        const double a = std::fabs(energy[i]);
        if (a != 0.0 && (a < 0x1p-440 || a > 0x1p+498)) exact = false;
    }
    *out = fa_super_cluster_out{};
    if (n == 0) return FA_SUCCESS;
    FA_HIP(hipSetDevice(ctx->device));
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t bd = al(sizeof(double) * n), bi = al(sizeof(int) * n);
    FA_HIP(hipDeviceSynchronize());  // the buffer may be regrown: no earlier clustering still reads it
    int rc = fa_ensure(ctx->cluster, 3 * bd + 2 * bi + 256);
    if (rc) return rc;
    char* base = static_cast<char*>(ctx->cluster.p);
    double* d_x = reinterpret_cast<double*>(base);
    double* d_sorted = reinterpret_cast<double*>(base + bd);
    double* d_cd = reinterpret_cast<double*>(base + 2 * bd);
    int* d_rank = reinterpret_cast<int*>(base + 3 * bd);
    int* d_out = reinterpret_cast<int*>(base + 3 * bd + bi);
    int* d_info = reinterpret_cast<int*>(base + 3 * bd + 2 * bi);
    hipStream_t s = nullptr;
    FA_HIP(hipMemcpyAsync(d_x, energy, sizeof(double) * n, hipMemcpyHostToDevice, s));
    hipEvent_t e0, e1;
    FA_HIP(hipEventCreate(&e0));
    FA_HIP(hipEventCreate(&e1));
    FA_HIP(hipEventRecord(e0, s));
    const int nb = (n + fa::SC_BLOCK - 1) / fa::SC_BLOCK;
    FA_HIP(hipMemsetAsync(d_rank, 0, sizeof(int) * n, s));
    hipLaunchKernelGGL(fa::sc_rank_kernel, dim3(nb, (n + fa::SC_RANK_CHUNK - 1) / fa::SC_RANK_CHUNK), dim3(fa::SC_BLOCK), 0,
                       s, d_x, n, d_rank);
    hipLaunchKernelGGL(fa::sc_scatter_kernel, dim3(nb), dim3(fa::SC_BLOCK), 0, s, d_x, n, d_rank, d_sorted);
    hipLaunchKernelGGL(fa::sc_core_kernel, dim3(nb), dim3(fa::SC_BLOCK), 0, s, d_x, d_rank, d_sorted, n, min_pts, d_cd);
    FA_HIP(hipGetLastError());
    // points per thread: the least K in {1, 2, 4, 8, 16} with SC_THREADS K >= n
    // (at most 8 waves: a step's cross-wave part shrinks with the waves)
    const int K = n <= fa::SC_THREADS ? 1 : n <= 2 * fa::SC_THREADS ? 2 : n <= 4 * fa::SC_THREADS ? 4 : n <= 8 * fa::SC_THREADS ? 8 : 16;
    const int threads = ((n + K - 1) / K + 63) / 64 * 64;
    hipError_t le = K == 1   ? fa::launch_optics<1>(exact, threads, d_x, d_cd, n, d_out, d_info, s)
                    : K == 2 ? fa::launch_optics<2>(exact, threads, d_x, d_cd, n, d_out, d_info, s)
                    : K == 4 ? fa::launch_optics<4>(exact, threads, d_x, d_cd, n, d_out, d_info, s)
                    : K == 8 ? fa::launch_optics<8>(exact, threads, d_x, d_cd, n, d_out, d_info, s)
                             : fa::launch_optics<16>(exact, threads, d_x, d_cd, n, d_out, d_info, s);
    FA_HIP(le);
    FA_HIP(hipEventRecord(e1, s));
    FA_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    FA_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    int info[2] = {0, 0};
    FA_HIP(hipMemcpy(info, d_info, sizeof(info), hipMemcpyDeviceToHost));
    if (info[0] < 1 || info[0] > n) return fa_fail(FA_ERR_INVALID, "super cluster: bad member count from the device");
    FA_HIP(hipMemcpy(members, d_out, sizeof(int) * info[0], hipMemcpyDeviceToHost));
    out->n_members = info[0];
    out->seed = info[1];
    out->device_ms = ms;
    return FA_SUCCESS;
}

}  // extern "C"
