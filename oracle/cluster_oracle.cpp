/*
 * cluster_oracle.cpp — CPU restatement of the reference's default post-GA
 * pose clustering (CLUSTA CF).
 *
 * TEST INFRASTRUCTURE ONLY (tests/, bench.py's CPU baseline); the shipped
 * library never links it.
 *
 *   fa_oracle_cluster_cf   cluster(), LIB/cluster.cpp:9-176, with
 *                          QuickSort_Clusters (:323-359) and the RMSDs of
 *                          calc_rmsd_chrom (LIB/calc_rmsd_chrom.cpp) and
 *                          flexaids::sum_sq_distances_f
 *                          (LIB/simd_distance.h:680-716)
 *   fa_oracle_super_cluster  FastOPTICS on one-dimensional points with
 *                          SUPER_CLUSTER_ONLY extraction (LIB/fast_optics.cpp:
 *                          computeOrdering, coreDist, extractSuperCluster), as
 *                          get_super_cluster_ensemble calls it
 *                          (LIB/BindingMode.cpp:150-180)
 *
 * Poses come as ligand coordinates (what calc_rmsd_chrom rebuilds from the
 * genes) in the order given.  dist_mode selects the squared-distance sum of
 * sum_sq_distances_f: 0 its scalar form (a build without AVX2), 1 its AVX2
 * form (eight fused multiply-add accumulators, its horizontal sum, a fused
 * tail: the reference's x86 release build).
 *
 * Parity: pinned against the reference's cluster() compiled from its
 * sources (oracle/_ref/libfa_ref.so and libfa_ref_fast.so, fa_ref_cluster
 * in oracle/ref/ref_cluster.cpp) on generated
 * pose sets, by the .cad file it writes (heads, sizes, TCF / ACF and the
 * RMSD between heads to its six printed decimals), and by brute-force
 * properties in tests/test_cluster_oracle.py.
 */
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <limits>
#include <queue>
#include <utility>
#include <vector>

namespace {

float sum_sq(const float* a, const float* b, int nf, int mode)
{
    if (mode == 0) {
        float s = 0.0f;
        for (int i = 0; i < nf; ++i) {
            const float d = a[i] - b[i];
            s += d * d;
        }
        return s;
    }
    float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    int i = 0;
    for (; i + 7 < nf; i += 8)
        for (int q = 0; q < 8; ++q) {
            const float d = a[i + q] - b[i + q];
            acc[q] = std::fma(d, d, acc[q]);
        }
    // _mm256_extractf128 / add, _mm_movehl / add, _mm_shuffle / add_ss
    const float s4[4] = {acc[0] + acc[4], acc[1] + acc[5], acc[2] + acc[6], acc[3] + acc[7]};
    const float s2[2] = {s4[0] + s4[2], s4[1] + s4[3]};
    float s = s2[0] + s2[1];
    for (; i < nf; ++i) {
        const float d = a[i] - b[i];
        s = std::fma(d, d, s);
    }
    return s;
}

void swap_rec(int* top, int* fre, double* tcf, double* acf, int a, int b)
{
    std::swap(top[a], top[b]);
    std::swap(fre[a], fre[b]);
    std::swap(tcf[a], tcf[b]);
    std::swap(acf[a], acf[b]);
}

// QuickSort_Clusters, recursive as written
void sort_clusters(int* top, int* fre, double* tcf, double* acf, int beg, int end)
{
    while (beg < end) {
        int l = beg, p = beg + (end - beg) / 2, r = end;
        const double piv = acf[p];
        while (true) {
            while ((l <= r) && (acf[l] - piv) <= 0) ++l;
            while ((l <= r) && (acf[r] - piv) > 0) --r;
            if (l > r) break;
            swap_rec(top, fre, tcf, acf, l, r);
            if (p == r) p = l;
            ++l;
            --r;
        }
        swap_rec(top, fre, tcf, acf, p, r);
        --r;
        if ((r - beg) < (end - l)) {
            sort_clusters(top, fre, tcf, acf, beg, r);
            beg = l;
        } else {
            sort_clusters(top, fre, tcf, acf, l, end);
            end = r;
        }
    }
}

}  // namespace

extern "C" {

/* Returns the number of output modes (num_of_clusters); top/fre/tcf/acf
 * [max_results]; member/rmsd_head [n]; top_rmsd [n_out * n_out] */
int fa_oracle_cluster_cf(int n, int n_atoms, const float* xyz, const double* energy, double T, float rmsd_thr,
                         int max_results, int dist_mode, int* top, int* fre, double* tcf, double* acf, int* member,
                         float* rmsd_head, float* top_rmsd)
{
    const int nf = 3 * n_atoms;
    const double beta = T > 0.0 ? 1.0 / T : 0.0;
    std::vector<int> gapop(n, -1);
    std::vector<float> rmsdt(n, 0.0f);
    for (int k = 0; k < max_results; ++k) {
        top[k] = 0;
        fre[k] = 0;
        tcf[k] = 0.0;
        acf[k] = 0.0;
    }
    double z = 0.0;
    if (T > 0.0)
        for (int j = 0; j < n; ++j) z += std::exp((-1.0) * beta * energy[j]);
    int n_unclus = n, nc = 0;
    while (n_unclus > 0) {
        int j = 0;
        for (; j < n; ++j)
            if (gapop[j] == -1) break;
        if (j >= n) break;
        gapop[j] = j;
        rmsdt[j] = 0.0f;
        --n_unclus;
        if (T > 0.0) {
            const double pj = std::exp((-1.0) * beta * energy[j]) / z;
            tcf[nc] = acf[nc] = (double)((pj * energy[j]) - (T * pj * std::log(pj)));
        } else {
            tcf[nc] = acf[nc] = energy[j];
        }
        top[nc] = j;
        fre[nc]++;
        const float* cj = xyz + (std::size_t)j * nf;
        for (int i = j + 1; i < n; ++i) {
            if (gapop[i] != -1) continue;
            const float d = sum_sq(xyz + (std::size_t)i * nf, cj, nf, dist_mode);
            const float r = std::sqrt(d / (float)n_atoms);
            if (r <= rmsd_thr) {
                gapop[i] = j;
                rmsdt[i] = r;
                --n_unclus;
                if (T > 0.0) {
                    const double pi = std::exp((-1.0) * beta * energy[i]) / z;
                    acf[nc] += (double)((pi * energy[i]) + (T * pi * std::log(pi)));
                } else {
                    acf[nc] += energy[i];
                }
                fre[nc]++;
            }
        }
        nc++;
        if (nc == max_results) break;
    }
    if (T > 0.0) sort_clusters(top, fre, tcf, acf, 0, max_results - 1);
    for (int i = 0; i < n; ++i) {
        if (member) member[i] = gapop[i];
        if (rmsd_head) rmsd_head[i] = rmsdt[i];
    }
    if (top_rmsd)
        for (int a = 0; a < nc; ++a)
            for (int b = 0; b < nc; ++b) {
                if (a == b) {
                    top_rmsd[a * nc + b] = 0.0f;
                    continue;
                }
                const int pa = top[a < b ? a : b], pb = top[a < b ? b : a];
                float s = 0.0f;
                for (int i = 0; i < n_atoms; ++i) {
                    const float* u = xyz + (std::size_t)pa * nf + 3 * i;
                    const float* v = xyz + (std::size_t)pb * nf + 3 * i;
                    const float d0 = u[0] - v[0], d1 = u[1] - v[1], d2 = u[2] - v[2];
                    s += d0 * d0 + d1 * d1 + d2 * d2;
                }
                top_rmsd[a * nc + b] = (float)std::sqrt(s / ((float)n_atoms));
            }
    return nc;
}

/* The FastOPTICS super-cluster of one-dimensional points: out_idx receives
 * the point indices in the flood fill's visiting order, *n_out their count.
 * Distances are sqrt(d * d) in double, as the reference's distance(). */
int fa_oracle_super_cluster(const double* e, int n, int min_pts, int* out_idx, int* n_out)
{
    const double inf = std::numeric_limits<double>::infinity();
    auto dist = [&](int a, int b) {
        const double d = e[a] - e[b];
        return std::sqrt(d * d);
    };
    auto core = [&](int v) {
        if (min_pts <= 1) return 0.0;
        if (n < min_pts) return inf;
        std::vector<double> d;
        for (int j = 0; j < n; ++j)
            if (j != v) d.push_back(dist(v, j));
        if ((int)d.size() < min_pts - 1) return inf;
        std::nth_element(d.begin(), d.begin() + (min_pts - 2), d.end());
        return d[min_pts - 2];
    };
    *n_out = 0;
    if (n <= 0) return 0;
    std::vector<double> cd(n);
    for (int v = 0; v < n; ++v) cd[v] = core(v);
    // computeOrdering's priority queue holds (max(cd[u], dist(u, j)), j) for
    // every processed u and every j unprocessed when u was processed; a pop
    // skips processed j, so it yields the unprocessed j of least (best[j], j),
    // best[j] the least entry for j so far.  The queue empties only when no
    // processed point of the run had a finite core distance (one that has
    // pushes every unprocessed point): every best[j] is then infinite and the
    // least unprocessed index opens the next run with an infinite reach.
    std::vector<char> done(n, 0);
    std::vector<double> best(n, inf);
    std::vector<std::pair<int, double>> order;
    int u = 0;
    done[0] = 1;
    order.push_back({0, inf});
    for (int step = 1; step < n; ++step) {
        int w = -1;
        for (int j = 0; j < n; ++j) {
            if (done[j]) continue;
            if (cd[u] < inf) best[j] = std::min(best[j], std::max(cd[u], dist(u, j)));
            if (w < 0 || best[j] < best[w]) w = j;
        }
        done[w] = 1;
        order.push_back({w, best[w]});
        u = w;
    }
    // the entry of least reach (ties: the smaller point index) seeds a flood
    // fill: v joins from u when dist(u, v) <= 0.8 coreDist(v)
    size_t sp = 0;
    for (size_t q = 1; q < order.size(); ++q)
        if (order[q].second < order[sp].second || (order[q].second == order[sp].second && order[q].first < order[sp].first))
            sp = q;
    const int seed = order[sp].first;
    std::vector<char> vis(n, 0);
    std::queue<int> fifo;
    fifo.push(seed);
    vis[seed] = 1;
    while (!fifo.empty()) {
        const int u = fifo.front();
        fifo.pop();
        out_idx[(*n_out)++] = u;
        for (int v = 0; v < n; ++v)
            if (!vis[v] && dist(u, v) <= 0.8 * cd[v]) {
                vis[v] = 1;
                fifo.push(v);
            }
    }
    return 0;
}

}  // extern "C"
