/*
 * ref_optics.cpp — harness that runs the reference's FastOPTICS
 * super-cluster extraction (LIB/fast_optics.cpp, as
 * BindingPopulation::get_super_cluster_ensemble calls it,
 * LIB/BindingMode.cpp:150-180) on one-dimensional points (pose energies).
 *
 * TEST INFRASTRUCTURE ONLY (linked with the reference sources compiled in
 * place, oracle/ref/Makefile).
 */
#include "fast_optics.hpp"

#include <vector>

extern "C" int fa_ref_super_cluster(const double* energy, int n, int min_pts, int* out_idx, int* n_out)
{
    if (n < 0 || (n > 0 && (!energy || !out_idx)) || !n_out) return -1;
    std::vector<fast_optics::Point> pts(n);
    for (int i = 0; i < n; ++i) pts[i].coords = {energy[i]};
    fast_optics::FastOPTICS fo(pts, min_pts);
    const std::vector<size_t> idx = fo.extractSuperCluster(fast_optics::ClusterMode::SUPER_CLUSTER_ONLY);
    for (size_t q = 0; q < idx.size(); ++q) out_idx[q] = (int)idx[q];
    *n_out = (int)idx.size();
    return 0;
}
