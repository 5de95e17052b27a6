/*
 * ref_cluster.cpp — harness that runs the reference's own post-GA pose
 * clustering (cluster(), LIB/cluster.cpp, CLUSTA CF) on poses of a system
 * described by include/fa_types.h.
 *
 * TEST INFRASTRUCTURE ONLY (linked with the reference sources compiled in
 * place, oracle/ref/Makefile).  The system is the one ref_driver.cpp sets
 * up; the chromosomes carry the given genes (to_ic) and energies.  cluster()
 * writes `<prefix>.cad` (one line per cluster: TOP, TCF, ACF, freq; then the
 * RMSD between cluster heads) and one PDB file per result, which the caller
 * reads.
 */
#include "flexaid.h"
#include "gaboom.h"
#include "fa_types.h"

#include <cstdio>
#include <vector>

extern "C" {
void* fa_ref_system_create(const fa_receptor_desc* rec, const fa_ligand_desc* lig, const fa_params* prm);
void fa_ref_system_parts(void* h, void** FA, void** VC, void** atoms, void** residue, void** grid);
void fa_ref_system_destroy(void* h);

int fa_ref_cluster(const fa_receptor_desc* rec, const fa_ligand_desc* lig, const fa_params* prm, int n,
                   const double* ic, const double* app, const double* evalue, unsigned temperature, float rmsd,
                   int max_results, const char* prefix)
{
    if (!rec || !lig || !prm || !ic || !app || !evalue || !prefix || n < 1 || max_results < 1 || max_results > n)
        return -1;
    void* h = fa_ref_system_create(rec, lig, prm);
    if (!h) return -1;
    void *pFA, *pVC, *pat, *pres, *pgrid;
    fa_ref_system_parts(h, &pFA, &pVC, &pat, &pres, &pgrid);
    FA_Global* FA = static_cast<FA_Global*>(pFA);
    const int G = lig->n_genes;
    std::vector<double> opt(G, 0.0);  // FA->opt_par as LIB/top.cpp allocates it
    FA->opt_par = opt.data();
    FA->temperature = temperature;
    FA->beta = temperature ? 1.0 / (double)temperature : 0.0;  // LIB/read_input.cpp:242-244
    FA->cluster_rmsd = rmsd;
    FA->max_results = max_results;
    FA->htpmode = false;
    FA->refstructure = 0;
    GB_Global GB{};
    GB.num_genes = G;
    std::vector<genlim> lim(G);
    std::vector<gene> genes((size_t)n * G);
    std::vector<chromosome> chrom(n);
    for (int i = 0; i < n; ++i) {
        chrom[i] = chromosome{};
        chrom[i].genes = &genes[(size_t)i * G];
        for (int g = 0; g < G; ++g) genes[(size_t)i * G + g].to_ic = ic[(size_t)i * G + g];
        chrom[i].app_evalue = app[i];
        chrom[i].evalue = evalue[i];
    }
    char end[MAX_PATH__], tmp[MAX_PATH__], dockinp[] = "dock.inp", gainp[] = "ga.inp";
    std::snprintf(end, sizeof end, "%s", prefix);
    cluster(FA, &GB, static_cast<VC_Global*>(pVC), chrom.data(), lim.data(), static_cast<atom*>(pat),
            static_cast<resid*>(pres), static_cast<gridpoint*>(pgrid), n, end, tmp, dockinp, gainp);
    FA->opt_par = nullptr;
    fa_ref_system_destroy(h);
    return 0;
}

}  // extern "C"
