/*
 * bridge_check.cpp — harness for include/bridge/flexaids_rocm_bridge.h, the
 * reference-side bridge, on the reference structures ref_driver.cpp builds
 * (what LIB/top.cpp sets up before the GA).
 *
 * TEST INFRASTRUCTURE ONLY.  Linked with the reference sources compiled in
 * place (oracle/ref/Makefile), ref_driver.o and libflexaidds_amd.so.
 *
 *   fa_bridge_compare     descriptors A → reference structures →
 *                         system_from_reference → descriptors, compared field
 *                         by field with descriptors B (A itself for a round
 *                         trip; no device work);
 *   fa_bridge_batch_eval  flexaids_rocm::batch_eval, voronoi_cf::batch_eval's
 *                         signature, over chromosomes (needs a GPU);
 *   fa_bridge_dispatch    the flexaids:: dispatch entries of the bridge
 *                         (needs a GPU).
 */
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "bridge/flexaids_rocm_bridge.h"

extern "C" {
void* fa_ref_system_create(const fa_receptor_desc* rec, const fa_ligand_desc* lig, const fa_params* prm);
void fa_ref_system_parts(void* h, void** FA, void** VC, void** atoms, void** residue, void** grid);
void fa_ref_system_destroy(void* h);
}

namespace {

struct Ref {
    void* h = nullptr;
    FA_Global* FA = nullptr;
    VC_Global* VC = nullptr;
    atom* atoms = nullptr;
    resid* residue = nullptr;
    gridpoint* grid = nullptr;
    std::vector<genlim> gl;
    Ref(const fa_receptor_desc* rec, const fa_ligand_desc* lig, const fa_params* prm)
    {
        h = fa_ref_system_create(rec, lig, prm);
        void *a, *b, *c, *d, *e;
        fa_ref_system_parts(h, &a, &b, &c, &d, &e);
        FA = (FA_Global*)a;
        VC = (VC_Global*)b;
        atoms = (atom*)c;
        residue = (resid*)d;
        grid = (gridpoint*)e;
        gl.assign((size_t)lig->n_genes, genlim{});
        for (int g = 0; g < lig->n_genes; ++g) {
            gl[(size_t)g].min = lig->gene_min[g];
            gl[(size_t)g].max = lig->gene_max[g];
        }
    }
    ~Ref() { fa_ref_system_destroy(h); }
};

struct Diff {
    std::string what;
    bool any = false;
    void note(const char* field, long long at)
    {
        if (any) return;
        any = true;
        what = std::string(field) + " [" + std::to_string(at) + "]";
    }
};

template <class T>
void cmp(Diff& d, const char* field, const T* a, const T* b, long long n)
{
    for (long long i = 0; i < n; ++i)
        if (std::memcmp(&a[i], &b[i], sizeof(T)) != 0) return d.note(field, i);
}

void cmp_flags(Diff& d, const char* field, const uint8_t* a, const uint8_t* b, long long n)
{
    for (long long i = 0; i < n; ++i)
        if ((a[i] != 0) != (b[i] != 0)) return d.note(field, i);
}

int report(const Diff& d, char* err, int len)
{
    if (!d.any) return 0;
    if (err && len > 0) std::snprintf(err, (size_t)len, "%s", d.what.c_str());
    return 1;
}

}  // namespace

extern "C" {

int fa_bridge_compare(const fa_receptor_desc* rec_a, const fa_ligand_desc* lig_a, const fa_params* prm_a,
                      const fa_receptor_desc* rec, const fa_ligand_desc* lig, const fa_params* prm, char* err, int len)
{
    try {
        Ref R(rec_a, lig_a, prm_a);
        flexaids_rocm::FaSystem S;
        flexaids_rocm::system_from_reference(R.FA, R.atoms, R.residue, R.grid, R.gl, S);
        Diff d;
        const fa_receptor_desc& r = S.rec;
        if (r.n_atoms != rec->n_atoms) d.note("receptor.n_atoms", r.n_atoms);
        if (r.n_flex != rec->n_flex) d.note("receptor.n_flex", r.n_flex);
        if (!d.any) {
            cmp(d, "receptor.xyz", r.xyz, rec->xyz, 3LL * rec->n_atoms);
            cmp(d, "receptor.radius", r.radius, rec->radius, rec->n_atoms);
            cmp(d, "receptor.type", r.type, rec->type, rec->n_atoms);
            for (int k = 0; k < rec->n_flex && !d.any; ++k) {
                const fa_flex_residue &a = r.flex[k], &b = rec->flex[k];
                if (a.first != b.first || a.n_atoms != b.n_atoms || a.n_rot != b.n_rot) {
                    d.note("flex.shape", k);
                    break;
                }
                cmp_flags(d, "flex.scorable", a.scorable, b.scorable, b.n_atoms);
                cmp(d, "flex.rot_xyz", a.rot_xyz, b.rot_xyz, 3LL * b.n_rot * b.n_atoms);
                cmp_flags(d, "flex.bonded", a.bonded, b.bonded, (long long)b.n_atoms * b.n_atoms);
            }
        }
        const fa_ligand_desc& l = S.lig;
        if (l.n_atoms != lig->n_atoms || l.n_genes != lig->n_genes || l.n_grid != lig->n_grid ||
            l.n_cons != lig->n_cons)
            d.note("ligand.shape", l.n_atoms);
        if (!d.any) {
            const long long L = lig->n_atoms, G = lig->n_genes;
            cmp(d, "ligand.rec", l.rec, lig->rec, 4 * L);
            cmp(d, "ligand.dis", l.dis, lig->dis, L);
            cmp(d, "ligand.ang", l.ang, lig->ang, L);
            cmp(d, "ligand.dih", l.dih, lig->dih, L);
            cmp(d, "ligand.shift", l.shift, lig->shift, L);
            cmp(d, "ligand.radius", l.radius, lig->radius, L);
            cmp(d, "ligand.type", l.type, lig->type, L);
            cmp(d, "ligand.mov", l.mov, lig->mov, L);
            cmp_flags(d, "ligand.bonded", l.bonded, lig->bonded, L * L);
            cmp(d, "ligand.gene_typ", l.gene_typ, lig->gene_typ, G);
            cmp(d, "ligand.gene_atm", l.gene_atm, lig->gene_atm, G);
            cmp(d, "ligand.gene_min", l.gene_min, lig->gene_min, G);
            cmp(d, "ligand.gene_max", l.gene_max, lig->gene_max, G);
            cmp(d, "ligand.grid_ic", l.grid_ic, lig->grid_ic, 3LL * lig->n_grid);
            cmp(d, "ligand.ori", l.ori, lig->ori, 3);
            cmp(d, "ligand.globalmin", l.globalmin, lig->globalmin, 3);
            cmp(d, "ligand.globalmax", l.globalmax, lig->globalmax, 3);
            cmp(d, "ligand.maxwidth", &l.maxwidth, &lig->maxwidth, 1);
            for (int q = 0; q < lig->n_cons && !d.any; ++q) {
                const fa_constraint &a = l.cons[q], &b = lig->cons[q];
                if (a.atom != b.atom || a.partner != b.partner || a.partner_is_ligand != b.partner_is_ligand ||
                    a.type != b.type || a.bond_len != b.bond_len || a.max_dist != b.max_dist)
                    d.note("ligand.cons", q);
            }
        }
        const fa_params& p = S.prm;
        if (p.permeability != prm->permeability || p.dee_clash != prm->dee_clash ||
            p.intramolecular != prm->intramolecular || p.solventterm != prm->solventterm ||
            p.ntypes != prm->ntypes || (p.normalize_area != 0) != (prm->normalize_area != 0))
            d.note("params.scalars", 0);
        if (!d.any) {
            const long long nt = prm->ntypes;
            for (long long t = 0; t < nt * nt && !d.any; ++t) {  // density pairs carry no single value
                const bool dens = prm->dens_off && prm->dens_off[t + 1] > prm->dens_off[t];
                if (!dens) cmp(d, "params.emat", p.emat + t, prm->emat + t, 1);
            }
            const bool da = p.dens_off != nullptr, db = prm->dens_off != nullptr && prm->dens_off[nt * nt] > 0;
            if (da != db)
                d.note("params.dens_off", -1);
            else if (da) {
                cmp(d, "params.dens_off", p.dens_off, prm->dens_off, nt * nt + 1);
                if (!d.any) cmp(d, "params.dens_xy", p.dens_xy, prm->dens_xy, 2LL * prm->dens_off[nt * nt]);
            }
        }
        return report(d, err, len);
    } catch (const std::exception& e) {
        if (err && len > 0) std::snprintf(err, (size_t)len, "exception: %s", e.what());
        return -1;
    }
}

/* cf: per chromosome the cfstr fields com, con, wal, sas, elec, gist, hbond,
 * gist_desolv, metal_coord, totsas, rclash (11 doubles); app: app_evalue. */
int fa_bridge_batch_eval(const fa_receptor_desc* rec, const fa_ligand_desc* lig, const fa_params* prm, int n,
                         const double* genes, const char* status, double* cf, double* app, double* wall_ms, char* err,
                         int len)
{
    try {
        Ref R(rec, lig, prm);
        const int ng = lig->n_genes;
        GB_Global GB{};
        GB.num_genes = ng;
        std::vector<gene> store((size_t)n * ng);
        std::vector<chromosome> chroms((size_t)n);
        for (int i = 0; i < n; ++i) {
            for (int k = 0; k < ng; ++k) store[(size_t)i * ng + k].to_ic = genes[(size_t)i * ng + k];
            chroms[(size_t)i].genes = &store[(size_t)i * ng];
            chroms[(size_t)i].status = status ? status[i] : 'o';
        }
        const voronoi_cf::BatchResult res =
            flexaids_rocm::batch_eval(chroms, R.FA, &GB, R.VC, R.gl, R.atoms, R.residue, R.grid, ic2cf);
        for (int i = 0; i < n; ++i) {
            const cfstr& c = res.cf[(size_t)i];
            const double v[11] = {c.com, c.con, c.wal, c.sas, c.elec, c.gist, c.hbond, c.gist_desolv,
                                  c.metal_coord, c.totsas, (double)c.rclash};
            for (int k = 0; k < 11; ++k) cf[(size_t)i * 11 + k] = v[k];
            app[i] = res.app_evalue[(size_t)i];
        }
        *wall_ms = res.wall_ms;
        flexaids_rocm::reset();
        return 0;
    } catch (const std::exception& e) {
        flexaids_rocm::reset();
        if (err && len > 0) std::snprintf(err, (size_t)len, "exception: %s", e.what());
        return -1;
    }
}

/* compute_boltzmann_batch, log_sum_exp_dispatch and select_backend of the
 * bridge; `weights` [n]; out3 = {log_Z, E_min, lse}; backend = the
 * HardwareBackend value; telemetry_backend = the result's telemetry backend. */
int fa_bridge_dispatch(const double* energies, int n, double beta, double* weights, double* out3, int* backend,
                       int* telemetry_backend, char* err, int len)
{
    try {
        const std::span<const double> e(energies, (size_t)n);
        const hw::BoltzmannBatchResult b = flexaids_rocm::compute_boltzmann_batch(e, beta);
        for (size_t i = 0; i < b.weights.size(); ++i) weights[i] = b.weights[i];
        out3[0] = b.log_Z;
        out3[1] = b.E_min;
        out3[2] = flexaids_rocm::log_sum_exp_dispatch(e);
        // the harness links no UnifiedHardwareDispatch.cpp: without a device it
        // would report the scalar backend
        *backend = (int)flexaids_rocm::select_backend([] { return flexaids::HardwareBackend::SCALAR; });
        *telemetry_backend = (int)b.telemetry.backend;
        flexaids_rocm::reset();
        return 0;
    } catch (const std::exception& e) {
        flexaids_rocm::reset();
        if (err && len > 0) std::snprintf(err, (size_t)len, "exception: %s", e.what());
        return -1;
    }
}

}  // extern "C"
