/*
 * ref_driver.cpp — harness that runs the reference's own CPU contact
 * function on a system described by include/fa_types.h.
 *
 * TEST INFRASTRUCTURE ONLY.  It is linked against the reference sources
 * compiled in place from LIB/ (see oracle/ref/Makefile) and is used to pin
 * the oracle restatement; the shipped library never loads it.
 *
 * The in-memory system mirrors what LIB/top.cpp sets up before the GA:
 *   - receptor residues (type 0) in calc order: each flexible side chain is
 *     a residue of its own, the rigid atoms between them one residue per
 *     stretch; then the ligand (type 1);
 *   - atoms are 1-based, atom number == internal index (num_atm identity);
 *   - rotamers as build_rotamers lays them out (LIB/build_rotamers.cpp:
 *     103-172): each rotamer's copy of the residue's atoms appended after the
 *     last atom, keeping the atoms' numbers, fatm[rot]/latm[rot] naming the
 *     copy, the residue's bonded matrix (update_bonded) and optres (type 0)
 *     appended after the ligand's, optres pointers on the non-backbone atoms
 *     (update_optres), one typ-4 gene per side chain on fatm[0];
 *   - one optimised residue rebuilt by buildcc (the ligand), FA->mov[0];
 *   - VC buffers as LIB/top.cpp:207-211 and :1011-1043 allocate them;
 *   - a first ic2cf call (the "initial structure", LIB/top.cpp:1132), then
 *     VC->recalc = 0, then the poses — the steady state the GA sees.
 */
#include "flexaid.h"
#include "gaboom.h"
#include "Vcontacts.h"
#include "fa_types.h"

#include <cstdint>
#include <cstring>
#include <vector>

namespace {

struct RefSystem {
    FA_Global* FA = nullptr;
    VC_Global* VC = nullptr;
    std::vector<atom> atoms;
    std::vector<resid> residue;
    std::vector<gridpoint> grid;
    std::vector<int> num_atm, contacts, mov, fatm, latm;
    std::vector<float> contributions;
    std::vector<int*> bonded_rows;
    std::vector<int> bonded_data;
    std::vector<optmap> map_par;
    std::vector<OptRes> optres;
    std::vector<energy_matrix> emat;
    std::vector<energy_values> evals;
    std::vector<constraint> cons;
    std::vector<std::vector<constraint*>> cons_of;  // per atom number
    std::vector<std::vector<int>> rfatm, rlatm;     // per residue: fatm/latm per rotamer
    std::vector<std::vector<int>> flex_bonded_data;
    std::vector<std::vector<int*>> flex_bonded_rows;
    int Rn = 0, L = 0, N = 0, lig_res = 0;

    ~RefSystem()
    {
        if (VC) {
            free(VC->Calc);
            free(VC->Calclist);
            free(VC->ca_index);
            free(VC->seed);
            free(VC->contlist);
            free(VC->ca_rec);
            free(VC->ptorder);
            free(VC->centerpt);
            free(VC->poly);
            free(VC->cont);
            free(VC->vedge);
            free(VC);
        }
        delete FA;
    }
};

void setup(RefSystem& S, const fa_receptor_desc* rec, const fa_ligand_desc* lig, const fa_params* prm)
{
    S.Rn = rec->n_atoms;
    S.L = lig->n_atoms;
    S.N = S.Rn + S.L;
    const int N = S.N, Rn = S.Rn, L = S.L;
    S.FA = new FA_Global();
    FA_Global* FA = S.FA;
    int n_copies = 0;
    for (int k = 0; k < rec->n_flex; ++k) n_copies += (rec->flex[k].n_rot - 1) * rec->flex[k].n_atoms;
    S.atoms.assign(N + n_copies + 1, atom{});
    // residues in calc order: rigid stretches and side chains, then the ligand
    std::vector<int> res_of(Rn, 0), flex_res(rec->n_flex, 0);
    int nres = 0, pos = 0;
    S.rfatm.assign(1, {});
    S.rlatm.assign(1, {});
    auto add_res = [&](int lo, int hi) {  // 0-based receptor atoms [lo, hi)
        ++nres;
        S.rfatm.push_back({lo + 1});
        S.rlatm.push_back({hi});
        for (int i = lo; i < hi; ++i) res_of[i] = nres;
    };
    for (int k = 0; k < rec->n_flex; ++k) {
        const fa_flex_residue& f = rec->flex[k];
        if (f.first > pos) add_res(pos, f.first);
        add_res(f.first, f.first + f.n_atoms);
        flex_res[k] = nres;
        pos = f.first + f.n_atoms;
    }
    if (pos < Rn) add_res(pos, Rn);
    S.lig_res = nres + 1;
    S.rfatm.push_back({Rn + 1});
    S.rlatm.push_back({N});
    S.residue.assign(S.lig_res + 1, resid{});
    // rotamer copies after the last atom (fatm[r] = start of copy r)
    int next = N + 1;
    for (int k = 0; k < rec->n_flex; ++k) {
        const fa_flex_residue& f = rec->flex[k];
        for (int r = 1; r < f.n_rot; ++r) {
            S.rfatm[flex_res[k]].push_back(next);
            S.rlatm[flex_res[k]].push_back(next + f.n_atoms - 1);
            next += f.n_atoms;
        }
    }
    for (int q = 1; q <= S.lig_res; ++q) {
        S.residue[q].type = (q == S.lig_res) ? 1 : 0;
        S.residue[q].fatm = S.rfatm[q].data();
        S.residue[q].latm = S.rlatm[q].data();
        S.residue[q].rot = 0;
        S.residue[q].trot = (int)S.rfatm[q].size() - 1;
    }
    S.bonded_data.assign((size_t)L * L, -1);
    S.bonded_rows.resize(L);
    for (int i = 0; i < L; ++i) {
        S.bonded_rows[i] = &S.bonded_data[(size_t)i * L];
        for (int j = 0; j < L; ++j)
            if (lig->bonded[i * L + j]) S.bonded_rows[i][j] = (i == j) ? 0 : 1;
    }
    S.residue[S.lig_res].bonded = S.bonded_rows.data();
    S.flex_bonded_data.resize(rec->n_flex);
    S.flex_bonded_rows.resize(rec->n_flex);
    for (int k = 0; k < rec->n_flex; ++k) {
        const fa_flex_residue& f = rec->flex[k];
        const int n = f.n_atoms;
        S.flex_bonded_data[k].assign((size_t)n * n, -1);
        S.flex_bonded_rows[k].resize(n);
        for (int i = 0; i < n; ++i) {
            S.flex_bonded_rows[k][i] = &S.flex_bonded_data[k][(size_t)i * n];
            for (int j = 0; j < n; ++j)
                if (f.bonded[i * n + j]) S.flex_bonded_rows[k][i][j] = (i == j) ? 0 : 1;
        }
        S.residue[flex_res[k]].bonded = S.flex_bonded_rows[k].data();
    }
    // optimised residues: the ligand, then the side chains (build_rotamers appends them)
    S.optres.assign(1 + rec->n_flex, OptRes{});
    S.optres[0].rnum = S.lig_res;
    S.optres[0].type = 1;
    S.optres[0].tot = L;
    for (int k = 0; k < rec->n_flex; ++k) {
        S.optres[1 + k].rnum = flex_res[k];
        S.optres[1 + k].type = 0;
        S.optres[1 + k].tot = rec->flex[k].n_atoms;
    }
    FA->optres = S.optres.data();
    FA->num_optres = 1 + rec->n_flex;
    // atoms
    std::vector<int> flex_of(Rn, -1);
    for (int k = 0; k < rec->n_flex; ++k)
        for (int i = 0; i < rec->flex[k].n_atoms; ++i) flex_of[rec->flex[k].first + i] = k;
    for (int i = 0; i < Rn; ++i) {
        atom& a = S.atoms[1 + i];
        a.coor[0] = rec->xyz[3 * i];
        a.coor[1] = rec->xyz[3 * i + 1];
        a.coor[2] = rec->xyz[3 * i + 2];
        a.number = 1 + i;
        a.radius = rec->radius[i];
        a.type = rec->type[i];
        a.ofres = res_of[i];
        a.recs = 'r';
        a.acs = -1.0f;
        a.optres = nullptr;
        const int k = flex_of[i];
        if (k >= 0) {
            const fa_flex_residue& f = rec->flex[k];
            const int w = i - f.first;
            a.isbb = f.scorable[w] ? 0 : 1;
            if (f.scorable[w]) {
                a.optres = &S.optres[1 + k];
                a.recs = 'm';
            }
            for (int r = 1; r < f.n_rot; ++r) {  // the rotamer copy keeps the number
                atom& c = S.atoms[S.rfatm[flex_res[k]][r] + w];
                c = a;
                for (int q = 0; q < 3; ++q) c.coor[q] = f.rot_xyz[((size_t)r * f.n_atoms + w) * 3 + q];
            }
        }
    }
    for (int i = 0; i < L; ++i) {
        atom& a = S.atoms[Rn + 1 + i];
        a.number = Rn + 1 + i;
        a.radius = lig->radius[i];
        a.type = lig->type[i];
        a.ofres = S.lig_res;
        a.recs = 'm';
        a.acs = -1.0f;
        a.dis = lig->dis[i];
        a.ang = lig->ang[i];
        a.dih = lig->dih[i];
        a.shift = lig->shift[i];
        for (int k = 0; k < 4; ++k) {
            const int r = lig->rec[4 * i + k];
            a.rec[k] = (r >= 0) ? Rn + 1 + r : 0;
        }
        a.optres = &S.optres[0];
    }
    // distance constraints (read_constraints' result): every atom of a
    // constraint points at it, in constraint order
    S.cons.assign(lig->n_cons, constraint{});
    S.cons_of.assign(N + 1, {});
    for (int q = 0; q < lig->n_cons; ++q) {
        const fa_constraint& k = lig->cons[q];
        constraint& c = S.cons[q];
        c.id = q;
        c.type = k.type;
        c.bond_len = k.bond_len;
        c.max_dist = k.max_dist;
        const int a1 = Rn + 1 + k.atom;
        const int a2 = k.partner_is_ligand ? Rn + 1 + k.partner : 1 + k.partner;
        c.anum1 = a1;  // PDB numbers equal internal ones here (num_atm is the identity)
        c.anum2 = a2;
        c.inum1 = a1;  // internal numbers (assign_constraint, LIB/assign_constraint.cpp:92-93)
        c.inum2 = a2;
        S.cons_of[a1].push_back(&c);
        S.cons_of[a2].push_back(&c);
    }
    for (int i = 1; i <= N; ++i) {
        S.atoms[i].ncons = (int)S.cons_of[i].size();
        S.atoms[i].cons = S.cons_of[i].empty() ? nullptr : S.cons_of[i].data();
    }
    FA->constraints = S.cons.empty() ? nullptr : S.cons.data();
    FA->num_constraints = (int)S.cons.size();
    S.num_atm.assign(N + 2, 0);
    for (int i = 1; i <= N; ++i) S.num_atm[i] = i;
    FA->num_atm = S.num_atm.data();
    S.contacts.assign(MAX_ATOM_NUMBER, 0);
    FA->contacts = S.contacts.data();
    FA->atm_cnt = N + n_copies;
    FA->atm_cnt_real = N;
    FA->res_cnt = S.lig_res;
    // gene map (rotamer genes name the side chain's fatm[0])
    S.map_par.assign(lig->n_genes, optmap{});
    for (int g = 0; g < lig->n_genes; ++g) {
        S.map_par[g].typ = lig->gene_typ[g];
        S.map_par[g].atm = (lig->gene_typ[g] == FA_GENE_ROT) ? S.rfatm[flex_res[lig->gene_atm[g]]][0]
                                                             : Rn + 1 + lig->gene_atm[g];
    }
    FA->map_par = S.map_par.data();
    FA->npar = lig->n_genes;
    S.mov.resize(L);
    for (int i = 0; i < L; ++i) S.mov[i] = Rn + 1 + lig->mov[i];
    FA->nors = 1;
    FA->nmov[0] = L;
    FA->mov[0] = S.mov.data();
    FA->opt_res[0] = S.lig_res;
    // cleft grid
    S.grid.assign(lig->n_grid, gridpoint{});
    for (int g = 0; g < lig->n_grid; ++g) {
        S.grid[g].index = g;
        S.grid[g].dis = lig->grid_ic[3 * g];
        S.grid[g].ang = lig->grid_ic[3 * g + 1];
        S.grid[g].dih = lig->grid_ic[3 * g + 2];
    }
    FA->num_grd = lig->n_grid;
    for (int j = 0; j < 3; ++j) {
        FA->ori[j] = lig->ori[j];
        FA->globalmin[j] = lig->globalmin[j];
        FA->globalmax[j] = lig->globalmax[j];
    }
    FA->maxwidth = lig->maxwidth;
    // scoring parameters
    const int nt = prm->ntypes;
    FA->ntypes = nt;
    // weight-1 pairs: one value (read_emat, LIB/read_emat.cpp:86-101);
    // density pairs: their (x, y) nodes as a linked list (:103-138)
    const int n_dens = prm->dens_off ? prm->dens_off[nt * nt] : 0;
    S.emat.assign((size_t)nt * nt, energy_matrix{});
    S.evals.assign((size_t)nt * nt + n_dens, energy_values{});
    for (int i = 0; i < nt * nt; ++i) {
        S.emat[i].type1 = i / nt + 1;
        S.emat[i].type2 = i % nt + 1;
        const int b = prm->dens_off ? prm->dens_off[i] : 0, e = prm->dens_off ? prm->dens_off[i + 1] : 0;
        if (e > b) {
            for (int q = b; q < e; ++q) {
                energy_values& v = S.evals[(size_t)nt * nt + q];
                v.x = prm->dens_xy[2 * q];
                v.y = prm->dens_xy[2 * q + 1];
                v.next_value = (q + 1 < e) ? &S.evals[(size_t)nt * nt + q + 1] : nullptr;
            }
            S.emat[i].weight = 0;
            S.emat[i].energy_values = &S.evals[(size_t)nt * nt + b];
        } else {
            S.evals[i].x = -1.0f;
            S.evals[i].y = prm->emat[i];
            S.evals[i].next_value = nullptr;
            S.emat[i].weight = 1;
            S.emat[i].energy_values = &S.evals[i];
        }
    }
    FA->energy_matrix = S.emat.data();
    S.contributions.assign((size_t)nt * nt, 0.0f);
    FA->contributions = S.contributions.data();
    FA->permeability = prm->permeability;
    FA->dee_clash = prm->dee_clash;
    FA->intramolecular = prm->intramolecular;
    FA->solventterm = prm->solventterm;
    FA->normalize_area = prm->normalize_area ? 1 : 0;
    FA->useacs = 0;
    FA->vindex = 0;
    FA->bloops = 2;
    FA->normal_modes = 0;
    FA->useflexdee = 0;
    FA->deelig_flex = 0;
    FA->use_gist = 0;
    FA->use_hbond = 0;
    FA->use_elec = 0;
    FA->use_metal_coord = 0;
    // Vcontacts workspace (LIB/top.cpp:207-211, :1011-1043)
    VC_Global* VC = (VC_Global*)calloc(1, sizeof(VC_Global));
    S.VC = VC;
    VC->planedef = 'X';
    VC->Calc = (atomsas*)calloc(N, sizeof(atomsas));
    VC->Calclist = (int*)calloc(N, sizeof(int));
    VC->ca_index = (int*)calloc(N, sizeof(int));
    VC->seed = (int*)calloc(3 * N, sizeof(int));
    VC->contlist = (contactlist*)calloc(10000, sizeof(contactlist));
    VC->ca_recsize = 5 * N;
    VC->ca_rec = (ca_struct*)calloc(VC->ca_recsize, sizeof(ca_struct));
    VC->ptorder = (ptindex*)calloc(MAX_PT, sizeof(ptindex));
    VC->centerpt = (vertex*)calloc(MAX_PT, sizeof(vertex));
    VC->poly = (vertex*)calloc(MAX_POLY, sizeof(vertex));
    VC->cont = (plane*)calloc(MAX_PT, sizeof(plane));
    VC->vedge = (edgevector*)calloc(MAX_POLY, sizeof(edgevector));
    for (int i = 0; i < N; ++i) {
        VC->Calc[i].atom = nullptr;
        VC->Calc[i].residue = nullptr;
        VC->Calc[i].exposed = true;
    }
    VC->recalc = 0;
}

cfstr run_pose(RefSystem& S, const fa_ligand_desc* lig, const double* genes)
{
    std::vector<double> icv(lig->n_genes);
    for (int g = 0; g < lig->n_genes; ++g) {
        double v = genes[g];
        if (v > lig->gene_max[g]) v = lig->gene_max[g];
        if (v < lig->gene_min[g]) v = lig->gene_min[g];
        icv[g] = v;
    }
    return ic2cf(S.FA, S.VC, S.atoms.data(), S.residue.data(), S.grid.data(), lig->n_genes, icv.data());
}

}  // namespace

extern "C" {

/* The reference structures of a system (what LIB/top.cpp sets up), for
 * harnesses that drive reference-side code on them: an opaque handle, its
 * parts (FA_Global*, VC_Global*, atom* 1-based, resid* 1-based, gridpoint*),
 * and its release. */
void* fa_ref_system_create(const fa_receptor_desc* rec, const fa_ligand_desc* lig, const fa_params* prm)
{
    if (!rec || !lig || !prm) return nullptr;
    RefSystem* S = new RefSystem();
    setup(*S, rec, lig, prm);
    return S;
}

void fa_ref_system_parts(void* h, void** FA, void** VC, void** atoms, void** residue, void** grid)
{
    RefSystem* S = static_cast<RefSystem*>(h);
    *FA = S->FA;
    *VC = S->VC;
    *atoms = S->atoms.data();
    *residue = S->residue.data();
    *grid = S->grid.data();
}

void fa_ref_system_destroy(void* h) { delete static_cast<RefSystem*>(h); }

/* Evaluate poses serially with the reference ic2cf.  `lig_xyz` (optional)
 * receives the rebuilt ligand coordinates per pose; `recs`/`rec_counts`
 * (optional) receive the contact records of each pose (ca_rec lists of the
 * scorable atoms, calc-order indices), at most `recs_cap_per_pose` each. */
int fa_ref_eval(const fa_receptor_desc* rec, const fa_ligand_desc* lig, const fa_params* prm, int n_poses,
                const double* genes, fa_cf_result* out, float* lig_xyz, fa_contact_rec* recs, int recs_cap_per_pose,
                int* rec_counts)
{
    if (!rec || !lig || !prm || !genes || !out || n_poses < 1) return -1;
    RefSystem S;
    setup(S, rec, lig, prm);
    run_pose(S, lig, genes);  // initial structure call (LIB/top.cpp:1132)
    S.VC->recalc = 0;
    for (int p = 0; p < n_poses; ++p) {
        const int fail0 = S.FA->skipped, clash0 = S.FA->clashed;
        cfstr cf = run_pose(S, lig, genes + (size_t)p * lig->n_genes);
        fa_cf_result& r = out[p];
        r.com = cf.com;
        r.wal = cf.wal;
        r.sas = cf.sas;
        r.con = cf.con;
        r.totsas = 0.0;
        r.rclash = 0;
        for (int q = 0; q < S.FA->num_optres; ++q) {  // every scored residue
            r.totsas += S.FA->optres[q].cf.totsas;
            r.rclash |= S.FA->optres[q].cf.rclash;
        }
        r.cf = get_cf_evalue(&cf);
        r.app = get_apparent_cf_evalue(&cf);
        r.status = (S.FA->skipped != fail0) ? FA_HULL_FAIL : (S.FA->clashed != clash0) ? FA_CLASH : FA_OK;
        if (r.status != FA_OK) {
            r.totsas = 0.0;
            r.rclash = 0;
        }
        if (lig_xyz)
            for (int i = 0; i < S.L; ++i)
                for (int k = 0; k < 3; ++k) lig_xyz[((size_t)p * S.L + i) * 3 + k] = S.atoms[S.Rn + 1 + i].coor[k];
        if (rec_counts) {
            int n = 0;
            if (r.status == FA_OK) {
                for (int i = 0; i < S.N; ++i) {  // scorable atoms in calc order
                    if (!S.VC->Calc[i].score) continue;  // receptor atoms hold reverse stubs only
                    for (int ci = S.VC->ca_index[i]; ci != -1; ci = S.VC->ca_rec[ci].prev) {
                        if (recs && n < recs_cap_per_pose) {
                            fa_contact_rec& o = recs[(size_t)p * recs_cap_per_pose + n];
                            o.from = S.VC->ca_rec[ci].from;
                            o.atom = S.VC->ca_rec[ci].atom;
                            o.area = S.VC->ca_rec[ci].area;
                            o.dist = S.VC->ca_rec[ci].dist;
                        }
                        ++n;
                    }
                }
            }
            rec_counts[p] = n;
        }
    }
    return 0;
}

}  // extern "C"
