/*
 * ref_reader.cpp — harness that runs the reference's own receptor-file chain
 * on a PDB file, in the order LIB/top.cpp:582-646 and :856 call it:
 * rna_structure, modify_pdb (into tmp_path), read_pdb, the last residue's
 * end, calc_center, residue_conect (proteins), assign_types,
 * assign_radii_types.
 *
 * TEST INFRASTRUCTURE ONLY.  Linked against the reference sources compiled
 * in place from LIB/ (oracle/ref/Makefile); it pins the receptor reader of
 * the package, and the shipped library never loads it.
 *
 * Returns the atom count (atoms copied out in the reference's order), or -1
 * when it exceeds max_atoms.  center = ori[3], globalmin[3], globalmax[3],
 * maxdst, maxwidth.
 */
#include "flexaid.h"

#include <cstring>

extern "C" int fa_ref_read_receptor(const char* pdb, const char* amino_def, const char* nucl_def,
                                    const char* tmp_path, int ntypes, int exclude_het, int remove_water,
                                    int keep_ions, int keep_structural_waters, float bfactor_max,
                                    int max_atoms, float* xyz, float* radius, int* type, int* number,
                                    int* ofres, int* nbond, int* is_protein, float* center)
{
    FA_Global* FA = new FA_Global();
    FA->MIN_NUM_ATOM = 1000;  // LIB/top.cpp:222-226
    FA->MIN_NUM_RESIDUE = 250;
    FA->MIN_ROTAMER = 1;
    FA->MIN_FLEX_BONDS = 5;
    FA->ntypes = ntypes;
    FA->exclude_het = exclude_het;
    FA->remove_water = remove_water;
    FA->keep_ions = keep_ions;
    FA->keep_structural_waters = keep_structural_waters;
    FA->structural_water_bfactor_max = bfactor_max;
    FA->is_protein = 1;

    char infile[1024], tmpfile[1024], deftyp[1024];
    std::strncpy(infile, pdb, sizeof infile - 1);
    infile[sizeof infile - 1] = '\0';
    std::strncpy(tmpfile, tmp_path, sizeof tmpfile - 1);
    tmpfile[sizeof tmpfile - 1] = '\0';
    if (rna_structure(infile)) FA->is_protein = 0;
    std::strncpy(deftyp, FA->is_protein ? amino_def : nucl_def, sizeof deftyp - 1);
    deftyp[sizeof deftyp - 1] = '\0';

    atom* atoms = nullptr;
    resid* residue = nullptr;
    modify_pdb(infile, tmpfile, FA->exclude_het, FA->remove_water, FA->is_protein, FA->keep_ions,
               FA->keep_structural_waters, FA->structural_water_bfactor_max);
    read_pdb(FA, &atoms, &residue, tmpfile);
    residue[FA->res_cnt].latm[0] = FA->atm_cnt;
    calc_center(FA, atoms, residue);
    if (FA->is_protein) residue_conect(FA, atoms, residue, deftyp);
    assign_types(FA, atoms, residue, deftyp);
    assign_radii_types(FA, atoms, residue);

    const int n = FA->atm_cnt;
    *is_protein = FA->is_protein;
    if (n > max_atoms) return -1;
    for (int i = 1; i <= n; ++i) {
        for (int j = 0; j < 3; ++j) xyz[3 * (i - 1) + j] = atoms[i].coor[j];
        radius[i - 1] = atoms[i].radius;
        type[i - 1] = atoms[i].type;
        number[i - 1] = atoms[i].number;
        ofres[i - 1] = atoms[i].ofres;
        nbond[i - 1] = atoms[i].bond[0];
    }
    for (int j = 0; j < 3; ++j) {
        center[j] = FA->ori[j];
        center[3 + j] = FA->globalmin[j];
        center[6 + j] = FA->globalmax[j];
    }
    center[9] = FA->maxdst;
    center[10] = FA->maxwidth;
    return n;
}

/*
 * The reference's automatic binding-site grid (LIB/top.cpp:904-916):
 * detect_cleft over the receptor atoms (1-based array, atm_cnt_real = n),
 * generate_grid over its spheres, calc_cleftic in the ori frame.  Built
 * without OpenMP, so the probes come in the serial branch's order.
 * Returns the number of grid points (cleftgrid[1 .. num_grd-1] copied out),
 * -1 when no cleft is found, -2 when the output arrays are too small.
 */
#include "CleftDetector.h"

#include <vector>

gridpoint* generate_grid(FA_Global* FA, sphere* spheres, atom* atoms, resid* residue);
void calc_cleftic(FA_Global* FA, gridpoint* cleftgrid);

extern "C" int fa_ref_cleft_grid(const float* xyz, const float* radius, int n, const float* ori, float spacer,
                                 int max_points, float* points, float* ic, int max_spheres, float* spheres,
                                 int* n_spheres)
{
    std::vector<atom> atoms(n + 1);
    std::memset(atoms.data(), 0, sizeof(atom) * (n + 1));
    for (int i = 1; i <= n; ++i) {
        for (int j = 0; j < 3; ++j) atoms[i].coor[j] = xyz[3 * (i - 1) + j];
        atoms[i].radius = radius[i - 1];
    }
    std::vector<resid> residue(2);
    std::memset(residue.data(), 0, sizeof(resid) * 2);
    int gpa[3] = {1, 1, 1};
    residue[1].gpa = gpa;
    FA_Global* FA = new FA_Global();
    FA->res_cnt = 1;
    FA->MIN_CLEFTGRID_POINTS = 100;
    FA->spacer_length = spacer;
    for (int j = 0; j < 3; ++j) FA->ori[j] = ori[j];

    sphere* sph = detect_cleft(atoms.data(), residue.data(), n, 1);
    if (!sph) return -1;
    int ns = 0;
    for (sphere* s = sph; s; s = s->prev, ++ns) {
        if (ns < max_spheres) {
            spheres[4 * ns + 0] = s->center[0];
            spheres[4 * ns + 1] = s->center[1];
            spheres[4 * ns + 2] = s->center[2];
            spheres[4 * ns + 3] = s->radius;
        }
    }
    *n_spheres = ns;
    gridpoint* g = generate_grid(FA, sph, atoms.data(), residue.data());
    calc_cleftic(FA, g);
    free_sphere_list(sph);
    const int ng = FA->num_grd - 1;
    if (ng > max_points || ns > max_spheres) return -2;
    for (int i = 1; i <= ng; ++i) {
        for (int j = 0; j < 3; ++j) points[3 * (i - 1) + j] = g[i].coor[j];
        ic[3 * (i - 1) + 0] = g[i].dis;
        ic[3 * (i - 1) + 1] = g[i].ang;
        ic[3 * (i - 1) + 2] = g[i].dih;
    }
    free(g);
    return ng;
}

/*
 * The reference's direct-mode ligand set-up from a MOL2 file, in the order
 * LIB/top.cpp:845-897 and :982-998 call it: read_mol2_ligand (atoms, SYBYL
 * types and radii, bonds, BFS reconstruction tree, buildic, bonded matrix
 * within bloops), assign_radii_types (which leaves the ligand's radii),
 * the GPA block (ori = the ligand's centroid, gpa = its first three atoms,
 * buildic_point of the first; restated from top.cpp:864-897, inline there),
 * ic_bounds for an n_grid-point cleft grid ("locclf"), add2_optimiz_vec for
 * the translation gene (opt[1] = -1) and the three rotation genes
 * (opt[1] = 0), and create_rebuild_list.  Ligand atoms come out in file
 * order, their rec[0..2] as ligand indices (-1 for the ori frame); bonded is
 * the residue's matrix (bonds away, -1 beyond bloops); mov the rebuild order;
 * the genes' typ, atom, min and max; grid0 the cleft grid's row 0 (the GPA
 * atom: coordinates, then dis/ang/dih).  Returns the atom count, -1 when
 * the reader fails, -2 when an output array is too small.
 */
int read_mol2_ligand(FA_Global* FA, atom** atoms, resid** residue, const char* mol2_file);
void buildic_point(FA_Global* FA, float coor[], float* dis, float* ang, float* dih);

extern "C" int fa_ref_read_mol2(const char* mol2, int n_grid, int max_atoms, float* xyz, float* radius, int* type,
                                int* rec, float* ic, int* bonded, int* mov, int* n_mov, int max_genes, int* gene_typ,
                                int* gene_atm, double* gene_min, double* gene_max, int* n_genes, float* ori,
                                float* grid0)
{
    FA_Global* FA = new FA_Global();
    FA->MIN_NUM_ATOM = 1000;  // LIB/top.cpp:222-230
    FA->MIN_NUM_RESIDUE = 250;
    FA->MIN_ROTAMER = 1;
    FA->MIN_FLEX_BONDS = 5;
    FA->MIN_OPTRES = 1;
    FA->MIN_PAR = 6;
    FA->ntypes = 40;
    FA->bloops = 2;           // LIB/top.cpp:296
    FA->htpmode = true;       // no per-atom printing in buildlist
    FA->num_grd = n_grid;
    FA->num_atm = (int*)calloc(MAX_ATOM_NUMBER, sizeof(int));
    atom* atoms = (atom*)calloc(FA->MIN_NUM_ATOM, sizeof(atom));
    resid* residue = (resid*)calloc(FA->MIN_NUM_RESIDUE, sizeof(resid));
    if (!read_mol2_ligand(FA, &atoms, &residue, mol2)) return -1;
    assign_radii_types(FA, atoms, residue);
    const int lr = FA->res_cnt;
    const int fa = residue[lr].fatm[0], la = residue[lr].latm[0], n = la - fa + 1;
    if (n > max_atoms) return -2;
    // LIB/top.cpp:864-897 (the GPA block of direct mode)
    FA->ori[0] = FA->ori[1] = FA->ori[2] = 0.0f;
    for (int ai = fa; ai <= la; ai++) {
        FA->ori[0] += atoms[ai].coor[0];
        FA->ori[1] += atoms[ai].coor[1];
        FA->ori[2] += atoms[ai].coor[2];
    }
    if (n > 0) {
        FA->ori[0] /= n;
        FA->ori[1] /= n;
        FA->ori[2] /= n;
    }
    int* gpa = (int*)malloc(3 * sizeof(int));
    gpa[0] = fa;
    gpa[1] = (n > 1) ? fa + 1 : fa;
    gpa[2] = (n > 2) ? fa + 2 : fa;
    residue[lr].gpa = gpa;
    buildic_point(FA, atoms[fa].coor, &atoms[fa].dis, &atoms[fa].ang, &atoms[fa].dih);
    // the cleft grid's row 0 (LIB/generate_grid.cpp:24-28)
    for (int j = 0; j < 3; ++j) grid0[j] = atoms[gpa[0]].coor[j];
    grid0[3] = atoms[gpa[0]].dis;
    grid0[4] = atoms[gpa[0]].ang;
    grid0[5] = atoms[gpa[0]].dih;
    // LIB/top.cpp:940-998: parameter vectors, bounds, translation and rotation genes
    FA->map_par = (optmap*)calloc(FA->MIN_PAR, sizeof(optmap));
    FA->opt_par = (double*)calloc(FA->MIN_PAR, sizeof(double));
    FA->del_opt_par = (double*)calloc(FA->MIN_PAR, sizeof(double));
    FA->min_opt_par = (double*)calloc(FA->MIN_PAR, sizeof(double));
    FA->max_opt_par = (double*)calloc(FA->MIN_PAR, sizeof(double));
    FA->map_opt_par = (int*)calloc(FA->MIN_PAR, sizeof(int));
    char rngopt[] = "locclf";
    ic_bounds(FA, rngopt);
    int opt[2] = {FA->resligand->number, -1};
    add2_optimiz_vec(FA, atoms, residue, opt, ' ', "");
    opt[1] = 0;
    add2_optimiz_vec(FA, atoms, residue, opt, ' ', "");
    create_rebuild_list(FA, atoms, residue);
    // outputs
    for (int i = 0; i < n; ++i) {
        const atom& a = atoms[fa + i];
        for (int j = 0; j < 3; ++j) {
            xyz[3 * i + j] = a.coor[j];
            rec[3 * i + j] = a.rec[j] == 0 ? -1 : a.rec[j] - fa;
        }
        radius[i] = a.radius;
        type[i] = a.type;
        ic[3 * i + 0] = a.dis;
        ic[3 * i + 1] = a.ang;
        ic[3 * i + 2] = a.dih;
        for (int j = 0; j < n; ++j) bonded[i * n + j] = residue[lr].bonded[i][j];
    }
    *n_mov = FA->nmov[0];
    for (int k = 0; k < FA->nmov[0]; ++k) mov[k] = FA->mov[0][k] - fa;
    *n_genes = FA->npar;
    if (FA->npar > max_genes) return -2;
    for (int g = 0; g < FA->npar; ++g) {
        gene_typ[g] = FA->map_par[g].typ;
        gene_atm[g] = FA->map_par[g].atm - fa;
        gene_min[g] = FA->min_opt_par[g];
        gene_max[g] = FA->max_opt_par[g];
    }
    for (int j = 0; j < 3; ++j) ori[j] = FA->ori[j];
    return n;
}

/*
 * The reference's SDF reader (LIB/SdfReader.cpp:63-263) on a file: atoms in
 * file order (coordinates, element-derived type and radius) and each atom's
 * bond list as read_sdf_ligand leaves it (ligand indices, at most six).
 * Returns the atom count, -1 when the reader fails, -2 when an output array
 * is too small.
 */
int read_sdf_ligand(FA_Global* FA, atom** atoms, resid** residue, const char* sdf_file);

extern "C" int fa_ref_read_sdf(const char* sdf, int max_atoms, float* xyz, float* radius, int* type, int* bonds7)
{
    FA_Global* FA = new FA_Global();
    FA->MIN_NUM_ATOM = 1000;
    FA->MIN_NUM_RESIDUE = 250;
    FA->MIN_FLEX_BONDS = 5;
    FA->MIN_OPTRES = 1;
    FA->num_atm = (int*)calloc(MAX_ATOM_NUMBER, sizeof(int));
    atom* atoms = (atom*)calloc(FA->MIN_NUM_ATOM, sizeof(atom));
    resid* residue = (resid*)calloc(FA->MIN_NUM_RESIDUE, sizeof(resid));
    if (!read_sdf_ligand(FA, &atoms, &residue, sdf)) return -1;
    const int fa = residue[FA->res_cnt].fatm[0], la = residue[FA->res_cnt].latm[0], n = la - fa + 1;
    if (n > max_atoms) return -2;
    for (int i = 0; i < n; ++i) {
        const atom& a = atoms[fa + i];
        for (int j = 0; j < 3; ++j) xyz[3 * i + j] = a.coor[j];
        radius[i] = a.radius;
        type[i] = a.type;
        bonds7[7 * i] = a.bond[0];
        for (int k = 1; k <= 6; ++k) bonds7[7 * i + k] = k <= a.bond[0] ? a.bond[k] - fa : -1;
    }
    return n;
}
