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
