/*
 * fa_types.h — plain-data descriptors of one docking system for the
 * FlexAID∆S population-evaluation path (pose build → Voronoi contact
 * function → ensemble statistics).
 *
 * These structs carry what the reference's setup stage produces and its hot
 * path consumes (FA_Global / atom / resid / gridpoint / energy_matrix in
 * LIB/flexaid.h:122-572).  They hold borrowed pointers only; every consumer
 * (the HIP library, the CPU oracle, the reference harness) copies what it
 * needs.  All indices are 0-based and ligand-local unless stated otherwise.
 *
 * Atom ordering ("calc order") follows LIB/Vcontacts.cpp:1594-1625
 * (index_protein walks residues 1..res_cnt): receptor atoms first, then the
 * ligand atoms in their input order.
 */
#ifndef FA_TYPES_H
#define FA_TYPES_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Hard limits of the device implementation (the reference allows 100 genes,
 * LIB/gaboom.h:33, and MAX_ATM_HET=200 hetero atoms, LIB/flexaid.h:62). */
#define FA_MAX_LIG_ATOMS 128
#define FA_MAX_GENES 64

/* gene → internal-coordinate mapping types (optmap.typ, LIB/flexaid.h:192-197,
 * decoded at LIB/ic2cf.cpp:53-105) */
enum fa_gene_type {
    FA_GENE_GRID = -1, /* index into the cleft grid: sets dis/ang/dih of atm */
    FA_GENE_DIS = 0,
    FA_GENE_ANG = 1,
    FA_GENE_DIH = 2,   /* dihedral; cascades along rec[3] shift chains */
    FA_GENE_ROT = 4    /* rotamer of flexible side chain `atm` (an index into
                          fa_receptor_desc.flex): (unsigned)(v + 0.5f),
                          LIB/ic2cf.cpp:90-103 */
};

/* Scorable atoms of one pose: every ligand atom plus the side-chain atoms of
 * the flexible residues, at most FA_MAX_SCORABLE in all; a flexible residue
 * has at most FA_MAX_FLEX_ATOMS atoms. */
#define FA_MAX_FLEX_RESIDUES 32
#define FA_MAX_FLEX_ATOMS 128
#define FA_MAX_SCORABLE 128

/* Per-pose evaluation status.  FA_OK, FA_HULL_FAIL and FA_CLASH carry the
 * reference's energies (the clash and hull-failure cfstr of LIB/ic2cf.cpp:129
 * with LIB/vcfunction.cpp:56-75).  The others mark poses the device did not
 * score: their cf and app are +inf, so they rank after every scored pose in
 * GA selection and carry Boltzmann weight 0 in every ensemble reduction. */
enum fa_status {
    FA_OK = 0,
    FA_HULL_FAIL = 1,     /* Voronoi hull did not converge → POLYHEDRON_PENALTY */
    FA_CLASH = 2,         /* clash pre-pass ≥ CLASH_THRESHOLD (decided before any capacity check) */
    FA_CAPACITY = 3,      /* a device capacity (contacts/vertices/edges) overflowed */
    FA_GRID_RANGE = 4,    /* an atom fell outside the cell grid (undefined in the reference) */
    FA_BAD_LIGAND = 5,    /* the pose names a ligand index the context does not hold */
    FA_UNDEFINED_FACE = 6 /* a face with no second point on its first edge's line: the
                             reference's order_faces reads past the face's point list
                             there (LIB/Vcontacts.cpp:794-805), its result is undefined */
};

/* One flexible side chain (a residue of FA->flex_res with its accepted
 * rotamers, LIB/build_rotamers.cpp:7-375).  Its atoms are receptor atoms
 * [first, first + n_atoms) in calc order.  Conformation 0 is the input one
 * (the receptor's xyz); conformations 1..n_rot-1 are the rotamers, whose
 * atoms the reference appends as copies that keep their PDB numbers, so
 * that num_atm maps them back to the input atoms for the bonded lookups
 * (LIB/vcfunction.cpp:189-240, LIB/Vcontacts.cpp:1813-1816).  The scorable
 * atoms are those outside the backbone (update_optres, LIB/update_optres.cpp:
 * 10-30); the residue is a protein residue (type 0). */
typedef struct fa_flex_residue {
    int32_t        first;     /* receptor index of the residue's first atom */
    int32_t        n_atoms;   /* atoms of the residue, backbone included */
    int32_t        n_rot;     /* conformations (>= 1) */
    const uint8_t* scorable;  /* [n_atoms] 1 = side-chain atom (scored) */
    const float*   rot_xyz;   /* [n_rot * n_atoms * 3]; row 0 equals the receptor's xyz */
    const uint8_t* bonded;    /* [n_atoms * n_atoms] 1 = within bloops bonds (residue.bonded >= 0) */
} fa_flex_residue;

/* Receptor atoms in calc order.  The rigid part is every atom that is not a
 * scorable atom of a flexible residue. */
typedef struct fa_receptor_desc {
    int32_t        n_atoms;
    const float*   xyz;      /* [n_atoms*3] Å */
    const float*   radius;   /* [n_atoms] van der Waals radius */
    const int32_t* type;     /* [n_atoms] 1-based energy-matrix type */
    int32_t        n_flex;   /* flexible side chains (0: rigid receptor) */
    const fa_flex_residue* flex; /* [n_flex], increasing `first`, not overlapping */
} fa_receptor_desc;

/* Distance constraint (constraint_str, LIB/flexaid.h:168-190) between ligand
 * atom `atom` and a receptor atom (partner_is_ligand = 0: receptor index)
 * or another ligand atom.  Covalent (type 1) constraints score in
 * vcfunction (LIB/vcfunction.cpp:146-166, :270-310): +KDIST for each
 * constrained scorable atom, -KDIST * GetValueFromGaussian(d, bond_len,
 * max_dist) when the pair is in contact, the clash distance shortened to
 * permeability * bond_len, and no contact energy for the pair.  Type 0
 * constraints do not score. */
typedef struct fa_constraint {
    int32_t atom;
    int32_t partner;
    int32_t partner_is_ligand;
    int32_t type;
    float   bond_len;
    float   max_dist;
} fa_constraint;

/* One flexible ligand with its gene map and cleft grid. */
typedef struct fa_ligand_desc {
    int32_t        n_atoms;
    int32_t        n_genes;
    int32_t        n_grid;     /* cleft-grid entries (index 0 is a valid slot) */
    const int32_t* rec;        /* [n_atoms*4] rebuild references, -1 = frame point / none */
    const float*   dis;        /* [n_atoms] base internal coordinates (Å, degrees) */
    const float*   ang;
    const float*   dih;
    const float*   shift;      /* [n_atoms] dihedral shift along the rec[3] chain */
    const float*   radius;     /* [n_atoms] */
    const int32_t* type;       /* [n_atoms] 1-based */
    const int32_t* mov;        /* [n_atoms] rebuild order (buildlist output) */
    const uint8_t* bonded;     /* [n_atoms*n_atoms] 1 = within bloops bonds (bonded[i][j] >= 0) */
    const int32_t* gene_typ;   /* [n_genes] fa_gene_type */
    const int32_t* gene_atm;   /* [n_genes] ligand-local atom (FA_GENE_ROT: flexible residue index) */
    const double*  gene_min;   /* [n_genes] clamp bounds (genlim.min/max) */
    const double*  gene_max;
    const float*   grid_ic;    /* [n_grid*3] (dis, ang, dih) of each cleft-grid point */
    float          ori[3];     /* FA->ori, frame origin for rec = -1 */
    float          globalmin[3];
    float          globalmax[3];
    float          maxwidth;
    int32_t        n_cons;     /* distance constraints (optional) */
    const fa_constraint* cons; /* [n_cons], in the order the scorer looks them up */
} fa_ligand_desc;

/* Scoring parameters (FA_Global fields used by vcfunction/Vcontacts). */
typedef struct fa_params {
    float        permeability;   /* FA->permeability */
    float        dee_clash;      /* FA->dee_clash */
    int32_t      intramolecular; /* FA->intramolecular */
    float        solventterm;    /* FA->solventterm (0 = use solvent column) */
    int32_t      ntypes;         /* FA->ntypes; type ntypes is the solvent column */
    const float* emat;           /* [ntypes*ntypes] y value per type pair, weight = 1 */
    int32_t      normalize_area; /* FA->normalize_area: weight-1 terms use area / (4 pi rado^2) */
    /* Optional density-function pairs (energy_matrix.weight = 0, read_emat,
     * LIB/read_emat.cpp:103-138; get_yval, LIB/vcfunction.cpp:576-621):
     * pair k = (t1-1)*ntypes + (t2-1) has nodes [dens_off[k], dens_off[k+1])
     * of dens_xy, (x, y) in file order; an empty range makes pair k a
     * weight-1 pair with value emat[k].  Null dens_off: every pair weight 1. */
    const int32_t* dens_off;     /* [ntypes*ntypes + 1], dens_off[0] = 0 */
    const float*   dens_xy;      /* [2 * dens_off[ntypes*ntypes]] */
} fa_params;

/* Per-pose result: the cfstr fields the path produces (LIB/flexaid.h:122-135)
 * plus get_cf_evalue / get_apparent_cf_evalue (LIB/ic2cf.cpp:391-402). */
typedef struct fa_cf_result {
    double  com;
    double  wal;
    double  sas;
    double  con;
    double  totsas;
    double  cf;        /* com + wal + sas + con + metal_coord */
    double  app;       /* com + wal + sas + metal_coord */
    int32_t rclash;
    int32_t status;    /* fa_status */
} fa_cf_result;

/* One directed contact record (ca_struct, LIB/Vcontacts.h:127-134), in calc
 * order indices: `from` is the scorable atom whose polyhedron produced it. */
typedef struct fa_contact_rec {
    int32_t from;
    int32_t atom;
    double  area;
    double  dist;
} fa_contact_rec;

/* Operation counts of the contact-function algorithm, summed over poses.
 * Each counter is one unit of the algorithmic FLOP model (weights in
 * parentheses, FLOP per unit): cull = neighbour candidates distance-tested
 * (11), plane = contact planes built (14), hull_test = plane intersections
 * evaluated in the hull walk (13), arc_solve = edge/sphere solves (82),
 * face_order = vertex cosines for face ordering (18), sph_triangle =
 * spherical triangles (85), sph_arc = spherical arc segments (60),
 * vc_contact = contact records scored (25), build_atom = atoms rebuilt from
 * internal coordinates (94). */
typedef struct fa_opcount {
    uint64_t cull;
    uint64_t plane;
    uint64_t hull_test;
    uint64_t arc_solve;
    uint64_t face_order;
    uint64_t sph_triangle;
    uint64_t sph_arc;
    uint64_t vc_contact;
    uint64_t build_atom;
} fa_opcount;

#define FA_OPCOUNT_N 9

/* Canonical-ensemble thermodynamics (statmech::Thermodynamics,
 * LIB/statmech.h:36-45). */
typedef struct fa_thermo {
    double temperature;
    double log_Z;
    double free_energy;
    double mean_energy;
    double mean_energy_sq;
    double heat_capacity;
    double entropy;
    double std_energy;
} fa_thermo;

#ifdef __cplusplus
}
#endif

#endif /* FA_TYPES_H */
