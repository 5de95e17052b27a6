// fa_device.h — device-side system tables shared by the kernels and the host
// context code.  Everything here is plain data resident in HBM.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fa_types.h"

namespace fa {

// Receptor atoms in calc order.  The scorable atoms of a pose ("movable"
// atoms, index m) are the side-chain atoms of the flexible residues, m < n_sc
// in calc order, then the ligand's atoms, m = n_sc + ligand index; they are
// placed per pose and never stored in the receptor cell lists.
struct DevReceptor {
    int n;
    const float4* xyzr;             // [n] x, y, z, radius (conformation 0 for side chains)
    const int* type;                // [n]
    int n_sc;                       // scorable side-chain atoms (0: rigid receptor)
    int n_flex;                     // flexible residues
    const int* sc_calc;             // [n_sc] calc index
    const int* sc_res;              // [n_sc] flexible residue
    const int* sc_rot_off;          // [n_sc] first of the atom's conformations in sc_rot
    const float4* sc_rot;           // conformations of each side-chain atom (x, y, z, radius)
    const int* flex_nrot;           // [n_flex] conformations per residue
    const int2* sc_range;           // [n_sc] calc range [lo, hi) of the atom's residue
    const unsigned long long* sc_bonded; // [2 * n_sc] bonded bits over calc index - lo
    const unsigned char* sc_flag;   // [n] 1 = scorable side-chain atom, or null (rigid receptor)
};

// Cell list of the receptor for one grid (the grid index_protein builds when
// no scorable atom widens it).  cell_xyzr / cell_atom are the receptor atoms
// sorted by cell, calc order inside a cell.
struct DevGrid {
    float gmin[3];
    float gmax[3];
    float maxwidth;
    int dim;
    const int* cell_start;   // [dim^3 + 1]
    const int* cell_atom;    // [n_receptor] calc index
    const float4* cell_xyzr; // [n_receptor] x, y, z, radius
    // Neighbourhood lists (grids with dim >= 3, else null): for every cell c
    // the receptor atoms of its 27 neighbourhood slots in scan order (slot,
    // then cell-list order), entries [nb_start[c], nb_start[c + 1]);
    // nb_atom = calc index | slot << 27; nb_bnd[28 c + s] = where slot s
    // starts, relative to nb_start[c] (nb_bnd[28 c + 27] = the list length).
    const int* nb_start;     // [dim^3 + 1]
    const int* nb_atom;
    const float4* nb_xyzr;
    const int* nb_bnd;       // [28 dim^3]
};

// One ligand topology (pointers into the context's device pool).
struct DevLigand {
    int n_atoms;
    int n_genes;
    int n_grid;
    int grid_id;
    float ori[3];
    float globalmin[3];
    float globalmax[3];
    float maxwidth;
    const int* rec;                 // [4n]
    const float* dis;               // [n]
    const float* ang;
    const float* dih;
    const float* shift;
    const float* radius;
    const int* type;
    const int* mov;                 // [n]
    const unsigned long long* bonded; // [n * 2] bitset rows
    const int* gene_typ;            // [g]
    const int* gene_atm;
    const double* gene_min;
    const double* gene_max;
    const float* grid_ic;           // [3 * n_grid]
    // distance constraints of each atom: [cons_off[a], cons_off[a + 1]) of
    // the arrays below, partner as a calc index (receptor i, ligand Rn + k)
    const int* cons_off;            // [n + 1], null without constraints
    const int* cons_part;
    const int* cons_type;
    const float* cons_bond;
    const float* cons_max;
    double con0;                    // KDIST per covalent constraint of each atom
};

struct DevParams {
    float permeability;
    float dee_clash;
    float solventterm;
    int intramolecular;
    int ntypes;
    const float* emat;              // [ntypes * ntypes]
    int normalize_area;
    const int* dens_off;            // [ntypes * ntypes + 1] or null (all pairs weight 1)
    const float2* dens_xy;          // density nodes (x, y)
};

// One launch over a batch of poses.
struct DevBatch {
    int n_poses;
    int gene_stride;
    const double* genes;            // [n_poses * gene_stride]
    const int* pose_lig;            // [n_poses] or null (ligand 0)
    fa_cf_result* out;              // [n_poses]
    fa_contact_rec* recs;           // optional: [n_poses * rec_cap]
    int rec_cap;
    int* rec_count;                 // optional: [n_poses]
    float* lig_xyz;                 // optional: [n_poses * xyz_stride * 3]
    int xyz_stride;
    int n_ligands;                  // ligands of the context (pose_lig entries outside [0, n) are rejected)
    unsigned int* counters;         // [4]: slow-grid poses, capacity overflows, work queue, spare
    unsigned long long* totals;     // [FA_TOTALS_N] never cleared by a launch: slow-grid, capacity,
                                    // grid-range, bad-ligand, undefined-face poses since the
                                    // context's last fa_ctx_reset_totals
    unsigned long long* ops;        // optional: [FA_OPCOUNT_N] operation counters (fa_opcount order)
    unsigned long long* prof;       // optional: [16] cycles per kernel phase (summed over waves)
    // per-wave neighbour lists (HBM scratch): wave w owns entries
    // [w * near_cap, (w + 1) * near_cap) of both arrays
    int2* near_idx;                 // calc index, squared distance (float bits)
    float4* near_xyz;               // x, y, z, radius of the neighbour
    int near_cap;
    // per-wave slow-path scratch: with R4 = receptor atoms rounded up to a
    // multiple of 4, wave w owns [2 * R4 * w, 2 * R4 * (w + 1)): the cell of
    // every receptor atom in a pose's widened grid, then the atoms gathered
    // for one ligand atom
    int* rbox;
    int cell_lists;                 // 1: scan the cell lists even where neighbourhood lists exist
};

// Per-pose rows (coordinates, internal coordinates) of many poses, tiled for
// streaming: tiles of COORD_TILE poses, each tile holding its poses' row k
// contiguously (row k of pose p at coord_at(rows, k, p)), so a block of
// COORD_TILE threads reads and writes one contiguous region.
constexpr int COORD_TILE = 256;
__host__ __device__ __forceinline__ size_t coord_at(int rows, int k, int pose)
{
    return ((size_t)(pose / COORD_TILE) * rows + k) * COORD_TILE + (size_t)(pose % COORD_TILE);
}
__host__ __device__ __forceinline__ size_t coord_floats(int rows, int n)
{
    return (size_t)((n + COORD_TILE - 1) / COORD_TILE) * rows * COORD_TILE;
}

}  // namespace fa
