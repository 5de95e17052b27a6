/* cell_slots.h — the 27-slot neighbourhood of Vcontacts' cell grid.
 *
 * index_protein puts atom (x, y, z) in cell
 *     bx * dim^2 + by * dim + bz,   b? = (int)((? - globalmin?) / CELLSIZE),
 * and get_contlist4 scans, for an atom in cell c, the cells
 *     c + (s / 9 - 1) dim^2 + ((s / 3) % 3 - 1) dim + s % 3 - 1,   s = 0..26,
 * that fall inside [0, dim^3) (LIB/Vcontacts.cpp:1738-1741).  The index
 * arithmetic has no per-axis bounds: a cell at z = dim - 1 names, through
 * dz = +1, the cell at z = 0 of the next row.  With dim >= 3 the 27 offsets
 * are distinct, so a cell is named by at most one slot.
 *
 * fa_slot_of_cells finds that slot from the two cells' coordinates without
 * a search: the offset plus (dim^2 + dim + 1), written in the grid's mixed
 * radix (digits z, then y, then x, each carry at most one since the
 * coordinates differ by less than dim), has digits 0..2 exactly when it is a
 * slot offset, and the slot is that digit string.  It equals the search over
 * the slots for every pair of cells (tests/cell_slots_check.cpp).
 */
#pragma once

#if defined(__HIPCC__)
#define FA_CELL_FN __host__ __device__ inline
#else
#define FA_CELL_FN static inline
#endif

/* cell named by slot s around cell c (not bounds-checked) */
FA_CELL_FN int fa_slot_box(int dim, int c, int s)
{
    return c + dim * dim * ((s / 9) - 1) + dim * (((s / 3) % 3) - 1) + (s % 3) - 1;
}

/* slot (0..26) whose offset is d, -1 if none (dim >= 3) */
FA_CELL_FN int fa_slot_of(int dim, int d)
{
    int slot = -1;
    for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy) {
            const int r = d - dim * dim * dx - dim * dy;
            if (r >= -1 && r <= 1) slot = (dx + 1) * 9 + (dy + 1) * 3 + (r + 1);
        }
    return slot;
}

/* how many of the 27 slots have offset d (coarse grids, dim < 3, repeat cells) */
FA_CELL_FN int fa_slot_multiplicity(int dim, int d)
{
    int mult = 0;
    for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy) {
            const int r = d - dim * dim * dx - dim * dy;
            mult += (r >= -1 && r <= 1);
        }
    return mult;
}

/* cell coordinates packed x | y << 10 | z << 20 (dim < 1024) */
FA_CELL_FN int fa_cell_pack(int x, int y, int z) { return x | (y << 10) | (z << 20); }

/* fa_slot_of for the offset between packed cells rc and kc (dim >= 3) */
FA_CELL_FN int fa_slot_of_cells(int dim, int kc, int rc)
{
    int ex = (rc & 1023) - (kc & 1023) + 1;
    int ey = ((rc >> 10) & 1023) - ((kc >> 10) & 1023) + 1;
    int ez = ((rc >> 20) & 1023) - ((kc >> 20) & 1023) + 1;
    if (ez < 0) {
        ez += dim;
        --ey;
    } else if (ez >= dim) {
        ez -= dim;
        ++ey;
    }
    if (ey < 0) {
        ey += dim;
        --ex;
    } else if (ey >= dim) {
        ey -= dim;
        ++ex;
    }
    return ((unsigned)ex <= 2u && (unsigned)ey <= 2u && (unsigned)ez <= 2u) ? ex * 9 + ey * 3 + ez : -1;
}
