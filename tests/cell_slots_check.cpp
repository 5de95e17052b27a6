// Host check of cell_slots.h: for every pair of cells of grids with 3..14
// cells a side (and random pairs up to 60 a side), the slot found by
// fa_slot_of (search over the offsets) and by fa_slot_of_cells (mixed radix
// from coordinates) equals the slot of the reference's enumeration
// (the first s with fa_slot_box(c_k, s) == c_r, LIB/Vcontacts.cpp:1738-1741);
// on coarse grids (1 and 2 cells a side) fa_slot_multiplicity equals the
// number of such s.  Prints the number of cases and of mismatches.
#include <cstdio>
#include <random>

#include "cell_slots.h"

static int truth(int dim, int ck, int cr, int* mult)
{
    int first = -1, n = 0;
    for (int s = 0; s < 27; ++s)
        if (fa_slot_box(dim, ck, s) == cr) {
            if (first < 0) first = s;
            ++n;
        }
    if (mult) *mult = n;
    return first;
}

int main()
{
    long long cases = 0, bad = 0;
    auto check = [&](int dim, int kx, int ky, int kz, int rx, int ry, int rz) {
        const int ck = kx * dim * dim + ky * dim + kz, cr = rx * dim * dim + ry * dim + rz;
        const int t = truth(dim, ck, cr, nullptr);
        const int a = fa_slot_of(dim, cr - ck);
        const int b = fa_slot_of_cells(dim, fa_cell_pack(kx, ky, kz), fa_cell_pack(rx, ry, rz));
        ++cases;
        if (a != t || b != t) {
            if (bad < 10) std::printf("dim %d k (%d %d %d) r (%d %d %d): truth %d search %d radix %d\n", dim, kx, ky, kz, rx,
                                      ry, rz, t, a, b);
            ++bad;
        }
    };
    for (int dim = 3; dim <= 14; ++dim)
        for (int kx = 0; kx < dim; ++kx)
            for (int ky = 0; ky < dim; ++ky)
                for (int kz = 0; kz < dim; ++kz)
                    for (int rx = 0; rx < dim; ++rx)
                        for (int ry = 0; ry < dim; ++ry)
                            for (int rz = 0; rz < dim; ++rz) check(dim, kx, ky, kz, rx, ry, rz);
    std::mt19937 rng(7);
    for (int dim = 15; dim <= 60; ++dim) {
        std::uniform_int_distribution<int> u(0, dim - 1), n(-2, 2);
        for (int q = 0; q < 200000; ++q) {
            const int kx = u(rng), ky = u(rng), kz = u(rng);
            // near pairs (where the slots are) and any pairs
            int rx = kx + n(rng), ry = ky + n(rng), rz = (q & 1) ? u(rng) : kz + n(rng);
            if (q % 3 == 0) rz = (kz == dim - 1) ? 0 : (kz == 0 ? dim - 1 : rz);  // row wraps
            if (rx < 0 || rx >= dim || ry < 0 || ry >= dim || rz < 0 || rz >= dim) continue;
            check(dim, kx, ky, kz, rx, ry, rz);
        }
    }
    for (int dim = 1; dim <= 2; ++dim) {
        const int d3 = dim * dim * dim;
        for (int ck = 0; ck < d3; ++ck)
            for (int cr = 0; cr < d3; ++cr) {
                int m = 0;
                truth(dim, ck, cr, &m);
                ++cases;
                if (fa_slot_multiplicity(dim, cr - ck) != m) ++bad;
            }
    }
    std::printf("%lld cases, %lld mismatches\n", cases, bad);
    return bad ? 1 : 0;
}
