/*
 * cf_oracle.cpp — CPU oracle for the FlexAID∆S population-evaluation path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates the reference CPU algorithm
 * (LIB/ic2cf.cpp, LIB/buildcc.cpp, LIB/Vcontacts.cpp, LIB/vcfunction.cpp) so
 * that the tests, the smoke check and the CPU-baseline leg of the benchmark
 * can check the HIP library against it.  Nothing in the shipped library
 * links, loads or calls this code.
 *
 * Semantics restated (each function cites the reference lines it follows):
 *   - gene decode + dihedral shift cascade      LIB/ic2cf.cpp:53-105
 *   - Cartesian rebuild (float arithmetic)       LIB/buildcc.cpp:6-134
 *   - cell grid + box order (steady state, i.e.
 *     after the first call, with only scorable
 *     atoms able to widen the grid)              LIB/Vcontacts.cpp:1570-1715
 *   - clash pre-pass (half list via `done`)      LIB/Vcontacts.cpp:36-57, 1725-1839
 *   - per-atom contact list with the `done`/stub
 *     rule for already-processed atoms           LIB/Vcontacts.cpp:1750-1836
 *   - hull walk, arc points, seeds               LIB/Vcontacts.cpp:147-640
 *   - face ordering and areas (+engulfing)       LIB/Vcontacts.cpp:699-1044, 1278-1557
 *   - contact records (save_areas)               LIB/Vcontacts.cpp:1149-1220
 *   - energy sum                                 LIB/vcfunction.cpp:14-574
 *
 * Flexible side chains (fa_flex_residue): the rotamer genes pick each
 * residue's conformation (LIB/ic2cf.cpp:90-103), the residue's non-backbone
 * atoms are scorable atoms among the receptor atoms in calc order
 * (LIB/Vcontacts.cpp:1598-1625, LIB/update_optres.cpp:10-30), bonded and
 * intramolecular tests follow the residue (LIB/vcfunction.cpp:189-240,
 * LIB/Vcontacts.cpp:1806-1821).
 *
 * Evaluation modes (bits of fa_oracle_set_mode_bits):
 *   0 (the default) — the reference's serial semantics above;
 *   4 — each cell's candidates taken rigid atoms first, then scorable atoms
 *     (both in calc order): the order the GPU scans a cell in; with side
 *     chains among the receptor atoms it changes only summation order;
 *   1, 2 — no done/stub exclusion, no hull seeds: every scorable atom's
 *     polyhedron on its own (independent atoms);
 *   8 — the speculative form: independent atoms from starts found by ray
 *     casting, and every pose whose reference start the checks of
 *     spec_check cannot vouch for evaluated in the serial semantics.
 * Bits 1, 2 and 8 exist to measure why the device keeps the serial order:
 * the reference's areas depend on the hull walk's order, so neither form
 * stays within the CF tolerance (tests/test_oracle_atom_order.py).
 *
 * The geometry type `R` is double for the oracle proper (the reference does
 * all Vcontacts geometry in double); the float instantiation exists only to
 * measure how far FP32 geometry drifts from it.
 *
 * Parity status: every function above is pinned against the reference
 * sources compiled by oracle/ref/Makefile (tests/test_oracle_vs_ref.py) and
 * the helper known-answer tests of tests/test_vcontacts.cpp in the reference
 * (tests/test_oracle_kat.py).
 */
#include "fa_types.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr double PI_REF = 3.141592;        // LIB/flexaid.h:86
constexpr float RW = 1.4f;                 // LIB/flexaid.h:88
constexpr float KWALL_F = 1.0e6f;          // LIB/flexaid.h:81
constexpr float POLY_PENALTY = 1e6f;       // LIB/flexaid.h:69
constexpr double CLASH_THRESHOLD = 1e4;    // LIB/flexaid.h:70
constexpr float CELLSIZE = 6.5f;           // LIB/Vcontacts.h:56
constexpr int EDGE_LIMIT = 200;            // LIB/Vcontacts.cpp:378

// Trig of the pose build: 0 = the host libm's cosf/sinf (what the reference
// binary calls), 1 = the double-precision value rounded to float (correctly
// rounded; what the GPU kernel computes).
int g_trig_mode = 0;

// 0 = reference semantics, 1 = independent atoms (see the header); the
// three parts of the independent form can be chosen alone (bits of
// g_mode_bits: 1 no done/stub rule, 2 no hull seeds, 4 rigid atoms first in
// a cell) to measure what each changes
int g_eval_mode = 0;
int g_mode_bits = 0;
// a pose re-evaluated in the serial form after the speculative form flagged
// it runs with the speculative bit cleared (this thread only)
thread_local int t_bits_override = -1;
inline int mode_bits() { return t_bits_override >= 0 ? t_bits_override : g_mode_bits; }
// bit 8: the speculative form, every atom's polyhedron from its full
// neighbour list and a start found by ray casting, poses that the checks of
// spec_check flag re-evaluated in the serial form (implies bits 1 and 2)
inline bool m_spec() { return mode_bits() & 8; }
inline bool m_nodone() { return (mode_bits() & 1) || m_spec(); }
inline bool m_noseed() { return (mode_bits() & 2) || m_spec(); }
inline bool m_rigidfirst() { return mode_bits() & 4; }

// speculative form: an atom whose hull has this many edges sends its pose to
// the serial form (the reference's reduced contact list can change the count
// near its 200-edge failure limit)
constexpr int SPEC_EDGE_MAX = 120;
// a start vertex more than this outside a plane is not a polyhedron vertex
constexpr double SPEC_VALID_EPS = 1e-10;
// per pose: set when the speculative form flagged the pose (poses evaluated
// in the serial form); per batch: the flags of the last eval_batch call
thread_local int t_fallback = 0;

inline float trig_cos(float x) { return g_trig_mode ? (float)std::cos((double)x) : std::cos(x); }
inline float trig_sin(float x) { return g_trig_mode ? (float)std::sin((double)x) : std::sin(x); }

// per-thread operation counters (see fa_opcount)
thread_local fa_opcount* t_ops = nullptr;
#define OPS(field, n)                        \
    do {                                     \
        if (t_ops) t_ops->field += (n);      \
    } while (0)

// capacity statistics: largest sizes seen (sizing of the GPU kernel's
// fixed per-wave arrays); null unless requested
struct CapStats {
    int near_pose;   // in-range candidates of all atoms of a pose, done or not
    int nc;          // contacts of one atom (main pass)
    int edges;       // hull edges of one atom
    int vn;          // hull vertices + arc points before the sphere cut
    int nv_opp;      // vertices after the opposite points are appended
    int face_slots;  // sum over faces of (points + 1 if more than 2 points)
    int face_slots_run;
};
thread_local CapStats* t_caps = nullptr;

// branches of calc_areas reached: atoms with one, two, three or more
// engulfing planes (LIB/Vcontacts.cpp:963-1042), atoms with a polyhedron
struct BranchStats {
    long long epi1, epi2, epi3, atoms, hull_fail;
};
thread_local BranchStats* t_branch = nullptr;
#define CAPMAX(field, v)                                      \
    do {                                                      \
        if (t_caps && (v) > t_caps->field) t_caps->field = (v); \
    } while (0)

template <class R> struct Plane {
    R Ai[4];
    R dist;
    int index;
    R area;
    char flag;
};
template <class R> struct Vert {
    R xi[3];
    R dist;
    int plane[3];
};
template <class R> struct Edge {
    R V[3];
    int startpt, endpt;
    int plane[2];
    int startplane, endplane;
    char arc;
};
template <class R> struct Cont {
    int index;
    R dist;
};

// ------------------------------------------------------------------ helpers
// LIB/Vcontacts.cpp:1384-1413
template <class R> int solve_3x3(const R* eq0, const R* eq1, const R* eq2, R* pt)
{
    R c00 = eq1[1] * eq2[2] - eq2[1] * eq1[2];
    R c01 = -eq1[0] * eq2[2] + eq2[0] * eq1[2];
    R c02 = eq1[0] * eq2[1] - eq2[0] * eq1[1];
    R c10 = -eq0[1] * eq2[2] + eq2[1] * eq0[2];
    R c11 = eq0[0] * eq2[2] - eq2[0] * eq0[2];
    R c12 = -eq0[0] * eq2[1] + eq2[0] * eq0[1];
    R c20 = eq0[1] * eq1[2] - eq1[1] * eq0[2];
    R c21 = -eq0[0] * eq1[2] + eq1[0] * eq0[2];
    R c22 = eq0[0] * eq1[1] - eq1[0] * eq0[1];
    R det = eq0[0] * c00 + eq0[1] * c01 + eq0[2] * c02;
    if (det == R(0)) return -1;
    pt[0] = -(eq0[3] * c00 + eq1[3] * c10 + eq2[3] * c20) / det;
    pt[1] = -(eq0[3] * c01 + eq1[3] * c11 + eq2[3] * c21) / det;
    pt[2] = -(eq0[3] * c02 + eq1[3] * c12 + eq2[3] * c22) / det;
    return 0;
}

// LIB/Vcontacts.cpp:1424-1476
template <class R> int solve_2xS(const R* a0, const R* a1, float rado, R* pt0, R* pt1)
{
    R e2[3];
    e2[0] = a0[1] * a1[2] - a0[2] * a1[1];
    e2[1] = a0[2] * a1[0] - a0[0] * a1[2];
    e2[2] = a0[0] * a1[1] - a0[1] * a1[0];
    R c00 = a1[1] * e2[2] - e2[1] * a1[2];
    R c01 = -a1[0] * e2[2] + e2[0] * a1[2];
    R c02 = a1[0] * e2[1] - e2[0] * a1[1];
    R c10 = -a0[1] * e2[2] + e2[1] * a0[2];
    R c11 = a0[0] * e2[2] - e2[0] * a0[2];
    R c12 = -a0[0] * e2[1] + e2[0] * a0[1];
    R det = e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2];
    if (det == R(0)) return -1;
    R avg[3];
    avg[0] = -(a0[3] * c00 + a1[3] * c10) / det;
    avg[1] = -(a0[3] * c01 + a1[3] * c11) / det;
    avg[2] = -(a0[3] * c02 + a1[3] * c12) / det;
    // rado*rado is a float product in the reference (float operands)
    R t = ((R)(rado * rado) - avg[0] * avg[0] - avg[1] * avg[1] - avg[2] * avg[2]) / det;
    if (t < R(0)) return -1;
    t = std::sqrt(t);
    for (int k = 0; k < 3; ++k) {
        pt0[k] = avg[k] + t * e2[k];
        pt1[k] = avg[k] - t * e2[k];
    }
    return 0;
}

// LIB/Vcontacts.cpp:1486-1505
template <class R> R cosPQR(const R* P, const R* Q, const R* Rr)
{
    R qp[3] = {P[0] - Q[0], P[1] - Q[1], P[2] - Q[2]};
    R qr[3] = {Rr[0] - Q[0], Rr[1] - Q[1], Rr[2] - Q[2]};
    return (qp[0] * qr[0] + qp[1] * qr[1] + qp[2] * qr[2]) /
           std::sqrt((qp[0] * qp[0] + qp[1] * qp[1] + qp[2] * qp[2]) *
                     (qr[0] * qr[0] + qr[1] * qr[1] + qr[2] * qr[2]));
}

// LIB/Vcontacts.cpp:1518-1557: cap sector at centre A between B and C minus
// the spherical triangle (A', B, C)
template <class R> R spherical_arc(const Vert<R>& A, const Vert<R>& B, const Vert<R>& C, float rado)
{
    OPS(sph_arc, 1);
    R cAOB = (A.xi[0] * B.xi[0] + A.xi[1] * B.xi[1] + A.xi[2] * B.xi[2]) / (A.dist * B.dist);
    R cBOC = (B.xi[0] * C.xi[0] + B.xi[1] * C.xi[1] + B.xi[2] * C.xi[2]) / (B.dist * C.dist);
    R U = (1 + cAOB) * std::sqrt((1 + cBOC) / R(8));
    R V = (1 - cAOB) * std::sqrt((1 + cBOC) / R(8));
    R W = std::sqrt((1 - cAOB) * (1 + cAOB) * (1 - cBOC) / R(8));
    R tS = (1 - U + V + W + W) / (1 + U - V - W - W);
    R tSB = (1 - U - V) / (1 + U + V);
    R tSA = (1 - U + V - W - W) / (1 + U - V + W + W);
    R tanprod;
    if ((tS * tSA) > R(0))
        tanprod = std::sqrt(tSB * tSB * tS * tSA);
    else
        tanprod = R(0);
    R cBAC = cosPQR(B.xi, A.xi, C.xi);
    R angBAC;
    if (cBAC > R(1))
        angBAC = R(0);
    else if (cBAC < R(-1))
        angBAC = (R)PI_REF;
    else
        angBAC = std::acos(cBAC);
    R ro = (R)rado;
    return ro * (angBAC * (ro - A.dist) - R(4) * ro * std::atan(std::sqrt(tanprod)));
}

// LIB/Vcontacts.cpp:1357-1372
template <class R>
char test_point(const R* p, const Plane<R>* cont, int NC, int pA, int pB, int pC)
{
    for (int cp = 0; cp < NC; ++cp) {
        if (cp != pA && cp != pB && cp != pC && cont[cp].flag != 'X') {
            if (cont[cp].Ai[0] * p[0] + cont[cp].Ai[1] * p[1] + cont[cp].Ai[2] * p[2] + cont[cp].Ai[3] > R(0))
                return 'N';
        }
    }
    return 'Y';
}

// LIB/Vcontacts.cpp:646-659
template <class R> void add_vertex(Vert<R>* poly, int vn, const R* c, int a, int b, int cc)
{
    poly[vn].xi[0] = c[0];
    poly[vn].xi[1] = c[1];
    poly[vn].xi[2] = c[2];
    poly[vn].plane[0] = a;
    poly[vn].plane[1] = b;
    poly[vn].plane[2] = cc;
    poly[vn].dist = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
}

// LIB/Vcontacts.cpp:599-637: edge direction along the body of the polyhedron
template <class R>
void add_vedge(Edge<R>* e, int k, const Plane<R>* cont, int p0, int p1, int tp, const Vert<R>* poly, int sp)
{
    e[k].V[0] = cont[p0].Ai[1] * cont[p1].Ai[2] - cont[p0].Ai[2] * cont[p1].Ai[1];
    e[k].V[1] = cont[p0].Ai[2] * cont[p1].Ai[0] - cont[p0].Ai[0] * cont[p1].Ai[2];
    e[k].V[2] = cont[p0].Ai[0] * cont[p1].Ai[1] - cont[p0].Ai[1] * cont[p1].Ai[0];
    e[k].startpt = sp;
    e[k].endpt = -1;
    e[k].plane[0] = p0;
    e[k].plane[1] = p1;
    e[k].startplane = tp;
    e[k].endplane = -1;
    e[k].arc = '.';
    R t[3] = {poly[sp].xi[0] + e[k].V[0], poly[sp].xi[1] + e[k].V[1], poly[sp].xi[2] + e[k].V[2]};
    R tv = cont[tp].Ai[0] * t[0] + cont[tp].Ai[1] * t[1] + cont[tp].Ai[2] * t[2] + cont[tp].Ai[3];
    if (tv > R(0)) {
        e[k].V[0] = -e[k].V[0];
        e[k].V[1] = -e[k].V[1];
        e[k].V[2] = -e[k].V[2];
    }
}

// ---------------------------------------------------------------- workspace
template <class R> struct Work {
    // per-atom (calc order)
    std::vector<float> x, y, z, rad;
    std::vector<int> type;
    std::vector<int> boxnum, calclist;
    std::vector<char> done;
    std::vector<int> seed;                 // 3 per atom
    std::vector<std::vector<int>> stubs;   // per scorable slot: earlier atoms with a record to it
    std::vector<std::vector<fa_contact_rec>> records; // per scorable slot: own valid contacts (list order)
    std::vector<int> rank;                 // per calc index: processing rank of a scorable atom
    std::vector<int> boxfirst, boxcount;
    int near_count = 0;
    // per atomzero scratch
    std::vector<Cont<R>> contlist;
    std::vector<Plane<R>> cont;
    std::vector<Vert<R>> poly, centerpt;
    std::vector<Edge<R>> vedge;
    std::vector<std::vector<int>> ptorder;
    int dim = 0;
    int last_edges = 0;   // hull edges of the last atom
};

struct System {
    const fa_receptor_desc* rec;
    const fa_ligand_desc* lig;
    const fa_params* prm;
    // per calc index: scorable slot (-1: not scorable) and residue (flexible
    // residue k >= 0, the ligand -1, any other receptor atom -2)
    std::vector<int> slot, res;
    std::vector<int> scorable;  // calc indices of the scorable atoms, calc order
    int R() const { return rec->n_atoms; }
    int L() const { return lig->n_atoms; }
    int N() const { return rec->n_atoms + lig->n_atoms; }

    System(const fa_receptor_desc* r, const fa_ligand_desc* l, const fa_params* p) : rec(r), lig(l), prm(p)
    {
        const int Rn = R(), n = N();
        slot.assign(n, -1);
        res.assign(n, -2);
        for (int k = 0; k < rec->n_flex; ++k) {
            const fa_flex_residue& f = rec->flex[k];
            for (int i = 0; i < f.n_atoms; ++i) res[f.first + i] = k;
        }
        for (int i = 0; i < n; ++i) {
            if (i >= Rn) res[i] = -1;
            const bool sc = i >= Rn || (res[i] >= 0 && rec->flex[res[i]].scorable[i - rec->flex[res[i]].first]);
            if (sc) {
                slot[i] = (int)scorable.size();
                scorable.push_back(i);
            }
        }
    }
    // both atoms in one residue (the ligand or one flexible side chain)
    bool same_res(int a, int b) const { return res[a] != -2 && res[a] == res[b]; }
    // bonded within bloops (residue.bonded >= 0); a and b in one residue
    bool bonded(int a, int b) const
    {
        const int Rn = R();
        if (res[a] == -1) return lig->bonded[(a - Rn) * L() + (b - Rn)] != 0;
        const fa_flex_residue& f = rec->flex[res[a]];
        return f.bonded[(a - f.first) * f.n_atoms + (b - f.first)] != 0;
    }
};

// ---------------------------------------------------------------- pose build
// LIB/ic2cf.cpp:53-105 (gene decode) then LIB/buildcc.cpp:6-134 (float rebuild).
// `rot` (optional, one entry per flexible residue of `rec`) receives the
// conformation each rotamer gene selects; residues without a gene keep 0.
void decode_and_build(const fa_ligand_desc* lig, const double* icv_in, float* xyz,
                      const fa_receptor_desc* rec = nullptr, int* rot = nullptr)
{
    if (rot && rec)
        for (int k = 0; k < rec->n_flex; ++k) rot[k] = 0;
    const int n = lig->n_atoms;
    std::vector<float> dis(lig->dis, lig->dis + n), ang(lig->ang, lig->ang + n), dih(lig->dih, lig->dih + n);
    for (int i = 0; i < lig->n_genes; ++i) {
        double v = icv_in[i];
        // clamp to the gene limits (LIB/gaboom.cpp:1851-1860, LIB/VoronoiCFBatch.h:58-64)
        if (v > lig->gene_max[i]) v = lig->gene_max[i];
        if (v < lig->gene_min[i]) v = lig->gene_min[i];
        const int atm = lig->gene_atm[i];
        switch (lig->gene_typ[i]) {
        case FA_GENE_GRID: {
            unsigned g = (unsigned)v;
            if ((int)g >= lig->n_grid) g = lig->n_grid - 1;
            dis[atm] = lig->grid_ic[3 * g + 0];
            ang[atm] = lig->grid_ic[3 * g + 1];
            dih[atm] = lig->grid_ic[3 * g + 2];
            break;
        }
        case FA_GENE_DIS: dis[atm] = (float)v; break;
        case FA_GENE_ANG: ang[atm] = (float)v; break;
        case FA_GENE_DIH: {
            dih[atm] = (float)v;
            int j = atm, cat = lig->rec[4 * j + 3];
            if (cat >= 0) {
                for (int guard = 0; cat != atm && guard < n; ++guard) {
                    dih[cat] = dih[j] + lig->shift[cat];
                    j = cat;
                    cat = lig->rec[4 * j + 3];
                }
            }
            break;
        }
        case FA_GENE_ROT: {  // rot_idx = (uint)(icv[i] + 0.5f), LIB/ic2cf.cpp:90-97
            if (rot && rec && atm >= 0 && atm < rec->n_flex) {
                unsigned r = (unsigned)(v + 0.5f);
                if ((int)r >= rec->flex[atm].n_rot) r = (unsigned)(rec->flex[atm].n_rot - 1);
                rot[atm] = (int)r;
            }
            break;
        }
        default: break;
        }
    }
    OPS(build_atom, n);
    for (int an = 0; an < n; ++an) {
        const int a = lig->mov[an];
        float x[4], y[4], z[4];
        for (int i = 1; i <= 3; ++i) {
            const int j = lig->rec[4 * a + i - 1];
            if (j >= 0) {
                x[i] = xyz[3 * j + 0];
                y[i] = xyz[3 * j + 1];
                z[i] = xyz[3 * j + 2];
            } else if (i == 1) {
                x[i] = 1.0f + lig->ori[0];
                y[i] = 0.0f + lig->ori[1];
                z[i] = 0.0f + lig->ori[2];
            } else if (i == 3) {
                x[i] = 0.0f + lig->ori[0];
                y[i] = 1.0f + lig->ori[1];
                z[i] = 0.0f + lig->ori[2];
            } else {
                x[i] = 0.0f + lig->ori[0];
                y[i] = 0.0f + lig->ori[1];
                z[i] = 0.0f + lig->ori[2];
            }
            x[i] += 1e-10f;
            y[i] += 1e-10f;
            z[i] += 1e-10f;
        }
        float A = y[1] * (z[2] - z[3]) + y[2] * (z[3] - z[1]) + y[3] * (z[1] - z[2]);
        float B = z[1] * (x[2] - x[3]) + z[2] * (x[3] - x[1]) + z[3] * (x[1] - x[2]);
        float C = x[1] * (y[2] - y[3]) + x[2] * (y[3] - y[1]) + x[3] * (y[1] - y[2]);
        float op = std::sqrt(A * A + B * B + C * C);
        float cx = A / op, cy = B / op, cz = C / op;
        A = x[2] - x[1];
        B = y[2] - y[1];
        C = z[2] - z[1];
        float d = 1.0f / std::sqrt(A * A + B * B + C * C);
        op = dis[a] * d;
        float xn = A * op, yn = B * op, zn = C * op;
        A = cx * cx;
        B = cy * cy;
        C = cz * cz;
        float angPI = (float)(ang[a] * PI_REF / 180.0f);
        float ct = trig_cos(angPI);
        float st = -trig_sin(angPI);
        op = 1.0f - ct;
        float xk = (cx * cz * op - cy * st) * zn + ((1.0f - A) * ct + A) * xn + (cx * cy * op + cz * st) * yn;
        float yk = (cy * cx * op - cz * st) * xn + ((1.0f - B) * ct + B) * yn + (cy * cz * op + cx * st) * zn;
        float zk = (cz * cy * op - cx * st) * yn + ((1.0f - C) * ct + C) * zn + (cz * cx * op + cy * st) * xn;
        float dihPI = (float)(dih[a] * PI_REF / 180.0f);
        ct = trig_cos(dihPI);
        st = trig_sin(dihPI);
        op = 1.0f - ct;
        cx = (x[2] - x[1]) * d;
        cy = (y[2] - y[1]) * d;
        cz = (z[2] - z[1]) * d;
        A = cx * cx;
        B = cy * cy;
        C = cz * cz;
        xyz[3 * a + 0] = (cx * cz * op - cy * st) * zk + ((1.0f - A) * ct + A) * xk + (cx * cy * op + cz * st) * yk + x[1];
        xyz[3 * a + 1] = (cy * cx * op - cz * st) * xk + ((1.0f - B) * ct + B) * yk + (cy * cz * op + cx * st) * zk + y[1];
        xyz[3 * a + 2] = (cz * cy * op - cx * st) * yk + ((1.0f - C) * ct + C) * zk + (cz * cx * op + cy * st) * xk + z[1];
    }
}

// ---------------------------------------------------------------- cell grid
// LIB/Vcontacts.cpp:1570-1715 in steady state (vindex = 0, only scorable
// atoms widen the grid).  Returns false when an atom lands outside the grid.
template <class R> bool index_atoms(const System& S, Work<R>& W)
{
    const fa_ligand_desc* lig = S.lig;
    const int N = S.N(), Rn = S.R();
    float gmin[3], gmax[3];
    for (int j = 0; j < 3; ++j) {
        gmin[j] = (float)((int)lig->globalmin[j] + 1.0);
        gmax[j] = (float)((int)lig->globalmax[j] + 1.0);
    }
    float max_width = lig->maxwidth;
    int alter = 0;
    (void)Rn;
    for (int i : S.scorable) {  // calc order: side-chain atoms, then the ligand
        const float c[3] = {W.x[i], W.y[i], W.z[i]};
        for (int j = 0; j < 3; ++j) {
            if (c[j] < gmin[j]) {
                gmin[j] = c[j];
                alter = 1;
            } else if (c[j] > gmax[j]) {
                gmax[j] = c[j];
                alter = 1;
            }
        }
    }
    if (alter) {
        for (int j = 0; j < 3; ++j) {
            gmin[j] = (float)(gmin[j] < 0.0f ? (int)gmin[j] - 1.0 : (int)gmin[j] + 1.0);
            gmax[j] = (float)(gmax[j] < 0.0f ? (int)gmax[j] - 1.0 : (int)gmax[j] + 1.0);
            float diff = gmax[j] - gmin[j];
            if (diff > max_width) max_width = diff;
        }
    }
    const int dim = (int)(max_width / CELLSIZE) + 1;
    const int dim2 = dim * dim, dim3 = dim2 * dim;
    W.dim = dim;
    W.boxfirst.assign(dim3, 0);
    W.boxcount.assign(dim3, 0);
    for (int i = 0; i < N; ++i) {
        int b = (int)((W.x[i] - gmin[0]) / CELLSIZE) * dim2 + (int)((W.y[i] - gmin[1]) / CELLSIZE) * dim +
                (int)((W.z[i] - gmin[2]) / CELLSIZE);
        if (b < 0 || b >= dim3) return false;
        W.boxnum[i] = b;
        ++W.boxcount[b];
    }
    int start = 0;
    for (int b = 0; b < dim3; ++b) {
        W.boxfirst[b] = start;
        start += W.boxcount[b];
        W.boxcount[b] = 0;
    }
    // a box lists its atoms in calc order (reference); independent mode
    // lists the rigid atoms first, then the scorable ones
    for (int pass = 0; pass < 2; ++pass) {
        for (int i = 0; i < N; ++i) {
            if (m_rigidfirst() && (pass == 0) == (S.slot[i] >= 0)) continue;
            if (!m_rigidfirst() && pass == 1) break;
            const int b = W.boxnum[i];
            W.calclist[W.boxfirst[b] + W.boxcount[b]] = i;
            ++W.boxcount[b];
        }
    }
    // processing rank of each scorable atom (its place in the box order)
    int k = 0;
    for (int q = 0; q < N; ++q) {
        const int i = W.calclist[q];
        if (S.slot[i] >= 0) W.rank[i] = k++;
    }
    return true;
}

// ------------------------------------------------------- neighbour culling
// LIB/Vcontacts.cpp:1725-1839.  `clash` non-null selects the pre-pass form.
template <class R>
int get_contacts(const System& S, Work<R>& W, int a, float rado, double* clash)
{
    const int dim = W.dim, dim2 = dim * dim, dim3 = dim2 * dim;
    const double permea = (double)S.prm->permeability;
    // the speculative form runs the serial pre-pass (done marks), then full
    // neighbour lists for the polyhedra
    const bool indep = m_nodone() && !(m_spec() && clash);
    // earlier atoms that recorded a contact with `a` stay eligible ('C')
    if (!indep)
        for (int j : W.stubs[S.slot[a]]) W.done[j] = 'C';
    int NC = 0;
    W.contlist.clear();
    const int bz = W.boxnum[a];
    for (int i = 0; i < 27; ++i) {
        const int bi = bz + dim2 * ((i / 9) - 1) + dim * (((i / 3) % 3) - 1) + (i % 3) - 1;
        if (bi < 0 || bi >= dim3) continue;
        for (int k = 0; k < W.boxcount[bi]; ++k) {
            const int j = W.calclist[W.boxfirst[bi] + k];
            if (t_caps && clash && j != a) {
                const float ex = W.x[a] - W.x[j], ey = W.y[a] - W.y[j], ez = W.z[a] - W.z[j];
                const double sq = (double)(ex * ex + ey * ey + ez * ez);
                const double nd = (double)(rado + W.rad[j] + RW);
                if (sq < nd * nd && sq != 0.0) ++W.near_count;
            }
            if (!indep && W.done[j] == 'Y') continue;  // every atom is exposed here
            // independent pre-pass: a pair of scorable atoms is counted by the
            // one processed first (the pairs the reference's done marks leave)
            if (indep && clash && S.slot[j] >= 0 && W.rank[j] < W.rank[a]) continue;
            if (j != a) OPS(cull, 1);        // the atom itself never makes a contact
            const double rAB = (double)(W.rad[a] + W.rad[j]);
            const float dx = W.x[a] - W.x[j], dy = W.y[a] - W.y[j], dz = W.z[a] - W.z[j];
            const double sqrdist = (double)(dx * dx + dy * dy + dz * dz);
            const double neardist = (double)(rado + W.rad[j] + RW);
            if (sqrdist < neardist * neardist && sqrdist != 0.0) {
                const double dist = std::sqrt(sqrdist);
                W.contlist.push_back({j, (R)dist});
                if (clash) {
                    const double clashdist = permea * rAB;
                    if (dist < clashdist) {
                        const bool intra = S.same_res(a, j);
                        if (!intra || !S.bonded(a, j))
                            *clash += (double)KWALL_F * (std::pow(dist, -12.0) - std::pow(clashdist, -12.0));
                    }
                    if (!indep) W.done[a] = 'Y';
                }
                ++NC;
            }
        }
    }
    if (!indep)
        for (int j : W.stubs[S.slot[a]]) W.done[j] = 'Y';
    return NC;
}

// ----------------------------------------------------------- first vertex
// LIB/Vcontacts.cpp:499-590
template <class R> void heuristic_firstvert(const Plane<R>* cont, int* pA, int* pB, int* pC, int NC);

template <class R> void get_firstvert(const Work<R>& W, const Plane<R>* cont, int* pA, int* pB, int* pC, int NC, int a)
{
    *pA = *pB = *pC = -1;
    const int* sd = &W.seed[3 * a];
    if (!m_noseed() && sd[0] != -1) {
        for (int c = 0; c < NC; ++c) {
            if (cont[c].index == sd[0])
                *pA = c;
            else if (cont[c].index == sd[1])
                *pB = c;
            else if (cont[c].index == sd[2])
                *pC = c;
        }
    }
    if (*pA != -1 && *pB != -1 && *pC != -1) return;
    heuristic_firstvert(cont, pA, pB, pC, NC);
}

// The reference's start without a seed (LIB/Vcontacts.cpp:516-590): A the
// closest plane, B the plane whose line with A passes closest to the atom, C
// the plane cutting that line nearest to that point.  Nothing guarantees the
// result is a vertex of the polyhedron.
template <class R> void heuristic_firstvert(const Plane<R>* cont, int* pA, int* pB, int* pC, int NC)
{
    *pA = *pB = *pC = -1;
    R mindist = (R)9.9e+9;
    for (int c = 0; c < NC; ++c) {
        if (cont[c].dist < mindist) {
            mindist = cont[c].dist;
            *pA = c;
        }
    }
    R ptA[3] = {0, 0, 0}, vA[4], tp[3];
    mindist = (R)9.9e+9;
    for (int c = 0; c < NC + 4; ++c) {
        if (c == *pA) continue;
        vA[0] = cont[*pA].Ai[1] * cont[c].Ai[2] - cont[*pA].Ai[2] * cont[c].Ai[1];
        vA[1] = cont[*pA].Ai[2] * cont[c].Ai[0] - cont[*pA].Ai[0] * cont[c].Ai[2];
        vA[2] = cont[*pA].Ai[0] * cont[c].Ai[1] - cont[*pA].Ai[1] * cont[c].Ai[0];
        vA[3] = 0;
        if (solve_3x3(cont[*pA].Ai, cont[c].Ai, vA, tp) != -1) {
            R d = std::sqrt(tp[0] * tp[0] + tp[1] * tp[1] + tp[2] * tp[2]);
            if (d < mindist) {
                *pB = c;
                mindist = d;
                ptA[0] = tp[0];
                ptA[1] = tp[1];
                ptA[2] = tp[2];
            }
        }
    }
    vA[0] = cont[*pA].Ai[1] * cont[*pB].Ai[2] - cont[*pA].Ai[2] * cont[*pB].Ai[1];
    vA[1] = cont[*pA].Ai[2] * cont[*pB].Ai[0] - cont[*pA].Ai[0] * cont[*pB].Ai[2];
    vA[2] = cont[*pA].Ai[0] * cont[*pB].Ai[1] - cont[*pA].Ai[1] * cont[*pB].Ai[0];
    R vtmin = (R)9.9e+9;
    for (int c = 0; c < NC + 4; ++c) {
        if (c == *pA || c == *pB) continue;
        R vtdiv = cont[c].Ai[0] * vA[0] + cont[c].Ai[1] * vA[1] + cont[c].Ai[2] * vA[2];
        if (vtdiv != R(0)) {
            R vt = -(cont[c].Ai[0] * ptA[0] + cont[c].Ai[1] * ptA[1] + cont[c].Ai[2] * ptA[2] + cont[c].Ai[3]) / vtdiv;
            if (std::fabs(vt) < vtmin) {
                vtmin = std::fabs(vt);
                *pC = c;
            }
        }
    }
}

// -------------------------------------------------- speculative start
// First plane of [0, n), other than x0 and x1, that the ray p + t d (t >= 0)
// leaves through: the smallest parameter among planes the ray moves towards
// (n.d > 0), the lowest index among equals.  -1 when none.
template <class R> int first_hit(const Plane<R>* cont, int n, const R* p, const R* d, int x0, int x1, R* tmin)
{
    int best = -1;
    R tb = (R)9.9e+9;
    for (int c = 0; c < n; ++c) {
        if (c == x0 || c == x1) continue;
        const R den = cont[c].Ai[0] * d[0] + cont[c].Ai[1] * d[1] + cont[c].Ai[2] * d[2];
        if (!(den > R(0))) continue;
        const R t = -(cont[c].Ai[0] * p[0] + cont[c].Ai[1] * p[1] + cont[c].Ai[2] * p[2] + cont[c].Ai[3]) / den;
        if (t < tb) {
            tb = t;
            best = c;
        }
    }
    *tmin = tb;
    return best;
}

// A vertex of the polyhedron, found by ray casting from the atom's centre
// (inside the polyhedron when no plane engulfs it): out along a fixed
// direction to the first plane, within that plane to the second, along their
// line to the third.  false when a plane engulfs the atom or a ray escapes.
template <class R> bool ray_firstvert(const Plane<R>* cont, int NC, int* pA, int* pB, int* pC)
{
    for (int c = 0; c < NC; ++c)
        if (cont[c].Ai[3] > R(0)) return false;
    const R o[3] = {0, 0, 0};
    const R d1[3] = {(R)0.48, (R)0.6, (R)0.64};
    R t1, t2, t3;
    const int P1 = first_hit(cont, NC + 4, o, d1, -1, -1, &t1);
    if (P1 < 0) return false;
    const R* n1 = cont[P1].Ai;
    const R p1[3] = {t1 * d1[0], t1 * d1[1], t1 * d1[2]};
    const R nn = n1[0] * n1[0] + n1[1] * n1[1] + n1[2] * n1[2];
    const R dn = (d1[0] * n1[0] + d1[1] * n1[1] + d1[2] * n1[2]) / nn;
    const R d2[3] = {d1[0] - dn * n1[0], d1[1] - dn * n1[1], d1[2] - dn * n1[2]};
    const int P2 = first_hit(cont, NC + 4, p1, d2, P1, -1, &t2);
    if (P2 < 0) return false;
    const R* n2 = cont[P2].Ai;
    const R p2[3] = {p1[0] + t2 * d2[0], p1[1] + t2 * d2[1], p1[2] + t2 * d2[2]};
    const R d3[3] = {n1[1] * n2[2] - n1[2] * n2[1], n1[2] * n2[0] - n1[0] * n2[2], n1[0] * n2[1] - n1[1] * n2[0]};
    const int P3 = first_hit(cont, NC + 4, p2, d3, P1, P2, &t3);
    if (P3 < 0) return false;
    *pA = P1;
    *pB = P2;
    *pC = P3;
    return true;
}

// Is the point of planes (pA, pB, pC) inside every other plane of [0, n)?
template <class R> bool vertex_valid(const Plane<R>* cont, int n, int pA, int pB, int pC)
{
    R tp[3];
    if (solve_3x3(cont[pA].Ai, cont[pB].Ai, cont[pC].Ai, tp) == -1) return false;
    for (int c = 0; c < n; ++c) {
        if (c == pA || c == pB || c == pC) continue;
        if (cont[c].Ai[0] * tp[0] + cont[c].Ai[1] * tp[1] + cont[c].Ai[2] * tp[2] + cont[c].Ai[3] > (R)SPEC_VALID_EPS)
            return false;
    }
    return true;
}

// ----------------------------------------------------------------- hull
// LIB/Vcontacts.cpp:147-490.  Returns the number of vertices kept inside the
// sphere or -1 when the edge walk does not converge (VC->recalc == 0).
template <class R> int voronoi_poly(const System& S, Work<R>& W, int a, float rado, int NC)
{
    Plane<R>* cont = W.cont.data();
    const double rado_d = (double)rado;
    for (int c = 0; c < NC; ++c) {
        const int j = W.contlist[c].index;
        const R atomdist = W.contlist[c].dist;
        OPS(plane, 1);
        if (atomdist < (R)1.0e-10) continue;  // degenerate, never produced by the cull
        // float products as in LIB/Vcontacts.cpp:202 (float operands)
        const float rj = RW + W.rad[j];
        const R planedist = (atomdist * atomdist + (R)(rado * rado) - (R)(rj * rj)) / (2 * atomdist);
        cont[c].Ai[0] = (R)(W.x[j] - W.x[a]) / atomdist;
        cont[c].Ai[1] = (R)(W.y[j] - W.y[a]) / atomdist;
        cont[c].Ai[2] = (R)(W.z[j] - W.z[a]) / atomdist;
        cont[c].Ai[3] = -planedist;
        cont[c].dist = std::fabs(planedist);
        cont[c].index = j;
        cont[c].flag = 'X';
    }
    (void)rado_d;
    // four bounding planes (LIB/Vcontacts.cpp:223-241)
    const R bp[4][4] = {{(R)0.707, 1, 0, -10}, {(R)0.707, -1, 0, -10}, {(R)-0.707, 0, 1, -10}, {(R)-0.707, 0, -1, -10}};
    for (int k = 0; k < 4; ++k) {
        for (int q = 0; q < 4; ++q) cont[NC + k].Ai[q] = bp[k][q];
        cont[NC + k].dist = 0;
        cont[NC + k].index = -1;
        cont[NC + k].flag = 'X';
        cont[NC + k].area = 0;
    }
    int pA, pB, pC;
    if (m_spec()) {
        // speculative form: a start that is a vertex, by ray casting, else
        // the heuristic start if it is one; otherwise the pose goes serial
        if (!ray_firstvert(cont, NC, &pA, &pB, &pC)) {
            heuristic_firstvert(cont, &pA, &pB, &pC, NC);
            if (pA < 0 || pB < 0 || pC < 0 || !vertex_valid(cont, NC + 4, pA, pB, pC)) return -3;
        }
    } else {
        get_firstvert(W, cont, &pA, &pB, &pC, NC, a);
    }
    Vert<R>* poly = W.poly.data();
    Edge<R>* vedge = W.vedge.data();
    R tp[3];
    solve_3x3(cont[pA].Ai, cont[pB].Ai, cont[pC].Ai, tp);
    add_vertex(poly, 0, tp, pA, pB, pC);
    cont[pA].flag = 'Y';
    cont[pB].flag = 'Y';
    cont[pC].flag = 'Y';
    add_vedge(vedge, 0, cont, pA, pB, pC, poly, 0);
    add_vedge(vedge, 1, cont, pB, pC, pA, poly, 0);
    add_vedge(vedge, 2, cont, pC, pA, pB, poly, 0);
    int startedge = 0, edgenum = 3, vn = 1;
    const R rr = (R)(rado * rado);  // float product (LIB/Vcontacts.cpp:310)
    while (true) {
        while (vedge[startedge].endpt >= 0 && (edgenum - startedge) > 0) ++startedge;
        if ((edgenum - startedge) <= 0) break;
        R vtmin = (R)9.9e+9;
        const Vert<R> vs = poly[vedge[startedge].startpt];
        const R* V = vedge[startedge].V;
        const int A = vedge[startedge].plane[0], B = vedge[startedge].plane[1];
        const int oldC = vedge[startedge].startplane;
        int C = -1;
        for (int c = 0; c < NC; ++c) {
            if (c != A && c != B && c != oldC) {
                OPS(hull_test, 1);
                R vtdiv = cont[c].Ai[0] * V[0] + cont[c].Ai[1] * V[1] + cont[c].Ai[2] * V[2];
                if (vtdiv != R(0)) {
                    R vt = -(cont[c].Ai[0] * vs.xi[0] + cont[c].Ai[1] * vs.xi[1] + cont[c].Ai[2] * vs.xi[2] + cont[c].Ai[3]) / vtdiv;
                    if (vt < vtmin && vt > R(0)) {
                        vtmin = vt;
                        C = c;
                    }
                }
            }
        }
        R np[3] = {vs.xi[0] + vtmin * V[0], vs.xi[1] + vtmin * V[1], vs.xi[2] + vtmin * V[2]};
        if (np[0] * np[0] + np[1] * np[1] + np[2] * np[2] > rr) {
            for (int c = NC; c < NC + 4; ++c) {
                if (c != A && c != B && c != oldC) {
                    OPS(hull_test, 1);
                    R vtdiv = cont[c].Ai[0] * V[0] + cont[c].Ai[1] * V[1] + cont[c].Ai[2] * V[2];
                    if (vtdiv != R(0)) {
                        R vt = -(cont[c].Ai[0] * vs.xi[0] + cont[c].Ai[1] * vs.xi[1] + cont[c].Ai[2] * vs.xi[2] + cont[c].Ai[3]) / vtdiv;
                        if (vt < vtmin && vt > R(0)) {
                            vtmin = vt;
                            C = c;
                        }
                    }
                }
            }
            np[0] = vs.xi[0] + vtmin * V[0];
            np[1] = vs.xi[1] + vtmin * V[1];
            np[2] = vs.xi[2] + vtmin * V[2];
        }
        if ((int)W.poly.size() <= vn + 2 || (int)W.vedge.size() <= edgenum + 2) return -2;  // capacity
        add_vertex(poly, vn, np, A, B, C);
        vedge[startedge].endpt = vn;
        vedge[startedge].endplane = C;
        cont[C].flag = 'Y';
        // edge (A, C)
        bool found = false;
        for (int e = startedge + 1; e < edgenum; ++e) {
            if ((vedge[e].plane[0] == A && vedge[e].plane[1] == C) || (vedge[e].plane[0] == C && vedge[e].plane[1] == A)) {
                vedge[e].endpt = vn;
                vedge[e].endplane = B;
                found = true;
                break;
            }
        }
        if (!found) {
            add_vedge(vedge, edgenum, cont, A, C, B, poly, vn);
            ++edgenum;
        }
        // edge (B, C)
        found = false;
        for (int e = startedge + 1; e < edgenum; ++e) {
            if ((vedge[e].plane[0] == B && vedge[e].plane[1] == C) || (vedge[e].plane[0] == C && vedge[e].plane[1] == B)) {
                vedge[e].endpt = vn;
                vedge[e].endplane = A;
                found = true;
                break;
            }
        }
        if (!found) {
            add_vedge(vedge, edgenum, cont, B, C, A, poly, vn);
            ++edgenum;
            if (edgenum >= EDGE_LIMIT) {  // hull failure; VC->recalc == 0
                CAPMAX(vn, vn + 1);
                CAPMAX(edges, edgenum);
                W.last_edges = edgenum;
                return -1;
            }
        }
        ++vn;
    }
    CAPMAX(edges, edgenum);
    W.last_edges = edgenum;
    // arc points on edges that cross the sphere (LIB/Vcontacts.cpp:419-457)
    for (int e = 0; e < edgenum; ++e)
        vedge[e].arc = ((R)rado < poly[vedge[e].startpt].dist || (R)rado < poly[vedge[e].endpt].dist) ? '?' : 'X';
    for (int e = 0; e < edgenum; ++e) {
        if (vedge[e].arc == 'X') continue;
        OPS(arc_solve, 1);
        R p0[3], p1[3];
        if (solve_2xS(cont[vedge[e].plane[0]].Ai, cont[vedge[e].plane[1]].Ai, rado, p0, p1) == -1) {
            vedge[e].arc = 'X';
            continue;
        }
        const int ta = vedge[e].startplane, tb = vedge[e].endplane;
        for (int s = 0; s < 2; ++s) {
            const R* p = s ? p1 : p0;
            R va = cont[ta].Ai[0] * p[0] + cont[ta].Ai[1] * p[1] + cont[ta].Ai[2] * p[2] + cont[ta].Ai[3];
            R vb = cont[tb].Ai[0] * p[0] + cont[tb].Ai[1] * p[1] + cont[tb].Ai[2] * p[2] + cont[tb].Ai[3];
            if (va < R(0) && vb < R(0)) {
                if ((int)W.poly.size() <= vn + 1) return -2;
                add_vertex(poly, vn, p, vedge[e].plane[0], vedge[e].plane[1], -1);
                poly[vn].dist = (R)rado;
                ++vn;
            }
        }
    }
    CAPMAX(vn, vn);
    // keep vertices inside the sphere (LIB/Vcontacts.cpp:460-467)
    int vj = 0;
    for (int vi = 0; vi < vn; ++vi)
        if (poly[vi].dist <= (R)rado) poly[vj++] = poly[vi];
    vn = vj;
    // centre points and engulfing flags (LIB/Vcontacts.cpp:470-478)
    for (int c = 0; c < NC; ++c) {
        W.centerpt[c].xi[0] = -cont[c].Ai[0] * cont[c].Ai[3];
        W.centerpt[c].xi[1] = -cont[c].Ai[1] * cont[c].Ai[3];
        W.centerpt[c].xi[2] = -cont[c].Ai[2] * cont[c].Ai[3];
        W.centerpt[c].dist = std::fabs(cont[c].Ai[3]);
        if (cont[c].Ai[3] > R(0)) cont[c].flag = 'E';
    }
    // seeds for atoms not yet done (LIB/Vcontacts.cpp:595-629)
    for (int vi = 0; vi < vn && !m_noseed(); ++vi) {
        if (poly[vi].plane[2] == -1) continue;
        const int i0 = cont[poly[vi].plane[0]].index, i1 = cont[poly[vi].plane[1]].index, i2 = cont[poly[vi].plane[2]].index;
        int* s0 = &W.seed[3 * i0];
        if (s0[0] == -1) { s0[0] = a; s0[1] = i1; s0[2] = i2; }
        int* s1 = &W.seed[3 * i1];
        if (s1[0] == -1) { s1[0] = a; s1[1] = i0; s1[2] = i2; }
        int* s2 = &W.seed[3 * i2];
        if (s2[0] == -1) { s2[0] = a; s2[1] = i0; s2[2] = i1; }
    }
    return vn;
}

// ----------------------------------------------------------- face order
// LIB/Vcontacts.cpp:699-831 (appends opposite arc points to poly)
template <class R> int order_faces(Work<R>& W, float rado, int NC, int NV)
{
    Plane<R>* cont = W.cont.data();
    std::vector<Vert<R>>& poly = W.poly;
    if (t_caps) t_caps->face_slots_run = 0;
    for (int X = 0; X < NC; ++X) {
        std::vector<int>& pt = W.ptorder[X];
        pt.clear();
        if (cont[X].flag == 'X') continue;
        for (int vi = 0; vi < NV; ++vi)
            if (poly[vi].plane[0] == X || poly[vi].plane[1] == X || poly[vi].plane[2] == X) pt.push_back(vi);
        int sc = (int)pt.size();
        if (t_caps) t_caps->face_slots_run += sc + (sc > 2 ? 1 : 0);
        if (sc > 2) {
            for (int vi = 0; vi < sc; ++vi) {
                if (poly[pt[vi]].plane[2] == -1) {
                    R t[3];
                    for (int k = 0; k < 3; ++k) t[k] = 2 * W.centerpt[X].xi[k] - poly[pt[vi]].xi[k];
                    if (test_point(t, cont, NC, X, -1, -1) == 'Y') {
                        if ((int)poly.size() <= NV + 1) return -2;
                        poly[NV].xi[0] = t[0];
                        poly[NV].xi[1] = t[1];
                        poly[NV].xi[2] = t[2];
                        poly[NV].plane[0] = X;
                        poly[NV].plane[1] = -1;
                        poly[NV].plane[2] = -1;
                        poly[NV].dist = (R)rado;
                        pt.push_back(NV);
                        ++sc;
                        ++NV;
                    }
                    break;
                }
            }
        }
        if (sc > 3) {
            const Vert<R>& p0 = poly[pt[0]];
            const int Y = (p0.plane[0] == X) ? p0.plane[1] : p0.plane[0];
            int vi;
            for (vi = 1; vi < sc; ++vi)
                if (poly[pt[vi]].plane[0] == Y || poly[pt[vi]].plane[1] == Y || poly[pt[vi]].plane[2] == Y) break;
            if (vi == sc) return -3;  // no second point on the common line (reference reads past the list)
            std::swap(pt[vi], pt[1]);
            std::vector<R> cs(sc);
            OPS(face_order, sc - 2);
            for (vi = 2; vi < sc; ++vi) cs[vi] = cosPQR(poly[pt[1]].xi, poly[pt[0]].xi, poly[pt[vi]].xi);
            for (vi = 2; vi < sc - 1; ++vi)
                for (int v2 = vi + 1; v2 < sc; ++v2)
                    if (cs[vi] < cs[v2]) {
                        std::swap(pt[vi], pt[v2]);
                        std::swap(cs[vi], cs[v2]);
                    }
        }
    }
    if (t_caps) {
        CAPMAX(face_slots, t_caps->face_slots_run);
        CAPMAX(nv_opp, NV);
    }
    return NV;
}

// ----------------------------------------------------------- areas
// LIB/Vcontacts.cpp:1278-1344 (project_points) and :838-1044 (calc_areas)
template <class R> void calc_areas(Work<R>& W, float rado, int NC, int NV, int a)
{
    Plane<R>* cont = W.cont.data();
    Vert<R>* poly = W.poly.data();
    const Vert<R>* ctr = W.centerpt.data();
    const R ro = (R)rado;
    bool eng = false;
    int epi = 0, engplane[4];
    for (int X = 0; X < NC; ++X) {
        cont[X].area = 0;
        if (cont[X].flag == 'E') {
            eng = true;
            engplane[epi++] = X;
            if (epi == 3) break;
        }
    }
    if (t_branch) {
        ++t_branch->atoms;
        if (epi == 1) ++t_branch->epi1;
        else if (epi == 2) ++t_branch->epi2;
        else if (epi >= 3) ++t_branch->epi3;
    }
    if (eng) {
        R proj[3], p0[3], p1[3];
        if (epi == 1) {
            for (int k = 0; k < 3; ++k) proj[k] = ctr[engplane[0]].xi[k];
        } else if (epi == 2) {
            if (solve_2xS(cont[engplane[0]].Ai, cont[engplane[1]].Ai, rado, p0, p1) == -1) {
                // the reference uses the atom's absolute coordinates here
                proj[0] = (R)(W.x[a] + rado);
                proj[1] = (R)W.y[a];
                proj[2] = (R)W.z[a];
            } else {
                for (int k = 0; k < 3; ++k) proj[k] = (p0[k] + p1[k]) / 2;
            }
        } else {
            solve_3x3(cont[engplane[0]].Ai, cont[engplane[1]].Ai, cont[engplane[2]].Ai, p0);
            for (int k = 0; k < 3; ++k) proj[k] = p0[k];
        }
        for (int vi = 0; vi < NV; ++vi) {
            if (poly[vi].plane[2] == -1) continue;
            R* P = poly[vi].xi;
            R V[3] = {P[0] - proj[0], P[1] - proj[1], P[2] - proj[2]};
            R qa = V[0] * V[0] + V[1] * V[1] + V[2] * V[2];
            R qb = 2 * (P[0] * V[0] + P[1] * V[1] + P[2] * V[2]);
            R qc = P[0] * P[0] + P[1] * P[1] + P[2] * P[2] - (R)(rado * rado);
            R k = (std::sqrt(qb * qb - 4 * qa * qc) - qb) / (2 * qa);
            P[0] += k * V[0];
            P[1] += k * V[1];
            P[2] += k * V[2];
            poly[vi].dist = ro;
        }
    }
    std::vector<R> cNN1, cN0;
    for (int X = 0; X < NC; ++X) {
        const std::vector<int>& pt = W.ptorder[X];
        const int NP = (int)pt.size();
        if (cont[X].flag == 'X') continue;
        R area = 0;
        if (NP == 0) {
            if (test_point(ctr[X].xi, cont, NC, X, -1, -1) == 'Y')
                cont[X].area = (R)(2.0f * PI_REF) * ro * (ro - ctr[X].dist);
        } else if (NP == 2) {
            if (test_point(ctr[X].xi, cont, NC, X, -1, -1) == 'Y')
                cont[X].area = (R)(2.0f * PI_REF) * ro * (ro - ctr[X].dist) - spherical_arc(ctr[X], poly[pt[0]], poly[pt[1]], rado);
            else
                cont[X].area = spherical_arc(ctr[X], poly[pt[0]], poly[pt[1]], rado);
        } else {
            cNN1.assign(NP, 0);
            cN0.assign(NP, 0);
            for (int vi = 0; vi < NP; ++vi) {
                const Vert<R>& v0 = poly[pt[0]];
                const Vert<R>& va = poly[pt[vi]];
                const Vert<R>& vb = poly[pt[(vi + 1) % NP]];
                cNN1[vi] = (va.xi[0] * vb.xi[0] + va.xi[1] * vb.xi[1] + va.xi[2] * vb.xi[2]) / (va.dist * vb.dist);
                if (vi != 0) cN0[vi] = (v0.xi[0] * va.xi[0] + v0.xi[1] * va.xi[1] + v0.xi[2] * va.xi[2]) / (v0.dist * va.dist);
            }
            if (NP > 2) OPS(sph_triangle, NP - 2);
            for (int vi = 1; vi < NP - 1; ++vi) {
                R U = std::sqrt((1 + cN0[vi]) * (1 + cNN1[vi]) * (1 + cN0[vi + 1]) / R(8));
                R V = std::sqrt((1 - cN0[vi]) * (1 - cNN1[vi]) * (1 + cN0[vi + 1]) / R(8));
                R Wv = std::sqrt((1 - cN0[vi]) * (1 + cNN1[vi]) * (1 - cN0[vi + 1]) / R(8));
                R Xv = std::sqrt((1 + cN0[vi]) * (1 - cNN1[vi]) * (1 - cN0[vi + 1]) / R(8));
                R tS = (1 - U + V + Wv + Xv) / (1 + U - V - Wv - Xv);
                R tSA = (1 - U - V - Wv + Xv) / (1 + U + V + Wv - Xv);
                R tSB = (1 - U - V + Wv - Xv) / (1 + U + V - Wv + Xv);
                R tSC = (1 - U + V - Wv - Xv) / (1 + U - V + Wv + Xv);
                R tp = std::sqrt(tS * tSA * tSB * tSC);
                if (tp > R(0)) area += 4 * ro * ro * std::atan(std::sqrt(tp));
            }
            for (int vi = 0; vi < NP; ++vi) {
                const Vert<R>& va = poly[pt[vi]];
                const Vert<R>& vb = poly[pt[(vi + 1) % NP]];
                if (va.plane[2] == -1 && vb.plane[2] == -1)
                    if (va.plane[0] + va.plane[1] != vb.plane[0] + vb.plane[1])
                        area += spherical_arc(ctr[X], va, vb, rado);
            }
            cont[X].area = area;
        }
    }
    if (!eng) return;
    for (int X = 0; X < NC; ++X) {
        if (cont[X].flag != 'E') continue;
        const std::vector<int>& pt = W.ptorder[X];
        const int NP = (int)pt.size();
        for (int vi = 0; vi < NP; ++vi) {
            const Vert<R>& cu = poly[pt[vi]];
            const Vert<R>& nx = poly[pt[(vi + 1) % NP]];
            int pa1, pa2;
            if (cu.plane[0] == X) {
                pa1 = cu.plane[1];
                pa2 = cu.plane[2];
            } else {
                pa1 = cu.plane[0];
                pa2 = (cu.plane[1] == X) ? cu.plane[2] : cu.plane[1];
            }
            int comm;
            if (pa1 == nx.plane[0] || pa1 == nx.plane[1] || pa1 == nx.plane[2])
                comm = pa1;
            else if (pa2 == nx.plane[0] || pa2 == nx.plane[1] || pa2 == nx.plane[2])
                comm = pa2;
            else
                continue;
            if (comm != -1 && cont[comm].flag != 'E') {
                cont[comm].area += spherical_arc(ctr[X], cu, nx, rado);
                if (NP == 2) break;
            }
        }
    }
    if (epi == 1) {
        cont[engplane[0]].area = (R)(2.0f * PI_REF) * ro * (ro + cont[engplane[0]].dist);
    } else if (epi == 2) {
        R b[3], c[3];
        if (solve_2xS(cont[engplane[0]].Ai, cont[engplane[1]].Ai, rado, b, c) == -1) {
            cont[engplane[0]].area = (R)(2.0f * PI_REF) * ro * ro;
            cont[engplane[1]].area = (R)(2.0f * PI_REF) * ro * ro;
        } else {
            Vert<R> B{{b[0], b[1], b[2]}, ro, {0, 0, 0}}, C{{c[0], c[1], c[2]}, ro, {0, 0, 0}};
            R m = spherical_arc(ctr[engplane[0]], B, C, rado);
            m += spherical_arc(ctr[engplane[1]], B, C, rado);
            cont[engplane[0]].area = (R)(2.0f * PI_REF) * ro * ro - R(0.5) * m;
            cont[engplane[1]].area = (R)(2.0f * PI_REF) * ro * ro - R(0.5) * m;
        }
    } else if (epi >= 3) {
        for (int X = 0; X < NC; ++X)
            cont[X].area = (cont[X].flag == 'E') ? (R)(4.0f * PI_REF) * ro * ro / epi : R(0);
    }
}

// ----------------------------------------------------------- one pose
constexpr double KDIST = 1.0e3;  // LIB/flexaid.h:83

// GetValueFromGaussian, LIB/geometry.cpp:3-5
double gaussian_bell(double x, double opt, double zero)
{
    return std::pow(-(x - zero) * (x - (2 * opt - zero)) / std::pow(zero - opt, 2.0), 50.0);
}

// One type pair's scored value for `area` of an atom whose whole sphere is
// `surf`: weight 1 -> y * area (or y * area / surf with normalize_area,
// LIB/vcfunction.cpp:343-353, :479-491); density -> get_yval(area / surf),
// LIB/vcfunction.cpp:576-621, walking the nodes as the linked list is walked.
double pair_term(const fa_params* prm, int k, double area, double surf)
{
    if (prm->dens_off && prm->dens_off[k + 1] > prm->dens_off[k]) {
        const int e = prm->dens_off[k + 1];
        const float* xy = prm->dens_xy;
        const double rel = area / surf;
        int i = prm->dens_off[k];
        while (i + 1 < e && rel > (double)xy[2 * (i + 1)]) ++i;
        const float x0 = xy[2 * i], y0 = xy[2 * i + 1];
        if ((double)x0 > rel) return 0.0;            // no left bound data
        if (i + 1 == e) return (double)y0;           // no right bound data
        const float x1 = xy[2 * (i + 1)], y1 = xy[2 * (i + 1) + 1];
        const float dx = x1 - x0, dy = y1 - y0;      // float fields, float differences
        return (double)y0 + (rel - (double)x0) / (double)dx * (double)dy;
    }
    const double y = (double)prm->emat[k];
    return prm->normalize_area ? y * area / surf : y * area;
}

// Speculative form: does the reference, atom `a`'s turn in the serial order
// (LIB/Vcontacts.cpp:74-139), compute the polyhedron just computed from the
// full neighbour list?  Within the sphere the polyhedron is the same
// whenever the reference starts from a true vertex: the planes it leaves
// out (earlier scorable atoms that recorded no contact with `a`) cut it only
// outside the sphere, since radical planes are symmetric.  It starts from
// a seed when an earlier scorable atom's polyhedron has a vertex inside its
// sphere on `a`'s plane (LIB/Vcontacts.cpp:595-629) -- the same point is
// then a vertex of `a`'s polyhedron inside `a`'s sphere, so the seed is
// found here -- and from the heuristic vertex of its reduced list
// otherwise, which has to be tested.  Checked: the edge count stays clear of
// the 200-edge failure, every candidate seed names atoms of the reduced
// list, and an unseeded atom's heuristic start is a vertex.
template <class R> bool spec_check(const System& S, const Work<R>& W, int a, int NC, int NV)
{
    if (W.last_edges >= SPEC_EDGE_MAX) return false;
    const Plane<R>* cont = W.cont.data();
    const int ra = W.rank[a];
    auto earlier = [&](int c) {
        const int j = cont[c].index;
        return S.slot[j] >= 0 && W.rank[j] < ra;
    };
    // left out of the reference's list: earlier, and no record of `a` (an
    // area of the shared face, which both sides compute)
    auto excluded = [&](int c) { return earlier(c) && !(cont[c].area > R(0) && cont[c].flag != 'X'); };
    bool seeded = false;
    for (int vi = 0; vi < NV; ++vi) {
        const Vert<R>& v = W.poly[vi];
        if (v.plane[2] == -1) continue;
        for (int k = 0; k < 3; ++k) {
            const int c = v.plane[k];
            if (c >= NC || !earlier(c)) continue;
            seeded = true;
            for (int k2 = 0; k2 < 3; ++k2)
                if (k2 != k && (v.plane[k2] >= NC || excluded(v.plane[k2]))) return false;
        }
    }
    if (seeded) return true;
    std::vector<Plane<R>> rl;
    rl.reserve(NC + 4);
    for (int c = 0; c < NC; ++c)
        if (!excluded(c)) rl.push_back(cont[c]);
    const int NR = (int)rl.size();
    if (NR == 0) return true;  // no contact in the reference either: the whole sphere is exposed
    for (int k = 0; k < 4; ++k) rl.push_back(cont[NC + k]);
    int pA, pB, pC;
    heuristic_firstvert(rl.data(), &pA, &pB, &pC, NR);
    if (pA < 0 || pB < 0 || pC < 0) return false;
    return vertex_valid(rl.data(), NR + 4, pA, pB, pC);
}

template <class R>
fa_cf_result eval_pose(const System& S, Work<R>& W, const double* icv, float* xyz_out,
                       std::vector<fa_contact_rec>* recs_out)
{
    if (t_bits_override < 0) t_fallback = 0;
    const fa_ligand_desc* lig = S.lig;
    const fa_params* prm = S.prm;
    const int Rn = S.R(), L = S.L(), N = S.N();
    fa_cf_result res{};
    std::vector<float> lxyz(3 * L);
    std::vector<int> rot(S.rec->n_flex > 0 ? S.rec->n_flex : 1, 0);
    decode_and_build(lig, icv, lxyz.data(), S.rec, rot.data());
    if (xyz_out) std::memcpy(xyz_out, lxyz.data(), sizeof(float) * 3 * L);
    for (int i = 0; i < L; ++i) {
        W.x[Rn + i] = lxyz[3 * i];
        W.y[Rn + i] = lxyz[3 * i + 1];
        W.z[Rn + i] = lxyz[3 * i + 2];
    }
    // side chains in the selected conformation (the atoms fatm[rot]..latm[rot]
    // index_protein takes, LIB/Vcontacts.cpp:1598-1606)
    for (int k = 0; k < S.rec->n_flex; ++k) {
        const fa_flex_residue& f = S.rec->flex[k];
        const float* c = f.rot_xyz + (size_t)rot[k] * f.n_atoms * 3;
        for (int i = 0; i < f.n_atoms; ++i) {
            W.x[f.first + i] = c[3 * i];
            W.y[f.first + i] = c[3 * i + 1];
            W.z[f.first + i] = c[3 * i + 2];
        }
    }
    if (!index_atoms(S, W)) {
        res.cf = res.app = HUGE_VAL;  // not scored
        res.status = FA_GRID_RANGE;
        return res;
    }
    // clash pre-pass (LIB/Vcontacts.cpp:36-57)
    std::fill(W.done.begin(), W.done.end(), 'N');
    for (auto& s : W.stubs) s.clear();
    double clash = 0.0;
    W.near_count = 0;
    for (int i = 0; i < N; ++i) {
        const int a = W.calclist[i];
        if (S.slot[a] < 0) continue;
        get_contacts(S, W, a, W.rad[a] + RW, &clash);
    }
    CAPMAX(near_pose, W.near_count);
    if (clash >= CLASH_THRESHOLD) {
        res.wal = clash;
        res.cf = clash + 1.0;  // metal_coord = 1 in the clash cfstr (LIB/ic2cf.cpp:129)
        res.app = clash + 1.0;
        res.status = FA_CLASH;
        return res;
    }
    std::fill(W.done.begin(), W.done.end(), 'N');
    std::fill(W.seed.begin(), W.seed.end(), -1);
    for (auto& r : W.records) r.clear();
    // calc_region (LIB/Vcontacts.cpp:74-139)
    bool spec_flag = false;  // speculative form: the pose goes to the serial form
    for (int i = 0; i < N; ++i) {
        const int a = W.calclist[i];
        if (S.slot[a] < 0) continue;
        const float rado = W.rad[a] + RW;
        const int NC = get_contacts(S, W, a, rado, nullptr);
        CAPMAX(nc, NC);
        if (NC == 0) continue;
        if ((int)W.cont.size() < NC + 4) W.cont.resize(NC + 4);
        if ((int)W.ptorder.size() < NC) W.ptorder.resize(NC);
        if ((int)W.centerpt.size() < NC) W.centerpt.resize(NC);
        int NV = voronoi_poly(S, W, a, rado, NC);
        if (m_spec() && NV < 0) {  // hull failure, capacity or no start: every atom still runs
            spec_flag = true;
            continue;
        }
        if (NV == -1) {
            if (t_branch) ++t_branch->hull_fail;
            res.wal = (double)POLY_PENALTY;
            res.cf = res.wal + 1.0;
            res.app = res.wal + 1.0;
            res.status = FA_HULL_FAIL;
            return res;
        }
        if (NV < 0) {
            res.cf = res.app = HUGE_VAL;
            res.status = FA_CAPACITY;
            return res;
        }
        int NV2 = order_faces(W, rado, NC, NV);
        if (m_spec() && NV2 < 0) {
            spec_flag = true;
            continue;
        }
        if (NV2 < 0) {  // -3: the reference reads past a face's point list; -2: capacity
            res.cf = res.app = HUGE_VAL;
            res.status = NV2 == -3 ? FA_UNDEFINED_FACE : FA_CAPACITY;
            return res;
        }
        calc_areas(W, rado, NC, NV, a);
        if (m_spec() && !spec_check(S, W, a, NC, NV)) spec_flag = true;
        // save_areas (LIB/Vcontacts.cpp:1149-1220): own list in reverse
        std::vector<fa_contact_rec>& rl = W.records[S.slot[a]];
        for (int c = NC - 1; c >= 0; --c) {
            if (W.cont[c].area > R(0) && W.cont[c].flag != 'X')
                rl.push_back({a, W.cont[c].index, (double)W.cont[c].area, (double)W.contlist[c].dist});
        }
        if (!m_nodone()) {
            W.done[a] = 'Y';
            for (int c = 0; c < NC; ++c) {
                const int j = W.cont[c].index;
                if (W.done[j] != 'Y' && W.cont[c].area > R(0) && W.cont[c].flag != 'X' && S.slot[j] >= 0)
                    W.stubs[S.slot[j]].push_back(a);
            }
        }
    }
    if (spec_flag) {
        // the serial form decides this pose (build, pre-pass and polyhedra again)
        t_fallback = 1;
        const int saved = t_bits_override;
        t_bits_override = mode_bits() & ~8;
        const fa_cf_result r2 = eval_pose(S, W, icv, nullptr, recs_out);
        t_bits_override = saved;
        return r2;
    }
    // vcfunction (LIB/vcfunction.cpp:77-527), scorable atoms in calc order,
    // each residue's terms in its own cfstr (optres[].cf); ic2cf then adds the
    // residues in optres order, the ligand first (LIB/ic2cf.cpp:143-245)
    const double permea = (double)prm->permeability;
    const double dee = (double)prm->dee_clash;
    const int nt = prm->ntypes;
    std::vector<fa_cf_result> acc(1 + S.rec->n_flex, fa_cf_result{});
    for (int i : S.scorable) {
        fa_cf_result& res = acc[S.res[i] + 1];
        const double radA = (double)W.rad[i];
        const double radoA = radA + RW;
        double SAS = 4.0 * PI_REF * radoA * radoA;
        const double surfA = SAS;
        // constraints of this atom in lookup order (atoms[].cons): a ligand
        // pair is listed on both of its atoms; side-chain atoms carry none
        std::vector<const fa_constraint*> cons;
        std::vector<int> cons_partner;  // calc index
        for (int q = 0; q < lig->n_cons && i >= Rn; ++q) {
            const fa_constraint& k = lig->cons[q];
            if (k.atom == i - Rn) {
                cons.push_back(&k);
                cons_partner.push_back(k.partner_is_ligand ? Rn + k.partner : k.partner);
            } else if (k.partner_is_ligand && k.partner == i - Rn) {
                cons.push_back(&k);
                cons_partner.push_back(Rn + k.atom);
            }
        }
        for (const fa_constraint* k : cons)  // LIB/vcfunction.cpp:146-166
            if (k->type == 1) res.con += KDIST;
        const std::vector<fa_contact_rec>& recs = W.records[S.slot[i]];
        OPS(vc_contact, recs.size());
        for (const fa_contact_rec& r : recs) {
            const int c = r.atom;
            const double rAB = radA + (double)W.rad[c];
            const double area = r.area;
            const int pk = (W.type[i] - 1) * nt + (W.type[c] - 1);
            SAS -= area;
            const bool intra = S.same_res(i, c);
            if (intra && S.bonded(c, i)) continue;
            if (S.slot[c] >= 0 && c < i) continue;  // FA->contacts: already visited
            // both in one residue, or both protein residues (LIB/vcfunction.cpp:219-229)
            const bool intramol = intra || (i < Rn && c < Rn);
            // covalent constraint (LIB/vcfunction.cpp:270-310)
            bool covalent = false;
            double dist_opt = 0.0;
            for (size_t q = 0; q < cons.size(); ++q) {
                if (cons_partner[q] != c) continue;
                if (cons[q]->type == 1) {
                    covalent = true;
                    dist_opt = cons[q]->bond_len;
                    res.con -= KDIST * gaussian_bell(r.dist, dist_opt, cons[q]->max_dist);
                }
                break;
            }
            float clash_distance = permea * rAB;
            if (covalent && permea * dist_opt < permea * rAB) clash_distance = permea * dist_opt;
            if (r.dist < (double)clash_distance) {
                const double d = r.dist, d2 = d * d, d4 = d2 * d2, d6 = d4 * d2;
                const double cr = permea * rAB, cr2 = cr * cr, cr4 = cr2 * cr2, cr6 = cr4 * cr2;
                res.wal += (double)KWALL_F * (1.0 / (d6 * d6) - 1.0 / (cr6 * cr6));
                if (r.dist <= dee * rAB) res.rclash = 1;
            }
            if (!covalent && (prm->intramolecular || !intramol)) res.com += pair_term(prm, pk, area, surfA);
        }
        if (SAS < 0.0) SAS = 0.0;
        res.totsas += SAS;
        if (prm->solventterm != 0.0f)
            res.sas += (double)prm->solventterm * SAS;
        else
            res.sas += pair_term(prm, (W.type[i] - 1) * nt + (nt - 1), SAS, surfA);
    }
    for (const fa_cf_result& a : acc) {
        res.com += a.com;
        res.wal += a.wal;
        res.sas += a.sas;
        res.con += a.con;
        res.totsas += a.totsas;
        res.rclash |= a.rclash;
    }
    res.cf = res.com + res.wal + res.sas + res.con;
    res.app = res.com + res.wal + res.sas;
    res.status = FA_OK;
    if (recs_out)  // atom by atom in calc order, as the ca_index lists are walked
        for (int i : S.scorable)
            for (const fa_contact_rec& r : W.records[S.slot[i]]) recs_out->push_back(r);
    return res;
}

template <class R> void init_work(const System& S, Work<R>& W)
{
    const int N = S.N(), Rn = S.R(), L = S.L();
    W.x.resize(N);
    W.y.resize(N);
    W.z.resize(N);
    W.rad.resize(N);
    W.type.resize(N);
    for (int i = 0; i < Rn; ++i) {
        W.x[i] = S.rec->xyz[3 * i];
        W.y[i] = S.rec->xyz[3 * i + 1];
        W.z[i] = S.rec->xyz[3 * i + 2];
        W.rad[i] = S.rec->radius[i];
        W.type[i] = S.rec->type[i];
    }
    for (int i = 0; i < L; ++i) {
        W.rad[Rn + i] = S.lig->radius[i];
        W.type[Rn + i] = S.lig->type[i];
    }
    W.boxnum.assign(N, -1);
    W.calclist.assign(N, 0);
    W.done.assign(N, 'N');
    W.seed.assign(3 * N, -1);
    W.rank.assign(N, -1);
    W.stubs.assign(S.scorable.size(), {});
    W.records.assign(S.scorable.size(), {});
    W.cont.resize(256);
    W.poly.resize(4096);
    W.centerpt.resize(256);
    W.vedge.resize(4096);
    W.ptorder.resize(256);
}

template <class R>
int eval_batch(const fa_receptor_desc* rec, const fa_ligand_desc* lig, const fa_params* prm, int n_poses,
               const double* genes, fa_cf_result* out, int n_threads, fa_opcount* ops, CapStats* caps = nullptr,
               BranchStats* branch = nullptr, int* fallback = nullptr)
{
    if (!rec || !lig || !prm || !genes || !out || n_poses < 0) return -1;
    System S{rec, lig, prm};
    if (ops) std::memset(ops, 0, sizeof(*ops));
#ifdef _OPENMP
    if (n_threads <= 0) n_threads = omp_get_max_threads();
#pragma omp parallel num_threads(n_threads)
#endif
    {
        Work<R> W;
        init_work(S, W);
        fa_opcount local{};
        CapStats lcaps{};
        BranchStats lbr{};
        t_ops = ops ? &local : nullptr;
        t_caps = caps ? &lcaps : nullptr;
        t_branch = branch ? &lbr : nullptr;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 4)
#endif
        for (int p = 0; p < n_poses; ++p) {
            out[p] = eval_pose(S, W, genes + (size_t)p * lig->n_genes, nullptr, nullptr);
            if (fallback) fallback[p] = t_fallback;
        }
        t_ops = nullptr;
        t_caps = nullptr;
        t_branch = nullptr;
        if (branch) {
#ifdef _OPENMP
#pragma omp critical
#endif
            {
                branch->epi1 += lbr.epi1;
                branch->epi2 += lbr.epi2;
                branch->epi3 += lbr.epi3;
                branch->atoms += lbr.atoms;
                branch->hull_fail += lbr.hull_fail;
            }
        }
        if (caps) {
#ifdef _OPENMP
#pragma omp critical
#endif
            {
                int* d = &caps->near_pose;
                const int* s0 = &lcaps.near_pose;
                for (int k = 0; k < 6; ++k) d[k] = std::max(d[k], s0[k]);
            }
        }
        if (ops) {
#ifdef _OPENMP
#pragma omp critical
#endif
            {
                uint64_t* d = &ops->cull;
                const uint64_t* s0 = &local.cull;
                for (int k = 0; k < FA_OPCOUNT_N; ++k) d[k] += s0[k];
            }
        }
    }
    (void)n_threads;
    return 0;
}

}  // namespace

// ======================================================================
// C ABI of the oracle (consumed by tests/ and bench.py's CPU baseline)
// ======================================================================
extern "C" {

int fa_oracle_eval(const fa_receptor_desc* rec, const fa_ligand_desc* lig, const fa_params* prm, int n_poses,
                   const double* genes, fa_cf_result* out, int n_threads)
{
    return eval_batch<double>(rec, lig, prm, n_poses, genes, out, n_threads, nullptr);
}

/* As fa_oracle_eval, also summing the operation counters over the batch. */
int fa_oracle_eval_counted(const fa_receptor_desc* rec, const fa_ligand_desc* lig, const fa_params* prm,
                           int n_poses, const double* genes, fa_cf_result* out, int n_threads, fa_opcount* ops)
{
    return eval_batch<double>(rec, lig, prm, n_poses, genes, out, n_threads, ops);
}

int fa_oracle_eval_f32(const fa_receptor_desc* rec, const fa_ligand_desc* lig, const fa_params* prm, int n_poses,
                       const double* genes, fa_cf_result* out, int n_threads)
{
    return eval_batch<float>(rec, lig, prm, n_poses, genes, out, n_threads, nullptr);
}

/* Largest per-atom / per-pose working sizes over a batch (6 ints: in-range
 * candidates per pose, contacts, hull edges, vertices before the sphere cut,
 * vertices with opposite points, face-list slots). */
int fa_oracle_cap_stats(const fa_receptor_desc* rec, const fa_ligand_desc* lig, const fa_params* prm, int n_poses,
                        const double* genes, fa_cf_result* out, int n_threads, int* stats6)
{
    CapStats c{};
    const int rc = eval_batch<double>(rec, lig, prm, n_poses, genes, out, n_threads, nullptr, &c);
    const int v[6] = {c.near_pose, c.nc, c.edges, c.vn, c.nv_opp, c.face_slots};
    for (int k = 0; k < 6; ++k) stats6[k] = v[k];
    return rc;
}

/* Counts of the calc_areas branches a batch reaches (5 values: atoms with one,
 * two, three or more engulfing planes, atoms with a polyhedron, poses whose
 * hull walk failed). */
int fa_oracle_branch_stats(const fa_receptor_desc* rec, const fa_ligand_desc* lig, const fa_params* prm, int n_poses,
                           const double* genes, fa_cf_result* out, int n_threads, long long* stats5)
{
    BranchStats b{};
    const int rc = eval_batch<double>(rec, lig, prm, n_poses, genes, out, n_threads, nullptr, nullptr, &b);
    const long long v[5] = {b.epi1, b.epi2, b.epi3, b.atoms, b.hull_fail};
    for (int k = 0; k < 5; ++k) stats5[k] = v[k];
    return rc;
}

/* 0: reference semantics (done/stub rule, hull seeds, calc order in a cell);
 * 1: independent atoms (no done/stub rule, no seeds); 2: the speculative
 * form (independent atoms from ray-cast starts, poses its checks flag in the
 * reference semantics). */
void fa_oracle_set_mode(int mode)
{
    g_eval_mode = mode == 1 ? 1 : 0;
    g_mode_bits = mode == 1 ? 7 : mode == 2 ? 8 : 0;
}

/* fa_oracle_eval in the current mode, with per pose 1 where the speculative
 * form handed the pose to the serial form. */
int fa_oracle_eval_fallback(const fa_receptor_desc* rec, const fa_ligand_desc* lig, const fa_params* prm, int n_poses,
                            const double* genes, fa_cf_result* out, int n_threads, int* fallback)
{
    return eval_batch<double>(rec, lig, prm, n_poses, genes, out, n_threads, nullptr, nullptr, nullptr, fallback);
}

/* The parts of the independent form one by one (bits as g_mode_bits). */
void fa_oracle_set_mode_bits(int bits)
{
    g_mode_bits = bits & 15;
    g_eval_mode = g_mode_bits == 7 ? 1 : 0;
}

/* 0: pose-build trig from the host libm (the reference binary's bits);
 * 1: correctly rounded float cos/sin (the GPU kernel's bits). */
void fa_oracle_set_trig(int mode) { g_trig_mode = mode ? 1 : 0; }

/* One pose with its Cartesian ligand coordinates and contact records.
 * Returns the number of records (capped at `cap`) or -1. */
int fa_oracle_eval_pose(const fa_receptor_desc* rec, const fa_ligand_desc* lig, const fa_params* prm,
                        const double* genes, fa_cf_result* out, float* lig_xyz, fa_contact_rec* recs, int cap)
{
    if (!rec || !lig || !prm || !genes || !out) return -1;
    System S{rec, lig, prm};
    Work<double> W;
    init_work(S, W);
    std::vector<fa_contact_rec> rv;
    *out = eval_pose(S, W, genes, lig_xyz, &rv);
    const int n = (int)rv.size();
    if (recs)
        for (int i = 0; i < n && i < cap; ++i) recs[i] = rv[i];
    return n;
}

/* Ligand pose build only (gene decode + buildcc). */
int fa_oracle_build_pose(const fa_ligand_desc* lig, const double* genes, float* lig_xyz)
{
    if (!lig || !genes || !lig_xyz) return -1;
    decode_and_build(lig, genes, lig_xyz);
    return 0;
}

/* Geometry helpers exposed for known-answer tests. */
int fa_oracle_solve_3x3(const double* e0, const double* e1, const double* e2, double* pt)
{
    return solve_3x3<double>(e0, e1, e2, pt);
}
int fa_oracle_solve_2xS(const double* a0, const double* a1, float rado, double* p0, double* p1)
{
    return solve_2xS<double>(a0, a1, rado, p0, p1);
}
double fa_oracle_cosPQR(const double* P, const double* Q, const double* Rr) { return cosPQR<double>(P, Q, Rr); }

/* the scored value of type pair k (get_yval + contribution) */
double fa_oracle_pair_term(const fa_params* prm, int k, double area, double surf)
{
    return pair_term(prm, k, area, surf);
}
double fa_oracle_spherical_arc(const double* A, double Ad, const double* B, double Bd, const double* C, double Cd,
                               float rado)
{
    Vert<double> a{{A[0], A[1], A[2]}, Ad, {0, 0, 0}}, b{{B[0], B[1], B[2]}, Bd, {0, 0, 0}},
        c{{C[0], C[1], C[2]}, Cd, {0, 0, 0}};
    return spherical_arc(a, b, c, rado);
}

}  // extern "C"
