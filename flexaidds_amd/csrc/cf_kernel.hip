// cf_kernel.hip — Voronoi contact-function population evaluation for gfx950.
//
// One 64-lane wavefront evaluates one pose: gene decode + Cartesian rebuild
// (LIB/ic2cf.cpp:53-123, LIB/buildcc.cpp:6-134), clash pre-pass and contact
// polyhedra (LIB/Vcontacts.cpp:6-1839), energy sum (LIB/vcfunction.cpp:14-574).
//
// Numerics: the pose build is FP32 operation-for-operation like the reference;
// every Vcontacts quantity is IEEE FP64 with no FMA contraction (the file is
// compiled with -ffp-contract=off), so contact sets and areas match the CPU
// path bit-for-bit except where a transcendental differs in its last ulp.
// The contact function has discontinuities (hull topology, arc/test-point
// decisions) that turn tiny arithmetic differences into kcal/mol changes, so
// FP32 geometry is not an option.
//
// Ligand atoms of a pose are processed in the reference's box order because
// the `done`/stub contact rule and the hull seeds carry state from one atom to
// the next (LIB/Vcontacts.cpp:1750-1836, :595-629).  Inside an atom, lanes work
// on candidates, planes, edges, vertices and faces.
#include <hip/hip_runtime.h>

#include "fa_device.h"

namespace fa {

constexpr int WAVE = 64;
constexpr int MAXC = 160;          // contacts per atom (planes)
constexpr int MAXP = MAXC + 4;     // + 4 bounding planes
constexpr int MAXV = 448;          // polyhedron vertices incl. arc and opposite points
constexpr int MAXE = 256;          // hull edges (the walk aborts near 200)
constexpr int MAXF = 1280;         // face point-list entries
constexpr int MAXL = FA_MAX_LIG_ATOMS;

constexpr double PI_REF = 3.141592;     // LIB/flexaid.h:86
constexpr float RW = 1.4f;              // LIB/flexaid.h:88
constexpr float KWALL_F = 1.0e6f;       // LIB/flexaid.h:81
constexpr float POLY_PENALTY = 1e6f;    // LIB/flexaid.h:69
constexpr double CLASH_THRESHOLD = 1e4; // LIB/flexaid.h:70
constexpr float CELLSIZE = 6.5f;        // LIB/Vcontacts.h:56
constexpr int EDGE_LIMIT = 200;         // LIB/Vcontacts.cpp:378
constexpr double BIG = 9.9e+9;          // sentinel of the reference scans

struct Smem {
    // planes / contacts (index c < NC are contacts, NC..NC+3 bounding planes)
    double A0[MAXP], A1[MAXP], A2[MAXP], A3[MAXP];
    double pdist[MAXP];
    double area[MAXP];
    double cdist[MAXC];
    int pidx[MAXP];
    char pflag[MAXP];
    // vertices
    double vx[MAXV], vy[MAXV], vz[MAXV], vd[MAXV];
    short vp0[MAXV], vp1[MAXV], vp2[MAXV];
    // edges
    double eV0[MAXE], eV1[MAXE], eV2[MAXE];
    short es[MAXE], ee[MAXE], ep0[MAXE], ep1[MAXE], esp[MAXE], eep[MAXE];
    char earc[MAXE];
    // faces (per plane point lists, CSR)
    short foff[MAXC + 1];
    short fcnt[MAXC];
    short fpt[MAXF];
    double fcos[MAXF];
    // ligand state
    float lx[MAXL], ly[MAXL], lz[MAXL];
    float ldis[MAXL], lang[MAXL], ldih[MAXL];
    int lbox[MAXL];
    unsigned char order[MAXL];
    int seed[MAXL * 3];
    unsigned long long stub[MAXL * 2];
    unsigned char done[MAXL];
    double acc_com[MAXL], acc_wal[MAXL], acc_sas[MAXL], acc_tot[MAXL];
    unsigned char acc_rcl[MAXL];
    // operation counters of the current pose (fa_opcount order)
    unsigned long long ops[FA_OPCOUNT_N];
};

// ------------------------------------------------------------ wave helpers
__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ int lanes_below(unsigned long long m)
{
    return __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

__device__ __forceinline__ int first_lane(unsigned long long m) { return __ffsll((long long)m) - 1; }

__device__ __forceinline__ int wave_sum_i(int v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

__device__ __forceinline__ double wave_sum_d(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

__device__ __forceinline__ int wave_excl_scan(int v)
{
    int x = v;
    const int l = lane_id();
    for (int off = 1; off < WAVE; off <<= 1) {
        int y = __shfl_up(x, off);
        if (l >= off) x += y;
    }
    return x - v;
}

// argmin over (value, index), ties to the lowest index: the result of a
// sequential scan that replaces only on a strictly smaller value
__device__ __forceinline__ void wave_argmin(double& v, int& i)
{
    for (int off = 32; off > 0; off >>= 1) {
        double ov = __shfl_xor(v, off);
        int oi = __shfl_xor(i, off);
        if (ov < v || (ov == v && oi < i)) {
            v = ov;
            i = oi;
        }
    }
}

__device__ __forceinline__ void wave_sync() { __syncthreads(); }

enum OpIdx { OP_CULL = 0, OP_PLANE, OP_HULL, OP_ARC, OP_ORDER, OP_TRI, OP_SARC, OP_VC, OP_BUILD };

// optional per-phase cycle accounting (debug builds of a launch only)
enum Phase { PH_BUILD = 0, PH_PREPASS, PH_CULL, PH_FIRST, PH_HULL, PH_ARC, PH_SEED, PH_FACE, PH_AREA, PH_REC, PH_N };

struct PhaseClock {
    unsigned long long* acc;  // null when disabled
    unsigned long long t;
    __device__ void start() { if (acc) t = __builtin_readcyclecounter(); }
    __device__ void stop(int ph)
    {
        if (acc) {
            const unsigned long long now = __builtin_readcyclecounter();
            if (lane_id() == 0) atomicAdd(&acc[ph], now - t);
            t = now;
        }
    }
};

// add a wave-uniform count to the pose's operation counter (lane 0 writes)
__device__ __forceinline__ void count_op(unsigned long long* ops, int k, long long n)
{
    if (lane_id() == 0) ops[k] += (unsigned long long)n;
}

// ------------------------------------------------------------ geometry (FP64)
// LIB/Vcontacts.cpp:1384-1413
__device__ __forceinline__ int solve_3x3(const double* e0, const double* e1, const double* e2, double* pt)
{
    double c00 = e1[1] * e2[2] - e2[1] * e1[2];
    double c01 = -e1[0] * e2[2] + e2[0] * e1[2];
    double c02 = e1[0] * e2[1] - e2[0] * e1[1];
    double c10 = -e0[1] * e2[2] + e2[1] * e0[2];
    double c11 = e0[0] * e2[2] - e2[0] * e0[2];
    double c12 = -e0[0] * e2[1] + e2[0] * e0[1];
    double c20 = e0[1] * e1[2] - e1[1] * e0[2];
    double c21 = -e0[0] * e1[2] + e1[0] * e0[2];
    double c22 = e0[0] * e1[1] - e1[0] * e0[1];
    double det = e0[0] * c00 + e0[1] * c01 + e0[2] * c02;
    if (det == 0.0) return -1;
    pt[0] = -(e0[3] * c00 + e1[3] * c10 + e2[3] * c20) / det;
    pt[1] = -(e0[3] * c01 + e1[3] * c11 + e2[3] * c21) / det;
    pt[2] = -(e0[3] * c02 + e1[3] * c12 + e2[3] * c22) / det;
    return 0;
}

// LIB/Vcontacts.cpp:1424-1476
__device__ __forceinline__ int solve_2xS(const double* a0, const double* a1, float rado, double* p0, double* p1)
{
    double e2[3];
    e2[0] = a0[1] * a1[2] - a0[2] * a1[1];
    e2[1] = a0[2] * a1[0] - a0[0] * a1[2];
    e2[2] = a0[0] * a1[1] - a0[1] * a1[0];
    double c00 = a1[1] * e2[2] - e2[1] * a1[2];
    double c01 = -a1[0] * e2[2] + e2[0] * a1[2];
    double c02 = a1[0] * e2[1] - e2[0] * a1[1];
    double c10 = -a0[1] * e2[2] + e2[1] * a0[2];
    double c11 = a0[0] * e2[2] - e2[0] * a0[2];
    double c12 = -a0[0] * e2[1] + e2[0] * a0[1];
    double det = e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2];
    if (det == 0.0) return -1;
    double av0 = -(a0[3] * c00 + a1[3] * c10) / det;
    double av1 = -(a0[3] * c01 + a1[3] * c11) / det;
    double av2 = -(a0[3] * c02 + a1[3] * c12) / det;
    double t = ((double)(rado * rado) - av0 * av0 - av1 * av1 - av2 * av2) / det;
    if (t < 0.0) return -1;
    t = sqrt(t);
    p0[0] = av0 + t * e2[0];
    p1[0] = av0 - t * e2[0];
    p0[1] = av1 + t * e2[1];
    p1[1] = av1 - t * e2[1];
    p0[2] = av2 + t * e2[2];
    p1[2] = av2 - t * e2[2];
    return 0;
}

// LIB/Vcontacts.cpp:1486-1505
__device__ __forceinline__ double cosPQR(const double* P, const double* Q, const double* R)
{
    double a0 = P[0] - Q[0], a1 = P[1] - Q[1], a2 = P[2] - Q[2];
    double b0 = R[0] - Q[0], b1 = R[1] - Q[1], b2 = R[2] - Q[2];
    return (a0 * b0 + a1 * b1 + a2 * b2) / sqrt((a0 * a0 + a1 * a1 + a2 * a2) * (b0 * b0 + b1 * b1 + b2 * b2));
}

// LIB/Vcontacts.cpp:1518-1557
__device__ double spherical_arc(const double* A, double Ad, const double* B, double Bd, const double* C, double Cd,
                                float rado)
{
    double cAOB = (A[0] * B[0] + A[1] * B[1] + A[2] * B[2]) / (Ad * Bd);
    double cBOC = (B[0] * C[0] + B[1] * C[1] + B[2] * C[2]) / (Bd * Cd);
    double U = (1 + cAOB) * sqrt((1 + cBOC) / 8.0);
    double V = (1 - cAOB) * sqrt((1 + cBOC) / 8.0);
    double W = sqrt((1 - cAOB) * (1 + cAOB) * (1 - cBOC) / 8.0);
    double tS = (1 - U + V + W + W) / (1 + U - V - W - W);
    double tSB = (1 - U - V) / (1 + U + V);
    double tSA = (1 - U + V - W - W) / (1 + U - V + W + W);
    double tanprod;
    if ((tS * tSA) > 0.0)
        tanprod = sqrt(tSB * tSB * tS * tSA);
    else
        tanprod = 0.0;
    double cBAC = cosPQR(B, A, C);
    double angBAC;
    if (cBAC > 1.0)
        angBAC = 0.0;
    else if (cBAC < -1.0)
        angBAC = PI_REF;
    else
        angBAC = acos(cBAC);
    double ro = (double)rado;
    return ro * (angBAC * (ro - Ad) - 4.0 * ro * atan(sqrt(tanprod)));
}

// LIB/Vcontacts.cpp:1357-1372, sequential over planes (one lane)
__device__ __forceinline__ bool test_point(const Smem& s, const double* p, int NC, int X)
{
    for (int cp = 0; cp < NC; ++cp) {
        if (cp != X && s.pflag[cp] != 'X') {
            if (s.A0[cp] * p[0] + s.A1[cp] * p[1] + s.A2[cp] * p[2] + s.A3[cp] > 0.0) return false;
        }
    }
    return true;
}

__device__ __forceinline__ void plane_of(const Smem& s, int c, double* a)
{
    a[0] = s.A0[c];
    a[1] = s.A1[c];
    a[2] = s.A2[c];
    a[3] = s.A3[c];
}

// LIB/Vcontacts.cpp:599-637 (all lanes compute, lane 0 stores)
__device__ void add_vedge(Smem& s, int k, int p0, int p1, int tp, int sp)
{
    double V0 = s.A1[p0] * s.A2[p1] - s.A2[p0] * s.A1[p1];
    double V1 = s.A2[p0] * s.A0[p1] - s.A0[p0] * s.A2[p1];
    double V2 = s.A0[p0] * s.A1[p1] - s.A1[p0] * s.A0[p1];
    double t0 = s.vx[sp] + V0, t1 = s.vy[sp] + V1, t2 = s.vz[sp] + V2;
    double tv = s.A0[tp] * t0 + s.A1[tp] * t1 + s.A2[tp] * t2 + s.A3[tp];
    if (tv > 0.0) {
        V0 = -V0;
        V1 = -V1;
        V2 = -V2;
    }
    if (lane_id() == 0) {
        s.eV0[k] = V0;
        s.eV1[k] = V1;
        s.eV2[k] = V2;
        s.es[k] = (short)sp;
        s.ee[k] = -1;
        s.ep0[k] = (short)p0;
        s.ep1[k] = (short)p1;
        s.esp[k] = (short)tp;
        s.eep[k] = -1;
        s.earc[k] = '.';
    }
}

__device__ __forceinline__ void put_vertex(Smem& s, int vn, double x, double y, double z, int a, int b, int c)
{
    if (lane_id() == 0) {
        s.vx[vn] = x;
        s.vy[vn] = y;
        s.vz[vn] = z;
        s.vd[vn] = sqrt(x * x + y * y + z * z);
        s.vp0[vn] = (short)a;
        s.vp1[vn] = (short)b;
        s.vp2[vn] = (short)c;
    }
}

// ------------------------------------------------------------ pose build
// LIB/ic2cf.cpp:53-105 then LIB/buildcc.cpp:6-134, lane 0 (FP32 like the
// reference; cos/sin are taken in double and rounded, which is what the host
// libm's correctly rounded cosf/sinf return)
__device__ void build_pose(Smem& s, const DevLigand& lig, const double* icv)
{
    const int n = lig.n_atoms;
    for (int i = lane_id(); i < n; i += WAVE) {
        s.ldis[i] = lig.dis[i];
        s.lang[i] = lig.ang[i];
        s.ldih[i] = lig.dih[i];
    }
    wave_sync();
    if (lane_id() == 0) {
        for (int g = 0; g < lig.n_genes; ++g) {
            double v = icv[g];
            if (v > lig.gene_max[g]) v = lig.gene_max[g];
            if (v < lig.gene_min[g]) v = lig.gene_min[g];
            const int atm = lig.gene_atm[g];
            const int typ = lig.gene_typ[g];
            if (typ == FA_GENE_GRID) {
                unsigned gi = (unsigned)v;
                if ((int)gi >= lig.n_grid) gi = lig.n_grid - 1;
                s.ldis[atm] = lig.grid_ic[3 * gi];
                s.lang[atm] = lig.grid_ic[3 * gi + 1];
                s.ldih[atm] = lig.grid_ic[3 * gi + 2];
            } else if (typ == FA_GENE_DIS) {
                s.ldis[atm] = (float)v;
            } else if (typ == FA_GENE_ANG) {
                s.lang[atm] = (float)v;
            } else if (typ == FA_GENE_DIH) {
                s.ldih[atm] = (float)v;
                int j = atm, cat = lig.rec[4 * j + 3];
                if (cat >= 0) {
                    for (int guard = 0; cat != atm && guard < n; ++guard) {
                        s.ldih[cat] = s.ldih[j] + lig.shift[cat];
                        j = cat;
                        cat = lig.rec[4 * j + 3];
                    }
                }
            }
        }
        for (int an = 0; an < n; ++an) {
            const int a = lig.mov[an];
            float x[4], y[4], z[4];
            for (int i = 1; i <= 3; ++i) {
                const int j = lig.rec[4 * a + i - 1];
                if (j >= 0) {
                    x[i] = s.lx[j];
                    y[i] = s.ly[j];
                    z[i] = s.lz[j];
                } else if (i == 1) {
                    x[i] = 1.0f + lig.ori[0];
                    y[i] = 0.0f + lig.ori[1];
                    z[i] = 0.0f + lig.ori[2];
                } else if (i == 3) {
                    x[i] = 0.0f + lig.ori[0];
                    y[i] = 1.0f + lig.ori[1];
                    z[i] = 0.0f + lig.ori[2];
                } else {
                    x[i] = 0.0f + lig.ori[0];
                    y[i] = 0.0f + lig.ori[1];
                    z[i] = 0.0f + lig.ori[2];
                }
                x[i] += 1e-10f;
                y[i] += 1e-10f;
                z[i] += 1e-10f;
            }
            float A = y[1] * (z[2] - z[3]) + y[2] * (z[3] - z[1]) + y[3] * (z[1] - z[2]);
            float B = z[1] * (x[2] - x[3]) + z[2] * (x[3] - x[1]) + z[3] * (x[1] - x[2]);
            float C = x[1] * (y[2] - y[3]) + x[2] * (y[3] - y[1]) + x[3] * (y[1] - y[2]);
            float op = sqrtf(A * A + B * B + C * C);
            float cx = A / op, cy = B / op, cz = C / op;
            A = x[2] - x[1];
            B = y[2] - y[1];
            C = z[2] - z[1];
            float d = 1.0f / sqrtf(A * A + B * B + C * C);
            op = s.ldis[a] * d;
            float xn = A * op, yn = B * op, zn = C * op;
            A = cx * cx;
            B = cy * cy;
            C = cz * cz;
            float angPI = (float)(s.lang[a] * PI_REF / 180.0f);
            float ct = (float)cos((double)angPI);
            float st = -(float)sin((double)angPI);
            op = 1.0f - ct;
            float xk = (cx * cz * op - cy * st) * zn + ((1.0f - A) * ct + A) * xn + (cx * cy * op + cz * st) * yn;
            float yk = (cy * cx * op - cz * st) * xn + ((1.0f - B) * ct + B) * yn + (cy * cz * op + cx * st) * zn;
            float zk = (cz * cy * op - cx * st) * yn + ((1.0f - C) * ct + C) * zn + (cz * cx * op + cy * st) * xn;
            float dihPI = (float)(s.ldih[a] * PI_REF / 180.0f);
            ct = (float)cos((double)dihPI);
            st = (float)sin((double)dihPI);
            op = 1.0f - ct;
            cx = (x[2] - x[1]) * d;
            cy = (y[2] - y[1]) * d;
            cz = (z[2] - z[1]) * d;
            A = cx * cx;
            B = cy * cy;
            C = cz * cz;
            s.lx[a] = (cx * cz * op - cy * st) * zk + ((1.0f - A) * ct + A) * xk + (cx * cy * op + cz * st) * yk + x[1];
            s.ly[a] = (cy * cx * op - cz * st) * xk + ((1.0f - B) * ct + B) * yk + (cy * cz * op + cx * st) * zk + y[1];
            s.lz[a] = (cz * cy * op - cx * st) * yk + ((1.0f - C) * ct + C) * zk + (cz * cx * op + cy * st) * xk + z[1];
        }
    }
    wave_sync();
}

// ------------------------------------------------------------ pose grid
struct PoseGrid {
    float gmin[3];
    int dim, dim2, dim3;
    bool slow;   // grid differs from the receptor's precomputed one
};

// LIB/Vcontacts.cpp:1594-1658 (steady state): only scorable atoms widen it
__device__ bool pose_grid(const Smem& s, const DevLigand& lig, const DevGrid& G, PoseGrid& pg)
{
    float gmin[3], gmax[3];
    for (int j = 0; j < 3; ++j) {
        gmin[j] = G.gmin[j];
        gmax[j] = G.gmax[j];
    }
    // sequential semantics: scan ligand atoms in calc order
    int alter = 0;
    for (int i = 0; i < lig.n_atoms; ++i) {
        const float c[3] = {s.lx[i], s.ly[i], s.lz[i]};
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
    float max_width = G.maxwidth;
    if (alter) {
        for (int j = 0; j < 3; ++j) {
            gmin[j] = (float)(gmin[j] < 0.0f ? (int)gmin[j] - 1.0 : (int)gmin[j] + 1.0);
            gmax[j] = (float)(gmax[j] < 0.0f ? (int)gmax[j] - 1.0 : (int)gmax[j] + 1.0);
            float diff = gmax[j] - gmin[j];
            if (diff > max_width) max_width = diff;
        }
    }
    pg.dim = (int)(max_width / CELLSIZE) + 1;
    pg.dim2 = pg.dim * pg.dim;
    pg.dim3 = pg.dim2 * pg.dim;
    for (int j = 0; j < 3; ++j) pg.gmin[j] = gmin[j];
    pg.slow = alter != 0;
    return true;
}

__device__ __forceinline__ int box_of(const PoseGrid& pg, float x, float y, float z)
{
    return (int)((x - pg.gmin[0]) / CELLSIZE) * pg.dim2 + (int)((y - pg.gmin[1]) / CELLSIZE) * pg.dim +
           (int)((z - pg.gmin[2]) / CELLSIZE);
}

// ------------------------------------------------------------ candidate scan
// Calls f(valid, calc_index, x, y, z, r) for every atom of the 27 boxes
// around box `bz`, in the reference's order (slot, then calc order in a box:
// receptor atoms first, then ligand atoms).  All lanes call f together.
template <class F>
__device__ void scan_boxes(const Smem& s, const DevReceptor& R, const DevGrid& G, const PoseGrid& pg, int L, int bz,
                           F&& f)
{
    const int l = lane_id();
    for (int slot = 0; slot < 27; ++slot) {
        const int bi = bz + pg.dim2 * ((slot / 9) - 1) + pg.dim * (((slot / 3) % 3) - 1) + (slot % 3) - 1;
        if (bi < 0 || bi >= pg.dim3) continue;
        if (!pg.slow) {
            const int b0 = G.cell_start[bi], b1 = G.cell_start[bi + 1];
            for (int base = b0; base < b1; base += WAVE) {
                const int k = base + l;
                const bool v = k < b1;
                float4 p = v ? G.cell_xyzr[k] : make_float4(0.f, 0.f, 0.f, 0.f);
                const int j = v ? G.cell_atom[k] : -1;
                f(v, j, p.x, p.y, p.z, p.w);
            }
        } else {
            for (int base = 0; base < R.n; base += WAVE) {
                const int j = base + l;
                bool v = j < R.n;
                float x = 0.f, y = 0.f, z = 0.f, r = 0.f;
                if (v) {
                    x = R.x[j];
                    y = R.y[j];
                    z = R.z[j];
                    r = R.rad[j];
                    v = box_of(pg, x, y, z) == bi;
                }
                if (__ballot(v) == 0ull) continue;
                f(v, j, x, y, z, r);
            }
        }
        for (int base = 0; base < L; base += WAVE) {
            const int m = base + l;
            const bool v = m < L && s.lbox[m] == bi;
            if (__ballot(v) == 0ull) continue;
            const int mm = v ? m : 0;
            f(v, R.n + m, s.lx[mm], s.ly[mm], s.lz[mm], 0.0f /* radius looked up by caller */);
        }
    }
}

// ------------------------------------------------------------ kernel
__global__ __launch_bounds__(WAVE) void cf_pose_kernel(DevReceptor R, const DevGrid* __restrict__ grids,
                                                        const DevLigand* __restrict__ ligs, DevParams prm, DevBatch B)
{
    __shared__ Smem s;
    const int l = lane_id();
    for (;;) {
        // dynamic work queue: pose costs vary a lot (clash exits, hull sizes)
        int pose = 0;
        if (l == 0) pose = (int)atomicAdd(&B.counters[2], 1u);
        pose = __shfl(pose, 0);
        if (pose >= B.n_poses) break;
        const int li = B.pose_lig ? B.pose_lig[pose] : 0;
        const DevLigand lig = ligs[li];
        const DevGrid G = grids[lig.grid_id];
        const int L = lig.n_atoms;
        const int Rn = R.n;
        fa_cf_result res;
        res.com = res.wal = res.sas = res.con = res.totsas = res.cf = res.app = 0.0;
        res.rclash = 0;
        res.status = FA_OK;
        int nrec_total = 0;

        if (l < FA_OPCOUNT_N) s.ops[l] = 0ull;
        wave_sync();
        PhaseClock clk{B.prof, 0ull};
        clk.start();
        build_pose(s, lig, B.genes + (size_t)pose * B.gene_stride);
        count_op(s.ops, OP_BUILD, L);
        if (B.lig_xyz) {
            for (int i = l; i < L; i += WAVE) {
                float* o = B.lig_xyz + ((size_t)pose * B.xyz_stride + i) * 3;
                o[0] = s.lx[i];
                o[1] = s.ly[i];
                o[2] = s.lz[i];
            }
        }
        PoseGrid pg;
        pose_grid(s, lig, G, pg);
        if (pg.slow && l == 0) atomicAdd(&B.counters[0], 1u);
        // ligand boxes, processing order (box, calc index)
        bool range_ok = true;
        for (int i = l; i < L; i += WAVE) {
            const int b = box_of(pg, s.lx[i], s.ly[i], s.lz[i]);
            s.lbox[i] = b;
            if (b < 0 || b >= pg.dim3) range_ok = false;
        }
        range_ok = __ballot(!range_ok) == 0ull;
        wave_sync();
        if (!range_ok) {
            res.status = FA_GRID_RANGE;
        } else {
            for (int i = l; i < L; i += WAVE) {
                const int bi = s.lbox[i];
                int rank = 0;
                for (int k = 0; k < L; ++k) {
                    const int bk = s.lbox[k];
                    rank += (bk < bi) || (bk == bi && k < i);
                }
                s.order[rank] = (unsigned char)i;
            }
            wave_sync();

            clk.stop(PH_BUILD);
            // ---- clash pre-pass (LIB/Vcontacts.cpp:36-51, :1803-1821): every
            // receptor contact and every ligand pair once (the earlier atom in
            // box order counts it), non-bonded, d < permeability * (rA + rB)
            double clash = 0.0;
            const double permea = (double)prm.permeability;
            for (int i = l; i < L; i += WAVE) s.done[i] = 0;
            wave_sync();
            for (int k = 0; k < L; ++k) {
                const int a = s.order[k];
                const float ax = s.lx[a], ay = s.ly[a], az = s.lz[a];
                const float ra = lig.radius[a];
                const float rado = ra + RW;
                int found = 0, tested = 0;
                scan_boxes(s, R, G, pg, L, s.lbox[a], [&](bool v, int j, float x, float y, float z, float r) {
                    if (!v) return;
                    const bool lj = j >= Rn;
                    const int m = j - Rn;
                    if (lj) {
                        if (s.done[m]) return;  // an earlier atom already counted the pair
                        r = lig.radius[m];
                    }
                    tested += (j != Rn + a);  // the atom itself never makes a contact
                    const float dx = ax - x, dy = ay - y, dz = az - z;
                    const double sq = (double)(dx * dx + dy * dy + dz * dz);
                    const double nd = (double)(rado + r + RW);
                    if (!(sq < nd * nd && sq != 0.0)) return;
                    ++found;
                    const double dist = sqrt(sq);
                    const double rAB = (double)(ra + r);
                    const double cd = permea * rAB;
                    if (dist < cd) {
                        bool bonded = false;
                        if (lj) bonded = (lig.bonded[2 * a + (m >> 6)] >> (m & 63)) & 1ull;
                        if (!lj || !bonded) clash += (double)KWALL_F * (pow(dist, -12.0) - pow(cd, -12.0));
                    }
                });
                found = wave_sum_i(found);
                count_op(s.ops, OP_CULL, wave_sum_i(tested));
                if (l == 0 && found > 0) s.done[a] = 1;  // marked done once it has a contact
                wave_sync();
            }
            clash = wave_sum_d(clash);
            if (clash >= CLASH_THRESHOLD) {
                res.wal = clash;
                res.cf = clash + 1.0;
                res.app = clash + 1.0;
                res.status = FA_CLASH;
            }
        }

        clk.stop(PH_PREPASS);
        if (res.status == FA_OK) {
            // ---- contact polyhedra in box order (calc_region, LIB/Vcontacts.cpp:74-139)
            for (int i = l; i < L; i += WAVE) {
                s.done[i] = 0;
                s.seed[3 * i] = -1;
                s.stub[2 * i] = 0ull;
                s.stub[2 * i + 1] = 0ull;
                s.acc_com[i] = 0.0;
                s.acc_wal[i] = 0.0;
                s.acc_sas[i] = 0.0;
                s.acc_tot[i] = 0.0;
                s.acc_rcl[i] = 0;
            }
            wave_sync();
            for (int k = 0; k < L && res.status == FA_OK; ++k) {
                const int a = s.order[k];
                const int acalc = Rn + a;
                const float ax = s.lx[a], ay = s.ly[a], az = s.lz[a];
                const float ra = lig.radius[a];
                const float rado = ra + RW;
                clk.stop(PH_REC);
                // -- contact list (get_contlist4, LIB/Vcontacts.cpp:1725-1839)
                int NC = 0, tested = 0;
                bool overflow = false;
                scan_boxes(s, R, G, pg, L, s.lbox[a], [&](bool v, int j, float x, float y, float z, float r) {
                    bool in = false;
                    double dist = 0.0;
                    if (v) {
                        const bool lj = j >= Rn;
                        const int m = j - Rn;
                        bool skip = false;
                        if (lj) {
                            r = lig.radius[m];
                            // done atoms are skipped unless they recorded a contact with a
                            skip = s.done[m] && !((s.stub[2 * a + (m >> 6)] >> (m & 63)) & 1ull);
                        }
                        if (!skip) {
                            tested += (j != acalc);
                            const float dx = ax - x, dy = ay - y, dz = az - z;
                            const double sq = (double)(dx * dx + dy * dy + dz * dz);
                            const double nd = (double)(rado + r + RW);
                            if (sq < nd * nd && sq != 0.0) {
                                in = true;
                                dist = sqrt(sq);
                            }
                        }
                    }
                    const unsigned long long m = __ballot(in);
                    if (in) {
                        const int pos = NC + lanes_below(m);
                        if (pos < MAXC) {
                            s.pidx[pos] = j;
                            s.cdist[pos] = dist;
                            // plane (voronoi_poly2, LIB/Vcontacts.cpp:181-220), 'X' extended radical
                            const float rj = RW + r;
                            const double pd = (dist * dist + (double)(rado * rado) - (double)(rj * rj)) / (2 * dist);
                            s.A0[pos] = (double)(x - ax) / dist;
                            s.A1[pos] = (double)(y - ay) / dist;
                            s.A2[pos] = (double)(z - az) / dist;
                            s.A3[pos] = -pd;
                            s.pdist[pos] = fabs(pd);
                            s.pflag[pos] = 'X';
                        }
                    }
                    NC += __popcll(m);
                });
                count_op(s.ops, OP_CULL, wave_sum_i(tested));
                if (NC > MAXC) overflow = true;
                if (overflow) {
                    res.status = FA_CAPACITY;
                    break;
                }
                if (NC == 0) continue;  // SAS = -1, atom not marked done
                count_op(s.ops, OP_PLANE, NC);
                // bounding planes (LIB/Vcontacts.cpp:223-241)
                if (l < 4) {
                    const double bp[4][4] = {{0.707, 1.0, 0.0, -10.0}, {0.707, -1.0, 0.0, -10.0},
                                             {-0.707, 0.0, 1.0, -10.0}, {-0.707, 0.0, -1.0, -10.0}};
                    s.A0[NC + l] = bp[l][0];
                    s.A1[NC + l] = bp[l][1];
                    s.A2[NC + l] = bp[l][2];
                    s.A3[NC + l] = bp[l][3];
                    s.pdist[NC + l] = 0.0;
                    s.pidx[NC + l] = -1;
                    s.pflag[NC + l] = 'X';
                }
                wave_sync();

                clk.stop(PH_CULL);
                // -- first vertex (get_firstvert, LIB/Vcontacts.cpp:499-590)
                int pA = -1, pB = -1, pC = -1;
                {
                    const int s0 = s.seed[3 * a];
                    if (s0 != -1) {
                        const int s1 = s.seed[3 * a + 1], s2 = s.seed[3 * a + 2];
                        int fa = -1, fb = -1, fc = -1;
                        for (int base = 0; base < NC; base += WAVE) {
                            const int c = base + l;
                            const int idx = c < NC ? s.pidx[c] : -2;
                            unsigned long long m0 = __ballot(c < NC && idx == s0);
                            unsigned long long m1 = __ballot(c < NC && idx != s0 && idx == s1);
                            unsigned long long m2 = __ballot(c < NC && idx != s0 && idx != s1 && idx == s2);
                            if (m0) fa = base + 63 - __clzll((long long)m0);
                            if (m1) fb = base + 63 - __clzll((long long)m1);
                            if (m2) fc = base + 63 - __clzll((long long)m2);
                        }
                        pA = fa;
                        pB = fb;
                        pC = fc;
                    }
                    if (!(pA != -1 && pB != -1 && pC != -1)) {
                        double bv = BIG;
                        int bi = 0x7fffffff;
                        for (int c = l; c < NC; c += WAVE) {
                            if (s.pdist[c] < bv) {
                                bv = s.pdist[c];
                                bi = c;
                            }
                        }
                        wave_argmin(bv, bi);
                        pA = bi;
                        double eA[4];
                        plane_of(s, pA, eA);
                        bv = BIG;
                        bi = 0x7fffffff;
                        double ptx = 0.0, pty = 0.0, ptz = 0.0;
                        for (int c = l; c < NC + 4; c += WAVE) {
                            if (c == pA) continue;
                            double eC[4], vA[4], tp[3];
                            plane_of(s, c, eC);
                            vA[0] = eA[1] * eC[2] - eA[2] * eC[1];
                            vA[1] = eA[2] * eC[0] - eA[0] * eC[2];
                            vA[2] = eA[0] * eC[1] - eA[1] * eC[0];
                            vA[3] = 0.0;
                            if (solve_3x3(eA, eC, vA, tp) != -1) {
                                double d = sqrt(tp[0] * tp[0] + tp[1] * tp[1] + tp[2] * tp[2]);
                                if (d < bv) {
                                    bv = d;
                                    bi = c;
                                    ptx = tp[0];
                                    pty = tp[1];
                                    ptz = tp[2];
                                }
                            }
                        }
                        {
                            double v0 = bv;
                            int i0 = bi;
                            wave_argmin(v0, i0);
                            const int src = first_lane(__ballot(bi == i0 && bv == v0));
                            ptx = __shfl(ptx, src);
                            pty = __shfl(pty, src);
                            ptz = __shfl(ptz, src);
                            pB = (i0 == 0x7fffffff) ? -1 : i0;
                            if (pB < 0) { ptx = pty = ptz = 0.0; }
                        }
                        double eB[4];
                        plane_of(s, pB < 0 ? 0 : pB, eB);
                        double vA0 = eA[1] * eB[2] - eA[2] * eB[1];
                        double vA1 = eA[2] * eB[0] - eA[0] * eB[2];
                        double vA2 = eA[0] * eB[1] - eA[1] * eB[0];
                        bv = BIG;
                        bi = 0x7fffffff;
                        for (int c = l; c < NC + 4; c += WAVE) {
                            if (c == pA || c == pB) continue;
                            double vtdiv = s.A0[c] * vA0 + s.A1[c] * vA1 + s.A2[c] * vA2;
                            if (vtdiv != 0.0) {
                                double vt = -(s.A0[c] * ptx + s.A1[c] * pty + s.A2[c] * ptz + s.A3[c]) / vtdiv;
                                if (fabs(vt) < bv) {
                                    bv = fabs(vt);
                                    bi = c;
                                }
                            }
                        }
                        wave_argmin(bv, bi);
                        pC = (bi == 0x7fffffff) ? -1 : bi;
                    }
                }
                if (pA < 0 || pB < 0 || pC < 0) {
                    res.status = FA_CAPACITY;
                    break;
                }
                clk.stop(PH_FIRST);
                // -- hull walk (LIB/Vcontacts.cpp:243-410)
                {
                    double eA[4], eB[4], eC[4], tp[3];
                    plane_of(s, pA, eA);
                    plane_of(s, pB, eB);
                    plane_of(s, pC, eC);
                    solve_3x3(eA, eB, eC, tp);
                    put_vertex(s, 0, tp[0], tp[1], tp[2], pA, pB, pC);
                    if (l == 0) {
                        s.pflag[pA] = 'Y';
                        s.pflag[pB] = 'Y';
                        s.pflag[pC] = 'Y';
                    }
                    wave_sync();
                    add_vedge(s, 0, pA, pB, pC, 0);
                    add_vedge(s, 1, pB, pC, pA, 0);
                    add_vedge(s, 2, pC, pA, pB, 0);
                    wave_sync();
                }
                int startedge = 0, edgenum = 3, vn = 1;
                bool hull_fail = false, cap = false;
                const double rr = (double)(rado * rado);
                while (true) {
                    while (s.ee[startedge] >= 0 && (edgenum - startedge) > 0) ++startedge;
                    if ((edgenum - startedge) <= 0) break;
                    const int sp = s.es[startedge];
                    const double sx = s.vx[sp], sy = s.vy[sp], sz = s.vz[sp];
                    const double V0 = s.eV0[startedge], V1 = s.eV1[startedge], V2 = s.eV2[startedge];
                    const int A = s.ep0[startedge], Bp = s.ep1[startedge], oldC = s.esp[startedge];
                    double bv = BIG;
                    int bi = 0x7fffffff;
                    for (int c = l; c < NC; c += WAVE) {
                        if (c != A && c != Bp && c != oldC) {
                            double vtdiv = s.A0[c] * V0 + s.A1[c] * V1 + s.A2[c] * V2;
                            if (vtdiv != 0.0) {
                                double vt = -(s.A0[c] * sx + s.A1[c] * sy + s.A2[c] * sz + s.A3[c]) / vtdiv;
                                if (vt < bv && vt > 0.0) {
                                    bv = vt;
                                    bi = c;
                                }
                            }
                        }
                    }
                    wave_argmin(bv, bi);
                    count_op(s.ops, OP_HULL, NC - (A < NC) - (Bp < NC) - (oldC < NC));
                    double vtmin = bv;
                    int C = (bi == 0x7fffffff) ? -1 : bi;
                    double nx = sx + vtmin * V0, ny = sy + vtmin * V1, nz = sz + vtmin * V2;
                    if (nx * nx + ny * ny + nz * nz > rr) {
                        count_op(s.ops, OP_HULL, 4 - (A >= NC) - (Bp >= NC) - (oldC >= NC));
                        for (int c = NC; c < NC + 4; ++c) {
                            if (c != A && c != Bp && c != oldC) {
                                double vtdiv = s.A0[c] * V0 + s.A1[c] * V1 + s.A2[c] * V2;
                                if (vtdiv != 0.0) {
                                    double vt = -(s.A0[c] * sx + s.A1[c] * sy + s.A2[c] * sz + s.A3[c]) / vtdiv;
                                    if (vt < vtmin && vt > 0.0) {
                                        vtmin = vt;
                                        C = c;
                                    }
                                }
                            }
                        }
                        nx = sx + vtmin * V0;
                        ny = sy + vtmin * V1;
                        nz = sz + vtmin * V2;
                    }
                    if (C < 0 || vn + 1 >= MAXV || edgenum + 2 >= MAXE) {
                        cap = true;
                        break;
                    }
                    put_vertex(s, vn, nx, ny, nz, A, Bp, C);
                    if (l == 0) {
                        s.ee[startedge] = (short)vn;
                        s.eep[startedge] = (short)C;
                        s.pflag[C] = 'Y';
                    }
                    wave_sync();
                    // edges (A, C) then (B, C)
                    for (int pass = 0; pass < 2; ++pass) {
                        const int P = pass ? Bp : A;
                        const int other = pass ? A : Bp;
                        int found = -1;
                        for (int base = startedge + 1; base < edgenum && found < 0; base += WAVE) {
                            const int e = base + l;
                            bool hit = false;
                            if (e < edgenum) {
                                const int q0 = s.ep0[e], q1 = s.ep1[e];
                                hit = (q0 == P && q1 == C) || (q0 == C && q1 == P);
                            }
                            const unsigned long long m = __ballot(hit);
                            if (m) found = base + first_lane(m);
                        }
                        if (found >= 0) {
                            if (l == 0) {
                                s.ee[found] = (short)vn;
                                s.eep[found] = (short)other;
                            }
                            wave_sync();
                        } else {
                            add_vedge(s, edgenum, P, C, other, vn);
                            ++edgenum;
                            wave_sync();
                            if (pass == 1 && edgenum >= EDGE_LIMIT) {
                                hull_fail = true;
                                break;
                            }
                        }
                    }
                    if (hull_fail) break;
                    ++vn;
                }
                if (cap) {
                    res.status = FA_CAPACITY;
                    break;
                }
                if (hull_fail) {
                    res.status = FA_HULL_FAIL;
                    res.wal = (double)POLY_PENALTY;
                    res.cf = res.wal + 1.0;
                    res.app = res.wal + 1.0;
                    break;
                }
                clk.stop(PH_HULL);
                // -- arc points (LIB/Vcontacts.cpp:419-457), in edge order
                for (int base = 0; base < edgenum; base += WAVE) {
                    const int e = base + l;
                    int cnt = 0;
                    double q0[3] = {0, 0, 0}, q1[3] = {0, 0, 0};
                    bool k0 = false, k1 = false;
                    int e0 = 0, e1 = 0, solved = 0;
                    if (e < edgenum) {
                        const int st = s.es[e], en = s.ee[e];
                        const double rd = (double)rado;
                        if (rd < s.vd[st] || rd < s.vd[en]) {
                            solved = 1;
                            double a0[4], a1[4];
                            e0 = s.ep0[e];
                            e1 = s.ep1[e];
                            plane_of(s, e0, a0);
                            plane_of(s, e1, a1);
                            if (solve_2xS(a0, a1, rado, q0, q1) != -1) {
                                const int ta = s.esp[e], tb = s.eep[e];
                                double va = s.A0[ta] * q0[0] + s.A1[ta] * q0[1] + s.A2[ta] * q0[2] + s.A3[ta];
                                double vb = s.A0[tb] * q0[0] + s.A1[tb] * q0[1] + s.A2[tb] * q0[2] + s.A3[tb];
                                k0 = (va < 0.0 && vb < 0.0);
                                va = s.A0[ta] * q1[0] + s.A1[ta] * q1[1] + s.A2[ta] * q1[2] + s.A3[ta];
                                vb = s.A0[tb] * q1[0] + s.A1[tb] * q1[1] + s.A2[tb] * q1[2] + s.A3[tb];
                                k1 = (va < 0.0 && vb < 0.0);
                            }
                        }
                        cnt = (int)k0 + (int)k1;
                    }
                    const int off = wave_excl_scan(cnt);
                    const int tot = wave_sum_i(cnt);
                    count_op(s.ops, OP_ARC, wave_sum_i(solved));
                    if (vn + tot >= MAXV) {
                        cap = true;
                    } else {
                        int pos = vn + off;
                        const double rd = (double)rado;
                        if (k0) {
                            s.vx[pos] = q0[0];
                            s.vy[pos] = q0[1];
                            s.vz[pos] = q0[2];
                            s.vd[pos] = rd;
                            s.vp0[pos] = (short)e0;
                            s.vp1[pos] = (short)e1;
                            s.vp2[pos] = -1;
                            ++pos;
                        }
                        if (k1) {
                            s.vx[pos] = q1[0];
                            s.vy[pos] = q1[1];
                            s.vz[pos] = q1[2];
                            s.vd[pos] = rd;
                            s.vp0[pos] = (short)e0;
                            s.vp1[pos] = (short)e1;
                            s.vp2[pos] = -1;
                        }
                    }
                    vn += tot;
                    wave_sync();
                    if (cap) break;
                }
                if (cap) {
                    res.status = FA_CAPACITY;
                    break;
                }
                // -- keep vertices inside the sphere (LIB/Vcontacts.cpp:460-467)
                {
                    int vj = 0;
                    const double rd = (double)rado;
                    for (int base = 0; base < vn; base += WAVE) {
                        const int vi = base + l;
                        const bool v = vi < vn;
                        double x = 0, y = 0, z = 0, d = 0;
                        short p0 = 0, p1 = 0, p2 = 0;
                        if (v) {
                            x = s.vx[vi];
                            y = s.vy[vi];
                            z = s.vz[vi];
                            d = s.vd[vi];
                            p0 = s.vp0[vi];
                            p1 = s.vp1[vi];
                            p2 = s.vp2[vi];
                        }
                        const bool keep = v && d <= rd;
                        const unsigned long long m = __ballot(keep);
                        wave_sync();
                        if (keep) {
                            const int pos = vj + lanes_below(m);
                            s.vx[pos] = x;
                            s.vy[pos] = y;
                            s.vz[pos] = z;
                            s.vd[pos] = d;
                            s.vp0[pos] = p0;
                            s.vp1[pos] = p1;
                            s.vp2[pos] = p2;
                        }
                        vj += __popcll(m);
                        wave_sync();
                    }
                    vn = vj;
                }
                // -- engulfing flags (LIB/Vcontacts.cpp:470-478)
                for (int c = l; c < NC; c += WAVE)
                    if (s.A3[c] > 0.0) s.pflag[c] = 'E';
                wave_sync();
                clk.stop(PH_ARC);
                // -- seeds for ligand atoms not seeded yet (LIB/Vcontacts.cpp:595-629)
                for (int m = l; m < L; m += WAVE) {
                    if (s.seed[3 * m] != -1) continue;
                    const int mc = Rn + m;
                    for (int vi = 0; vi < vn; ++vi) {
                        if (s.vp2[vi] == -1) continue;
                        const int i0 = s.pidx[s.vp0[vi]], i1 = s.pidx[s.vp1[vi]], i2 = s.pidx[s.vp2[vi]];
                        if (i0 == mc) {
                            s.seed[3 * m] = acalc;
                            s.seed[3 * m + 1] = i1;
                            s.seed[3 * m + 2] = i2;
                            break;
                        } else if (i1 == mc) {
                            s.seed[3 * m] = acalc;
                            s.seed[3 * m + 1] = i0;
                            s.seed[3 * m + 2] = i2;
                            break;
                        } else if (i2 == mc) {
                            s.seed[3 * m] = acalc;
                            s.seed[3 * m + 1] = i0;
                            s.seed[3 * m + 2] = i1;
                            break;
                        }
                    }
                }
                wave_sync();
                const int NV = vn;
                clk.stop(PH_SEED);
                // -- face point lists (order_faces, LIB/Vcontacts.cpp:699-831)
                {
                    // count points per face (+1 slot for an opposite point)
                    int off_total = 0;
                    for (int base = 0; base < NC; base += WAVE) {
                        const int X = base + l;
                        int cnt = 0;
                        if (X < NC && s.pflag[X] != 'X') {
                            for (int vi = 0; vi < NV; ++vi)
                                cnt += (s.vp0[vi] == X) | (s.vp1[vi] == X) | (s.vp2[vi] == X);
                        }
                        const int slots = cnt + (cnt > 2 ? 1 : 0);
                        const int off = wave_excl_scan(slots);
                        if (X < NC) {
                            s.foff[X] = (short)(off_total + off);
                            s.fcnt[X] = (short)cnt;
                        }
                        off_total += wave_sum_i(slots);
                    }
                    if (off_total > MAXF) {
                        res.status = FA_CAPACITY;
                        break;
                    }
                    wave_sync();
                    for (int X = l; X < NC; X += WAVE) {
                        if (s.pflag[X] == 'X') continue;
                        int o = s.foff[X];
                        for (int vi = 0; vi < NV; ++vi)
                            if (s.vp0[vi] == X || s.vp1[vi] == X || s.vp2[vi] == X) s.fpt[o++] = (short)vi;
                    }
                    wave_sync();
                    // opposite point of the first arc point, appended in plane order
                    int nv_extra = NV;
                    for (int base = 0; base < NC; base += WAVE) {
                        const int X = base + l;
                        bool add = false;
                        double t[3] = {0, 0, 0};
                        if (X < NC && s.pflag[X] != 'X' && s.fcnt[X] > 2) {
                            const int o = s.foff[X], sc = s.fcnt[X];
                            const double cx = -s.A0[X] * s.A3[X], cy = -s.A1[X] * s.A3[X], cz = -s.A2[X] * s.A3[X];
                            for (int q = 0; q < sc; ++q) {
                                const int vi = s.fpt[o + q];
                                if (s.vp2[vi] == -1) {
                                    t[0] = 2.0 * cx - s.vx[vi];
                                    t[1] = 2.0 * cy - s.vy[vi];
                                    t[2] = 2.0 * cz - s.vz[vi];
                                    add = test_point(s, t, NC, X);
                                    break;
                                }
                            }
                        }
                        const unsigned long long m = __ballot(add);
                        if (nv_extra + __popcll(m) > MAXV) {
                            cap = true;
                            break;
                        }
                        if (add) {
                            const int vi = nv_extra + lanes_below(m);
                            s.vx[vi] = t[0];
                            s.vy[vi] = t[1];
                            s.vz[vi] = t[2];
                            s.vd[vi] = (double)rado;
                            s.vp0[vi] = (short)X;
                            s.vp1[vi] = -1;
                            s.vp2[vi] = -1;
                            s.fpt[s.foff[X] + s.fcnt[X]] = (short)vi;
                            s.fcnt[X] = s.fcnt[X] + 1;
                        }
                        nv_extra += __popcll(m);
                    }
                    if (cap) {
                        res.status = FA_CAPACITY;
                        break;
                    }
                    wave_sync();
                    // order points around each face by cosine
                    bool noline = false;
                    int ncos = 0;
                    for (int X = l; X < NC; X += WAVE) {
                        const int sc = s.fcnt[X];
                        if (s.pflag[X] == 'X' || sc <= 3) continue;
                        ncos += sc - 2;
                        short* pt = &s.fpt[s.foff[X]];
                        double* cs = &s.fcos[s.foff[X]];
                        const int v0 = pt[0];
                        const int Y = (s.vp0[v0] == X) ? s.vp1[v0] : s.vp0[v0];
                        int vi;
                        for (vi = 1; vi < sc; ++vi) {
                            const int w = pt[vi];
                            if (s.vp0[w] == Y || s.vp1[w] == Y || s.vp2[w] == Y) break;
                        }
                        if (vi == sc) {  // no second point on the common line (reference reads past its list)
                            noline = true;
                            continue;
                        }
                        short tsw = pt[vi];
                        pt[vi] = pt[1];
                        pt[1] = tsw;
                        const double P1[3] = {s.vx[pt[1]], s.vy[pt[1]], s.vz[pt[1]]};
                        const double P0[3] = {s.vx[pt[0]], s.vy[pt[0]], s.vz[pt[0]]};
                        for (vi = 2; vi < sc; ++vi) {
                            const double PX[3] = {s.vx[pt[vi]], s.vy[pt[vi]], s.vz[pt[vi]]};
                            cs[vi] = cosPQR(P1, P0, PX);
                        }
                        for (vi = 2; vi < sc - 1; ++vi)
                            for (int v2 = vi + 1; v2 < sc; ++v2)
                                if (cs[vi] < cs[v2]) {
                                    short ts = pt[vi];
                                    pt[vi] = pt[v2];
                                    pt[v2] = ts;
                                    double tc = cs[vi];
                                    cs[vi] = cs[v2];
                                    cs[v2] = tc;
                                }
                    }
                    wave_sync();
                    count_op(s.ops, OP_ORDER, wave_sum_i(ncos));
                    if (__ballot(noline) != 0ull) {
                        res.status = FA_CAPACITY;
                        break;
                    }
                }
                clk.stop(PH_FACE);
                // -- areas (calc_areas + project_points, LIB/Vcontacts.cpp:838-1044, 1278-1344)
                {
                    int epi = 0, eng[3] = {-1, -1, -1};
                    for (int base = 0; base < NC && epi < 3; base += WAVE) {
                        const int c = base + l;
                        unsigned long long m = __ballot(c < NC && s.pflag[c] == 'E');
                        while (m && epi < 3) {
                            eng[epi++] = base + first_lane(m);
                            m &= m - 1;
                        }
                    }
                    for (int c = l; c < NC; c += WAVE) s.area[c] = 0.0;
                    const double ro = (double)rado;
                    if (epi > 0) {
                        double proj[3];
                        if (epi == 1) {
                            const int X = eng[0];
                            proj[0] = -s.A0[X] * s.A3[X];
                            proj[1] = -s.A1[X] * s.A3[X];
                            proj[2] = -s.A2[X] * s.A3[X];
                        } else if (epi == 2) {
                            double a0[4], a1[4], p0[3], p1[3];
                            plane_of(s, eng[0], a0);
                            plane_of(s, eng[1], a1);
                            if (solve_2xS(a0, a1, rado, p0, p1) == -1) {
                                proj[0] = (double)(ax + rado);
                                proj[1] = (double)ay;
                                proj[2] = (double)az;
                            } else {
                                proj[0] = (p0[0] + p1[0]) / 2;
                                proj[1] = (p0[1] + p1[1]) / 2;
                                proj[2] = (p0[2] + p1[2]) / 2;
                            }
                        } else {
                            double a0[4], a1[4], a2[4], p0[3];
                            plane_of(s, eng[0], a0);
                            plane_of(s, eng[1], a1);
                            plane_of(s, eng[2], a2);
                            solve_3x3(a0, a1, a2, p0);
                            proj[0] = p0[0];
                            proj[1] = p0[1];
                            proj[2] = p0[2];
                        }
                        for (int vi = l; vi < NV; vi += WAVE) {
                            if (s.vp2[vi] == -1) continue;
                            double P0 = s.vx[vi], P1 = s.vy[vi], P2 = s.vz[vi];
                            double V0 = P0 - proj[0], V1 = P1 - proj[1], V2 = P2 - proj[2];
                            double qa = V0 * V0 + V1 * V1 + V2 * V2;
                            double qb = 2 * (P0 * V0 + P1 * V1 + P2 * V2);
                            double qc = P0 * P0 + P1 * P1 + P2 * P2 - (double)(rado * rado);
                            double kk = (sqrt(qb * qb - 4.0 * qa * qc) - qb) / (2 * qa);
                            s.vx[vi] = P0 + kk * V0;
                            s.vy[vi] = P1 + kk * V1;
                            s.vz[vi] = P2 + kk * V2;
                            s.vd[vi] = ro;
                        }
                    }
                    wave_sync();
                    int ntri = 0, narc = 0;
                    for (int X = l; X < NC; X += WAVE) {
                        if (s.pflag[X] == 'X') continue;
                        const int NP = s.fcnt[X];
                        const short* pt = &s.fpt[s.foff[X]];
                        const double ctr[3] = {-s.A0[X] * s.A3[X], -s.A1[X] * s.A3[X], -s.A2[X] * s.A3[X]};
                        const double cd = fabs(s.A3[X]);
                        double ar = 0.0;
                        if (NP == 0) {
                            if (test_point(s, ctr, NC, X)) ar = (2.0f * PI_REF) * ro * (ro - cd);
                        } else if (NP == 2) {
                            const double Pa[3] = {s.vx[pt[0]], s.vy[pt[0]], s.vz[pt[0]]};
                            const double Pb[3] = {s.vx[pt[1]], s.vy[pt[1]], s.vz[pt[1]]};
                            const double arc = spherical_arc(ctr, cd, Pa, s.vd[pt[0]], Pb, s.vd[pt[1]], rado);
                            ++narc;
                            if (test_point(s, ctr, NC, X))
                                ar = (2.0f * PI_REF) * ro * (ro - cd) - arc;
                            else
                                ar = arc;
                        } else {
                            const int w0 = pt[0];
                            const double x0 = s.vx[w0], y0 = s.vy[w0], z0 = s.vz[w0], d0 = s.vd[w0];
                            // cos to vertex 0 and to the next vertex, on the fly
                            auto cN0 = [&](int q) {
                                const int w = pt[q];
                                return (x0 * s.vx[w] + y0 * s.vy[w] + z0 * s.vz[w]) / (d0 * s.vd[w]);
                            };
                            auto cNN = [&](int q) {
                                const int wa = pt[q], wb = pt[(q + 1) % NP];
                                return (s.vx[wa] * s.vx[wb] + s.vy[wa] * s.vy[wb] + s.vz[wa] * s.vz[wb]) /
                                       (s.vd[wa] * s.vd[wb]);
                            };
                            double c0 = NP > 2 ? cN0(1) : 0.0;
                            if (NP > 2) ntri += NP - 2;
                            for (int q = 1; q < NP - 1; ++q) {
                                const double cq = cNN(q);
                                const double c1 = cN0(q + 1);
                                double U = sqrt((1 + c0) * (1 + cq) * (1 + c1) / 8.0);
                                double V = sqrt((1 - c0) * (1 - cq) * (1 + c1) / 8.0);
                                double Wv = sqrt((1 - c0) * (1 + cq) * (1 - c1) / 8.0);
                                double Xv = sqrt((1 + c0) * (1 - cq) * (1 - c1) / 8.0);
                                double tS = (1 - U + V + Wv + Xv) / (1 + U - V - Wv - Xv);
                                double tSA = (1 - U - V - Wv + Xv) / (1 + U + V + Wv - Xv);
                                double tSB = (1 - U - V + Wv - Xv) / (1 + U + V - Wv + Xv);
                                double tSC = (1 - U + V - Wv - Xv) / (1 + U - V + Wv + Xv);
                                double tp = sqrt(tS * tSA * tSB * tSC);
                                if (tp > 0.0) ar += 4.0 * ro * ro * atan(sqrt(tp));
                                c0 = c1;
                            }
                            for (int q = 0; q < NP; ++q) {
                                const int wa = pt[q], wb = pt[(q + 1) % NP];
                                if (s.vp2[wa] == -1 && s.vp2[wb] == -1 &&
                                    (s.vp0[wa] + s.vp1[wa]) != (s.vp0[wb] + s.vp1[wb])) {
                                    const double Pa[3] = {s.vx[wa], s.vy[wa], s.vz[wa]};
                                    const double Pb[3] = {s.vx[wb], s.vy[wb], s.vz[wb]};
                                    ar += spherical_arc(ctr, cd, Pa, s.vd[wa], Pb, s.vd[wb], rado);
                                    ++narc;
                                }
                            }
                        }
                        s.area[X] = ar;
                    }
                    count_op(s.ops, OP_TRI, wave_sum_i(ntri));
                    count_op(s.ops, OP_SARC, wave_sum_i(narc));
                    wave_sync();
                    if (epi > 0) {
                        // corrections onto the planes adjacent to engulfing faces, sequential
                        if (l == 0) {
                            for (int X = 0; X < NC; ++X) {
                                if (s.pflag[X] != 'E') continue;
                                const int NP = s.fcnt[X];
                                const short* pt = &s.fpt[s.foff[X]];
                                const double ctr[3] = {-s.A0[X] * s.A3[X], -s.A1[X] * s.A3[X], -s.A2[X] * s.A3[X]};
                                const double cd = fabs(s.A3[X]);
                                for (int q = 0; q < NP; ++q) {
                                    const int cu = pt[q], nx = pt[(q + 1) % NP];
                                    int pa1, pa2;
                                    if (s.vp0[cu] == X) {
                                        pa1 = s.vp1[cu];
                                        pa2 = s.vp2[cu];
                                    } else {
                                        pa1 = s.vp0[cu];
                                        pa2 = (s.vp1[cu] == X) ? s.vp2[cu] : s.vp1[cu];
                                    }
                                    int comm;
                                    if (pa1 == s.vp0[nx] || pa1 == s.vp1[nx] || pa1 == s.vp2[nx])
                                        comm = pa1;
                                    else if (pa2 == s.vp0[nx] || pa2 == s.vp1[nx] || pa2 == s.vp2[nx])
                                        comm = pa2;
                                    else
                                        continue;
                                    if (comm != -1 && s.pflag[comm] != 'E') {
                                        const double Pa[3] = {s.vx[cu], s.vy[cu], s.vz[cu]};
                                        const double Pb[3] = {s.vx[nx], s.vy[nx], s.vz[nx]};
                                        s.area[comm] += spherical_arc(ctr, cd, Pa, s.vd[cu], Pb, s.vd[nx], rado);
                                        s.ops[OP_SARC] += 1;
                                        if (NP == 2) break;
                                    }
                                }
                            }
                            if (epi == 1) {
                                s.area[eng[0]] = (2.0f * PI_REF) * ro * (ro + s.pdist[eng[0]]);
                            } else if (epi == 2) {
                                double a0[4], a1[4], b[3], c[3];
                                plane_of(s, eng[0], a0);
                                plane_of(s, eng[1], a1);
                                if (solve_2xS(a0, a1, rado, b, c) == -1) {
                                    s.area[eng[0]] = (2.0f * PI_REF) * ro * ro;
                                    s.area[eng[1]] = (2.0f * PI_REF) * ro * ro;
                                } else {
                                    const double c0[3] = {-a0[0] * a0[3], -a0[1] * a0[3], -a0[2] * a0[3]};
                                    const double c1[3] = {-a1[0] * a1[3], -a1[1] * a1[3], -a1[2] * a1[3]};
                                    double mx = spherical_arc(c0, fabs(a0[3]), b, ro, c, ro, rado);
                                    mx += spherical_arc(c1, fabs(a1[3]), b, ro, c, ro, rado);
                                    s.ops[OP_SARC] += 2;
                                    s.area[eng[0]] = (2.0f * PI_REF) * ro * ro - 0.5 * mx;
                                    s.area[eng[1]] = (2.0f * PI_REF) * ro * ro - 0.5 * mx;
                                }
                            }
                        }
                        wave_sync();
                        if (epi >= 3) {
                            for (int X = l; X < NC; X += WAVE)
                                s.area[X] = (s.pflag[X] == 'E') ? (4.0f * PI_REF) * ro * ro / epi : 0.0;
                            wave_sync();
                        }
                    }
                }
                clk.stop(PH_AREA);
                // -- records (save_areas) and the energy terms of atom a (vcfunction)
                {
                    const double radA = (double)ra;
                    const double radoA = radA + RW;
                    const double permeaD = (double)prm.permeability;
                    const double dee = (double)prm.dee_clash;
                    const int nt = prm.ntypes;
                    const int ta = lig.type[a];
                    double sas_sub = 0.0, com = 0.0, wal = 0.0;
                    int rcl = 0;
                    int recbase = nrec_total;
                    int nvalid_total = 0;
                    // count valid to place records in reverse order
                    for (int base = 0; base < NC; base += WAVE) {
                        const int c = base + l;
                        nvalid_total += __popcll(__ballot(c < NC && s.area[c] > 0.0 && s.pflag[c] != 'X'));
                    }
                    int seen = 0;
                    for (int base = 0; base < NC; base += WAVE) {
                        const int c = base + l;
                        const bool valid = c < NC && s.area[c] > 0.0 && s.pflag[c] != 'X';
                        const unsigned long long m = __ballot(valid);
                        if (valid) {
                            const int j = s.pidx[c];
                            const double ar = s.area[c];
                            const double dist = s.cdist[c];
                            sas_sub += ar;
                            const bool lj = j >= Rn;
                            const int mj = j - Rn;
                            float rb;
                            int tb;
                            if (lj) {
                                rb = lig.radius[mj];
                                tb = lig.type[mj];
                            } else {
                                rb = R.rad[j];
                                tb = R.type[j];
                            }
                            if (B.recs) {
                                const int r = recbase + (nvalid_total - 1 - (seen + lanes_below(m)));
                                if (r < B.rec_cap) {
                                    fa_contact_rec& o = B.recs[(size_t)pose * B.rec_cap + r];
                                    o.from = acalc;
                                    o.atom = j;
                                    o.area = ar;
                                    o.dist = dist;
                                }
                            }
                            bool skip = false;
                            if (lj) {
                                const bool bonded = (lig.bonded[2 * mj + (a >> 6)] >> (a & 63)) & 1ull;
                                skip = bonded || (mj < a);
                                if (!s.done[mj]) s.stub[2 * mj + (a >> 6)] |= (1ull << (a & 63));
                            }
                            if (!skip) {
                                const double rAB = radA + (double)rb;
                                const float clash_distance = (float)(permeaD * rAB);
                                if (dist < (double)clash_distance) {
                                    const double d2 = dist * dist, d4 = d2 * d2, d6 = d4 * d2;
                                    const double cr = permeaD * rAB, cr2 = cr * cr, cr4 = cr2 * cr2, cr6 = cr4 * cr2;
                                    wal += (double)KWALL_F * (1.0 / (d6 * d6) - 1.0 / (cr6 * cr6));
                                    if (dist <= dee * rAB) rcl = 1;
                                }
                                if (prm.intramolecular || !lj) com += (double)prm.emat[(ta - 1) * nt + (tb - 1)] * ar;
                            }
                        }
                        seen += __popcll(m);
                    }
                    nrec_total += nvalid_total;
                    count_op(s.ops, OP_VC, nvalid_total);
                    sas_sub = wave_sum_d(sas_sub);
                    com = wave_sum_d(com);
                    wal = wave_sum_d(wal);
                    rcl = wave_sum_i(rcl);
                    if (l == 0) {
                        double SAS = 4.0 * PI_REF * radoA * radoA - sas_sub;
                        if (SAS < 0.0) SAS = 0.0;
                        s.acc_com[a] = com;
                        s.acc_wal[a] = wal;
                        s.acc_tot[a] = SAS;
                        s.acc_sas[a] = (prm.solventterm != 0.0f)
                                           ? (double)prm.solventterm * SAS
                                           : (double)prm.emat[(ta - 1) * nt + (nt - 1)] * SAS;
                        s.acc_rcl[a] = rcl ? 1 : 0;
                        s.done[a] = 1;
                    }
                    wave_sync();
                }
                clk.stop(PH_REC);
            }
            if (res.status == FA_OK) {
                // atoms with no contact still contribute their full SAS
                for (int i = l; i < L; i += WAVE) {
                    if (!s.done[i]) {
                        const double radoA = (double)lig.radius[i] + RW;
                        const double SAS = 4.0 * PI_REF * radoA * radoA;
                        const int ti = lig.type[i];
                        s.acc_com[i] = 0.0;
                        s.acc_wal[i] = 0.0;
                        s.acc_tot[i] = SAS;
                        s.acc_sas[i] = (prm.solventterm != 0.0f)
                                           ? (double)prm.solventterm * SAS
                                           : (double)prm.emat[(ti - 1) * prm.ntypes + (prm.ntypes - 1)] * SAS;
                        s.acc_rcl[i] = 0;
                    }
                }
                wave_sync();
                if (l == 0) {
                    for (int i = 0; i < L; ++i) {
                        res.com += s.acc_com[i];
                        res.wal += s.acc_wal[i];
                        res.sas += s.acc_sas[i];
                        res.totsas += s.acc_tot[i];
                        res.rclash |= s.acc_rcl[i];
                    }
                    res.cf = res.com + res.wal + res.sas + res.con;
                    res.app = res.com + res.wal + res.sas;
                }
            }
        }
        if (res.status == FA_CAPACITY && l == 0) atomicAdd(&B.counters[1], 1u);
        wave_sync();
        if (B.ops && l < FA_OPCOUNT_N && s.ops[l]) atomicAdd(&B.ops[l], s.ops[l]);
        if (l == 0) {
            B.out[pose] = res;
            if (B.rec_count) B.rec_count[pose] = (res.status == FA_OK) ? nrec_total : 0;
        }
        wave_sync();
    }
}

// host-side launcher (same translation unit as the kernel)
hipError_t launch_cf(const DevReceptor& R, const DevGrid* grids, const DevLigand* ligs, const DevParams& prm,
                     const DevBatch& B, int n_blocks, hipStream_t stream)
{
    if (B.n_poses <= 0) return hipSuccess;
    if (n_blocks <= 0 || n_blocks > B.n_poses) n_blocks = B.n_poses;
    hipLaunchKernelGGL(cf_pose_kernel, dim3(n_blocks), dim3(WAVE), 0, stream, R, grids, ligs, prm, B);
    return hipGetLastError();
}

int smem_bytes() { return (int)sizeof(Smem); }

}  // namespace fa
