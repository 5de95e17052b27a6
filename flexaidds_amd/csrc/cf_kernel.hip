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
// the next (LIB/Vcontacts.cpp:1750-1836, :595-629).  Work inside an atom is
// spread over lanes:
//   - neighbour cull, once per atom: the 27 cell ranges are fetched in one
//     round (the next atom's while the current one streams), candidates go
//     through in 64-wide chunks with the next chunk's loads in flight, and the
//     in-range ones land in a per-wave neighbour list in HBM.  The clash
//     pre-pass sums its walls on the way; the polyhedron pass re-reads the
//     list instead of scanning the cells again;
//   - hull walk: lane l keeps planes l, l+64 and edges l, l+64, ... in
//     registers; the closest-plane search is a DPP reduction, the open-edge
//     queue is a set of wave-uniform bit masks, so a step touches LDS only to
//     store the new vertex;
//   - arc points, faces, areas and records: lanes over edges / faces /
//     contacts, compactions with ballots.
// The LDS working set is ~19 KB, so eight waves share a CU.
#include <hip/hip_runtime.h>

#include <climits>
#include <stdint.h>

#include "cell_slots.h"
#include "fa_device.h"

namespace fa {

constexpr int WAVE = 64;
constexpr int MAXC = 128;                      // contacts (planes) of one atom
constexpr int MAXP = MAXC + 4;                 // + 4 bounding planes
constexpr int PSL = MAXC / WAVE;               // plane register slots per lane
constexpr int MAXV = 160;                      // hull vertices, arc points, opposite points
constexpr int VSL = (MAXV + WAVE - 1) / WAVE;  // vertex slots per lane
constexpr int MAXE = 256;                      // hull edges (the walk fails at 200)
constexpr int ESL = MAXE / WAVE;               // edge register slots per lane
constexpr int MAXF = 384;                      // face point-list slots
constexpr int MAXL = FA_MAX_SCORABLE;       // scorable atoms of a pose (side chains + ligand)
constexpr int LSL = MAXL / WAVE;               // ligand atom slots per lane
constexpr unsigned NONE = 0xFFu;               // packed plane id "none"

constexpr double PI_REF = 3.141592;     // LIB/flexaid.h:86
constexpr float RW = 1.4f;              // LIB/flexaid.h:88
constexpr float KWALL_F = 1.0e6f;       // LIB/flexaid.h:81
constexpr float POLY_PENALTY = 1e6f;    // LIB/flexaid.h:69
constexpr double CLASH_THRESHOLD = 1e4; // LIB/flexaid.h:70
constexpr float CELLSIZE = 6.5f;        // LIB/Vcontacts.h:56
constexpr int EDGE_LIMIT = 200;         // LIB/Vcontacts.cpp:378
constexpr double BIG = 9.9e+9;          // sentinel of the reference scans

static_assert(MAXC % WAVE == 0 && MAXE % WAVE == 0 && MAXL % WAVE == 0, "slot layout");
static_assert(FA_MAX_LIG_ATOMS <= MAXL, "a ligand fits the scorable-atom slots");
static_assert(MAXP < (int)NONE, "plane ids are packed in 8 bits");
static_assert(MAXE >= EDGE_LIMIT + 2, "the edge limit must be reached before the edge capacity");

constexpr int N_PHASES = 10;  // phases of the optional cycle accounting

struct Smem {
    // planes of the current atom: c < NC contacts, NC..NC+3 bounding planes
    double A0[MAXP], A1[MAXP], A2[MAXP], A3[MAXP];
    // polyhedron vertices (coordinates relative to the atom)
    double vx[MAXV], vy[MAXV], vz[MAXV];
    // storage of phases that never overlap
    union {
        double fcos[MAXF];                  // face ordering cosines
        double area[MAXC];                  // areas while engulfing corrections run
        struct {
            double cd[MAXC];                // contact distance, radius and index
            float rb[MAXC];                 //   on their way into registers
            int idx[MAXC];
        } c;
        struct {
            float dis[MAXL], ang[MAXL], dih[MAXL];
            int rot[FA_MAX_FLEX_RESIDUES];  // conformation of each flexible residue
        } b;                                // internal coordinates (pose build)
        struct {
            int span[64];                   // per-slot scratch of one atom's scan
            int rcell[MAXL];                // cell of each rank, x | y << 10 | z << 20
            int calc[MAXL];                 // calc index of each scorable atom
        } p;                                // neighbour pre-pass
    } u;
    unsigned long long stub[2 * MAXL];      // stub[2m..2m+1]: atoms with a record of m
    unsigned long long ops[FA_OPCOUNT_N];   // operation counters of the pose
    unsigned long long phase[N_PHASES];     // diagnostic launches: cycles per phase, this wave
    unsigned vpl[MAXV];                     // vertex planes p0 | p1 << 8 | p2 << 16
    float lx[MAXL], ly[MAXL], lz[MAXL];     // scorable atoms: side-chain atoms, then the ligand's
    float lr[MAXL];                         //   radius
    int sbox[MAXL];                         // cell of the atom of each rank
    int nstart[MAXL + 1];                   // neighbour list of rank k: [nstart[k], nstart[k+1])
    int seed[3 * MAXL];                     // hull seeds (calc indices)
    short fpt[MAXF];                        // face point lists (CSR)
    short foff[MAXC];
    unsigned char fcnt[MAXC];
    char pflag[MAXP];                       // 'X' off the polyhedron, 'Y' on it, 'E' engulfing
    signed char pmv[MAXC];                  // scorable index of each contact, -1 for rigid atoms
    unsigned char order[MAXL], rank[MAXL];  // box order: calc index of rank k, rank of atom i
};

// ------------------------------------------------------------ wave helpers
__device__ __forceinline__ int lane_id() { return __lane_id(); }

// lane mask of a predicate (bool-typed: an int-typed predicate makes the
// compiler turn the condition into 0/1 and compare it again first)
__device__ __forceinline__ unsigned long long ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

__device__ __forceinline__ int lanes_below(unsigned long long m)
{
    return __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

__device__ __forceinline__ int first_lane(unsigned long long m) { return __ffsll((long long)m) - 1; }

// a wave-uniform value moved to a scalar register, so that branches on it
// compile to scalar branches instead of exec-mask save/restore sequences
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

template <int CTRL>
__device__ __forceinline__ int dpp_i(int v)
{
    // lanes whose source is outside their row read 0
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, true);
}

template <int CTRL>
__device__ __forceinline__ double dpp_d(double v)
{
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), CTRL, 0xF, 0xF, true);
    return __longlong_as_double((long long)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo));
}

// Wave reductions and scans: DPP inside each 16-lane row (quad_perm,
// row_half_mirror, row_mirror butterflies; row_shr for scans), then the four
// row results through readlane.  All lanes must be active.
__device__ __forceinline__ int wave_sum_i(int v)
{
    v += dpp_i<0xB1>(v);   // quad_perm [1,0,3,2]
    v += dpp_i<0x4E>(v);   // quad_perm [2,3,0,1]
    v += dpp_i<0x141>(v);  // row_half_mirror
    v += dpp_i<0x140>(v);  // row_mirror
    return __builtin_amdgcn_readlane(v, 0) + __builtin_amdgcn_readlane(v, 16) + __builtin_amdgcn_readlane(v, 32) +
           __builtin_amdgcn_readlane(v, 48);
}

__device__ __forceinline__ double readlane_d(double v, int lane);

__device__ __forceinline__ double wave_sum_d(double v)
{
    v += dpp_d<0xB1>(v);
    v += dpp_d<0x4E>(v);
    v += dpp_d<0x141>(v);
    v += dpp_d<0x140>(v);
    return (readlane_d(v, 0) + readlane_d(v, 16)) + (readlane_d(v, 32) + readlane_d(v, 48));
}

__device__ __forceinline__ int wave_incl_scan(int v)
{
    v += dpp_i<0x111>(v);  // row_shr:1
    v += dpp_i<0x112>(v);  // row_shr:2
    v += dpp_i<0x114>(v);  // row_shr:4
    v += dpp_i<0x118>(v);  // row_shr:8
    const int r0 = __builtin_amdgcn_readlane(v, 15), r1 = __builtin_amdgcn_readlane(v, 31);
    const int r2 = __builtin_amdgcn_readlane(v, 47);
    const int l = __lane_id();
    return v + (l >= 16 ? r0 : 0) + (l >= 32 ? r1 : 0) + (l >= 48 ? r2 : 0);
}

__device__ __forceinline__ int wave_excl_scan(int v) { return wave_incl_scan(v) - v; }

// LDS visibility point for the whole (single-wave) work-group
__device__ __forceinline__ void wave_sync() { __syncthreads(); }

__device__ __forceinline__ int readlane_i(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }

__device__ __forceinline__ unsigned readlane_u(unsigned v, int lane)
{
    return (unsigned)__builtin_amdgcn_readlane((int)v, lane);
}

__device__ __forceinline__ float readlane_f(float v, int lane)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

__device__ __forceinline__ double readlane_d(double v, int lane)
{
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, lane);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), lane);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

__device__ __forceinline__ uint64_t readlane_u64(uint64_t v, int lane)
{
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, lane);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), lane);
    return ((uint64_t)hi << 32) | lo;
}

template <int CTRL>
__device__ __forceinline__ uint64_t dpp_u64(uint64_t v)
{
    // every control used below reads a lane of the same row, so no lane reads
    // out of range; `old` (all ones) is the identity of the minimum anyway
    const int lo = __builtin_amdgcn_update_dpp(-1, (int)(uint32_t)v, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(-1, (int)(uint32_t)(v >> 32), CTRL, 0xF, 0xF, false);
    return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
}

__device__ __forceinline__ uint64_t umin64(uint64_t a, uint64_t b) { return a < b ? a : b; }

// one DPP step of a 32-bit minimum: lanes of rows outside ROWS, or without a
// source, combine with themselves
template <int CTRL, int ROWS>
__device__ __forceinline__ unsigned dpp_min_u32(unsigned v)
{
    const unsigned o = (unsigned)__builtin_amdgcn_update_dpp((int)v, (int)v, CTRL, ROWS, 0xF, false);
    return v < o ? v : o;
}

// one in-row DPP step of a 32-bit minimum: every lane has a source in its
// row for the controls used, so bound_ctrl (read 0 without a source) never
// applies and the move can fuse into the minimum
template <int CTRL>
__device__ __forceinline__ unsigned dpp_row_min_u32(unsigned v)
{
    const unsigned o = (unsigned)__builtin_amdgcn_mov_dpp((int)v, CTRL, 0xF, 0xF, true);
    return v < o ? v : o;
}

// Minimum over the wave of an unsigned 32-bit value, all lanes active:
// butterflies inside each row, then row_bcast:15 (rows 1, 3) and
// row_bcast:31 (rows 2, 3) carry the lower rows up; lane 63 ends with the
// wave minimum.
__device__ __forceinline__ unsigned wave_min_u32(unsigned v)
{
    v = dpp_row_min_u32<0xB1>(v);    // quad_perm [1,0,3,2]
    v = dpp_row_min_u32<0x4E>(v);    // quad_perm [2,3,0,1]
    v = dpp_row_min_u32<0x141>(v);   // row_half_mirror
    v = dpp_row_min_u32<0x140>(v);   // row_mirror
    v = dpp_min_u32<0x142, 0xA>(v);  // row_bcast:15
    v = dpp_min_u32<0x143, 0xC>(v);  // row_bcast:31
    return readlane_u(v, 63);
}

// Minimum over the wave of an unsigned 64-bit key, all lanes active: the
// minimum high word, then the minimum low word among the lanes that hold it.
__device__ __forceinline__ uint64_t wave_min_u64(uint64_t v)
{
    const unsigned hi = (unsigned)(v >> 32);
    const unsigned mh = wave_min_u32(hi);
    const unsigned ml = wave_min_u32(hi == mh ? (unsigned)v : 0xFFFFFFFFu);
    return ((uint64_t)mh << 32) | ml;
}

// non-negative doubles order like their bit patterns
__device__ __forceinline__ uint64_t key_of(double v) { return (uint64_t)__double_as_longlong(v); }

enum OpIdx { OP_CULL = 0, OP_PLANE, OP_HULL, OP_ARC, OP_ORDER, OP_TRI, OP_SARC, OP_VC, OP_BUILD };

// add a wave-uniform count to the pose's operation counter (lane 0 writes)
__device__ __forceinline__ void count_op(unsigned long long* ops, int k, long long n)
{
    if (lane_id() == 0) ops[k] += (unsigned long long)n;
}

// optional per-phase cycle accounting (diagnostic launches only)
enum Phase { PH_BUILD = 0, PH_PREPASS, PH_CULL, PH_FIRST, PH_HULL, PH_ARC, PH_SEED, PH_FACE, PH_AREA, PH_REC, PH_N };
static_assert(PH_N == N_PHASES, "phase slots");
constexpr int PROF_POSE0 = 16;  // per-pose cycles follow the 16 phase slots

// Phase cycles are summed per wave in LDS (no global traffic between
// phases, which would hold up the wave's next global load behind a
// contended atomic) and added to the launch totals once, at the wave's exit.
struct PhaseClock {
    unsigned long long* acc;  // the wave's LDS slots, null when disabled
    unsigned long long t;
    __device__ __forceinline__ void start()
    {
        if (acc) t = __builtin_readcyclecounter();
    }
    __device__ __forceinline__ void stop(int ph)
    {
        if (acc) {
            const unsigned long long now = __builtin_readcyclecounter();
            if (lane_id() == 0) atomicAdd(&acc[ph], now - t);
            t = now;
        }
    }
};

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

// out-of-line copy for the rare engulfing corrections (keeps the inlined
// transcendental code to one site of low register pressure)
__device__ __noinline__ double spherical_arc_call(double ax, double ay, double az, double ad, double bx, double by,
                                                  double bz, double bd, double cx, double cy, double cz, double cd,
                                                  float rado)
{
    const double A[3] = {ax, ay, az}, B[3] = {bx, by, bz}, C[3] = {cx, cy, cz};
    return spherical_arc(A, ad, B, bd, C, cd, rado);
}

// one spherical triangle of a face fan (LIB/Vcontacts.cpp:926-940): the term
// the reference adds for the triangle with cosines c0 (vertex 0 to q),
// cq (q to q+1) and c1 (vertex 0 to q+1); zero when it adds nothing
constexpr double KDIST = 1.0e3;  // LIB/flexaid.h:83

// GetValueFromGaussian (LIB/geometry.cpp:3-5): 1 at x = opt, 0 at x = zero
// and 2 opt - zero, an even power outside
__device__ __forceinline__ double gaussian_bell(double x, double opt, double zero)
{
    return pow(-(x - zero) * (x - (2 * opt - zero)) / pow(zero - opt, 2.0), 50.0);
}

// Scored value of one type pair k for `area` of an atom whose whole sphere
// is `surf`: weight-1 pairs give y * area (y * area / surf with
// normalize_area, LIB/vcfunction.cpp:343-353, :479-491); density pairs give
// get_yval(area / surf) (LIB/vcfunction.cpp:576-621): 0 left of the first
// node, the last y right of the last node, linear in between, the node
// differences taken in float as the reference's float fields are.
__device__ __forceinline__ double pair_term(const DevParams& prm, int k, double area, double surf)
{
    if (prm.dens_off) {
        const int b = prm.dens_off[k], e = prm.dens_off[k + 1];
        if (e > b) {
            const double rel = area / surf;
            int i = b;
            while (i + 1 < e && rel > (double)prm.dens_xy[i + 1].x) ++i;
            const float2 p0 = prm.dens_xy[i];
            if ((double)p0.x > rel) return 0.0;
            if (i + 1 == e) return (double)p0.y;
            const float2 p1 = prm.dens_xy[i + 1];
            return (double)p0.y + (rel - (double)p0.x) / (double)(p1.x - p0.x) * (double)(p1.y - p0.y);
        }
    }
    const double y = (double)prm.emat[k];
    return prm.normalize_area ? y * area / surf : y * area;
}

__device__ __forceinline__ double sph_triangle(double c0, double cq, double c1, double ro)
{
    const double U = sqrt((1 + c0) * (1 + cq) * (1 + c1) / 8.0);
    const double V = sqrt((1 - c0) * (1 - cq) * (1 + c1) / 8.0);
    const double Wv = sqrt((1 - c0) * (1 + cq) * (1 - c1) / 8.0);
    const double Xv = sqrt((1 + c0) * (1 - cq) * (1 - c1) / 8.0);
    const double tS = (1 - U + V + Wv + Xv) / (1 + U - V - Wv - Xv);
    const double tSA = (1 - U - V - Wv + Xv) / (1 + U + V + Wv - Xv);
    const double tSB = (1 - U - V + Wv - Xv) / (1 + U + V - Wv + Xv);
    const double tSC = (1 - U + V - Wv - Xv) / (1 + U - V + Wv + Xv);
    const double tp = sqrt(tS * tSA * tSB * tSC);
    return (tp > 0.0) ? 4.0 * ro * ro * atan(sqrt(tp)) : 0.0;
}

// test_point (LIB/Vcontacts.cpp:1357-1372) for the points t of the lanes
// with `need` (a point of the face of plane X = lane + 64 q): a point is kept
// unless a polyhedron plane other than X has it on its outer side.  The
// points are taken one at a time and every plane is tested on its own lane
// (planes c < NC in registers: lane c & 63, slot c >> 6), with the
// reference's sums.
__device__ __forceinline__ bool points_inside(bool need, const double* t, int q, int NC, const double* pa0,
                                              const double* pa1, const double* pa2, const double* pa3,
                                              const char* pfl)
{
    const int l = lane_id();
    NC = uni(NC);
    unsigned long long kept = 0ull;
    for (unsigned long long nm = ballot(need); nm; nm &= nm - 1) {
        const int xl = first_lane(nm);
        const int Xc = WAVE * q + xl;
        const double t0 = readlane_d(t[0], xl), t1 = readlane_d(t[1], xl), t2 = readlane_d(t[2], xl);
        bool outside = false;
#pragma unroll
        for (int q2 = 0; q2 < PSL; ++q2) {
            if (q2 * WAVE >= NC) break;  // wave-uniform
            const int c = l + WAVE * q2;
            const bool above = pa0[q2] * t0 + pa1[q2] * t1 + pa2[q2] * t2 + pa3[q2] > 0.0;
            outside = outside || (c < NC && c != Xc && pfl[q2] != 'X' && above);
        }
        if (ballot(outside) == 0ull) kept |= 1ull << xl;
    }
    return need && ((kept >> l) & 1ull);
}

__device__ __forceinline__ void plane_of(const Smem& s, int c, double* a)
{
    a[0] = s.A0[c];
    a[1] = s.A1[c];
    a[2] = s.A2[c];
    a[3] = s.A3[c];
}

// the four bounding planes (LIB/Vcontacts.cpp:223-241)
__device__ __forceinline__ void bounding_plane(int k, double* a)
{
    a[0] = (k < 2) ? 0.707 : -0.707;
    a[1] = (k == 0) ? 1.0 : (k == 1) ? -1.0 : 0.0;
    a[2] = (k == 2) ? 1.0 : (k == 3) ? -1.0 : 0.0;
    a[3] = -10.0;
}

// vertex planes, packed 8 bits each (NONE = -1)
__device__ __forceinline__ unsigned pack_planes(int a, int b, int c)
{
    return ((unsigned)a & NONE) | (((unsigned)b & NONE) << 8) | (((unsigned)c & NONE) << 16);
}

__device__ __forceinline__ int plane_at(unsigned pk, int k)
{
    const unsigned v = (pk >> (8 * k)) & NONE;
    return v == NONE ? -1 : (int)v;
}

__device__ __forceinline__ bool has_plane(unsigned pk, int X)
{
    return (int)(pk & NONE) == X || (int)((pk >> 8) & NONE) == X || (int)((pk >> 16) & NONE) == X;
}

// consecutive face points that bound an arc of the face (LIB/Vcontacts.cpp:942-954):
// two arc points (no third plane) on different plane pairs
__device__ __forceinline__ int arc_pair(unsigned ka, unsigned kb)
{
    return (plane_at(ka, 2) == -1 && plane_at(kb, 2) == -1 &&
            (plane_at(ka, 0) + plane_at(ka, 1)) != (plane_at(kb, 0) + plane_at(kb, 1)))
               ? 1
               : 0;
}

// distance of a hull vertex from the atom (the value the reference stores)
__device__ __forceinline__ double vnorm(const Smem& s, int v)
{
    const double x = s.vx[v], y = s.vy[v], z = s.vz[v];
    return sqrt(x * x + y * y + z * z);
}

// ------------------------------------------------------------ edge registers
// Edge e lives in lane (e & 63), slot (e >> 6) as two packed words; its
// direction is recomputed from its planes when the walk takes it up.  An
// edge is open (still needs its end point) while its end plane is NONE.
struct Edges {
    unsigned pl[ESL];                   // plane0 | plane1 << 8 | start plane << 16 | end plane << 24
    unsigned vt[ESL];                   // start vertex | end vertex << 16
};

// The slot of an edge is wave-uniform, but every update is written as a
// select over all slots with constant indices (no branch per slot): a
// store or read under a switch on the slot gets merged into one access at
// a computed index, which moves the whole array to scratch memory.

// new edge along planes p0 x p1 from vertex sv, turned away from plane tp
__device__ __forceinline__ void edge_new(Edges& E, int e, int p0, int p1, int tp, int sv)
{
    const unsigned pl = ((unsigned)p0 & NONE) | (((unsigned)p1 & NONE) << 8) | (((unsigned)tp & NONE) << 16) |
                        (NONE << 24);
    const unsigned vt = (unsigned)sv | 0xFFFF0000u;
    e = uni(e);
    const bool me = lane_id() == (e & 63);
    const int k = e >> 6;
    if (k == 0) {  // the common case: the first 64 edges
        E.pl[0] = me ? pl : E.pl[0];
        E.vt[0] = me ? vt : E.vt[0];
        return;
    }
#pragma unroll
    for (int K = 1; K < ESL; ++K) {
        const bool hit = me && k == K;
        E.pl[K] = hit ? pl : E.pl[K];
        E.vt[K] = hit ? vt : E.vt[K];
    }
}

// end point and end plane of edge e; the edge leaves the open queue
__device__ __forceinline__ void edge_close(Edges& E, int e, int end_plane, int end_vertex)
{
    e = uni(e);
    const bool me = lane_id() == (e & 63);
    const int k = e >> 6;
    const unsigned ep = ((unsigned)end_plane & NONE) << 24, ev = (unsigned)end_vertex << 16;
    if (k == 0) {  // the common case: the first 64 edges
        E.pl[0] = me ? ((E.pl[0] & 0x00FFFFFFu) | ep) : E.pl[0];
        E.vt[0] = me ? ((E.vt[0] & 0xFFFFu) | ev) : E.vt[0];
        return;
    }
#pragma unroll
    for (int K = 1; K < ESL; ++K) {
        const bool hit = me && k == K;
        E.pl[K] = hit ? ((E.pl[K] & 0x00FFFFFFu) | ep) : E.pl[K];
        E.vt[K] = hit ? ((E.vt[K] & 0xFFFFu) | ev) : E.vt[K];
    }
}

// packed planes and vertices of edge e (wave-uniform)
__device__ __forceinline__ void edge_get(const Edges& E, int e, unsigned& pl, unsigned& vt)
{
    e = uni(e);
    const int k = e >> 6;
    unsigned p = E.pl[0], v = E.vt[0];
#pragma unroll
    for (int K = 1; K < ESL; ++K) {
        p = (k == K) ? E.pl[K] : p;
        v = (k == K) ? E.vt[K] : v;
    }
    pl = readlane_u(p, e & 63);
    vt = readlane_u(v, e & 63);
}

// Lowest-numbered open edge among edges [from, edgenum), -1 when none.  The
// walk always takes the lowest open edge, new edges get the next number and
// edges only ever close, so the lowest open edge never moves down: the
// caller passes the previous one as `from`.
__device__ __forceinline__ int first_open(const Edges& E, int from, int edgenum)
{
    from = uni(from);
    edgenum = uni(edgenum);
    const int l = lane_id();
#pragma unroll
    for (int k = 0; k < ESL; ++k) {
        if (k * 64 >= edgenum) break;
        if ((k + 1) * 64 <= from) continue;
        const int e = k * 64 + l;
        const unsigned long long m = ballot(e >= from && e < edgenum && (E.pl[k] >> 24) == NONE);
        if (m) return k * 64 + first_lane(m);
    }
    return -1;
}

// first edge e in (lo, hi) between planes P and C (either order), -1 if none
__device__ __forceinline__ int find_edge(const Edges& E, int lo, int hi, int P, int C)
{
    lo = uni(lo);
    hi = uni(hi);
    P = uni(P);
    C = uni(C);
    const int l = lane_id();
    const unsigned pc = (unsigned)P | ((unsigned)C << 8), cp = (unsigned)C | ((unsigned)P << 8);
#pragma unroll
    for (int k = 0; k < ESL; ++k) {
        if (k * 64 >= hi) break;
        const int e = k * 64 + l;
        const unsigned both = E.pl[k] & 0xFFFFu;  // plane0 | plane1 << 8
        const bool hit = e > lo && e < hi && (both == pc || both == cp);
        const unsigned long long m = ballot(hit);
        if (m) return k * 64 + first_lane(m);
    }
    return -1;
}

// LIB/Vcontacts.cpp:599-637: direction of the edge along planes p0 x p1 from
// the point (x, y, z), turned away from plane tp (the same arithmetic at
// every call, so the walk sees the values the reference stores)
__device__ __forceinline__ void edge_dir(const Smem& s, int p0, int p1, int tp, double x, double y, double z,
                                         double& V0, double& V1, double& V2)
{
    double a[4], b[4], t[4];
    plane_of(s, p0, a);
    plane_of(s, p1, b);
    plane_of(s, tp, t);
    V0 = a[1] * b[2] - a[2] * b[1];
    V1 = a[2] * b[0] - a[0] * b[2];
    V2 = a[0] * b[1] - a[1] * b[0];
    const double t0 = x + V0, t1 = y + V1, t2 = z + V2;
    const double tv = t[0] * t0 + t[1] * t1 + t[2] * t2 + t[3];
    if (tv > 0.0) {
        V0 = -V0;
        V1 = -V1;
        V2 = -V2;
    }
}

// One step of the hull walk (LIB/Vcontacts.cpp:271-335) for edge e: the
// search over the contact planes (lane-wise quotients, DPP minimum), then the
// bounding planes when the point leaves the sphere, and the new vertex.
struct HullProbe {
    int A, B, oldC, C;
    double dx, dy, dz, px, py, pz;
    double vtmin;
    int tests;
};

struct HullStep {
    int A, B, C;
    double nx, ny, nz;
    int tests;
};

__device__ __forceinline__ HullProbe hull_search(const Smem& s, const Edges& E, int e, int NC, const double* pa0,
                                                 const double* pa1, const double* pa2, const double* pa3)
{
    HullProbe h;
    unsigned epl, evt;
    e = uni(e);
    NC = uni(NC);
    edge_get(E, e, epl, evt);
    h.A = uni((int)(epl & NONE));
    h.B = uni((int)((epl >> 8) & NONE));
    h.oldC = uni((int)((epl >> 16) & NONE));
    const int sv = uni((int)(evt & 0xFFFFu));
    h.px = s.vx[sv];
    h.py = s.vy[sv];
    h.pz = s.vz[sv];
    edge_dir(s, h.A, h.B, h.oldC, h.px, h.py, h.pz, h.dx, h.dy, h.dz);
    // closest plane cut ahead along the edge: lane-wise quotients, DPP minimum
    const int l = lane_id();
    uint64_t kq[PSL];
    uint64_t kb = ~0ull;
#pragma unroll
    for (int q = 0; q < PSL; ++q) {
        kq[q] = ~0ull;
        if (q * WAVE < NC) {  // wave-uniform: skip empty slots
            // every lane computes its quotient (no divergent branches); lanes
            // without a candidate plane, a zero divisor (the quotient is then
            // inf or NaN) or a cut behind the point keep the empty key
            const int c = l + WAVE * q;
            const double vtdiv = pa0[q] * h.dx + pa1[q] * h.dy + pa2[q] * h.dz;
            const double vt = -(pa0[q] * h.px + pa1[q] * h.py + pa2[q] * h.pz + pa3[q]) / vtdiv;
            const bool ok = c < NC && c != h.A && c != h.B && c != h.oldC && vtdiv != 0.0 && vt < BIG && vt > 0.0;
            kq[q] = ok ? key_of(vt) : ~0ull;
            kb = umin64(kb, kq[q]);
        }
    }
    h.tests = NC - (h.A < NC) - (h.B < NC) - (h.oldC < NC);
    const uint64_t kmin = wave_min_u64(kb);
    h.C = -1;
    h.vtmin = BIG;
    if (kmin != ~0ull) {
        h.vtmin = __longlong_as_double((long long)kmin);
#pragma unroll
        for (int q = 0; q < PSL; ++q) {
            const unsigned long long m = ballot(kq[q] == kmin);
            if (h.C < 0 && m) h.C = WAVE * q + first_lane(m);
        }
    }
    return h;
}

__device__ __forceinline__ HullStep hull_finish(const HullProbe& h, int NC, double rr)
{
    NC = uni(NC);
    HullStep w;
    w.A = h.A;
    w.B = h.B;
    w.C = h.C;
    w.tests = h.tests;
    double vtmin = h.vtmin;
    w.nx = h.px + vtmin * h.dx;
    w.ny = h.py + vtmin * h.dy;
    w.nz = h.pz + vtmin * h.dz;
    // the point and cut values are the same on every lane (loaded from
    // uniform LDS addresses), but the compiler cannot tell: the decisions
    // go through scalar registers so that the walk stays a scalar loop
    if (uni(w.nx * w.nx + w.ny * w.ny + w.nz * w.nz > rr ? 1 : 0)) {
        w.tests += 4 - (h.A >= NC) - (h.B >= NC) - (h.oldC >= NC);
        // the four bounding planes NC + k on lanes k = 0..3 (one quotient
        // each), then taken in plane order: the first strictly smaller cut
        // ahead wins, as in a scan over them
        const int l = lane_id();
        const int c = NC + l;
        double b[4];
        bounding_plane(l & 3, b);
        const double vtdiv = b[0] * h.dx + b[1] * h.dy + b[2] * h.dz;
        const double vt = -(b[0] * h.px + b[1] * h.py + b[2] * h.pz + b[3]) / vtdiv;
        const bool ok = l < 4 && c != h.A && c != h.B && c != h.oldC && vtdiv != 0.0 && vt > 0.0;
        const unsigned long long okm = ballot(ok);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (!((okm >> k) & 1ull)) continue;
            const double v = readlane_d(vt, k);
            if (uni(v < vtmin ? 1 : 0)) {
                vtmin = v;
                w.C = NC + k;
            }
        }
        w.nx = h.px + vtmin * h.dx;
        w.ny = h.py + vtmin * h.dy;
        w.nz = h.pz + vtmin * h.dz;
    }
    w.C = uni(w.C);
    return w;
}

__device__ __forceinline__ void put_vertex(Smem& s, int vn, double x, double y, double z, int a, int b, int c)
{
    if (lane_id() == 0) {
        s.vx[vn] = x;
        s.vy[vn] = y;
        s.vz[vn] = z;
        s.vpl[vn] = pack_planes(a, b, c);
    }
}

// ------------------------------------------------------------ pose build
// One atom from its internal coordinates: the plane of reference atoms 1-3,
// the bond of length dis along 1 -> 2 turned by the angle and the dihedral
// (LIB/buildcc.cpp:40-120, FP32 in the reference's operation order)
__device__ __forceinline__ void place_atom(const float* x, const float* y, const float* z, float dis, float cta,
                                           float sta, float ctd, float std_, float& ox, float& oy, float& oz)
{
    float A = y[1] * (z[2] - z[3]) + y[2] * (z[3] - z[1]) + y[3] * (z[1] - z[2]);
    float B = z[1] * (x[2] - x[3]) + z[2] * (x[3] - x[1]) + z[3] * (x[1] - x[2]);
    float C = x[1] * (y[2] - y[3]) + x[2] * (y[3] - y[1]) + x[3] * (y[1] - y[2]);
    float op = sqrtf(A * A + B * B + C * C);
    float cx = A / op, cy = B / op, cz = C / op;
    A = x[2] - x[1];
    B = y[2] - y[1];
    C = z[2] - z[1];
    float d = 1.0f / sqrtf(A * A + B * B + C * C);
    op = dis * d;
    float xn = A * op, yn = B * op, zn = C * op;
    A = cx * cx;
    B = cy * cy;
    C = cz * cz;
    float ct = cta;
    float st = -sta;
    op = 1.0f - ct;
    float xk = (cx * cz * op - cy * st) * zn + ((1.0f - A) * ct + A) * xn + (cx * cy * op + cz * st) * yn;
    float yk = (cy * cx * op - cz * st) * xn + ((1.0f - B) * ct + B) * yn + (cy * cz * op + cx * st) * zn;
    float zk = (cz * cy * op - cx * st) * yn + ((1.0f - C) * ct + C) * zn + (cz * cx * op + cy * st) * xn;
    ct = ctd;
    st = std_;
    op = 1.0f - ct;
    cx = (x[2] - x[1]) * d;
    cy = (y[2] - y[1]) * d;
    cz = (z[2] - z[1]) * d;
    A = cx * cx;
    B = cy * cy;
    C = cz * cz;
    ox = (cx * cz * op - cy * st) * zk + ((1.0f - A) * ct + A) * xk + (cx * cy * op + cz * st) * yk + x[1];
    oy = (cy * cx * op - cz * st) * xk + ((1.0f - B) * ct + B) * yk + (cy * cz * op + cx * st) * zk + y[1];
    oz = (cz * cy * op - cx * st) * yk + ((1.0f - C) * ct + C) * zk + (cz * cx * op + cy * st) * xk + z[1];
}

// LIB/ic2cf.cpp:53-105 then LIB/buildcc.cpp:6-134 (FP32 like the reference;
// cos/sin are taken in double and rounded, i.e. correctly rounded floats).
// The gene decode is serial (shift chains); the trig runs one atom per lane;
// the rebuild is serial over atoms with the coordinates of atom i held by
// lane i & 63, slot i >> 6.
__device__ void build_pose(Smem& s, const DevReceptor& R, const DevLigand& lig, const double* icv)
{
    const int n = lig.n_atoms;
    const int ns = R.n_sc;  // ligand atom i is scorable atom ns + i
    const int l = lane_id();
    if (l < FA_MAX_FLEX_RESIDUES) s.u.b.rot[l] = 0;
    for (int i = l; i < n; i += WAVE) {
        s.u.b.dis[i] = lig.dis[i];
        s.u.b.ang[i] = lig.ang[i];
        s.u.b.dih[i] = lig.dih[i];
    }
    // the genes' values, clamps, atoms and types on lanes (one load round),
    // then the decode in gene order on lane 0 from those registers
    double gv = 0.0;
    int gatm = 0, gtyp = 0;
    if (l < lig.n_genes) {
        gv = icv[l];
        const double gmax = lig.gene_max[l], gmin = lig.gene_min[l];
        if (gv > gmax) gv = gmax;
        if (gv < gmin) gv = gmin;
        gatm = lig.gene_atm[l];
        gtyp = lig.gene_typ[l];
    }
    wave_sync();
    const int ng = lig.n_genes;
    for (int g = 0; g < ng; ++g) {
        const double v = readlane_d(gv, g);
        const int atm = readlane_i(gatm, g);
        const int typ = readlane_i(gtyp, g);
        if (l == 0) {
            if (typ == FA_GENE_GRID) {
                unsigned gi = (unsigned)v;
                if ((int)gi >= lig.n_grid) gi = lig.n_grid - 1;
                s.u.b.dis[atm] = lig.grid_ic[3 * gi];
                s.u.b.ang[atm] = lig.grid_ic[3 * gi + 1];
                s.u.b.dih[atm] = lig.grid_ic[3 * gi + 2];
            } else if (typ == FA_GENE_DIS) {
                s.u.b.dis[atm] = (float)v;
            } else if (typ == FA_GENE_ANG) {
                s.u.b.ang[atm] = (float)v;
            } else if (typ == FA_GENE_ROT) {  // rot_idx = (uint)(icv + 0.5f), LIB/ic2cf.cpp:90-97
                unsigned r = (unsigned)(v + 0.5f);
                const int nr = R.flex_nrot[atm];
                if ((int)r >= nr) r = (unsigned)(nr - 1);
                s.u.b.rot[atm] = (int)r;
            } else if (typ == FA_GENE_DIH) {
                s.u.b.dih[atm] = (float)v;
                int j = atm, cat = lig.rec[4 * j + 3];
                if (cat >= 0) {
                    for (int guard = 0; cat != atm && guard < n; ++guard) {
                        s.u.b.dih[cat] = s.u.b.dih[j] + lig.shift[cat];
                        j = cat;
                        cat = lig.rec[4 * j + 3];
                    }
                }
            }
        }
    }
    wave_sync();
    float ca[LSL], sa[LSL], cd[LSL], sd[LSL], di[LSL];
    float X[LSL], Y[LSL], Z[LSL];
#pragma unroll
    for (int q = 0; q < LSL; ++q) {
        const int a = l + WAVE * q;
        ca[q] = sa[q] = cd[q] = sd[q] = di[q] = 0.0f;
        X[q] = Y[q] = Z[q] = 0.0f;
        if (a < n) {
            const float angPI = (float)(s.u.b.ang[a] * PI_REF / 180.0f);
            const float dihPI = (float)(s.u.b.dih[a] * PI_REF / 180.0f);
            ca[q] = (float)cos((double)angPI);
            sa[q] = (float)sin((double)angPI);
            cd[q] = (float)cos((double)dihPI);
            sd[q] = (float)sin((double)dihPI);
            di[q] = s.u.b.dis[a];
        }
    }
    static_assert(LSL == 2, "two ligand slots per lane");
    // the rebuild order and the reference atoms of every atom on lanes (atom
    // i: lane i & 63, slot i >> 6), read back by lane broadcast in the loop
    int mv[LSL], r0[LSL], r1[LSL], r2[LSL];
#pragma unroll
    for (int q = 0; q < LSL; ++q) {
        const int i = l + WAVE * q;
        mv[q] = r0[q] = r1[q] = r2[q] = -1;
        if (i < n) {
            mv[q] = lig.mov[i];
            r0[q] = lig.rec[4 * i];
            r1[q] = lig.rec[4 * i + 1];
            r2[q] = lig.rec[4 * i + 2];
        }
    }
    const float ox0 = lig.ori[0], oy0 = lig.ori[1], oz0 = lig.ori[2];
    for (int an = 0; an < n; ++an) {
        const int a = readlane_i(an < WAVE ? mv[0] : mv[1], an & 63);
        const int la = a & 63;
        const bool hi = a >= WAVE;
        const float cta = readlane_f(hi ? ca[1] : ca[0], la);
        const float sta = readlane_f(hi ? sa[1] : sa[0], la);
        const float ctd = readlane_f(hi ? cd[1] : cd[0], la);
        const float std_ = readlane_f(hi ? sd[1] : sd[0], la);
        const float dis = readlane_f(hi ? di[1] : di[0], la);
        const int rj[3] = {readlane_i(hi ? r0[1] : r0[0], la), readlane_i(hi ? r1[1] : r1[0], la),
                           readlane_i(hi ? r2[1] : r2[0], la)};
        float x[4], y[4], z[4];
#pragma unroll
        for (int i = 1; i <= 3; ++i) {
            const int j = rj[i - 1];
            if (j >= 0) {
                const bool hj = j >= WAVE;
                x[i] = readlane_f(hj ? X[1] : X[0], j & 63);
                y[i] = readlane_f(hj ? Y[1] : Y[0], j & 63);
                z[i] = readlane_f(hj ? Z[1] : Z[0], j & 63);
            } else if (i == 1) {
                x[i] = 1.0f + ox0;
                y[i] = 0.0f + oy0;
                z[i] = 0.0f + oz0;
            } else if (i == 3) {
                x[i] = 0.0f + ox0;
                y[i] = 1.0f + oy0;
                z[i] = 0.0f + oz0;
            } else {
                x[i] = 0.0f + ox0;
                y[i] = 0.0f + oy0;
                z[i] = 0.0f + oz0;
            }
            x[i] += 1e-10f;
            y[i] += 1e-10f;
            z[i] += 1e-10f;
        }
        float ox, oy, oz;
        place_atom(x, y, z, dis, cta, sta, ctd, std_, ox, oy, oz);
        if (l == la) {
            if (hi) {
                X[1] = ox;
                Y[1] = oy;
                Z[1] = oz;
            } else {
                X[0] = ox;
                Y[0] = oy;
                Z[0] = oz;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < LSL; ++q) {
        const int a = l + WAVE * q;
        if (a < n) {
            s.lx[ns + a] = X[q];
            s.ly[ns + a] = Y[q];
            s.lz[ns + a] = Z[q];
            s.lr[ns + a] = lig.radius[a];
        }
    }
    // side chains in the conformation their genes select (the atoms
    // fatm[rot]..latm[rot] that index_protein takes, LIB/Vcontacts.cpp:1598-1606)
    for (int m = l; m < ns; m += WAVE) {
        const float4 p = R.sc_rot[R.sc_rot_off[m] + s.u.b.rot[R.sc_res[m]]];
        s.lx[m] = p.x;
        s.ly[m] = p.y;
        s.lz[m] = p.z;
        s.lr[m] = p.w;
    }
    wave_sync();
}

// calc index of scorable atom m: side-chain atoms (m < n_sc) sit among the
// receptor atoms, the ligand's follow the receptor
__device__ __forceinline__ int calc_of(const DevReceptor& R, int m)
{
    return m < R.n_sc ? R.sc_calc[m] : R.n + (m - R.n_sc);
}

// ------------------------------------------------------------ pose grid
struct PoseGrid {
    float gmin[3];
    int dim, dim2, dim3;
    bool slow;   // grid differs from the receptor's precomputed one
};

// LIB/Vcontacts.cpp:1594-1658 (steady state): only scorable atoms widen it
__device__ void pose_grid(const Smem& s, int M, const DevGrid& G, PoseGrid& pg)
{
    float gmin[3], gmax[3];
    for (int j = 0; j < 3; ++j) {
        gmin[j] = G.gmin[j];
        gmax[j] = G.gmax[j];
    }
    int alter = 0;
    for (int i = 0; i < M; ++i) {  // calc order, like the reference
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
}

// packed cell coordinates of a point (fa_cell_pack; dim < 1024 where used)
__device__ __forceinline__ int cell_packed(const PoseGrid& pg, float x, float y, float z)
{
    return fa_cell_pack((int)((x - pg.gmin[0]) / CELLSIZE), (int)((y - pg.gmin[1]) / CELLSIZE),
                        (int)((z - pg.gmin[2]) / CELLSIZE));
}

__device__ __forceinline__ int box_of(const PoseGrid& pg, float x, float y, float z)
{
    return (int)((x - pg.gmin[0]) / CELLSIZE) * pg.dim2 + (int)((y - pg.gmin[1]) / CELLSIZE) * pg.dim +
           (int)((z - pg.gmin[2]) / CELLSIZE);
}

// the 27-slot neighbourhood (cell_slots.h, LIB/Vcontacts.cpp:1738-1741)
__device__ __forceinline__ int slot_box(const PoseGrid& pg, int bz, int slot) { return fa_slot_box(pg.dim, bz, slot); }
__device__ __forceinline__ int slot_of(const PoseGrid& pg, int d) { return fa_slot_of(pg.dim, d); }
__device__ __forceinline__ int slot_multiplicity(const PoseGrid& pg, int d) { return fa_slot_multiplicity(pg.dim, d); }

__device__ __forceinline__ bool bit_of(unsigned long long b0, unsigned long long b1, int m)
{
    return ((m < 64 ? (b0 >> (m & 63)) : (b1 >> (m & 63))) & 1ull) != 0ull;
}

// ------------------------------------------------------------ neighbour cull
// One candidate of the 27-cell scan, in the reference's order (slot, then
// receptor atoms of the cell, then ligand atoms of the cell; calc order).
struct Cand {
    float x, y, z, r;
    int j;      // calc index, -1 when the lane has no candidate
    int m;      // scorable-atom index for scorable candidates
    bool lj;    // a scorable candidate (side-chain or ligand atom)
};

// Per-atom cull state and results (wave-uniform except `clash`).
struct CullAcc {
    int n_near;
    int total, tested, found;   // lane partials until reduced
    double clash;               // lane partial
    bool over;
};

struct CullCtx {
    int acalc, a;
    float ax, ay, az, ra, rado;
    int rlo, rhi;                      // calc range of a's residue (the ligand or a side chain)
    unsigned long long brow0, brow1;   // bonded row of atom a over calc index - rlo
    unsigned long long dpre0, dpre1;   // atoms already marked by the pre-pass
    double permea;
    int2* nidx;
    float4* nxyz;
    int cap;
};

// x^-12 (the reference calls pow(x, -12.0); this agrees to a few ulps)
__device__ __forceinline__ double inv12(double x)
{
    const double x2 = x * x, x4 = x2 * x2, x12 = x4 * x4 * x4;
    return 1.0 / x12;
}

// Tests one candidate: counts, clash term, and whether it is in contact
// range (sq returns the squared distance as the reference computes it).
__device__ __forceinline__ bool cand_test(const CullCtx& cc, const Cand& c, CullAcc& acc, float& sq)
{
    // the tests are evaluated on every lane and combined (no nested
    // divergent branches); only the rare clash term branches
    const bool valid = c.j >= 0 && c.j != cc.acalc;
    const float dx = cc.ax - c.x, dy = cc.ay - c.y, dz = cc.az - c.z;
    const float d2 = dx * dx + dy * dy + dz * dz;
    const double nd = (double)(cc.rado + c.r + RW);
    const bool in = valid && (double)d2 < nd * nd && d2 != 0.0f;
    // LIB/Vcontacts.cpp:36-51 (pre-pass): a ligand atom counted once by
    // the earlier atom in box order, i.e. skipped once it is marked
    const bool pskip = c.lj && bit_of(cc.dpre0, cc.dpre1, c.m);
    const bool tested = valid && !pskip;
    const bool found = tested && in;
    acc.total += valid ? 1 : 0;
    acc.tested += tested ? 1 : 0;
    acc.found += found ? 1 : 0;
    sq = valid ? d2 : 0.0f;
    const double cdist = cc.permea * (double)(cc.ra + c.r);
    // bonded: same residue and within bloops (LIB/Vcontacts.cpp:1813-1816)
    const bool bonded = c.j >= cc.rlo && c.j < cc.rhi && bit_of(cc.brow0, cc.brow1, c.j - cc.rlo);
    // the distance (a square root) only where the wall can apply: the
    // rounded root is below cdist only if d2 < cdist^2, which the widened
    // bound keeps; most chunks have no such lane
    const bool near = found && !bonded && (double)d2 < cdist * cdist * (1.0 + 1e-9);
    if (ballot(near) != 0ull) {
        const double dist = sqrt((double)d2);
        if (near && dist < cdist) acc.clash += (double)KWALL_F * (inv12(dist) - inv12(cdist));
    }
    return in;
}

// Appends the lanes with `in` set to the neighbour list, in lane order.
__device__ __forceinline__ void cand_emit(const CullCtx& cc, const Cand& c, bool in, float sq, CullAcc& acc)
{
    const unsigned long long mk = ballot(in);
    if (in) {
        const int pos = acc.n_near + lanes_below(mk);
        if (pos < cc.cap) {
            // a scorable neighbour is kept as -1 - (its scorable index)
            cc.nidx[pos] = make_int2(c.lj ? -1 - c.m : c.j, __float_as_int(sq));
            cc.nxyz[pos] = make_float4(c.x, c.y, c.z, c.r);
        }
    }
    acc.n_near += __popcll(mk);
    if (acc.n_near > cc.cap) acc.over = true;
}

__device__ __forceinline__ void visit(const CullCtx& cc, const Cand& c, CullAcc& acc)
{
    float sq;
    const bool in = cand_test(cc, c, acc, sq);
    cand_emit(cc, c, in, sq, acc);
}

// Cells of the ligand atoms by rank, which are nondecreasing: ranks l and
// l + 64 on lane l (past the last rank: INT_MAX).
struct RankCells {
    int b0, b1;
};

__device__ __forceinline__ RankCells rank_cells(const Smem& s, int L)
{
    const int l = lane_id();
    return RankCells{l < L ? s.sbox[l] : INT_MAX, l + WAVE < L ? s.sbox[l + WAVE] : INT_MAX};
}

// ligand atoms of the wave-uniform cell sb: ranks [lo, lo + nl)
__device__ __forceinline__ void cell_span(const RankCells& rc, int L, int sb, int& lo, int& nl)
{
    lo = __popcll(ballot(rc.b0 < sb));
    nl = __popcll(ballot(rc.b0 == sb));
    if (L > WAVE) {
        lo += __popcll(ballot(rc.b1 < sb));
        nl += __popcll(ballot(rc.b1 == sb));
    }
}

// Segment table of one atom's stream (lanes 0..26 hold slot 0..26).
struct SlotTab {
    int bi;       // cell of the slot, -1 if outside the grid
    int rs, re;   // receptor range of the cell in the cell list
};

__device__ __forceinline__ SlotTab slot_table(const Smem& s, const DevGrid& G, const PoseGrid& pg, int k)
{
    SlotTab t{-1, 0, 0};
    const int l = lane_id();
    if (l < 27) {
        const int b = slot_box(pg, s.sbox[k], l);
        if (b >= 0 && b < pg.dim3) t.bi = b;
    }
    if (t.bi >= 0 && !pg.slow) {
        t.rs = G.cell_start[t.bi];
        t.re = G.cell_start[t.bi + 1];
    }
    return t;
}

// receptor cells from the precomputed cell list
__device__ void cull_fast(const Smem& s, const DevReceptor& R, const DevGrid& G, const DevLigand& lig, const PoseGrid& pg,
                          int Rn, int L,
                          const SlotTab& tab, int bz, int* span, const CullCtx& cc, CullAcc& acc)
{
    const int l = lane_id();
    int lo = 0, nl = 0;  // ligand atoms of the slot's cell: ranks [lo, lo + nl)
    if (pg.dim >= 3) {
        // one pass over the ranks: with three or more cells a side a cell
        // offset names at most one slot, and the ranks of a cell are
        // consecutive (count and first rank per slot, LDS atomics)
        if (l < 27) {
            span[l] = 0;
            span[32 + l] = INT_MAX;
        }
        wave_sync();
        for (int r = l; r < L; r += WAVE) {
            const int sl = slot_of(pg, s.sbox[r] - bz);
            if (sl >= 0) {
                atomicAdd(&span[sl], 1);
                atomicMin(&span[32 + sl], r);
            }
        }
        wave_sync();
        if (l < 27 && tab.bi >= 0) {
            nl = span[l];
            lo = nl > 0 ? span[32 + l] : 0;
        }
        wave_sync();  // the slots are read before the next atom clears them
    } else {
        // coarse grids: a cell can be named by several slots
        const RankCells rc = rank_cells(s, L);
        for (unsigned long long m = ballot(tab.bi >= 0); m; m &= m - 1) {
            const int q = first_lane(m);
            int a, n;
            cell_span(rc, L, readlane_i(tab.bi, q), a, n);
            if (l == q) {
                lo = a;
                nl = n;
            }
        }
    }
    const int nr = (tab.bi >= 0) ? tab.re - tab.rs : 0;
    const int seg = nr + nl;
    const int off = wave_excl_scan(seg);
    const int T = readlane_i(off, 63);  // lanes >= 27 hold empty segments
    // per segment (lane < 27): stream position t maps to receptor cell-list
    // entry t + rbase while t < rend, else to ligand rank t + lbase
    const int rbase = tab.rs - off, rend = off + nr, lbase = lo - off - nr;
    auto fetch = [&](int base) {
        Cand c{0.f, 0.f, 0.f, 0.f, -1, 0, false};
        const int t = base + l;
        // the segment of t: the last non-empty one starting at or before t
        // among those overlapping the chunk, then its fields by lane shuffle
        unsigned long long ov = ballot(l < 27 && seg > 0 && off < base + WAVE && off + seg > base);
        int sidx = 0;
        while (ov) {
            const int q = first_lane(ov);
            ov &= ov - 1;
            if (t >= readlane_i(off, q)) sidx = q;
        }
        const int rb = __shfl(rbase, sidx), re = __shfl(rend, sidx), lb = __shfl(lbase, sidx);
        if (t < T) {
            if (t < re) {
                const int k = t + rb;
                const float4 p = G.cell_xyzr[k];
                c.x = p.x;
                c.y = p.y;
                c.z = p.z;
                c.r = p.w;
                c.j = G.cell_atom[k];
            } else {
                const int m = s.order[t + lb];
                c.m = m;
                c.x = s.lx[m];
                c.y = s.ly[m];
                c.z = s.lz[m];
                c.r = s.lr[m];
                c.j = calc_of(R, m);
                c.lj = true;
            }
        }
        return c;
    };
    if (T <= 0) return;
    // four chunks in flight: the loads of the next ones are issued before a
    // chunk is tested, so an atom waits about once for L2 instead of per chunk
    const Cand none{0.f, 0.f, 0.f, 0.f, -1, 0, false};
    Cand c0 = fetch(0), c1 = none, c2 = none, c3 = none;
    if (WAVE < T) c1 = fetch(WAVE);
    if (2 * WAVE < T) c2 = fetch(2 * WAVE);
    if (3 * WAVE < T) c3 = fetch(3 * WAVE);
    for (int base = 0; base < T; base += 4 * WAVE) {
        visit(cc, c0, acc);
        if (base + 4 * WAVE < T) c0 = fetch(base + 4 * WAVE);
        if (base + WAVE >= T) break;
        visit(cc, c1, acc);
        if (base + 5 * WAVE < T) c1 = fetch(base + 5 * WAVE);
        if (base + 2 * WAVE >= T) break;
        visit(cc, c2, acc);
        if (base + 6 * WAVE < T) c2 = fetch(base + 6 * WAVE);
        if (base + 3 * WAVE >= T) break;
        visit(cc, c3, acc);
        if (base + 7 * WAVE < T) c3 = fetch(base + 7 * WAVE);
    }
}

// Inclusive prefix minimum over the wave (all lanes active).
template <int CTRL>
__device__ __forceinline__ int dpp_min_shr(int v)
{
    const int o = __builtin_amdgcn_update_dpp(INT_MAX, v, CTRL, 0xF, 0xF, false);  // no source: INT_MAX
    return v < o ? v : o;
}

__device__ __forceinline__ int wave_incl_min(int v)
{
    v = dpp_min_shr<0x111>(v);  // row_shr:1
    v = dpp_min_shr<0x112>(v);  // row_shr:2
    v = dpp_min_shr<0x114>(v);  // row_shr:4
    v = dpp_min_shr<0x118>(v);  // row_shr:8
    const int r0 = readlane_i(v, 15), r1 = readlane_i(v, 31), r2 = readlane_i(v, 47);
    const int l = lane_id();
    if (l >= 16) v = v < r0 ? v : r0;
    if (l >= 32) v = v < r1 ? v : r1;
    if (l >= 48) v = v < r2 ? v : r2;
    return v;
}

// Residue of scorable atom a for the pre-pass, by vector loads (lane 0 and
// 1: the bonded row's two words over calc index - lo, lane 2: lo | hi << 32)
// so that it can be fetched an atom ahead without holding up LDS waits.
__device__ __forceinline__ uint64_t residue_fetch(const DevReceptor& R, const DevLigand& lig, int a)
{
    const int l = lane_id();
    const int nsc = R.n_sc;
    if (a >= nsc && l == 2) return (uint64_t)(uint32_t)R.n | ((uint64_t)(uint32_t)(R.n + lig.n_atoms) << 32);
    // a lane-dependent pointer: the load is a vector load
    const unsigned long long* p = a < nsc ? (l < 2 ? R.sc_bonded + 2 * a + l
                                                    : reinterpret_cast<const unsigned long long*>(R.sc_range + a))
                                          : lig.bonded + 2 * (a - nsc) + (l & 1);
    return l < 3 ? (uint64_t)*p : 0ull;  // int2 {lo, hi} reads as lo | hi << 32
}

// The neighbourhood list of one atom's cell (grids with neighbourhood
// lists): wave-uniform start and length, and on lane s < 27 where slot s's
// receptor atoms end in it.
struct NbTab {
    int se;    // lane 0: the list's start, lane 1: its end
    int bend;
};

// Vector loads only (lane-varying addresses): a scalar load would share the
// LDS wait counter and hold the atom's first LDS access until it returns.
__device__ __forceinline__ NbTab nb_table(const Smem& s, const DevGrid& G, int k)
{
    const int b = uni(s.sbox[k]);
    const int l = lane_id();
    NbTab t;
    t.se = G.nb_start[b + (l & 1)];
    t.bend = l < 27 ? G.nb_bnd[28 * b + l + 1] : INT_MAX;
    return t;
}

// The 27-cell scan of one atom over the cell's neighbourhood list (the
// receptor atoms, slot by slot) and the scorable atoms, emitted in the
// reference's order: per slot its receptor atoms, then its scorable atoms
// in rank order.  With three or more cells a side a slot names one cell and
// the slots' cells increase with the slot, so the ranks (sorted by cell)
// list the scorable candidates in scan order.  Positions follow from two
// per-slot prefix counts: in-range scorable candidates of the slots before
// (lin, built from the first in-range candidate of each slot and a suffix
// minimum) and in-range receptor atoms of the slots up to each slot (rin,
// read off the receptor chunks at the slot ends).
__device__ void cull_nb(const Smem& s, const DevReceptor& R, const DevGrid& G, const PoseGrid& pg, int L, int k,
                        const NbTab& tab, int* tbl, const CullCtx& cc, CullAcc& acc)
{
    const int l = lane_id();
    const int base0 = acc.n_near;
    const int kc = uni(s.u.p.rcell[k]);
    // -- scorable candidates, in rank order
    int sig[LSL];
    float ssq[LSL];
    unsigned long long im[LSL];
#pragma unroll
    for (int q = 0; q < LSL; ++q) {
        const int r = l + WAVE * q;
        Cand c{0.f, 0.f, 0.f, 0.f, -1, 0, false};
        sig[q] = -1;
        if (q * WAVE < L && r < L) {
            const int sl = fa_slot_of_cells(pg.dim, kc, s.u.p.rcell[r]);
            if (sl >= 0) {
                const int m = s.order[r];
                c.m = m;
                c.x = s.lx[m];
                c.y = s.ly[m];
                c.z = s.lz[m];
                c.r = s.lr[m];
                c.j = s.u.p.calc[m];
                c.lj = true;
                sig[q] = sl;
            }
        }
        ssq[q] = 0.0f;
        const bool in = (q * WAVE < L) ? cand_test(cc, c, acc, ssq[q]) : false;
        im[q] = ballot(in);
    }
    static_assert(LSL == 2, "two rank slots per lane");
    const int n0 = __popcll(im[0]);
    const int nL = n0 + __popcll(im[1]);
    // first in-range candidate of each slot: an in-range candidate whose
    // previous in-range candidate (in rank order) has a smaller slot
    if (l < 32) tbl[l] = nL;
    wave_sync();
#pragma unroll
    for (int q = 0; q < LSL; ++q) {
        const bool in = (im[q] >> l) & 1ull;
        const unsigned long long below = im[q] & ((1ull << l) - 1ull);
        int prev_lane = -1, prev_q = q;
        if (below) {
            prev_lane = 63 - __clzll((long long)below);
        } else if (q == 1 && im[0]) {
            prev_lane = 63 - __clzll((long long)im[0]);
            prev_q = 0;
        }
        const int p0 = __shfl(sig[0], prev_lane < 0 ? 0 : prev_lane);
        const int p1 = __shfl(sig[1], prev_lane < 0 ? 0 : prev_lane);
        const int psig = prev_lane < 0 ? -1 : (prev_q == 0 ? p0 : p1);
        if (in && psig < sig[q]) tbl[sig[q]] = (q ? n0 : 0) + lanes_below(im[q]);
    }
    wave_sync();
    // lin(s) = in-range scorable candidates of slots < s: the first index of
    // slot s or of the next slot that has one (a suffix minimum, run as a
    // prefix minimum over lanes holding slots 26 - l)
    const int lin_rev = wave_incl_min(l < 27 ? tbl[26 - l] : nL);
    wave_sync();  // the table is read before the next atom writes it

    // -- receptor atoms of the 27 slots
    const int nb0 = readlane_i(tab.se, 0);
    const int T = readlane_i(tab.se, 1) - nb0;
    int nR = 0;
    int rin = -1;  // lane s < 27: in-range receptor atoms of slots <= s
    auto fetch = [&](int b, float4& p, int& a) {
        const int t = b + l;
        a = 0;
        p = make_float4(0.f, 0.f, 0.f, 0.f);
        if (t < T) {
            p = G.nb_xyzr[nb0 + t];
            a = G.nb_atom[nb0 + t];
        }
    };
    auto visit_r = [&](int b, const float4& p, int a) {
        const bool ok = b + l < T;  // entries pack calc index | slot << 27 (any sign)
        Cand c{p.x, p.y, p.z, p.w, ok ? (a & 0x7FFFFFF) : -1, 0, false};
        float sq;
        const bool in = cand_test(cc, c, acc, sq);
        const unsigned long long mk = ballot(in);
        const int sl = ok ? (int)((unsigned)a >> 27) : 0;
        const int lin = __shfl(lin_rev, 26 - sl);
        if (in) {
            const int pos = base0 + nR + lanes_below(mk) + lin;
            if (pos < cc.cap) {
                cc.nidx[pos] = make_int2(c.j, __float_as_int(sq));
                cc.nxyz[pos] = make_float4(c.x, c.y, c.z, c.r);
            }
        }
        const int e = tab.bend - b;
        if (e >= 0 && e < WAVE) rin = nR + __popcll(mk & ((1ull << e) - 1ull));
        nR += __popcll(mk);
    };
    if (T > 0) {
        // four chunks in flight, as the cell-list scan
        float4 p0, p1 = make_float4(0.f, 0.f, 0.f, 0.f), p2 = p1, p3 = p1;
        int a0, a1 = 0, a2 = 0, a3 = 0;
        fetch(0, p0, a0);
        if (WAVE < T) fetch(WAVE, p1, a1);
        if (2 * WAVE < T) fetch(2 * WAVE, p2, a2);
        if (3 * WAVE < T) fetch(3 * WAVE, p3, a3);
        for (int b = 0; b < T; b += 4 * WAVE) {
            visit_r(b, p0, a0);
            if (b + 4 * WAVE < T) fetch(b + 4 * WAVE, p0, a0);
            if (b + WAVE >= T) break;
            visit_r(b + WAVE, p1, a1);
            if (b + 5 * WAVE < T) fetch(b + 5 * WAVE, p1, a1);
            if (b + 2 * WAVE >= T) break;
            visit_r(b + 2 * WAVE, p2, a2);
            if (b + 6 * WAVE < T) fetch(b + 6 * WAVE, p2, a2);
            if (b + 3 * WAVE >= T) break;
            visit_r(b + 3 * WAVE, p3, a3);
            if (b + 7 * WAVE < T) fetch(b + 7 * WAVE, p3, a3);
        }
    }
    if (rin < 0) rin = nR;  // slots ending at the list's end

    // -- the scorable candidates in range, after the receptor atoms of their slot
#pragma unroll
    for (int q = 0; q < LSL; ++q) {
        const bool in = (im[q] >> l) & 1ull;
        const int rs = __shfl(rin, sig[q] < 0 ? 0 : sig[q]);
        if (in) {
            const int pos = base0 + rs + (q ? n0 : 0) + lanes_below(im[q]);
            if (pos < cc.cap) {
                const int m = s.order[l + WAVE * q];
                cc.nidx[pos] = make_int2(-1 - m, __float_as_int(ssq[q]));
                cc.nxyz[pos] = make_float4(s.lx[m], s.ly[m], s.lz[m], s.lr[m]);
            }
        }
    }
    acc.n_near = base0 + nR + nL;
    if (acc.n_near > cc.cap) acc.over = true;
}

// Ligand atoms of cell sb (ranks [lo, lo + nl) of the box order), visited.
__device__ __forceinline__ void visit_ligand_cell(const Smem& s, const DevReceptor& R, const DevLigand& lig, int Rn, int L,
                                                  const RankCells& rc, int sb, const CullCtx& cc, CullAcc& acc)
{
    int lo, nl;
    cell_span(rc, L, sb, lo, nl);
    for (int b = 0; b < nl; b += WAVE) {
        const int r = b + lane_id();
        Cand c{0.f, 0.f, 0.f, 0.f, -1, 0, false};
        if (r < nl) {
            const int m = s.order[lo + r];
            c.m = m;
            c.x = s.lx[m];
            c.y = s.ly[m];
            c.z = s.lz[m];
            c.r = s.lr[m];
            c.j = calc_of(R, m);
            c.lj = true;
        }
        visit(cc, c, acc);
    }
}

constexpr int SLOW_TMP = 3 * WAVE;  // in-range receptor atoms of one ligand atom (slow path)
// receptor rows of the widened grid that no slot names (side-chain atoms)
constexpr int FAR_CELL = -1;

// slot of packed cell rc around packed cell kc, -1 for FAR_CELL
__device__ __forceinline__ int slot_rc(int dim, int kc, int rc) { return rc == FAR_CELL ? -1 : fa_slot_of_cells(dim, kc, rc); }

// Per-wave slow-path scratch: the cell of every receptor atom in the pose's
// widened grid, then the receptor atoms of one ligand atom's 27 cells.
__device__ __forceinline__ size_t slow_rows(int Rn) { return (size_t)((Rn + 3) & ~3); }

// Grid widened by this pose: the receptor cells are not precomputed.  The
// packed cell of every receptor atom (computed once per pose) is streamed
// four per lane and the atoms in the 27 cells around the ligand atom's cell
// gathered in index order with their slot -- once per ligand cell, since
// the atoms of a cell are consecutive in box order and share the gather
// (`gcell` and `ncand` carry it from one atom to the next).  The gathered
// atoms are tested (the in-range entries parked, with their slot, at the
// tail of the wave's neighbour buffer) and a last pass emits them slot by
// slot, each slot's ligand atoms after its receptor atoms, as the reference
// scans.
__device__ void cull_slow(const Smem& s, const DevReceptor& R, const DevLigand& lig, const PoseGrid& pg, int L,
                          const SlotTab& tab, int kc, const CullCtx& cc, CullAcc& acc, const int* rbox, int* rcand,
                          int& gcell, int& ncand)
{
    const int l = lane_id();
    const int Rn = R.n;
    if (kc != gcell) {
        const int Rq = (int)(slow_rows(Rn) >> 2);
        const int4* const rbox4 = reinterpret_cast<const int4*>(rbox);
        ncand = 0;
        int4 nxt = make_int4(-1, -1, -1, -1);
        if (l < Rq) nxt = rbox4[l];
        for (int base = 0; base < Rq; base += WAVE) {
            const int4 bx = nxt;
            if (base + WAVE < Rq) nxt = (base + WAVE + l < Rq) ? rbox4[base + WAVE + l] : make_int4(-1, -1, -1, -1);
            const int t0 = 4 * (base + l);
            // slots of the lane's four atoms (cells of rows past Rn are never read)
            const int s0 = (t0 < Rn) ? slot_rc(pg.dim, kc, bx.x) : -1;
            const int s1 = (t0 + 1 < Rn) ? slot_rc(pg.dim, kc, bx.y) : -1;
            const int s2 = (t0 + 2 < Rn) ? slot_rc(pg.dim, kc, bx.z) : -1;
            const int s3 = (t0 + 3 < Rn) ? slot_rc(pg.dim, kc, bx.w) : -1;
            const int cnt = (s0 >= 0) + (s1 >= 0) + (s2 >= 0) + (s3 >= 0);
            if (ballot(cnt > 0) == 0ull) continue;
            int pos = ncand + wave_excl_scan(cnt);
            if (s0 >= 0) rcand[pos++] = t0 | (s0 << 26);
            if (s1 >= 0) rcand[pos++] = (t0 + 1) | (s1 << 26);
            if (s2 >= 0) rcand[pos++] = (t0 + 2) | (s2 << 26);
            if (s3 >= 0) rcand[pos++] = (t0 + 3) | (s3 << 26);
            ncand += wave_sum_i(cnt);
        }
        gcell = kc;
        wave_sync();  // gathered atoms visible to every lane
    }
    int2* const tidx = cc.nidx + cc.cap;   // SLOW_TMP entries past the list capacity
    float4* const txyz = cc.nxyz + cc.cap;
    int ntmp = 0;
    // the gathered atoms in chunks, the next chunk's loads in flight
    int v = (l < ncand) ? rcand[l] : -1;
    float4 p = (v >= 0) ? R.xyzr[v & ((1 << 26) - 1)] : make_float4(0.f, 0.f, 0.f, 0.f);
    for (int base = 0; base < ncand; base += WAVE) {
        const int cv = v;
        const float4 cp = p;
        if (base + WAVE < ncand) {
            v = (base + WAVE + l < ncand) ? rcand[base + WAVE + l] : -1;
            p = (v >= 0) ? R.xyzr[v & ((1 << 26) - 1)] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
        const int t = (cv >= 0) ? (cv & ((1 << 26) - 1)) : -1;
        Cand c{cp.x, cp.y, cp.z, cp.w, t, 0, false};
        float sq;
        const bool in = cand_test(cc, c, acc, sq);
        const unsigned long long mk = ballot(in);
        if (in) {
            const int q = ntmp + lanes_below(mk);
            if (q < SLOW_TMP) {
                tidx[q] = make_int2(cv, __float_as_int(sq));
                txyz[q] = cp;
            }
        }
        ntmp += __popcll(mk);
    }
    if (ntmp > SLOW_TMP) {
        acc.over = true;
        ntmp = SLOW_TMP;
    }
    wave_sync();  // parked entries visible to every lane
    int2 ti[SLOW_TMP / WAVE];
    float4 tx[SLOW_TMP / WAVE];
    int tslot[SLOW_TMP / WAVE];
#pragma unroll
    for (int q = 0; q < SLOW_TMP / WAVE; ++q) {
        const int e = q * WAVE + l;
        ti[q] = make_int2(-1, 0);
        tx[q] = make_float4(0.f, 0.f, 0.f, 0.f);
        tslot[q] = -1;
        if (e < ntmp) {
            ti[q] = tidx[e];
            tx[q] = txyz[e];
            tslot[q] = ti[q].x >> 26;
            ti[q].x &= (1 << 26) - 1;
        }
    }
    const RankCells rc = rank_cells(s, L);
    for (int slot = 0; slot < 27; ++slot) {
        const int sb = readlane_i(tab.bi, slot);
        if (sb < 0) continue;
#pragma unroll
        for (int q = 0; q < SLOW_TMP / WAVE; ++q) {
            const bool mine = tslot[q] == slot;
            if (ballot(mine) == 0ull) continue;
            const Cand c{tx[q].x, tx[q].y, tx[q].z, tx[q].w, ti[q].x, 0, false};
            cand_emit(cc, c, mine, __int_as_float(ti[q].y), acc);
        }
        visit_ligand_cell(s, R, lig, Rn, L, rc, sb, cc, acc);
    }
}

// Coarse grids (fewer than three cells a side) name a cell from several
// slots; there every slot is scanned over all receptor atoms.
__device__ void cull_slow_coarse(const Smem& s, const DevReceptor& R, const DevLigand& lig, const PoseGrid& pg, int L,
                                 const SlotTab& tab, const CullCtx& cc, CullAcc& acc)
{
    const int l = lane_id();
    const int Rn = R.n;
    const RankCells rc = rank_cells(s, L);
    for (int slot = 0; slot < 27; ++slot) {
        const int sb = readlane_i(tab.bi, slot);
        if (sb < 0) continue;
        for (int base = 0; base < Rn; base += WAVE) {
            const int t = base + l;
            Cand c{0.f, 0.f, 0.f, 0.f, -1, 0, false};
            bool v = false;
            if (t < Rn && !(R.sc_flag && R.sc_flag[t])) {  // side-chain atoms come with the pose
                const float4 p = R.xyzr[t];
                v = box_of(pg, p.x, p.y, p.z) == sb;
                if (v) c = Cand{p.x, p.y, p.z, p.w, t, 0, false};
            }
            if (ballot(v) == 0ull) continue;
            visit(cc, c, acc);
        }
        visit_ligand_cell(s, R, lig, Rn, L, rc, sb, cc, acc);
    }
}

// ------------------------------------------------------------ kernel
__global__ __launch_bounds__(WAVE, 2) void cf_pose_kernel(DevReceptor R, const DevGrid* __restrict__ grids,
                                                           const DevLigand* __restrict__ ligs, DevParams prm,
                                                           DevBatch B)
{
    __shared__ Smem s;
    const int l = lane_id();
    int2* const nidx = B.near_idx + (size_t)blockIdx.x * B.near_cap;
    float4* const nxyz = B.near_xyz + (size_t)blockIdx.x * B.near_cap;
    if (B.prof && l < N_PHASES) s.phase[l] = 0ull;
    for (;;) {
        // dynamic work queue: pose costs vary a lot (clash exits, hull sizes)
        int pose = 0;
        if (l == 0) pose = (int)atomicAdd(&B.counters[2], 1u);
        pose = __builtin_amdgcn_readfirstlane(pose);
        if (pose >= B.n_poses) break;
        const int li = uni(B.pose_lig ? B.pose_lig[pose] : 0);
        if (li < 0 || li >= B.n_ligands) {  // a ligand the context does not hold: not scored
            if (l == 0) {
                fa_cf_result bad{};
                bad.cf = bad.app = __longlong_as_double(0x7ff0000000000000ll);
                bad.status = FA_BAD_LIGAND;
                B.out[pose] = bad;
                if (B.rec_count) B.rec_count[pose] = 0;
                atomicAdd(&B.totals[3], 1ull);
            }
            continue;
        }
        const DevLigand lig = ligs[li];
        const DevGrid G = grids[lig.grid_id];
        const int NSC = R.n_sc;          // scorable side-chain atoms come first
        const int L = NSC + lig.n_atoms;  // scorable atoms of the pose
        const int Rn = R.n;
        // residue of scorable atom a: calc range and bonded row over calc - lo
        auto residue_of = [&](int a, int& lo, int& hi, unsigned long long& b0, unsigned long long& b1) {
            if (a < NSC) {
                const int2 rg = R.sc_range[a];
                lo = rg.x;
                hi = rg.y;
                b0 = R.sc_bonded[2 * a];
                b1 = R.sc_bonded[2 * a + 1];
            } else {
                lo = Rn;
                hi = Rn + lig.n_atoms;
                b0 = lig.bonded[2 * (a - NSC)];
                b1 = lig.bonded[2 * (a - NSC) + 1];
            }
        };
        int status = FA_OK;
        double r_com = 0.0, r_wal = 0.0, r_sas = 0.0, r_tot = 0.0, r_con = 0.0;
        int r_rclash = 0;
        int nrec_total = 0;
        double clash = 0.0;

        if (l < FA_OPCOUNT_N) s.ops[l] = 0ull;
        PhaseClock clk{B.prof ? s.phase : nullptr, 0ull};
        clk.start();
        const unsigned long long pose_t0 = B.prof ? __builtin_readcyclecounter() : 0ull;
        build_pose(s, R, lig, B.genes + (size_t)pose * B.gene_stride);
        count_op(s.ops, OP_BUILD, lig.n_atoms);
        if (B.lig_xyz) {
            for (int i = l; i < lig.n_atoms; i += WAVE) {
                float* o = B.lig_xyz + ((size_t)pose * B.xyz_stride + i) * 3;
                o[0] = s.lx[NSC + i];
                o[1] = s.ly[NSC + i];
                o[2] = s.lz[NSC + i];
            }
        }
        PoseGrid pg;
        pose_grid(s, L, G, pg);
        if (pg.slow && l == 0) {
            atomicAdd(&B.counters[0], 1u);
            atomicAdd(&B.totals[0], 1ull);
        }
        // cells of the ligand atoms and the processing order (cell, calc index)
        int lb[LSL];
        bool bad = false;
#pragma unroll
        for (int q = 0; q < LSL; ++q) {
            const int i = l + WAVE * q;
            lb[q] = 0x7fffffff;
            if (i < L) {
                lb[q] = box_of(pg, s.lx[i], s.ly[i], s.lz[i]);
                if (lb[q] < 0 || lb[q] >= pg.dim3) bad = true;
            }
        }
        const bool off_grid = ballot(bad) != 0ull;
        if (pg.dim >= 1024) {  // cell coordinates are packed in 10 bits
            status = FA_CAPACITY;
        } else if (off_grid) {
            status = FA_GRID_RANGE;
        } else {
            int rk[LSL] = {0, 0};
            for (int k = 0; k < L; ++k) {
                const int bk = readlane_i(k < WAVE ? lb[0] : lb[1], k & 63);
#pragma unroll
                for (int q = 0; q < LSL; ++q) {
                    const int i = l + WAVE * q;
                    rk[q] += (bk < lb[q]) || (bk == lb[q] && k < i);
                }
            }
#pragma unroll
            for (int q = 0; q < LSL; ++q) {
                const int i = l + WAVE * q;
                if (i < L) {
                    s.order[rk[q]] = (unsigned char)i;
                    s.rank[i] = (unsigned char)rk[q];
                    s.sbox[rk[q]] = lb[q];
                    s.u.p.rcell[rk[q]] = cell_packed(pg, s.lx[i], s.ly[i], s.lz[i]);
                    s.u.p.calc[i] = calc_of(R, i);
                }
            }
        }
        wave_sync();
        clk.stop(PH_BUILD);

        // ---- neighbour lists + clash pre-pass (LIB/Vcontacts.cpp:36-51, :1725-1839):
        // every receptor contact and every ligand pair once, non-bonded,
        // d < permeability * (rA + rB)
        int totq[LSL] = {0, 0};  // candidates of rank k (self excluded): lane k & 63, slot k >> 6
        if (status == FA_OK) {
            unsigned long long dpre0 = 0ull, dpre1 = 0ull;
            long long pre_tested = 0;
            CullAcc acc{0, 0, 0, 0, 0.0, false};
            // widened grid: every receptor atom's cell once for the pose
            int* const wb = B.rbox + 2 * slow_rows(R.n) * blockIdx.x;
            int* const rcand = wb + slow_rows(R.n);
            if (pg.slow && pg.dim >= 3) {
                for (int t = l; t < R.n; t += WAVE) {
                    const float4 p = R.xyzr[t];
                    // packed cell; side-chain atoms come with the pose: a cell no slot names
                    wb[t] = (R.sc_flag && R.sc_flag[t]) ? FAR_CELL : cell_packed(pg, p.x, p.y, p.z);
                }
                wave_sync();
            }
            // neighbourhood lists when the pose keeps the receptor's grid
            const bool use_nb = !pg.slow && G.nb_start != nullptr && !B.cell_lists;
            int gcell = -1, ncand = 0;  // slow path: the cell whose 27-cell gather rcand holds
            SlotTab tab{-1, 0, 0};
            NbTab ntab{0, INT_MAX};
            uint64_t rres = residue_fetch(R, lig, uni(s.order[0]));
            if (use_nb)
                ntab = nb_table(s, G, 0);
            else
                tab = slot_table(s, G, pg, 0);
            for (int k = 0; k < L; ++k) {
                SlotTab next{-1, 0, 0};
                NbTab nnext{0, INT_MAX};
                uint64_t rnext = 0ull;
                if (k + 1 < L) {  // in flight during this atom
                    if (use_nb)
                        nnext = nb_table(s, G, k + 1);
                    else
                        next = slot_table(s, G, pg, k + 1);
                    rnext = residue_fetch(R, lig, uni(s.order[k + 1]));
                }
                const int a = uni(s.order[k]);
                CullCtx cc;
                cc.a = a;
                cc.acalc = uni(s.u.p.calc[a]);
                cc.ax = s.lx[a];
                cc.ay = s.ly[a];
                cc.az = s.lz[a];
                cc.ra = s.lr[a];
                cc.rado = cc.ra + RW;
                cc.brow0 = readlane_u64(rres, 0);
                cc.brow1 = readlane_u64(rres, 1);
                {
                    const uint64_t rr = readlane_u64(rres, 2);
                    cc.rlo = (int)(uint32_t)rr;
                    cc.rhi = (int)(uint32_t)(rr >> 32);
                }
                cc.dpre0 = dpre0;
                cc.dpre1 = dpre1;
                cc.permea = (double)prm.permeability;
                cc.nidx = nidx;
                cc.nxyz = nxyz;
                cc.cap = B.near_cap - SLOW_TMP;  // the tail parks slow-path entries
                if (l == 0) s.nstart[k] = acc.n_near;
                acc.total = acc.tested = acc.found = 0;
                if (use_nb)
                    cull_nb(s, R, G, pg, L, k, ntab, s.u.p.span, cc, acc);
                else if (!pg.slow)
                    cull_fast(s, R, G, lig, pg, Rn, L, tab, s.sbox[k], s.u.p.span, cc, acc);
                else if (pg.dim >= 3)
                    cull_slow(s, R, lig, pg, L, tab, uni(s.u.p.rcell[k]), cc, acc, wb, rcand, gcell, ncand);
                else
                    cull_slow_coarse(s, R, lig, pg, L, tab, cc, acc);
                const int total = wave_sum_i(acc.total);
                const int tested = wave_sum_i(acc.tested);
                const int found = wave_sum_i(acc.found);
                pre_tested += tested;
                if (found > 0) {  // marked once it has a contact
                    if (a < WAVE)
                        dpre0 |= 1ull << a;
                    else
                        dpre1 |= 1ull << (a - WAVE);
                }
                if (l == (k & 63)) {
                    if (k < WAVE)
                        totq[0] = total;
                    else
                        totq[1] = total;
                }
                tab = next;
                ntab = nnext;
                rres = rnext;
            }
            if (l == 0) s.nstart[L] = acc.n_near;
            count_op(s.ops, OP_CULL, pre_tested);
            clash = wave_sum_d(acc.clash);
            // the clash sum is complete even when the neighbour list overflowed
            // (every candidate is tested), so a clash is decided first
            if (clash >= CLASH_THRESHOLD)
                status = FA_CLASH;
            else if (acc.over)
                status = FA_CAPACITY;
        }
        wave_sync();
        clk.stop(PH_PREPASS);

        if (status == FA_OK) {
            // ---- contact polyhedra in box order (calc_region, LIB/Vcontacts.cpp:74-139)
            for (int i = l; i < L; i += WAVE) {
                s.seed[3 * i] = -1;
                s.stub[2 * i] = 0ull;
                s.stub[2 * i + 1] = 0ull;
            }
            wave_sync();
            unsigned long long dm0 = 0ull, dm1 = 0ull;  // atoms done in this pass
            const double permea = (double)prm.permeability;
            for (int k = 0; k < L && status == FA_OK; ++k) {
                const int a = s.order[k];
                const int acalc = calc_of(R, a);
                const float ax = s.lx[a], ay = s.ly[a], az = s.lz[a];
                const float ra = s.lr[a];
                const float rado = ra + RW;
                const double ro = (double)rado;
                clk.stop(PH_REC);
                // -- contact list (get_contlist4, LIB/Vcontacts.cpp:1725-1839): the
                // neighbour list without done atoms, except those that recorded
                // a contact with `a` (the 'C' marks)
                const unsigned long long sk0 = dm0 & ~s.stub[2 * a];
                const unsigned long long sk1 = dm1 & ~s.stub[2 * a + 1];
                const int ns = s.nstart[k], ne = s.nstart[k + 1];
                int NC = 0;
                for (int base = ns; base < ne; base += WAVE) {
                    const int q = base + l;
                    bool keep = false;
                    int j = -1, mv = -1;
                    float sq = 0.0f;
                    float4 p = make_float4(0.f, 0.f, 0.f, 0.f);
                    if (q < ne) {
                        const int2 e = nidx[q];
                        p = nxyz[q];
                        mv = e.x < 0 ? -1 - e.x : -1;  // scorable neighbours are stored as -1 - index
                        j = mv >= 0 ? calc_of(R, mv) : e.x;
                        sq = __int_as_float(e.y);
                        keep = mv < 0 || !bit_of(sk0, sk1, mv);
                    }
                    const unsigned long long mk = ballot(keep);
                    if (keep) {
                        const int pos = NC + lanes_below(mk);
                        if (pos < MAXC) {
                            const double dist = sqrt((double)sq);
                            // plane (voronoi_poly2, LIB/Vcontacts.cpp:181-220), 'X' until the hull touches it
                            const float rj = RW + p.w;
                            const double pd = (dist * dist + (double)(rado * rado) - (double)(rj * rj)) / (2 * dist);
                            s.A0[pos] = (double)(p.x - ax) / dist;
                            s.A1[pos] = (double)(p.y - ay) / dist;
                            s.A2[pos] = (double)(p.z - az) / dist;
                            s.A3[pos] = -pd;
                            s.pflag[pos] = 'X';
                            s.u.c.cd[pos] = dist;
                            s.u.c.rb[pos] = p.w;
                            s.u.c.idx[pos] = j;
                            s.pmv[pos] = (signed char)mv;
                        }
                    }
                    NC += __popcll(mk);
                }
                // cull count of this pass: every candidate but the atom itself and
                // the done atoms it skips (all of them in the 27 cells)
                {
                    int sk = 0;
#pragma unroll
                    for (int q = 0; q < LSL; ++q) {
                        const int m = l + WAVE * q;
                        if (m < L && bit_of(sk0, sk1, m)) sk += slot_multiplicity(pg, s.sbox[s.rank[m]] - s.sbox[k]);
                    }
                    const int tot_k = readlane_i(k < WAVE ? totq[0] : totq[1], k & 63);
                    count_op(s.ops, OP_CULL, tot_k - wave_sum_i(sk));
                }
                NC = uni(NC);
                if (NC > MAXC) {
                    status = FA_CAPACITY;
                    break;
                }
                if (NC == 0) continue;  // no contact: SAS stays whole, atom not done
                count_op(s.ops, OP_PLANE, NC);
                if (l < 4) {
                    double bp[4];
                    bounding_plane(l, bp);
                    s.A0[NC + l] = bp[0];
                    s.A1[NC + l] = bp[1];
                    s.A2[NC + l] = bp[2];
                    s.A3[NC + l] = bp[3];
                    s.pflag[NC + l] = 'X';
                }
                wave_sync();
                // planes into registers: lane l holds contacts l and l + 64
                double pa0[PSL], pa1[PSL], pa2[PSL], pa3[PSL], pcd[PSL];
                float prb[PSL];
                int pix[PSL];
#pragma unroll
                for (int q = 0; q < PSL; ++q) {
                    const int c = l + WAVE * q;
                    pa0[q] = pa1[q] = pa2[q] = pa3[q] = pcd[q] = 0.0;
                    prb[q] = 0.0f;
                    pix[q] = -1;
                    if (c < NC) {
                        pa0[q] = s.A0[c];
                        pa1[q] = s.A1[c];
                        pa2[q] = s.A2[c];
                        pa3[q] = s.A3[c];
                        pcd[q] = s.u.c.cd[c];
                        prb[q] = s.u.c.rb[c];
                        pix[q] = s.u.c.idx[c];
                    }
                }
                static_assert(PSL == 2, "two plane slots per lane");
                // calc index of contact c (wave-uniform c < NC)
                auto atom_of = [&](int c) { return readlane_i(c < WAVE ? pix[0] : pix[1], c & 63); };
                clk.stop(PH_CULL);

                // -- first vertex (get_firstvert, LIB/Vcontacts.cpp:499-590)
                int pA = -1, pB = -1, pC = -1;
                {
                    const int s0 = s.seed[3 * a];
                    if (s0 != -1) {
                        const int s1 = s.seed[3 * a + 1], s2 = s.seed[3 * a + 2];
#pragma unroll
                        for (int q = 0; q < PSL; ++q) {
                            const int c = l + WAVE * q;
                            const bool v = c < NC;
                            const unsigned long long m0 = ballot(v && pix[q] == s0);
                            const unsigned long long m1 = ballot(v && pix[q] != s0 && pix[q] == s1);
                            const unsigned long long m2 = ballot(v && pix[q] != s0 && pix[q] != s1 && pix[q] == s2);
                            // the reference keeps the last match
                            if (m0) pA = WAVE * q + 63 - __clzll((long long)m0);
                            if (m1) pB = WAVE * q + 63 - __clzll((long long)m1);
                            if (m2) pC = WAVE * q + 63 - __clzll((long long)m2);
                        }
                    }
                    if (!(pA != -1 && pB != -1 && pC != -1)) {
                        // a partial seed match is discarded: all three are searched again
                        pA = pB = pC = -1;
                        // A: the plane closest to the atom (first of equals)
                        uint64_t kb = ~0ull;
#pragma unroll
                        for (int q = 0; q < PSL; ++q) {
                            const int c = l + WAVE * q;
                            const double d = fabs(pa3[q]);
                            if (c < NC && d < BIG) kb = umin64(kb, key_of(d));
                        }
                        const uint64_t kA = wave_min_u64(kb);
#pragma unroll
                        for (int q = 0; q < PSL; ++q) {
                            const int c = l + WAVE * q;
                            const unsigned long long m = ballot(c < NC && kA != ~0ull && key_of(fabs(pa3[q])) == kA);
                            if (pA < 0 && m) pA = WAVE * q + first_lane(m);
                        }
                        if (pA >= 0) {
                            double eA[4];
                            plane_of(s, pA, eA);
                            // B: the plane whose line with A passes closest to the atom;
                            // slot PSL holds the bounding planes (lanes 0..3)
                            double dq[PSL + 1], tx[PSL + 1], ty[PSL + 1], tz[PSL + 1];
                            uint64_t kbb = ~0ull;
#pragma unroll
                            for (int q = 0; q <= PSL; ++q) {
                                const int c = (q < PSL) ? l + WAVE * q : NC + l;
                                const bool ok = (q < PSL) ? (c < NC) : (l < 4);
                                dq[q] = BIG;
                                tx[q] = ty[q] = tz[q] = 0.0;
                                if (ok && c != pA) {
                                    double eC[4], vA[4], tp[3];
                                    if (q < PSL) {
                                        eC[0] = pa0[q];
                                        eC[1] = pa1[q];
                                        eC[2] = pa2[q];
                                        eC[3] = pa3[q];
                                    } else {
                                        bounding_plane(l, eC);
                                    }
                                    vA[0] = eA[1] * eC[2] - eA[2] * eC[1];
                                    vA[1] = eA[2] * eC[0] - eA[0] * eC[2];
                                    vA[2] = eA[0] * eC[1] - eA[1] * eC[0];
                                    vA[3] = 0.0;
                                    if (solve_3x3(eA, eC, vA, tp) != -1) {
                                        const double d = sqrt(tp[0] * tp[0] + tp[1] * tp[1] + tp[2] * tp[2]);
                                        if (d < BIG) {
                                            dq[q] = d;
                                            tx[q] = tp[0];
                                            ty[q] = tp[1];
                                            tz[q] = tp[2];
                                            kbb = umin64(kbb, key_of(d));
                                        }
                                    }
                                }
                            }
                            const uint64_t kB = wave_min_u64(kbb);
                            double ptx = 0.0, pty = 0.0, ptz = 0.0;
#pragma unroll
                            for (int q = 0; q <= PSL; ++q) {
                                const unsigned long long m = ballot(kB != ~0ull && dq[q] < BIG && key_of(dq[q]) == kB);
                                if (pB < 0 && m) {
                                    const int ln = first_lane(m);
                                    pB = (q < PSL) ? WAVE * q + ln : NC + ln;
                                    ptx = readlane_d(tx[q], ln);
                                    pty = readlane_d(ty[q], ln);
                                    ptz = readlane_d(tz[q], ln);
                                }
                            }
                            if (pB >= 0) {
                                double eB[4];
                                plane_of(s, pB, eB);
                                const double vA0 = eA[1] * eB[2] - eA[2] * eB[1];
                                const double vA1 = eA[2] * eB[0] - eA[0] * eB[2];
                                const double vA2 = eA[0] * eB[1] - eA[1] * eB[0];
                                // C: the plane cutting the line A x B closest to that point
                                double fq[PSL + 1];
                                uint64_t kcc = ~0ull;
#pragma unroll
                                for (int q = 0; q <= PSL; ++q) {
                                    const int c = (q < PSL) ? l + WAVE * q : NC + l;
                                    const bool ok = (q < PSL) ? (c < NC) : (l < 4);
                                    fq[q] = BIG;
                                    if (ok && c != pA && c != pB) {
                                        double eC[4];
                                        if (q < PSL) {
                                            eC[0] = pa0[q];
                                            eC[1] = pa1[q];
                                            eC[2] = pa2[q];
                                            eC[3] = pa3[q];
                                        } else {
                                            bounding_plane(l, eC);
                                        }
                                        const double vtdiv = eC[0] * vA0 + eC[1] * vA1 + eC[2] * vA2;
                                        if (vtdiv != 0.0) {
                                            const double vt = -(eC[0] * ptx + eC[1] * pty + eC[2] * ptz + eC[3]) / vtdiv;
                                            const double av = fabs(vt);
                                            if (av < BIG) {
                                                fq[q] = av;
                                                kcc = umin64(kcc, key_of(av));
                                            }
                                        }
                                    }
                                }
                                const uint64_t kC = wave_min_u64(kcc);
#pragma unroll
                                for (int q = 0; q <= PSL; ++q) {
                                    const unsigned long long m = ballot(kC != ~0ull && fq[q] < BIG && key_of(fq[q]) == kC);
                                    if (pC < 0 && m) pC = (q < PSL) ? WAVE * q + first_lane(m) : NC + first_lane(m);
                                }
                            }
                        }
                    }
                }
                if (pA < 0 || pB < 0 || pC < 0) {
                    status = FA_CAPACITY;
                    break;
                }
                clk.stop(PH_FIRST);

                // -- hull walk (LIB/Vcontacts.cpp:243-410)
                Edges E;
#pragma unroll
                for (int q = 0; q < ESL; ++q) {
                    E.pl[q] = 0xFFFFFFFFu;
                    E.vt[q] = 0xFFFFFFFFu;
                }
                int vn = 1, edgenum = 3;
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
                    edge_new(E, 0, pA, pB, pC, 0);
                    edge_new(E, 1, pB, pC, pA, 0);
                    edge_new(E, 2, pC, pA, pB, 0);
                }
                bool hull_fail = false, cap = false;
                const double rr = (double)(rado * rado);
                long long hull_tests = 0;
                int se = 0;
                while (true) {
                    se = first_open(E, se, edgenum);
                    if (se < 0) break;
                    const HullStep w = hull_finish(hull_search(s, E, se, NC, pa0, pa1, pa2, pa3), NC, rr);
                    hull_tests += w.tests;
                    if (w.C < 0 || vn + 1 >= MAXV) {
                        cap = true;
                        break;
                    }
                    put_vertex(s, vn, w.nx, w.ny, w.nz, w.A, w.B, w.C);
                    if (l == 0) s.pflag[w.C] = 'Y';
                    edge_close(E, se, w.C, vn);
                    // edges (A, C) then (B, C) among the later edges
                    for (int pass = 0; pass < 2; ++pass) {
                        const int P = pass ? w.B : w.A;
                        const int other = pass ? w.A : w.B;
                        const int f = find_edge(E, se, edgenum, P, w.C);
                        if (f >= 0) {
                            edge_close(E, f, other, vn);
                        } else {
                            edge_new(E, edgenum, P, w.C, other, vn);
                            ++edgenum;
                            if (pass == 1 && edgenum >= EDGE_LIMIT) hull_fail = true;
                        }
                    }
                    if (hull_fail) break;
                    ++vn;
                }
                count_op(s.ops, OP_HULL, hull_tests);
                if (cap) {
                    status = FA_CAPACITY;
                    break;
                }
                if (hull_fail) {
                    status = FA_HULL_FAIL;
                    break;
                }
                edgenum = uni(edgenum);
                vn = uni(vn);
                wave_sync();
                clk.stop(PH_HULL);

                // -- arc points (LIB/Vcontacts.cpp:419-457), in edge order
                {
                    int nsolved = 0;
#pragma unroll
                    for (int q = 0; q < ESL; ++q) {
                        if (q * WAVE >= edgenum) break;
                        const int e = q * WAVE + l;
                        double q0[3] = {0, 0, 0}, q1[3] = {0, 0, 0};
                        bool k0 = false, k1 = false;
                        const int e0 = (int)(E.pl[q] & NONE), e1 = (int)((E.pl[q] >> 8) & NONE);
                        if (e < edgenum) {
                            const int st = (int)(E.vt[q] & 0xFFFFu), en = (int)(E.vt[q] >> 16);
                            if (ro < vnorm(s, st) || ro < vnorm(s, en)) {
                                ++nsolved;
                                double a0[4], a1[4];
                                plane_of(s, e0, a0);
                                plane_of(s, e1, a1);
                                if (solve_2xS(a0, a1, rado, q0, q1) != -1) {
                                    const int ta = (int)((E.pl[q] >> 16) & NONE), tb = (int)(E.pl[q] >> 24);
                                    double pa[4], pb[4];
                                    plane_of(s, ta, pa);
                                    plane_of(s, tb, pb);
                                    double va = pa[0] * q0[0] + pa[1] * q0[1] + pa[2] * q0[2] + pa[3];
                                    double vb = pb[0] * q0[0] + pb[1] * q0[1] + pb[2] * q0[2] + pb[3];
                                    k0 = (va < 0.0 && vb < 0.0);
                                    va = pa[0] * q1[0] + pa[1] * q1[1] + pa[2] * q1[2] + pa[3];
                                    vb = pb[0] * q1[0] + pb[1] * q1[1] + pb[2] * q1[2] + pb[3];
                                    k1 = (va < 0.0 && vb < 0.0);
                                }
                            }
                        }
                        const int cnt = (int)k0 + (int)k1;
                        const int off = wave_excl_scan(cnt);
                        const int tot = wave_sum_i(cnt);
                        if (vn + tot > MAXV) {
                            cap = true;
                            break;
                        }
                        int pos = vn + off;
                        if (k0) {
                            s.vx[pos] = q0[0];
                            s.vy[pos] = q0[1];
                            s.vz[pos] = q0[2];
                            s.vpl[pos] = pack_planes(e0, e1, -1);
                            ++pos;
                        }
                        if (k1) {
                            s.vx[pos] = q1[0];
                            s.vy[pos] = q1[1];
                            s.vz[pos] = q1[2];
                            s.vpl[pos] = pack_planes(e0, e1, -1);
                        }
                        vn += tot;
                    }
                    count_op(s.ops, OP_ARC, wave_sum_i(nsolved));
                }
                if (cap) {
                    status = FA_CAPACITY;
                    break;
                }
                wave_sync();
                // -- keep the vertices inside the sphere (LIB/Vcontacts.cpp:460-467);
                // arc points lie on it (distance rado by definition)
                {
                    int vj = 0;
                    for (int base = 0; base < vn; base += WAVE) {
                        const int vi = base + l;
                        double x = 0, y = 0, z = 0;
                        unsigned pk = 0;
                        bool keep = false;
                        if (vi < vn) {
                            x = s.vx[vi];
                            y = s.vy[vi];
                            z = s.vz[vi];
                            pk = s.vpl[vi];
                            const double d = (plane_at(pk, 2) == -1) ? ro : sqrt(x * x + y * y + z * z);
                            keep = d <= ro;
                        }
                        const unsigned long long m = ballot(keep);
                        wave_sync();
                        if (keep) {
                            const int pos = vj + lanes_below(m);
                            s.vx[pos] = x;
                            s.vy[pos] = y;
                            s.vz[pos] = z;
                            s.vpl[pos] = pk;
                        }
                        vj += __popcll(m);
                        wave_sync();
                    }
                    vn = vj;
                }
                // -- engulfing planes (LIB/Vcontacts.cpp:470-478)
#pragma unroll
                for (int q = 0; q < PSL; ++q) {
                    const int c = l + WAVE * q;
                    if (c < NC && pa3[q] > 0.0) s.pflag[c] = 'E';
                }
                wave_sync();
                char pfl[PSL];
#pragma unroll
                for (int q = 0; q < PSL; ++q) {
                    const int c = l + WAVE * q;
                    pfl[q] = (c < NC) ? s.pflag[c] : 'X';
                }
                clk.stop(PH_ARC);
                const int NV = uni(vn);

                // -- seeds of scorable atoms not seeded yet (save_seeds, LIB/Vcontacts.cpp:595-629):
                // the first hull vertex (in order) with a plane of that atom
#pragma unroll
                for (int q = 0; q < PSL; ++q) {
                    const int c = l + WAVE * q;
                    const int cm = c < NC ? (int)s.pmv[c] : -1;
                    const bool cand = cm >= 0 && s.seed[3 * cm] == -1;
                    unsigned long long mc = ballot(cand);
                    while (mc) {
                        const int cc = WAVE * q + first_lane(mc);
                        mc &= mc - 1;
                        int hit_v = -1, hit_pos = 0;
                        for (int vb = 0; vb < NV && hit_v < 0; vb += WAVE) {
                            const int vi = vb + l;
                            int pos = -1;
                            if (vi < NV) {
                                const unsigned pk = s.vpl[vi];
                                if (plane_at(pk, 2) != -1) {
                                    if (plane_at(pk, 0) == cc)
                                        pos = 0;
                                    else if (plane_at(pk, 1) == cc)
                                        pos = 1;
                                    else if (plane_at(pk, 2) == cc)
                                        pos = 2;
                                }
                            }
                            const unsigned long long mv = ballot(pos >= 0);
                            if (mv) {
                                const int ln = first_lane(mv);
                                hit_v = vb + ln;
                                hit_pos = readlane_i(pos, ln);
                            }
                        }
                        if (hit_v >= 0) {
                            const unsigned pk = s.vpl[hit_v];
                            const int i0 = atom_of(plane_at(pk, 0));
                            const int i1 = atom_of(plane_at(pk, 1));
                            const int i2 = atom_of(plane_at(pk, 2));
                            const int m = s.pmv[cc];
                            if (l == 0) {
                                s.seed[3 * m] = acalc;
                                s.seed[3 * m + 1] = hit_pos == 0 ? i1 : i0;
                                s.seed[3 * m + 2] = hit_pos == 2 ? i1 : i2;
                            }
                        }
                    }
                }
                wave_sync();
                clk.stop(PH_SEED);

                // -- face point lists (order_faces, LIB/Vcontacts.cpp:699-831)
                {
                    unsigned vpk[VSL];
#pragma unroll
                    for (int k2 = 0; k2 < VSL; ++k2) {
                        const int vi = l + WAVE * k2;
                        vpk[k2] = (vi < NV) ? s.vpl[vi] : 0xFFFFFFFFu;
                    }
                    for (int c = l; c < NC; c += WAVE) s.fcnt[c] = 0;
                    int run = 0;
#pragma unroll
                    for (int q = 0; q < PSL; ++q) {
                        const int c = l + WAVE * q;
                        unsigned long long mX = ballot(c < NC && pfl[q] != 'X');
                        while (mX) {
                            const int X = WAVE * q + first_lane(mX);
                            mX &= mX - 1;
                            int cnt = 0;
#pragma unroll
                            for (int k2 = 0; k2 < VSL; ++k2) {
                                if (k2 * WAVE >= NV) break;
                                const int vi = l + WAVE * k2;
                                const bool on = vi < NV && has_plane(vpk[k2], X);
                                const unsigned long long m = ballot(on);
                                if (on) {
                                    const int pos = run + cnt + lanes_below(m);
                                    if (pos < MAXF) s.fpt[pos] = (short)vi;
                                }
                                cnt += __popcll(m);
                            }
                            if (l == 0) {
                                s.foff[X] = (short)run;
                                s.fcnt[X] = (unsigned char)cnt;
                            }
                            run += cnt + (cnt > 2 ? 1 : 0);
                        }
                    }
                    if (run > MAXF) {
                        status = FA_CAPACITY;
                        break;
                    }
                    wave_sync();
                    // opposite point of the first arc point, appended in plane order
                    int nv_extra = NV;
#pragma unroll
                    for (int q = 0; q < PSL; ++q) {
                        const int X = l + WAVE * q;
                        bool need = false;
                        double t[3] = {0, 0, 0};
                        if (X < NC && pfl[q] != 'X' && s.fcnt[X] > 2) {
                            const int o = s.foff[X], sc = s.fcnt[X];
                            const double cx = -pa0[q] * pa3[q], cy = -pa1[q] * pa3[q], cz = -pa2[q] * pa3[q];
                            for (int w = 0; w < sc; ++w) {
                                const int vi = s.fpt[o + w];
                                if (plane_at(s.vpl[vi], 2) == -1) {
                                    t[0] = 2.0 * cx - s.vx[vi];
                                    t[1] = 2.0 * cy - s.vy[vi];
                                    t[2] = 2.0 * cz - s.vz[vi];
                                    need = true;
                                    break;
                                }
                            }
                        }
                        const bool add = points_inside(need, t, q, NC, pa0, pa1, pa2, pa3, pfl);
                        const unsigned long long m = ballot(add);
                        if (nv_extra + __popcll(m) > MAXV) {
                            cap = true;
                            break;
                        }
                        if (add) {
                            const int vi = nv_extra + lanes_below(m);
                            s.vx[vi] = t[0];
                            s.vy[vi] = t[1];
                            s.vz[vi] = t[2];
                            s.vpl[vi] = pack_planes(X, -1, -1);
                            s.fpt[s.foff[X] + s.fcnt[X]] = (short)vi;
                            s.fcnt[X] = (unsigned char)(s.fcnt[X] + 1);
                        }
                        nv_extra += __popcll(m);
                    }
                    if (cap) {
                        status = FA_CAPACITY;
                        break;
                    }
                    wave_sync();
                    // points of each face ordered by cosine around it
                    bool noline = false;
                    int ncos = 0;
#pragma unroll
                    for (int q = 0; q < PSL; ++q) {
                        const int X = l + WAVE * q;
                        if (X >= NC || pfl[q] == 'X') continue;
                        const int sc = s.fcnt[X];
                        if (sc <= 3) continue;
                        short* pt = &s.fpt[s.foff[X]];
                        double* cs = &s.u.fcos[s.foff[X]];
                        const unsigned pk0 = s.vpl[pt[0]];
                        const int Y = (plane_at(pk0, 0) == X) ? plane_at(pk0, 1) : plane_at(pk0, 0);
                        int vi;
                        for (vi = 1; vi < sc; ++vi) {
                            const unsigned pk = s.vpl[pt[vi]];
                            if (plane_at(pk, 0) == Y || plane_at(pk, 1) == Y || plane_at(pk, 2) == Y) break;
                        }
                        if (vi == sc) {  // no second point on the common line (the reference reads past its list)
                            noline = true;
                            continue;
                        }
                        ncos += sc - 2;
                        const short tsw = pt[vi];
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
                                    const short ts = pt[vi];
                                    pt[vi] = pt[v2];
                                    pt[v2] = ts;
                                    const double tc = cs[vi];
                                    cs[vi] = cs[v2];
                                    cs[v2] = tc;
                                }
                    }
                    wave_sync();
                    count_op(s.ops, OP_ORDER, wave_sum_i(ncos));
                    if (ballot(noline) != 0ull) {
                        status = FA_UNDEFINED_FACE;
                        break;
                    }
                }
                clk.stop(PH_FACE);

                // -- areas (calc_areas + project_points, LIB/Vcontacts.cpp:838-1044, 1278-1344)
                double par[PSL] = {0.0, 0.0};
                {
                    int epi = 0, eng[3] = {-1, -1, -1};
#pragma unroll
                    for (int q = 0; q < PSL; ++q) {
                        const int c = l + WAVE * q;
                        unsigned long long m = ballot(c < NC && pfl[q] == 'E');
                        while (m && epi < 3) {
                            eng[epi++] = WAVE * q + first_lane(m);
                            m &= m - 1;
                        }
                    }
                    const bool projected = epi > 0;
                    if (projected) {
                        double proj[3];
                        if (epi == 1) {
                            double a0[4];
                            plane_of(s, eng[0], a0);
                            proj[0] = -a0[0] * a0[3];
                            proj[1] = -a0[1] * a0[3];
                            proj[2] = -a0[2] * a0[3];
                        } else if (epi == 2) {
                            double a0[4], a1[4], p0[3], p1[3];
                            plane_of(s, eng[0], a0);
                            plane_of(s, eng[1], a1);
                            if (solve_2xS(a0, a1, rado, p0, p1) == -1) {
                                // the reference uses the atom's absolute coordinates here
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
                            if (plane_at(s.vpl[vi], 2) == -1) continue;
                            const double P0 = s.vx[vi], P1 = s.vy[vi], P2 = s.vz[vi];
                            const double V0 = P0 - proj[0], V1 = P1 - proj[1], V2 = P2 - proj[2];
                            const double qa = V0 * V0 + V1 * V1 + V2 * V2;
                            const double qb = 2 * (P0 * V0 + P1 * V1 + P2 * V2);
                            const double qc = P0 * P0 + P1 * P1 + P2 * P2 - (double)(rado * rado);
                            const double kk = (sqrt(qb * qb - 4.0 * qa * qc) - qb) / (2 * qa);
                            s.vx[vi] = P0 + kk * V0;
                            s.vy[vi] = P1 + kk * V1;
                            s.vz[vi] = P2 + kk * V2;
                        }
                        wave_sync();
                    }
                    // distance of a face point: hull vertices keep |v| until projected
                    // onto the sphere; arc and opposite points lie on it
                    auto vdist = [&](int w) {
                        return (projected || plane_at(s.vpl[w], 2) == -1) ? ro : vnorm(s, w);
                    };
                    // Jobs: the spherical triangles of each face's fan (NP - 2 of them)
                    // and its boundary arcs are evaluated one per lane into u.fcos;
                    // each face's lane then sums its own terms in the reference's
                    // order (triangles, then arcs), so the sums are bit-identical.
                    int ntq[PSL], naq[PSL], toff[PSL], aoff[PSL];
                    unsigned amk[PSL];  // boundary arcs of a face: bit qq for the pair (qq, qq + 1), faces of <= 32 points
                    int NT = 0, NA = 0;
#pragma unroll
                    for (int q = 0; q < PSL; ++q) {
                        const int X = l + WAVE * q;
                        ntq[q] = naq[q] = 0;
                        amk[q] = 0u;
                        if (X < NC && pfl[q] != 'X') {
                            const int NP = s.fcnt[X];
                            if (NP == 2) {
                                naq[q] = 1;
                            } else if (NP > 2) {
                                ntq[q] = NP - 2;
                                const short* pt = &s.fpt[s.foff[X]];
                                // consecutive point pairs, the next point's index read
                                // one pair ahead so each pair waits for one LDS read
                                const unsigned kfirst = s.vpl[pt[0]];
                                unsigned kcur = kfirst;
                                int wn = pt[1];
                                for (int qq = 0; qq < NP; ++qq) {
                                    const int wnn = (qq + 2 < NP) ? pt[qq + 2] : 0;
                                    const unsigned knext = (qq + 1 < NP) ? s.vpl[wn] : kfirst;
                                    const int ap = arc_pair(kcur, knext);
                                    naq[q] += ap;
                                    if (qq < 32) amk[q] |= (unsigned)ap << qq;
                                    kcur = knext;
                                    wn = wnn;
                                }
                            }
                        }
                        toff[q] = NT + wave_excl_scan(ntq[q]);
                        aoff[q] = NA + wave_excl_scan(naq[q]);
                        NT += wave_sum_i(ntq[q]);
                        NA += wave_sum_i(naq[q]);
                    }
                    count_op(s.ops, OP_TRI, NT);
                    count_op(s.ops, OP_SARC, NA);
                    wave_sync();  // u.fcos is free again
                    // face and index of job t in a list laid out by (toff, ntq)
                    auto job_face = [&](int base, int t, const int* off, const int* cnt, int& X, int& u) {
                        X = -1;
                        u = 0;
#pragma unroll
                        for (int q = 0; q < PSL; ++q) {
                            unsigned long long ov = ballot(cnt[q] > 0 && off[q] < base + WAVE && off[q] + cnt[q] > base);
                            while (ov) {
                                const int fl = first_lane(ov);
                                ov &= ov - 1;
                                const int o_ = readlane_i(off[q], fl), n_ = readlane_i(cnt[q], fl);
                                if (t >= o_ && t < o_ + n_) {
                                    X = WAVE * q + fl;
                                    u = t - o_;
                                }
                            }
                        }
                    };
                    for (int base = 0; base < NT; base += WAVE) {
                        const int t = base + l;
                        int X, u;
                        job_face(base, t, toff, ntq, X, u);
                        if (t < NT) {
                            const short* pt = &s.fpt[s.foff[X]];
                            const int w0 = pt[0], wq = pt[u + 1], wn = pt[u + 2];
                            const double x0 = s.vx[w0], y0 = s.vy[w0], z0 = s.vz[w0], d0 = vdist(w0);
                            const double dq = vdist(wq), dn = vdist(wn);
                            const double c0 = (x0 * s.vx[wq] + y0 * s.vy[wq] + z0 * s.vz[wq]) / (d0 * dq);
                            const double cq = (s.vx[wq] * s.vx[wn] + s.vy[wq] * s.vy[wn] + s.vz[wq] * s.vz[wn]) / (dq * dn);
                            const double c1 = (x0 * s.vx[wn] + y0 * s.vy[wn] + z0 * s.vz[wn]) / (d0 * dn);
                            s.u.fcos[t] = sph_triangle(c0, cq, c1, ro);
                        }
                    }
                    wave_sync();
#pragma unroll
                    for (int q = 0; q < PSL; ++q)
                        for (int u = 0; u < ntq[q]; ++u) par[q] += s.u.fcos[toff[q] + u];
                    wave_sync();
                    for (int base = 0; base < NA; base += WAVE) {
                        const int t = base + l;
                        int X, u;
                        job_face(base, t, aoff, naq, X, u);
                        // the arc mask of the job's face, from the face's lane
                        const unsigned am0 = (unsigned)__shfl((int)amk[0], X & 63);
                        const unsigned am1 = (unsigned)__shfl((int)amk[1], X & 63);
                        static_assert(PSL == 2, "two face slots per lane");
                        if (t < NA) {
                            const short* pt = &s.fpt[s.foff[X]];
                            const int NP = s.fcnt[X];
                            int wa = pt[0], wb = pt[1];
                            if (NP > 32) {  // the u-th qualifying boundary arc, by scan
                                int k2 = -1;
                                for (int qq = 0; qq < NP && k2 < u; ++qq) {
                                    wa = pt[qq];
                                    wb = pt[(qq + 1) % NP];
                                    k2 += arc_pair(s.vpl[wa], s.vpl[wb]);
                                }
                            } else if (NP > 2) {  // the u-th set bit of the face's arc mask
                                unsigned m = (X < WAVE) ? am0 : am1;
                                for (int i = 0; i < u; ++i) m &= m - 1u;
                                const int qq = __ffs((int)m) - 1;
                                wa = pt[qq];
                                wb = pt[(qq + 1) % NP];
                            }
                            double pX[4];
                            plane_of(s, X, pX);
                            const double ctr[3] = {-pX[0] * pX[3], -pX[1] * pX[3], -pX[2] * pX[3]};
                            const double Pa[3] = {s.vx[wa], s.vy[wa], s.vz[wa]};
                            const double Pb[3] = {s.vx[wb], s.vy[wb], s.vz[wb]};
                            s.u.fcos[t] = spherical_arc(ctr, fabs(pX[3]), Pa, vdist(wa), Pb, vdist(wb), rado);
                        }
                    }
                    wave_sync();
#pragma unroll
                    for (int q = 0; q < PSL; ++q) {
                        const int X = l + WAVE * q;
                        const bool face = X < NC && pfl[q] != 'X';
                        const int NP = face ? s.fcnt[X] : -1;
                        const double ctr[3] = {-pa0[q] * pa3[q], -pa1[q] * pa3[q], -pa2[q] * pa3[q]};
                        const double cd = fabs(pa3[q]);
                        // faces with no point or one arc keep their cap when the
                        // plane's centre is inside the polyhedron
                        const bool kept = points_inside(NP == 0 || NP == 2, ctr, q, NC, pa0, pa1, pa2, pa3, pfl);
                        if (!face) continue;
                        if (NP == 0) {
                            if (kept) par[q] = (2.0f * PI_REF) * ro * (ro - cd);
                        } else if (NP == 2) {
                            const double arc = s.u.fcos[aoff[q]];
                            if (kept)
                                par[q] = (2.0f * PI_REF) * ro * (ro - cd) - arc;
                            else
                                par[q] = arc;
                        } else {
                            for (int u = 0; u < naq[q]; ++u) par[q] += s.u.fcos[aoff[q] + u];
                        }
                    }
                    if (projected) {
                        // corrections onto the planes next to engulfing faces, sequential
                        wave_sync();
#pragma unroll
                        for (int q = 0; q < PSL; ++q) {
                            const int c = l + WAVE * q;
                            if (c < NC) s.u.area[c] = par[q];
                        }
                        wave_sync();
                        if (l == 0) {
                            for (int X = 0; X < NC; ++X) {
                                if (s.pflag[X] != 'E') continue;
                                const int NP = s.fcnt[X];
                                const short* pt = &s.fpt[s.foff[X]];
                                double pX[4];
                                plane_of(s, X, pX);
                                const double ctr[3] = {-pX[0] * pX[3], -pX[1] * pX[3], -pX[2] * pX[3]};
                                const double cd = fabs(pX[3]);
                                for (int qq = 0; qq < NP; ++qq) {
                                    const int cu = pt[qq], nx = pt[(qq + 1) % NP];
                                    const unsigned kc = s.vpl[cu], kn = s.vpl[nx];
                                    int pa1, pa2;
                                    if (plane_at(kc, 0) == X) {
                                        pa1 = plane_at(kc, 1);
                                        pa2 = plane_at(kc, 2);
                                    } else {
                                        pa1 = plane_at(kc, 0);
                                        pa2 = (plane_at(kc, 1) == X) ? plane_at(kc, 2) : plane_at(kc, 1);
                                    }
                                    const int n0 = plane_at(kn, 0), n1 = plane_at(kn, 1), n2 = plane_at(kn, 2);
                                    int comm;
                                    if (pa1 == n0 || pa1 == n1 || pa1 == n2)
                                        comm = pa1;
                                    else if (pa2 == n0 || pa2 == n1 || pa2 == n2)
                                        comm = pa2;
                                    else
                                        continue;
                                    if (comm != -1 && s.pflag[comm] != 'E') {
                                        s.u.area[comm] += spherical_arc_call(ctr[0], ctr[1], ctr[2], cd, s.vx[cu], s.vy[cu],
                                                                             s.vz[cu], ro, s.vx[nx], s.vy[nx], s.vz[nx],
                                                                             ro, rado);
                                        s.ops[OP_SARC] += 1;
                                        if (NP == 2) break;
                                    }
                                }
                            }
                            if (epi == 1) {
                                s.u.area[eng[0]] = (2.0f * PI_REF) * ro * (ro + fabs(s.A3[eng[0]]));
                            } else if (epi == 2) {
                                double a0[4], a1[4], b[3], c[3];
                                plane_of(s, eng[0], a0);
                                plane_of(s, eng[1], a1);
                                if (solve_2xS(a0, a1, rado, b, c) == -1) {
                                    s.u.area[eng[0]] = (2.0f * PI_REF) * ro * ro;
                                    s.u.area[eng[1]] = (2.0f * PI_REF) * ro * ro;
                                } else {
                                    const double c0[3] = {-a0[0] * a0[3], -a0[1] * a0[3], -a0[2] * a0[3]};
                                    const double c1[3] = {-a1[0] * a1[3], -a1[1] * a1[3], -a1[2] * a1[3]};
                                    double mx = spherical_arc_call(c0[0], c0[1], c0[2], fabs(a0[3]), b[0], b[1], b[2], ro,
                                                                   c[0], c[1], c[2], ro, rado);
                                    mx += spherical_arc_call(c1[0], c1[1], c1[2], fabs(a1[3]), b[0], b[1], b[2], ro, c[0],
                                                             c[1], c[2], ro, rado);
                                    s.ops[OP_SARC] += 2;
                                    s.u.area[eng[0]] = (2.0f * PI_REF) * ro * ro - 0.5 * mx;
                                    s.u.area[eng[1]] = (2.0f * PI_REF) * ro * ro - 0.5 * mx;
                                }
                            }
                        }
                        wave_sync();
#pragma unroll
                        for (int q = 0; q < PSL; ++q) {
                            const int c = l + WAVE * q;
                            if (c < NC) par[q] = s.u.area[c];
                            if (epi >= 3 && c < NC) par[q] = (pfl[q] == 'E') ? (4.0f * PI_REF) * ro * ro / epi : 0.0;
                        }
                    }
                }
                clk.stop(PH_AREA);

                // -- records (save_areas, LIB/Vcontacts.cpp:1149-1220) and the energy
                // terms of atom a (vcfunction, LIB/vcfunction.cpp:77-527)
                {
                    const double radA = (double)ra;
                    const double radoA = radA + RW;
                    const double surfA = 4.0 * PI_REF * radoA * radoA;
                    const double dee = (double)prm.dee_clash;
                    const int nt = prm.ntypes;
                    const int ta = acalc < Rn ? R.type[acalc] : lig.type[acalc - Rn];
                    int rlo, rhi;  // a's residue (the ligand or its side chain) and bonded row
                    unsigned long long rb0, rb1;
                    residue_of(a, rlo, rhi, rb0, rb1);
                    double sas_sub = 0.0, com = 0.0, wal = 0.0, con = 0.0;
                    int rcl = 0;
                    bool valid[PSL];
                    int nvalid_total = 0;
#pragma unroll
                    for (int q = 0; q < PSL; ++q) {
                        const int c = l + WAVE * q;
                        valid[q] = c < NC && par[q] > 0.0 && pfl[q] != 'X';
                        nvalid_total += __popcll(ballot(valid[q]));
                    }
                    int seen = 0;
#pragma unroll
                    for (int q = 0; q < PSL; ++q) {
                        const unsigned long long m = ballot(valid[q]);
                        if (valid[q]) {
                            const int j = pix[q];
                            const double ar = par[q];
                            const double dist = pcd[q];
                            sas_sub += ar;
                            const int mj = s.pmv[l + WAVE * q];  // scorable index, -1 for a rigid atom
                            const bool same = j >= rlo && j < rhi;
                            const int tb = j < Rn ? R.type[j] : lig.type[j - Rn];
                            if (B.recs) {  // own list in reverse plane order
                                const int r = nrec_total + (nvalid_total - 1 - (seen + lanes_below(m)));
                                if (r < B.rec_cap) {
                                    fa_contact_rec& o = B.recs[(size_t)pose * B.rec_cap + r];
                                    o.from = acalc;
                                    o.atom = j;
                                    o.area = ar;
                                    o.dist = dist;
                                }
                            }
                            // bonded atoms of the residue never score; a pair of scorable
                            // atoms scores once, at the lower calc index (FA->contacts,
                            // LIB/vcfunction.cpp:230-250); scorable indices follow calc order
                            const bool bonded = same && bit_of(rb0, rb1, j - rlo);
                            const bool skip = bonded || (mj >= 0 && mj < a);
                            if (mj >= 0 && !bit_of(dm0, dm1, mj)) s.stub[2 * mj + (a >> 6)] |= (1ull << (a & 63));
                            if (!skip) {
                                const double rAB = radA + (double)prb[q];
                                // covalent constraint of the pair (LIB/vcfunction.cpp:270-310): the
                                // first constraint of `a` naming this partner decides
                                bool covalent = false;
                                double dist_opt = 0.0;
                                if (lig.cons_off && a >= NSC) {  // constraints are the ligand's
                                    const int c1 = lig.cons_off[a - NSC + 1];
                                    for (int k = lig.cons_off[a - NSC]; k < c1; ++k) {
                                        if (lig.cons_part[k] != j) continue;
                                        if (lig.cons_type[k] == 1) {
                                            covalent = true;
                                            dist_opt = (double)lig.cons_bond[k];
                                            con -= KDIST * gaussian_bell(dist, dist_opt, (double)lig.cons_max[k]);
                                        }
                                        break;
                                    }
                                }
                                float clash_distance = (float)(permea * rAB);
                                if (covalent && permea * dist_opt < permea * rAB) clash_distance = (float)(permea * dist_opt);
                                if (dist < (double)clash_distance) {
                                    // the wall keeps permeability * rAB as its reference distance
                                    const double d2 = dist * dist, d4 = d2 * d2, d6 = d4 * d2;
                                    const double cr = permea * rAB, cr2 = cr * cr, cr4 = cr2 * cr2, cr6 = cr4 * cr2;
                                    wal += (double)KWALL_F * (1.0 / (d6 * d6) - 1.0 / (cr6 * cr6));
                                    if (dist <= dee * rAB) rcl = 1;
                                }
                                // intramolecular: one residue, or both protein residues
                                // (a side-chain atom and any receptor atom), LIB/vcfunction.cpp:219-229
                                const bool intramol = same || (acalc < Rn && j < Rn);
                                if (!covalent && (prm.intramolecular || !intramol))
                                    com += pair_term(prm, (ta - 1) * nt + (tb - 1), ar, surfA);
                            }
                        }
                        seen += __popcll(m);
                    }
                    nrec_total += nvalid_total;
                    count_op(s.ops, OP_VC, nvalid_total);
                    sas_sub = wave_sum_d(sas_sub);
                    com = wave_sum_d(com);
                    wal = wave_sum_d(wal);
                    if (lig.cons_off) r_con += wave_sum_d(con);
                    rcl = wave_sum_i(rcl);
                    double SAS = surfA - sas_sub;
                    if (SAS < 0.0) SAS = 0.0;
                    r_com += com;
                    r_wal += wal;
                    r_tot += SAS;
                    r_sas += (prm.solventterm != 0.0f) ? (double)prm.solventterm * SAS
                                                        : pair_term(prm, (ta - 1) * nt + (nt - 1), SAS, surfA);
                    r_rclash |= rcl ? 1 : 0;
                    if (a < WAVE)
                        dm0 |= 1ull << a;
                    else
                        dm1 |= 1ull << (a - WAVE);
                    wave_sync();
                }
                clk.stop(PH_REC);
            }
            if (status == FA_OK) {
                // atoms with no contact at all keep their whole sphere as SAS
                double add_tot = 0.0, add_sas = 0.0;
                for (int i = l; i < L; i += WAVE) {
                    if (!bit_of(dm0, dm1, i)) {
                        const double radoA = (double)s.lr[i] + RW;
                        const double SAS = 4.0 * PI_REF * radoA * radoA;
                        const int ci = calc_of(R, i);
                        const int ti = ci < Rn ? R.type[ci] : lig.type[ci - Rn];
                        add_tot += SAS;
                        add_sas += (prm.solventterm != 0.0f)
                                       ? (double)prm.solventterm * SAS
                                       : pair_term(prm, (ti - 1) * prm.ntypes + (prm.ntypes - 1), SAS, SAS);
                    }
                }
                r_tot += wave_sum_d(add_tot);
                r_sas += wave_sum_d(add_sas);
            }
        }
        fa_cf_result res;
        res.com = res.wal = res.sas = res.con = res.totsas = res.cf = res.app = 0.0;
        res.rclash = 0;
        res.status = status;
        if (status == FA_OK) {
            res.com = r_com;
            res.wal = r_wal;
            res.sas = r_sas;
            res.totsas = r_tot;
            res.rclash = r_rclash;
            res.con = lig.con0 + r_con;  // KDIST per covalent constraint, less the satisfied ones
            res.cf = res.com + res.wal + res.sas + res.con;
            res.app = res.com + res.wal + res.sas;
        } else if (status == FA_CLASH) {
            res.wal = clash;
            res.cf = clash + 1.0;  // metal_coord = 1 in the clash cfstr (LIB/ic2cf.cpp:129)
            res.app = clash + 1.0;
        } else if (status == FA_HULL_FAIL) {
            res.wal = (double)POLY_PENALTY;
            res.cf = res.wal + 1.0;
            res.app = res.wal + 1.0;
        } else {
            // not scored (a device capacity or an atom outside the cell grid):
            // +inf ranks after every scored pose and weighs 0 in ensembles
            res.cf = res.app = __longlong_as_double(0x7ff0000000000000ll);
        }
        if (status == FA_CAPACITY && l == 0) {
            atomicAdd(&B.counters[1], 1u);
            atomicAdd(&B.totals[1], 1ull);
        }
        if (status == FA_GRID_RANGE && l == 0) atomicAdd(&B.totals[2], 1ull);
        if (status == FA_UNDEFINED_FACE && l == 0) atomicAdd(&B.totals[4], 1ull);
        wave_sync();
        if (B.ops && l < FA_OPCOUNT_N && s.ops[l]) atomicAdd(&B.ops[l], s.ops[l]);
        if (l == 0) {
            B.out[pose] = res;
            if (B.rec_count) B.rec_count[pose] = (status == FA_OK) ? nrec_total : 0;
            // diagnostic launches: the pose's cycles, top bit set on a widened grid
            if (B.prof)
                B.prof[PROF_POSE0 + pose] = (__builtin_readcyclecounter() - pose_t0) | (pg.slow ? (1ull << 63) : 0ull);
        }
        wave_sync();
    }
    if (B.prof && l < N_PHASES) atomicAdd(&B.prof[l], s.phase[l]);
}

// host-side launcher (same translation unit as the kernel)
hipError_t launch_cf(const DevReceptor& R, const DevGrid* grids, const DevLigand* ligs, const DevParams& prm,
                     const DevBatch& B, int n_blocks, hipStream_t stream)
{
    if (B.n_poses <= 0) return hipSuccess;
    if (n_blocks <= 0) n_blocks = 1;
    hipLaunchKernelGGL(cf_pose_kernel, dim3(n_blocks), dim3(WAVE), 0, stream, R, grids, ligs, prm, B);
    return hipGetLastError();
}

// Ligand coordinates of poses without scoring them, for the pose
// clustering (the coordinates calc_rmsd_chrom rebuilds,
// LIB/calc_rmsd_chrom.cpp): build_pose's decode and rebuild (LIB/ic2cf.cpp:
// 53-105, LIB/buildcc.cpp) with one pose per thread instead of one per wave,
// the same operations in the same order (place_atom), so the coordinates
// are the pose kernel's bit for bit.  Internal coordinates and coordinates
// live in global memory in tiles of COORD_TILE poses (coord_at), so the
// threads of a block touch consecutive words of one region (keeping a
// thread's coordinates in LDS instead measured slower: the LDS it needs
// leaves too few waves to hide the trigonometry).
__global__ __launch_bounds__(COORD_TILE) void pose_coords_kernel(const DevLigand* __restrict__ ligs, int ligand,
                                                          const double* __restrict__ genes, int stride, int n,
                                                          float* __restrict__ xyzT, float* __restrict__ icT)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const DevLigand lig = ligs[ligand];
    const int L = lig.n_atoms, R3 = 3 * L;
    // rows of icT: dis of atom a at a, ang at L + a, dih at 2L + a
    auto D = [&](int a) -> float& { return icT[coord_at(R3, a, t)]; };
    auto A = [&](int a) -> float& { return icT[coord_at(R3, L + a, t)]; };
    auto H = [&](int a) -> float& { return icT[coord_at(R3, 2 * L + a, t)]; };
    auto X = [&](int a, int k) -> float& { return xyzT[coord_at(R3, 3 * a + k, t)]; };
    for (int a = 0; a < L; ++a) {
        D(a) = lig.dis[a];
        A(a) = lig.ang[a];
        H(a) = lig.dih[a];
        // atoms not yet built read as 0, as build_pose's registers start
        X(a, 0) = 0.0f;
        X(a, 1) = 0.0f;
        X(a, 2) = 0.0f;
    }
    const double* icv = genes + (size_t)t * stride;
    for (int g = 0; g < lig.n_genes; ++g) {
        double v = icv[g];
        const double gmax = lig.gene_max[g], gmin = lig.gene_min[g];
        if (v > gmax) v = gmax;
        if (v < gmin) v = gmin;
        const int atm = lig.gene_atm[g], typ = lig.gene_typ[g];
        if (typ == FA_GENE_GRID) {
            unsigned gi = (unsigned)v;
            if ((int)gi >= lig.n_grid) gi = lig.n_grid - 1;
            D(atm) = lig.grid_ic[3 * gi];
            A(atm) = lig.grid_ic[3 * gi + 1];
            H(atm) = lig.grid_ic[3 * gi + 2];
        } else if (typ == FA_GENE_DIS) {
            D(atm) = (float)v;
        } else if (typ == FA_GENE_ANG) {
            A(atm) = (float)v;
        } else if (typ == FA_GENE_DIH) {
            H(atm) = (float)v;
            int j = atm, cat = lig.rec[4 * j + 3];
            if (cat >= 0) {
                for (int guard = 0; cat != atm && guard < L; ++guard) {
                    H(cat) = H(j) + lig.shift[cat];
                    j = cat;
                    cat = lig.rec[4 * j + 3];
                }
            }
        }  // rotamer genes move side chains only
    }
    const float ox0 = lig.ori[0], oy0 = lig.ori[1], oz0 = lig.ori[2];
    for (int an = 0; an < L; ++an) {
        const int a = lig.mov[an];
        const float angPI = (float)(A(a) * PI_REF / 180.0f);
        const float dihPI = (float)(H(a) * PI_REF / 180.0f);
        const float cta = (float)cos((double)angPI);
        const float sta = (float)sin((double)angPI);
        const float ctd = (float)cos((double)dihPI);
        const float std_ = (float)sin((double)dihPI);
        const float dis = D(a);
        float x[4], y[4], z[4];
        x[0] = y[0] = z[0] = 0.0f;
        for (int i = 1; i <= 3; ++i) {
            const int j = lig.rec[4 * a + i - 1];
            if (j >= 0) {
                x[i] = X(j, 0);
                y[i] = X(j, 1);
                z[i] = X(j, 2);
            } else if (i == 1) {
                x[i] = 1.0f + ox0;
                y[i] = 0.0f + oy0;
                z[i] = 0.0f + oz0;
            } else if (i == 3) {
                x[i] = 0.0f + ox0;
                y[i] = 1.0f + oy0;
                z[i] = 0.0f + oz0;
            } else {
                x[i] = 0.0f + ox0;
                y[i] = 0.0f + oy0;
                z[i] = 0.0f + oz0;
            }
            x[i] += 1e-10f;
            y[i] += 1e-10f;
            z[i] += 1e-10f;
        }
        float px, py, pz;
        place_atom(x, y, z, dis, cta, sta, ctd, std_, px, py, pz);
        X(a, 0) = px;
        X(a, 1) = py;
        X(a, 2) = pz;
    }
}

hipError_t launch_coords(const DevLigand* ligs, int ligand, int n_atoms, const double* genes, int stride, int n,
                         float* xyzT, float* icT, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    if (n_atoms < 1 || n_atoms > FA_MAX_LIG_ATOMS) return hipErrorInvalidValue;
    hipLaunchKernelGGL(pose_coords_kernel, dim3((n + COORD_TILE - 1) / COORD_TILE), dim3(COORD_TILE), 0, stream, ligs,
                       ligand, genes, stride, n, xyzT, icT);
    return hipGetLastError();
}

int smem_bytes() { return (int)sizeof(Smem); }

// resident waves per CU for the pose kernel (LDS and registers), 0 on error
int cf_waves_per_cu()
{
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, cf_pose_kernel, WAVE, 0) != hipSuccess) return 0;
    return n;
}

}  // namespace fa
