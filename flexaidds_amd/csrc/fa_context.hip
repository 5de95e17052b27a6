// fa_context.hip — device context, system upload and the C ABI.
//
// A context owns one GPU's resident state: the receptor (SoA, calc order),
// one receptor cell list per distinct grid, the ligand topologies, the
// scoring parameters, and growable pose buffers.  Uploads happen once; the
// evaluation entry points only move genes in and results out.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "fa_ctx.h"

namespace {
thread_local std::string g_err;
constexpr float CELLSIZE = 6.5f;  // LIB/Vcontacts.h:56
constexpr int NEAR_PER_ATOM = 160;
constexpr int NEAR_TAIL = 192;  // the pose kernel parks slow-path candidates past the lists (SLOW_TMP)
constexpr long long NB_MAX_CELLS = 1LL << 18;  // grids with more cells scan the cell lists instead
}  // namespace

int fa_fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

int fa_ensure(DevBuf& b, size_t bytes)
{
    if (b.n >= bytes) return FA_SUCCESS;
    if (b.p) FA_HIP(hipFree(b.p));
    b.p = nullptr;
    b.n = 0;
    FA_HIP(hipMalloc(&b.p, bytes));
    b.n = bytes;
    return FA_SUCCESS;
}

namespace {

template <class T> int upload(fa_ctx* c, const T* h, size_t n, const T** d)
{
    void* p = nullptr;
    const size_t bytes = sizeof(T) * (n ? n : 1);
    FA_HIP(hipMalloc(&p, bytes));
    c->allocs.push_back(p);
    if (n) FA_HIP(hipMemcpy(p, h, sizeof(T) * n, hipMemcpyHostToDevice));
    *d = static_cast<const T*>(p);
    return FA_SUCCESS;
}

struct GridKey {
    float gmin[3], gmax[3], maxwidth;
    bool operator<(const GridKey& o) const
    {
        return std::make_tuple(gmin[0], gmin[1], gmin[2], gmax[0], gmax[1], gmax[2], maxwidth) <
               std::make_tuple(o.gmin[0], o.gmin[1], o.gmin[2], o.gmax[0], o.gmax[1], o.gmax[2], o.maxwidth);
    }
};

int validate_flex(const fa_receptor_desc& r)
{
    if (r.n_flex < 0 || r.n_flex > FA_MAX_FLEX_RESIDUES || (r.n_flex > 0 && !r.flex))
        return fa_fail(FA_ERR_LIMIT, "between 0 and FA_MAX_FLEX_RESIDUES flexible residues");
    int end = 0, n_sc = 0;
    for (int k = 0; k < r.n_flex; ++k) {
        const fa_flex_residue& f = r.flex[k];
        const std::string tag = "flexible residue " + std::to_string(k) + ": ";
        if (f.n_atoms < 1 || f.n_atoms > FA_MAX_FLEX_ATOMS || f.n_rot < 1)
            return fa_fail(FA_ERR_LIMIT, tag + "1..FA_MAX_FLEX_ATOMS atoms and at least one conformation");
        if (f.first < end || f.first + f.n_atoms > r.n_atoms)
            return fa_fail(FA_ERR_INVALID, tag + "atoms must lie in the receptor, after the previous residue's");
        if (!f.scorable || !f.rot_xyz || !f.bonded) return fa_fail(FA_ERR_INVALID, tag + "null array");
        for (int i = 0; i < f.n_atoms; ++i) {
            for (int q = 0; q < 3; ++q)
                if (f.rot_xyz[3 * i + q] != r.xyz[3 * (f.first + i) + q])
                    return fa_fail(FA_ERR_INVALID, tag + "conformation 0 must equal the receptor coordinates");
            n_sc += f.scorable[i] ? 1 : 0;
        }
        end = f.first + f.n_atoms;
    }
    if (n_sc >= FA_MAX_SCORABLE) return fa_fail(FA_ERR_LIMIT, "too many scorable side-chain atoms");
    return FA_SUCCESS;
}

int validate_ligand(const fa_ligand_desc& L, int idx, const fa_receptor_desc& R, int n_sc)
{
    const std::string tag = "ligand " + std::to_string(idx) + ": ";
    if (L.n_atoms < 3 || L.n_atoms > FA_MAX_LIG_ATOMS)
        return fa_fail(FA_ERR_LIMIT, tag + "n_atoms must be in [3, " + std::to_string(FA_MAX_LIG_ATOMS) + "]");
    if (L.n_atoms + n_sc > FA_MAX_SCORABLE)
        return fa_fail(FA_ERR_LIMIT, tag + "ligand and side-chain atoms above FA_MAX_SCORABLE");
    if (L.n_genes < 1 || L.n_genes > FA_MAX_GENES)
        return fa_fail(FA_ERR_LIMIT, tag + "n_genes must be in [1, " + std::to_string(FA_MAX_GENES) + "]");
    if (L.n_grid < 1) return fa_fail(FA_ERR_INVALID, tag + "empty cleft grid");
    if (!L.rec || !L.dis || !L.ang || !L.dih || !L.shift || !L.radius || !L.type || !L.mov || !L.bonded ||
        !L.gene_typ || !L.gene_atm || !L.gene_min || !L.gene_max || !L.grid_ic)
        return fa_fail(FA_ERR_INVALID, tag + "null array");
    const int n = L.n_atoms;
    std::vector<int> pos(n, -1);
    for (int k = 0; k < n; ++k) {
        const int a = L.mov[k];
        if (a < 0 || a >= n || pos[a] != -1) return fa_fail(FA_ERR_INVALID, tag + "mov is not a permutation");
        pos[a] = k;
    }
    for (int a = 0; a < n; ++a) {
        for (int i = 0; i < 3; ++i) {
            const int r = L.rec[4 * a + i];
            if (r < -1 || r >= n) return fa_fail(FA_ERR_INVALID, tag + "rec out of range");
            if (r >= 0 && pos[r] >= pos[a]) return fa_fail(FA_ERR_INVALID, tag + "rec atom rebuilt after its dependant");
        }
        const int r3 = L.rec[4 * a + 3];
        if (r3 < -1 || r3 >= n) return fa_fail(FA_ERR_INVALID, tag + "rec[3] out of range");
        if (L.type[a] < 1) return fa_fail(FA_ERR_INVALID, tag + "types are 1-based");
    }
    for (int g = 0; g < L.n_genes; ++g) {
        const int t = L.gene_typ[g], a = L.gene_atm[g];
        if (t == FA_GENE_ROT) {  // rotamer of flexible residue a
            if (a < 0 || a >= R.n_flex) return fa_fail(FA_ERR_INVALID, tag + "rotamer gene names no flexible residue");
            if (L.gene_min[g] < 0.0 || L.gene_max[g] > (double)(R.flex[a].n_rot - 1))
                return fa_fail(FA_ERR_INVALID, tag + "rotamer gene limits outside the residue's conformations");
            continue;
        }
        if (t < FA_GENE_GRID || t > FA_GENE_DIH)
            return fa_fail(FA_ERR_INVALID, tag + "unsupported gene type (normal-mode genes are not implemented)");
        if (a < 0 || a >= n) return fa_fail(FA_ERR_INVALID, tag + "gene atom out of range");
        if (t == FA_GENE_DIH && L.rec[4 * a + 3] >= 0) {  // shift chain must return to the gene atom
            int cat = L.rec[4 * a + 3], steps = 0;
            while (cat != a && steps <= n) {
                cat = L.rec[4 * cat + 3];
                ++steps;
                if (cat < 0) break;
            }
            if (cat != a) return fa_fail(FA_ERR_INVALID, tag + "rec[3] shift chain does not close");
        }
        if (t == FA_GENE_GRID && (L.gene_min[g] < 0.0 || L.gene_max[g] > (double)(L.n_grid - 1)))
            return fa_fail(FA_ERR_INVALID, tag + "grid gene limits outside the cleft grid");
    }
    if (L.n_cons < 0 || (L.n_cons > 0 && !L.cons)) return fa_fail(FA_ERR_INVALID, tag + "bad constraint list");
    for (int q = 0; q < L.n_cons; ++q) {
        const fa_constraint& k = L.cons[q];
        if (k.atom < 0 || k.atom >= n) return fa_fail(FA_ERR_INVALID, tag + "constraint atom out of range");
        if (k.partner < 0 || (k.partner_is_ligand && (k.partner >= n || k.partner == k.atom)))
            return fa_fail(FA_ERR_INVALID, tag + "constraint partner out of range");
        if (k.type != 0 && k.type != 1) return fa_fail(FA_ERR_INVALID, tag + "constraint type must be 0 or 1");
    }
    return FA_SUCCESS;
}

}  // namespace

int fa_launch_poses(fa_ctx* c, int n_poses, const double* d_genes, int stride, const int* d_lig, fa_cf_result* d_out,
                  fa_contact_rec* d_recs, int rec_cap, int* d_reccnt, float* d_xyz, int xyz_stride,
                  hipStream_t stream)
{
    fa::DevBatch B{};
    B.n_poses = n_poses;
    B.gene_stride = stride;
    B.genes = d_genes;
    B.pose_lig = d_lig;
    B.out = d_out;
    B.recs = d_recs;
    B.rec_cap = rec_cap;
    B.rec_count = d_reccnt;
    B.lig_xyz = d_xyz;
    B.xyz_stride = xyz_stride;
    B.n_ligands = c->n_ligands;
    B.counters = static_cast<unsigned int*>(c->counters.p);
    B.totals = static_cast<unsigned long long*>(c->totals.p);
    B.ops = static_cast<unsigned long long*>(c->ops.p);
    FA_HIP(fa::launch_zero_u32(static_cast<unsigned int*>(c->counters.p), 4, stream));
    FA_HIP(fa::launch_zero_u32(static_cast<unsigned int*>(c->ops.p), (int)(sizeof(fa_opcount) / 4), stream));
    B.prof = nullptr;
    if (c->profile) {
        // 16 phase slots, then one cycle count per pose
        int rc = fa_ensure(c->prof, sizeof(unsigned long long) * (16 + (size_t)std::max(n_poses, 1)));
        if (rc) return rc;
        c->prof_poses = n_poses;
        B.prof = static_cast<unsigned long long*>(c->prof.p);
        FA_HIP(fa::launch_zero_u32(static_cast<unsigned int*>(c->prof.p), 32, stream));
    }
    B.near_idx = static_cast<int2*>(c->near_idx.p);
    B.near_xyz = static_cast<float4*>(c->near_xyz.p);
    B.rbox = static_cast<int*>(c->rbox.p);
    B.near_cap = c->near_cap;
    B.cell_lists = (c->shape_flags & FA_SHAPE_CELL_LISTS) ? 1 : 0;
    // one wave per resident slot (or the shape's cap); each wave loops over
    // poses (work queue)
    const int waves = c->shape_waves > 0 ? std::min(c->shape_waves, c->n_waves) : c->n_waves;
    FA_HIP(fa::launch_cf(c->rec, c->d_grids, c->d_ligs, c->prm, B, waves, stream));
    return FA_SUCCESS;
}

extern "C" {

const char* fa_last_error(void) { return g_err.c_str(); }

int fa_kernel_lds_bytes(void) { return fa::smem_bytes(); }

int fa_ctx_launch_waves(fa_ctx* c) { return c ? c->n_waves : 0; }

int fa_ctx_set_launch_shape(fa_ctx* c, int waves, int flags)
{
    if (!c) return fa_fail(FA_ERR_INVALID, "null context");
    if (waves < 0 || (flags & ~FA_SHAPE_CELL_LISTS)) return fa_fail(FA_ERR_INVALID, "bad launch shape");
    c->shape_waves = waves;
    c->shape_flags = flags;
    return FA_SUCCESS;
}

int fa_ctx_get_launch_shape(fa_ctx* c, int* waves, int* flags)
{
    if (!c || !waves || !flags) return fa_fail(FA_ERR_INVALID, "null argument");
    *waves = c->shape_waves > 0 ? std::min(c->shape_waves, c->n_waves) : c->n_waves;
    *flags = c->shape_flags;
    return FA_SUCCESS;
}

int fa_get_device_info(int device_id, fa_device_info* out)
{
    if (!out) return fa_fail(FA_ERR_INVALID, "null output");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fa_fail(FA_ERR_NODEVICE, "no HIP device");
    if (device_id < 0 || device_id >= n) return fa_fail(FA_ERR_INVALID, "device id out of range");
    hipDeviceProp_t p;
    FA_HIP(hipGetDeviceProperties(&p, device_id));
    std::memset(out, 0, sizeof(*out));
    std::snprintf(out->name, sizeof(out->name), "%s", p.name);
    out->compute_units = p.multiProcessorCount;
    out->global_mem = p.totalGlobalMem;
    out->gcn_arch = 0;
    std::sscanf(p.gcnArchName, "gfx%d", &out->gcn_arch);
    out->wavefront_size = p.warpSize;
    out->max_threads_per_cu = p.maxThreadsPerMultiProcessor;
    out->shared_mem_per_cu = (int)p.maxSharedMemoryPerMultiProcessor;
    return FA_SUCCESS;
}

int fa_ctx_create(int device_id, int max_pop, fa_ctx** out)
{
    if (!out) return fa_fail(FA_ERR_INVALID, "null output");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fa_fail(FA_ERR_NODEVICE, "no HIP device");
    if (device_id < 0 || device_id >= n) return fa_fail(FA_ERR_INVALID, "device id out of range");
    FA_HIP(hipSetDevice(device_id));
    hipDeviceProp_t p;
    FA_HIP(hipGetDeviceProperties(&p, device_id));
    if (int rc = fa_noalloc_init_device(device_id)) return rc;
    fa_ctx* c = new fa_ctx();
    c->device = device_id;
    c->n_cu = p.multiProcessorCount;
    c->max_pop = max_pop > 0 ? max_pop : 4096;
    int rc;
    if ((rc = fa_ensure(c->counters, 4 * sizeof(uint32_t))) != FA_SUCCESS ||
        (rc = fa_ensure(c->scratch, fa::lse_scratch_bytes())) != FA_SUCCESS ||
        (rc = fa_ensure(c->thermo, sizeof(fa_thermo))) != FA_SUCCESS ||
        (rc = fa_ensure(c->ops, sizeof(fa_opcount))) != FA_SUCCESS ||
        (rc = fa_ensure(c->prof, 16 * sizeof(unsigned long long))) != FA_SUCCESS ||
        (rc = fa_ensure(c->genes, sizeof(double) * (size_t)c->max_pop * 32)) != FA_SUCCESS ||
        (rc = fa_ensure(c->out, sizeof(fa_cf_result) * (size_t)c->max_pop)) != FA_SUCCESS ||
        (rc = fa_ensure(c->totals, FA_TOTALS_N * sizeof(unsigned long long))) != FA_SUCCESS) {
        delete c;
        return rc;
    }
    if (hipMemset(c->totals.p, 0, FA_TOTALS_N * sizeof(unsigned long long)) != hipSuccess) {
        delete c;
        return fa_fail(FA_ERR_HIP, "clear totals");
    }
    *out = c;
    return FA_SUCCESS;
}

int fa_ctx_destroy(fa_ctx* c)
{
    if (!c) return FA_SUCCESS;
    (void)hipSetDevice(c->device);
    (void)hipDeviceSynchronize();
    c->free_system();
    for (DevBuf* b : {&c->genes, &c->out, &c->counters, &c->recs, &c->reccnt, &c->xyz, &c->scratch, &c->cluster, &c->thermo,
                      &c->ops, &c->prof, &c->near_idx, &c->near_xyz, &c->rbox, &c->totals})
        if (b->p) (void)hipFree(b->p);
    delete c;
    return FA_SUCCESS;
}

int fa_upload_system(fa_ctx* c, const fa_receptor_desc* r, const fa_params* prm, int n_ligands,
                     const fa_ligand_desc* ligs)
{
    if (!c || !r || !prm || !ligs || n_ligands < 1) return fa_fail(FA_ERR_INVALID, "null argument");
    if (c->live_users > 0)
        return fa_fail(FA_ERR_INVALID, "a GA or docking pool still uses the resident system; destroy it first");
    if (r->n_atoms < 1 || !r->xyz || !r->radius || !r->type) return fa_fail(FA_ERR_INVALID, "empty receptor");
    if (prm->ntypes < 1 || !prm->emat) return fa_fail(FA_ERR_INVALID, "energy matrix missing");
    const size_t npairs = (size_t)prm->ntypes * prm->ntypes;
    if (prm->dens_off) {
        if (prm->dens_off[0] != 0) return fa_fail(FA_ERR_INVALID, "density offsets must start at 0");
        for (size_t k = 0; k < npairs; ++k)
            if (prm->dens_off[k + 1] < prm->dens_off[k]) return fa_fail(FA_ERR_INVALID, "density offsets must not decrease");
        if (prm->dens_off[npairs] > 0 && !prm->dens_xy) return fa_fail(FA_ERR_INVALID, "density nodes missing");
    }
    for (int i = 0; i < r->n_atoms; ++i)
        if (r->type[i] < 1 || r->type[i] > prm->ntypes) return fa_fail(FA_ERR_INVALID, "receptor type out of range");
    if (int rc = validate_flex(*r)) return rc;
    int n_sc = 0;
    for (int k = 0; k < r->n_flex; ++k)
        for (int i = 0; i < r->flex[k].n_atoms; ++i) n_sc += r->flex[k].scorable[i] ? 1 : 0;
    for (int li = 0; li < n_ligands; ++li) {
        int rc = validate_ligand(ligs[li], li, *r, n_sc);
        if (rc != FA_SUCCESS) return rc;
        for (int a = 0; a < ligs[li].n_atoms; ++a)
            if (ligs[li].type[a] > prm->ntypes) return fa_fail(FA_ERR_INVALID, "ligand type out of range");
    }
    FA_HIP(hipSetDevice(c->device));
    FA_HIP(hipDeviceSynchronize());
    c->free_system();
    c->max_genes = 0;
    int rc;
    const int Rn = r->n_atoms;
    // receptor atoms (calc order): x, y, z, radius packed for one 16-byte load
    std::vector<float> x(Rn), y(Rn), z(Rn);
    std::vector<float4> rxyzr(Rn);
    for (int i = 0; i < Rn; ++i) {
        x[i] = r->xyz[3 * i];
        y[i] = r->xyz[3 * i + 1];
        z[i] = r->xyz[3 * i + 2];
        rxyzr[i] = make_float4(x[i], y[i], z[i], r->radius[i]);
    }
    c->rec.n = Rn;
    if ((rc = upload(c, rxyzr.data(), Rn, &c->rec.xyzr)) || (rc = upload(c, r->type, Rn, &c->rec.type)))
        return rc;
    // flexible side chains: scorable atoms in calc order with their
    // conformations, residue ranges and bonded bits; they stay out of the
    // receptor cell lists (placed per pose)
    std::vector<unsigned char> scflag(Rn, 0);
    {
        std::vector<int> sc_calc, sc_res, sc_off, nrot;
        std::vector<float4> sc_rot;
        std::vector<int2> sc_range;
        std::vector<unsigned long long> sc_bits;
        for (int k = 0; k < r->n_flex; ++k) {
            const fa_flex_residue& f = r->flex[k];
            nrot.push_back(f.n_rot);
            for (int i = 0; i < f.n_atoms; ++i) {
                if (!f.scorable[i]) continue;
                const int ci = f.first + i;
                scflag[ci] = 1;
                sc_calc.push_back(ci);
                sc_res.push_back(k);
                sc_off.push_back((int)sc_rot.size());
                for (int q = 0; q < f.n_rot; ++q) {
                    const float* p = f.rot_xyz + ((size_t)q * f.n_atoms + i) * 3;
                    sc_rot.push_back(make_float4(p[0], p[1], p[2], r->radius[ci]));
                }
                sc_range.push_back(make_int2(f.first, f.first + f.n_atoms));
                unsigned long long b[2] = {0ull, 0ull};
                for (int j = 0; j < f.n_atoms; ++j)
                    if (f.bonded[i * f.n_atoms + j]) b[j >> 6] |= 1ull << (j & 63);
                sc_bits.push_back(b[0]);
                sc_bits.push_back(b[1]);
            }
        }
        c->rec.n_sc = (int)sc_calc.size();
        c->rec.n_flex = r->n_flex;
        c->n_sc = c->rec.n_sc;
        if (c->rec.n_sc > 0) {
            if ((rc = upload(c, sc_calc.data(), sc_calc.size(), &c->rec.sc_calc)) ||
                (rc = upload(c, sc_res.data(), sc_res.size(), &c->rec.sc_res)) ||
                (rc = upload(c, sc_off.data(), sc_off.size(), &c->rec.sc_rot_off)) ||
                (rc = upload(c, sc_rot.data(), sc_rot.size(), &c->rec.sc_rot)) ||
                (rc = upload(c, nrot.data(), nrot.size(), &c->rec.flex_nrot)) ||
                (rc = upload(c, sc_range.data(), sc_range.size(), &c->rec.sc_range)) ||
                (rc = upload(c, sc_bits.data(), sc_bits.size(), &c->rec.sc_bonded)) ||
                (rc = upload(c, scflag.data(), scflag.size(), &c->rec.sc_flag)))
                return rc;
        } else {
            c->rec.sc_calc = c->rec.sc_res = c->rec.sc_rot_off = c->rec.flex_nrot = nullptr;
            c->rec.sc_rot = nullptr;
            c->rec.sc_range = nullptr;
            c->rec.sc_bonded = nullptr;
            c->rec.sc_flag = nullptr;
        }
    }
    // parameters
    c->prm.permeability = prm->permeability;
    c->prm.dee_clash = prm->dee_clash;
    c->prm.solventterm = prm->solventterm;
    c->prm.intramolecular = prm->intramolecular;
    c->prm.ntypes = prm->ntypes;
    if ((rc = upload(c, prm->emat, npairs, &c->prm.emat))) return rc;
    c->prm.normalize_area = prm->normalize_area ? 1 : 0;
    c->prm.dens_off = nullptr;
    c->prm.dens_xy = nullptr;
    if (prm->dens_off && prm->dens_off[npairs] > 0) {
        const size_t nn = (size_t)prm->dens_off[npairs];
        std::vector<float2> xy(nn);
        for (size_t i = 0; i < nn; ++i) xy[i] = make_float2(prm->dens_xy[2 * i], prm->dens_xy[2 * i + 1]);
        if ((rc = upload(c, prm->dens_off, npairs + 1, &c->prm.dens_off)) || (rc = upload(c, xy.data(), nn, &c->prm.dens_xy)))
            return rc;
    }
    // grids: one receptor cell list per distinct (rounded bounds, width)
    // (LIB/Vcontacts.cpp:1594-1715 with no scorable atom outside the bounds)
    std::map<GridKey, int> grid_ids;
    std::vector<fa::DevGrid> grids;
    c->h_ligs.resize(n_ligands);
    c->lig_genes.resize(n_ligands);
    for (int li = 0; li < n_ligands; ++li) {
        const fa_ligand_desc& L = ligs[li];
        GridKey k;
        for (int j = 0; j < 3; ++j) {
            k.gmin[j] = (float)((int)L.globalmin[j] + 1.0);
            k.gmax[j] = (float)((int)L.globalmax[j] + 1.0);
        }
        k.maxwidth = L.maxwidth;
        auto it = grid_ids.find(k);
        int gid;
        if (it != grid_ids.end()) {
            gid = it->second;
        } else {
            fa::DevGrid G{};
            for (int j = 0; j < 3; ++j) {
                G.gmin[j] = k.gmin[j];
                G.gmax[j] = k.gmax[j];
            }
            G.maxwidth = k.maxwidth;
            G.dim = (int)(k.maxwidth / CELLSIZE) + 1;
            const long long dim2 = (long long)G.dim * G.dim, dim3 = dim2 * G.dim;
            if (G.dim < 1 || dim3 > (1LL << 26)) return fa_fail(FA_ERR_LIMIT, "cell grid too large");
            std::vector<int> box(Rn), cnt(dim3 + 1, 0);
            int n_rigid = 0;
            for (int i = 0; i < Rn; ++i) {
                if (scflag[i]) continue;  // side-chain atoms are placed per pose
                const long long b = (long long)(int)((x[i] - G.gmin[0]) / CELLSIZE) * dim2 +
                                    (long long)(int)((y[i] - G.gmin[1]) / CELLSIZE) * G.dim +
                                    (int)((z[i] - G.gmin[2]) / CELLSIZE);
                if (b < 0 || b >= dim3) return fa_fail(FA_ERR_INVALID, "receptor atom outside the cell grid");
                box[i] = (int)b;
                ++cnt[b + 1];
                ++n_rigid;
            }
            for (long long b = 0; b < dim3; ++b) cnt[b + 1] += cnt[b];
            std::vector<int> fill(cnt.begin(), cnt.end() - 1), atom(n_rigid);
            std::vector<float4> xyzr(n_rigid);
            for (int i = 0; i < Rn; ++i) {  // stable: calc order inside a cell
                if (scflag[i]) continue;
                const int pos = fill[box[i]]++;
                atom[pos] = i;
                xyzr[pos] = make_float4(x[i], y[i], z[i], r->radius[i]);
            }
            if ((rc = upload(c, cnt.data(), (size_t)dim3 + 1, &G.cell_start)) ||
                (rc = upload(c, atom.data(), n_rigid, &G.cell_atom)) ||
                (rc = upload(c, xyzr.data(), n_rigid, &G.cell_xyzr)))
                return rc;
            // neighbourhood lists: the 27-slot scan of every cell laid out once
            // (slot s names cell c + (s/9 - 1) dim^2 + ((s/3)%3 - 1) dim + s%3 - 1
            // when that index is inside the grid, LIB/Vcontacts.cpp:1738-1741)
            if (G.dim >= 3 && dim3 <= NB_MAX_CELLS && Rn < (1 << 27)) {  // entries pack index | slot << 27
                std::vector<int> nb_start((size_t)dim3 + 1, 0), nb_bnd((size_t)dim3 * 28);
                for (long long b = 0; b < dim3; ++b) {
                    long long n = 0;
                    for (int sl = 0; sl < 27; ++sl) {
                        nb_bnd[(size_t)b * 28 + sl] = (int)n;
                        const long long nbx = b + dim2 * (sl / 9 - 1) + G.dim * ((sl / 3) % 3 - 1) + sl % 3 - 1;
                        if (nbx >= 0 && nbx < dim3) n += cnt[nbx + 1] - cnt[nbx];
                    }
                    nb_bnd[(size_t)b * 28 + 27] = (int)n;
                    if ((long long)nb_start[b] + n > (1LL << 30)) return fa_fail(FA_ERR_LIMIT, "neighbourhood lists too large");
                    nb_start[b + 1] = nb_start[b] + (int)n;
                }
                const size_t total = (size_t)nb_start[dim3];
                std::vector<int> nb_atom(total);
                std::vector<float4> nb_xyzr(total);
                size_t w = 0;
                for (long long b = 0; b < dim3; ++b)
                    for (int sl = 0; sl < 27; ++sl) {
                        const long long nbx = b + dim2 * (sl / 9 - 1) + G.dim * ((sl / 3) % 3 - 1) + sl % 3 - 1;
                        if (nbx < 0 || nbx >= dim3) continue;
                        for (int q = cnt[nbx]; q < cnt[nbx + 1]; ++q) {
                            nb_atom[w] = atom[q] | (sl << 27);
                            nb_xyzr[w] = xyzr[q];
                            ++w;
                        }
                    }
                if ((rc = upload(c, nb_start.data(), nb_start.size(), &G.nb_start)) ||
                    (rc = upload(c, nb_bnd.data(), nb_bnd.size(), &G.nb_bnd)) ||
                    (total && ((rc = upload(c, nb_atom.data(), total, &G.nb_atom)) ||
                               (rc = upload(c, nb_xyzr.data(), total, &G.nb_xyzr)))))
                    return rc;
            }
            gid = (int)grids.size();
            grids.push_back(G);
            grid_ids[k] = gid;
        }
        // ligand topology
        fa::DevLigand D{};
        D.n_atoms = L.n_atoms;
        D.n_genes = L.n_genes;
        D.n_grid = L.n_grid;
        D.grid_id = gid;
        for (int j = 0; j < 3; ++j) {
            D.ori[j] = L.ori[j];
            D.globalmin[j] = L.globalmin[j];
            D.globalmax[j] = L.globalmax[j];
        }
        D.maxwidth = L.maxwidth;
        const int n = L.n_atoms;
        std::vector<unsigned long long> bits((size_t)n * 2, 0ull);
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j)
                if (L.bonded[i * n + j]) bits[2 * i + (j >> 6)] |= 1ull << (j & 63);
        if ((rc = upload(c, L.rec, (size_t)4 * n, &D.rec)) || (rc = upload(c, L.dis, n, &D.dis)) ||
            (rc = upload(c, L.ang, n, &D.ang)) || (rc = upload(c, L.dih, n, &D.dih)) ||
            (rc = upload(c, L.shift, n, &D.shift)) || (rc = upload(c, L.radius, n, &D.radius)) ||
            (rc = upload(c, L.type, n, &D.type)) || (rc = upload(c, L.mov, n, &D.mov)) ||
            (rc = upload(c, bits.data(), bits.size(), &D.bonded)) ||
            (rc = upload(c, L.gene_typ, L.n_genes, &D.gene_typ)) ||
            (rc = upload(c, L.gene_atm, L.n_genes, &D.gene_atm)) ||
            (rc = upload(c, L.gene_min, L.n_genes, &D.gene_min)) ||
            (rc = upload(c, L.gene_max, L.n_genes, &D.gene_max)) ||
            (rc = upload(c, L.grid_ic, (size_t)3 * L.n_grid, &D.grid_ic)))
            return rc;
        // constraints per scorable atom: a ligand pair is listed on both of
        // its atoms (each counts its own starting penalty)
        D.cons_off = nullptr;
        D.con0 = 0.0;
        if (L.n_cons > 0) {
            std::vector<std::vector<int>> at(n);
            for (int q = 0; q < L.n_cons; ++q) {
                if (L.cons[q].partner >= Rn && !L.cons[q].partner_is_ligand)
                    return fa_fail(FA_ERR_INVALID, "ligand " + std::to_string(li) + ": constraint partner out of range");
                at[L.cons[q].atom].push_back(q);
                if (L.cons[q].partner_is_ligand) at[L.cons[q].partner].push_back(q);
            }
            std::vector<int> off(n + 1, 0), part, typ;
            std::vector<float> bond, mx;
            for (int a = 0; a < n; ++a) {
                for (int q : at[a]) {
                    const fa_constraint& k = L.cons[q];
                    const int other = (k.partner_is_ligand && k.partner == a) ? k.atom : k.partner;
                    part.push_back(k.partner_is_ligand ? Rn + other : other);
                    typ.push_back(k.type);
                    bond.push_back(k.bond_len);
                    mx.push_back(k.max_dist);
                    if (k.type == 1) D.con0 += 1.0e3;  // KDIST, LIB/flexaid.h:83
                }
                off[a + 1] = (int)part.size();
            }
            const size_t m = part.size();
            if ((rc = upload(c, off.data(), off.size(), &D.cons_off)) || (rc = upload(c, part.data(), m, &D.cons_part)) ||
                (rc = upload(c, typ.data(), m, &D.cons_type)) || (rc = upload(c, bond.data(), m, &D.cons_bond)) ||
                (rc = upload(c, mx.data(), m, &D.cons_max)))
                return rc;
        }
        c->h_ligs[li] = D;
        c->lig_genes[li] = L.n_genes;
        c->max_genes = std::max(c->max_genes, L.n_genes);
    }
    const fa::DevGrid* dg = nullptr;
    const fa::DevLigand* dl = nullptr;
    if ((rc = upload(c, grids.data(), grids.size(), &dg)) || (rc = upload(c, c->h_ligs.data(), n_ligands, &dl)))
        return rc;
    c->d_grids = const_cast<fa::DevGrid*>(dg);
    c->d_ligs = const_cast<fa::DevLigand*>(dl);
    c->n_ligands = n_ligands;
    // neighbour-list scratch: NEAR_PER_ATOM entries per ligand atom and wave
    // (in-range candidates of all atoms of a pose; ~77 per atom at most on
    // dense receptors, overflow is reported as a capacity status)
    int max_l = 0;
    for (int li = 0; li < n_ligands; ++li) max_l = std::max(max_l, ligs[li].n_atoms + c->n_sc);
    c->max_lig_atoms = max_l;
    const int per_cu = std::max(1, fa::cf_waves_per_cu());
    c->n_waves = c->n_cu * per_cu;
    c->near_cap = max_l * NEAR_PER_ATOM + NEAR_TAIL;
    const size_t n_near = (size_t)c->n_waves * c->near_cap;
    if ((rc = fa_ensure(c->near_idx, sizeof(int2) * n_near)) || (rc = fa_ensure(c->near_xyz, sizeof(float4) * n_near)) ||
        (rc = fa_ensure(c->rbox, sizeof(int) * (size_t)c->n_waves * 2 * (size_t)((Rn + 3) & ~3))))
        return rc;
    c->have_system = true;
    return FA_SUCCESS;
}


int fa_eval_batch_device(fa_ctx* c, int n_poses, const double* genes, int gene_stride, const int* pose_lig,
                         fa_cf_result* out, void* stream)
{
    if (!c) return fa_fail(FA_ERR_INVALID, "null context");
    if (!c->have_system) return fa_fail(FA_ERR_NOSYSTEM, "no system uploaded");
    if (n_poses < 0 || (n_poses > 0 && (!genes || !out))) return fa_fail(FA_ERR_INVALID, "bad batch");
    if (gene_stride < c->max_genes)
        return fa_fail(FA_ERR_INVALID, "gene_stride below the largest gene count of the uploaded ligands");
    FA_HIP(hipSetDevice(c->device));
    return fa_launch_poses(c, n_poses, genes, gene_stride, pose_lig, out, nullptr, 0, nullptr, nullptr, 0,
                  static_cast<hipStream_t>(stream));
}

static int eval_host(fa_ctx* c, int ligand, int n_poses, const double* genes, fa_cf_result* out, float* lig_xyz,
                     fa_contact_rec* recs, int rec_cap, int* rec_count)
{
    if (!c) return fa_fail(FA_ERR_INVALID, "null context");
    if (!c->have_system) return fa_fail(FA_ERR_NOSYSTEM, "no system uploaded");
    if (ligand < 0 || ligand >= c->n_ligands) return fa_fail(FA_ERR_INVALID, "ligand index out of range");
    if (n_poses < 0 || (n_poses > 0 && (!genes || !out))) return fa_fail(FA_ERR_INVALID, "bad batch");
    if (n_poses == 0) return FA_SUCCESS;
    FA_HIP(hipSetDevice(c->device));
    const int G = c->lig_genes[ligand];
    int rc;
    if ((rc = fa_ensure(c->genes, sizeof(double) * (size_t)n_poses * G)) ||
        (rc = fa_ensure(c->out, sizeof(fa_cf_result) * (size_t)n_poses)))
        return rc;
    std::vector<int> lig(n_poses, ligand);
    DevBuf ligbuf;
    if ((rc = fa_ensure(ligbuf, sizeof(int) * (size_t)n_poses))) return rc;
    fa_contact_rec* d_recs = nullptr;
    int* d_cnt = nullptr;
    float* d_xyz = nullptr;
    const int L = c->h_ligs[ligand].n_atoms;
    if (recs && rec_count && rec_cap > 0) {
        if ((rc = fa_ensure(c->recs, sizeof(fa_contact_rec) * (size_t)n_poses * rec_cap)) ||
            (rc = fa_ensure(c->reccnt, sizeof(int) * (size_t)n_poses)))
            return rc;
        d_recs = static_cast<fa_contact_rec*>(c->recs.p);
        d_cnt = static_cast<int*>(c->reccnt.p);
    }
    if (lig_xyz) {
        if ((rc = fa_ensure(c->xyz, sizeof(float) * 3 * (size_t)n_poses * L))) return rc;
        d_xyz = static_cast<float*>(c->xyz.p);
    }
    hipError_t e = hipMemcpy(c->genes.p, genes, sizeof(double) * (size_t)n_poses * G, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(ligbuf.p, lig.data(), sizeof(int) * (size_t)n_poses, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(ligbuf.p);
        return fa_fail(FA_ERR_HIP, std::string("upload genes: ") + hipGetErrorString(e));
    }
    rc = fa_launch_poses(c, n_poses, static_cast<const double*>(c->genes.p), G, static_cast<const int*>(ligbuf.p),
                static_cast<fa_cf_result*>(c->out.p), d_recs, rec_cap, d_cnt, d_xyz, L, nullptr);
    if (rc == FA_SUCCESS) {
        e = hipMemcpy(out, c->out.p, sizeof(fa_cf_result) * (size_t)n_poses, hipMemcpyDeviceToHost);
        if (e == hipSuccess && d_recs)
            e = hipMemcpy(recs, d_recs, sizeof(fa_contact_rec) * (size_t)n_poses * rec_cap, hipMemcpyDeviceToHost);
        if (e == hipSuccess && d_cnt)
            e = hipMemcpy(rec_count, d_cnt, sizeof(int) * (size_t)n_poses, hipMemcpyDeviceToHost);
        if (e == hipSuccess && d_xyz)
            e = hipMemcpy(lig_xyz, d_xyz, sizeof(float) * 3 * (size_t)n_poses * L, hipMemcpyDeviceToHost);
        if (e == hipSuccess)
            e = hipMemcpy(c->last_counters.data(), c->counters.p, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = fa_fail(FA_ERR_HIP, std::string("download results: ") + hipGetErrorString(e));
    }
    (void)hipFree(ligbuf.p);
    return rc;
}

int fa_eval_batch(fa_ctx* c, int ligand, int n_poses, const double* genes, fa_cf_result* out)
{
    return eval_host(c, ligand, n_poses, genes, out, nullptr, nullptr, 0, nullptr);
}

int fa_eval_batch_debug(fa_ctx* c, int ligand, int n_poses, const double* genes, fa_cf_result* out, float* lig_xyz,
                        fa_contact_rec* recs, int rec_cap, int* rec_count)
{
    return eval_host(c, ligand, n_poses, genes, out, lig_xyz, recs, rec_cap, rec_count);
}

int fa_thermo_device(fa_ctx* c, const double* energy, int stride, const int* count, int64_t n, double T,
                     fa_thermo* out, void* stream)
{
    if (!c || !energy || !out || n < 1 || stride < 1) return fa_fail(FA_ERR_INVALID, "bad thermo arguments");
    if (!(T > 0.0)) return fa_fail(FA_ERR_INVALID, "temperature must be > 0");
    FA_HIP(hipSetDevice(c->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    FA_HIP(fa::launch_thermo(energy, stride, count, n, T, c->scratch.p, static_cast<fa_thermo*>(c->thermo.p), s));
    FA_HIP(hipMemcpyAsync(out, c->thermo.p, sizeof(fa_thermo), hipMemcpyDeviceToHost, s));
    FA_HIP(hipStreamSynchronize(s));
    return FA_SUCCESS;
}

int fa_thermo_host(fa_ctx* c, const double* energy, const int* count, int64_t n, double T, fa_thermo* out)
{
    if (!c || !energy || !out || n < 1) return fa_fail(FA_ERR_INVALID, "bad thermo arguments");
    FA_HIP(hipSetDevice(c->device));
    DevBuf e, k;
    int rc = fa_ensure(e, sizeof(double) * (size_t)n);
    if (rc) return rc;
    if (count && (rc = fa_ensure(k, sizeof(int) * (size_t)n))) {
        (void)hipFree(e.p);
        return rc;
    }
    hipError_t he = hipMemcpy(e.p, energy, sizeof(double) * (size_t)n, hipMemcpyHostToDevice);
    if (he == hipSuccess && count) he = hipMemcpy(k.p, count, sizeof(int) * (size_t)n, hipMemcpyHostToDevice);
    if (he != hipSuccess)
        rc = fa_fail(FA_ERR_HIP, std::string("thermo upload: ") + hipGetErrorString(he));
    else
        rc = fa_thermo_device(c, static_cast<const double*>(e.p), 1, static_cast<const int*>(k.p), n, T, out, nullptr);
    (void)hipFree(e.p);
    if (k.p) (void)hipFree(k.p);
    return rc;
}

int fa_upload_emat(fa_ctx* c, const float* emat, int ntypes)
{
    if (!c || !emat) return fa_fail(FA_ERR_INVALID, "null argument");
    if (!c->have_system) return fa_fail(FA_ERR_NOSYSTEM, "no system uploaded");
    if (ntypes != c->prm.ntypes) return fa_fail(FA_ERR_INVALID, "ntypes differs from the uploaded system");
    FA_HIP(hipSetDevice(c->device));
    FA_HIP(hipDeviceSynchronize());
    FA_HIP(hipMemcpy(const_cast<float*>(c->prm.emat), emat, sizeof(float) * (size_t)ntypes * ntypes,
                     hipMemcpyHostToDevice));
    return FA_SUCCESS;
}

int fa_boltzmann_device(fa_ctx* c, const double* energy, int stride, int64_t n, double T, double* weights,
                        double* log_Z, void* stream)
{
    if (!c || !energy || !weights || n < 1 || stride < 1) return fa_fail(FA_ERR_INVALID, "bad boltzmann arguments");
    if (!(T > 0.0)) return fa_fail(FA_ERR_INVALID, "temperature must be > 0");
    FA_HIP(hipSetDevice(c->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    FA_HIP(fa::launch_boltzmann(energy, stride, n, T, c->scratch.p, weights, s));
    if (log_Z) {
        double ms[2];
        FA_HIP(hipMemcpyAsync(ms, c->scratch.p, sizeof(ms), hipMemcpyDeviceToHost, s));
        FA_HIP(hipStreamSynchronize(s));
        *log_Z = ms[0] + std::log(ms[1]);
    }
    return FA_SUCCESS;
}

int fa_log_sum_exp_device(fa_ctx* c, const double* values, int stride, int64_t n, double* out, void* stream)
{
    if (!c || !out || stride < 1 || (n > 0 && !values)) return fa_fail(FA_ERR_INVALID, "bad log_sum_exp arguments");
    if (n < 1) {
        *out = -HUGE_VAL;
        return FA_SUCCESS;
    }
    FA_HIP(hipSetDevice(c->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    FA_HIP(fa::launch_lse_values(values, stride, n, c->scratch.p, s));
    double ms[2];
    FA_HIP(hipMemcpyAsync(ms, c->scratch.p, sizeof(ms), hipMemcpyDeviceToHost, s));
    FA_HIP(hipStreamSynchronize(s));
    *out = std::isfinite(ms[0]) ? ms[0] + std::log(ms[1]) : ms[0];
    return FA_SUCCESS;
}

int fa_grand_device(fa_ctx* c, const double* log_Z, int stride, const double* log_c, int n, double* out, void* stream)
{
    if (!c || !log_Z || !log_c || !out || n < 0 || stride < 1) return fa_fail(FA_ERR_INVALID, "bad grand arguments");
    FA_HIP(hipSetDevice(c->device));
    FA_HIP(fa::launch_grand(log_Z, stride, log_c, n, out, static_cast<hipStream_t>(stream)));
    return FA_SUCCESS;
}

int fa_ctx_totals(fa_ctx* c, uint64_t* totals4)
{
    if (!c || !totals4) return fa_fail(FA_ERR_INVALID, "null argument");
    FA_HIP(hipSetDevice(c->device));
    FA_HIP(hipDeviceSynchronize());
    FA_HIP(hipMemcpy(totals4, c->totals.p, 4 * sizeof(uint64_t), hipMemcpyDeviceToHost));
    return FA_SUCCESS;
}

int fa_ctx_totals_n(fa_ctx* c, uint64_t* totals, int n)
{
    if (!c || (!totals && n > 0)) return fa_fail(FA_ERR_INVALID, "null argument");
    if (n < 0 || n > FA_TOTALS_N) return fa_fail(FA_ERR_INVALID, "n outside [0, FA_TOTALS_N]");
    FA_HIP(hipSetDevice(c->device));
    FA_HIP(hipDeviceSynchronize());
    if (n > 0) FA_HIP(hipMemcpy(totals, c->totals.p, (size_t)n * sizeof(uint64_t), hipMemcpyDeviceToHost));
    return FA_SUCCESS;
}

int fa_ctx_reset_totals(fa_ctx* c)
{
    if (!c) return fa_fail(FA_ERR_INVALID, "null context");
    FA_HIP(hipSetDevice(c->device));
    FA_HIP(hipDeviceSynchronize());
    FA_HIP(hipMemset(c->totals.p, 0, FA_TOTALS_N * sizeof(uint64_t)));
    return FA_SUCCESS;
}

int fa_last_counters(fa_ctx* c, uint32_t* slow_grid, uint32_t* capacity)
{
    if (!c) return fa_fail(FA_ERR_INVALID, "null context");
    FA_HIP(hipSetDevice(c->device));
    FA_HIP(hipDeviceSynchronize());
    FA_HIP(hipMemcpy(c->last_counters.data(), c->counters.p, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (slow_grid) *slow_grid = c->last_counters[0];
    if (capacity) *capacity = c->last_counters[1];
    return FA_SUCCESS;
}

int fa_set_phase_profiling(fa_ctx* c, int enable)
{
    if (!c) return fa_fail(FA_ERR_INVALID, "null context");
    if (c->live_users > 0 && (enable != 0) != c->profile)
        return fa_fail(FA_ERR_INVALID, "a GA or docking pool uses this context; destroy it before changing profiling");
    c->profile = enable != 0;
    return FA_SUCCESS;
}

int fa_last_phase_cycles(fa_ctx* c, uint64_t* cycles16)
{
    if (!c || !cycles16) return fa_fail(FA_ERR_INVALID, "null argument");
    FA_HIP(hipSetDevice(c->device));
    FA_HIP(hipDeviceSynchronize());
    FA_HIP(hipMemcpy(cycles16, c->prof.p, 16 * sizeof(uint64_t), hipMemcpyDeviceToHost));
    return FA_SUCCESS;
}

int fa_last_pose_cycles(fa_ctx* c, int n, uint64_t* cycles)
{
    if (!c || !cycles || n < 0) return fa_fail(FA_ERR_INVALID, "null argument");
    if (!c->profile || n > c->prof_poses) return fa_fail(FA_ERR_INVALID, "no profiled launch of that many poses");
    FA_HIP(hipSetDevice(c->device));
    FA_HIP(hipDeviceSynchronize());
    FA_HIP(hipMemcpy(cycles, static_cast<unsigned long long*>(c->prof.p) + 16, sizeof(uint64_t) * n,
                     hipMemcpyDeviceToHost));
    return FA_SUCCESS;
}

int fa_last_opcount(fa_ctx* c, fa_opcount* out)
{
    if (!c || !out) return fa_fail(FA_ERR_INVALID, "null argument");
    FA_HIP(hipSetDevice(c->device));
    FA_HIP(hipDeviceSynchronize());
    FA_HIP(hipMemcpy(out, c->ops.p, sizeof(fa_opcount), hipMemcpyDeviceToHost));
    return FA_SUCCESS;
}

}  // extern "C"
