// flexaids_rocm_bridge.h — reference-side bridge: the code a FlexAID∆S
// maintainer adds under LIB/ to score on the GPU through this library with
// the reference's own data structures and call signatures.
//
// It includes the reference's headers (flexaid.h, gaboom.h, Vcontacts.h,
// VoronoiCFBatch.h, UnifiedHardwareDispatch.h) and this library's C++ layer
// (hip_eval.h), so it compiles only inside the reference's tree, with
// -D__HIP_PLATFORM_AMD__ -I/opt/rocm/include and this repository's include/
// on the include path, and links libflexaidds_amd.so.
//
//   system_from_reference  — the reference's set-up (LIB/top.cpp,
//       add2_optimiz_vec.cpp:89-197, create_rebuild_list.cpp:8-50,
//       calc_cleftic.cpp:5-34, update_bonded.cpp:8-94, build_rotamers.cpp:
//       103-172, update_optres.cpp:10-30, read_constraints.cpp:100-125,
//       read_emat.cpp:86-138) copied once into the descriptors of fa_types.h;
//   batch_eval             — voronoi_cf::batch_eval's exact signature
//       (LIB/VoronoiCFBatch.h:184-247) scored on the device;
//   compute_boltzmann_batch / log_sum_exp_dispatch / select_backend — the
//       flexaids:: dispatch entries (LIB/UnifiedHardwareDispatch.cpp:1051-1073)
//       with their signatures and results, for a ROCm case in the reference's
//       dispatch (see the patch at the end of this file).
#pragma once

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <limits>
#include <span>
#include <stdexcept>
#include <vector>

#include "flexaid.h"
#include "gaboom.h"
#include "Vcontacts.h"
#include "VoronoiCFBatch.h"
#include "UnifiedHardwareDispatch.h"
#include "../hip_eval.h"  // this library's (the reference tree has its own hip_eval.h)

namespace flexaids_rocm {

// Owns every array the descriptors point to.
struct FaSystem {
    std::vector<float> rxyz, rrad;
    std::vector<int> rtype;
    std::vector<std::vector<uint8_t>> flex_scorable, flex_bonded;
    std::vector<std::vector<float>> flex_xyz;
    std::vector<fa_flex_residue> flex;
    std::vector<float> ldis, lang, ldih, lshift, lrad, grid_ic;
    std::vector<int> lrec, ltype, lmov, gtyp, gatm;
    std::vector<uint8_t> lbonded;
    std::vector<double> gmin, gmax;
    std::vector<fa_constraint> cons;
    std::vector<float> emat, dens_xy;
    std::vector<int> dens_off;
    fa_receptor_desc rec{};
    fa_ligand_desc lig{};
    fa_params prm{};
};

// The optimised ligand (FA->optres entry of type 1) and the flexible side
// chains (type 0 entries, build_rotamers' layout: conformation r of residue q
// at atoms fatm[r]..latm[r], r = 0 the input one) become the ligand and the
// flexible residues of the descriptors.  Receptor atoms are the atoms of the
// other residues in residue order at conformation 0 (index_protein's calc
// order, LIB/Vcontacts.cpp:1594-1625).
inline void system_from_reference(const FA_Global* FA, const atom* atoms, const resid* residue,
                                  const gridpoint* cleftgrid, std::span<const genlim> gene_lim, FaSystem& S)
{
    S = FaSystem{};
    int lig_res = -1;
    std::vector<int> sc_res;  // residue numbers of the flexible side chains
    for (int k = 0; k < FA->num_optres; ++k) {
        if (FA->optres[k].type == 1 && lig_res < 0)
            lig_res = FA->optres[k].rnum;
        else if (FA->optres[k].type == 0)
            sc_res.push_back(FA->optres[k].rnum);
    }
    if (lig_res < 0) throw std::invalid_argument("system_from_reference: no optimised ligand residue");
    if ((int)gene_lim.size() != FA->npar) throw std::invalid_argument("system_from_reference: gene_lim size");
    const int first = residue[lig_res].fatm[0], last = residue[lig_res].latm[0], L = last - first + 1;
    auto lidx = [&](int a) { return a >= first && a <= last ? a - first : -1; };

    // receptor atoms in calc order, conformation 0
    std::vector<int> to_rec(FA->atm_cnt + 2, -1);
    for (int q = 1; q <= FA->res_cnt; ++q) {
        if (q == lig_res) continue;
        for (int i = residue[q].fatm[0]; i <= residue[q].latm[0]; ++i) {
            to_rec[i] = (int)S.rtype.size();
            S.rxyz.insert(S.rxyz.end(), {atoms[i].coor[0], atoms[i].coor[1], atoms[i].coor[2]});
            S.rrad.push_back(atoms[i].radius);
            S.rtype.push_back(atoms[i].type);
        }
    }
    // flexible side chains, by position in calc order
    std::sort(sc_res.begin(), sc_res.end(),
              [&](int a, int b) { return to_rec[residue[a].fatm[0]] < to_rec[residue[b].fatm[0]]; });
    const int nf = (int)sc_res.size();
    S.flex_scorable.resize(nf);
    S.flex_bonded.resize(nf);
    S.flex_xyz.resize(nf);
    for (int k = 0; k < nf; ++k) {
        const resid& r = residue[sc_res[k]];
        const int f0 = r.fatm[0], n = r.latm[0] - f0 + 1, nrot = r.trot + 1;
        for (int w = 0; w < n; ++w) S.flex_scorable[k].push_back(atoms[f0 + w].optres != nullptr ? 1 : 0);
        for (int rr = 0; rr < nrot; ++rr)
            for (int w = 0; w < n; ++w) {
                const float* c = atoms[r.fatm[rr] + w].coor;
                S.flex_xyz[k].insert(S.flex_xyz[k].end(), {c[0], c[1], c[2]});
            }
        S.flex_bonded[k].assign((size_t)n * n, 0);
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) S.flex_bonded[k][(size_t)i * n + j] = r.bonded[i][j] >= 0 ? 1 : 0;
        S.flex.push_back(fa_flex_residue{to_rec[f0], n, nrot, S.flex_scorable[k].data(), S.flex_xyz[k].data(),
                                         S.flex_bonded[k].data()});
    }
    auto flex_of_atom = [&](int a) {
        for (int k = 0; k < nf; ++k)
            if (residue[sc_res[k]].fatm[0] == a) return k;
        return -1;
    };

    // ligand, atom order: internal coordinates, rebuild references (0, the
    // frame, becomes -1), rebuild order, bloops-excluded pairs
    for (int i = first; i <= last; ++i) {
        const atom& a = atoms[i];
        for (int k = 0; k < 4; ++k) S.lrec.push_back(a.rec[k] ? lidx(a.rec[k]) : -1);
        S.ldis.push_back(a.dis);
        S.lang.push_back(a.ang);
        S.ldih.push_back(a.dih);
        S.lshift.push_back(a.shift);
        S.lrad.push_back(a.radius);
        S.ltype.push_back(a.type);
    }
    for (int k = 0; k < FA->nmov[0]; ++k) S.lmov.push_back(lidx(FA->mov[0][k]));
    S.lbonded.assign((size_t)L * L, 0);
    for (int i = 0; i < L; ++i)
        for (int j = 0; j < L; ++j) S.lbonded[(size_t)i * L + j] = residue[lig_res].bonded[i][j] >= 0 ? 1 : 0;
    // gene map and limits (rotamer genes name a side chain's fatm[0])
    for (int g = 0; g < FA->npar; ++g) {
        const int typ = FA->map_par[g].typ;
        S.gtyp.push_back(typ);
        S.gatm.push_back(typ == FA_GENE_ROT ? flex_of_atom(FA->map_par[g].atm) : lidx(FA->map_par[g].atm));
        S.gmin.push_back(gene_lim[(size_t)g].min);
        S.gmax.push_back(gene_lim[(size_t)g].max);
    }
    for (int g = 0; g < FA->num_grd; ++g)
        S.grid_ic.insert(S.grid_ic.end(), {cleftgrid[g].dis, cleftgrid[g].ang, cleftgrid[g].dih});
    // distance constraints in FA->constraints order, named from their ligand end
    for (int q = 0; q < FA->num_constraints; ++q) {
        const constraint& c = FA->constraints[q];
        const bool one = lidx(c.inum1) >= 0;
        const int a = one ? c.inum1 : c.inum2, p = one ? c.inum2 : c.inum1;
        if (lidx(a) < 0) throw std::invalid_argument("system_from_reference: a constraint without a ligand atom");
        const bool pl = lidx(p) >= 0;
        S.cons.push_back(fa_constraint{lidx(a), pl ? lidx(p) : to_rec[p], pl ? 1 : 0, c.type, c.bond_len, c.max_dist});
    }
    // weight-1 pairs and density functions, in file order
    const int nt = FA->ntypes;
    S.emat.assign((size_t)nt * nt, 0.0f);
    S.dens_off.assign((size_t)nt * nt + 1, 0);
    for (int t = 0; t < nt * nt; ++t) {
        const energy_matrix& m = FA->energy_matrix[t];
        S.dens_off[(size_t)t + 1] = S.dens_off[(size_t)t];
        if (m.weight) {
            S.emat[(size_t)t] = m.energy_values->y;
            continue;
        }
        for (const energy_values* v = m.energy_values; v; v = v->next_value) {
            S.dens_xy.push_back(v->x);
            S.dens_xy.push_back(v->y);
            ++S.dens_off[(size_t)t + 1];
        }
    }

    S.rec = fa_receptor_desc{(int)S.rtype.size(), S.rxyz.data(), S.rrad.data(), S.rtype.data(), nf,
                             nf ? S.flex.data() : nullptr};
    fa_ligand_desc& d = S.lig;
    d.n_atoms = L;
    d.n_genes = FA->npar;
    d.n_grid = FA->num_grd;
    d.rec = S.lrec.data();
    d.dis = S.ldis.data();
    d.ang = S.lang.data();
    d.dih = S.ldih.data();
    d.shift = S.lshift.data();
    d.radius = S.lrad.data();
    d.type = S.ltype.data();
    d.mov = S.lmov.data();
    d.bonded = S.lbonded.data();
    d.gene_typ = S.gtyp.data();
    d.gene_atm = S.gatm.data();
    d.gene_min = S.gmin.data();
    d.gene_max = S.gmax.data();
    d.grid_ic = S.grid_ic.data();
    for (int j = 0; j < 3; ++j) {
        d.ori[j] = FA->ori[j];
        d.globalmin[j] = FA->globalmin[j];
        d.globalmax[j] = FA->globalmax[j];
    }
    d.maxwidth = FA->maxwidth;
    d.n_cons = (int)S.cons.size();
    d.cons = S.cons.empty() ? nullptr : S.cons.data();
    const bool dens = S.dens_off.back() > 0;
    S.prm = fa_params{FA->permeability, FA->dee_clash, FA->intramolecular, FA->solventterm, nt, S.emat.data(),
                      FA->normalize_area, dens ? S.dens_off.data() : nullptr, dens ? S.dens_xy.data() : nullptr};
}

// The device session behind the bridge: one system, uploaded at the first
// call and kept until reset() (the reference scores one system per run).
struct Session {
    FaSystem sys;
    bool uploaded = false;
};
inline Session& session()
{
    static Session s;
    return s;
}
inline void reset()
{
    hip_eval::shutdown();
    session().uploaded = false;
}

inline void ensure_uploaded(FA_Global* FA, const atom* atoms, const resid* residue, const gridpoint* cleftgrid,
                            std::span<const genlim> gene_lim, int max_pop)
{
    Session& s = session();
    if (s.uploaded) return;
    hip_eval::init(0, max_pop);
    system_from_reference(FA, atoms, residue, cleftgrid, gene_lim, s.sys);
    hip_eval::upload_system(s.sys.rec, s.sys.prm, &s.sys.lig, 1);
    s.uploaded = true;
}

// fa_cf_result → the reference's cfstr.  Clash and hull-failure poses carry
// ic2cf's clash record (penalty in wal, metal_coord 1, LIB/ic2cf.cpp:128-131);
// a pose the device did not score gets metal_coord +inf, so that both
// get_cf_evalue and get_apparent_cf_evalue rank it last.
inline cfstr to_cfstr(const fa_cf_result& r)
{
    cfstr c{};
    c.com = r.com;
    c.con = r.con;
    c.wal = r.wal;
    c.sas = r.sas;
    c.totsas = r.totsas;
    c.rclash = r.rclash;
    if (r.status == FA_CLASH || r.status == FA_HULL_FAIL)
        c.metal_coord = 1.0;
    else if (r.status != FA_OK)
        c.metal_coord = std::numeric_limits<double>::infinity();
    return c;
}

// voronoi_cf::batch_eval with its signature and semantics: rows with status
// 'n' are skipped (their cf stays value-initialised), genes are clamped to
// gene_lim, app_evalue is get_apparent_cf_evalue of each record, wall_ms the
// evaluation's wall clock.  The device scores ic2cf's contact function only;
// VC and the reference atoms stay untouched (the device keeps its own copy).
inline voronoi_cf::BatchResult batch_eval(std::span<const chromosome> chroms, FA_Global* FA, GB_Global* GB,
                                          VC_Global* VC, std::span<const genlim> gene_lim, const atom* ref_atoms,
                                          const resid* ref_residue, gridpoint* cleftgrid,
                                          cfstr (*function)(FA_Global*, VC_Global*, atom*, resid*, gridpoint*, int,
                                                            double*))
{
    (void)VC;
    if (function != &ic2cf) throw std::invalid_argument("flexaids_rocm::batch_eval scores ic2cf only");
    const int n = (int)chroms.size(), ng = GB->num_genes;
    if ((int)gene_lim.size() != ng) throw std::invalid_argument("flexaids_rocm::batch_eval: gene_lim size");
    ensure_uploaded(FA, ref_atoms, ref_residue, cleftgrid, gene_lim, n);
    std::vector<double> genes((size_t)n * ng);
    std::vector<char> status((size_t)n);
    std::vector<voronoi_cf_hip::GeneLimit> gl;
    for (int k = 0; k < ng; ++k) gl.push_back({gene_lim[(size_t)k].max, gene_lim[(size_t)k].min});
    for (int i = 0; i < n; ++i) {
        status[(size_t)i] = chroms[(size_t)i].status;
        for (int k = 0; k < ng; ++k) genes[(size_t)i * ng + k] = chroms[(size_t)i].genes[k].to_ic;
    }
    const voronoi_cf_hip::BatchResult r = voronoi_cf_hip::batch_eval(genes, ng, status, gl, 0);
    voronoi_cf::BatchResult out;
    out.cf.resize((size_t)n);
    out.app_evalue.resize((size_t)n);
    for (int i = 0; i < n; ++i) {
        if (status[(size_t)i] == 'n') continue;
        out.cf[(size_t)i] = to_cfstr(r.cf[(size_t)i]);
        out.app_evalue[(size_t)i] = get_apparent_cf_evalue(&out.cf[(size_t)i]);
    }
    out.wall_ms = r.wall_ms;
    return out;
}

inline hw::DispatchTelemetry telemetry(std::chrono::steady_clock::time_point t0, std::size_t n)
{
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return hw::DispatchTelemetry{hw::Backend::ROCM, ms, (int64_t)n, ms > 0.0 ? (double)n / (ms * 1e3) : 0.0};
}

// flexaids::compute_boltzmann_batch (LIB/UnifiedHardwareDispatch.cpp:768-794):
// weights exp(-beta (E - E_min)), log_Z = ln sum w - beta E_min, E_min.
inline hw::BoltzmannBatchResult compute_boltzmann_batch(std::span<const double> energies, double beta)
{
    if (energies.empty()) return {{}, 0.0, 0.0, {hw::Backend::SCALAR, 0.0, 0, 0.0}};
    const auto t0 = std::chrono::steady_clock::now();
    hip_eval::init(0);  // idempotent on device 0
    rocm_dispatch::BoltzmannBatch b = rocm_dispatch::boltzmann_batch(energies, beta);
    return {std::move(b.weights), b.log_Z, b.E_min, telemetry(t0, energies.size())};
}

// flexaids::log_sum_exp_dispatch (LIB/UnifiedHardwareDispatch.cpp:800-843):
// -inf for no values, the maximum itself when it is not finite.
inline double log_sum_exp_dispatch(std::span<const double> values)
{
    if (values.empty()) return -std::numeric_limits<double>::infinity();
    const double x_max = *std::max_element(values.begin(), values.end());
    if (!std::isfinite(x_max)) return x_max;
    hip_eval::init(0);
    return rocm_dispatch::log_sum_exp(values);
}

// flexaids::select_backend: ROCM when a gfx9 device is present, else the
// reference's own choice (`otherwise`, flexaids::select_backend by default).
inline flexaids::HardwareBackend select_backend(flexaids::HardwareBackend (*otherwise)() = &flexaids::select_backend)
{
    return rocm_dispatch::is_available() ? flexaids::HardwareBackend::ROCM : otherwise();
}

}  // namespace flexaids_rocm

// The ROCm case in the reference's dispatch (LIB/UnifiedHardwareDispatch.cpp:
// 1051-1073) then reads:
//
//   BoltzmannBatchResult compute_boltzmann_batch(std::span<const double> energies, double beta) {
//       if (rocm_dispatch::is_available()) return flexaids_rocm::compute_boltzmann_batch(energies, beta);
//       return hw::UnifiedHardwareDispatch::instance().compute_boltzmann_batch(energies, beta);
//   }
//   double log_sum_exp_dispatch(std::span<const double> values) {
//       if (rocm_dispatch::is_available()) return flexaids_rocm::log_sum_exp_dispatch(values);
//       return hw::UnifiedHardwareDispatch::instance().log_sum_exp_dispatch(values);
//   }
//
// and a caller of voronoi_cf::batch_eval switches by backend:
//
//   auto res = flexaids::select_backend() == flexaids::HardwareBackend::ROCM
//                  ? flexaids_rocm::batch_eval(chroms, FA, GB, VC, gene_lim, atoms, residue, cleftgrid, ic2cf)
//                  : voronoi_cf::batch_eval(chroms, FA, GB, VC, gene_lim, atoms, residue, cleftgrid, ic2cf);
