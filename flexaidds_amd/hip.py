"""ctypes binding of the HIP library's C ABI (include/flexaidds_amd.h).

The library is required: importing this module raises when
`libflexaidds_amd.so` is missing, and every call raises on a non-zero status.
There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import abi

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libflexaidds_amd.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is not built: run `python -m flexaidds_amd.build`")

# One HIP runtime per process: when PyTorch is present its bundled runtime is
# loaded first so that torch tensors and this library share it.
try:
    import torch  # noqa: F401
except ImportError:  # pragma: no cover
    torch = None

_lib = C.CDLL(LIB_PATH)
_P = C.c_void_p


class OpCount(C.Structure):
    _fields_ = [(f, C.c_uint64) for f in ("cull", "plane", "hull_test", "arc_solve", "face_order", "sph_triangle",
                                          "sph_arc", "vc_contact", "build_atom")]


#: FLOP per counted unit (the algorithmic FLOP model of the contact function)
FLOP_WEIGHTS = {"cull": 11, "plane": 14, "hull_test": 13, "arc_solve": 82, "face_order": 18,
                "sph_triangle": 85, "sph_arc": 60, "vc_contact": 25, "build_atom": 94}


def flops(ops: dict) -> int:
    return int(sum(FLOP_WEIGHTS[k] * int(v) for k, v in ops.items()))


class DeviceInfo(C.Structure):
    _fields_ = [("name", C.c_char * 256), ("compute_units", C.c_int32), ("global_mem", C.c_uint64),
                ("gcn_arch", C.c_int32), ("wavefront_size", C.c_int32), ("max_threads_per_cu", C.c_int32),
                ("shared_mem_per_cu", C.c_int32)]


_SIGS = {
    "fa_get_device_info": ([C.c_int, _P], C.c_int),
    "fa_last_error": ([], C.c_char_p),
    "fa_ctx_create": ([C.c_int, C.c_int, _P], C.c_int),
    "fa_ctx_destroy": ([_P], C.c_int),
    "fa_upload_system": ([_P, _P, _P, C.c_int, _P], C.c_int),
    "fa_eval_batch": ([_P, C.c_int, C.c_int, _P, _P], C.c_int),
    "fa_eval_batch_device": ([_P, C.c_int, _P, C.c_int, _P, _P, _P], C.c_int),
    "fa_eval_batch_device_noalloc": ([_P, C.c_int, _P, C.c_int, _P, _P, _P], C.c_int),
    "fa_noalloc_scratch_bytes": ([], C.c_size_t),
    "fa_eval_batch_debug": ([_P, C.c_int, C.c_int, _P, _P, _P, _P, C.c_int, _P], C.c_int),
    "fa_thermo_device": ([_P, _P, C.c_int, _P, C.c_int64, C.c_double, _P, _P], C.c_int),
    "fa_thermo_host": ([_P, _P, _P, C.c_int64, C.c_double, _P], C.c_int),
    "fa_last_counters": ([_P, _P, _P], C.c_int),
    "fa_last_opcount": ([_P, _P], C.c_int),
    "fa_ctx_totals": ([_P, _P], C.c_int),
    "fa_ctx_totals_n": ([_P, _P, C.c_int], C.c_int),
    "fa_ctx_reset_totals": ([_P], C.c_int),
    "fa_upload_emat": ([_P, _P, C.c_int], C.c_int),
    "fa_set_phase_profiling": ([_P, C.c_int], C.c_int),
    "fa_last_phase_cycles": ([_P, _P], C.c_int),
    "fa_last_pose_cycles": ([_P, C.c_int, _P], C.c_int),
    "fa_boltzmann_device": ([_P, _P, C.c_int, C.c_int64, C.c_double, _P, _P, _P], C.c_int),
    "fa_log_sum_exp_device": ([_P, _P, C.c_int, C.c_int64, _P, _P], C.c_int),
    "fa_kernel_lds_bytes": ([], C.c_int),
    "fa_ctx_launch_waves": ([_P], C.c_int),
    "fa_ctx_set_launch_shape": ([_P, C.c_int, C.c_int], C.c_int),
    "fa_ctx_get_launch_shape": ([_P, _P, _P], C.c_int),
    "fa_ga_create": ([_P, _P, _P, _P], C.c_int),
    "fa_ga_create_batch": ([_P, _P, C.c_int, _P, _P, _P], C.c_int),
    "fa_ga_shape": ([_P, _P, _P], C.c_int),
    "fa_ga_fitness": ([_P, _P, _P], C.c_int),
    "fa_ga_dedup_stats": ([_P, _P], C.c_int),
    "fa_ga_snapshot_enable": ([_P, C.c_int], C.c_int),
    "fa_ga_snapshot": ([_P, C.c_int, _P, _P, _P, _P, _P], C.c_int),
    "fa_ga_grand": ([_P, _P, _P, _P, _P], C.c_int),
    "fa_grand_device": ([_P, _P, C.c_int, _P, C.c_int, _P, _P], C.c_int),
    "fa_serve_run": ([_P, C.c_int, _P, C.c_int, C.c_int, _P, _P], C.c_int),
    "fa_cluster_cf": ([_P, C.c_int, C.c_int, _P, C.c_int, _P, _P, _P, _P, _P, _P, _P], C.c_int),
    "fa_super_cluster": ([_P, _P, C.c_int, C.c_int, _P, _P], C.c_int),
    "fa_ga_run": ([_P, C.c_int, C.c_int, _P], C.c_int),
    "fa_pool_create": ([_P, _P, _P], C.c_int),
    "fa_pool_shape": ([_P, _P, _P, _P], C.c_int),
    "fa_pool_submit": ([_P, _P, _P], C.c_int),
    "fa_pool_step": ([_P, _P, _P], C.c_int),
    "fa_pool_collect": ([_P, _P, C.c_int, _P], C.c_int),
    "fa_pool_population": ([_P, C.c_int64, _P, _P, _P], C.c_int),
    "fa_pool_set_max_running": ([_P, C.c_int], C.c_int),
    "fa_pool_busy": ([_P, _P], C.c_int),
    "fa_pool_destroy": ([_P], C.c_int),
    "fa_ga_population": ([_P, _P, _P, _P], C.c_int),
    "fa_ga_thermo": ([_P, _P, _P], C.c_int),
    "fa_ga_generation": ([_P, _P], C.c_int),
    "fa_ga_destroy": ([_P], C.c_int),
}

EXPORTED = tuple(_SIGS)

for _name, (_args, _res) in _SIGS.items():
    _f = getattr(_lib, _name)
    _f.argtypes = _args
    _f.restype = _res


class FAError(RuntimeError):
    pass


def _check(rc):
    if rc != 0:
        raise FAError(f"status {rc}: {_lib.fa_last_error().decode()}")


def device_info(device=0) -> dict:
    d = DeviceInfo()
    _check(_lib.fa_get_device_info(device, C.byref(d)))
    return {"name": d.name.decode(), "compute_units": d.compute_units, "global_mem": d.global_mem,
            "gcn_arch": d.gcn_arch, "wavefront_size": d.wavefront_size,
            "max_threads_per_cu": d.max_threads_per_cu, "shared_mem_per_cu": d.shared_mem_per_cu}


def kernel_lds_bytes() -> int:
    return int(_lib.fa_kernel_lds_bytes())


def _addr(s):
    return C.cast(C.pointer(s), C.c_void_p)


class Context:
    """One GPU's resident docking system (receptor, ligands, parameters)."""

    def __init__(self, device: int = 0, max_pop: int = 0):
        self._ctx = C.c_void_p()
        _check(_lib.fa_ctx_create(device, max_pop, C.byref(self._ctx)))
        self.device = int(device)
        self._desc = None
        self.n_genes = []

    def close(self):
        if self._ctx:
            _lib.fa_ctx_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def upload(self, receptor, ligands, params):
        desc = abi.Descriptors(receptor, ligands, params)
        arr = (abi.LigandDesc * len(desc.ligands))(*desc.ligands)
        _check(_lib.fa_upload_system(self._ctx, _addr(desc.receptor), _addr(desc.params), len(desc.ligands),
                                     C.cast(arr, C.c_void_p)))
        self._desc = desc
        self.n_genes = [l.n_genes for l in ligands]
        self.n_atoms = [l.n_atoms for l in ligands]

    def upload_system(self, system):
        self.upload(system.receptor, system.ligands, system.params)

    def eval(self, ligand: int, genes: np.ndarray) -> np.ndarray:
        genes = np.ascontiguousarray(genes, np.float64)
        assert genes.ndim == 2 and genes.shape[1] == self.n_genes[ligand]
        out = np.zeros(genes.shape[0], abi.CF_DTYPE)
        _check(_lib.fa_eval_batch(self._ctx, ligand, genes.shape[0], genes.ctypes.data, out.ctypes.data))
        return out

    def eval_debug(self, ligand: int, genes: np.ndarray, rec_cap: int = 2048):
        genes = np.ascontiguousarray(genes, np.float64)
        P = genes.shape[0]
        L = self.n_atoms[ligand]
        out = np.zeros(P, abi.CF_DTYPE)
        xyz = np.zeros((P, L, 3), np.float32)
        recs = np.zeros((P, rec_cap), abi.REC_DTYPE)
        cnt = np.zeros(P, np.int32)
        _check(_lib.fa_eval_batch_debug(self._ctx, ligand, P, genes.ctypes.data, out.ctypes.data, xyz.ctypes.data,
                                        recs.ctypes.data, rec_cap, cnt.ctypes.data))
        return out, xyz, [recs[p, :min(cnt[p], rec_cap)] for p in range(P)]

    def eval_device(self, n_poses: int, genes_ptr: int, gene_stride: int, pose_lig_ptr: int, out_ptr: int,
                    stream: int = 0):
        """Enqueue an evaluation over device buffers (raw device pointers)."""
        _check(_lib.fa_eval_batch_device(self._ctx, n_poses, C.c_void_p(genes_ptr), gene_stride,
                                         C.c_void_p(pose_lig_ptr) if pose_lig_ptr else None,
                                         C.c_void_p(out_ptr), C.c_void_p(stream) if stream else None))

    def eval_device_noalloc(self, n_poses: int, genes_ptr: int, gene_stride: int, pose_lig_ptr: int, out_ptr: int,
                            stream: int = 0):
        """eval_device with the library's static scratch (allocates nothing)."""
        _check(_lib.fa_eval_batch_device_noalloc(self._ctx, n_poses, C.c_void_p(genes_ptr), gene_stride,
                                                 C.c_void_p(pose_lig_ptr) if pose_lig_ptr else None,
                                                 C.c_void_p(out_ptr), C.c_void_p(stream) if stream else None))

    def thermo(self, energy: np.ndarray, temperature: float = 300.0, count=None) -> dict:
        e = np.ascontiguousarray(energy, np.float64)
        cnt = None if count is None else np.ascontiguousarray(count, np.int32)
        th = abi.Thermo()
        _check(_lib.fa_thermo_host(self._ctx, e.ctypes.data, None if cnt is None else cnt.ctypes.data,
                                   e.shape[0], temperature, C.byref(th)))
        return {f: getattr(th, f) for f, _ in abi.Thermo._fields_}

    def thermo_device(self, energy_ptr: int, stride: int, n: int, temperature: float = 300.0,
                      count_ptr: int = 0, stream: int = 0) -> dict:
        th = abi.Thermo()
        _check(_lib.fa_thermo_device(self._ctx, C.c_void_p(energy_ptr), stride,
                                     C.c_void_p(count_ptr) if count_ptr else None, n, temperature, C.byref(th),
                                     C.c_void_p(stream) if stream else None))
        return {f: getattr(th, f) for f, _ in abi.Thermo._fields_}

    def boltzmann_device(self, energy_ptr: int, stride: int, n: int, temperature: float, weights_ptr: int,
                         want_log_z: bool = True, stream: int = 0):
        """compute_boltzmann_batch on the device: weights[i] = exp(-beta (E_i - E_min))
        into `weights_ptr` (device); returns ln Z when asked (then synchronous)."""
        lz = C.c_double()
        _check(_lib.fa_boltzmann_device(self._ctx, C.c_void_p(energy_ptr), stride, n, temperature,
                                        C.c_void_p(weights_ptr), C.byref(lz) if want_log_z else None,
                                        C.c_void_p(stream) if stream else None))
        return lz.value if want_log_z else None

    def log_sum_exp_device(self, values_ptr: int, stride: int, n: int, stream: int = 0) -> float:
        """log_sum_exp_dispatch over device values (synchronous)."""
        out = C.c_double()
        _check(_lib.fa_log_sum_exp_device(self._ctx, C.c_void_p(values_ptr) if values_ptr else None, stride, n,
                                          C.byref(out), C.c_void_p(stream) if stream else None))
        return out.value

    def grand_device(self, log_z_ptr: int, stride: int, log_c_ptr: int, n: int, out_ptr: int, stream: int = 0):
        """Grand partition function on the device: out[0] = ln Xi, out[1] = p_empty,
        out[2 + i] = p_bound of ligand i (device pointers, asynchronous)."""
        _check(_lib.fa_grand_device(self._ctx, C.c_void_p(log_z_ptr), stride, C.c_void_p(log_c_ptr), n,
                                    C.c_void_p(out_ptr), C.c_void_p(stream) if stream else None))

    def cluster_cf(self, ligand: int, genes, energy, temperature: float = 0.0, rmsd: float = 2.0,
                   max_results: int = 10) -> dict:
        """Binding modes of poses of `ligand` (fa_cluster_cf): genes [n, >= genes of
        the ligand] internal coordinates in the order to cluster (by increasing
        energy, as the reference passes them), energy [n] (app energies)."""
        g = np.ascontiguousarray(genes, np.float64)
        e = np.ascontiguousarray(energy, np.float64)
        n = int(e.shape[0])
        if g.ndim != 2 or g.shape[0] != n:
            raise ValueError("genes must be [n, genes] with one row per energy")
        p = ClusterParams(float(temperature), float(rmsd), int(max_results))
        modes = (ClusterMode * max(1, int(max_results)))()
        member = np.zeros(max(n, 1), np.int32)
        rh = np.zeros(max(n, 1), np.float32)
        trm = np.zeros(max(1, int(max_results)) ** 2, np.float32)
        out = ClusterOut()
        _check(_lib.fa_cluster_cf(self._ctx, ligand, n, g.ctypes.data, g.shape[1], e.ctypes.data, C.byref(p), modes,
                                  member.ctypes.data, rh.ctypes.data, trm.ctypes.data, C.byref(out)))
        k = int(out.n_modes)
        return {"modes": [{"top": int(m.top), "freq": int(m.freq), "tcf": float(m.tcf), "acf": float(m.acf)}
                          for m in modes[:k]],
                "member": member[:n], "rmsd_head": rh[:n], "top_rmsd": trm[:k * k].reshape(k, k),
                "n_clustered": int(out.n_clustered), "device_ms": float(out.device_ms)}

    def super_cluster(self, energy, min_pts: int) -> dict:
        """FastOPTICS super-cluster of the energies as one-dimensional points
        (fa_super_cluster): {"members": point indices in the flood fill's
        visiting order, "seed", "device_ms"}."""
        e = np.ascontiguousarray(energy, np.float64)
        n = int(e.shape[0])
        members = np.zeros(max(n, 1), np.int32)
        out = SuperClusterOut()
        _check(_lib.fa_super_cluster(self._ctx, e.ctypes.data, n, int(min_pts), members.ctypes.data, C.byref(out)))
        return {"members": members[:int(out.n_members)], "seed": int(out.seed), "device_ms": float(out.device_ms)}

    def opcount(self) -> dict:
        o = OpCount()
        _check(_lib.fa_last_opcount(self._ctx, C.byref(o)))
        return {f: getattr(o, f) for f, _ in OpCount._fields_}

    PHASES = ("build", "prepass", "cull", "first_vertex", "hull_walk", "arc_points", "seeds", "faces", "areas",
              "records")

    def set_phase_profiling(self, on: bool):
        _check(_lib.fa_set_phase_profiling(self._ctx, 1 if on else 0))

    def phase_cycles(self) -> dict:
        a = np.zeros(16, np.uint64)
        _check(_lib.fa_last_phase_cycles(self._ctx, a.ctypes.data))
        return {k: int(a[i]) for i, k in enumerate(self.PHASES)}

    def pose_cycles(self, n: int):
        """(cycles per pose, widened-grid flag per pose) of the last profiled launch."""
        raw = np.zeros(n, np.uint64)
        _check(_lib.fa_last_pose_cycles(self._ctx, n, raw.ctypes.data))
        return (raw & np.uint64((1 << 63) - 1)).astype(np.int64), (raw >> np.uint64(63)).astype(bool)

    def totals(self) -> dict:
        """Unscored and widened-grid poses over every launch since the last reset."""
        a = np.zeros(5, np.uint64)
        _check(_lib.fa_ctx_totals_n(self._ctx, a.ctypes.data, 5))
        return {"slow_grid": int(a[0]), "capacity": int(a[1]), "grid_range": int(a[2]), "bad_ligand": int(a[3]),
                "undefined_face": int(a[4])}

    def reset_totals(self):
        _check(_lib.fa_ctx_reset_totals(self._ctx))

    def set_launch_shape(self, waves: int = 0, flags: int = 0):
        """Cap on resident waves (0: every slot) and scan variant (abi.FA_SHAPE_*)."""
        _check(_lib.fa_ctx_set_launch_shape(self._ctx, int(waves), int(flags)))

    def launch_shape(self) -> tuple:
        """(waves a launch uses, flags)."""
        w, f = C.c_int(), C.c_int()
        _check(_lib.fa_ctx_get_launch_shape(self._ctx, C.byref(w), C.byref(f)))
        return w.value, f.value

    def launch_waves(self) -> int:
        """Waves one pose launch runs (resident slots on the device)."""
        return int(_lib.fa_ctx_launch_waves(self._ctx))

    def counters(self):
        a, b = C.c_uint32(), C.c_uint32()
        _check(_lib.fa_last_counters(self._ctx, C.byref(a), C.byref(b)))
        return {"slow_grid": a.value, "capacity": b.value}


class GeneLim(C.Structure):
    _fields_ = [("min", C.c_double), ("max", C.c_double), ("del_", C.c_double), ("map", C.c_int32),
                ("pad", C.c_int32)]


class ClusterParams(C.Structure):
    _fields_ = [("temperature", C.c_double), ("rmsd", C.c_float), ("max_results", C.c_int32)]


class ClusterMode(C.Structure):
    _fields_ = [("top", C.c_int32), ("freq", C.c_int32), ("tcf", C.c_double), ("acf", C.c_double)]


class ClusterOut(C.Structure):
    _fields_ = [("n_modes", C.c_int32), ("n_clustered", C.c_int32), ("device_ms", C.c_float), ("pad", C.c_int32)]


SUPER_CLUSTER_MAX = 8192   # FA_SUPER_CLUSTER_MAX


class SuperClusterOut(C.Structure):
    _fields_ = [("n_members", C.c_int32), ("seed", C.c_int32), ("device_ms", C.c_float), ("pad", C.c_int32)]


class GAParams(C.Structure):
    _fields_ = [("ligand", C.c_int32), ("population", C.c_int32), ("offspring", C.c_int32),
                ("intragenes", C.c_int32), ("cross_rate", C.c_double), ("mut_rate", C.c_double),
                ("temperature", C.c_double), ("seed", C.c_uint64), ("fitness", C.c_int32), ("dedup_rounds", C.c_int32),
                ("share_alpha", C.c_double), ("share_peaks", C.c_double), ("share_scale", C.c_double),
                ("sig_share", C.c_double), ("entropy_weight", C.c_double), ("reproduction", C.c_int32),
                ("pad", C.c_int32)]


FITNESS = {"LINEAR": 0, "PSHARE": 1, "SMFREE": 2}
REPRODUCTION = {"BOOM": 0, "STEADY": 1}


class GA:
    """Genetic algorithm running on the device (fa_ga_* of the C ABI).

    `limits` is a sequence of (min, max, del, map) per gene of the ligand.
    With `ligands` (a sequence of ligand indices) the GA is batched: one job
    per ligand, `limits` then a sequence of per-job limit sequences."""

    def __init__(self, ctx: Context, limits, ligand=0, population=2000, offspring=None, cross_rate=0.8,
                 mut_rate=0.03, temperature=300.0, seed=1, intragenes=False, ligands=None, fitness="LINEAR",
                 share_alpha=1.0, share_peaks=5.0, share_scale=10.0, sig_share=0.0, entropy_weight=0.5,
                 dedup_rounds=0, reproduction="BOOM"):
        self.ctx = ctx
        per_job = [limits] if ligands is None else list(limits)
        flat = [x for job in per_job for x in job]
        self._lim = (GeneLim * len(flat))(*[GeneLim(float(a), float(b), float(d), int(m), 0)
                                             for a, b, d, m in flat])
        model = FITNESS[fitness] if isinstance(fitness, str) else int(fitness)
        self.params = GAParams(ligand, population, population if offspring is None else offspring,
                               1 if intragenes else 0, cross_rate, mut_rate, temperature, seed, model,
                               int(dedup_rounds), share_alpha, share_peaks, share_scale, sig_share, entropy_weight,
                               REPRODUCTION[reproduction] if isinstance(reproduction, str) else int(reproduction), 0)
        self.job_genes = [len(j) for j in per_job]
        self._ga = C.c_void_p()
        if ligands is None:
            _check(_lib.fa_ga_create(ctx._ctx, C.byref(self.params), self._lim, C.byref(self._ga)))
        else:
            if len(ligands) != len(per_job):
                raise ValueError("one limit sequence per ligand")
            self._ligs = (C.c_int * len(ligands))(*ligands)
            _check(_lib.fa_ga_create_batch(ctx._ctx, C.byref(self.params), len(ligands), self._ligs, self._lim,
                                           C.byref(self._ga)))
        j, g = C.c_int(), C.c_int()
        _check(_lib.fa_ga_shape(self._ga, C.byref(j), C.byref(g)))
        self.jobs, self.genes = int(j.value), int(g.value)

    def run(self, generations: int, graph: bool = True) -> float:
        ms = C.c_float()
        _check(_lib.fa_ga_run(self._ga, generations, 1 if graph else 0, C.byref(ms)))
        return float(ms.value)

    def population(self):
        """(genes int32 [rows, Gmax], ic [rows, Gmax], results [rows]); rows are the
        survivors of each job (population each), then the children of each job."""
        rows = self.jobs * (self.params.population + self.params.offspring)
        g = np.zeros((rows, self.genes), np.int32)
        ic = np.zeros((rows, self.genes), np.float64)
        res = np.zeros(rows, abi.CF_DTYPE)
        _check(_lib.fa_ga_population(self._ga, g.ctypes.data, ic.ctypes.data, res.ctypes.data))
        return g, ic, res

    def fitness(self):
        """(fitness of the parent rows [jobs * population], niche radius per job)."""
        f = np.zeros(self.jobs * self.params.population)
        sig = np.zeros(self.jobs)
        _check(_lib.fa_ga_fitness(self._ga, f.ctypes.data, sig.ctypes.data))
        return f, sig

    def job_rows(self, k: int):
        """(survivor rows, child rows) of job k in population()'s arrays."""
        pop, new = self.params.population, self.params.offspring
        return (slice(k * pop, (k + 1) * pop),
                slice(self.jobs * pop + k * new, self.jobs * pop + (k + 1) * new))

    def thermo(self):
        """Job 0's ensemble (dict) and the poses evaluated over all jobs."""
        t, n = self.thermo_jobs()
        return t[0], n

    def thermo_jobs(self):
        """Each job's ensemble thermodynamics (list of dicts), poses evaluated."""
        t = (abi.Thermo * self.jobs)()
        n = C.c_uint64()
        _check(_lib.fa_ga_thermo(self._ga, t, C.byref(n)))
        return [{f: getattr(x, f) for f, _ in abi.Thermo._fields_} for x in t], int(n.value)

    def grand(self, conc_M):
        """Grand partition function over the jobs at concentrations conc_M (mol/L):
        {"log_xi", "p_bound" [jobs], "p_empty"}."""
        c = np.ascontiguousarray(conc_M, np.float64)
        if c.shape != (self.jobs,):
            raise ValueError("one concentration per job")
        lx, pe = C.c_double(), C.c_double()
        pb = np.zeros(self.jobs)
        _check(_lib.fa_ga_grand(self._ga, c.ctypes.data, C.byref(lx), pb.ctypes.data, C.byref(pe)))
        return {"log_xi": float(lx.value), "p_bound": pb, "p_empty": float(pe.value)}

    def generation(self) -> int:
        g = C.c_uint64()
        _check(_lib.fa_ga_generation(self._ga, C.byref(g)))
        return int(g.value)

    def snapshot_enable(self, max_generations: int):
        """Record each job's population after every later generation, for up
        to max_generations generations (fa_ga_snapshot_enable)."""
        _check(_lib.fa_ga_snapshot_enable(self._ga, int(max_generations)))

    def snapshot(self, job: int = 0) -> dict:
        """Job `job`'s records sorted by energy without duplicates
        (fa_ga_snapshot): {"ic" [kept, genes], "evalue", "app", "n_raw"}."""
        nr, nk = C.c_int32(), C.c_int32()
        _check(_lib.fa_ga_snapshot(self._ga, job, None, None, None, C.byref(nr), C.byref(nk)))
        n, G = int(nr.value), self.job_genes[job]
        ic = np.zeros((max(n, 1), G))
        e = np.zeros(max(n, 1))
        a = np.zeros(max(n, 1))
        _check(_lib.fa_ga_snapshot(self._ga, job, ic.ctypes.data, e.ctypes.data, a.ctypes.data, C.byref(nr),
                                   C.byref(nk)))
        k = int(nk.value)
        return {"ic": ic[:k], "evalue": e[:k], "app": a[:k], "n_raw": n}

    def snapshot_thermo(self, job: int = 0, super_cluster: bool = False) -> dict:
        """The end-of-run ensemble: every kept record's energy (cf) once
        (LIB/gaboom.cpp:535-540), on the device.  With super_cluster (the
        reference's use_super_cluster, :542-560) and more than 4 records,
        only the FastOPTICS super-cluster of the energies (minPts
        max(4, n / 20), fa_super_cluster: at most 8,192 records) when it
        leaves some record out."""
        e = self.snapshot(job)["evalue"]
        T = self.params.temperature
        if super_cluster and len(e) > 4:
            members = self.ctx.super_cluster(e, max(4, len(e) // 20))["members"]
            if 0 < len(members) < len(e):
                return self.ctx.thermo(e[members], T)
        return self.ctx.thermo(e, T)

    def dedup_stats(self) -> np.ndarray:
        """[jobs, 2]: duplicate candidates drawn, rows that kept a duplicate
        at the batch cap (fa_ga_dedup_stats)."""
        out = np.zeros((self.jobs, 2), np.uint64)
        _check(_lib.fa_ga_dedup_stats(self._ga, out.ctypes.data))
        return out

    def close(self):
        if self._ga:
            _lib.fa_ga_destroy(self._ga)
            self._ga = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


MAX_GENES = 64   # FA_MAX_GENES


class PoolParams(C.Structure):
    _fields_ = [("population", C.c_int32), ("offspring", C.c_int32), ("intragenes", C.c_int32),
                ("fitness", C.c_int32), ("cross_rate", C.c_double), ("mut_rate", C.c_double),
                ("temperature", C.c_double), ("share_alpha", C.c_double), ("share_peaks", C.c_double),
                ("share_scale", C.c_double), ("sig_share", C.c_double), ("entropy_weight", C.c_double),
                ("max_slots", C.c_int32), ("max_genes", C.c_int32), ("hbm_budget", C.c_uint64),
                ("max_running", C.c_int32), ("reproduction", C.c_int32), ("dedup_rounds", C.c_int32),
                ("dedup_max_generations", C.c_int32)]


class Job(C.Structure):
    _fields_ = [("ligand", C.c_int32), ("generations", C.c_int32), ("seed", C.c_uint64),
                ("limits", C.POINTER(GeneLim))]


class JobResult(C.Structure):
    _fields_ = [("id", C.c_int64), ("ligand", C.c_int32), ("generations", C.c_int32), ("n_genes", C.c_int32),
                ("pad", C.c_int32), ("best", abi.CFResult), ("best_genes", C.c_int32 * MAX_GENES),
                ("best_ic", C.c_double * MAX_GENES), ("thermo", abi.Thermo), ("n_poses", C.c_uint64)]


def _job_dict(r) -> dict:
    return {"id": int(r.id), "ligand": int(r.ligand), "generations": int(r.generations),
            "best": {f: getattr(r.best, f) for f, _ in abi.CFResult._fields_},
            "best_genes": np.array(r.best_genes[:r.n_genes], np.int32),
            "best_ic": np.array(r.best_ic[:r.n_genes]),
            "thermo": {f: getattr(r.thermo, f) for f, _ in abi.Thermo._fields_},
            "n_poses": int(r.n_poses)}


class Request(C.Structure):
    _fields_ = [("t", C.c_double), ("pool", C.c_int32), ("pad", C.c_int32), ("job", Job)]


class ServedRec(C.Structure):
    _fields_ = [("request", C.c_int64), ("latency_s", C.c_double), ("result", JobResult)]


class ServeStats(C.Structure):
    _fields_ = [("wall_s", C.c_double), ("steps", C.c_int64), ("poses", C.c_uint64)]


def serve_run(pools, requests, realtime=True):
    """The served request loop in C++ (fa_serve_run).  requests: dicts with
    t, pool, ligand, generations, seed and limits (a sequence of
    (min, max, del, map)).  Returns ([(request index, latency_s, result
    dict)] in completion order, {"wall_s", "steps", "poses"})."""
    n = len(requests)
    keep = []
    arr = (Request * max(n, 1))()
    for i, r in enumerate(requests):
        lim = (GeneLim * len(r["limits"]))(*[GeneLim(float(a), float(b), float(d), int(m), 0)
                                              for a, b, d, m in r["limits"]])
        keep.append(lim)
        arr[i] = Request(float(r["t"]), int(r["pool"]), 0,
                         Job(int(r["ligand"]), int(r["generations"]), int(r["seed"]), lim))
    handles = (C.c_void_p * len(pools))(*[p._pool for p in pools])
    out = (ServedRec * max(n, 1))()
    st = ServeStats()
    _check(_lib.fa_serve_run(handles, len(pools), arr, n, 1 if realtime else 0, out, C.byref(st)))
    res = [(int(o.request), float(o.latency_s), _job_dict(o.result)) for o in out[:n]]
    return res, {"wall_s": float(st.wall_s), "steps": int(st.steps), "poses": int(st.poses)}


class Pool:
    """Continuous-batching docking pool on one context (fa_pool_* of the C ABI).

    Jobs are submitted as they arrive; every step() admits queued jobs into
    free slots and advances all resident jobs by one generation in one pose
    launch; collect() returns the jobs that step finished."""

    def __init__(self, ctx: Context, population=1024, offspring=None, max_slots=64, hbm_budget=8 << 30,
                 cross_rate=0.8, mut_rate=0.03, temperature=300.0, intragenes=False, fitness="LINEAR",
                 share_alpha=1.0, share_peaks=5.0, share_scale=10.0, sig_share=0.0, entropy_weight=0.5,
                 max_genes=0, max_running=0, reproduction="BOOM", dedup_rounds=0, dedup_max_generations=0):
        self.ctx = ctx
        model = FITNESS[fitness] if isinstance(fitness, str) else int(fitness)
        self.params = PoolParams(population, population if offspring is None else offspring,
                                 1 if intragenes else 0, model, cross_rate, mut_rate, temperature, share_alpha,
                                 share_peaks, share_scale, sig_share, entropy_weight, max_slots, max_genes,
                                 int(hbm_budget), max_running,
                                 REPRODUCTION[reproduction] if isinstance(reproduction, str) else int(reproduction),
                                 int(dedup_rounds), int(dedup_max_generations))
        self._pool = C.c_void_p()
        _check(_lib.fa_pool_create(ctx._ctx, C.byref(self.params), C.byref(self._pool)))
        s, b, g = C.c_int(), C.c_uint64(), C.c_int()
        _check(_lib.fa_pool_shape(self._pool, C.byref(s), C.byref(b), C.byref(g)))
        self.slots, self.slot_bytes, self.genes = int(s.value), int(b.value), int(g.value)
        self._out = (JobResult * self.slots)()

    def submit(self, ligand: int, limits, generations: int, seed: int) -> int:
        lim = (GeneLim * len(limits))(*[GeneLim(float(a), float(b), float(d), int(m), 0) for a, b, d, m in limits])
        job = Job(ligand, generations, seed, lim)
        jid = C.c_int64()
        _check(_lib.fa_pool_submit(self._pool, C.byref(job), C.byref(jid)))
        return int(jid.value)

    def step(self):
        """(jobs still resident after this step, jobs queued)."""
        r, q = C.c_int(), C.c_int()
        _check(_lib.fa_pool_step(self._pool, C.byref(r), C.byref(q)))
        return int(r.value), int(q.value)

    def collect(self):
        n = C.c_int()
        _check(_lib.fa_pool_collect(self._pool, self._out, self.slots, C.byref(n)))
        return [_job_dict(r) for r in self._out[:n.value]]

    def set_max_running(self, n: int):
        _check(_lib.fa_pool_set_max_running(self._pool, int(n)))

    def busy(self) -> bool:
        """True while the last step still runs on the device (never waits)."""
        b = C.c_int()
        _check(_lib.fa_pool_busy(self._pool, C.byref(b)))
        return bool(b.value)

    def population(self, job_id: int):
        pop = self.params.population
        g = np.zeros((pop, self.genes), np.int32)
        ic = np.zeros((pop, self.genes))
        res = np.zeros(pop, abi.CF_DTYPE)
        _check(_lib.fa_pool_population(self._pool, job_id, g.ctypes.data, ic.ctypes.data, res.ctypes.data))
        return g, ic, res

    def close(self):
        if self._pool:
            _lib.fa_pool_destroy(self._pool)
            self._pool = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
