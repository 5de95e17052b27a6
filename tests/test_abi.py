"""The C-ABI library loads and exports every entry point include/*.h declares
(no GPU needed: only symbol resolution, no compute calls)."""
import ctypes as C
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in ("flexaidds_amd.h",):
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(fa_\w+)\s*\(", src, re.M):
            names.add(m.group(1))
    return names


def test_header_declares_entry_points():
    names = declared_functions()
    for n in ("fa_ctx_create", "fa_ctx_destroy", "fa_upload_system", "fa_eval_batch", "fa_eval_batch_device",
              "fa_thermo_device", "fa_get_device_info"):
        assert n in names


def test_library_exports_all_declared():
    from flexaidds_amd import hip
    lib = C.CDLL(hip.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(hip.EXPORTED) <= declared_functions()


def test_hip_eval_cxx_symbols():
    """The C++ drop-in (namespace hip_eval) is exported with C++ linkage."""
    import subprocess
    from flexaidds_amd import hip
    out = subprocess.run(["nm", "-D", "--defined-only", "-C", hip.LIB_PATH], capture_output=True, text=True).stdout
    for sym in ("hip_eval::init", "hip_eval::shutdown", "hip_eval::batch_evaluate", "hip_eval::upload_system",
                "hip_eval::get_device_info", "hip_eval::upload_emat"):
        assert sym in out, sym


def test_struct_layouts_match_header():
    from flexaidds_amd import abi
    assert C.sizeof(abi.CFResult) == 64
    assert C.sizeof(abi.ContactRec) == 24
    assert C.sizeof(abi.Thermo) == 64


def test_struct_sizes_against_the_c_compiler(tmp_path):
    """Every ctypes mirror has the size and field offsets the C compiler gives
    the header's struct."""
    import subprocess
    from flexaidds_amd import abi, hip
    mirrors = {"fa_params": abi.Params, "fa_ligand_desc": abi.LigandDesc, "fa_constraint": abi.Constraint, "fa_receptor_desc": abi.ReceptorDesc,
               "fa_cf_result": abi.CFResult, "fa_thermo": abi.Thermo, "fa_gene_lim": hip.GeneLim,
               "fa_ga_params": hip.GAParams, "fa_pool_params": hip.PoolParams, "fa_job": hip.Job,
               "fa_job_result": hip.JobResult, "fa_cluster_params": hip.ClusterParams,
               "fa_cluster_mode": hip.ClusterMode, "fa_cluster_out": hip.ClusterOut, "fa_super_cluster_out": hip.SuperClusterOut,
               "fa_request": hip.Request,
               "fa_served": hip.ServedRec, "fa_serve_stats": hip.ServeStats}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "flexaidds_amd.h"', "int main(void) {"]
    for cname, py in mirrors.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            cf = {"del_": "del", "from_": "from"}.get(f, f)
            lines.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {cf}));')
    lines.append("return 0; }")
    src = tmp_path / "sizes.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "sizes"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = dict(l.rsplit(" ", 1) for l in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split("\n") if l)
    for cname, py in mirrors.items():
        assert int(got[cname]) == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert int(got[f"{cname}.{f}"]) == getattr(py, f).offset, (cname, f)


def test_super_cluster_limit_matches_header():
    import re
    from flexaidds_amd import hip
    text = open(os.path.join(ROOT, "include", "flexaidds_amd.h")).read()
    assert int(re.search(r"#define FA_SUPER_CLUSTER_MAX (\d+)", text).group(1)) == hip.SUPER_CLUSTER_MAX
