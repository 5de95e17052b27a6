"""Builds the drop-in test object tests/dropin/_build/libfa_dropin.so from
dropin.cpp against include/hip_eval.h and flexaidds_amd/libflexaidds_amd.so
(found at run time through its RUNPATH).  Run after the library is built:
    python tests/dropin/build.py
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "_build", "libfa_dropin.so")


def build(verbose: bool = False) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    pkg = os.path.join(ROOT, "flexaidds_amd")
    if not os.path.exists(os.path.join(pkg, "libflexaidds_amd.so")):
        raise RuntimeError("flexaidds_amd/libflexaidds_amd.so is not built")
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    tmp = OUT + ".tmp"
    cmd = [hipcc, "-O2", "-std=c++20", "-fPIC", "-shared", "--offload-arch=gfx950", "-Wall",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           f"-I{os.path.join(ROOT, 'include')}", os.path.join(HERE, "dropin.cpp"),
           f"-L{pkg}", "-lflexaidds_amd", "-Wl,-rpath,$ORIGIN/../../../flexaidds_amd", "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
