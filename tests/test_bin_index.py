"""The device GA's bin count (flexaidds_amd/csrc/bin_index.h) against the
reference's genetoic accumulation loop, on the host: bins 1/n and random
bins, u at and around every sum the loop reaches, random u and GA genes."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "flexaidds_amd", "csrc")


def test_bin_index_equals_the_loop(tmp_path):
    exe = str(tmp_path / "bin_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", f"-I{CSRC}", "-o", exe,
                    os.path.join(HERE, "bin_index_check.cpp")], check=True)
    r = subprocess.run([exe, "400"], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert " 0 mismatches" in r.stdout
