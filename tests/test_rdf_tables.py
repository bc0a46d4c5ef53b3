"""The RDF kernel's host-built lookup tables (image-index thresholds of the
reference's CPU and GPU rules, bin thresholds of the three binning rules) and
its table arithmetic, against the direct arithmetic of each rule on the CPU:
tests/cpp/rdf_tables_check.cpp, built with the host code that builds the
tables (lennard-jones-cuda_amd/csrc/lj_host.cpp).  No GPU involved."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(shutil.which(HIPCC) is None and not os.path.exists(HIPCC),
                    reason="hipcc (host compiler for the HIP headers) not found")
def test_rdf_tables_match_direct_rules(tmp_path):
    exe = tmp_path / "rdf_tables_check"
    cmd = [HIPCC, "-O2", "-std=c++17", "-w", f"-I{ROOT}/include",
           f"-I{ROOT}/lennard-jones-cuda_amd/csrc",
           f"{ROOT}/tests/cpp/rdf_tables_check.cpp",
           f"{ROOT}/lennard-jones-cuda_amd/csrc/lj_host.cpp", "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True, text=True, timeout=300)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    lines = [l for l in r.stdout.splitlines() if "mismatches" in l]
    assert len(lines) == 5 and all(l.split(":")[1].strip().startswith("0 ") for l in lines), r.stdout
    assert r.returncode == 0
