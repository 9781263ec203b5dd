"""Analytic Newton Jacobian on the host (no GPU): the device headers' host path compiled with hipcc, every entry of
jacobian_analytic_row checked against forward differences of the residual in the reference's operation order
(cell_res) on random physical states, with the tolerance stated in tests/jacobian_host_check.cpp."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def _hipcc():
    p = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "bin", "hipcc")
    return p if os.path.exists(p) else shutil.which("hipcc")


def test_analytic_jacobian_matches_forward_differences_on_the_host(tmp_path):
    cc = _hipcc()
    if cc is None:
        pytest.skip("hipcc not available")
    exe = str(tmp_path / "jacobian_host_check")
    cmd = [cc, "-O2", "-std=c++17", "-ffp-contract=off", "--cuda-host-only",
           "-I", os.path.join(ROOT, "diffusionimex_amd", "csrc"), "-I", os.path.join(ROOT, "include"),
           os.path.join(HERE, "jacobian_host_check.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    r = subprocess.run([exe, "20000"], capture_output=True, text=True, timeout=120)
    last = r.stdout.strip().splitlines()[-1]
    assert r.returncode == 0, r.stdout
    fields = last.split()
    entries, bad, kinks = int(fields[1]), int(fields[3]), int(fields[5])
    assert bad == 0 and entries > 80000 and kinks > 1000, last
