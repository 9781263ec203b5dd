"""The reference-side binding of INTEGRATION.md, compiled: tests/abi_replay.cpp drives the reference's time loop
(PC:1230-1296) through the C ABI, replaying every attempt stage by stage (spb6_attempt_begin, spb6_stage_combine,
spb6_stage_update, spb6_stage_pressure, spb6_attempt_accept with the caller's own Butcher tables), and compares it
with spb6_run on a second context bit for bit. The build is a CPU test; the runs need the GPU."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIBDIR = os.path.join(ROOT, "diffusionimex_amd")


def _build(dst):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("g++ not available")
    if not os.path.exists(os.path.join(LIBDIR, "libspb6_hip.so")):
        pytest.skip("HIP library not built")
    exe = os.path.join(dst, "abi_replay")
    cmd = [cxx, "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), os.path.join(HERE, "abi_replay.cpp"),
           "-L", LIBDIR, "-lspb6_hip", f"-Wl,-rpath,{LIBDIR}", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return exe


def test_binding_compiles_and_links(tmp_path):
    assert os.path.exists(_build(str(tmp_path)))


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    return _build(str(tmp_path_factory.mktemp("abi")))


@pytest.mark.gpu
@pytest.mark.parametrize("case,rk,steps", [(0, 2, 40), (1, 2, 25), (0, 3, 30), (1, 3, 20)])
def test_stage_by_stage_replay_equals_spb6_run_bitwise(exe, case, rk, steps):
    r = subprocess.run([exe, str(case), str(rk), str(steps)], capture_output=True, text=True, timeout=240)
    print(r.stdout.strip())
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("identical") and f"steps {steps}" in r.stdout
