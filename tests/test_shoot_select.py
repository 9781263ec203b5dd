"""The strong-shock kernel's select form of the shooting state machine (spb6_device.h: shoot_next_sel) against the
branch form shoot_next, on the CPU (tests/shoot_check.hip, built with the library into tests/bin/shoot_check; compiled
here if it is missing): single transitions from random states over every phase, bracket edge, sign, zero, NaN and
infinity and the Illinois iteration cap, and whole shootings on a synthetic F. Every returned t and every field of the
state must agree bit for bit. The GPU run of the same cases is tests/test_gpu_shoot_select.py."""
import os
import re
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "bin", "shoot_check")


def run_check(mode):
    if not os.path.exists(EXE):
        from diffusionimex_amd import build
        os.makedirs(os.path.dirname(EXE), exist_ok=True)
        subprocess.run([build.hipcc()] + build.flags() + ["-o", EXE, os.path.join(HERE, "shoot_check.hip")], check=True)
    r = subprocess.run([EXE, mode], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    out = {}
    for ln in r.stdout.splitlines():
        m = re.match(r"(\w+) cases (\d+)(?: steps (\d+))? differ (\d+)", ln)
        if m:
            out[m.group(1)] = (int(m.group(2)), int(m.group(3) or 0), int(m.group(4)))
    assert set(out) == {"single", "chain"}, r.stdout
    return out


def test_select_form_equals_branch_form_on_the_host():
    out = run_check("host")
    assert out["single"][0] == 1 << 20 and out["single"][2] == 0, out
    assert out["chain"][0] == 1 << 16 and out["chain"][1] > 1 << 18 and out["chain"][2] == 0, out
