"""The strong-shock kernel's shared-reciprocal quotients against the compiler's division, on the GPU
(tests/quotient_check.hip, built with the library into tests/bin/quotient_check): random operand pairs over the whole
band the shortcut is used in, the profile right-hand side on states at scale and on states built to fail the band test,
and whole profile integrations step by step. Every comparison is bit for bit; the shock-budget parity tests then check
the kernel itself against the oracle."""
import os
import re
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "bin", "quotient_check")

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lines():
    if not os.path.exists(EXE):
        pytest.fail("tests/bin/quotient_check is not built (the package build compiles it)")
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    out = {}
    for ln in r.stdout.splitlines():
        m = re.match(r"(\w+) cases (\d+) fast (\d+) fallback (\d+) differ (\d+)", ln)
        if m:
            out[m.group(1)] = dict(zip(("cases", "fast", "fallback", "differ"), map(int, m.groups()[1:])))
    assert set(out) == {"quot", "rhs", "step"}, r.stdout
    return out


def test_quotient_equals_division(lines):
    q = lines["quot"]
    assert q["cases"] == 2048 * 256 * 128 and q["differ"] == 0


def test_rhs_equals_plain_wherever_the_band_test_passes(lines):
    q = lines["rhs"]
    assert q["differ"] == 0
    # both outcomes of the band test occur (a quarter of the cases use a phase pair whose g0 - 1 is out of band, and
    # a third of the others are built to fail it)
    assert q["fast"] > 0.3 * q["cases"] and q["fallback"] > 0.3 * q["cases"]


def test_step_equals_plain_step(lines):
    q = lines["step"]
    assert q["differ"] == 0
    assert q["fast"] > 0 and q["fallback"] > 0 and q["cases"] == q["fast"] + q["fallback"]
