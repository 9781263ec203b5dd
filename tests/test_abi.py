"""The C-ABI library loads and exports every entry point include/spb6_imex.h declares (no GPU calls)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = []
    for fn in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if fn.endswith(".h"):
            src = open(os.path.join(ROOT, "include", fn)).read()
            src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
            names += re.findall(r"\b(spb6_[a-z0-9_]+)\s*\(", src)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = _declared()
    assert "spb6_create" in names and "spb6_run" in names and len(names) >= 20


def test_library_exports_every_declared_symbol():
    from diffusionimex_amd import spb6
    lib = spb6.load()
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, f"missing exports: {missing}"
    assert set(spb6.EXPORTS) <= set(_declared())


def test_python_structs_match_header_sizes():
    """ctypes mirrors of the header structs have the C layout (checked against a compiled probe)."""
    from diffusionimex_amd import spb6
    src = os.path.join(ROOT, "include", "spb6_imex.h")
    import subprocess
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "p.c")
        open(c, "w").write('#include "%s"\n#include <stdio.h>\nint main(){printf("%%zu %%zu %%zu %%zu %%zu\\n", sizeof(spb6_config_t), '
                           'sizeof(spb6_stats_t), sizeof(spb6_kernel_times_t), sizeof(spb6_batch_problem_t), '
                           'sizeof(spb6_batch_tune_t));return 0;}\n' % src)
        exe = os.path.join(d, "p")
        subprocess.run(["gcc", c, "-o", exe], check=True)
        out = subprocess.run([exe], stdout=subprocess.PIPE, check=True).stdout.decode().split()
    assert int(out[0]) == ctypes.sizeof(spb6.Config)
    assert int(out[1]) == ctypes.sizeof(spb6.Stats)
    assert int(out[2]) == ctypes.sizeof(spb6.KernelTimes)
    from diffusionimex_amd import batch
    assert int(out[3]) == ctypes.sizeof(batch.BatchProblem)
    assert int(out[4]) == ctypes.sizeof(batch.Tune)


def test_missing_library_fails_loudly(monkeypatch):
    from diffusionimex_amd import spb6
    monkeypatch.setattr(spb6, "_lib", None)
    monkeypatch.setattr(spb6, "LIB_PATH", "/nonexistent/libspb6_hip.so")
    with pytest.raises(RuntimeError):
        spb6.load()
