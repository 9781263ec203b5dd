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


@pytest.mark.parametrize("mg", [False, True], ids=["nasg", "mie_grueneisen"])
def test_library_exports_every_declared_symbol(mg):
    """both builds, libspb6_hip.so and its Mie-Grueneisen build libspb6_hip_mg.so, export the whole C ABI"""
    from diffusionimex_amd import spb6
    lib = spb6.load(mg=mg)
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
    monkeypatch.setattr(spb6, "_libs", {})
    monkeypatch.setattr(spb6, "LIB_PATH", "/nonexistent/libspb6_hip.so")
    with pytest.raises(RuntimeError):
        spb6.load()


def test_nasg_library_rejects_mie_grueneisen_phases():
    """libspb6_hip.so is built without the Mie-Grueneisen branches: spb6_create refuses an MG phase (before any device
    call) and names the library that runs it"""
    from diffusionimex_amd import spb6
    lib = spb6.load(mg=False)
    cfg = spb6.make_config(16, eos_kind=(2, 0), mg=((0.25, 854.5e9, 20.5e9, 4.6, 1.35, 1840.0, 0.0), (0.0,) * 7))
    assert not lib.spb6_create(ctypes.byref(cfg), 0)
    assert b"libspb6_hip_mg.so" in lib.spb6_last_error()
    bad = spb6.make_config(16, eos_kind=(2, 0), mg=((0.25, 854.5e9, 20.5e9, 4.6, 1.35, 1840.0, 0.0), (0.0,) * 7), relax=1)
    assert not spb6.load(mg=True).spb6_create(ctypes.byref(bad), 0)
