"""Builds the gfx950 HIP library libspb6_hip.so in-tree with hipcc, its Mie-Grueneisen build libspb6_hip_mg.so (the same
sources with SPB6_MG=1: the EOS branches of Cochran-Chan and JWL phases compiled in, which the NASG-only library folds
away to keep its kernels' registers), and beside them the GPU self-check programs of the test suite (tests/bin/), with
the same flags.

Flags: -ffp-contract=off keeps every FP64 operation separately rounded (the reference's results are reproduced bit for bit);
the packed-FP32 target feature is disabled so that no v_pk_*_f32 instruction is generated.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libspb6_hip.so")
LIB_MG = os.path.join(HERE, "libspb6_hip_mg.so")
SOURCES = ["spb6_kernels.hip", "spb6_variants.hip", "spb6_hllc.hip", "spb6_mlp.hip", "spb6_batch.hip", "spb6_api.hip"]
ARCH = os.environ.get("SPB6_ARCH", "gfx950")
TESTS = os.path.join(HERE, "..", "tests")
CHECKS = [("quotient_check.hip", "quotient_check"), ("shoot_check.hip", "shoot_check")]   # tests/<source> -> tests/bin/<program>


def hipcc():
    p = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "bin", "hipcc")
    return p if os.path.exists(p) else "hipcc"


def flags():
    return ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-result", "-I", os.path.join(HERE, "..", "include")]


def build(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    libs = [(LIB, [], ""), (LIB_MG, ["-DSPB6_MG=1"], ".mg")]
    objs = {lib: [] for lib, _, _ in libs}
    procs = []
    for lib, defs, tag in libs:
        for s in srcs:
            o = os.path.join(CSRC, os.path.basename(s) + tag + ".o")
            cmd = [hipcc()] + flags() + defs + ["-c", s, "-o", o]
            if verbose:
                print(" ".join(cmd))
            procs.append((subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT), s))
            objs[lib].append(o)
    os.makedirs(os.path.join(TESTS, "bin"), exist_ok=True)
    for src, exe in CHECKS:
        s = os.path.join(TESTS, src)
        if os.path.exists(s):
            cmd = [hipcc()] + flags() + ["-o", os.path.join(TESTS, "bin", exe), s]
            if verbose:
                print(" ".join(cmd))
            procs.append((subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT), s))
    for p, s in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out.decode(errors="replace"))
            raise RuntimeError(f"hipcc failed on {s}")
    for lib, _, _ in libs:
        tmp = lib + ".tmp"
        cmd = [hipcc()] + flags() + ["-shared", "-o", tmp] + objs[lib]
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
        if r.returncode != 0:
            sys.stderr.write(r.stdout.decode(errors="replace"))
            raise RuntimeError("hipcc link failed")
        os.replace(tmp, lib)
        for o in objs[lib]:
            os.remove(o)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
