"""Builds libjsi_mi355x.so for gfx950 with hipcc (in-tree)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libjsi_mi355x.so")
SOURCES = ["jsi_kernels.hip", "jsi_runtime.hip", "jsi_legacy.hip", "jsi_opabi.hip", "jsi_serve.hip", "jsi_refine.hip", "jsi_sort.hip",
           "jsi_noalloc.hip", "jsi_batch.hip", "jsi_isect_cells.hip"]

# -ffp-contract=off: no fused multiply-add, so float results match VERSION5
# (built without FMA on x86-64) and double results match the FP64 oracle.
# Packed FP32 VALU ops are disabled: some of their forms have given wrong results.
# -amdgpu-mfma-vgpr-form: matrix-core results go to ordinary VGPRs, so the
# comparisons read them without accumulator-register copies.
HIPFLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-mllvm", "-amdgpu-mfma-vgpr-form",
            "-Wno-unused-command-line-argument"]


def build(verbose=False, jobs=3):
    objs = []
    procs = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(CSRC, src.replace(".hip", ".o"))
        cmd = ["hipcc", *HIPFLAGS, "-c", path, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((subprocess.Popen(cmd), cmd))
        objs.append(obj)
    bad = [cmd for p, cmd in procs if p.wait() != 0]
    if bad:
        raise RuntimeError("hipcc failed: " + " ".join(bad[0]))
    cmd = ["hipcc", "--offload-arch=gfx950", "-shared", "-fPIC", "-Wl,--no-undefined", "-o", LIB + ".tmp", *objs]
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    build_legacy_driver(verbose)
    return LIB


def build_legacy_driver(verbose=False):
    """tests/legacy_driver: a C++ program written like the VERSION5 example
    drivers, compiled with g++ against include/ alone (no HIP headers) and
    linked to the library, as a maintainer's relinked driver would be."""
    src = os.path.join(ROOT, "tests", "legacy_driver.cpp")
    out = os.path.join(ROOT, "tests", "legacy_driver")
    cmd = ["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), src, "-o", out + ".tmp",
           "-L", HERE, "-ljsi_mi355x", "-Wl,-rpath,$ORIGIN/../jsi_cad_next_versions_amd"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
