"""Compile libgscore.so (the HIP kernels + C ABI) in-tree for gfx950.

    python -m imaged11_amd.build

hipcc cross-compiles without a GPU.  The library lands next to this file,
where ``imaged11_amd._lib`` loads it.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libgscore.so")
SOURCES = [os.path.join(CSRC, "gscore.hip")]

ARCH = os.environ.get("GSCORE_ARCH", "gfx950")

FLAGS = [
    "--offload-arch=" + ARCH,
    "-O3", "-fPIC", "-shared", "-std=c++17",
    # parity: no FMA contraction anywhere; FMA only where written explicitly
    "-ffp-contract=off",
    "-mcode-object-version=5",
    # packed FP32 instructions are never wanted in this library
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build_library(verbose=False):
    """Always recompiles; writes atomically so a running process never sees
    a half-written library."""
    tmp = LIB + ".tmp%d" % os.getpid()
    cmd = [hipcc()] + FLAGS + ["-I", os.path.join(ROOT, "include")] + SOURCES + ["-o", tmp]
    if verbose:
        print(" ".join(cmd))
    out = subprocess.run(cmd, capture_output=True, text=True)
    noise = [ln for ln in (out.stdout + out.stderr).splitlines() if "packed-fp32-ops" not in ln]
    if out.returncode != 0:
        raise RuntimeError("hipcc failed (%d):\n%s" % (out.returncode, "\n".join(noise)))
    if verbose and noise:
        print("\n".join(noise))
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build_library(verbose="-v" in sys.argv))
