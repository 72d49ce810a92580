"""Builds libpfem2gpu.so in-tree with hipcc for gfx950.

    python -m pfem2dealii_amd.build

Flags: -ffp-contract=off keeps every product and sum separately rounded (the
parity contract with the reference's arithmetic); packed FP32 instructions are
disabled target-wide.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libpfem2gpu.so")
SOURCES = ["pfem2gpu.hip"]
HEADERS = ["pfem2gpu_device.h", "pfem2gpu_kernels.h", "pfem2gpu_adjoint.h", "pfem2gpu_output.h", "pfem2gpu_wire.h",
           "pfem2gpu_fem.h"]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def flags():
    return ["-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared", "-ffp-contract=off",
            "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-command-line-argument", "-I" + os.path.join(HERE, "..", "include")]


def build(verbose=False):
    """Compiles every HIP source into one shared library (always rebuilds)."""
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + flags() + srcs + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed for libpfem2gpu.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
