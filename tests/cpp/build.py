"""Builds tests/cpp/handler_run.bin, the C++ program that drives the particle
path through include/pfem2gpu_handler.hpp (linked against the in-tree
libpfem2gpu.so, which must be built first).

    python tests/cpp/build.py
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
SRC = os.path.join(HERE, "handler_run.cpp")
OUT = os.path.join(HERE, "handler_run.bin")


def compile_command(out=OUT, syntax_only=False, sanitize=False):
    """g++ command for the program (host code only; `sanitize` adds the
    address and undefined-behaviour sanitizers)."""
    cxx = shutil.which("g++") or "g++"
    cmd = [cxx, "-O2", "-std=c++17", "-Wall", "-Wextra", "-I" + os.path.join(ROOT, "include"), SRC]
    if sanitize:
        cmd[1:2] = ["-O1", "-g", "-fsanitize=address,undefined", "-fno-omit-frame-pointer"]
    if syntax_only:
        return cmd + ["-fsyntax-only"]
    return cmd + ["-L" + os.path.join(ROOT, "pfem2dealii_amd"), "-lpfem2gpu",
                  "-Wl,-rpath,$ORIGIN/../../pfem2dealii_amd", "-Wl,-rpath-link,/opt/rocm/lib", "-o", out]


def build(verbose=False):
    tmp = OUT + ".tmp"
    r = subprocess.run(compile_command(tmp), capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stderr)
        raise RuntimeError("g++ failed for handler_run.bin")
    os.replace(tmp, OUT)
    if verbose:
        print("built", OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
