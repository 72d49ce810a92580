"""Host code under the address and undefined-behaviour sanitizers (CPU, g++,
no GPU code): the oracle's entry points driven by oracle/san_driver.cpp
(`make -C oracle sanitize`: particle oracle in both modes on 2D/3D box and
perturbed meshes, insertion with rejects, correction, moves with inflow and
outflow, projection, removal, download, and the three FEM assemblies), and
the C++ handler program built and linked with the sanitizers."""
import os
import subprocess

import pytest

from tests.cpp.build import compile_command

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_oracle_clean_under_sanitizers():
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "sanitized oracle run: ok" in r.stdout


def test_handler_program_builds_with_sanitizers(tmp_path):
    if not os.path.exists(os.path.join(ROOT, "pfem2dealii_amd", "libpfem2gpu.so")):
        pytest.skip("libpfem2gpu.so not built")
    out = str(tmp_path / "handler_run_san.bin")
    r = subprocess.run(compile_command(out, sanitize=True), capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    assert os.path.exists(out)
