"""The C++ handler class (include/pfem2gpu_handler.hpp, the reference's
pfem2ParticleHandler<dim> methods over the C ABI) driven by
tests/cpp/handler_run.cpp: the header compiles (CPU), and on the GPU the
program's particles and projected field after a few steps of
move -> project -> correct equal, bit for bit, those of the same sequence
through the Python binding (whose results the parity tests hold against the
oracle)."""
import os
import struct
import subprocess
import sys

import numpy as np
import pytest

from pfem2dealii_amd import workloads as W
from tests.cpp.build import OUT as BIN, compile_command


def test_handler_header_compiles():
    r = subprocess.run(compile_command(syntax_only=True), capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr


def _write_case(path, cfg, steps):
    mesh, prm = cfg["mesh"], cfg["params"]
    dim = mesh["dim"]
    X = np.ascontiguousarray(mesh["X"], dtype=np.float64)
    cv = np.ascontiguousarray(mesh["cell_vertices"], dtype=np.int32)
    cells = np.ascontiguousarray(cfg["cells"], dtype=np.int32)
    x = np.ascontiguousarray(cfg["x"], dtype=np.float64)
    V = np.ascontiguousarray(cfg["V"], dtype=np.float64)
    q = list(prm["particles_per_cell"]) + [1] * (3 - dim)
    with open(path, "wb") as f:
        f.write(struct.pack("<ii", dim, steps))
        f.write(struct.pack("<qqqq", X.shape[1], cv.shape[0], mesh["n_dofs"], cells.shape[0]))
        f.write(struct.pack("<IdI3I", prm["n_sub"], prm["dt"], prm["max_per_part"], *q))
        for a in (X, cv, cells, x, V):
            f.write(a.tobytes())


def _read_result(path, dim, n_dofs):
    b = open(path, "rb").read()
    n, rej = struct.unpack_from("<qq", b, 0)
    o = 16
    out = {"rejected": rej}
    for key in ("x", "xi", "v"):
        out[key] = np.frombuffer(b, dtype=np.float64, count=dim * n, offset=o).reshape(dim, n)
        o += 8 * dim * n
    out["cell"] = np.frombuffer(b, dtype=np.int32, count=n, offset=o)
    o += 4 * n
    out["P"] = np.frombuffer(b, dtype=np.float64, count=dim * n_dofs, offset=o).reshape(dim, n_dofs)
    o += 8 * dim * n_dofs
    assert o == len(b)
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("name,scale,steps", [("channel2d", 0.5, 3), ("abc3d", 1.0 / 16, 2)])
def test_cpp_handler_matches_python_binding(gpu_available, tmp_path, name, scale, steps):
    from pfem2dealii_amd import ParticleHandler
    print("[cpp] %s at %g, %d steps" % (name, scale, steps), flush=True)
    sys.stdout.flush()
    if not os.path.exists(BIN):
        pytest.fail("tests/cpp/handler_run.bin is not built (python tests/cpp/build.py)")
    cfg = W.config(name, scale=scale)
    inp, outp = str(tmp_path / "case.bin"), str(tmp_path / "result.bin")
    _write_case(inp, cfg, steps)
    r = subprocess.run([BIN, inp, outp], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    mesh = cfg["mesh"]
    got = _read_result(outp, mesh["dim"], mesh["n_dofs"])

    h = ParticleHandler(mesh, cfg["params"])
    h.initialize_maps()
    rej = h.insert_particles(cfg["x"], cfg["cells"])
    P = None
    for _ in range(steps):
        h.move_particles(cfg["V"])
        P = h.project_particle_fields()
        h.correct_particle_velocity(cfg["V"], P)
    want = h.download(xi=True)
    h.close()
    assert got["rejected"] == rej
    for key in ("x", "xi", "v", "cell"):
        assert np.array_equal(got[key], want[key]), key
    assert np.array_equal(np.isnan(got["P"]), np.isnan(P))
    assert np.array_equal(np.nan_to_num(got["P"], nan=7.0), np.nan_to_num(P, nan=7.0))
