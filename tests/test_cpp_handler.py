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


def _write_case(path, cfg, steps, order=0):
    mesh, prm = cfg["mesh"], cfg["params"]
    dim = mesh["dim"]
    X = np.ascontiguousarray(mesh["X"], dtype=np.float64)
    cv = np.ascontiguousarray(mesh["cell_vertices"], dtype=np.int32)
    cells = np.ascontiguousarray(cfg["cells"], dtype=np.int32)
    x = np.ascontiguousarray(cfg["x"], dtype=np.float64)
    V = np.ascontiguousarray(cfg["V"], dtype=np.float64)
    q = list(prm["particles_per_cell"]) + [1] * (3 - dim)
    with open(path, "wb") as f:
        f.write(struct.pack("<iii", dim, steps, order))
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


@pytest.mark.gpu
@pytest.mark.parametrize("name,scale,steps", [("channel2d", 0.5, 3), ("abc3d", 1.0 / 16, 2)])
def test_cpp_handler_in_the_reference_call_order(gpu_available, tmp_path, name, scale, steps):
    """pfem2Solver::run's order (src/pfem2solver.cpp:151-166): seed_particles
    before initialize_maps, then correct -> move -> project per step, with the
    FEM step replaced by V_old := projected field, V := the prescribed field
    (V_old starts at zero as the FEM's initial old solution). The C++ class
    equals the oracle (cells, positions and reference positions bitwise,
    velocities and the projection within 1e-11) and the Python binding run in
    the same order (bitwise), and reports the reference's four timer sections
    with their call counts."""
    from oracle import Oracle
    from pfem2dealii_amd import ParticleHandler
    from tests.harness import assert_fields_match, assert_particles_match
    print("[cpp] %s at %g, %d steps in the reference order" % (name, scale, steps), flush=True)
    sys.stdout.flush()
    if not os.path.exists(BIN):
        pytest.fail("tests/cpp/handler_run.bin is not built (python tests/cpp/build.py)")
    cfg = W.config(name, scale=scale)
    mesh, V = cfg["mesh"], cfg["V"]
    inp, outp = str(tmp_path / "case.bin"), str(tmp_path / "result.bin")
    _write_case(inp, dict(cfg, x=np.zeros((mesh["dim"], 0)), cells=np.zeros(0, dtype=np.int32)), steps, order=1)
    r = subprocess.run([BIN, inp, outp], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = _read_result(outp, mesh["dim"], mesh["n_dofs"])
    for sec, calls in (("Particle seeding", 1), ("Particle velocity correction", steps),
                       ("Particle transport", steps), ("Particle fields projection", steps)):
        line = [ln for ln in r.stdout.splitlines() if ln.startswith("| " + sec)]
        assert len(line) == 1 and int(line[0].split("|")[2]) == calls, r.stdout

    o = Oracle(mesh, cfg["params"])
    o.seed_lattice(V)
    o.initialize_maps()
    h = ParticleHandler(mesh, cfg["params"])
    h.seed_particles(V)
    h.initialize_maps()
    Vold_o = np.zeros_like(V)
    Vold_h = np.zeros_like(V)
    for _ in range(steps):
        o.correct(V, Vold_o)
        o.move(V)
        Vold_o = o.project()
        h.correct_particle_velocity(V, Vold_h)
        h.move_particles(V)
        Vold_h = h.project_particle_fields()
    want = h.download(xi=True)
    h.close()
    for key in ("x", "xi", "v", "cell"):
        assert np.array_equal(got[key], want[key]), key
    assert np.array_equal(np.nan_to_num(got["P"], nan=7.0), np.nan_to_num(Vold_h, nan=7.0))
    assert_particles_match({k: np.array(got[k]) for k in ("x", "xi", "v", "cell")}, o.download())
    assert_fields_match(np.array(got["P"]), Vold_o)
