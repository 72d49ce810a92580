"""Particle output files (output_particle_solution,
src/pfem2particlehandler.cpp:217-317): the library's writer against the
oracle's restatement of the reference's text, the restatement pinned by
known answers of the C/C++ formatting rules it relies on and by parsing the
files back as VTK XML. The CPU tests use the host-array writers; the GPU test
writes a handle's particles after a move."""
import math
import os
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import pfem2dealii_amd as P
from oracle import vtk_text as VT


def test_format_rules_known_answers():
    # printf("%.6g") and printf("%f") on values whose text is fixed by the C
    # standard's rules (6 significant digits, trailing zeros removed, exponent
    # of at least two digits when the exponent is < -4 or >= 6)
    assert [VT.stream_double(a) for a in (0.1, 0.5, 1e-05, 100000.0, 1000000.0, 123456789.0, -0.0, 2.0 / 3.0)] == \
        ["0.1", "0.5", "1e-05", "100000", "1e+06", "1.23457e+08", "-0", "0.666667"]
    assert [VT.to_string(a) for a in (0.1, 1e-07, 123456789.0, -0.0, 2.0 / 3.0, 1.0)] == \
        ["0.100000", "0.000000", "123456789.000000", "-0.000000", "0.666667", "1.000000"]
    assert VT.stream_double(float("nan")) == "nan" and VT.stream_double(-float("nan")) == "-nan"
    assert [VT.int_to_string(v, 2) for v in (0, 5, 12, 123)] == ["00", "05", "12", "123"]
    assert VT.int_to_string(7, 3) == "007"
    assert VT.file_names(5, 2) == ("particles-05.002.vtu", "particles-05.pvtu")


def _values(dim, n, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-3.0, 3.0, (dim, n)) * 10.0 ** rng.integers(-7, 9, (dim, n))
    v = rng.normal(size=(dim, n))
    special = [0.0, -0.0, 1e-05, 1e+06, 123456789.0, 0.1, 1.0, 2.0 / 3.0]
    for k, a in enumerate(special[: n]):
        x[0, k] = a
        x[-1, k] = -a
        v[0, k] = a
    return x, v


@pytest.mark.parametrize("dim", [2, 3])
def test_vtu_text_matches_the_reference_format(tmp_path, dim):
    x, v = _values(dim, 40, dim)
    path = tmp_path / "piece.vtu"
    P.write_particles_vtu(str(path), x, v, subdomain=3)
    assert path.read_text() == VT.vtu_text(x, v, subdomain=3)


def test_vtu_parses_back_as_vtk(tmp_path):
    x, v = _values(2, 25, 9)
    path = tmp_path / "piece.vtu"
    P.write_particles_vtu(str(path), x, v, subdomain=0)
    root = ET.parse(str(path)).getroot()
    assert root.tag == "VTKFile" and root.get("type") == "UnstructuredGrid"
    piece = root.find("UnstructuredGrid/Piece")
    assert piece.get("NumberOfPoints") == "25" and piece.get("NumberOfCells") == "25"
    pts = np.array(piece.find("Points/DataArray").text.split(), dtype=float).reshape(-1, 3)
    # first component: 6 significant digits; the others: 6 decimals
    assert np.allclose(pts[:, 0], x[0], rtol=1e-5, atol=0.0)
    assert np.allclose(pts[:, 1], x[1], rtol=0.0, atol=5.1e-7)
    assert (pts[:, 2] == 0.0).all()
    arrays = {a.get("Name"): a.text.split() for a in piece.iter("DataArray") if a.get("Name")}
    assert arrays["connectivity"] == [str(i) for i in range(25)]
    assert arrays["offsets"] == [str(i + 1) for i in range(25)]
    assert arrays["types"] == ["1"] * 25
    assert arrays["subdomain"] == ["0"] * 25
    vel = np.array(arrays["velocity"], dtype=float).reshape(-1, 3)
    assert np.allclose(vel[:, 0], v[0], rtol=1e-5) and np.allclose(vel[:, 1], v[1], atol=5.1e-7)


def test_empty_piece_and_index_file(tmp_path):
    e = np.zeros((3, 0))
    path = tmp_path / "empty.vtu"
    P.write_particles_vtu(str(path), e, e)
    assert path.read_text() == VT.vtu_text(e, e)
    for step, ranks in ((5, 1), (123, 3)):
        path = tmp_path / ("index-%d.pvtu" % step)
        P.write_particles_pvtu(str(path), step, ranks)
        text = path.read_text()
        assert text == VT.pvtu_text(step, ranks)
        assert text.count("<Piece Source=") == ranks
    root = ET.fromstring(VT.pvtu_text(7, 2))
    assert [p.get("Source") for p in root.iter("Piece")] == ["particles-07.000.vtu", "particles-07.001.vtu"]


def test_writer_errors(tmp_path):
    lib = P.load_library()
    x = np.zeros((2, 1))
    bad = os.fsencode(str(tmp_path / "missing-dir" / "p.vtu"))
    assert lib.pfem2gpu_write_particles_vtu(bad, 2, 1, x.ctypes.data, x.ctypes.data, 0) == 5
    assert b"cannot write" in lib.pfem2gpu_last_error()
    assert lib.pfem2gpu_write_particles_vtu(bad, 4, 1, x.ctypes.data, x.ctypes.data, 0) == 1
    assert lib.pfem2gpu_write_particles_pvtu(os.fsencode(str(tmp_path / "i.pvtu")), 1, 0) == 1
    assert lib.pfem2gpu_output_particle_solution(None, 0, 0, 1, 0, b".") == 1


@pytest.mark.gpu
def test_output_after_a_move(gpu_available, tmp_path):
    import torch
    from pfem2dealii_amd import ParticleHandler, workloads as W
    cfg = W.config("channel2d")
    h = ParticleHandler(cfg["mesh"], cfg["params"])
    h.initialize_maps()
    h.insert_particles(cfg["x"], cfg["cells"])
    V = torch.tensor(cfg["V"], dtype=torch.float64, device="cuda")
    h.set_velocity_from_field(V)
    h.move_particles(V)
    d = h.download(xi=False)
    h.output_particle_solution(3, directory=str(tmp_path), rank=1, n_ranks=2, subdomain=1)
    piece, index = VT.file_names(3, 1)
    assert (tmp_path / piece).read_text() == VT.vtu_text(d["x"], d["v"], subdomain=1)
    assert not (tmp_path / index).exists()  # only rank 0 writes the index
    h.output_particle_solution(3, directory=str(tmp_path), rank=0, n_ranks=2, subdomain=0)
    piece0, _ = VT.file_names(3, 0)
    assert (tmp_path / piece0).read_text() == VT.vtu_text(d["x"], d["v"], subdomain=0)
    assert (tmp_path / index).read_text() == VT.pvtu_text(3, 2)
    assert not math.isnan(float(d["x"].sum()))
    h.close()
