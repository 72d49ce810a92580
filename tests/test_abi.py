"""CPU checks of the C ABI and the host-side harness (no GPU needed)."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

import pfem2dealii_amd as P
from pfem2dealii_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "pfem2gpu.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|const char\*|void\*)\s+(pfem2gpu_\w+)\s*\(", txt, re.M)))


def test_library_exports_every_declared_symbol():
    lib = P.load_library()
    names = _declared()
    assert len(names) >= 18
    missing = [n for n in names if not hasattr(lib, n)]
    assert missing == []
    assert set(names) == set(P.EXPORTS)


def test_create_rejects_bad_arguments_before_touching_the_device():
    lib = P.load_library()
    h = ctypes.c_void_p()
    assert lib.pfem2gpu_create(None, None, None, ctypes.byref(h)) == 1
    assert b"null" in lib.pfem2gpu_last_error()
    X = np.zeros((2, 4))
    cv = np.array([[0, 1, 2, 3]], dtype=np.int32)
    m = P._Mesh(4, 4, X.ctypes.data, 1, cv.ctypes.data, None, 4)
    p = P._Params(5, 0.1, 5, (ctypes.c_uint32 * 3)(2, 2, 2))
    assert lib.pfem2gpu_create(ctypes.byref(m), ctypes.byref(p), None, ctypes.byref(h)) == 1
    assert b"dim" in lib.pfem2gpu_last_error()
    bad = np.array([[0, 1, 2, 7]], dtype=np.int32)
    m = P._Mesh(2, 4, X.ctypes.data, 1, bad.ctypes.data, None, 4)
    assert lib.pfem2gpu_create(ctypes.byref(m), ctypes.byref(p), None, ctypes.byref(h)) == 1
    assert b"out of range" in lib.pfem2gpu_last_error()
    m = P._Mesh(2, 4, X.ctypes.data, 1, cv.ctypes.data, None, 4)
    p0 = P._Params(5, 0.1, 5, (ctypes.c_uint32 * 3)(0, 2, 2))
    assert lib.pfem2gpu_create(ctypes.byref(m), ctypes.byref(p0), None, ctypes.byref(h)) == 1
    assert lib.pfem2gpu_initialize_maps(None) == 1
    assert lib.pfem2gpu_move_particles(None, None, 0) == 1


def test_splitmix_stream_is_fixed():
    # first outputs of splitmix64 from state 0 (published test values)
    assert int(W.splitmix64(np.array([0], dtype=np.uint64))[0]) == 0xE220A8397B1DCDAF
    a = W.uniform(7, 1, 1000)
    b = W.uniform(7, 1, 1000)
    assert np.array_equal(a, b) and (a >= 0).all() and (a < 1).all()
    assert abs(a.mean() - 0.5) < 0.05
    assert not np.array_equal(W.uniform(7, 2, 1000), a)


def test_poisson_counts():
    c = W.poisson(3, 7, 200000, 8.0, 32)
    assert c.min() >= 0 and c.max() <= 32
    assert abs(c.mean() - 8.0) < 0.05
    assert abs(c.var() - 8.0) < 0.2


def test_structured_mesh_numbering():
    m = W.structured_mesh([3, 2], [0.0, 0.0], [3.0, 2.0])
    assert m["X"].shape == (2, 12)
    # vertex (i, j) = i + 4 j ; cell (i, j) = i + 3 j ; vertex order bit0 = +x, bit1 = +y
    assert m["cell_vertices"][4].tolist() == [5, 6, 9, 10]
    assert m["X"][:, 6].tolist() == [2.0, 1.0]
    m3 = W.structured_mesh([2, 2, 2], [0.0] * 3, [1.0] * 3)
    assert m3["cell_vertices"][7].tolist() == [13, 14, 16, 17, 22, 23, 25, 26]


def test_config_shapes_and_dt():
    c1 = W.config("channel2d")
    assert c1["mesh"]["cell_vertices"].shape[0] == 2048 and c1["mesh"]["n_dofs"] == 2145
    assert c1["x"].shape == (2, 32768)
    assert c1["params"]["dt"] == 0.015625  # 0.5 * (1/32) / max|u| with max|u| = 1 at y = 0.5
    c2 = W.config("taylor_green2d", scale=1.0 / 64)
    assert c2["mesh"]["n"] == [32, 16]
    assert abs(c2["params"]["dt"] - 0.5 * (2 * math.pi / 32)) < 1e-15
    c3 = W.config("abc3d", scale=1.0 / 16)
    assert c3["mesh"]["n"] == [8, 8, 8] and c3["x"].shape[0] == 3
    inst = W.batched_instances(8, scale=0.25)
    assert len(inst) == 8 and all(i["mesh"]["n"][0] == 2 * i["mesh"]["n"][1] for i in inst)


def test_bench_refuses_more_than_one_gpu():
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True)
    assert r.returncode == 2 and "one GPU" in r.stderr
