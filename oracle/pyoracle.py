"""ctypes wrapper of the C++ oracle (oracle/pfem2_oracle.cpp). Test infrastructure.

Arrays cross the boundary as numpy float64/int32, vector fields as [dim][n].
"""
import ctypes
import os
import subprocess

import numpy as np

MODE_CANONICAL = 0
MODE_REFERENCE = 1

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liborc.so")
_lib = None

_d = ctypes.POINTER(ctypes.c_double)
_i = ctypes.POINTER(ctypes.c_int)
_u = ctypes.POINTER(ctypes.c_uint)
_l = ctypes.POINTER(ctypes.c_int64)


def build_oracle(force=False):
    """Compile liborc.so with oracle/Makefile (g++ -O2 -ffp-contract=off)."""
    if force or not os.path.exists(_LIB):
        subprocess.run(["make", "-s"] + (["-B"] if force else []) + ["-C", _HERE, "liborc.so"], check=True)
    return _LIB


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build_oracle()
        lib = ctypes.CDLL(_LIB)
        lib.orc_create.restype = ctypes.c_void_p
        lib.orc_create.argtypes = [ctypes.c_int, ctypes.c_int64, _d, ctypes.c_int64, _i, _i, ctypes.c_int64,
                                   ctypes.c_uint, ctypes.c_double, ctypes.c_uint, _u, ctypes.c_int]
        lib.orc_destroy.argtypes = [ctypes.c_void_p]
        lib.orc_initialize_maps.argtypes = [ctypes.c_void_p]
        lib.orc_insert.restype = ctypes.c_int64
        lib.orc_insert.argtypes = [ctypes.c_void_p, ctypes.c_int64, _d, _i, _d]
        lib.orc_seed_lattice.argtypes = [ctypes.c_void_p, _d]
        lib.orc_set_velocity_from_field.argtypes = [ctypes.c_void_p, _d]
        lib.orc_correct.argtypes = [ctypes.c_void_p, _d, _d]
        lib.orc_move.restype = ctypes.c_int
        lib.orc_move.argtypes = [ctypes.c_void_p, _d]
        lib.orc_project.argtypes = [ctypes.c_void_p, _d]
        lib.orc_project_weights.argtypes = [ctypes.c_void_p, _d, _d]
        lib.orc_remove.argtypes = [ctypes.c_void_p, ctypes.c_int64, _l]
        lib.orc_count.restype = ctypes.c_int64
        lib.orc_count.argtypes = [ctypes.c_void_p]
        lib.orc_download.argtypes = [ctypes.c_void_p, _d, _d, _d, _i]
        lib.orc_stats.argtypes = [ctypes.c_void_p, _l]
        lib.orc_shape.argtypes = [ctypes.c_void_p, _d, _d]
        lib.orc_fwd.argtypes = [ctypes.c_void_p, ctypes.c_int, _d, _d]
        lib.orc_inv.restype = ctypes.c_int
        lib.orc_inv.argtypes = [ctypes.c_void_p, ctypes.c_int, _d, _d]
        lib.orc_is_box.restype = ctypes.c_int
        lib.orc_is_box.argtypes = [ctypes.c_void_p, ctypes.c_int]
        lib.orc_closest_vertex.restype = ctypes.c_int
        lib.orc_closest_vertex.argtypes = [ctypes.c_void_p, ctypes.c_int, _d, _d]
        lib.orc_v2c.restype = ctypes.c_int
        lib.orc_v2c.argtypes = [ctypes.c_void_p, ctypes.c_int, _i, _d]
        lib.orc_key.restype = ctypes.c_uint64
        lib.orc_key.argtypes = [_d, ctypes.c_int]
        _u8 = ctypes.POINTER(ctypes.c_uint8)
        lib.orc_fem_pattern.restype = ctypes.c_int64
        lib.orc_fem_pattern.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_int64, _i, _l, _i]
        lib.orc_fem_pressure.restype = ctypes.c_int
        lib.orc_fem_pressure.argtypes = [ctypes.c_int, ctypes.c_int64, _d, ctypes.c_int64, _i, _i, ctypes.c_int64, _d,
                                         _u8, ctypes.c_double, ctypes.c_double, _l, _i, _d, _d]
        lib.orc_fem_prediction.restype = ctypes.c_int
        lib.orc_fem_prediction.argtypes = [ctypes.c_int, ctypes.c_int64, _d, ctypes.c_int64, _i, _i, ctypes.c_int64,
                                           _d, _d, _u8, ctypes.c_double, ctypes.c_double, ctypes.c_double, _l, _i,
                                           _d, _d]
        lib.orc_fem_vcorr.restype = ctypes.c_int
        lib.orc_fem_vcorr.argtypes = [ctypes.c_int, ctypes.c_int64, _d, ctypes.c_int64, _i, _i, ctypes.c_int64, _d,
                                      _d, ctypes.c_double, ctypes.c_double, _l, _i, _d, _d]
        _lib = lib
    return _lib


def _dp(a):
    return a.ctypes.data_as(_d) if a is not None else None


def _ip(a):
    return a.ctypes.data_as(_i) if a is not None else None


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Oracle:
    """CPU restatement of pfem2ParticleHandler<dim> (src/pfem2particlehandler.h:18-71)."""

    def __init__(self, mesh, params, mode=MODE_CANONICAL):
        lib = _load()
        self.lib = lib
        self.dim = int(mesh["dim"])
        self.ndofs = int(mesh["n_dofs"])
        self._X = _f64(mesh["X"])
        self._cv = np.ascontiguousarray(mesh["cell_vertices"], dtype=np.int32)
        self._cd = np.ascontiguousarray(mesh.get("cell_dofs", mesh["cell_vertices"]), dtype=np.int32)
        q = np.array(list(params["particles_per_cell"]) + [1] * (3 - len(params["particles_per_cell"])),
                     dtype=np.uint32)
        self.h = lib.orc_create(self.dim, self._X.shape[1], _dp(self._X), self._cv.shape[0], _ip(self._cv),
                                _ip(self._cd), self.ndofs, int(params["n_sub"]), float(params["dt"]),
                                int(params["max_per_part"]), q.ctypes.data_as(_u), int(mode))
        if not self.h:
            raise ValueError("oracle: unsupported mesh")

    def close(self):
        if self.h:
            self.lib.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def initialize_maps(self):
        self.lib.orc_initialize_maps(self.h)

    def insert(self, x, cell, v=None):
        x = _f64(x)
        cell = np.ascontiguousarray(cell, dtype=np.int32)
        v = _f64(v) if v is not None else None
        return self.lib.orc_insert(self.h, cell.shape[0], _dp(x), _ip(cell), _dp(v))

    def seed_lattice(self, V):
        self.lib.orc_seed_lattice(self.h, _dp(_f64(V)))

    def set_velocity_from_field(self, V):
        self.lib.orc_set_velocity_from_field(self.h, _dp(_f64(V)))

    def correct(self, V, Vold):
        self.lib.orc_correct(self.h, _dp(_f64(V)), _dp(_f64(Vold)))

    def move(self, V):
        if self.lib.orc_move(self.h, _dp(_f64(V))) != 0:
            raise RuntimeError("oracle: initialize_maps() not called")

    def project(self):
        out = np.empty((self.dim, self.ndofs), dtype=np.float64)
        self.lib.orc_project(self.h, _dp(out))
        return out

    def project_weights(self):
        """(projected V, nodal weights W_j = sum_p N_j(xi_p))."""
        out = np.empty((self.dim, self.ndofs), dtype=np.float64)
        w = np.empty(self.ndofs, dtype=np.float64)
        self.lib.orc_project_weights(self.h, _dp(out), _dp(w))
        return out, w

    def remove(self, index):
        idx = np.ascontiguousarray(index, dtype=np.int64).ravel()
        self.lib.orc_remove(self.h, idx.size, idx.ctypes.data_as(_l))

    def count(self):
        return int(self.lib.orc_count(self.h))

    def download(self):
        n = self.count()
        x = np.empty((self.dim, n))
        xi = np.empty((self.dim, n))
        v = np.empty((self.dim, n))
        cell = np.empty(n, dtype=np.int32)
        self.lib.orc_download(self.h, _dp(x), _dp(xi), _dp(v), _ip(cell))
        return {"x": x, "xi": xi, "v": v, "cell": cell}

    def stats(self):
        out = np.zeros(4, dtype=np.int64)
        self.lib.orc_stats(self.h, out.ctypes.data_as(_l))
        return dict(zip(["walks", "deleted_locate", "deleted_overflow", "refilled"], out.tolist()))

    # restated deal.II primitives
    def shape(self, xi):
        xi = _f64(xi)
        N = np.empty(1 << self.dim)
        self.lib.orc_shape(self.h, _dp(xi), _dp(N))
        return N

    def fwd(self, cell, xi):
        x = np.empty(self.dim)
        self.lib.orc_fwd(self.h, int(cell), _dp(_f64(xi)), _dp(x))
        return x

    def inv(self, cell, x):
        """(status, xi): status 1 inside, 0 outside, -1 transformation failed."""
        xi = np.zeros(self.dim)
        st = self.lib.orc_inv(self.h, int(cell), _dp(_f64(x)), _dp(xi))
        return st, xi

    def is_box(self, cell):
        return bool(self.lib.orc_is_box(self.h, int(cell)))

    def closest_vertex(self, cell, xi_old, u):
        return self.lib.orc_closest_vertex(self.h, int(cell), _dp(_f64(xi_old)), _dp(_f64(u)))

    def v2c(self, v):
        n = self.lib.orc_v2c(self.h, int(v), None, None)
        cells = np.empty(n, dtype=np.int32)
        dirs = np.empty((n, self.dim))
        self.lib.orc_v2c(self.h, int(v), _ip(cells), _dp(dirs))
        return cells, dirs


def key(x):
    x = _f64(x)
    return int(_load().orc_key(_dp(x), x.shape[0]))


class FemOracle:
    """pfem2Fem's pressure-equation and velocity-correction assembly on the
    CPU (oracle/fem_oracle.cpp) over a mesh dict (dim, X [dim][n_vertices],
    cell_vertices, cell_dofs, n_dofs). Matrices are CSR over pattern()."""

    def __init__(self, mesh):
        self.lib = _load()
        self.dim = int(mesh["dim"])
        self.X = _f64(mesh["X"])
        self.cv = np.ascontiguousarray(mesh["cell_vertices"], dtype=np.int32)
        self.cd = np.ascontiguousarray(mesh.get("cell_dofs", mesh["cell_vertices"]), dtype=np.int32)
        self.nd = int(mesh["n_dofs"])
        self.nc = self.cv.shape[0]
        self.row_ptr = np.zeros(self.nd + 1, dtype=np.int64)
        nnz = self.lib.orc_fem_pattern(self.dim, self.nc, self.nd, _ip(self.cd), self.row_ptr.ctypes.data_as(_l), None)
        self.cols = np.empty(nnz, dtype=np.int32)
        self.lib.orc_fem_pattern(self.dim, self.nc, self.nd, _ip(self.cd), self.row_ptr.ctypes.data_as(_l),
                                 _ip(self.cols))

    def pattern(self):
        return self.row_ptr, self.cols

    def pressure(self, V, rho, dt, faces=None):
        V = _f64(V)
        vals = np.empty(self.cols.shape[0])
        rhs = np.empty(self.nd)
        f = None
        if faces is not None:
            faces = np.ascontiguousarray(faces, dtype=np.uint8)
            f = faces.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
        rc = self.lib.orc_fem_pressure(self.dim, self.X.shape[1], _dp(self.X), self.nc, _ip(self.cv), _ip(self.cd),
                                       self.nd, _dp(V), f, float(rho), float(dt), self.row_ptr.ctypes.data_as(_l),
                                       _ip(self.cols), _dp(vals), _dp(rhs))
        assert rc == 0
        return vals, rhs

    def prediction(self, Vold, V, rho, mu, dt, faces=None):
        Vold, V = _f64(Vold), _f64(V)
        vals = np.empty((self.dim, self.cols.shape[0]))
        rhs = np.empty((self.dim, self.nd))
        f = None
        if faces is not None:
            faces = np.ascontiguousarray(faces, dtype=np.uint8)
            f = faces.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
        rc = self.lib.orc_fem_prediction(self.dim, self.X.shape[1], _dp(self.X), self.nc, _ip(self.cv),
                                         _ip(self.cd), self.nd, _dp(Vold), _dp(V), f, float(rho), float(mu),
                                         float(dt), self.row_ptr.ctypes.data_as(_l), _ip(self.cols), _dp(vals),
                                         _dp(rhs))
        assert rc == 0
        return vals, rhs

    def velocity_correction(self, V, P, rho, dt):
        V, P = _f64(V), _f64(P)
        vals = np.empty(self.cols.shape[0])
        rhs = np.empty((self.dim, self.nd))
        rc = self.lib.orc_fem_vcorr(self.dim, self.X.shape[1], _dp(self.X), self.nc, _ip(self.cv), _ip(self.cd),
                                    self.nd, _dp(V), _dp(P), float(rho), float(dt), self.row_ptr.ctypes.data_as(_l),
                                    _ip(self.cols), _dp(vals), _dp(rhs))
        assert rc == 0
        return vals, rhs

    def dense(self, vals):
        A = np.zeros((self.nd, self.nd))
        for r in range(self.nd):
            a, b = self.row_ptr[r], self.row_ptr[r + 1]
            A[r, self.cols[a:b]] = vals[a:b]
        return A
