"""pfem2dealii_amd — MI355X particle hot path of a PFEM-2 solver.

Python mirror of the reference's pfem2ParticleHandler<dim> interface
(src/pfem2particlehandler.h:18-71) over the C ABI in include/pfem2gpu.h.
Every method calls the HIP library libpfem2gpu.so; there is no CPU fallback:
if the library is missing, importing the handler raises.

Vector fields are [dim][n_dofs] float64 arrays: numpy arrays are host memory,
torch CUDA tensors are device memory (used in place on the handle's stream).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpfem2gpu.so")

HOST = 0
DEVICE = 1

_lib = None


class Pfem2GpuError(RuntimeError):
    pass


_vp = ctypes.c_void_p
_i64p = ctypes.POINTER(ctypes.c_int64)


class _Mesh(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int32), ("n_vertices", ctypes.c_int64), ("vertices", _vp),
                ("n_cells", ctypes.c_int64), ("cell_vertices", _vp), ("cell_dofs", _vp),
                ("n_dofs", ctypes.c_int64)]


class _Params(ctypes.Structure):
    _fields_ = [("particle_integration_steps", ctypes.c_uint32), ("time_step", ctypes.c_double),
                ("max_particles_per_cell_part", ctypes.c_uint32), ("particles_per_cell", ctypes.c_uint32 * 3)]


class Tuning(ctypes.Structure):
    """Launch shape of the move passes and the settle kernel (pfem2gpu_tuning)."""
    _fields_ = [("walk_grid", ctypes.c_int32), ("settle_grid", ctypes.c_int32), ("move_grid", ctypes.c_int32),
                ("move_block", ctypes.c_int32), ("walk_block", ctypes.c_int32)]

    def as_dict(self):
        return {"walk_grid": self.walk_grid, "settle_grid": self.settle_grid, "move_grid": self.move_grid,
                "move_block": self.move_block, "walk_block": self.walk_block}


EXPORTS = [
    "pfem2gpu_create", "pfem2gpu_destroy", "pfem2gpu_last_error", "pfem2gpu_initialize_maps",
    "pfem2gpu_seed_particles", "pfem2gpu_insert_particles", "pfem2gpu_clear_particles", "pfem2gpu_reserve_particles",
    "pfem2gpu_set_velocity_from_field", "pfem2gpu_correct_particle_velocity", "pfem2gpu_move_particles",
    "pfem2gpu_project_particle_fields", "pfem2gpu_particle_count", "pfem2gpu_particle_count_async",
    "pfem2gpu_download", "pfem2gpu_cell_counts",
    "pfem2gpu_stream", "pfem2gpu_synchronize", "pfem2gpu_set_timing", "pfem2gpu_timings",
    "pfem2gpu_batch_create", "pfem2gpu_batch_layout", "pfem2gpu_step", "pfem2gpu_set_tuning", "pfem2gpu_get_tuning",
    "pfem2gpu_adjoint", "pfem2gpu_move_particles_traced",
    "pfem2gpu_output_particle_solution", "pfem2gpu_write_particles_vtu", "pfem2gpu_write_particles_pvtu",
    "pfem2gpu_vertex_to_cells", "pfem2gpu_particle_record_size", "pfem2gpu_pack_particles",
    "pfem2gpu_unpack_particles", "pfem2gpu_pack_particles_noalloc", "pfem2gpu_remove_particles",
    "pfem2gpu_fem_sparsity", "pfem2gpu_assemble_pressure_equation", "pfem2gpu_assemble_velocity_correction",
    "pfem2gpu_assemble_velocity_prediction",
]


def load_library(path=None):
    """Loads libpfem2gpu.so (raises if it has not been built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise Pfem2GpuError("libpfem2gpu.so not built (run python -m pfem2dealii_amd.build): " + p)
    lib = ctypes.CDLL(p)
    lib.pfem2gpu_last_error.restype = ctypes.c_char_p
    lib.pfem2gpu_create.argtypes = [ctypes.POINTER(_Mesh), ctypes.POINTER(_Params), _vp, ctypes.POINTER(_vp)]
    lib.pfem2gpu_destroy.argtypes = [_vp]
    lib.pfem2gpu_initialize_maps.argtypes = [_vp]
    lib.pfem2gpu_seed_particles.argtypes = [_vp, _vp, ctypes.c_int]
    lib.pfem2gpu_insert_particles.argtypes = [_vp, ctypes.c_int64, _vp, _vp, _vp, ctypes.c_int, _i64p]
    lib.pfem2gpu_clear_particles.argtypes = [_vp]
    lib.pfem2gpu_reserve_particles.argtypes = [_vp, ctypes.c_int64]
    lib.pfem2gpu_set_velocity_from_field.argtypes = [_vp, _vp, ctypes.c_int]
    lib.pfem2gpu_correct_particle_velocity.argtypes = [_vp, _vp, _vp, ctypes.c_int]
    lib.pfem2gpu_move_particles.argtypes = [_vp, _vp, ctypes.c_int]
    lib.pfem2gpu_project_particle_fields.argtypes = [_vp, _vp, ctypes.c_int]
    lib.pfem2gpu_particle_count.argtypes = [_vp, _i64p]
    lib.pfem2gpu_particle_count_async.argtypes = [_vp, _vp]
    lib.pfem2gpu_download.argtypes = [_vp, _vp, _vp, _vp, _vp]
    lib.pfem2gpu_cell_counts.argtypes = [_vp, _vp]
    lib.pfem2gpu_stream.restype = _vp
    lib.pfem2gpu_stream.argtypes = [_vp]
    lib.pfem2gpu_synchronize.argtypes = [_vp]
    lib.pfem2gpu_set_timing.argtypes = [_vp, ctypes.c_int]
    lib.pfem2gpu_timings.argtypes = [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_char_p), _i64p,
                                     ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)]
    lib.pfem2gpu_batch_create.argtypes = [ctypes.c_int32, ctypes.POINTER(_Mesh), ctypes.POINTER(_Params), _vp,
                                          ctypes.POINTER(_vp)]
    lib.pfem2gpu_batch_layout.argtypes = [_vp, _i64p, _i64p, ctypes.POINTER(ctypes.c_int32)]
    lib.pfem2gpu_step.argtypes = [_vp, _vp, _vp, ctypes.c_int]
    lib.pfem2gpu_set_tuning.argtypes = [_vp, ctypes.POINTER(Tuning)]
    lib.pfem2gpu_get_tuning.argtypes = [_vp, ctypes.POINTER(Tuning)]
    lib.pfem2gpu_adjoint.argtypes = [_vp, ctypes.c_int32, ctypes.c_int32, _vp, _vp, _vp, _vp,
                                     ctypes.POINTER(ctypes.c_double)]
    lib.pfem2gpu_move_particles_traced.argtypes = [_vp, _vp, ctypes.c_int, _vp, ctypes.c_int64, _i64p]
    lib.pfem2gpu_output_particle_solution.argtypes = [_vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                                      ctypes.c_int32, ctypes.c_char_p]
    lib.pfem2gpu_write_particles_vtu.argtypes = [ctypes.c_char_p, ctypes.c_int32, ctypes.c_int64, _vp, _vp,
                                                 ctypes.c_int32]
    lib.pfem2gpu_write_particles_pvtu.argtypes = [ctypes.c_char_p, ctypes.c_int32, ctypes.c_int32]
    lib.pfem2gpu_vertex_to_cells.argtypes = [_vp, _vp, _vp, _vp]
    lib.pfem2gpu_particle_record_size.argtypes = [ctypes.c_int32]
    lib.pfem2gpu_particle_record_size.restype = ctypes.c_int64
    lib.pfem2gpu_pack_particles.argtypes = [_vp, ctypes.c_int32, _vp, _vp, _vp, ctypes.c_int64, ctypes.c_int, _vp,
                                            ctypes.c_int]
    lib.pfem2gpu_unpack_particles.argtypes = [_vp, _vp, ctypes.c_int64, ctypes.c_int, _i64p]
    lib.pfem2gpu_remove_particles.argtypes = [_vp, ctypes.c_int64, _vp]
    lib.pfem2gpu_pack_particles_noalloc.argtypes = [_vp, ctypes.c_int32, _vp, _vp, _vp, ctypes.c_int64, _vp,
                                                    ctypes.c_int]
    lib.pfem2gpu_fem_sparsity.argtypes = [_vp, _i64p, _vp, _vp]
    lib.pfem2gpu_assemble_pressure_equation.argtypes = [_vp, _vp, _vp, ctypes.c_double, ctypes.c_double, _vp, _vp,
                                                        ctypes.c_int]
    lib.pfem2gpu_assemble_velocity_correction.argtypes = [_vp, _vp, _vp, ctypes.c_double, ctypes.c_double, _vp,
                                                          _vp, ctypes.c_int]
    lib.pfem2gpu_assemble_velocity_prediction.argtypes = [_vp, _vp, _vp, _vp, ctypes.c_double, ctypes.c_double,
                                                          ctypes.c_double, _vp, _vp, ctypes.c_int]
    if path is None:
        _lib = lib
    return lib


def _check(lib, rc, what):
    if rc != 0:
        msg = lib.pfem2gpu_last_error().decode(errors="replace")
        raise Pfem2GpuError("%s failed (code %d): %s" % (what, rc, msg))


def write_particles_vtu(path, x, v, subdomain=0):
    """One rank's particle file from host arrays x, v ([dim][n]); the text of
    src/pfem2particlehandler.cpp:219-294."""
    lib = load_library()
    x = np.ascontiguousarray(x, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    if x.ndim != 2 or v.shape != x.shape:
        raise ValueError("x and v must be [dim][n] arrays of the same shape")
    _check(lib, lib.pfem2gpu_write_particles_vtu(os.fsencode(path), x.shape[0], x.shape[1], x.ctypes.data,
                                                 v.ctypes.data, int(subdomain)), "write_particles_vtu")


def write_particles_pvtu(path, timestep_number, n_ranks=1):
    """The index file of n_ranks pieces (src/pfem2particlehandler.cpp:296-316)."""
    lib = load_library()
    _check(lib, lib.pfem2gpu_write_particles_pvtu(os.fsencode(path), int(timestep_number), int(n_ranks)),
           "write_particles_pvtu")


def particle_record_size(dim):
    """Bytes of one particle record of the reference's MPI exchange
    (pfem2Particle::serialized_size_in_bytes, src/pfem2particle.cpp:213-219)."""
    return int(load_library().pfem2gpu_particle_record_size(int(dim)))


def _is_device(a):
    return hasattr(a, "is_cuda") and a.is_cuda


class _FieldPtrs:
    """Array of `dim` component pointers for a [dim][n] field, plus keep-alive."""

    def __init__(self, F, dim, n, writable=False):
        if _is_device(F):
            import torch
            if F.dtype != torch.float64 or tuple(F.shape) != (dim, n):
                raise ValueError("device field must be float64 [%d, %d]" % (dim, n))
            if not F.is_contiguous():
                raise ValueError("device field must be contiguous")
            self.keep = F
            base = F.data_ptr()
            self.where = DEVICE
        else:
            if writable:
                if not (isinstance(F, np.ndarray) and F.dtype == np.float64 and F.flags.c_contiguous
                        and F.shape == (dim, n)):
                    raise ValueError("host output must be a C-contiguous float64 [%d, %d] array" % (dim, n))
                A = F
            else:
                A = np.ascontiguousarray(F, dtype=np.float64)
                if A.shape != (dim, n):
                    raise ValueError("field must be [%d, %d]" % (dim, n))
            self.keep = A
            base = A.ctypes.data
            self.where = HOST
        self.arr = (_vp * 3)(*[base + 8 * n * k for k in range(dim)] + [None] * (3 - dim))


def _pack_mesh(mesh, keep):
    X = np.ascontiguousarray(mesh["X"], dtype=np.float64)
    cv = np.ascontiguousarray(mesh["cell_vertices"], dtype=np.int32)
    cd = np.ascontiguousarray(mesh.get("cell_dofs", cv), dtype=np.int32)
    keep.extend([X, cv, cd])
    return _Mesh(int(mesh["dim"]), X.shape[1], X.ctypes.data, cv.shape[0], cv.ctypes.data, cd.ctypes.data,
                 int(mesh["n_dofs"]))


def _pack_params(params):
    q = list(params["particles_per_cell"]) + [1] * (3 - len(params["particles_per_cell"]))
    return _Params(int(params["n_sub"]), float(params["dt"]), int(params["max_per_part"]), (ctypes.c_uint32 * 3)(*q))


class ParticleHandler:
    """pfem2ParticleHandler<dim> on one MI355X (src/pfem2particlehandler.h:18-71)."""

    def __init__(self, mesh, params, stream=None):
        self.lib = load_library()
        self.dim = int(mesh["dim"])
        self.n_dofs = int(mesh["n_dofs"])
        self.n_cells = int(np.asarray(mesh["cell_vertices"]).shape[0])
        self.n_vertices = int(np.asarray(mesh["X"]).shape[1])
        keep = []
        m = _pack_mesh(mesh, keep)
        p = _pack_params(params)
        h = _vp()
        _check(self.lib, self.lib.pfem2gpu_create(ctypes.byref(m), ctypes.byref(p), stream, ctypes.byref(h)),
               "pfem2gpu_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.pfem2gpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- the reference interface ----
    def initialize_maps(self):
        _check(self.lib, self.lib.pfem2gpu_initialize_maps(self.h), "initialize_maps")

    def seed_particles(self, V):
        f = _FieldPtrs(V, self.dim, self.n_dofs)
        _check(self.lib, self.lib.pfem2gpu_seed_particles(self.h, f.arr, f.where), "seed_particles")

    def insert_particles(self, x, cell, v=None):
        """Inserts particles at x ([dim][n]) in `cell`; returns the number rejected."""
        if _is_device(x):
            where = DEVICE
            n = int(cell.shape[0])
            keep = (x, cell, v)
            xp, cp = x.data_ptr(), cell.data_ptr()
            vp = v.data_ptr() if v is not None else None
        else:
            where = HOST
            xa = np.ascontiguousarray(x, dtype=np.float64)
            ca = np.ascontiguousarray(cell, dtype=np.int32)
            va = np.ascontiguousarray(v, dtype=np.float64) if v is not None else None
            n = int(ca.shape[0])
            if xa.shape != (self.dim, n) or (va is not None and va.shape != (self.dim, n)):
                raise ValueError("x and v must be [dim][n]")
            keep = (xa, ca, va)
            xp, cp = xa.ctypes.data, ca.ctypes.data
            vp = va.ctypes.data if va is not None else None
        rej = ctypes.c_int64(0)
        _check(self.lib, self.lib.pfem2gpu_insert_particles(self.h, n, xp, cp, vp, where, ctypes.byref(rej)),
               "insert_particles")
        del keep
        return int(rej.value)

    def remove_particles(self, index):
        """remove_particle for the particles at the given positions of the
        handle's order (src/pfem2particlehandler.cpp:332-336); the others keep
        their order."""
        idx = np.ascontiguousarray(index, dtype=np.int64).ravel()
        _check(self.lib, self.lib.pfem2gpu_remove_particles(self.h, idx.size, idx.ctypes.data if idx.size else None),
               "remove_particles")

    def clear_particles(self):
        _check(self.lib, self.lib.pfem2gpu_clear_particles(self.h), "clear_particles")

    def reserve_particles(self, n):
        """Device capacity for n particles (moves of up to n allocate nothing)."""
        _check(self.lib, self.lib.pfem2gpu_reserve_particles(self.h, int(n)), "reserve_particles")

    def set_velocity_from_field(self, V):
        f = _FieldPtrs(V, self.dim, self.n_dofs)
        _check(self.lib, self.lib.pfem2gpu_set_velocity_from_field(self.h, f.arr, f.where),
               "set_velocity_from_field")

    def correct_particle_velocity(self, V, V_old):
        a = _FieldPtrs(V, self.dim, self.n_dofs)
        b = _FieldPtrs(V_old, self.dim, self.n_dofs)
        if a.where != b.where:
            raise ValueError("V and V_old must both be host or both be device arrays")
        _check(self.lib, self.lib.pfem2gpu_correct_particle_velocity(self.h, a.arr, b.arr, a.where),
               "correct_particle_velocity")

    def move_particles(self, V):
        f = _FieldPtrs(V, self.dim, self.n_dofs)
        _check(self.lib, self.lib.pfem2gpu_move_particles(self.h, f.arr, f.where), "move_particles")

    def project_particle_fields(self, out=None):
        """Projected nodal velocity; into `out` (numpy or torch CUDA) if given."""
        if out is None:
            out = np.empty((self.dim, self.n_dofs), dtype=np.float64)
        f = _FieldPtrs(out, self.dim, self.n_dofs, writable=True)
        _check(self.lib, self.lib.pfem2gpu_project_particle_fields(self.h, f.arr, f.where),
               "project_particle_fields")
        return out

    # ---- FEM assembly (pfem2Fem, src/pfem2fem.cpp) ----

    def fem_sparsity(self):
        """CSR pattern of the DoFs that share a cell (src/pfem2fem.cpp:78-95):
        (row_ptr int64 [n_dofs + 1], cols int32 [nnz], ascending per row)."""
        nnz = ctypes.c_int64()
        _check(self.lib, self.lib.pfem2gpu_fem_sparsity(self.h, ctypes.byref(nnz), None, None), "fem_sparsity")
        rp = np.empty(self.n_dofs + 1, dtype=np.int64)
        cols = np.empty(nnz.value, dtype=np.int32)
        _check(self.lib, self.lib.pfem2gpu_fem_sparsity(self.h, None, rp.ctypes.data, cols.ctypes.data),
               "fem_sparsity")
        self._fem_nnz = nnz.value
        return rp, cols

    def _fem_out(self, dev, shape):
        if dev is not None:
            import torch
            return torch.empty(shape, dtype=torch.float64, device=dev)
        return np.empty(shape, dtype=np.float64)

    def _fem_call(self, dev, fn, *args):
        """Device arrays: the library reads and writes them on the handle's
        stream, so torch's stream is drained first and the handle's stream
        after (the results are complete when this returns)."""
        if dev is not None:
            import torch
            torch.cuda.current_stream(dev).synchronize()
        rc = fn(*args)
        if dev is not None and rc == 0:
            rc = self.lib.pfem2gpu_synchronize(self.h)
        return rc

    @staticmethod
    def _ptr(a):
        if a is None:
            return None
        return a.data_ptr() if _is_device(a) else a.ctypes.data

    def assemble_pressure_equation(self, V_pred, rho, time_step, boundary_faces=None, matrix=True):
        """assemble_pressure_equation (src/pfem2fem.cpp:392-458): returns
        (values [nnz] over fem_sparsity() or None, rhs [n_dofs]), on the device
        (torch) when V_pred is a device tensor. boundary_faces: [n_cells] uint8,
        bit f for face f at a boundary without a pressure Dirichlet condition."""
        if getattr(self, "_fem_nnz", None) is None:
            self.fem_sparsity()
        f = _FieldPtrs(V_pred, self.dim, self.n_dofs)
        dev = V_pred.device if f.where == DEVICE else None
        faces = None
        if boundary_faces is not None:
            if dev is not None:
                import torch
                faces = torch.as_tensor(np.asarray(boundary_faces.cpu() if _is_device(boundary_faces)
                                                   else boundary_faces, dtype=np.uint8)).to(dev)
            else:
                faces = np.ascontiguousarray(boundary_faces, dtype=np.uint8)
            if tuple(faces.shape) != (self.n_cells,):
                raise ValueError("boundary_faces must have one entry per cell")
        vals = self._fem_out(dev, (self._fem_nnz,)) if matrix else None
        rhs = self._fem_out(dev, (self.n_dofs,))
        _check(self.lib, self._fem_call(dev, self.lib.pfem2gpu_assemble_pressure_equation, self.h, f.arr,
                                        self._ptr(faces), float(rho), float(time_step), self._ptr(vals),
                                        self._ptr(rhs), f.where), "assemble_pressure_equation")
        return vals, rhs

    def _faces_arg(self, boundary_faces, dev):
        if boundary_faces is None:
            return None
        if dev is not None:
            import torch
            f = torch.as_tensor(np.asarray(boundary_faces.cpu() if _is_device(boundary_faces) else boundary_faces,
                                           dtype=np.uint8)).to(dev)
        else:
            f = np.ascontiguousarray(boundary_faces, dtype=np.uint8)
        if tuple(f.shape) != (self.n_cells,):
            raise ValueError("boundary_faces must have one entry per cell")
        return f

    def assemble_velocity_prediction(self, V_old, V, rho, mu, time_step, boundary_faces=None, matrix=True):
        """assemble_velocity_prediction (src/pfem2fem.cpp:246-390): returns
        (values [dim, nnz] over fem_sparsity() or None, rhs [dim, n_dofs]),
        one system per component, on the device (torch) when V_old is a
        device tensor. boundary_faces: [n_cells] uint8, bit f for face f at a
        boundary without a velocity Dirichlet condition."""
        if getattr(self, "_fem_nnz", None) is None:
            self.fem_sparsity()
        a = _FieldPtrs(V_old, self.dim, self.n_dofs)
        b = _FieldPtrs(V, self.dim, self.n_dofs)
        if a.where != b.where:
            raise ValueError("V_old and V must both be host or both be device arrays")
        dev = V_old.device if a.where == DEVICE else None
        faces = self._faces_arg(boundary_faces, dev)
        vals = self._fem_out(dev, (self.dim, self._fem_nnz)) if matrix else None
        rhs = self._fem_out(dev, (self.dim, self.n_dofs))
        mp = _FieldPtrs(vals, self.dim, self._fem_nnz, writable=True) if matrix else None
        r = _FieldPtrs(rhs, self.dim, self.n_dofs, writable=True)
        _check(self.lib, self._fem_call(dev, self.lib.pfem2gpu_assemble_velocity_prediction, self.h, a.arr, b.arr,
                                        self._ptr(faces), float(rho), float(mu), float(time_step),
                                        mp.arr if mp else None, r.arr, a.where), "assemble_velocity_prediction")
        return vals, rhs

    def assemble_velocity_correction(self, V_pred, P, rho, time_step, matrix=True):
        """assemble_velocity_correction (src/pfem2fem.cpp:460-517): returns
        (values [nnz] or None, rhs [dim, n_dofs]); the matrix is that of every
        component."""
        if getattr(self, "_fem_nnz", None) is None:
            self.fem_sparsity()
        f = _FieldPtrs(V_pred, self.dim, self.n_dofs)
        dev = V_pred.device if f.where == DEVICE else None
        if dev is not None:
            if not (_is_device(P) and P.is_contiguous() and tuple(P.shape) == (self.n_dofs,)
                    and str(P.dtype) == "torch.float64"):
                raise ValueError("P must be a contiguous float64 device tensor [n_dofs] with a device V_pred")
            Pk = P
        else:
            Pk = np.ascontiguousarray(P, dtype=np.float64)
            if Pk.shape != (self.n_dofs,):
                raise ValueError("P must be [n_dofs]")
        vals = self._fem_out(dev, (self._fem_nnz,)) if matrix else None
        rhs = self._fem_out(dev, (self.dim, self.n_dofs))
        r = _FieldPtrs(rhs, self.dim, self.n_dofs, writable=True)
        _check(self.lib, self._fem_call(dev, self.lib.pfem2gpu_assemble_velocity_correction, self.h, f.arr,
                                        self._ptr(Pk), float(rho), float(time_step), self._ptr(vals), r.arr,
                                        f.where), "assemble_velocity_correction")
        return vals, rhs

    # ---- readback ----
    def count(self):
        n = ctypes.c_int64(0)
        _check(self.lib, self.lib.pfem2gpu_particle_count(self.h, ctypes.byref(n)), "particle_count")
        return int(n.value)

    def count_async(self, dst_ptr):
        """Enqueues a copy of the particle count into device int32 at dst_ptr."""
        _check(self.lib, self.lib.pfem2gpu_particle_count_async(self.h, dst_ptr), "particle_count_async")

    def download(self, xi=True):
        n = self.count()
        x = np.empty((self.dim, n))
        v = np.empty((self.dim, n))
        r = np.empty((self.dim, n)) if xi else None
        cell = np.empty(n, dtype=np.int32)
        _check(self.lib, self.lib.pfem2gpu_download(self.h, x.ctypes.data, r.ctypes.data if xi else None,
                                                    v.ctypes.data, cell.ctypes.data), "download")
        out = {"x": x, "v": v, "cell": cell}
        if xi:
            out["xi"] = r
        return out

    def output_particle_solution(self, timestep_number, directory=".", rank=0, n_ranks=1, subdomain=0):
        """Writes particles-TT.RRR.vtu (and particles-TT.pvtu on rank 0) into
        `directory`, in the text format of src/pfem2particlehandler.cpp:217-317."""
        _check(self.lib, self.lib.pfem2gpu_output_particle_solution(
            self.h, int(timestep_number), int(rank), int(n_ranks), int(subdomain), os.fsencode(directory)),
            "output_particle_solution")

    def vertex_to_cells(self):
        """(offsets [n_vertices + 1], cells, dirs [n, dim]) of the map built by
        initialize_maps (src/pfem2particlehandler.cpp:319-324)."""
        off = np.empty(self.n_vertices + 1, dtype=np.int32)
        _check(self.lib, self.lib.pfem2gpu_vertex_to_cells(self.h, off.ctypes.data, None, None), "vertex_to_cells")
        cells = np.empty(int(off[-1]), dtype=np.int32)
        dirs = np.empty((int(off[-1]), self.dim))
        _check(self.lib, self.lib.pfem2gpu_vertex_to_cells(self.h, off.ctypes.data, cells.ctypes.data,
                                                           dirs.ctypes.data), "vertex_to_cells")
        return off, cells, dirs

    def pack_particles(self, cells_per_dest, remove=False, out=None):
        """Serialises the particles of the listed cells into records of the
        reference's wire format (the send side of send_recv_particles,
        src/pfem2particlehandler.cpp:528-548): cells_per_dest[i] lists the
        cells of destination i. Returns (data, bytes_per_dest); data is a new
        host uint8 array, or `out` (a torch CUDA uint8 tensor) when given.
        remove=True takes the packed particles out of the handle."""
        lists = [np.ascontiguousarray(c, dtype=np.int32).ravel() for c in cells_per_dest]
        off = np.zeros(len(lists) + 1, dtype=np.int32)
        off[1:] = np.cumsum([len(c) for c in lists])
        cells = np.concatenate(lists) if lists else np.zeros(0, dtype=np.int32)
        rec = particle_record_size(self.dim)
        if cells.size and (cells.min() < 0 or cells.max() >= self.n_cells):
            raise ValueError("cell id out of range")
        need = int(self.cell_counts()[cells].astype(np.int64).sum()) * rec if cells.size else 0
        nbytes = np.zeros(max(len(lists), 1), dtype=np.int64)
        if out is None:
            data = np.empty(need, dtype=np.uint8)
            ptr, cap, where = (data.ctypes.data if need else None), need, HOST
        else:
            if not _is_device(out):
                raise ValueError("out must be a torch CUDA tensor")
            data = out
            ptr, cap, where = out.data_ptr(), out.numel() * out.element_size(), DEVICE
        _check(self.lib, self.lib.pfem2gpu_pack_particles(self.h, len(lists), off.ctypes.data,
                                                          cells.ctypes.data if cells.size else None, ptr, cap,
                                                          where, nbytes.ctypes.data, 1 if remove else 0),
               "pack_particles")
        return data, nbytes[:len(lists)].copy()

    def unpack_particles(self, data):
        """Appends the particles of a buffer of records (the receive side,
        src/pfem2particlehandler.cpp:607-615): a host bytes/uint8 array or a
        torch CUDA uint8 tensor. Returns the number of rejected records
        (location not inside the record's cell)."""
        rej = ctypes.c_int64(0)
        if _is_device(data):
            n = data.numel() * data.element_size()
            _check(self.lib, self.lib.pfem2gpu_unpack_particles(self.h, data.data_ptr(), n, DEVICE,
                                                                ctypes.byref(rej)), "unpack_particles")
        else:
            buf = np.ascontiguousarray(np.frombuffer(data, dtype=np.uint8) if isinstance(data, (bytes, bytearray))
                                       else data, dtype=np.uint8)
            _check(self.lib, self.lib.pfem2gpu_unpack_particles(self.h, buf.ctypes.data if buf.size else None,
                                                                buf.size, HOST, ctypes.byref(rej)),
                   "unpack_particles")
        return rej.value

    def cell_counts(self):
        c = np.empty(self.n_cells, dtype=np.int32)
        _check(self.lib, self.lib.pfem2gpu_cell_counts(self.h, c.ctypes.data), "cell_counts")
        return c

    def stream(self):
        return self.lib.pfem2gpu_stream(self.h)

    def step(self, V, V_proj, use_graph=True):
        """move -> project(V_proj) -> correct(V, V_proj) with device arrays; a
        hipGraph of the sequence is captured once and replayed."""
        a = _FieldPtrs(V, self.dim, self.n_dofs)
        b = _FieldPtrs(V_proj, self.dim, self.n_dofs, writable=True)
        if a.where != DEVICE or b.where != DEVICE:
            raise ValueError("step() needs device (torch CUDA) fields")
        _check(self.lib, self.lib.pfem2gpu_step(self.h, a.arr, b.arr, 1 if use_graph else 0), "step")

    def move_particles_traced(self, V):
        """move_particles(V) that returns, per particle after the move, its
        index before the move (or -1 - part for refills)."""
        f = _FieldPtrs(V, self.dim, self.n_dofs)
        cap = self.count() + self.n_cells * 128 + 1024
        src = np.empty(cap, dtype=np.int32)
        n = ctypes.c_int64(0)
        _check(self.lib, self.lib.pfem2gpu_move_particles_traced(self.h, f.arr, f.where, src.ctypes.data, cap,
                                                                 ctypes.byref(n)), "move_particles_traced")
        return src[:n.value].copy()

    def adjoint(self, n_steps, checkpoint_every, V, target):
        """(loss, dL/dv0 [dim][n0], dL/dV [dim][n_dofs]) of n_steps of move and
        a final projection against `target` (device tensors)."""
        import torch
        n0 = self.count()
        a = _FieldPtrs(V, self.dim, self.n_dofs)
        b = _FieldPtrs(target, self.dim, self.n_dofs)
        if a.where != DEVICE or b.where != DEVICE:
            raise ValueError("adjoint() needs device (torch CUDA) fields")
        gv0 = torch.zeros((self.dim, max(n0, 1)), dtype=torch.float64, device=V.device)
        gV = torch.zeros((self.dim, self.n_dofs), dtype=torch.float64, device=V.device)
        pv = (_vp * 3)(*[gv0.data_ptr() + 8 * gv0.shape[1] * k for k in range(self.dim)] + [None] * (3 - self.dim))
        pV = (_vp * 3)(*[gV.data_ptr() + 8 * self.n_dofs * k for k in range(self.dim)] + [None] * (3 - self.dim))
        loss = ctypes.c_double(0.0)
        torch.cuda.synchronize()
        _check(self.lib, self.lib.pfem2gpu_adjoint(self.h, int(n_steps), int(checkpoint_every), a.arr, b.arr, pv, pV,
                                                   ctypes.byref(loss)), "adjoint")
        self.synchronize()
        return loss.value, gv0[:, :n0], gV

    def set_tuning(self, walk_grid=4096, settle_grid=2048, move_grid=0, move_block=256, walk_block=256):
        t = Tuning(int(walk_grid), int(settle_grid), int(move_grid), int(move_block), int(walk_block))
        _check(self.lib, self.lib.pfem2gpu_set_tuning(self.h, ctypes.byref(t)), "set_tuning")

    def get_tuning(self):
        t = Tuning()
        _check(self.lib, self.lib.pfem2gpu_get_tuning(self.h, ctypes.byref(t)), "get_tuning")
        return t.as_dict()

    def synchronize(self):
        _check(self.lib, self.lib.pfem2gpu_synchronize(self.h), "synchronize")

    def set_timing(self, enable):
        _check(self.lib, self.lib.pfem2gpu_set_timing(self.h, 1 if enable else 0), "set_timing")

    def timings(self):
        """{kernel name: (launches, total ms)} since set_timing(True)."""
        cap = 64
        names = (ctypes.c_char_p * cap)()
        launches = (ctypes.c_int64 * cap)()
        ms = (ctypes.c_double * cap)()
        n = ctypes.c_int(0)
        _check(self.lib, self.lib.pfem2gpu_timings(self.h, cap, names, launches, ms, ctypes.byref(n)), "timings")
        return {names[i].decode(): (int(launches[i]), float(ms[i])) for i in range(min(n.value, cap))}


class BatchHandler(ParticleHandler):
    """Many independent instances in one handle (pfem2gpu_batch_create): one
    kernel launch per phase for the whole batch. Cell ids and DoF-indexed
    fields are batch-global; see cell_offsets / dof_offsets."""

    def __init__(self, meshes, params, stream=None):
        self.lib = load_library()
        n = len(meshes)
        if n == 0 or len(params) != n:
            raise ValueError("need one params dict per mesh")
        self.dim = int(meshes[0]["dim"])
        keep = []
        marr = (_Mesh * n)(*[_pack_mesh(m, keep) for m in meshes])
        parr = (_Params * n)(*[_pack_params(p) for p in params])
        h = _vp()
        _check(self.lib, self.lib.pfem2gpu_batch_create(n, marr, parr, stream, ctypes.byref(h)), "batch_create")
        self.h = h
        co = np.zeros(n + 1, dtype=np.int64)
        do = np.zeros(n + 1, dtype=np.int64)
        ni = ctypes.c_int32(0)
        _check(self.lib, self.lib.pfem2gpu_batch_layout(self.h, co.ctypes.data_as(_i64p), do.ctypes.data_as(_i64p),
                                                        ctypes.byref(ni)), "batch_layout")
        self.cell_offsets = co
        self.dof_offsets = do
        self.n_cells = int(co[-1])
        self.n_dofs = int(do[-1])
        self.n_vertices = int(sum(np.asarray(m["X"]).shape[1] for m in meshes))
