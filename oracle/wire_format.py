"""CPU restatement of the reference's particle wire format. Test infrastructure.

send_recv_particles (src/pfem2particlehandler.cpp:505-618) ships particles
between MPI ranks as records that pfem2Particle::write_data writes
(src/pfem2particle.cpp:221-256) and the constructor from a data pointer reads
back (:17-44). The writer advances one pointer over the fields in this order,
with nothing between them:
  id (unsigned int, 4 bytes), tria_position (int, 4),
  cell_dofs (GeometryInfo<dim>::vertices_per_cell ints: 4 in 2D, 8 in 3D),
  location, reference_location, velocity, velocity_ext (dim doubles each).
Its size is serialized_size_in_bytes (:213-219), the sum of the member sizes:
4 + 4 + 4*2^dim + 4*8*dim = 88 (2D) / 136 (3D). Records of one destination
are written back to back, destinations in ascending neighbour order
(src/pfem2particlehandler.cpp:535-548); the receiver parses records until the
buffer ends and requires that it ends on a record boundary (:607-615).
Integers and doubles are in the host's (little-endian) byte order.

Pinning: the reference holds no golden bytes for this format (it has no
tests). The record size is pinned by the member types; field order and
offsets are the writer's sequence above, checked in the tests by decoding
fields at hand-computed offsets of records whose fields all differ, so that a
swapped or shifted field fails.
"""
import numpy as np


def record_dtype(dim):
    """numpy structured dtype of one record (packed, little-endian)."""
    nv = 1 << dim
    return np.dtype([("id", "<u4"), ("tria_position", "<i4"), ("cell_dofs", "<i4", (nv,)),
                     ("location", "<f8", (dim,)), ("reference_location", "<f8", (dim,)),
                     ("velocity", "<f8", (dim,)), ("velocity_ext", "<f8", (dim,))], align=False)


def record_size(dim):
    return record_dtype(dim).itemsize


def pack(ids, cells, cell_dofs, x, xi, v, v_ext=None):
    """Records of n particles: ids [n], cells [n] (tria_position), cell_dofs
    [n][2^dim], x, xi, v, v_ext as [dim][n] (v_ext None = zeros). Returns
    bytes."""
    x = np.asarray(x, dtype=np.float64)
    dim, n = x.shape
    r = np.zeros(n, dtype=record_dtype(dim))
    r["id"] = np.asarray(ids, dtype=np.uint32)
    r["tria_position"] = np.asarray(cells, dtype=np.int32)
    r["cell_dofs"] = np.asarray(cell_dofs, dtype=np.int32).reshape(n, 1 << dim)
    r["location"] = x.T
    r["reference_location"] = np.asarray(xi, dtype=np.float64).T
    r["velocity"] = np.asarray(v, dtype=np.float64).T
    if v_ext is not None:
        r["velocity_ext"] = np.asarray(v_ext, dtype=np.float64).T
    return r.tobytes()


def unpack(data, dim):
    """Fields of a buffer of records as a dict of arrays ([dim][n] for the
    vectors). Raises if the buffer does not end on a record boundary."""
    b = np.frombuffer(bytes(data), dtype=np.uint8)
    size = record_size(dim)
    if b.size % size:
        raise ValueError("%d bytes is not a whole number of %d-byte records" % (b.size, size))
    r = b.view(record_dtype(dim))
    return {"id": r["id"].copy(), "cell": r["tria_position"].copy(), "cell_dofs": r["cell_dofs"].copy(),
            "x": r["location"].T.copy(), "xi": r["reference_location"].T.copy(), "v": r["velocity"].T.copy(),
            "v_ext": r["velocity_ext"].T.copy()}
