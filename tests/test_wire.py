"""Particle records of the reference's MPI exchange (send_recv_particles,
src/pfem2particlehandler.cpp:505-618; record layout pfem2Particle::write_data
and its reading constructor, src/pfem2particle.cpp:17-44,213-256).

CPU: the record size against the member types, the oracle's field offsets
against hand-computed ones, the receive side's whole-record rule. GPU: the
library's records against the oracle's from the same particles, host and
device buffers, removal of the packed particles and their return, and the
argument checks."""
import struct

import numpy as np
import pytest

import pfem2dealii_amd as P
from oracle import wire_format as WF
from pfem2dealii_amd import workloads as W


@pytest.mark.parametrize("dim", [2, 3])
def test_record_size_is_the_sum_of_the_member_sizes(dim):
    # id (unsigned int), tria_position (int), cell_dofs (2^dim ints),
    # location, reference_location, velocity, velocity_ext (dim doubles each)
    want = 4 + 4 + 4 * (1 << dim) + 4 * 8 * dim
    assert want == {2: 88, 3: 136}[dim]
    assert WF.record_size(dim) == want
    assert P.particle_record_size(dim) == want


def test_record_size_of_other_dimensions_is_zero():
    assert P.particle_record_size(1) == 0 and P.particle_record_size(4) == 0


@pytest.mark.parametrize("dim", [2, 3])
def test_oracle_fields_at_the_writer_offsets(dim):
    nv = 1 << dim
    n = 3
    ids = np.array([7, 0xFFFFFFF0, 12], dtype=np.uint32)
    cells = np.array([5, 100000, 2 ** 31 - 1], dtype=np.int32)
    dofs = (np.arange(n * nv, dtype=np.int32) * 11 + 1).reshape(n, nv)
    vals = np.arange(4 * dim * n, dtype=np.float64).reshape(4, dim, n) * 0.5 + 0.25
    b = WF.pack(ids, cells, dofs, vals[0], vals[1], vals[2], vals[3])
    size = 4 + 4 + 4 * nv + 32 * dim
    assert len(b) == n * size
    for p in range(n):
        o = p * size
        assert struct.unpack_from("<I", b, o)[0] == ids[p]
        assert struct.unpack_from("<i", b, o + 4)[0] == cells[p]
        assert list(struct.unpack_from("<%di" % nv, b, o + 8)) == list(dofs[p])
        f = o + 8 + 4 * nv
        for a in range(4):  # location, reference_location, velocity, velocity_ext
            got = struct.unpack_from("<%dd" % dim, b, f + a * 8 * dim)
            assert list(got) == list(vals[a][:, p]), (a, p)
    back = WF.unpack(b, dim)
    assert np.array_equal(back["id"], ids) and np.array_equal(back["cell"], cells)
    assert np.array_equal(back["cell_dofs"], dofs)
    for a, key in enumerate(("x", "xi", "v", "v_ext")):
        assert np.array_equal(back[key], vals[a])


def test_oracle_unpack_requires_whole_records():
    b = WF.pack([1], [2], [[0, 1, 2, 3]], [[0.5], [0.5]], [[0.5], [0.5]], [[1.0], [2.0]])
    with pytest.raises(ValueError):
        WF.unpack(b[:-1], 2)


# ------------------------------------------------------------------ GPU


def _moved_handle(name, scale):
    import torch
    cfg = W.config(name, scale=scale)
    h = P.ParticleHandler(cfg["mesh"], cfg["params"])
    h.initialize_maps()
    assert h.insert_particles(cfg["x"], cfg["cells"]) == 0
    V = torch.tensor(cfg["V"], dtype=torch.float64, device="cuda")
    h.set_velocity_from_field(V)
    h.move_particles(V)  # a redistribution, so that refilled particles are present
    return cfg, h


def _dest_lists(nc, seed):
    """Three destinations of disjoint shuffled cell sets (some cells unlisted)."""
    rng = np.random.default_rng(seed)
    perm = rng.permutation(nc)
    a, b = nc // 5, nc // 5 + nc // 7
    return [perm[:a], perm[a:b], perm[b:b + nc // 3]]


def _oracle_records(cfg, d, lists):
    """The records the reference's writer produces for the listed cells, in
    destination order, particles in the handle's order within each cell."""
    mesh = cfg["mesh"]
    cs = np.concatenate([[0], np.cumsum(np.bincount(d["cell"], minlength=mesh["cell_vertices"].shape[0]))])
    cd = mesh["cell_dofs"]
    parts, sizes = [], []
    for lst in lists:
        idx = np.concatenate([np.arange(cs[c], cs[c + 1]) for c in lst]) if len(lst) else np.zeros(0, int)
        idx = idx.astype(np.int64)
        b = WF.pack(idx.astype(np.uint32), d["cell"][idx], cd[d["cell"][idx]], d["x"][:, idx], d["xi"][:, idx],
                    d["v"][:, idx])
        parts.append(b)
        sizes.append(len(b))
    return b"".join(parts), sizes


@pytest.mark.gpu
@pytest.mark.parametrize("name,scale", [("channel2d", 0.5), ("abc3d", 1.0 / 16)])
def test_pack_matches_the_oracle_records(gpu_available, name, scale):
    import torch
    print("[wire] pack vs oracle", name, scale, flush=True)
    cfg, h = _moved_handle(name, scale)
    d = h.download(xi=True)
    lists = _dest_lists(h.n_cells, 1)
    want, sizes = _oracle_records(cfg, d, lists)
    data, nbytes = h.pack_particles(lists)
    assert list(nbytes) == sizes
    assert data.tobytes() == want
    out = torch.empty(len(want) + 64, dtype=torch.uint8, device="cuda")
    _, nb2 = h.pack_particles(lists, out=out)
    assert list(nb2) == sizes
    assert out[:len(want)].cpu().numpy().tobytes() == want
    assert h.count() == d["cell"].shape[0]  # remove=False keeps them
    h.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name,scale", [("channel2d", 0.5), ("abc3d", 1.0 / 16)])
def test_remove_and_return(gpu_available, name, scale):
    import torch
    print("[wire] remove and return", name, scale, flush=True)
    cfg, h = _moved_handle(name, scale)
    d0 = h.download(xi=False)
    lists = _dest_lists(h.n_cells, 2)
    data, nbytes = h.pack_particles(lists, remove=True)
    listed = np.zeros(h.n_cells, dtype=bool)
    listed[np.concatenate(lists)] = True
    keep = ~listed[d0["cell"]]
    d1 = h.download(xi=False)
    assert h.count() == int(keep.sum())
    for k in ("x", "v"):
        assert np.array_equal(d1[k], d0[k][:, keep])
    assert np.array_equal(d1["cell"], d0["cell"][keep])
    counts = h.cell_counts()
    assert not counts[listed].any()
    # a second handle on the same mesh receives them
    h2 = P.ParticleHandler(cfg["mesh"], cfg["params"])
    h2.initialize_maps()
    assert h2.unpack_particles(torch.from_numpy(data).cuda()) == 0
    d2 = h2.download(xi=False)
    gone = ~keep
    assert h2.count() == int(gone.sum())
    assert np.array_equal(d2["cell"], d0["cell"][gone])
    for k in ("x", "v"):
        assert np.array_equal(d2[k], d0[k][:, gone])
    # and back to the first: the listed cells were emptied and refilled in
    # record order, the others kept their order, so the canonical order (part,
    # then position before the re-bin) is the original one
    assert h.unpack_particles(data) == 0
    d3 = h.download(xi=False)
    for k in ("x", "v", "cell"):
        assert np.array_equal(d3[k], d0[k]), k
    h.close()
    h2.close()


@pytest.mark.gpu
def test_wire_argument_checks(gpu_available):
    import torch
    print("[wire] argument checks", flush=True)
    cfg, h = _moved_handle("channel2d", 0.25)
    n = h.count()
    with pytest.raises(P.Pfem2GpuError):  # not a whole number of records
        h.unpack_particles(np.zeros(87, dtype=np.uint8))
    with pytest.raises(P.Pfem2GpuError):  # a cell listed twice
        h.pack_particles([[1, 2], [2]])
    small = torch.empty(8, dtype=torch.uint8, device="cuda")
    with pytest.raises(P.Pfem2GpuError):  # output too small
        h.pack_particles([np.arange(h.n_cells)], out=small)
    # a record whose location is outside its cell is rejected
    d = h.download(xi=True)
    cd = cfg["mesh"]["cell_vertices"]
    c = d["cell"][:2]
    x = d["x"][:, :2].copy()
    x[0, 1] += 10.0
    b = WF.pack([0, 1], c, cd[c], x, d["xi"][:, :2], d["v"][:, :2])
    assert h.unpack_particles(b) == 1
    assert h.count() == n + 1
    empty, nb = h.pack_particles([[], []])
    assert empty.size == 0 and list(nb) == [0, 0]
    h.close()
