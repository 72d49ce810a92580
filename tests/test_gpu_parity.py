"""GPU vs oracle parity through the C ABI (ctypes). Marked gpu.

Each test prints the case name before touching the GPU so a failing run names
its case."""
import sys

import numpy as np
import pytest

from pfem2dealii_amd import workloads as W
from tests.harness import (GpuStepper, OracleStepper, assert_fields_match, assert_particles_match,
                           perturbed_mesh, seed_in_cells)

pytestmark = pytest.mark.gpu


def _say(msg):
    print("[case] " + msg, flush=True)
    sys.stdout.flush()


def _run_pair(cfg, steps, mesh=None, x=None, cells=None):
    mesh = mesh or cfg["mesh"]
    x = cfg["x"] if x is None else x
    cells = cfg["cells"] if cells is None else cells
    V = cfg["V"]
    o = OracleStepper(cfg, mesh=mesh)
    g = GpuStepper(cfg, mesh=mesh)
    assert o.insert(x, cells, V) == g.insert(x, cells, V)
    assert_particles_match(g.download(), o.download())
    for s in range(steps):
        Vo = o.step(V)
        Vg = g.step(V)
        assert_fields_match(Vg, Vo)
        # canonical order makes the projection bitwise identical
        assert np.array_equal(np.isnan(Vg), np.isnan(Vo))
        assert np.array_equal(np.nan_to_num(Vg, nan=7.0), np.nan_to_num(Vo, nan=7.0)), "projection not bitwise"
        assert_particles_match(g.download(), o.download())
    return o, g


def test_channel2d_one_step(gpu_available):
    _say("channel2d 64x32 jittered lattice, 1 step")
    cfg = W.config("channel2d")
    o, g = _run_pair(cfg, 1)
    assert np.array_equal(g.h.cell_counts(), np.bincount(o.download()["cell"], minlength=2048))


def test_channel2d_three_steps(gpu_available):
    _say("channel2d 64x32, 3 steps")
    _run_pair(W.config("channel2d"), 3)


def test_taylor_green_small(gpu_available):
    _say("taylor_green2d scale 1/16 (128x64), 3 steps")
    _run_pair(W.config("taylor_green2d", scale=1.0 / 16), 3)


def test_abc3d_small(gpu_available):
    _say("abc3d scale 1/8 (16^3), 3 steps")
    _run_pair(W.config("abc3d", scale=1.0 / 8), 3)


def test_sixteen_substeps(gpu_available):
    """n_sub = 16: walks at late substep indices, 16 rounds of walk queues."""
    _say("channel 16x8, 16 substeps, 2 steps")
    cfg = W.config("channel2d", scale=0.25)
    cfg = dict(cfg, params=dict(cfg["params"], n_sub=16))
    _run_pair(cfg, 2)


def test_perturbed_2d_newton(gpu_available):
    _say("perturbed 2D mesh 24x12 (Newton inverse map), 3 steps")
    cfg = W.config("channel2d", scale=0.375)
    mesh = perturbed_mesh(cfg["mesh"], 0.2, seed=11)
    x, cells = seed_in_cells(mesh, 9, seed=12)
    _run_pair(cfg, 3, mesh=mesh, x=x, cells=cells)


def test_perturbed_3d_newton(gpu_available):
    _say("perturbed 3D mesh 8^3 (Newton inverse map), 2 steps")
    cfg = W.config("abc3d", scale=1.0 / 16)
    mesh = perturbed_mesh(cfg["mesh"], 0.15, seed=13)
    x, cells = seed_in_cells(mesh, 6, seed=14)
    _run_pair(cfg, 2, mesh=mesh, x=x, cells=cells)


def _star_mesh(n):
    """Five non-affine quads around a centre vertex (centre, spoke k, spoke k+1,
    outer corner), each split into n x n sub-quads: the centre vertex touches
    five cells, so the 2D walk takes its general path there."""
    K = 5
    th = [2.0 * np.pi * k / K for k in range(K)]
    C = np.zeros(2)
    S = [np.array([np.cos(t), np.sin(t)]) for t in th]
    M = [1.25 * np.array([np.cos(t + np.pi / K), np.sin(t + np.pi / K)]) for t in th]
    verts, index, cells = [], {}, []

    def vid(p):
        key = (round(p[0], 9), round(p[1], 9))
        if key not in index:
            index[key] = len(verts)
            verts.append(p)
        return index[key]

    for k in range(K):
        corners = [C, S[k], S[(k + 1) % K], M[k]]
        lat = {}
        for a in range(n + 1):
            for b in range(n + 1):
                u, w = a / n, b / n
                N = [(1 - u) * (1 - w), u * (1 - w), (1 - u) * w, u * w]
                p = sum(Ni * ci for Ni, ci in zip(N, corners))
                lat[a, b] = vid(p)
        for b in range(n):
            for a in range(n):
                cells.append([lat[a, b], lat[a + 1, b], lat[a, b + 1], lat[a + 1, b + 1]])
    X = np.array(verts).T.copy()
    cv = np.array(cells, dtype=np.int32)
    return {"dim": 2, "n": [n, n], "X": X, "cell_vertices": cv, "cell_dofs": cv, "n_dofs": X.shape[1]}


def test_vertex_of_five_cells(gpu_available):
    _say("star mesh: five non-affine quads around a vertex, 8x8 each, rotation, 3 steps")
    mesh = _star_mesh(8)
    X = mesh["X"]
    V = np.stack([-X[1], X[0]])
    cfg = {"mesh": mesh, "V": V,
           "params": dict(n_sub=5, dt=0.05, max_per_part=5, particles_per_cell=[3, 3])}
    centre = int(np.argmin(np.hypot(X[0], X[1])))
    assert (mesh["cell_vertices"] == centre).sum() == 5
    x, cells = seed_in_cells(mesh, 9, seed=21)
    o, g = _run_pair(cfg, 3, mesh=mesh, x=x, cells=cells)
    # particles crossed into and out of the five cells around the centre
    around = np.nonzero((mesh["cell_vertices"] == centre).any(axis=1))[0]
    assert np.isin(g.download()["cell"], around).sum() > 0


def test_seed_lattice(gpu_available):
    _say("seed_particles lattice 2D and 3D")
    for name, scale in (("channel2d", 0.5), ("abc3d", 1.0 / 16)):
        cfg = W.config(name, scale=scale)
        from oracle import Oracle
        from pfem2dealii_amd import ParticleHandler
        o = Oracle(cfg["mesh"], cfg["params"])
        o.initialize_maps()
        o.seed_lattice(cfg["V"])
        h = ParticleHandler(cfg["mesh"], cfg["params"])
        h.initialize_maps()
        h.seed_particles(cfg["V"])
        assert_particles_match(h.download(), o.download())
        for _ in range(2):
            o.move(cfg["V"])
            h.move_particles(cfg["V"])
            assert_particles_match(h.download(), o.download())


def _placed_meshes():
    """Box meshes whose part centres do not map back exactly (non-dyadic
    spacing and origin) and a perturbed (Newton) mesh."""
    a = W.structured_mesh([10, 7], [0.1, 0.0], [1.3, 0.7])
    b = W.structured_mesh([12, 6], [0.0, 0.0], [2.2, 0.41])
    c = perturbed_mesh(W.structured_mesh([10, 7], [0.1, 0.0], [1.3, 0.7]), 0.2, seed=31)
    return {"box_10x7": a, "box_12x6": b, "perturbed_10x7": c}


@pytest.mark.parametrize("name", ["box_10x7", "box_12x6", "perturbed_10x7"])
def test_seeded_particles_relocated_by_a_step_at_rest(gpu_available, name):
    """The reference's cases start from rest: the lattice is seeded at part
    centres (xi = centre exactly), the first move under a zero field leaves
    every position unchanged, and its locate stores xi = F^-1(x), which can
    differ from the centre by an ulp; the next step's gathers and corrections
    use that xi. Positions, reference positions and velocities are compared
    bitwise with the oracle after seeding, after the step at rest and after
    two steps of flow."""
    _say("seed -> step at rest -> 2 flow steps on %s" % name)
    from oracle import Oracle
    from pfem2dealii_amd import ParticleHandler
    mesh = _placed_meshes()[name]
    V = W.poiseuille(mesh)
    V0 = np.zeros_like(V)
    params = dict(n_sub=5, dt=W.cfl_dt(mesh, V), max_per_part=5, particles_per_cell=[2, 2])
    o = Oracle(mesh, params)
    o.initialize_maps()
    o.seed_lattice(V0)
    h = ParticleHandler(mesh, params)
    h.initialize_maps()
    h.seed_particles(V0)
    assert_particles_match(h.download(), o.download())
    prev = V0
    for s, F in enumerate((V0, V, V)):
        o.correct(F, prev)
        h.correct_particle_velocity(F, prev)
        o.move(F)
        h.move_particles(F)
        Po = o.project()
        Pg = h.project_particle_fields()
        assert_fields_match(Pg, Po)
        d_o = o.download()
        assert_particles_match(h.download(), d_o)
        if s == 0:  # the case is live: some relocated xi are not the centres
            q = np.array(params["particles_per_cell"], dtype=float)[:, None]
            centre = (np.floor(d_o["xi"] * q) + 0.5) / q
            assert np.any(d_o["xi"] != centre), "no xi moved off its centre: the case tests nothing"
        prev = Po


def test_insert_rejects_outside(gpu_available):
    _say("insert with particles outside their cell")
    cfg = W.config("channel2d", scale=0.25)
    x = cfg["x"].copy()
    cells = cfg["cells"].copy()
    cells[::7] = (cells[::7] + 3) % cells.max()
    cells[5] = -1
    cells[9] = 10 ** 6
    o = OracleStepper(cfg)
    g = GpuStepper(cfg)
    ro = o.insert(x, cells, cfg["V"])
    rg = g.insert(x, cells, cfg["V"])
    assert ro == rg and rg > 0
    assert_particles_match(g.download(), o.download())


@pytest.mark.parametrize("dim", [2, 3])
def test_box_reference_position_is_the_host_quotient(gpu_available, dim):
    """The box inverse map xi_k = (x_k - X0_k) / (X1_k - X0_k) divides with
    the edge's reciprocal and one FMA residual step; the result must be the
    IEEE quotient bit for bit (the reference's box case). Edge lengths that
    are not powers of two (a mesh spacing of (hi - lo) / n with awkward ends),
    positions at random reference coordinates and exactly on the faces."""
    _say("box inverse map against the host quotient, dim %d" % dim)
    from pfem2dealii_amd import ParticleHandler
    n = [7, 5, 3][:dim]
    mesh = W.structured_mesh(n, [0.1, -0.3, 1.0 / 7.0][:dim], [2.0 + 1.0 / 3.0, 0.7, 1e3 / 3.0][:dim])
    nc = int(np.prod(n))
    per = 400
    cells = np.repeat(np.arange(nc, dtype=np.int32), per)
    rng = np.random.default_rng(5 + dim)
    xi = rng.random((dim, cells.shape[0]))
    xi[:, ::9] = 0.0
    xi[:, 4::9] = 1.0
    xi[:, 2::11] = np.nextafter(1.0, 0.0)
    cv = mesh["cell_vertices"]
    X = mesh["X"]
    X0 = np.stack([X[k][cv[cells, 0]] for k in range(dim)])
    X1 = np.stack([X[k][cv[cells, (1 << dim) - 1]] for k in range(dim)])
    x = X0 + xi * (X1 - X0)
    x = np.minimum(np.maximum(x, X0), X1)  # inside the cell after rounding
    params = dict(n_sub=5, dt=1e-3, max_per_part=5, particles_per_cell=[3, 3, 3][:dim])
    h = ParticleHandler(mesh, params)
    h.initialize_maps()
    assert h.insert_particles(x, cells) == 0
    d = h.download(xi=True)
    order = np.lexsort([d["x"][k] for k in range(dim)][::-1] + [d["cell"]])
    ref = np.lexsort([x[k] for k in range(dim)][::-1] + [cells])
    xs, cs = x[:, ref], cells[ref]
    assert np.array_equal(d["cell"][order], cs) and np.array_equal(d["x"][:, order], xs)
    c0 = np.stack([X[k][cv[cs, 0]] for k in range(dim)])
    c1 = np.stack([X[k][cv[cs, (1 << dim) - 1]] for k in range(dim)])
    want = (xs - c0) / (c1 - c0)
    # no position is near a sub-cell centre, where the library snaps a
    # particle placed exactly at the centre's image back to the centre
    q = np.array(params["particles_per_cell"], dtype=float)[:, None]
    centre = (np.floor(want * q) + 0.5) / q
    assert not np.all(np.abs(want - centre) <= 1e-9, axis=0).any()
    got = d["xi"][:, order]
    assert np.array_equal(got.view(np.int64), want.view(np.int64))
    h.close()


def test_empty_handler(gpu_available):
    _say("empty handler: move and project with no particles")
    cfg = W.config("channel2d", scale=0.25)
    g = GpuStepper(cfg)
    g.h.move_particles(cfg["V"])
    assert g.h.count() == 0
    Vp = g.h.project_particle_fields()
    assert np.isnan(Vp).all()


def test_device_pointers(gpu_available):
    _say("device-resident fields through torch tensors")
    import torch
    cfg = W.config("channel2d", scale=0.5)
    o = OracleStepper(cfg)
    g = GpuStepper(cfg)
    o.insert(cfg["x"], cfg["cells"], cfg["V"])
    g.insert(cfg["x"], cfg["cells"], cfg["V"])
    Vd = torch.tensor(cfg["V"], dtype=torch.float64, device="cuda")
    out = torch.empty_like(Vd)
    for _ in range(2):
        Vo = o.step(cfg["V"])
        torch.cuda.synchronize()
        g.h.move_particles(Vd)
        g.h.project_particle_fields(out)
        g.h.correct_particle_velocity(Vd, out)
        g.h.synchronize()
        assert_fields_match(out.cpu().numpy(), Vo)
        assert_particles_match(g.download(), o.download())


def _lattice_on_faces(mesh, m):
    """(m+1)^2 particles per cell at xi = (i/m, j/m), i, j = 0..m: on faces,
    at vertices and at sub-cell centres, all exactly representable."""
    nc = mesh["cell_vertices"].shape[0]
    t = np.arange(m + 1) / float(m)
    a, b = np.meshgrid(t, t, indexing="ij")
    xi = np.stack([np.tile(a.ravel(), nc), np.tile(b.ravel(), nc)])
    cells = np.repeat(np.arange(nc, dtype=np.int32), (m + 1) ** 2)
    return W.seed_from_unit(mesh, cells, xi), cells


@pytest.mark.parametrize("u", [(1.0, 1.0), (1.0, 0.0), (-1.0, 0.5), (0.0, -1.0), (0.0, 0.0)])
def test_walk_ties_on_faces_and_vertices(gpu_available, u):
    """Constant fields with power-of-two substeps on a 16x8 mesh: particles
    start and land exactly on faces and vertices (two or more containing
    candidates), and axis-aligned or diagonal velocities give exact ties in
    the closest-vertex cosines, so every exact branch of the walk runs."""
    _say("walk ties, constant field %s, 16x8 mesh, 3 steps" % (u,))
    mesh = W.structured_mesh([16, 8], [0.0, 0.0], [2.0, 1.0])
    nd = mesh["X"].shape[1]
    V = np.stack([np.full(nd, u[0]), np.full(nd, u[1])])
    cfg = {"mesh": mesh, "V": V, "params": dict(n_sub=4, dt=0.125, max_per_part=5, particles_per_cell=[4, 4])}
    x, cells = _lattice_on_faces(mesh, 4)
    _run_pair(cfg, 3, x=x, cells=cells)


def test_dense_chunks_exceed_lds_stage(gpu_available):
    """150 particles in each of the first 40 cells of a 16x8 channel: their
    chunks hold more particles than the settle kernel stages in LDS, so the
    insertion re-binning, the capped redistribution and the projection of
    those chunks run through the global-memory path."""
    _say("dense chunks beyond the LDS stage, 16x8 channel, 2 steps")
    cfg = W.config("channel2d", scale=0.25)
    mesh = cfg["mesh"]
    nc = mesh["cell_vertices"].shape[0]
    rng = np.random.default_rng(21)
    counts = np.zeros(nc, dtype=np.int64)
    counts[:40] = 150
    counts[40:] = rng.integers(0, 12, nc - 40)
    cells = np.repeat(np.arange(nc, dtype=np.int32), counts)
    xi = rng.random((2, cells.shape[0]))
    x = W.seed_from_unit(mesh, cells, xi)
    _run_pair(cfg, 2, x=x, cells=cells)


def test_remove_particles_keeps_the_order(gpu_available):
    """remove_particle (src/pfem2particlehandler.cpp:332-336, an erase): the
    particles at listed positions leave the handle, the others keep their
    order, cell segments and placed flags; the next steps equal the oracle's
    after the same erase."""
    _say("remove particles after a step, channel 32x16, then 2 steps")
    cfg = W.config("channel2d", scale=0.5)
    o = OracleStepper(cfg)
    g = GpuStepper(cfg)
    o.insert(cfg["x"], cfg["cells"], cfg["V"])
    g.insert(cfg["x"], cfg["cells"], cfg["V"])
    o.step(cfg["V"])
    g.step(cfg["V"])
    before = g.download()
    n = before["cell"].shape[0]
    rng = np.random.default_rng(17)
    idx = rng.choice(n, size=n // 3, replace=False)
    idx = np.concatenate([idx, idx[:5]])  # repeats are removed once
    g.h.remove_particles(idx)
    o.o.remove(np.unique(idx))
    keep = np.ones(n, dtype=bool)
    keep[idx] = False
    after = g.download()
    for k in ("x", "xi", "v"):
        assert np.array_equal(after[k], before[k][:, keep]), k
    assert np.array_equal(after["cell"], before["cell"][keep])
    assert np.array_equal(g.h.cell_counts(), np.bincount(before["cell"][keep], minlength=g.h.n_cells))
    assert_particles_match(after, o.download())
    for _ in range(2):
        assert_fields_match(g.step(cfg["V"]), o.step(cfg["V"]))
        assert_particles_match(g.download(), o.download())
    with pytest.raises(Exception):
        g.h.remove_particles([g.h.count()])
