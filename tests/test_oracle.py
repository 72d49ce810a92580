"""Pins the CPU oracle to the mathematics of the reference's particle path.

The reference ships no tests, golden vectors or fixtures for this path, so
each oracle function is checked against a closed form, an invariant or a
brute-force search chosen so that a dropped term, a wrong sign or index or a
transposed operand fails at least one check.
"""
import math

import numpy as np
import pytest

from oracle import MODE_CANONICAL, MODE_REFERENCE, Oracle
from oracle.pyoracle import key as orc_key
from pfem2dealii_amd import workloads as W
from tests.harness import perturbed_mesh, seed_in_cells


def _mesh2(nx=4, ny=3, lo=(0.0, 0.0), hi=(2.0, 1.0)):
    return W.structured_mesh([nx, ny], list(lo), list(hi))


def _params(dim, q=4, max_pp=5, n_sub=5, dt=0.01):
    return dict(n_sub=n_sub, dt=dt, max_per_part=max_pp, particles_per_cell=[q] * dim)


# ---------------------------------------------------------------- A2 shape

@pytest.mark.parametrize("dim", [2, 3])
def test_shape_partition_of_unity_and_kronecker(dim):
    m = W.structured_mesh([2] * dim, [0.0] * dim, [1.0] * dim)
    o = Oracle(m, _params(dim))
    rng = np.random.default_rng(0)
    for _ in range(200):
        xi = rng.random(dim)
        N = o.shape(xi)
        assert abs(N.sum() - 1.0) <= 4e-16 * (1 << dim)
        assert (N >= 0).all()
    for i in range(1 << dim):
        xi = np.array([(i >> k) & 1 for k in range(dim)], dtype=float)
        N = o.shape(xi)
        e = np.zeros(1 << dim)
        e[i] = 1.0
        assert np.array_equal(N, e)  # exact Kronecker property at the vertices


def test_shape_vertex_order_and_product_form():
    """Vertex i has bit k -> +x_k; N_i = L(xi_0)*L(xi_1) with L0 = 1-xi, L1 = xi."""
    m = _mesh2()
    o = Oracle(m, _params(2))
    xi = np.array([0.3, 0.8])
    N = o.shape(xi)
    a, b = xi
    assert N[0] == (1 - a) * (1 - b)
    assert N[1] == a * (1 - b)
    assert N[2] == (1 - a) * b
    assert N[3] == a * b


# --------------------------------------------------- A5 forward/inverse map

def test_box_inverse_map_closed_form():
    m = W.structured_mesh([8, 4], [0.0, -1.0], [2.0, 3.0])
    o = Oracle(m, _params(2))
    rng = np.random.default_rng(1)
    for c in range(32):
        assert o.is_box(c)
        X0 = m["X"][:, m["cell_vertices"][c, 0]]
        h = np.array([m["X"][0, m["cell_vertices"][c, 1]] - X0[0], m["X"][1, m["cell_vertices"][c, 2]] - X0[1]])
        x = X0 + rng.random(2) * h
        st, xi = o.inv(c, x)
        assert st == 1
        assert np.array_equal(xi, (x - X0) / h)
        st, xi = o.inv(c, X0 + np.array([1.5, 0.5]) * h)
        assert st == 0 and abs(xi[0] - 1.5) < 1e-14


@pytest.mark.parametrize("dim", [2, 3])
def test_newton_round_trip_on_perturbed_cells(dim):
    base = W.structured_mesh([4] * dim, [0.0] * dim, [1.0] * dim)
    m = perturbed_mesh(base, 0.25, seed=3)
    o = Oracle(m, _params(dim))
    rng = np.random.default_rng(2)
    nonbox = 0
    for c in range(m["cell_vertices"].shape[0]):
        if not o.is_box(c):
            nonbox += 1
        for _ in range(5):
            xi = rng.random(dim)
            x = o.fwd(c, xi)
            st, xr = o.inv(c, x)
            assert st == 1
            assert np.abs(xr - xi).max() <= 1e-13
    assert nonbox > 0


def test_forward_map_of_bilinear_cell():
    """x = sum N_i X_i reproduces the bilinear map of the vertices."""
    m = perturbed_mesh(_mesh2(3, 3, hi=(1.0, 1.0)), 0.3, seed=5)
    o = Oracle(m, _params(2))
    c = 4  # interior cell, all 4 vertices moved
    X = m["X"][:, m["cell_vertices"][c]]
    for xi in ([0, 0], [1, 0], [0, 1], [1, 1]):
        i = xi[0] + 2 * xi[1]
        assert np.array_equal(o.fwd(c, np.array(xi, float)), X[:, i])
    xi = np.array([0.25, 0.5])
    ref = (0.75 * 0.5) * X[:, 0] + (0.25 * 0.5) * X[:, 1] + (0.75 * 0.5) * X[:, 2] + (0.25 * 0.5) * X[:, 3]
    assert np.abs(o.fwd(c, xi) - ref).max() < 1e-15


# ------------------------------------------------------- A7 maps, A6 vertex

def test_vertex_to_cell_map_and_directions():
    m = _mesh2(4, 3)
    o = Oracle(m, _params(2))
    o.initialize_maps()
    cv = m["cell_vertices"]
    for v in range(m["X"].shape[1]):
        cells, dirs = o.v2c(v)
        expect = sorted(set(np.nonzero((cv == v).any(axis=1))[0].tolist()))
        assert cells.tolist() == expect
        for c, d in zip(cells, dirs):
            ctr = m["X"][:, cv[c]].mean(axis=1)
            u = ctr - m["X"][:, v]
            u = u / np.linalg.norm(u)
            assert abs(np.linalg.norm(d) - 1.0) < 1e-15
            assert np.abs(d - u).max() < 1e-15


def test_closest_vertex_hand_cases():
    m = _mesh2(4, 4, hi=(1.0, 1.0))
    o = Oracle(m, _params(2))
    c = 5
    centre = np.array([0.5, 0.5])
    assert o.closest_vertex(c, centre, np.array([1.0, 1.0])) == 3
    assert o.closest_vertex(c, centre, np.array([-1.0, -1.0])) == 0
    assert o.closest_vertex(c, centre, np.array([-1.0, 1.0])) == 2
    # +x from the centre ties vertices 1 and 3: the strict '>' scan keeps 1
    assert o.closest_vertex(c, centre, np.array([1.0, 0.0])) == 1
    assert o.closest_vertex(c, centre, np.array([0.0, 1.0])) == 2
    # zero velocity -> NaN cosines -> vertex 0
    assert o.closest_vertex(c, centre, np.array([0.0, 0.0])) == 0
    # off-centre: heading +x from near the bottom edge favours vertex 1
    assert o.closest_vertex(c, np.array([0.5, 0.1]), np.array([1.0, 0.0])) == 1
    assert o.closest_vertex(c, np.array([0.5, 0.9]), np.array([1.0, 0.0])) == 3


# --------------------------------------------- A3/A4/A12 gather, Euler, correct

def _single_particle_run(m, params, x, cell, V, v=None):
    o = Oracle(m, params)
    o.initialize_maps()
    assert o.insert(np.asarray(x, float).reshape(-1, 1), np.array([cell]),
                    None if v is None else np.asarray(v, float).reshape(-1, 1)) == 0
    o.move(V)
    return o


def test_constant_field_advection_exact_sum():
    m = _mesh2(8, 4)
    V = np.zeros_like(m["X"])
    V[0] = 0.5
    V[1] = -0.25
    p = _params(2, n_sub=5, dt=0.02)
    o = _single_particle_run(m, p, [0.6, 0.6], 18, V)
    d = o.download()
    dts = 0.02 / 5.0
    x = np.array([0.6, 0.6])
    for _ in range(5):
        x = x + V[:, 0] * dts   # interpolation of a constant field is exact
    # the lone particle plus the refills of its cell's 15 other parts
    i = int(np.nonzero((np.abs(d["x"][0] - x[0]) < 1e-15) & (np.abs(d["x"][1] - x[1]) < 1e-15))[0][0])
    assert d["x"][0, i] == x[0] and d["x"][1, i] == x[1]


def test_bilinear_field_interpolated_exactly():
    """Q1 interpolation is isoparametric: on any Q1 cell, perturbed or not, a
    nodal field affine in the physical coordinates, V_i = a + B X_i, is
    reproduced at x = F(xi) (sum_i N_i (a + B X_i) = a + B F(xi)); on a box
    mesh a field bilinear in x and y is reproduced too."""
    m = perturbed_mesh(_mesh2(4, 4, hi=(1.0, 1.0)), 0.2, seed=7)
    X = m["X"]
    V = np.stack([1.0 + 2.0 * X[0] - 0.5 * X[1], -0.5 + 0.25 * X[0] + 3.0 * X[1]])
    o = Oracle(m, _params(2))
    o.initialize_maps()
    x, cells = seed_in_cells(m, 4, seed=8)
    assert not all(o.is_box(int(c)) for c in np.unique(cells))  # non-affine cells present
    assert o.insert(x, cells) == 0
    o.set_velocity_from_field(V)
    d = o.download()
    xx, yy = d["x"]
    ex = np.stack([1.0 + 2.0 * xx - 0.5 * yy, -0.5 + 0.25 * xx + 3.0 * yy])
    # Newton stops at |dxi| <= 1e-12, so the located xi maps to x within ~1e-12
    assert np.abs(d["v"] - ex).max() < 5e-11
    m2 = _mesh2(4, 4, hi=(1.0, 1.0))
    X2 = m2["X"]
    V2 = np.stack([1.0 + 2.0 * X2[0] - 0.5 * X2[1] + 0.75 * X2[0] * X2[1], 0.5 * X2[0] * X2[1]])
    o2 = Oracle(m2, _params(2))
    x2, c2 = W.seed_uniform(m2, np.full(16, 3), seed=9)
    o2.insert(x2, c2)
    o2.set_velocity_from_field(V2)
    d2 = o2.download()
    xx, yy = d2["x"]
    ex = np.stack([1.0 + 2.0 * xx - 0.5 * yy + 0.75 * xx * yy, 0.5 * xx * yy])
    assert np.abs(d2["v"] - ex).max() < 1e-14


def test_correction_zero_and_constant_increment():
    cfg = W.config("channel2d", scale=0.25)
    o = Oracle(cfg["mesh"], cfg["params"])
    o.initialize_maps()
    o.insert(cfg["x"], cfg["cells"])
    o.set_velocity_from_field(cfg["V"])
    v0 = o.download()["v"]
    o.correct(cfg["V"], cfg["V"])  # dV = 0 leaves v bitwise unchanged
    assert np.array_equal(o.download()["v"], v0)
    Vold = cfg["V"] - np.array([[0.125], [-0.5]])
    o.correct(cfg["V"], Vold)      # constant dV adds exactly dV (partition of unity, dyadic values)
    v1 = o.download()["v"]
    assert np.abs(v1[0] - (v0[0] + 0.125)).max() <= 1e-15
    assert np.abs(v1[1] - (v0[1] - 0.5)).max() <= 1e-15


def test_poiseuille_step_invariants():
    cfg = W.config("channel2d")
    m = cfg["mesh"]
    o = Oracle(m, cfg["params"])
    o.initialize_maps()
    o.insert(cfg["x"], cfg["cells"])
    o.set_velocity_from_field(cfg["V"])
    before = o.download()
    o.move(cfg["V"])
    st = o.stats()
    after = o.download()
    nx = m["n"][0]
    # v = 0 at every node: y is bitwise invariant and the cell row never changes
    key_b = {(a, b): c for a, b, c in zip(before["x"][0], before["x"][1], before["cell"])}
    dts = cfg["params"]["dt"] / 5.0
    h = m["h"]
    moved = 0
    for i in range(after["cell"].shape[0]):
        y = after["x"][1, i]
        row = after["cell"][i] // nx
        assert m["X"][1, m["cell_vertices"][after["cell"][i], 0]] <= y <= m["X"][1, m["cell_vertices"][after["cell"][i], 2]]
        assert row == int(min(y / h[1], m["n"][1] - 1)) or y / h[1] == row + 1
        # x(T) = x0 + 5 dt_sub u_h(y), u_h the linear interpolant of 4y(1-y) in y
        j0 = int(min(y // h[1], m["n"][1] - 1))
        y0, y1 = j0 * h[1], (j0 + 1) * h[1]
        uh = ((y1 - y) * 4 * y0 * (1 - y0) + (y - y0) * 4 * y1 * (1 - y1)) / h[1]
        xs = [k for k in key_b if k[1] == y]
        if len(xs) == 1 and uh > 0:
            assert abs(after["x"][0, i] - (xs[0][0] + 5 * dts * uh)) < 1e-13
            moved += 1
    assert moved > 30000
    # outflow at x = 2 deletes particles; nothing enters, so inflow cells refill
    assert st["deleted_locate"] > 0 and st["refilled"] > 0 and st["deleted_overflow"] == 0


def test_taylor_green_closed_box_conserves_particles():
    cfg = W.config("taylor_green2d", scale=1.0 / 32)
    o = Oracle(cfg["mesh"], cfg["params"])
    o.initialize_maps()
    o.insert(cfg["x"], cfg["cells"])
    o.set_velocity_from_field(cfg["V"])
    for _ in range(3):
        o.move(cfg["V"])
        assert o.stats()["deleted_locate"] == 0


# ------------------------------------------------- A5-A9 walk vs brute force

@pytest.mark.parametrize("dim,perturb", [(2, 0.0), (2, 0.25), (3, 0.0), (3, 0.2)])
def test_walk_matches_brute_force_containment(dim, perturb):
    base = W.structured_mesh([6] * dim, [0.0] * dim, [1.0] * dim)
    m = perturbed_mesh(base, perturb, seed=21) if perturb else base
    X = m["X"]
    V = np.stack([np.sin(3.0 * X[(k + 1) % dim]) + 0.3 for k in range(dim)])
    p = _params(dim, q=2, n_sub=5, dt=0.04)
    o = Oracle(m, p)
    o.initialize_maps()
    x, cells = seed_in_cells(m, 5, seed=22)
    o.insert(x, cells)
    o.set_velocity_from_field(V)
    o.move(V)
    d = o.download()
    nc = m["cell_vertices"].shape[0]
    for i in range(d["cell"].shape[0]):
        inside = [c for c in range(nc) if o.inv(c, d["x"][:, i])[0] == 1]
        assert d["cell"][i] in inside
        if len(inside) == 1:
            assert inside[0] == d["cell"][i]


def test_particle_leaving_domain_is_deleted():
    m = _mesh2(4, 2)
    V = np.zeros_like(m["X"])
    V[0] = 1.0
    p = _params(2, q=1, n_sub=1, dt=0.6)
    o = Oracle(m, p)
    o.initialize_maps()
    o.insert(np.array([[1.9], [0.25]]), np.array([3]))
    o.move(V)
    assert o.stats()["deleted_locate"] == 1
    assert o.count() == 0  # the cell became empty: no refill (empty cells stay empty)


# ------------------------------------------------------------ A10 redistribution

@pytest.mark.parametrize("name,scale", [("channel2d", 1.0), ("taylor_green2d", 1.0 / 16), ("abc3d", 1.0 / 8)])
def test_redistribution_invariants(name, scale):
    cfg = W.config(name, scale=scale)
    p = cfg["params"]
    q = p["particles_per_cell"]
    dim = len(q)
    res = {}
    for mode in (MODE_CANONICAL, MODE_REFERENCE):
        o = Oracle(cfg["mesh"], p, mode)
        o.initialize_maps()
        o.insert(cfg["x"], cfg["cells"])
        o.set_velocity_from_field(cfg["V"])
        before = np.unique(cfg["cells"])
        o.move(cfg["V"])
        d = o.download()
        res[mode] = d
        parts = np.floor(d["xi"] / (1.0 / np.array(q))[:, None]).astype(int)
        lin = np.zeros(d["cell"].shape[0], dtype=np.int64)
        for k in range(dim):
            lin = lin * (q[k] + 1) + parts[k]
        cells = np.unique(d["cell"])
        for c in cells[:400]:
            sel = d["cell"] == c
            cnt = np.bincount(lin[sel], minlength=int(np.prod(np.array(q) + 1)))
            valid = [int(np.ravel_multi_index(tuple(ix), [a + 1 for a in q]))
                     for ix in np.ndindex(*q)]
            assert (cnt[valid] >= 1).all()
            assert (cnt <= p["max_per_part"] + 1).all()
        # refilled particles sit exactly at sub-cell centres
        centre = (np.floor(d["xi"] * np.array(q)[:, None]) + 0.5) / np.array(q)[:, None]
        assert (d["xi"] == centre).all(axis=0).sum() > 0
        assert set(cells.tolist()) <= set(before.tolist()) | set(cells.tolist())
    # counts per cell do not depend on which particles survive
    a = np.bincount(res[MODE_CANONICAL]["cell"], minlength=cfg["mesh"]["cell_vertices"].shape[0])
    b = np.bincount(res[MODE_REFERENCE]["cell"], minlength=cfg["mesh"]["cell_vertices"].shape[0])
    assert np.array_equal(a, b)


def test_overflow_keeps_max_plus_one_smallest_keys():
    m = _mesh2(2, 1, hi=(2.0, 1.0))
    p = _params(2, q=1, max_pp=2, n_sub=1, dt=1e-9)
    o = Oracle(m, p)
    o.initialize_maps()
    rng = np.random.default_rng(4)
    x = np.stack([0.1 + 0.8 * rng.random(10), 0.1 + 0.8 * rng.random(10)])
    o.insert(x, np.zeros(10, dtype=np.int32))
    V = np.zeros_like(m["X"])
    o.move(V)
    d = o.download()
    assert d["cell"].shape[0] == 3  # max+1 survivors, no refill (the single part is occupied)
    keys = sorted(orc_key(x[:, i]) for i in range(10))[:3]
    got = sorted(orc_key(d["x"][:, i]) for i in range(3))
    assert got == keys


# ----------------------------------------------------------- A11 projection

def test_projection_constant_velocity_and_nan_nodes():
    m = _mesh2(6, 4)
    o = Oracle(m, _params(2))
    o.initialize_maps()
    cells = np.array([0, 0, 7, 7, 7, 13], dtype=np.int32)
    X0 = m["X"][:, m["cell_vertices"][cells, 0]]
    x = X0 + np.array([[0.1, 0.2, 0.3, 0.05, 0.2, 0.25], [0.2, 0.1, 0.1, 0.2, 0.15, 0.05]])
    v = np.stack([np.full(6, 2.0), np.full(6, -0.5)])  # powers of two scale sums exactly
    o.insert(x, cells, v)
    Vp = o.project()
    touched = np.unique(m["cell_vertices"][cells].ravel())
    untouched = np.setdiff1d(np.arange(m["X"].shape[1]), touched)
    assert (Vp[0, touched] == 2.0).all() and (Vp[1, touched] == -0.5).all()
    assert np.isnan(Vp[:, untouched]).all()   # 0/0 at nodes no particle reaches


def test_projection_linear_field_on_symmetric_lattice():
    m = W.structured_mesh([8, 8], [0.0, 0.0], [1.0, 1.0])
    o = Oracle(m, _params(2))
    o.initialize_maps()
    o.seed_lattice(np.stack([0.5 + 2.0 * m["X"][0] - 1.0 * m["X"][1], 0.25 * m["X"][1]]))
    Vp = o.project()
    X = m["X"]
    interior = (X[0] > 0) & (X[0] < 1) & (X[1] > 0) & (X[1] < 1)
    assert np.abs(Vp[0, interior] - (0.5 + 2.0 * X[0, interior] - X[1, interior])).max() < 1e-14
    assert np.abs(Vp[1, interior] - 0.25 * X[1, interior]).max() < 1e-14


def test_projection_weights_sum_to_particle_count():
    """sum_j W_j = N by partition of unity (a dropped vertex or a wrong DoF
    loses weight); projecting v = 1 gives exactly 1 at every touched node;
    the canonical and reference orders agree to rounding."""
    cfg = W.config("taylor_green2d", scale=1.0 / 32)
    outs = []
    n = len(cfg["cells"])
    for mode in (MODE_CANONICAL, MODE_REFERENCE):
        o = Oracle(cfg["mesh"], cfg["params"], mode)
        o.initialize_maps()
        assert o.insert(cfg["x"], cfg["cells"], np.stack([np.ones(n), np.zeros(n)])) == 0
        Vp, Wn = o.project_weights()
        assert abs(Wn.sum() - n) <= 1e-12 * n
        assert (Wn >= 0).all() and (Wn > 0).sum() > 0.9 * Wn.shape[0]
        assert np.array_equal(Vp[0][Wn > 0], np.ones(int((Wn > 0).sum())))
        o.set_velocity_from_field(cfg["V"])
        outs.append(o.project())
    fin = ~np.isnan(outs[1])
    assert np.abs(outs[0][fin] - outs[1][fin]).max() <= 1e-13 * np.abs(outs[1][fin]).max()


def test_step_at_rest_relocates_seeded_particles():
    """Seeding stores xi = the exact part centre; a move under a zero field
    leaves x = F(centre) unchanged and its locate stores xi = F^-1(x) (the
    reference's current-cell test, src/pfem2particlehandler.cpp:420-436).
    On a mesh whose spacing is not dyadic that differs from the centre for
    some particles, and always equals the oracle's own inverse map."""
    m = W.structured_mesh([10, 7], [0.1, 0.0], [1.3, 0.7])
    p = _params(2, q=2)
    o = Oracle(m, p)
    o.initialize_maps()
    V0 = np.zeros_like(m["X"])
    o.seed_lattice(V0)
    d0 = o.download()
    q = 2.0
    centre = (np.floor(d0["xi"] * q) + 0.5) / q
    assert np.array_equal(d0["xi"], centre)
    o.move(V0)
    d1 = o.download()
    assert np.array_equal(d1["x"], d0["x"]) and np.array_equal(d1["cell"], d0["cell"])
    for i in range(d1["cell"].shape[0]):
        code, xi = o.inv(int(d1["cell"][i]), d1["x"][:, i])
        assert code == 1 and np.array_equal(xi, d1["xi"][:, i])
    assert (d1["xi"] != d0["xi"]).any(axis=0).sum() > 0
