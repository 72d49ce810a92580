"""Pins the FP64 CPU reference gradient (oracle/adjoint_ref.py) on the CPU.

The frozen decisions (survivor maps, refill parts, end cells) come from the
oracle's own moves: each particle carries its index in v[0] and a sentinel in
v[1] (advection reads only the field, so the carried velocity does not steer
the move); after the move a survivor shows its old index, a refill shows the
field's value and its sub-cell part. The reference gradient is then checked
against central finite differences of the reference's own loss and against
closed forms that hold for the same model: the loss of a zero residual, and
the v0 gradient of a single step without refills (a projection-only term)."""
import numpy as np
import pytest

from oracle.adjoint_ref import FrozenModel
from oracle.pyoracle import Oracle
from pfem2dealii_amd import workloads as W

SENTINEL = -7.0


def trace(cfg, V, steps):
    """(x0, cells0, srcs, end_cells) of `steps` oracle moves under V."""
    mesh, params = cfg["mesh"], cfg["params"]
    D = mesh["dim"]
    q = params["particles_per_cell"]
    x, cells = cfg["x"], cfg["cells"]
    x0, c0 = x.copy(), cells.copy()
    srcs, ecs = [], []
    for _ in range(steps):
        o = Oracle(mesh, params)
        o.initialize_maps()
        n = cells.shape[0]
        v = np.zeros((D, n))
        v[0] = np.arange(n, dtype=np.float64)
        v[1] = SENTINEL
        assert o.insert(x, cells, v) == 0
        o.move(V)
        d = o.download()
        surv = d["v"][1] == SENTINEL
        src = np.empty(d["cell"].shape[0], dtype=np.int64)
        src[surv] = d["v"][0][surv].astype(np.int64)
        lin = np.zeros(int((~surv).sum()), dtype=np.int64)
        for k in range(D):  # part number, x slowest
            lin = lin * (q[k] + 1) + np.floor(d["xi"][k][~surv] * q[k]).astype(np.int64)
        src[~surv] = -1 - lin
        srcs.append(src)
        ecs.append(d["cell"].astype(np.int64))
        x, cells = d["x"], d["cell"]
        o.close()
    return x0, c0, srcs, ecs


@pytest.fixture(scope="module", params=["adjoint2d", "adjoint3d"])
def case(request):
    if request.param == "adjoint2d":  # 16x8 channel, Poiseuille, 4 steps
        cfg = W.config("adjoint2d", scale=1.0 / 16)
        V = W.poiseuille(cfg["mesh"], 1.0)
        Vt = W.poiseuille(cfg["mesh"], 1.05)
        steps = 4
    else:  # 8^3 unit cube, ABC flow, 3 steps; particles leave through the faces
        cfg = W.config("adjoint3d", scale=1.0 / 2)
        V = W.abc_flow(cfg["mesh"])
        Vt = W.abc_flow(cfg["mesh"], A=1.05)
        steps = 3
    x0, c0, srcs, ecs = trace(cfg, V, steps)
    assert any((s < 0).any() for s in srcs), "workload must exercise refills"
    m = FrozenModel(cfg["mesh"], cfg["params"], x0, c0, srcs, ecs)
    v0 = np.stack([np.interp(x0[1], [0, 1], [0.3, 0.9]), 0.1 * np.sin(3 * x0[0])] +
                  [0.2 * np.cos(2 * x0[k]) for k in range(2, cfg["mesh"]["dim"])])
    # target: the projection of a perturbed run along the same frozen path
    m.forward(V, v0, record=True)
    xt, vt, ct, pt = m.forward(Vt, 1.02 * v0)
    P, _, _ = m.project(xt, vt, ct, pt)
    return {"cfg": cfg, "V": V, "v0": v0, "target": P, "m": m}


def test_reference_forward_reproduces_oracle_cells(case):
    m = case["m"]
    x, v, c, _ = m.forward(case["V"], case["v0"], record=True)
    assert np.array_equal(c, m.end_cells[-1])


def test_reference_gradient_matches_central_differences(case):
    m, V, v0, T = case["m"], case["V"], case["v0"], case["target"]
    L, gv0, gV = m.grad(V, v0, T)
    assert L > 0
    rng = np.random.default_rng(11)
    eps = 1e-6
    for _ in range(6):
        dV = rng.standard_normal(V.shape)
        dv = rng.standard_normal(v0.shape)
        g = float((gV * dV).sum() + (gv0 * dv).sum())
        fd = (m.loss(V + eps * dV, v0 + eps * dv, T) - m.loss(V - eps * dV, v0 - eps * dv, T)) / (2 * eps)
        assert abs(fd - g) <= 1e-6 * abs(g), (fd, g)


def test_reference_gradient_of_zero_residual_is_zero(case):
    m, V, v0 = case["m"], case["V"], case["v0"]
    x, v, c, placed = m.forward(V, v0, record=True)
    P, _, _ = m.project(x, v, c, placed)
    L, gv0, gV = m.grad(V, v0, P)
    assert L == 0.0 and not gv0.any() and not gV.any()


def test_reference_v0_gradient_is_projection_weight_without_refills():
    """One step whose survivors are every particle: dL/dv0 of particle p is
    sum_i N_i(xi_p) r_i / W_i (the transpose of the projection), and dL/dV is
    driven only by the positions."""
    cfg = W.config("channel2d", scale=0.25)
    mesh = cfg["mesh"]
    V = np.zeros((2, mesh["n_dofs"]))
    x0, c0, srcs, ecs = trace(cfg, V, 1)
    # one particle per sub-cell and a zero field: nothing moves, nothing refills
    assert (srcs[0] >= 0).all() and srcs[0].shape[0] == x0.shape[1]
    m = FrozenModel(mesh, cfg["params"], x0, c0, srcs, ecs)
    rng = np.random.default_rng(3)
    v0 = rng.standard_normal(x0.shape)
    T = rng.standard_normal((2, mesh["n_dofs"]))
    L, gv0, gV = m.grad(V, v0, T)
    x, v, c, placed = m.forward(V, v0)
    P, Wt, xi = m.project(x, v, c, placed)
    r = P - T
    N = np.stack([(1 - xi[0]) * (1 - xi[1]), xi[0] * (1 - xi[1]), (1 - xi[0]) * xi[1], xi[0] * xi[1]])
    want = np.zeros_like(v)
    for i in range(4):
        d = mesh["cell_vertices"][c, i]
        want += N[i] * r[:, d] / Wt[d]
    inv = np.empty_like(srcs[0])
    inv[srcs[0]] = np.arange(srcs[0].shape[0])
    assert np.allclose(gv0, want[:, inv], rtol=1e-13, atol=1e-15)
