"""Adjoint of advect + project on the GPU against an FP64 CPU reference
gradient of the same frozen-decision steps and against central finite
differences. Marked gpu.

Workloads:
  - 2D: the adjoint channel (Q1 mesh of [0,2]x[0,1] at 1/8 of 256x128 per
    direction, 16 seeded uniform particles per cell), Poiseuille amplitude
    1.0, 20 steps of move, final projection, loss 1/2 |P - P(amplitude
    1.05)|^2, checkpoints every 5 steps;
  - 3D: the 16^3 Q1 unit cube with 8 seeded uniform particles per cell, ABC
    flow (A, B, C) = (1, 0.7, 0.4), 12 steps (particles leave through the
    faces, refills fill the emptied parts), target from A = 1.05,
    checkpoints every 5 steps (a last segment of 2 steps).
Central differences along 64 (2D) or 16 (3D) seeded random directions,
epsilon 1e-6."""
import sys

import numpy as np
import pytest

from pfem2dealii_amd import workloads as W

pytestmark = pytest.mark.gpu

CASES = {
    "2d": dict(workload="adjoint2d", scale=1.0 / 8, steps=20, ckpt=5, fd_dirs=64,
               field=lambda mesh, a: W.poiseuille(mesh, a)),
    "3d": dict(workload="adjoint3d", scale=1.0, steps=12, ckpt=5, fd_dirs=16,
               field=lambda mesh, a: W.abc_flow(mesh, A=a)),
}


def _say(msg):
    print("[case] " + msg, flush=True)
    sys.stdout.flush()


def _handle(cfg, V, x=None, cells=None, v=None):
    from pfem2dealii_amd import ParticleHandler
    h = ParticleHandler(cfg["mesh"], cfg["params"])
    h.initialize_maps()
    x = cfg["x"] if x is None else x
    cells = cfg["cells"] if cells is None else cells
    assert h.insert_particles(x, cells, v) == 0
    if v is None:
        h.set_velocity_from_field(V)
    return h


@pytest.fixture(scope="module", params=sorted(CASES))
def adjoint_case(request):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from oracle.adjoint_ref import FrozenModel
    cs = CASES[request.param]
    steps, ckpt = cs["steps"], cs["ckpt"]
    _say("%s: %s at scale %g, %d steps, checkpoints every %d" % (request.param, cs["workload"], cs["scale"], steps,
                                                                  ckpt))
    cfg = W.config(cs["workload"], scale=cs["scale"])
    mesh = cfg["mesh"]
    V1 = cs["field"](mesh, 1.0)
    V105 = cs["field"](mesh, 1.05)
    ht = _handle(cfg, V105)
    for _ in range(steps):
        ht.move_particles(V105)
    target = ht.project_particle_fields()
    ht.close()
    assert np.isfinite(target).all()
    h = _handle(cfg, V1)
    d0 = h.download(xi=False)
    Vd = torch.tensor(V1, dtype=torch.float64, device="cuda")
    Td = torch.tensor(target, dtype=torch.float64, device="cuda")
    loss, gv0, gV = h.adjoint(steps, ckpt, Vd, Td)
    hr = _handle(cfg, V1)
    d0r = hr.download(xi=False)
    assert all(np.array_equal(d0[k], d0r[k]) for k in ("x", "v", "cell"))
    srcs, ecs = [], []
    for _ in range(steps):
        srcs.append(hr.move_particles_traced(V1))
        ecs.append(hr.download(xi=False)["cell"])
    fin = h.download(xi=False)
    assert np.array_equal(fin["cell"], ecs[-1]) and np.array_equal(fin["x"], hr.download(xi=False)["x"])
    assert any((s < 0).any() for s in srcs), "workload must exercise refills"
    h.close()
    hr.close()
    ref = FrozenModel(mesh, cfg["params"], d0["x"], d0["cell"], srcs, ecs)
    return {"name": request.param, "steps": steps, "ckpt": ckpt, "fd_dirs": cs["fd_dirs"], "cfg": cfg, "V": V1,
            "target": target, "v0": d0["v"], "x0": d0["x"], "c0": d0["cell"], "loss": loss, "gv0": gv0.cpu().numpy(),
            "gV": gV.cpu().numpy(), "ref": ref}


def test_adjoint_matches_cpu_reference_gradient(adjoint_case):
    c = adjoint_case
    _say("%s: adjoint vs FP64 CPU reference gradient" % c["name"])
    L, gv0, gV = c["ref"].grad(c["V"], c["v0"], c["target"])
    assert abs(c["loss"] - L) <= 1e-11 * abs(L)
    assert np.abs(c["gv0"] - gv0).max() <= 1e-10 * np.abs(gv0).max()
    assert np.abs(c["gV"] - gV).max() <= 1e-10 * np.abs(gV).max()
    assert np.abs(gV).max() > 0 and np.abs(gv0).max() > 0


def test_adjoint_matches_finite_differences(adjoint_case):
    c = adjoint_case
    _say("%s: adjoint vs central finite differences, %d directions, eps 1e-6" % (c["name"], c["fd_dirs"]))
    ref = c["ref"]
    ref.grad(c["V"], c["v0"], c["target"])  # records the frozen path
    rng = np.random.default_rng(7)
    eps = 1e-6
    worst = 0.0
    for _ in range(c["fd_dirs"]):
        dV = rng.standard_normal(c["V"].shape)
        dv = rng.standard_normal(c["v0"].shape)
        gdot = float((c["gV"] * dV).sum() + (c["gv0"] * dv).sum())
        Lp = ref.loss(c["V"] + eps * dV, c["v0"] + eps * dv, c["target"])
        Lm = ref.loss(c["V"] - eps * dV, c["v0"] - eps * dv, c["target"])
        fd = (Lp - Lm) / (2 * eps)
        worst = max(worst, abs(fd - gdot) / abs(gdot))
    assert worst <= 1e-6, worst


def test_adjoint_through_autograd(adjoint_case):
    import torch
    c = adjoint_case
    _say("%s: adjoint wrapped in torch.autograd.Function" % c["name"])
    cfg = c["cfg"]

    class AdvectProject(torch.autograd.Function):
        @staticmethod
        def forward(ctx, V, v0):
            h = _handle(cfg, None, x=c["x0"], cells=c["c0"], v=v0.detach().cpu().numpy())
            T = torch.tensor(c["target"], device="cuda")
            loss, gv0, gV = h.adjoint(c["steps"], c["ckpt"], V.detach().contiguous(), T)
            ctx.save_for_backward(gv0, gV)
            h.close()
            return torch.tensor(loss, dtype=torch.float64, device=V.device)

        @staticmethod
        def backward(ctx, g):
            gv0, gV = ctx.saved_tensors
            return g * gV, g * gv0

    V = torch.tensor(c["V"], dtype=torch.float64, device="cuda", requires_grad=True)
    v0 = torch.tensor(c["v0"], dtype=torch.float64, device="cuda", requires_grad=True)
    loss = AdvectProject.apply(V, v0)
    (3.0 * loss).backward()
    assert abs(loss.item() - c["loss"]) <= 1e-12 * abs(c["loss"])
    assert np.abs(V.grad.cpu().numpy() - 3.0 * c["gV"]).max() <= 1e-10 * np.abs(c["gV"]).max()
    assert np.abs(v0.grad.cpu().numpy() - 3.0 * c["gv0"]).max() <= 1e-10 * np.abs(c["gv0"]).max()


def test_adjoint_rejects_non_box_cells_and_too_many_substeps():
    """The adjoint differentiates the box-cell inverse map and holds at most 8
    substeps of path per particle; other cases are refused before any
    device work."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from pfem2dealii_amd import ParticleHandler, Pfem2GpuError
    from tests.harness import perturbed_mesh, seed_in_cells
    _say("adjoint argument checks: non-box mesh, 9 substeps")
    cfg = W.config("channel2d", scale=0.25)
    mesh = perturbed_mesh(cfg["mesh"], 0.2, seed=3)
    x, cells = seed_in_cells(mesh, 4, seed=4)
    h = ParticleHandler(mesh, cfg["params"])
    h.initialize_maps()
    h.insert_particles(x, cells)
    V = torch.tensor(cfg["V"], dtype=torch.float64, device="cuda")
    with pytest.raises(Pfem2GpuError, match="box"):
        h.adjoint(2, 1, V, V)
    h.close()
    params = dict(cfg["params"], n_sub=9)
    h = ParticleHandler(cfg["mesh"], params)
    h.initialize_maps()
    h.insert_particles(cfg["x"], cfg["cells"])
    with pytest.raises(Pfem2GpuError, match="substeps"):
        h.adjoint(2, 1, V, V)
    h.close()


def test_adjoint_is_deterministic(adjoint_case):
    """The gradient accumulators are integer (fixed-point) sums, so a second
    call on the same state returns the same bits."""
    import torch
    c = adjoint_case
    _say("%s: adjoint called again gives the same bits" % c["name"])
    cfg = c["cfg"]
    h = _handle(cfg, None, x=c["x0"], cells=c["c0"], v=c["v0"])
    V = torch.tensor(c["V"], dtype=torch.float64, device="cuda")
    T = torch.tensor(c["target"], dtype=torch.float64, device="cuda")
    loss, gv0, gV = h.adjoint(c["steps"], c["ckpt"], V, T)
    h.close()
    assert loss == c["loss"]
    assert np.array_equal(gv0.cpu().numpy().view(np.int64), c["gv0"].view(np.int64))
    assert np.array_equal(gV.cpu().numpy().view(np.int64), c["gV"].view(np.int64))
