"""The FEM assembly oracle (oracle/fem_oracle.cpp, src/pfem2fem.cpp:392-517)
against closed forms that fix it independently of its own formulas:

  - on box meshes, 2-point Gauss integrates Q1 products exactly, so the
    Laplacian is the tensor-product sum of 1D stiffness and mass matrices
    (K1 = [1 -1; -1 1]/h, M1 = h [2 1; 1 2]/6) and the mass matrix their
    product, under any DoF numbering;
  - on perturbed (non-affine) meshes the Laplacian has zero row sums (constants
    are in its kernel) and is symmetric bit for bit;
  - the pressure right-hand side with every boundary face flagged is the
    discrete divergence theorem: a constant field gives 0 and a linear field
    V = B x gives -coeff tr(B) int N_i (the row sums of the unit mass matrix),
    exact for 2-point Gauss on bilinear and trilinear cells;
  - with no face flagged, a constant field (1, 0) on a box mesh gives
    coeff int dN_i/dx: -h_y w_j on the left boundary, +h_y w_j on the right,
    0 inside (w_j = 1, 1/2 at the corners);
  - the velocity-correction right-hand side with a linear pressure a . x is
    M V_k - dt a_k int N_i, on perturbed meshes too.
A dropped term, a wrong sign or normal, a transposed gradient product or a
wrong Gauss point fails one of these."""
import numpy as np
import pytest

from oracle import FemOracle
from pfem2dealii_amd import workloads as W
from tests.harness import perturbed_mesh


def _k1(n, h):
    K = np.zeros((n + 1, n + 1))
    for e in range(n):
        K[e:e + 2, e:e + 2] += np.array([[1.0, -1.0], [-1.0, 1.0]]) / h
    return K


def _m1(n, h):
    M = np.zeros((n + 1, n + 1))
    for e in range(n):
        M[e:e + 2, e:e + 2] += h * np.array([[2.0, 1.0], [1.0, 2.0]]) / 6.0
    return M


def _tensor(mesh, kind):
    """Laplacian or mass matrix of a box mesh in vertex numbering (x fastest)."""
    n, h = mesh["n"], mesh["h"]
    K = [_k1(n[k], h[k]) for k in range(mesh["dim"])]
    M = [_m1(n[k], h[k]) for k in range(mesh["dim"])]
    if mesh["dim"] == 2:
        if kind == "mass":
            return np.kron(M[1], M[0])
        return np.kron(M[1], K[0]) + np.kron(K[1], M[0])
    if kind == "mass":
        return np.kron(M[2], np.kron(M[1], M[0]))
    return (np.kron(M[2], np.kron(M[1], K[0])) + np.kron(M[2], np.kron(K[1], M[0]))
            + np.kron(K[2], np.kron(M[1], M[0])))


def _renumbered(mesh, seed):
    """The mesh with DoFs numbered by a random permutation of the vertices."""
    perm = np.random.default_rng(seed).permutation(mesh["X"].shape[1]).astype(np.int32)
    m = dict(mesh)
    m["cell_dofs"] = perm[mesh["cell_vertices"]]
    return m, perm


def _all_faces(mesh):
    """Bit f = 2k + side for every cell face on the box boundary."""
    n, dim = mesh["n"], mesh["dim"]
    nc = int(np.prod(n))
    ids = np.arange(nc)
    fm = np.zeros(nc, dtype=np.uint8)
    rem = ids.copy()
    for k in range(dim):
        ck = rem % n[k]
        rem = rem // n[k]
        fm |= np.where(ck == 0, 1 << (2 * k), 0).astype(np.uint8)
        fm |= np.where(ck == n[k] - 1, 1 << (2 * k + 1), 0).astype(np.uint8)
    return fm


BOXES = [((6, 5), (0.0, -1.0), (1.5, 1.0)), ((3, 4, 2), (0.0, 0.0, 0.0), (1.0, 0.8, 0.5))]
RHO, DT = 1.3, 0.02


@pytest.mark.parametrize("n,lo,hi", BOXES)
def test_matrices_equal_tensor_products(n, lo, hi):
    mesh = W.structured_mesh(n, lo, hi)
    m, perm = _renumbered(mesh, 5)
    fo = FemOracle(m)
    nd = m["n_dofs"]
    V = np.zeros((mesh["dim"], nd))
    A = fo.dense(fo.pressure(V, RHO, DT)[0])
    Mv = fo.dense(fo.velocity_correction(V, np.zeros(nd), RHO, DT)[0])
    inv = np.argsort(perm)
    for got, kind, scale in ((A, "laplace", 1.0), (Mv, "mass", RHO)):
        want = scale * _tensor(mesh, kind)
        back = got[np.ix_(perm, perm)]  # vertex numbering
        assert np.abs(back - want).max() <= 1e-13 * np.abs(want).max(), kind
    # the pattern is exactly the DoFs that share a cell
    rp, cols = fo.pattern()
    near = (np.abs(_tensor(mesh, "mass")) > 0)[np.ix_(inv, inv)]
    for r in range(nd):
        assert np.array_equal(cols[rp[r]:rp[r + 1]], np.nonzero(near[r])[0])


@pytest.mark.parametrize("n,lo,hi", BOXES)
def test_perturbed_laplacian_row_sums_and_symmetry(n, lo, hi):
    mesh = perturbed_mesh(W.structured_mesh(n, lo, hi), 0.25, seed=9)
    fo = FemOracle(mesh)
    A = fo.dense(fo.pressure(np.zeros((mesh["dim"], mesh["n_dofs"])), RHO, DT)[0])
    assert np.array_equal(A, A.T)
    assert np.abs(A.sum(axis=1)).max() <= 1e-13 * np.abs(np.diag(A)).max()
    assert (np.diag(A) > 0).all()


@pytest.mark.parametrize("n,lo,hi", BOXES)
def test_pressure_rhs_is_the_discrete_divergence_theorem(n, lo, hi):
    mesh = perturbed_mesh(W.structured_mesh(n, lo, hi), 0.25, seed=4)
    dim, nd = mesh["dim"], mesh["n_dofs"]
    fo = FemOracle(mesh)
    faces = _all_faces(mesh)
    coeff = RHO / DT
    c = np.array([0.7, -1.1, 0.4][:dim])
    _, b = fo.pressure(np.repeat(c[:, None], nd, axis=1), RHO, DT, faces)
    lumped = fo.dense(fo.velocity_correction(np.zeros((dim, nd)), np.zeros(nd), 1.0, DT)[0]).sum(axis=1)
    scale = coeff * np.abs(c).max() * lumped.max() / min(mesh["h"])
    assert np.abs(b).max() <= 1e-12 * scale
    B = np.array([[0.3, -0.8, 0.2], [1.1, 0.5, -0.4], [0.6, 0.9, -0.2]])[:dim, :dim]
    V = B @ mesh["X"]
    _, b = fo.pressure(V, RHO, DT, faces)
    want = -coeff * np.trace(B) * lumped
    assert np.abs(b - want).max() <= 1e-12 * np.abs(want).max()
    # only part of the boundary flagged: the residual is the unflagged faces' flux
    _, b0 = fo.pressure(V, RHO, DT, None)
    assert np.abs(b0 - want).max() > 1e-3 * np.abs(want).max()


def test_pressure_rhs_without_faces_is_the_boundary_flux():
    n, lo, hi = (6, 5), (0.0, -1.0), (1.5, 1.0)
    mesh = W.structured_mesh(n, lo, hi)
    fo = FemOracle(mesh)
    nd = mesh["n_dofs"]
    V = np.zeros((2, nd))
    V[0] = 1.0
    _, b = fo.pressure(V, RHO, DT)
    hy = mesh["h"][1]
    i = np.arange(nd) % (n[0] + 1)
    j = np.arange(nd) // (n[0] + 1)
    w = np.where((j == 0) | (j == n[1]), 0.5, 1.0)
    want = (RHO / DT) * hy * w * np.where(i == 0, -1.0, np.where(i == n[0], 1.0, 0.0))
    assert np.abs(b - want).max() <= 1e-12 * np.abs(want).max()


@pytest.mark.parametrize("n,lo,hi", BOXES)
def test_velocity_correction_rhs_with_a_linear_pressure(n, lo, hi):
    mesh = perturbed_mesh(W.structured_mesh(n, lo, hi), 0.25, seed=6)
    dim, nd = mesh["dim"], mesh["n_dofs"]
    fo = FemOracle(mesh)
    rng = np.random.default_rng(3)
    V = rng.standard_normal((dim, nd))
    a = np.array([2.0, -0.5, 1.5][:dim])
    P = a @ mesh["X"] + 0.25
    vals, b = fo.velocity_correction(V, P, RHO, DT)
    M = fo.dense(vals)
    lumped = M.sum(axis=1) / RHO
    for k in range(dim):
        want = M @ V[k] - DT * a[k] * lumped
        assert np.abs(b[k] - want).max() <= 1e-12 * np.abs(M).max() * (np.abs(V).max() + DT * np.abs(a).max())
    # a constant pressure leaves the mass term alone
    _, b = fo.velocity_correction(V, np.full(nd, 3.0), RHO, DT)
    for k in range(dim):
        assert np.abs(b[k] - M @ V[k]).max() <= 1e-13 * np.abs(M).max() * np.abs(V).max() * 8
