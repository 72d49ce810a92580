"""The FEM assembly oracle (oracle/fem_oracle.cpp, src/pfem2fem.cpp:246-517)
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
    M V_k - dt a_k int N_i, on perturbed meshes too;
  - the velocity-prediction matrices are rho M + mu dt (4/3 on the
    component's own second derivative, 1 on the others) as tensor products,
    and with every boundary face flagged the boundary terms close the
    viscous traction: linear fields give rho M applied to the field plus
    boundary fluxes in closed form, in 2D (matrix and right-hand side terms)
    and 3D (right-hand side terms, with the reference's first velocity
    component taken as zero there).
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


# ------------------------------------------------- velocity prediction


def _flux(mesh, k):
    """int N_i n_k over the box boundary per vertex: -/+ the face measure
    share of vertex i on the low/high k-faces (tensor weights, 1/2 at face
    edges)."""
    n, h, dim = mesh["n"], mesh["h"], mesh["dim"]
    nv = mesh["X"].shape[1]
    ids = np.arange(nv)
    coord, rem = [], ids.copy()
    for d in range(dim):
        coord.append(rem % (n[d] + 1))
        rem = rem // (n[d] + 1)
    w = np.ones(nv)
    for d in range(dim):
        if d == k:
            continue
        w = w * h[d] * np.where((coord[d] == 0) | (coord[d] == n[d]), 0.5, 1.0)
    return w * np.where(coord[k] == 0, -1.0, np.where(coord[k] == n[k], 1.0, 0.0))


def _kron_visc(mesh, comp):
    n, h = mesh["n"], mesh["h"]
    dim = mesh["dim"]
    K = [_k1(n[k], h[k]) for k in range(dim)]
    M = [_m1(n[k], h[k]) for k in range(dim)]
    terms = []
    for d in range(dim):  # d/dx_d d/dx_d term
        mats = [K[e] if e == d else M[e] for e in range(dim)]
        t = mats[0]
        for e in range(1, dim):
            t = np.kron(mats[e], t)
        terms.append((4.0 / 3.0 if d == comp else 1.0) * t)
    return sum(terms)


MU = 0.7


@pytest.mark.parametrize("n,lo,hi", BOXES)
def test_prediction_matrices_equal_tensor_products(n, lo, hi):
    mesh = W.structured_mesh(n, lo, hi)
    m, perm = _renumbered(mesh, 8)
    fo = FemOracle(m)
    dim, nd = mesh["dim"], m["n_dofs"]
    vals, _ = fo.prediction(np.zeros((dim, nd)), np.zeros((dim, nd)), RHO, MU, DT)
    Mm = _tensor(mesh, "mass")
    for k in range(dim):
        want = RHO * Mm + MU * DT * _kron_visc(mesh, k)
        got = fo.dense(vals[k])[np.ix_(perm, perm)]
        assert np.abs(got - want).max() <= 1e-13 * np.abs(want).max(), k


def test_prediction_2d_boundary_terms_close_the_traction():
    """With every boundary face flagged, the face terms cancel the volume
    terms' boundary flux: A^k x = rho M x (for both components, also on a
    perturbed mesh), and A^0 y = rho M y + mu dt int N_i n_y (no x-normal
    term for y); for V = (c x + d y, a x + b y) the right-hand sides are
    b^0 = rho M Vo_0 - mu dt a flux_y and b^1 = rho M Vo_1 + 2/3 mu dt c flux_y."""
    box = W.structured_mesh((6, 5), (0.0, -1.0), (1.5, 1.0))
    for mesh in (box, perturbed_mesh(box, 0.25, seed=3)):
        fo = FemOracle(mesh)
        nd = mesh["n_dofs"]
        faces = _all_faces(mesh)
        rng = np.random.default_rng(5)
        Vo = rng.standard_normal((2, nd))
        x, y = mesh["X"]
        a, b, c, d = 0.3, -0.8, 1.1, 0.5
        V = np.stack([c * x + d * y + 0.2, a * x + b * y - 0.4])
        vals, rhs = fo.prediction(Vo, V, RHO, MU, DT, faces)
        Mm = fo.dense(fo.velocity_correction(np.zeros((2, nd)), np.zeros(nd), 1.0, DT)[0])
        scale = RHO * np.abs(Mm).max() * 4 + MU * DT * 8
        for k in range(2):
            Ak = fo.dense(vals[k])
            assert np.abs(Ak @ x - RHO * Mm @ x).max() <= 1e-12 * scale, k
        if mesh is box:
            fy = _flux(mesh, 1)
            assert np.abs(fo.dense(vals[0]) @ y - (RHO * Mm @ y + MU * DT * fy)).max() <= 1e-12 * scale
            want0 = RHO * Mm @ Vo[0] - MU * DT * a * fy
            want1 = RHO * Mm @ Vo[1] + 2.0 / 3.0 * MU * DT * c * fy
            assert np.abs(rhs[0] - want0).max() <= 1e-12 * scale
            assert np.abs(rhs[1] - want1).max() <= 1e-12 * scale


def test_prediction_3d_boundary_terms_close_the_traction():
    """3D, every boundary face flagged, V = (U0, V1, V2) with V1, V2 linear:
    b^0 = rho M Vo_0, b^1 = rho M Vo_1 + mu dt (a1 fx + 4/3 b1 fy + c1 fz),
    b^2 = rho M Vo_2 + mu dt (a2 fx + b2 fy + 4/3 c2 fz) for U0 = 0; a linear
    U0 = c x + d y + e z then changes only the volume cross terms
    (b^1 by -mu dt (d fx - 2/3 c fy), b^2 by -mu dt (e fx - 2/3 c fz)): the
    reference's boundary traction takes the first velocity component as zero."""
    mesh = W.structured_mesh((3, 4, 2), (0.0, 0.0, 0.0), (1.0, 0.8, 0.5))
    fo = FemOracle(mesh)
    nd = mesh["n_dofs"]
    faces = _all_faces(mesh)
    Vo = np.random.default_rng(2).standard_normal((3, nd))
    x, y, z = mesh["X"]
    a1, b1, c1, a2, b2, c2 = 0.3, -0.8, 0.2, 1.1, 0.5, -0.4
    V = np.stack([np.zeros(nd), a1 * x + b1 * y + c1 * z, a2 * x + b2 * y + c2 * z + 1.0])
    _, rhs = fo.prediction(Vo, V, RHO, MU, DT, faces)
    Mm = fo.dense(fo.velocity_correction(np.zeros((3, nd)), np.zeros(nd), 1.0, DT)[0])
    fx, fy, fz = (_flux(mesh, k) for k in range(3))
    scale = RHO * np.abs(Mm).max() * 8 + MU * DT
    want = [RHO * Mm @ Vo[0],
            RHO * Mm @ Vo[1] + MU * DT * (a1 * fx + 4.0 / 3.0 * b1 * fy + c1 * fz),
            RHO * Mm @ Vo[2] + MU * DT * (a2 * fx + b2 * fy + 4.0 / 3.0 * c2 * fz)]
    for k in range(3):
        assert np.abs(rhs[k] - want[k]).max() <= 1e-12 * scale, k
    c, d, e = 0.9, -0.6, 0.35
    V2 = V.copy()
    V2[0] = c * x + d * y + e * z
    _, rhs2 = fo.prediction(Vo, V2, RHO, MU, DT, faces)
    assert np.abs(rhs2[0] - rhs[0]).max() <= 1e-12 * scale
    assert np.abs(rhs2[1] - rhs[1] + MU * DT * (d * fx - 2.0 / 3.0 * c * fy)).max() <= 1e-12 * scale
    assert np.abs(rhs2[2] - rhs[2] + MU * DT * (e * fx - 2.0 / 3.0 * c * fz)).max() <= 1e-12 * scale
