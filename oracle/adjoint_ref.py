"""FP64 CPU reference of the differentiated advect + project model and of its
gradient (test infrastructure only).

The model is the one the GPU adjoint differentiates, with every discrete
choice frozen: the cell of each particle at each substep (recorded from the
unperturbed forward run and reused for perturbed evaluations), the survivor
set and refill list of each redistribution (the identity maps returned by
pfem2gpu_move_particles_traced). Per step: n_sub explicit-Euler substeps
x <- x + dt u(x) with u = sum_i N_i(xi) V[dof_i] on box cells
(src/pfem2particlehandler.cpp:150-163), then survivors keep their state and
refills are placed at sub-cell centres with v = sum_i N_i(xi_c) V[dof_i]
(src/pfem2particlehandler.cpp:381-409); after the last step the projection
P = S / W (src/pfem2particlehandler.cpp:191-209) and L = 1/2 |P - target|^2.
Structured box meshes in 2D and 3D (the 2D adjoint workload and a 3D check).
"""
import numpy as np


def _shape(xi):
    """N_i(xi) = prod_k (bit_k(i) ? xi_k : 1 - xi_k), the product taken in k
    order (as the GPU's shape values), as [2^D][n]."""
    D = xi.shape[0]
    out = []
    for i in range(1 << D):
        v = xi[0] if (i & 1) else (1.0 - xi[0])
        for k in range(1, D):
            v = v * (xi[k] if (i >> k) & 1 else (1.0 - xi[k]))
        out.append(v)
    return np.stack(out)


def _dshape(xi):
    """dN_i/dxi_d as [i][d][n]."""
    D = xi.shape[0]
    out = []
    for i in range(1 << D):
        row = []
        for d in range(D):
            v = np.ones_like(xi[0])
            for k in range(D):
                if k == d:
                    v = v * (1.0 if (i >> k) & 1 else -1.0)
                else:
                    v = v * (xi[k] if (i >> k) & 1 else (1.0 - xi[k]))
            row.append(v)
        out.append(np.stack(row))
    return np.stack(out)


class FrozenModel:
    def __init__(self, mesh, params, x0, cells0, srcs, end_cells):
        self.D = int(mesh["dim"])
        assert self.D in (2, 3)
        self.NV = 1 << self.D
        self.mesh = mesh
        self.n = list(mesh["n"])
        self.X = mesh["X"]
        self.cv = mesh["cell_vertices"]
        self.n_sub = int(params["n_sub"])
        self.dt = float(params["dt"]) / float(self.n_sub)
        self.q = list(params["particles_per_cell"])[: self.D]
        self.hp = [1.0 / q for q in self.q]
        self.x0 = np.array(x0, dtype=np.float64)
        self.c0 = np.array(cells0, dtype=np.int64)
        self.srcs = [np.asarray(s, dtype=np.int64) for s in srcs]
        self.end_cells = [np.asarray(c, dtype=np.int64) for c in end_cells]
        self.paths = None

    # geometry of cells (box): vertex 0 and edge lengths
    def _X0h(self, c):
        v0 = self.cv[c, 0]
        X0 = self.X[:, v0]
        h = np.stack([self.X[k, self.cv[c, 1 << k]] - X0[k] for k in range(self.D)])
        return X0, h

    def _interp(self, V, c, xi):
        """u_k = sum_i N_i V_k[dof_i], vertex-outer, from 0.0"""
        N = _shape(xi)
        u = np.zeros((self.D, c.shape[0]))
        for i in range(self.NV):
            d = self.cv[c, i]
            for k in range(self.D):
                u[k] = u[k] + N[i] * V[k][d]
        return u

    def _fwd(self, c, xi):
        """x = sum_i N_i X_i (the forward map of a refill position)"""
        N = _shape(xi)
        x = np.zeros((self.D, c.shape[0]))
        for i in range(self.NV):
            for k in range(self.D):
                x[k] = x[k] + N[i] * self.X[k, self.cv[c, i]]
        return x

    def _parts(self, pp):
        """part indices (x slowest) of linear part numbers pp"""
        pk = [None] * self.D
        rem = pp.copy()
        for k in range(self.D - 1, -1, -1):
            Q = self.q[k] + 1
            pk[k] = rem % Q
            rem = rem // Q
        return pk

    def _locate(self, x, c):
        """current cell if it contains x (faces inclusive), else the cell that
        contains x on the structured grid; -1 outside the mesh"""
        X0, h = self._X0h(c)
        xi = (x - X0) / h
        keep = np.ones(c.shape[0], dtype=bool)
        for k in range(self.D):
            keep &= (xi[k] >= 0) & (xi[k] <= 1)
        lo = self.X[:, 0]
        hx = self.mesh["h"]
        out = np.zeros(c.shape[0], dtype=bool)
        cn = np.zeros(c.shape[0], dtype=np.int64)
        stride = 1
        for k in range(self.D):
            ik = np.floor((x[k] - lo[k]) / hx[k])
            ok = np.isfinite(x[k])
            ik = np.where(ok, ik, -1).astype(np.int64)
            out |= (ik < 0) | (ik >= self.n[k]) | ~ok
            cn += np.clip(ik, 0, self.n[k] - 1) * stride
            stride *= self.n[k]
        cn = np.where(out, -1, cn)
        return np.where(keep, c, cn)

    def forward(self, V, v0, record=False):
        """Final (x, v, cell) after all steps. record=True runs the locate and
        stores the cell, reference position and liveness of every particle at
        every substep (the frozen path); otherwise the stored path is reused."""
        D = self.D
        if record:
            self.paths = []
        x, v, c = self.x0.copy(), np.array(v0, dtype=np.float64).copy(), self.c0.copy()
        placed = None  # inserted particles: placed when at the exact image of their part centre
        for s, (src, ec) in enumerate(zip(self.srcs, self.end_cells)):
            alive = np.ones(c.shape[0], dtype=bool)
            cc = c.copy()
            path = []
            for k in range(self.n_sub):
                if not record:
                    cc, _, alive = self.paths[s][k]
                cs = np.where(alive, cc, 0)
                X0, h = self._X0h(cs)
                xi = (x - X0) / h
                if k == 0:
                    xi = self._snap(x, cs, xi) if placed is None else self._centre(xi, placed)
                if record:
                    path.append((cc.copy(), xi.copy(), alive.copy()))
                u = self._interp(V, cs, xi)
                x = np.where(alive, x + u * self.dt, x)
                if record:
                    nc_ = self._locate(x, cs)
                    alive = alive & (nc_ >= 0)
                    cc = np.where(alive, nc_, cc)
            if record:
                path.append((cc.copy(), None, alive.copy()))
                self.paths.append(path)
            else:
                cc, _, alive = self.paths[s][self.n_sub]
            # redistribution with the recorded identity map
            nq = src.shape[0]
            surv = src >= 0
            xn = np.zeros((D, nq))
            vn = np.zeros((D, nq))
            xn[:, surv] = x[:, src[surv]]
            vn[:, surv] = v[:, src[surv]]
            if record:
                assert alive[src[surv]].all(), "a survivor died in the reference"
                assert np.array_equal(cc[src[surv]], ec[surv]), "reference cells differ from the recorded ones"
            ref = ~surv
            if ref.any():
                pk = self._parts(-1 - src[ref])
                xi = np.stack([(pk[k] + 0.5) * self.hp[k] for k in range(D)])
                cr = ec[ref]
                xn[:, ref] = self._fwd(cr, xi)
                vn[:, ref] = self._interp(V, cr, xi)
            x, v, c = xn, vn, ec.copy()
            placed = ~surv  # refills sit at their part centre until located
        return x, v, c, placed

    def _centre(self, xi, placed):
        """the exact centre of the part holding xi where `placed` (a refill's
        stored reference position), xi elsewhere"""
        pk = [(xi[k] * self.q[k]).astype(np.int64) for k in range(self.D)]
        xc = np.stack([(pk[k] + 0.5) * self.hp[k] for k in range(self.D)])
        return np.where(placed, xc, xi)

    def _snap(self, x, c, xi):
        """exact sub-cell centre for inserted particles placed exactly at its
        image (the insertion rule)"""
        D = self.D
        pk = [(xi[k] * self.q[k]).astype(np.int64) for k in range(D)]
        xc = np.stack([(pk[k] + 0.5) * self.hp[k] for k in range(D)])
        F = self._fwd(c, xc)
        hit = np.ones(x.shape[1], dtype=bool)
        for k in range(D):
            hit &= (F[k] == x[k]) & (pk[k] >= 0) & (pk[k] < self.q[k])
        return np.where(hit, xc, xi)

    def project(self, x, v, c, placed=None):
        D = self.D
        X0, h = self._X0h(c)
        xi = (x - X0) / h
        xi = self._snap(x, c, xi) if placed is None else self._centre(xi, placed)
        N = _shape(xi)
        nd = self.X.shape[1]
        S = np.zeros((D, nd))
        W = np.zeros(nd)
        for i in range(self.NV):
            d = self.cv[c, i]
            for k in range(D):
                S[k] += np.bincount(d, weights=N[i] * v[k], minlength=nd)
            W += np.bincount(d, weights=N[i], minlength=nd)
        return S / W, W, xi

    def loss(self, V, v0, target):
        x, v, c, placed = self.forward(V, v0, record=self.paths is None)
        P, _, _ = self.project(x, v, c, placed)
        r = P - target
        return 0.5 * float((r * r).sum())

    def grad(self, V, v0, target):
        """(L, dL/dv0, dL/dV) by the reverse sweep over the recorded path."""
        D = self.D
        x, v, c, placed = self.forward(V, v0, record=True)
        P, W, xi = self.project(x, v, c, placed)
        r = P - target
        L = 0.5 * float((r * r).sum())
        N = _shape(xi)
        dN = _dshape(xi)
        _, h = self._X0h(c)
        nd = self.X.shape[1]
        av = np.zeros_like(v)
        ax = np.zeros_like(x)
        for i in range(self.NV):
            d = self.cv[c, i]
            a = r[:, d] / W[d]
            av += N[i] * a
            coef = (a * (v - P[:, d])).sum(axis=0)
            ax += dN[i] * coef
        ax = ax / h
        gV = np.zeros((D, nd))
        for s in range(len(self.srcs) - 1, -1, -1):
            src, ec = self.srcs[s], self.end_cells[s]
            path = self.paths[s]
            n_start = path[0][0].shape[0]
            xs = np.zeros((D, n_start))
            vs = np.zeros((D, n_start))
            surv = src >= 0
            xs[:, src[surv]] = ax[:, surv]
            vs[:, src[surv]] = av[:, surv]
            ref = ~surv
            if ref.any():
                pk = self._parts(-1 - src[ref])
                xi_r = np.stack([(pk[k] + 0.5) * self.hp[k] for k in range(D)])
                Nr = _shape(xi_r)
                cr = ec[ref]
                for i in range(self.NV):
                    dd = self.cv[cr, i]
                    for k in range(D):
                        gV[k] += np.bincount(dd, weights=Nr[i] * av[k, ref], minlength=nd)
            alive_end = path[self.n_sub][2]
            a = np.where(alive_end, xs, 0.0)
            for k in range(self.n_sub - 1, -1, -1):
                cc, xik, alive = path[k]
                cc = np.where(alive, cc, 0)
                Nk = _shape(xik)
                dNk = _dshape(xik)
                _, hk = self._X0h(cc)
                gx = np.zeros_like(a)
                for i in range(self.NV):
                    dd = self.cv[cc, i]
                    m = alive_end
                    for kk in range(D):
                        gV[kk] += np.bincount(dd[m], weights=self.dt * Nk[i][m] * a[kk][m], minlength=nd)
                    av_i = a[0] * V[0][dd]
                    for kk in range(1, D):
                        av_i = av_i + a[kk] * V[kk][dd]
                    gx += dNk[i] * av_i
                a = np.where(alive_end, a + self.dt * gx / hk, 0.0)
            ax, av = a, vs
        return L, av, gV
