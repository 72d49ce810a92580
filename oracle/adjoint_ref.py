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
2D structured box meshes only (the adjoint workload).
"""
import numpy as np


def _shape(xi):
    a, b = xi[0], xi[1]
    l0a, l0b = 1.0 - a, 1.0 - b
    return np.stack([l0a * l0b, a * l0b, l0a * b, a * b])


def _dshape(xi):
    """dN_i/dxi_d as [i][d][n]."""
    a, b = xi[0], xi[1]
    return np.stack([np.stack([-(1.0 - b), -(1.0 - a)]), np.stack([1.0 - b, -a]),
                     np.stack([-b, 1.0 - a]), np.stack([b, a])])


class FrozenModel:
    def __init__(self, mesh, params, x0, cells0, srcs, end_cells):
        assert mesh["dim"] == 2
        self.mesh = mesh
        self.nx, self.ny = mesh["n"]
        self.X = mesh["X"]
        self.cv = mesh["cell_vertices"]
        self.n_sub = int(params["n_sub"])
        self.dt = float(params["dt"]) / float(self.n_sub)
        self.q = list(params["particles_per_cell"])
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
        h = np.stack([self.X[0, self.cv[c, 1]] - X0[0], self.X[1, self.cv[c, 2]] - X0[1]])
        return X0, h

    def _interp(self, V, c, xi):
        N = _shape(xi)
        u = np.zeros((2, c.shape[0]))
        for i in range(4):
            d = self.cv[c, i]
            u[0] = u[0] + N[i] * V[0][d]
            u[1] = u[1] + N[i] * V[1][d]
        return u

    def _locate(self, x, c):
        """current cell if it contains x (faces inclusive), else the cell that
        contains x on the structured grid; -1 outside the mesh"""
        X0, h = self._X0h(c)
        xi = (x - X0) / h
        keep = (xi[0] >= 0) & (xi[0] <= 1) & (xi[1] >= 0) & (xi[1] <= 1)
        lo = self.X[:, 0]
        hx = self.mesh["h"]
        i = np.floor((x[0] - lo[0]) / hx[0]).astype(np.int64)
        j = np.floor((x[1] - lo[1]) / hx[1]).astype(np.int64)
        out = (i < 0) | (i >= self.nx) | (j < 0) | (j >= self.ny) | ~np.isfinite(x[0]) | ~np.isfinite(x[1])
        cn = np.where(out, -1, i + self.nx * np.clip(j, 0, self.ny - 1))
        return np.where(keep, c, cn)

    def forward(self, V, v0, record=False):
        """Final (x, v, cell) after all steps. record=True runs the locate and
        stores the cell, reference position and liveness of every particle at
        every substep (the frozen path); otherwise the stored path is reused."""
        if record:
            self.paths = []
        x, v, c = self.x0.copy(), np.array(v0, dtype=np.float64).copy(), self.c0.copy()
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
                    xi = self._snap(x, cs, xi)
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
            xn = np.zeros((2, nq))
            vn = np.zeros((2, nq))
            xn[:, surv] = x[:, src[surv]]
            vn[:, surv] = v[:, src[surv]]
            if record:
                assert alive[src[surv]].all(), "a survivor died in the reference"
                assert np.array_equal(cc[src[surv]], ec[surv]), "reference cells differ from the recorded ones"
            ref = ~surv
            if ref.any():
                pp = -1 - src[ref]
                Q1 = self.q[1] + 1
                pi0, pi1 = pp // Q1, pp % Q1
                xi = np.stack([(pi0 + 0.5) * self.hp[0], (pi1 + 0.5) * self.hp[1]])
                cr = ec[ref]
                N = _shape(xi)
                xr = np.zeros((2, cr.shape[0]))
                for i in range(4):
                    xr[0] = xr[0] + N[i] * self.X[0, self.cv[cr, i]]
                    xr[1] = xr[1] + N[i] * self.X[1, self.cv[cr, i]]
                xn[:, ref] = xr
                vn[:, ref] = self._interp(V, cr, xi)
            x, v, c = xn, vn, ec.copy()
        return x, v, c

    def _snap(self, x, c, xi):
        """exact sub-cell centre for particles placed there (refills)"""
        pk0 = (xi[0] * self.q[0]).astype(np.int64)
        pk1 = (xi[1] * self.q[1]).astype(np.int64)
        xc = np.stack([(pk0 + 0.5) * self.hp[0], (pk1 + 0.5) * self.hp[1]])
        N = _shape(xc)
        F = np.zeros_like(x)
        for i in range(4):
            F[0] = F[0] + N[i] * self.X[0, self.cv[c, i]]
            F[1] = F[1] + N[i] * self.X[1, self.cv[c, i]]
        hit = (F[0] == x[0]) & (F[1] == x[1]) & (pk0 >= 0) & (pk0 < self.q[0]) & (pk1 >= 0) & (pk1 < self.q[1])
        return np.where(hit, xc, xi)

    def project(self, x, v, c):
        X0, h = self._X0h(c)
        xi = (x - X0) / h
        xi = self._snap(x, c, xi)
        N = _shape(xi)
        nd = self.X.shape[1]
        S = np.zeros((2, nd))
        W = np.zeros(nd)
        for i in range(4):
            d = self.cv[c, i]
            np.add.at(S[0], d, N[i] * v[0])
            np.add.at(S[1], d, N[i] * v[1])
            np.add.at(W, d, N[i])
        return S / W, W, xi

    def loss(self, V, v0, target):
        x, v, c = self.forward(V, v0, record=self.paths is None)
        P, _, _ = self.project(x, v, c)
        r = P - target
        return 0.5 * float((r * r).sum())

    def grad(self, V, v0, target):
        """(L, dL/dv0, dL/dV) by the reverse sweep over the recorded path."""
        x, v, c = self.forward(V, v0, record=True)
        P, W, xi = self.project(x, v, c)
        r = P - target
        L = 0.5 * float((r * r).sum())
        N = _shape(xi)
        dN = _dshape(xi)
        _, h = self._X0h(c)
        nd = self.X.shape[1]
        av = np.zeros_like(v)
        ax = np.zeros_like(x)
        for i in range(4):
            d = self.cv[c, i]
            a = r[:, d] / W[d]
            av += N[i] * a
            coef = (a * (v - P[:, d])).sum(axis=0)
            ax += dN[i] * coef
        ax = ax / h
        gV = np.zeros((2, nd))
        for s in range(len(self.srcs) - 1, -1, -1):
            src, ec = self.srcs[s], self.end_cells[s]
            path = self.paths[s]
            n_start = path[0][0].shape[0]
            xs = np.zeros((2, n_start))
            vs = np.zeros((2, n_start))
            surv = src >= 0
            xs[:, src[surv]] = ax[:, surv]
            vs[:, src[surv]] = av[:, surv]
            ref = ~surv
            if ref.any():
                pp = -1 - src[ref]
                Q1 = self.q[1] + 1
                xi_r = np.stack([(pp // Q1 + 0.5) * self.hp[0], (pp % Q1 + 0.5) * self.hp[1]])
                Nr = _shape(xi_r)
                cr = ec[ref]
                for i in range(4):
                    dd = self.cv[cr, i]
                    np.add.at(gV[0], dd, Nr[i] * av[0, ref])
                    np.add.at(gV[1], dd, Nr[i] * av[1, ref])
            alive_end = path[self.n_sub][2]
            a = np.where(alive_end, xs, 0.0)
            for k in range(self.n_sub - 1, -1, -1):
                cc, xik, alive = path[k]
                cc = np.where(alive, cc, 0)
                Nk = _shape(xik)
                dNk = _dshape(xik)
                _, hk = self._X0h(cc)
                gx = np.zeros_like(a)
                for i in range(4):
                    dd = self.cv[cc, i]
                    m = alive_end
                    np.add.at(gV[0], dd[m], self.dt * Nk[i][m] * a[0][m])
                    np.add.at(gV[1], dd[m], self.dt * Nk[i][m] * a[1][m])
                    av_i = a[0] * V[0][dd] + a[1] * V[1][dd]
                    gx += dNk[i] * av_i
                a = np.where(alive_end, a + self.dt * gx / hk, 0.0)
            ax, av = a, vs
        return L, av, gV
