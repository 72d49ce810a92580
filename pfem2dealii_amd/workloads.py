"""Synthetic inputs of the benchmark shapes: structured Q1 meshes, nodal fields,
seeded particle sets and CFL time steps.

All randomness comes from a counter-based splitmix64 stream written here, so
every array is a pure function of (seed, index) and is identical on any host.

Mesh numbering follows deal.II's lexicographic conventions for structured
meshes: vertex (i, j, k) has id i + (nx+1)*(j + (ny+1)*k), cell (i, j, k) has id
i + nx*(j + ny*k), and the 2^dim vertices of a cell are listed with bit 0 ->
+x, bit 1 -> +y, bit 2 -> +z (GeometryInfo<dim> vertex order). The vertex id
doubles as the Q1 DoF id.
"""
import math

import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(counter):
    """splitmix64 output for an array of uint64 counters."""
    with np.errstate(over="ignore"):
        z = (np.asarray(counter, dtype=np.uint64) + np.uint64(0x9E3779B97F4A7C15)) & _M64
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed, stream, n, offset=0):
    """n doubles in [0, 1) from stream `stream` of seed `seed`."""
    base = (np.uint64(seed) << np.uint64(40)) ^ (np.uint64(stream) << np.uint64(32))
    c = base + np.arange(offset, offset + n, dtype=np.uint64)
    return (splitmix64(c) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def poisson(seed, stream, n, mean, cap):
    """Poisson(mean) counts clipped to [0, cap] by inverse-CDF on uniform()."""
    u = uniform(seed, stream, n)
    pmf = [math.exp(-mean)]
    for k in range(1, cap + 1):
        pmf.append(pmf[-1] * mean / k)
    cdf = np.cumsum(pmf)
    return np.minimum(np.searchsorted(cdf, u, side="right"), cap).astype(np.int32)


def structured_mesh(n, lo, hi):
    """Q1 box mesh with n[k] cells per direction on [lo, hi]."""
    dim = len(n)
    nvx = [n[k] + 1 for k in range(dim)]
    axes = []
    for k in range(dim):
        h = (hi[k] - lo[k]) / n[k]
        axes.append(lo[k] + np.arange(nvx[k], dtype=np.float64) * h)
    grids = np.meshgrid(*axes, indexing="ij")
    # vertex id = i + nvx0*(j + nvx1*k): x fastest
    X = np.stack([g.transpose(tuple(reversed(range(dim)))).reshape(-1) for g in grids])
    idx = [np.arange(n[k], dtype=np.int64) for k in range(dim)]
    cg = np.meshgrid(*idx, indexing="ij")
    cg = [g.transpose(tuple(reversed(range(dim)))).reshape(-1) for g in cg]  # cell ids x fastest
    ncell = int(np.prod(n))
    cv = np.empty((ncell, 1 << dim), dtype=np.int32)
    for b in range(1 << dim):
        vid = np.zeros(ncell, dtype=np.int64)
        stride = 1
        for k in range(dim):
            vid += (cg[k] + ((b >> k) & 1)) * stride
            stride *= nvx[k]
        cv[:, b] = vid
    h = [(hi[k] - lo[k]) / n[k] for k in range(dim)]
    return {"dim": dim, "n": list(n), "lo": list(lo), "hi": list(hi), "h": h, "X": X,
            "cell_vertices": cv, "cell_dofs": cv, "n_dofs": X.shape[1]}


# ---- nodal velocity fields ----

def poiseuille(mesh, amplitude=1.0):
    X = mesh["X"]
    y = X[1]
    V = np.zeros_like(X)
    V[0] = amplitude * (4.0 * y * (1.0 - y))
    return V


def taylor_green(mesh):
    X = mesh["X"]
    V = np.zeros_like(X)
    V[0] = np.sin(X[0]) * np.cos(X[1])
    V[1] = -np.cos(X[0]) * np.sin(X[1])
    return V


def abc_flow(mesh, A=1.0, B=0.7, C=0.4):
    X = mesh["X"] * (2.0 * math.pi)
    x, y, z = X
    V = np.empty_like(X)
    V[0] = A * np.sin(z) + C * np.cos(y)
    V[1] = B * np.sin(x) + A * np.cos(z)
    V[2] = C * np.sin(y) + B * np.cos(x)
    return V


def cfl_dt(mesh, V, cfl=0.5):
    """dt = cfl * h_min / max_nodes |u|_2."""
    hmin = min(mesh["h"])
    umax = float(np.sqrt((V * V).sum(axis=0)).max())
    return cfl * hmin / umax


# ---- particle seeding (positions in real coordinates, grouped by cell) ----

def _cell_origin_h(mesh, cells):
    X = mesh["X"]
    cv = mesh["cell_vertices"]
    dim = mesh["dim"]
    X0 = np.stack([X[k][cv[cells, 0]] for k in range(dim)])
    h = np.stack([X[k][cv[cells, 1 << k]] - X[k][cv[cells, 0]] for k in range(dim)])
    return X0, h


def seed_from_unit(mesh, cells, xi):
    """x = X0 + xi*h on box cells (harness placement; the library locates it)."""
    X0, h = _cell_origin_h(mesh, cells)
    return X0 + xi * h


def seed_jittered_lattice(mesh, q, jitter, seed):
    """q[k] particles per direction per cell at sub-cell centres + U(-j, j) sub-cell widths."""
    dim = mesh["dim"]
    nc = mesh["cell_vertices"].shape[0]
    ppc = int(np.prod(q))
    cells = np.repeat(np.arange(nc, dtype=np.int32), ppc)
    lat = np.stack(np.meshgrid(*[np.arange(q[k]) for k in range(dim)], indexing="ij")).reshape(dim, -1)
    xi = np.empty((dim, nc * ppc))
    for k in range(dim):
        u = uniform(seed, k, nc * ppc)
        xi[k] = (np.tile(lat[k], nc) + 0.5 + (2.0 * u - 1.0) * jitter) / q[k]
    return seed_from_unit(mesh, cells, xi), cells


def seed_uniform(mesh, counts, seed):
    """counts[c] particles at U[0,1)^dim reference positions in each cell c."""
    dim = mesh["dim"]
    counts = np.asarray(counts, dtype=np.int64)
    cells = np.repeat(np.arange(counts.shape[0], dtype=np.int32), counts)
    n = cells.shape[0]
    xi = np.stack([uniform(seed, k, n) for k in range(dim)])
    return seed_from_unit(mesh, cells, xi), cells


# ---- the benchmark configurations ----

def config(name, scale=1.0):
    """Mesh, field, particles and parameters of a named workload.

    scale < 1 shrinks the mesh (per direction) for CPU-sized parity runs with
    the same physics; scale = 1 is the full size.
    """
    def sc(nn):
        return [max(2, int(round(a * scale))) for a in nn]

    if name == "channel2d":      # 64x32 Q1 on [0,2]x[0,1], jittered 4x4 lattice, Poiseuille, 1 step
        mesh = structured_mesh(sc([64, 32]), [0.0, 0.0], [2.0, 1.0])
        V = poiseuille(mesh)
        x, cells = seed_jittered_lattice(mesh, [4, 4], 0.1, seed=1)
        params = dict(n_sub=5, dt=cfl_dt(mesh, V), max_per_part=5, particles_per_cell=[4, 4])
        steps = 1
    elif name == "taylor_green2d":  # 2048x1024 Q1 on [0,2pi]x[0,pi], 16 random/cell, 10 steps
        mesh = structured_mesh(sc([2048, 1024]), [0.0, 0.0], [2.0 * math.pi, math.pi])
        V = taylor_green(mesh)
        nc = mesh["cell_vertices"].shape[0]
        x, cells = seed_uniform(mesh, np.full(nc, 16), seed=2)
        params = dict(n_sub=5, dt=cfl_dt(mesh, V), max_per_part=5, particles_per_cell=[4, 4])
        steps = 10
    elif name == "abc3d":        # 128^3 Q1 unit cube, Poisson(8) clipped to [0,32], ABC, 10 steps
        mesh = structured_mesh(sc([128, 128, 128]), [0.0, 0.0, 0.0], [1.0, 1.0, 1.0])
        V = abc_flow(mesh)
        nc = mesh["cell_vertices"].shape[0]
        counts = poisson(3, 7, nc, 8.0, 32)
        x, cells = seed_uniform(mesh, counts, seed=3)
        params = dict(n_sub=5, dt=cfl_dt(mesh, V), max_per_part=5, particles_per_cell=[2, 2, 2])
        steps = 10
    elif name == "adjoint2d":    # 256x128 Q1 on [0,2]x[0,1], 16 random/cell, Poiseuille a=1
        mesh = structured_mesh(sc([256, 128]), [0.0, 0.0], [2.0, 1.0])
        V = poiseuille(mesh, 1.0)
        nc = mesh["cell_vertices"].shape[0]
        x, cells = seed_uniform(mesh, np.full(nc, 16), seed=5)
        params = dict(n_sub=5, dt=cfl_dt(mesh, V), max_per_part=5, particles_per_cell=[4, 4])
        steps = 20
    elif name == "adjoint3d":    # 16^3 Q1 unit cube, 8 random/cell, ABC, 10 steps (3D adjoint check)
        mesh = structured_mesh(sc([16, 16, 16]), [0.0, 0.0, 0.0], [1.0, 1.0, 1.0])
        V = abc_flow(mesh)
        nc = mesh["cell_vertices"].shape[0]
        x, cells = seed_uniform(mesh, np.full(nc, 8), seed=6)
        params = dict(n_sub=5, dt=cfl_dt(mesh, V), max_per_part=5, particles_per_cell=[2, 2, 2])
        steps = 10
    else:
        raise KeyError(name)
    return {"name": name, "mesh": mesh, "V": V, "x": x, "cells": cells, "params": params, "steps": steps}


def batched_instances(n_instances=256, seed=4, scale=1.0):
    """Independent 2D channel instances with log-uniform mesh sizes 32x16..512x256,
    Poisson(12) particles per cell and Poiseuille amplitude U[0.5, 2]."""
    out = []
    for i in range(n_instances):
        s = math.exp(uniform(seed + i, 0, 1)[0] * math.log(16.0))
        ny = max(2, int(round(16 * s * scale)))
        nx = 2 * ny
        mesh = structured_mesh([nx, ny], [0.0, 0.0], [2.0, 1.0])
        amp = 0.5 + 1.5 * uniform(seed + i, 1, 1)[0]
        V = poiseuille(mesh, amp)
        nc = mesh["cell_vertices"].shape[0]
        counts = poisson(seed + i, 2, nc, 12.0, 64)
        x, cells = seed_uniform(mesh, counts, seed=seed + i)
        params = dict(n_sub=5, dt=cfl_dt(mesh, V), max_per_part=5, particles_per_cell=[4, 4])
        out.append({"name": "channel2d_%d" % i, "mesh": mesh, "V": V, "x": x, "cells": cells,
                    "params": params, "steps": 20})
    return out
