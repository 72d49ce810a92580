"""Benchmark of the particle hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload taylor_green2d]

Workload (default): 2048x1024 Q1 quads on [0,2pi]x[0,pi], 16 particles per
cell at seeded uniform positions (33.5M particles), Taylor-Green nodal
velocity, CFL 0.5, 5 Euler substeps, FP64. One step is the harness order
move (correct fused) -> project -> FEM stand-in (V_old := projected, V :=
prescribed) -> correct, with every input resident in HBM.

Run A times W warmup and K measured steps from the seeded state (HIP events
on the handle's stream around the K steps; particles counted at the start of
each step). Run B repeats the same W + K steps from the same state with HIP
events around every launch, for the per-kernel durations and the roofline of
the dominant kernel (algorithmic bytes per launch / average launch duration;
the traffic figure comes from the committed counter summary). The same two
runs on the 128^3 ABC configuration (16.8M particles, 3D) give the `abc3d`
object. Then the 256-instance batched form against plain launches (20
steps), the adjoint's cost against its forward, the pack and unpack of the
reference's particle records for half the workload's cells, and the CPU
baseline: the oracle in reference-order mode, one thread, one full step of
each configuration, timed on this host (CPU model reported). Prints one JSON
line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def alg_bytes_per_particle_step(dim, n_nodes, n_particles):
    """B_alg = 3*S_p + 24*d*(N_nodes/N_particles), S_p = 16*d + 4."""
    s_p = 16 * dim + 4
    return 3 * s_p + 24 * dim * (n_nodes / float(n_particles))


def kernel_alg_bytes(dim, n_nodes, n_in, n_out):
    """Algorithmic bytes of one step split over the kernels that carry them:
    the move reads and writes the particle state once (2 S_p per particle at
    the start of the step) and reads V and V_old (16 d per node); the
    redistribution/projection kernel reads the final state once (S_p per
    particle after the step); the nodal division writes the projected field
    (8 d per node). Redistribution and sorting traffic is overhead, not
    algorithmic, and counts for no kernel."""
    s_p = 16 * dim + 4
    return {"k_move": 2.0 * s_p * n_in + 16.0 * dim * n_nodes, "k_settle": float(s_p) * n_out,
            "k_node": 8.0 * dim * n_nodes}


def hbm_traffic(workload, scale, kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by what the same
    counter reads for a plain copy of known size), or None."""
    if scale != 1.0:
        return None
    path = os.path.join(ROOT, "profiles", "hbm_traffic_%s.txt" % workload)
    try:
        lines = open(path).read().splitlines()
    except OSError:
        return None
    for ln in lines:
        f = ln.split()
        if len(f) >= 3 and f[0] == "traffic_bytes_per_launch" and f[1] == kernel:
            return float(f[2])
    return None


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), or the platform string."""
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or platform.machine()


def cpu_baseline(workload, steps=1):
    """The oracle in reference-order mode (AoS vector, swap-with-back
    deletion, std::map redistribution, vector-order projection), one thread,
    on the full workload: `steps` full steps (move, project, correct) from
    the seeded state, particles counted at the start of each step."""
    from oracle import Oracle, MODE_REFERENCE
    from pfem2dealii_amd import workloads as W
    cfg = W.config(workload)
    o = Oracle(cfg["mesh"], cfg["params"], MODE_REFERENCE)
    o.initialize_maps()
    o.insert(cfg["x"], cfg["cells"])
    n_seeded = len(cfg["cells"])
    V = cfg["V"]
    o.set_velocity_from_field(V)
    del cfg["x"], cfg["cells"]
    psteps = 0
    t0 = time.perf_counter()
    for _ in range(steps):
        psteps += o.count()
        o.move(V)
        Vp = o.project()
        o.correct(V, Vp)
    dt = time.perf_counter() - t0
    o.close()
    m = cfg["mesh"]
    return {"value": psteps / dt, "unit": "particle-steps/s", "cores": 1, "kind": "port",
            "sample": "%s full size (%s cells, %d particles seeded), %d full step(s) in %.1f s, oracle "
                      "reference-order mode, g++ -O2 -ffp-contract=off, one thread"
                      % (workload, "x".join(str(a) for a in m["n"]), n_seeded, steps, dt),
            "cpu_model": cpu_model(), "host_nproc": os.cpu_count()}


def batched_comparison(steps, warmup, n_inst=256, scale=1.0, log=None, force_tune=False):
    """256 independent 2D channel instances (32x16..512x256 Q1, Poisson(12)
    particles per cell, Poiseuille amplitude U[0.5, 2]): one batch handle
    (one launch per phase, hipGraph per step, autotuned launch shape) against
    the same instances as plain handles launched one after another on one
    stream. Particles are counted at the start of each step."""
    import torch
    from pfem2dealii_amd import BatchHandler, ParticleHandler, workloads as W
    from pfem2dealii_amd.autotune import shipped, tune
    inst = W.batched_instances(n_inst, seed=4, scale=scale)
    Vh = np.concatenate([i["V"] for i in inst], axis=1)
    V = torch.tensor(Vh, dtype=torch.float64, device="cuda")
    Vp = torch.empty_like(V)

    def make():
        b = BatchHandler([i["mesh"] for i in inst], [i["params"] for i in inst])
        b.initialize_maps()
        co = b.cell_offsets
        x = np.concatenate([i["x"] for i in inst], axis=1)
        cells = np.concatenate([i["cells"] + co[k] for k, i in enumerate(inst)]).astype(np.int32)
        assert b.insert_particles(x, cells) == 0
        b.set_velocity_from_field(V)
        b.reserve_particles(3 * len(cells))
        return b

    sig = {"workload": "batched", "n": n_inst, "scale": scale}
    tuning = None if force_tune else shipped(sig)
    if tuning is None:  # on a scratch batch: tuning steps advance the simulation
        bt = make()
        tuning, _ = tune(bt, lambda: bt.step(V, Vp, use_graph=True), signature=sig, log=log, force=True)
        bt.close()
    b = make()
    b.set_tuning(**tuning)
    n0 = b.count()
    torch.cuda.synchronize()
    bstep = lambda: b.step(V, Vp, use_graph=True)  # noqa: E731
    bstep()
    for _ in range(warmup):
        bstep()
    stream = torch.cuda.ExternalStream(b.stream())
    counts = torch.zeros(steps, dtype=torch.int32, device="cuda")
    b.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for s in range(steps):
        b.count_async(counts.data_ptr() + 4 * s)
        bstep()
    e1.record(stream)
    b.synchronize()
    b_ms = e0.elapsed_time(e1)
    b_ps = int(counts.sum().item())
    b.close()
    del V, Vp

    pstream = torch.cuda.Stream()
    hs, Vs, Vps = [], [], []
    for i in inst:
        h = ParticleHandler(i["mesh"], i["params"], stream=pstream.cuda_stream)
        h.initialize_maps()
        h.insert_particles(i["x"], i["cells"])
        h.reserve_particles(3 * len(i["cells"]))
        Vi = torch.tensor(i["V"], dtype=torch.float64, device="cuda")
        h.set_velocity_from_field(Vi)
        hs.append(h)
        Vs.append(Vi)
        Vps.append(torch.empty_like(Vi))
    torch.cuda.synchronize()
    for _ in range(1 + warmup):
        for h, Vi, Vpi in zip(hs, Vs, Vps):
            h.step(Vi, Vpi, use_graph=False)
    pc = torch.zeros((steps, n_inst), dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    e0.record(pstream)
    for s in range(steps):
        for k, (h, Vi, Vpi) in enumerate(zip(hs, Vs, Vps)):
            h.count_async(pc.data_ptr() + 4 * (s * n_inst + k))
            h.step(Vi, Vpi, use_graph=False)
    e1.record(pstream)
    torch.cuda.synchronize()
    p_ms = e0.elapsed_time(e1)
    p_ps = int(pc.sum().item())
    for h in hs:
        h.close()
    b_rate = b_ps / (b_ms * 1e-3)
    p_rate = p_ps / (p_ms * 1e-3)
    return {"instances": n_inst, "particles_at_start": n0, "steps": steps,
            "batched_particle_steps_per_s": b_rate, "batched_ms_per_step": b_ms / steps,
            "plain_particle_steps_per_s": p_rate, "plain_ms_per_step": p_ms / steps,
            "speedup": b_rate / p_rate, "tuning": tuning}


def adjoint_cost(reps=3, scale=1.0, workload="adjoint2d"):
    """Cost of the adjoint against its forward. adjoint2d: the adjoint channel
    (256x128 Q1 on [0,2]x[0,1], 16 seeded uniform particles per cell,
    Poiseuille amplitude 1, 20 steps of 5 substeps, target from amplitude
    1.05); adjoint3d: the unit cube (16^3 Q1 times `scale` per direction, 8
    seeded uniform particles per cell, ABC flow with A = 1, 10 steps, target
    from A = 1.05). Final projection, checkpoints every 5 steps. Forward: the
    moves and the projection. Adjoint: one
    pfem2gpu_adjoint call (its own checkpointed forward, the loss, the reverse
    sweep with segment recompute). Wall clock with the stream drained on both
    sides; each repetition re-inserts the same initial particles into the
    same two handles (so the second and later repetitions reuse the device
    buffers of the first); median of `reps` after one warm-up repetition."""
    import torch
    from pfem2dealii_amd import ParticleHandler, workloads as W
    cfg = W.config(workload, scale=scale)
    steps, K = cfg["steps"], 5
    field = (lambda a: W.poiseuille(cfg["mesh"], a)) if workload == "adjoint2d" else \
        (lambda a: W.abc_flow(cfg["mesh"], A=a))

    def fresh(amp):
        h = ParticleHandler(cfg["mesh"], cfg["params"])
        h.initialize_maps()
        h.insert_particles(cfg["x"], cfg["cells"])
        Vd = torch.tensor(field(amp), dtype=torch.float64, device="cuda")
        h.set_velocity_from_field(Vd)
        h.synchronize()
        return h, Vd

    ht, Vt = fresh(1.05)
    target = torch.empty_like(Vt)
    for _ in range(steps):
        ht.move_particles(Vt)
    ht.project_particle_fields(target)
    ht.synchronize()
    ht.close()
    hf, V = fresh(1.0)
    ha, _ = fresh(1.0)
    for h in (hf, ha):
        h.reserve_particles(3 * len(cfg["cells"]))

    def reset(h):
        h.clear_particles()
        h.insert_particles(cfg["x"], cfg["cells"])
        h.set_velocity_from_field(V)
        h.synchronize()

    fwd, adj = [], []
    n0 = 0
    loss = None
    Vp = torch.empty_like(V)
    for r in range(reps + 1):  # the first repetition warms the buffers up
        reset(hf)
        n0 = hf.count()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            hf.move_particles(V)
        hf.project_particle_fields(Vp)
        hf.synchronize()
        t1 = time.perf_counter()
        reset(ha)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        loss, _, _ = ha.adjoint(steps, K, V, target)
        ha.synchronize()
        t3 = time.perf_counter()
        if r > 0:
            fwd.append(t1 - t0)
            adj.append(t3 - t2)
    hf.close()
    ha.close()
    f, a = float(np.median(fwd)), float(np.median(adj))
    return {"workload": workload, "mesh": "x".join(str(a_) for a_ in cfg["mesh"]["n"]), "particles_at_start": n0,
            "steps": steps, "checkpoint_every": K, "forward_ms": round(f * 1e3, 3), "adjoint_ms": round(a * 1e3, 3),
            "ratio": round(a / f, 3), "loss": loss}


def calibrate_copy(n=1 << 27, reps=5):
    """Known-size references for the FETCH_SIZE/WRITE_SIZE counters: a 1 GiB
    device copy of n doubles (8n bytes read, 8n written) run `reps` times,
    first as the runtime's blit copy (copy_), then as an ordinary elementwise
    kernel with plain loads and stores (mul by 1.0 into another tensor)."""
    import torch
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    for _ in range(reps):
        torch.mul(a, 1.0, out=b)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": 8 * n, "copy_bytes_written": 8 * n, "reps": reps,
                      "kernels": ["blit copy", "elementwise mul"]}), flush=True)


def wire_cost(workload="taylor_green2d", scale=1.0, reps=3):
    """Particle records of the reference's MPI exchange (88 bytes in 2D, 136
    in 3D) on the full workload: after one move, the particles of the upper
    half of the cells (one subdomain's worth) are packed into a device
    buffer, then unpacked into an empty handle on the same mesh. Kernel
    times from HIP events around the launches (median over `reps` calls after
    one warm-up call); algorithmic bytes per record: pack reads x and v and
    writes the record (16*dim + record bytes), unpack reads the record and
    writes x, v and cell to the staging arrays (record + 16*dim + 4 bytes).
    The unpack call also re-bins the received particles (the insert path),
    which its wall time includes."""
    import torch
    from pfem2dealii_amd import ParticleHandler, particle_record_size, workloads as W
    cfg = W.config(workload, scale=scale)
    mesh = cfg["mesh"]
    dim = mesh["dim"]
    rec = particle_record_size(dim)
    V = torch.tensor(cfg["V"], dtype=torch.float64, device="cuda")
    h = ParticleHandler(mesh, cfg["params"])
    h.initialize_maps()
    h.insert_particles(cfg["x"], cfg["cells"])
    h.set_velocity_from_field(V)
    h.move_particles(V)
    nc = h.n_cells
    lists = [np.arange(nc // 2, nc, dtype=np.int32)]
    n_rec = int(h.cell_counts()[nc // 2:].astype(np.int64).sum())
    buf = torch.empty(n_rec * rec, dtype=torch.uint8, device="cuda")
    h.pack_particles(lists, out=buf)
    walls, kms = [], []
    for _ in range(reps):
        h.set_timing(True)
        t0 = time.perf_counter()
        h.pack_particles(lists, out=buf)
        walls.append(time.perf_counter() - t0)
        kms.append(h.timings()["k_wire_pack"][1])
        h.set_timing(False)
    h.close()
    pack_ms, pack_wall = float(np.median(kms)), float(np.median(walls)) * 1e3
    h2 = ParticleHandler(mesh, cfg["params"])
    h2.initialize_maps()
    h2.reserve_particles(n_rec)
    h2.unpack_particles(buf)
    uwalls, ukms = [], []
    for _ in range(reps):
        h2.clear_particles()
        h2.synchronize()
        h2.set_timing(True)
        t0 = time.perf_counter()
        rej = h2.unpack_particles(buf)
        uwalls.append(time.perf_counter() - t0)
        ukms.append(h2.timings()["k_wire_unpack"][1])
        h2.set_timing(False)
        assert rej == 0 and h2.count() == n_rec
    h2.close()
    unpack_ms, unpack_wall = float(np.median(ukms)), float(np.median(uwalls)) * 1e3
    pack_b = n_rec * (16 * dim + rec)
    unpack_b = n_rec * (rec + 16 * dim + 4)
    return {"workload": workload, "records": n_rec, "record_bytes": rec,
            "pack_kernel_ms": round(pack_ms, 4), "pack_call_ms": round(pack_wall, 3),
            "pack_GBps": round(pack_b / (pack_ms * 1e-3) / 1e9, 1),
            "pack_frac": round(pack_b / (pack_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "unpack_kernel_ms": round(unpack_ms, 4), "unpack_call_ms": round(unpack_wall, 3),
            "unpack_GBps": round(unpack_b / (unpack_ms * 1e-3) / 1e9, 1),
            "unpack_frac": round(unpack_b / (unpack_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}


FP64_PEAK_TFLOPS = 78.6


def fem_work_per_cell(dim):
    """Algorithmic HBM bytes and FP64 operations per cell of the cell
    kernels (k_fem_pressure, k_fem_vcorr) as written: reads of the cell's
    vertex and DoF ids (2 * 4 * 2^d) and of the places of its matrix and
    right-hand side entries (4 * 4^d + 4 * 2^d), one vertex's coordinates and
    one DoF's field values per cell (8 d each; + 8 for the pressure of
    k_fem_vcorr); writes of the cell matrix (8 * 4^d) and right-hand side
    (8 * 2^d, times d for k_fem_vcorr). Operations: per quadrature point the Jacobian
    (2^(d+1) d^2), its inverse (7 in 2D, 41 in 3D), JxW and the 2^d real
    gradients (2^d d (2d - 1)); per (row, point, column) 4d + 5 for the
    pressure (gradient product, matrix and right-hand side updates) and
    3 + 4d for the velocity correction."""
    nv = 1 << dim
    geo = nv * (2 * nv * dim * dim + (7 if dim == 2 else 41) + 1 + nv * dim * (2 * dim - 1))
    places = 4 * nv * nv + 4 * nv
    p_bytes = 8 * nv + places + 8 * dim + 8 * dim + 8 * nv * nv + 8 * nv
    v_bytes = 8 * nv + places + 8 * dim + 8 * dim + 8 + 8 * nv * nv + 8 * nv * dim
    # velocity prediction: two vector fields in, dim matrices and right-hand
    # sides out; per (row, point, column) the mass terms (3 + 3 dim), the
    # viscous stencils (dim (2 dim + 2)) and the cross terms (2D 2 x 7, 3D
    # 3 x 14)
    pr_bytes = 8 * nv + places * dim + 16 * dim + 8 * nv * nv * dim + 8 * nv * dim
    pr_ops = 3 + 3 * dim + dim * (2 * dim + 2) + (14 if dim == 2 else 42)
    return {"k_fem_pressure": (p_bytes, geo + nv * nv * nv * (4 * dim + 5)),
            "k_fem_vcorr": (v_bytes, geo + nv * nv * nv * (3 + 4 * dim)),
            "k_fem_predict": (pr_bytes, geo + nv * nv * nv * pr_ops)}


def fem_cost(reps=5, cpu=True):
    """FEM assembly (pressure equation with every boundary face flagged, and
    velocity correction, matrices and right-hand sides) on the meshes of the
    2D and 3D workloads at full size, device arrays: per-kernel HIP-event
    durations (median over `reps` calls after one warm-up call), algorithmic
    bytes and FP64 operations of the cell kernels (fem_work_per_cell) and of
    the segmented sums (matrix: 8 bytes read per contribution, 8 bytes of
    offset read and 8 written per entry; right-hand side: 8 bytes read per
    (cell, vertex) and component, 4 bytes of offset read per DoF and 8 written
    per DoF and component). CPU baseline: the
    oracle's assembly (one thread) of the 2D mesh and of the 3D mesh at
    half the cells per direction."""
    import torch
    from pfem2dealii_amd import ParticleHandler, workloads as W
    out = {}
    for name, n, lo, hi in (("taylor_green2d", [2048, 1024], [0.0, 0.0], [2.0 * np.pi, np.pi]),
                            ("abc3d", [128, 128, 128], [0.0] * 3, [1.0] * 3)):
        mesh = W.structured_mesh(n, lo, hi)
        dim, nd = mesh["dim"], mesh["n_dofs"]
        nc = mesh["cell_vertices"].shape[0]
        nv = 1 << dim
        prm = dict(n_sub=1, dt=0.01, max_per_part=4, particles_per_cell=[2] * dim)
        h = ParticleHandler(mesh, prm)
        rp, cols = h.fem_sparsity()
        nnz = cols.shape[0]
        rng = np.random.default_rng(1)
        V = torch.tensor(rng.standard_normal((dim, nd)), device="cuda")
        P = torch.tensor(rng.standard_normal(nd), device="cuda")
        faces = torch.tensor(box_faces(n), device="cuda")
        Vo = torch.tensor(rng.standard_normal((dim, nd)), device="cuda")
        h.assemble_pressure_equation(V, 1.0, 0.01, boundary_faces=faces)
        h.assemble_velocity_correction(V, P, 1.0, 0.01)
        h.assemble_velocity_prediction(Vo, V, 1.0, 0.01, 0.01, boundary_faces=faces)
        per = {}
        walls = {"pressure": [], "velocity_correction": [], "velocity_prediction": []}
        for _ in range(reps):
            h.set_timing(True)
            t0 = time.perf_counter()
            h.assemble_pressure_equation(V, 1.0, 0.01, boundary_faces=faces)
            walls["pressure"].append(time.perf_counter() - t0)
            t0 = time.perf_counter()
            h.assemble_velocity_correction(V, P, 1.0, 0.01)
            walls["velocity_correction"].append(time.perf_counter() - t0)
            t0 = time.perf_counter()
            h.assemble_velocity_prediction(Vo, V, 1.0, 0.01, 0.01, boundary_faces=faces)
            walls["velocity_prediction"].append(time.perf_counter() - t0)
            for k, (cnt, ms) in h.timings().items():
                if k.startswith("k_fem"):
                    per.setdefault(k, []).append(ms / cnt)
            h.set_timing(False)
        h.close()
        work = fem_work_per_cell(dim)
        kern = {}
        for k, v in sorted(per.items()):
            ms = float(np.median(v))
            e = {"ms": round(ms, 4)}
            if k in work:
                b, f = work[k][0] * nc, work[k][1] * nc
            elif k in ("k_fem_gather_matrix", "k_fem_gather_mass", "k_fem_gather_pred"):
                nm = dim if k == "k_fem_gather_pred" else 1
                b, f = nm * (8.0 * nc * nv * nv + 8.0 * nnz) + 8.0 * (nnz + 1), float(nm * nc * nv * nv)
            elif k in ("k_fem_gather_rhs", "k_fem_gather_rhs_v", "k_fem_gather_rhs_p"):
                ncomp = 1 if k == "k_fem_gather_rhs" else dim
                b, f = 8.0 * ncomp * nc * nv + 8.0 * ncomp * nd + 4.0 * (nd + 1), float(ncomp * nc * nv)
            else:
                b, f = None, None
            if b is not None:
                gbs = b / (ms * 1e-3) / 1e9
                tf = f / (ms * 1e-3) / 1e12
                e.update({"alg_bytes": b, "GBps": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
                          "fp64_ops": f, "TFLOPs": round(tf, 2), "fp64_frac": round(tf / FP64_PEAK_TFLOPS, 4)})
            kern[k] = e
        res = {"mesh": "x".join(str(a) for a in n), "cells": nc, "dofs": nd, "nnz": nnz, "kernels": kern,
               "pressure_call_ms": round(float(np.median(walls["pressure"])) * 1e3, 3),
               "velocity_correction_call_ms": round(float(np.median(walls["velocity_correction"])) * 1e3, 3),
               "velocity_prediction_call_ms": round(float(np.median(walls["velocity_prediction"])) * 1e3, 3)}
        if cpu:
            from oracle import FemOracle
            cs = 1 if dim == 2 else 2
            cm = W.structured_mesh([a // cs for a in n], lo, hi)
            fo = FemOracle(cm)
            Vc = np.ones((dim, cm["n_dofs"]))
            t0 = time.perf_counter()
            fo.pressure(Vc, 1.0, 0.01, box_faces(cm["n"]))
            tp = time.perf_counter() - t0
            t0 = time.perf_counter()
            fo.velocity_correction(Vc, Vc[0], 1.0, 0.01)
            tv = time.perf_counter() - t0
            ccells = cm["cell_vertices"].shape[0]
            t0 = time.perf_counter()
            fo.prediction(Vc, Vc, 1.0, 0.01, 0.01, box_faces(cm["n"]))
            tq = time.perf_counter() - t0
            res["cpu_baseline"] = {"pressure_cells_per_s": ccells / tp, "velocity_correction_cells_per_s": ccells / tv,
                                   "velocity_prediction_cells_per_s": ccells / tq,
                                   "cores": 1, "kind": "port",
                                   "sample": "oracle assembly (matrix + right-hand side) of the %s mesh, one thread"
                                             % cm_name(cm)}
        gp = sum(kern[k]["ms"] for k in ("k_fem_pressure", "k_fem_faces", "k_fem_gather_matrix", "k_fem_gather_rhs")
                 if k in kern)
        gv = sum(kern[k]["ms"] for k in ("k_fem_vcorr", "k_fem_gather_mass", "k_fem_gather_rhs_v") if k in kern)
        gq = sum(kern[k]["ms"] for k in ("k_fem_predict", "k_fem_predict_faces", "k_fem_gather_pred",
                                         "k_fem_gather_rhs_p") if k in kern)
        res["velocity_prediction_gpu_cells_per_s"] = nc / (gq * 1e-3)
        res["pressure_gpu_cells_per_s"] = nc / (gp * 1e-3)
        res["velocity_correction_gpu_cells_per_s"] = nc / (gv * 1e-3)
        out[name] = res
    return out


def box_faces(n):
    """Boundary face flags of a box mesh (bit 2k + side on the box faces)."""
    nc = int(np.prod(n))
    fm = np.zeros(nc, dtype=np.uint8)
    rem = np.arange(nc)
    for k in range(len(n)):
        ck = rem % n[k]
        rem = rem // n[k]
        fm |= np.where(ck == 0, 1 << (2 * k), 0).astype(np.uint8)
        fm |= np.where(ck == n[k] - 1, 1 << (2 * k + 1), 0).astype(np.uint8)
    return fm


def cm_name(mesh):
    return "x".join(str(a) for a in mesh["n"])


def plain_run(workload, scale, steps, warmup, autotune=False):
    """Run A: `warmup` + `steps` steps of the plain handle from the seeded
    state, HIP events around the timed steps on the handle's stream
    (particles counted on the device at the start of each step). Run B: the
    same steps from the same state with HIP events around every launch, for
    the per-kernel durations and the roofline of the dominant kernel
    (algorithmic bytes per launch / average launch duration; `traffic` from
    the committed counter summary)."""
    import torch
    from pfem2dealii_amd import ParticleHandler, workloads as W

    t_setup = time.perf_counter()
    cfg = W.config(workload, scale=scale)
    mesh = cfg["mesh"]
    dim = mesh["dim"]
    V = torch.tensor(cfg["V"], dtype=torch.float64, device="cuda")
    Vp = torch.empty_like(V)

    def make():
        h = ParticleHandler(mesh, cfg["params"])
        h.initialize_maps()
        assert h.insert_particles(cfg["x"], cfg["cells"]) == 0
        h.set_velocity_from_field(V)
        # capacity for the particle growth of the measured steps (refills),
        # so that no step reallocates device arrays
        h.reserve_particles(3 * len(cfg["cells"]))
        return h

    def step(h):
        h.move_particles(V)
        h.project_particle_fields(Vp)
        h.correct_particle_velocity(V, Vp)

    from pfem2dealii_amd.autotune import shipped, tune
    sig = {"workload": workload, "scale": scale}
    tuning = None if autotune else shipped(sig)
    if tuning is None:  # on a scratch handle: tuning steps advance the simulation
        ht = make()
        tuning, _ = tune(ht, lambda: step(ht), signature=sig, force=True)
        ht.close()
    h = make()
    h.set_tuning(**tuning)
    n0 = h.count()
    stream = torch.cuda.ExternalStream(h.stream())
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t_setup

    # A. timed run: warmup steps, then `steps` steps between two HIP events
    for _ in range(warmup):
        step(h)
    h.synchronize()
    counts = torch.zeros(steps, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    h.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    for s in range(steps):
        h.count_async(counts.data_ptr() + 4 * s)
        step(h)
    ev1.record(stream)
    h.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = ev0.elapsed_time(ev1)
    psteps = int(counts.sum().item())
    ms_per_step = gpu_ms / steps
    value = psteps / (gpu_ms * 1e-3)
    h.close()

    # B. the same warmup + timed steps again from the same initial state with
    #    HIP events around every launch on the handle's stream: per-kernel
    #    durations over exactly the launches of run A (a kernel trace of this
    #    program sees runs A and B, which launch identical work)
    h = make()
    h.set_tuning(**tuning)
    nb_steps = warmup + steps
    pcounts = torch.zeros(nb_steps + 1, dtype=torch.int32, device="cuda")
    h.set_timing(True)
    for s in range(nb_steps):
        h.count_async(pcounts.data_ptr() + 4 * s)
        step(h)
    h.count_async(pcounts.data_ptr() + 4 * nb_steps)
    tim = h.timings()
    h.set_timing(False)
    h.close()
    pcl = pcounts.cpu().tolist()
    nodes = mesh["X"].shape[1]
    b_alg = alg_bytes_per_particle_step(dim, nodes, psteps / steps)
    achieved_step = b_alg * psteps / steps / (ms_per_step * 1e-3) / 1e9
    # kernels only: the reference's timer sections ("Particle ...") bracket whole calls
    kern = {k: {"launches": n, "avg_ms": ms / n, "ms_per_step": ms / nb_steps} for k, (n, ms) in tim.items()
            if not k.startswith("Particle ")}
    dom = max(kern.items(), key=lambda kv: kv[1]["ms_per_step"]) if kern else ("none", {"avg_ms": 0.0, "launches": 0})
    n_in = sum(pcl[:-1]) / float(nb_steps)
    n_out = sum(pcl[1:]) / float(nb_steps)
    alg = kernel_alg_bytes(dim, nodes, n_in, n_out)
    dname = dom[0]
    dbytes = alg.get(dname)
    per_launch = dbytes / (dom[1]["launches"] / float(nb_steps)) if dbytes is not None else None
    achieved = per_launch / (dom[1]["avg_ms"] * 1e-3) / 1e9 if per_launch else None
    roofline = {"bound": "hbm", "achieved": round(achieved, 1) if achieved else None, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
                "traffic": hbm_traffic(workload, scale, dname), "kernel": dname,
                "alg_bytes_per_launch": per_launch, "avg_launch_ms": round(dom[1]["avg_ms"], 4),
                "launches": dom[1]["launches"],
                "step": {"achieved": round(achieved_step, 1), "frac": round(achieved_step / HBM_PEAK_GBS, 4),
                         "alg_bytes_per_particle_step": round(b_alg, 2), "particles_per_step": psteps / steps}}
    return {"value": value, "unit": "particle-steps/s", "ms_per_step": ms_per_step,
            "config": {"workload": workload if scale == 1.0 else "%s@%g" % (workload, scale),
                       "mesh": "x".join(str(a) for a in mesh["n"]), "particles_at_start": n0,
                       "n_sub": cfg["params"]["n_sub"], "dt": cfg["params"]["dt"], "parallelism": "1 GPU",
                       "form": "plain", "tuning": tuning},
            "roofline": roofline,
            "kernels_ms_per_step": {k: round(v["ms_per_step"], 4) for k, v in sorted(kern.items())},
            "wall_s": round(wall, 3), "setup_s": round(setup_s, 1)}


def tune_all():
    """Runs the launch-shape search (autotune.tune, forced) for the plain
    2D and 3D workloads and the 256-instance batch, each on a scratch handle,
    and returns the table entries (signature -> shape and times)."""
    import torch
    from pfem2dealii_amd import BatchHandler, ParticleHandler, workloads as W
    from pfem2dealii_amd.autotune import tune
    out = {}
    for workload in ("taylor_green2d", "abc3d"):
        cfg = W.config(workload)
        V = torch.tensor(cfg["V"], dtype=torch.float64, device="cuda")
        Vp = torch.empty_like(V)
        h = ParticleHandler(cfg["mesh"], cfg["params"])
        h.initialize_maps()
        h.insert_particles(cfg["x"], cfg["cells"])
        h.set_velocity_from_field(V)
        h.reserve_particles(3 * len(cfg["cells"]))

        def step():
            h.move_particles(V)
            h.project_particle_fields(Vp)
            h.correct_particle_velocity(V, Vp)

        sig = {"workload": workload, "scale": 1.0}
        _, info = tune(h, step, signature=sig, force=True, log=progress)
        out[json.dumps(sig, sort_keys=True)] = info
        progress("tuned %s: %s" % (workload, json.dumps(info)))
        h.close()
        del V, Vp
    inst = W.batched_instances(256, seed=4)
    V = torch.tensor(np.concatenate([i["V"] for i in inst], axis=1), dtype=torch.float64, device="cuda")
    Vp = torch.empty_like(V)
    b = BatchHandler([i["mesh"] for i in inst], [i["params"] for i in inst])
    b.initialize_maps()
    co = b.cell_offsets
    x = np.concatenate([i["x"] for i in inst], axis=1)
    cells = np.concatenate([i["cells"] + co[k] for k, i in enumerate(inst)]).astype(np.int32)
    b.insert_particles(x, cells)
    b.set_velocity_from_field(V)
    b.reserve_particles(3 * len(cells))
    sig = {"workload": "batched", "n": 256, "scale": 1.0}
    _, info = tune(b, lambda: b.step(V, Vp, use_graph=True), signature=sig, force=True, log=progress)
    out[json.dumps(sig, sort_keys=True)] = info
    progress("tuned batched: %s" % json.dumps(info))
    b.close()
    return out


def progress(msg):
    """One line per finished leg on stderr (stdout carries only the JSON)."""
    print("[bench] " + msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="taylor_green2d")
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batched-steps", type=int, default=20, help="steps of the 256-instance comparison (0: skip)")
    ap.add_argument("--no-3d", action="store_true", help="skip the 128^3 ABC leg")
    ap.add_argument("--adjoint-reps", type=int, default=3, help="repetitions of the adjoint cost measurement (0: skip)")
    ap.add_argument("--autotune", action="store_true", help="tune launch shapes instead of the shipped table")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run 1 GiB device copies (counter calibration) and exit")
    ap.add_argument("--wire-reps", type=int, default=3,
                    help="repetitions of the particle-record pack/unpack measurement (0: skip)")
    ap.add_argument("--fem-reps", type=int, default=5,
                    help="repetitions of the FEM assembly measurement (0: skip)")
    ap.add_argument("--section", choices=("batched", "adjoint", "wire", "fem", "tune"), default=None,
                    help="only run the batched comparison or the adjoint cost measurement (for profiling) and "
                         "print its result")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (the multi-GPU design is documented, not built); "
              "use --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.section == "batched":
        print(json.dumps({"section": "batched",
                          "batched": batched_comparison(max(1, args.batched_steps), 1, force_tune=args.autotune)}),
              flush=True)
        return
    if args.section == "adjoint":
        print(json.dumps({"section": "adjoint", "adjoint": adjoint_cost(max(1, args.adjoint_reps)),
                          "adjoint3d": adjoint_cost(max(1, args.adjoint_reps), 4.0, "adjoint3d")}), flush=True)
        return
    if args.section == "tune":
        print(json.dumps({"section": "tune", "table": tune_all()}), flush=True)
        return
    if args.section == "fem":
        print(json.dumps({"section": "fem", "fem": fem_cost(max(1, args.fem_reps), cpu=not args.no_cpu)}),
              flush=True)
        return
    if args.section == "wire":
        print(json.dumps({"section": "wire", "wire": wire_cost(args.workload, args.scale, max(1, args.wire_reps))}),
              flush=True)
        return

    plain = plain_run(args.workload, args.scale, args.steps, args.warmup, args.autotune)
    progress("plain %s: %.4g particle-steps/s" % (args.workload, plain["value"]))
    out = {
        "metric": "particle-steps/s plain vs batched-256 on 1 GPU at 1e-11 rel err; %HBM peak",
        "value": plain["value"], "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": plain["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic", "config": plain["config"], "roofline": plain["roofline"],
        "kernels_ms_per_step": plain["kernels_ms_per_step"], "wall_s": plain["wall_s"], "setup_s": plain["setup_s"],
    }
    if args.workload == "taylor_green2d" and args.scale == 1.0 and not args.no_3d:
        d3 = plain_run("abc3d", 1.0, args.steps, args.warmup, args.autotune)
        progress("plain abc3d: %.4g particle-steps/s" % d3["value"])
        out["abc3d"] = {k: d3[k] for k in ("value", "unit", "ms_per_step", "config", "roofline",
                                            "kernels_ms_per_step")}
    if args.batched_steps > 0:
        out["batched"] = batched_comparison(args.batched_steps, 1, force_tune=args.autotune)
        progress("batched comparison done")
    if args.adjoint_reps > 0:
        out["adjoint"] = adjoint_cost(args.adjoint_reps)
        out["adjoint3d"] = adjoint_cost(args.adjoint_reps, 4.0, "adjoint3d")
        progress("adjoint cost done")
    if args.wire_reps > 0:
        out["wire"] = wire_cost(args.workload, args.scale, args.wire_reps)
        progress("record pack/unpack done")
    if args.fem_reps > 0 and args.scale == 1.0:
        out["fem"] = fem_cost(args.fem_reps, cpu=not args.no_cpu)
        progress("FEM assembly done")
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.workload)
        progress("CPU baseline %s done" % args.workload)
        if "abc3d" in out:
            out["abc3d"]["cpu_baseline"] = cpu_baseline("abc3d")
            progress("CPU baseline abc3d done")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
