"""Batched form and graph-captured step on the GPU. Marked gpu.

A batch must reproduce, instance by instance and bit for bit, what a plain
handle over that instance alone computes (which the parity tests tie to the
oracle); a replayed hipGraph step must reproduce the eager step."""
import sys

import numpy as np
import pytest

from pfem2dealii_amd import workloads as W

pytestmark = pytest.mark.gpu


def _say(msg):
    print("[case] " + msg, flush=True)
    sys.stdout.flush()


def _concat(instances, offsets_c, offsets_d):
    x = np.concatenate([i["x"] for i in instances], axis=1)
    cells = np.concatenate([i["cells"] + offsets_c[k] for k, i in enumerate(instances)]).astype(np.int32)
    V = np.concatenate([i["V"] for i in instances], axis=1)
    return x, cells, V


def _same(a, b):
    return np.array_equal(np.nan_to_num(a, nan=7.0), np.nan_to_num(b, nan=7.0)) and \
        np.array_equal(np.isnan(a), np.isnan(b))


def test_batch_matches_plain_instances(gpu_available):
    _say("batch of 6 channel instances (mesh scale 1/4) vs 6 plain handles, 3 steps")
    import torch
    from pfem2dealii_amd import BatchHandler, ParticleHandler
    inst = W.batched_instances(6, seed=4, scale=0.25)
    b = BatchHandler([i["mesh"] for i in inst], [i["params"] for i in inst])
    b.initialize_maps()
    co, do = b.cell_offsets, b.dof_offsets
    x, cells, V = _concat(inst, co, do)
    assert b.insert_particles(x, cells) == 0
    b.set_velocity_from_field(V)
    plains = []
    for i in inst:
        h = ParticleHandler(i["mesh"], i["params"])
        h.initialize_maps()
        h.insert_particles(i["x"], i["cells"])
        h.set_velocity_from_field(i["V"])
        plains.append(h)
    for s in range(3):
        b.move_particles(V)
        Vb = b.project_particle_fields()
        b.correct_particle_velocity(V, Vb)
        gb = b.download()
        for k, (i, h) in enumerate(zip(inst, plains)):
            h.move_particles(i["V"])
            Vh = h.project_particle_fields()
            h.correct_particle_velocity(i["V"], Vh)
            assert _same(Vb[:, do[k]:do[k + 1]], Vh), "instance %d step %d projection" % (k, s)
            gh = h.download()
            sel = (gb["cell"] >= co[k]) & (gb["cell"] < co[k + 1])
            assert np.array_equal(gb["cell"][sel] - co[k], gh["cell"]), "instance %d cells" % k
            assert np.array_equal(gb["x"][:, sel], gh["x"]), "instance %d positions" % k
            assert np.array_equal(gb["v"][:, sel], gh["v"]), "instance %d velocities" % k
    torch.cuda.synchronize()


def test_graph_step_matches_eager(gpu_available):
    _say("hipGraph step replay vs eager step, channel 64x32, 5 steps")
    import torch
    from pfem2dealii_amd import ParticleHandler
    cfg = W.config("channel2d")
    hs = []
    for _ in range(2):
        h = ParticleHandler(cfg["mesh"], cfg["params"])
        h.initialize_maps()
        h.insert_particles(cfg["x"], cfg["cells"])
        h.set_velocity_from_field(cfg["V"])
        hs.append(h)
    V = torch.tensor(cfg["V"], dtype=torch.float64, device="cuda")
    outs = [torch.empty_like(V), torch.empty_like(V)]
    torch.cuda.synchronize()
    for s in range(5):
        hs[0].step(V, outs[0], use_graph=False)
        hs[1].step(V, outs[1], use_graph=True)
        hs[0].synchronize()
        hs[1].synchronize()
        a, b = outs[0].cpu().numpy(), outs[1].cpu().numpy()
        assert _same(a, b), "step %d projection" % s
        da, db = hs[0].download(), hs[1].download()
        assert np.array_equal(da["cell"], db["cell"]) and np.array_equal(da["x"], db["x"])
        assert np.array_equal(da["v"], db["v"]), "step %d velocities" % s
        # eager work between replays
        torch.cuda.synchronize()
        assert hs[1].count() == hs[0].count()


def test_graph_replays_while_the_count_grows(gpu_available):
    """Particles start in the first column of cells of a 64x32 channel and
    flow right at up to one cell per step (CFL 1), so every step fills a new
    column (the count grows by ~500 per step) and most particles cross a face
    in some substep. The graph is captured right after an exact count was
    read (~1000 particles) and then replayed seven times with no host reads
    between the replays; the capacity is reserved so that no replay moves a
    buffer. The captured launch shape and walk-queue stripes must hold every
    count the capacity allows: the replays must equal eager steps."""
    _say("hipGraph replays with a growing count, channel 64x32 CFL 1, 8 steps")
    import torch
    from pfem2dealii_amd import ParticleHandler
    mesh = W.structured_mesh([64, 32], [0.0, 0.0], [2.0, 1.0])
    V = W.poiseuille(mesh)
    params = dict(n_sub=5, dt=W.cfl_dt(mesh, V, cfl=1.0), max_per_part=5, particles_per_cell=[4, 4])
    x, cells = W.seed_jittered_lattice(mesh, [4, 4], 0.1, seed=41)
    sel = cells % 64 == 0
    x, cells = x[:, sel], cells[sel]
    hs = []
    for _ in range(2):
        h = ParticleHandler(mesh, params)
        h.initialize_maps()
        assert h.insert_particles(x, cells) == 0
        h.set_velocity_from_field(V)
        h.reserve_particles(6000)
        hs.append(h)
    Vd = torch.tensor(V, dtype=torch.float64, device="cuda")
    outs = [torch.empty_like(Vd), torch.empty_like(Vd)]
    torch.cuda.synchronize()
    counts = []
    for s in range(8):
        hs[0].step(Vd, outs[0], use_graph=False)
        hs[1].step(Vd, outs[1], use_graph=True)
        if s == 1:  # what the capture saw
            counts.append(hs[0].count())
    hs[0].synchronize()
    hs[1].synchronize()
    counts.append(hs[0].count())
    assert counts[-1] > 3 * counts[0], "the count did not grow: %s" % counts
    assert _same(outs[0].cpu().numpy(), outs[1].cpu().numpy())
    da, db = hs[0].download(), hs[1].download()
    assert np.array_equal(da["cell"], db["cell"]) and np.array_equal(da["x"], db["x"])
    assert np.array_equal(da["v"], db["v"])


def test_tuning_does_not_change_results(gpu_available):
    _say("launch tunings give bitwise identical steps")
    from pfem2dealii_amd import ParticleHandler
    cfg = W.config("taylor_green2d", scale=1.0 / 16)
    res = []
    for tune in ({"walk_grid": 4096, "settle_grid": 2048, "move_grid": 0, "move_block": 256, "walk_block": 256},
                 {"walk_grid": 1024, "settle_grid": 7, "move_grid": 256, "move_block": 64, "walk_block": 512},
                 {"walk_grid": 65536, "settle_grid": 300, "move_grid": 1000, "move_block": 512, "walk_block": 64},
                 {"walk_grid": 2048, "settle_grid": 4096, "move_grid": 0, "move_block": 128, "walk_block": 128}):
        h = ParticleHandler(cfg["mesh"], cfg["params"])
        h.set_tuning(**tune)
        assert h.get_tuning() == tune
        h.initialize_maps()
        h.insert_particles(cfg["x"], cfg["cells"])
        h.set_velocity_from_field(cfg["V"])
        for _ in range(2):
            h.move_particles(cfg["V"])
            Vp = h.project_particle_fields()
            h.correct_particle_velocity(cfg["V"], Vp)
        res.append((h.download(), Vp))
    for d, Vp in res[1:]:
        assert np.array_equal(d["cell"], res[0][0]["cell"]) and np.array_equal(d["x"], res[0][0]["x"])
        assert np.array_equal(d["v"], res[0][0]["v"]) and _same(Vp, res[0][1])
