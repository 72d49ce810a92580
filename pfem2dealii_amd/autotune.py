"""Autotuner for the launch shape of the particle step.

Tunes the knobs of pfem2gpu_set_tuning by coordinate descent on the live
handle (the measured steps advance the simulation; the launch shape does not
change results, only timing):
  walk_grid   workgroups of a walk round of the move
  settle_grid workgroups of the redistribution/projection kernel
  move_grid   workgroups of the in-cell move pass (0: one particle per thread)
Each candidate is timed over `steps` steps with HIP events on the handle's
stream after one untimed step. Results are cached per workload signature;
`tuning_mi355x.json` beside this file holds the shapes found on MI355X for the
benchmark workloads and is consulted first unless `force` is set.
"""
import json
import os

TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning_mi355x.json")

WALK_GRIDS = (1024, 2048, 4096, 8192, 16384)
SETTLE_GRIDS = (512, 1024, 2048, 4096, 8192)
MOVE_GRIDS = (0, 4096, 8192, 16384, 32768)

_cache = {}


def _time_steps(handle, step, steps):
    import torch
    stream = torch.cuda.ExternalStream(handle.stream())
    step()
    handle.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record(stream)
    for _ in range(steps):
        step()
    b.record(stream)
    handle.synchronize()
    return a.elapsed_time(b) / steps


def shipped(signature):
    """The launch shape recorded for `signature` in the MI355X table, or None."""
    try:
        with open(TABLE) as f:
            table = json.load(f)
    except (OSError, ValueError):
        return None
    e = table.get(json.dumps(signature, sort_keys=True))
    return dict(e["tuning"]) if e else None


def tune(handle, step, signature=None, steps=2, cache_path=None, log=None, force=False):
    """Chooses and applies the fastest launch shape; returns (tuning, ms/step).
    ms/step is None when the shape came from the shipped table."""
    key = json.dumps(signature, sort_keys=True) if signature is not None else None
    if key is not None and not force:
        t = shipped(signature)
        if t is not None:
            handle.set_tuning(**t)
            return t, None
    if key is not None and not force:
        if cache_path and not _cache and os.path.exists(cache_path):
            try:
                _cache.update(json.load(open(cache_path)))
            except (OSError, ValueError):
                pass
        if key in _cache:
            handle.set_tuning(**_cache[key]["tuning"])
            return _cache[key]["tuning"], _cache[key]["ms"]
    best = dict(handle.get_tuning())
    handle.set_tuning(**best)
    best_ms = _time_steps(handle, step, steps)
    trials = [dict(best, walk_grid=g) for g in WALK_GRIDS]
    for stage in ("walk", "settle", "move"):
        if stage == "settle":
            trials = [dict(best, settle_grid=g) for g in SETTLE_GRIDS]
        elif stage == "move":
            trials = [dict(best, move_grid=g) for g in MOVE_GRIDS]
        for t in trials:
            if t == best:
                continue
            handle.set_tuning(**t)
            ms = _time_steps(handle, step, steps)
            if log:
                log("tune %s: %.3f ms/step" % (t, ms))
            if ms < best_ms:
                best, best_ms = dict(t), ms
    handle.set_tuning(**best)
    if key is not None:
        _cache[key] = {"tuning": best, "ms": best_ms}
        if cache_path:
            try:
                with open(cache_path, "w") as f:
                    json.dump(_cache, f)
            except OSError:
                pass
    return best, best_ms
