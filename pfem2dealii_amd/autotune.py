"""Autotuner for the launch shape of the particle step.

Tunes the knobs of pfem2gpu_set_tuning by coordinate descent on the live
handle (the launch shape does not change results, only timing):
  move_block  threads per workgroup of the in-cell move pass (64..512)
  walk_block  threads per workgroup of a walk round (64..512)
  walk_grid   workgroups of a walk round of the move
  settle_grid workgroups of the redistribution/projection kernel
  move_grid   workgroups of the in-cell move pass (0: one particle per thread)
The measured steps advance the simulation, and the particle count grows
while refills fill empty sub-cell parts, so the handle is first stepped
until the count changes by less than 1 % per step, and candidates are
compared by time per particle-step: `steps` steps timed with HIP events on
the handle's stream after one untimed step, divided by the particles at the
start of those steps. Results are cached per workload signature;
`tuning_mi355x.json` beside this file holds the shapes found on MI355X for the
benchmark workloads (with the time per step of the shape found and of the
compiled-in default, measured in the same search) and is consulted first
unless `force` is set.
"""
import json
import os

TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning_mi355x.json")

DEFAULT = {"walk_grid": 4096, "settle_grid": 2048, "move_grid": 0, "move_block": 256, "walk_block": 256}
STAGES = (
    ("move_block", (64, 128, 256, 512)),
    ("walk_block", (64, 128, 256, 512)),
    ("walk_grid", (1024, 2048, 4096, 8192, 16384)),
    ("settle_grid", (512, 1024, 2048, 4096, 8192)),
    ("move_grid", (0, 4096, 8192, 16384, 32768)),
)

_cache = {}


def _time_steps(handle, step, steps):
    """(ms per step, ns per particle-step) of `steps` steps after one untimed step."""
    import torch
    stream = torch.cuda.ExternalStream(handle.stream())
    step()
    n0 = handle.count()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record(stream)
    for _ in range(steps):
        step()
    b.record(stream)
    handle.synchronize()
    n1 = handle.count()
    ms = a.elapsed_time(b) / steps
    return ms, ms * 1e6 / max(1.0, 0.5 * (n0 + n1))


def _settle_count(handle, step, max_steps=12):
    """Steps until the particle count changes by less than 1 % per step."""
    n = handle.count()
    for _ in range(max_steps):
        step()
        m = handle.count()
        if abs(m - n) <= 0.01 * max(n, 1):
            break
        n = m


def shipped(signature):
    """The launch shape recorded for `signature` in the MI355X table, or None."""
    try:
        with open(TABLE) as f:
            table = json.load(f)
    except (OSError, ValueError):
        return None
    e = table.get(json.dumps(signature, sort_keys=True))
    return dict(DEFAULT, **e["tuning"]) if e else None


def tune(handle, step, signature=None, steps=3, cache_path=None, log=None, force=False):
    """Chooses and applies the fastest launch shape; returns (tuning, info).
    info is None when the shape came from the shipped table, else a dict with
    the ms per step and ns per particle-step of the shape found and of the
    default shape."""
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
            return _cache[key]["tuning"], _cache[key]
    handle.set_tuning(**DEFAULT)
    _settle_count(handle, step)
    best = dict(DEFAULT)
    best_ms, best_ns = _time_steps(handle, step, steps)
    default_ms, default_ns = best_ms, best_ns
    for name, values in STAGES:
        for v in values:
            t = dict(best, **{name: v})
            if t == best:
                continue
            handle.set_tuning(**t)
            ms, ns = _time_steps(handle, step, steps)
            if log:
                log("tune %s: %.3f ms/step, %.4f ns/particle-step" % (t, ms, ns))
            if ns < best_ns:
                best, best_ms, best_ns = dict(t), ms, ns
    # the default and the winner once more, back to back
    handle.set_tuning(**DEFAULT)
    default_ms, default_ns = _time_steps(handle, step, steps)
    handle.set_tuning(**best)
    best_ms, best_ns = _time_steps(handle, step, steps)
    if best_ns >= default_ns:
        best, best_ms, best_ns = dict(DEFAULT), default_ms, default_ns
        handle.set_tuning(**best)
    info = {"tuning": best, "ms": round(best_ms, 4), "ns_per_particle_step": round(best_ns, 5),
            "default_ms": round(default_ms, 4), "default_ns_per_particle_step": round(default_ns, 5)}
    if key is not None:
        _cache[key] = info
        if cache_path:
            try:
                with open(cache_path, "w") as f:
                    json.dump(_cache, f)
            except OSError:
                pass
    return best, info
