"""Tile-size autotuner for the wake-induction launches (this project's own; the reference has no
tunables here).

Two knobs set how the wake induction is tiled into work-groups:
- the lattice work-group target of a fast-mode UVLM plan (tvlm_uvlm_set_lattice_target; library
  default 1024): above it a launch takes 2 or 4 wake rows per tile, below it the spans are split;
- the ring-by-ring form's target (tvlm_set_tile_targets, default 2048) for
  Backend::rhs_assemble_wake_influence and the plain UVLM loop.

The tuner times each candidate on the caller's own workload (whole simulations replayed from the
plan's graph, or the wake-influence call between HIP events) and keeps the fastest; results are
cached in a JSON file keyed by the workload's shape and the device, so a served workload is tuned
once per device.  Every candidate computes the same sums in a different tile order: results differ
only by rounding.
"""
from __future__ import annotations

import ctypes
import json
import os
import tempfile
import time

from . import capi as H

LATTICE_CANDIDATES = (512, 1024, 2048)
WAKE_CANDIDATES = (1024, 2048, 4096)


def default_cache_path() -> str:
    return os.environ.get("TVLM_AUTOTUNE_CACHE") or os.path.join(os.path.expanduser("~"), ".cache", "tinyvlm_amd",
                                                                 "autotune.json")


class Cache:
    """JSON file of tuned choices: {key: entry}.  path None keeps the choices in memory only.  A
    missing or unreadable file starts empty; writes replace the file atomically."""

    def __init__(self, path: str | None = None):
        self.path = path
        self.data: dict = {}
        if path and os.path.exists(path):
            try:
                with open(path) as f:
                    d = json.load(f)
                self.data = d if isinstance(d, dict) else {}
            except (OSError, ValueError):
                self.data = {}

    def get(self, key: str):
        return self.data.get(key)

    def put(self, key: str, entry: dict):
        self.data[key] = entry
        if not self.path:
            return
        os.makedirs(os.path.dirname(os.path.abspath(self.path)), exist_ok=True)
        fd, tmp = tempfile.mkstemp(dir=os.path.dirname(os.path.abspath(self.path)), suffix=".tmp")
        with os.fdopen(fd, "w") as f:
            json.dump(self.data, f, indent=1, sort_keys=True)
        os.replace(tmp, self.path)


def device_tag() -> str:
    import torch

    p = torch.cuda.get_device_properties(torch.cuda.current_device())
    return f"{p.name}|{getattr(p, 'gcnArchName', '')}|cu{p.multi_processor_count}"


def uvlm_key(plan, device: str) -> str:
    shapes = ",".join(f"{v.shape[0] - 1}x{v.shape[1] - 1}" for v in plan._verts)
    far = getattr(plan, "far_wake", None) or []
    sims = getattr(plan, "sims", 1)
    return f"uvlm-lattice|{shapes}|steps={plan.t_steps}|far={far}|sims={sims}|{device}"


def tune_uvlm(plan, candidates=LATTICE_CANDIDATES, cache: Cache | None = None, reps: int = 2,
              force: bool = False, apply_globally: bool = False) -> dict:
    """Picks the lattice work-group target of a fast-mode UVLMPlan (whose motion is set) by timing
    `reps` whole simulations per candidate (graph replay, after one capture run), sets it on the
    plan and records it in `cache`.  A cached choice for the same workload and device is applied
    without timing unless force.  apply_globally also makes it the library-wide lattice target, which
    UVLM plans created afterwards (a Server's, for instance) start with.  Returns the entry
    {"best", "sim_ms", "cached"}."""
    cache = cache if cache is not None else Cache(default_cache_path())
    key = uvlm_key(plan, device_tag())
    hit = cache.get(key)
    if hit is not None and not force:
        plan.set_lattice_target(hit["best"])
        if apply_globally:
            set_targets(get_targets()[0], hit["best"])
        return dict(hit, cached=True)
    sim_ms = {}
    for c in candidates:
        plan.set_lattice_target(c)
        plan.run(use_graph=True)
        plan.results()  # synchronizes
        t0 = time.perf_counter()
        for _ in range(reps):
            plan.run(use_graph=True)
        plan.results()
        sim_ms[str(c)] = round((time.perf_counter() - t0) / reps * 1e3, 4)
    best = int(min(sim_ms, key=sim_ms.get))
    plan.set_lattice_target(best)
    if apply_globally:
        set_targets(get_targets()[0], best)
    entry = {"best": best, "sim_ms": sim_ms, "candidates": list(candidates)}
    cache.put(key, entry)
    return dict(entry, cached=False)


def get_targets() -> tuple[int, int]:
    w, l = ctypes.c_int32(), ctypes.c_int32()
    H.check(H.lib().tvlm_get_tile_targets(ctypes.byref(w), ctypes.byref(l)), "tvlm_get_tile_targets")
    return w.value, l.value


def set_targets(wake_groups: int = 0, lattice_groups: int = 0):
    """Library-wide targets (0: the default)."""
    H.check(H.lib().tvlm_set_tile_targets(ctypes.c_int32(wake_groups), ctypes.c_int32(lattice_groups)),
            "tvlm_set_tile_targets")


def tune_wake_influence(rhs, gamma_wake, colloc, normals, verts_wake, lifting, iteration,
                        candidates=WAKE_CANDIDATES, cache: Cache | None = None, reps: int = 5,
                        force: bool = False) -> dict:
    """Picks the library-wide target of the ring-by-ring wake induction for this call's shape (the
    receivers, wake surfaces and rows of `iteration`) by timing Backend::rhs_assemble_wake_influence
    between HIP events; rhs is restored after timing.  Sets the target (keeping the lattice one)."""
    import torch

    cache = cache if cache is not None else Cache(default_cache_path())
    shapes = ",".join(f"{c.shape[0]}x{c.shape[1]}" for c in colloc)
    wakes = ",".join(f"{w.shape[0] - 1}x{w.shape[1]}" for w in verts_wake)
    key = f"wake-ring|{shapes}|{wakes}|it={iteration}|{str(rhs.dtype)}|{device_tag()}"
    _, lat = get_targets()
    hit = cache.get(key)
    if hit is not None and not force:
        set_targets(hit["best"], lat)
        return dict(hit, cached=True)
    saved = rhs.clone()
    ms = {}
    for c in candidates:
        set_targets(c, lat)
        H.rhs_assemble_wake_influence(rhs, gamma_wake, colloc, normals, verts_wake, lifting, iteration)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            H.rhs_assemble_wake_influence(rhs, gamma_wake, colloc, normals, verts_wake, lifting, iteration)
        e1.record()
        torch.cuda.synchronize()
        ms[str(c)] = round(e0.elapsed_time(e1) / reps, 5)
    rhs.copy_(saved)
    best = int(min(ms, key=ms.get))
    set_targets(best, lat)
    entry = {"best": best, "call_ms": ms, "candidates": list(candidates)}
    cache.put(key, entry)
    return dict(entry, cached=False)


__all__ = ["Cache", "tune_uvlm", "tune_wake_influence", "set_targets", "get_targets", "default_cache_path",
           "LATTICE_CANDIDATES", "WAKE_CANDIDATES"]
