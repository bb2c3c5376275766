"""Launch-shape autotuner for the time-evolution kernels and the batch scans.

Chooses between the phased default launch and a number of lanes sharing one
coarse step (threads per simulation step), for batches how many rows a wave
takes per grab of the work counter (the tile of work), and the rows per
workgroup chunk of the per-simulation state scans, by timing the candidates
on the device with HIP events.  Results are cached per (kernel, shape,
device) in memory and in a JSON file next to the code-object cache.
"""
from __future__ import annotations

import hashlib
import json
import os
import statistics
import threading

# Phased launches timed by tune_lanes: the resident waves, and 1-4 waves per
# SIMD (fewer, longer items against more waves to hide latency).
PHASED_CANDIDATES = [0, -1, -2, -3, -4]

_lock = threading.Lock()
_memo: dict = {}


def lane_candidates(n_fine: int, max_lanes: int = 64):
    """Divisors of n_fine (balanced chunks) and powers of two, up to 64."""
    top = max(1, min(int(n_fine), max_lanes))
    c = {g for g in range(1, top + 1) if n_fine % g == 0}
    g = 1
    while g <= top:
        c.add(g)
        g *= 2
    return sorted(c)


def _cache_path():
    from .simulator import _cache_dir
    return os.path.join(_cache_dir(), "autotune.json")


def _load_disk():
    try:
        with open(_cache_path()) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def _store_disk(key, value):
    try:
        data = _load_disk()
        data[key] = value
        os.makedirs(os.path.dirname(_cache_path()), exist_ok=True)
        tmp = _cache_path() + f".tmp{os.getpid()}"
        with open(tmp, "w") as fh:
            json.dump(data, fh)
        os.replace(tmp, _cache_path())
    except OSError:
        pass


def _key(sim, *shape):
    import torch

    h = hashlib.sha256(sim._code).hexdigest()[:16]
    name = torch.cuda.get_device_name(torch.cuda.current_device())
    return "|".join([h, name, *map(str, shape)])


def time_interleaved(fns: dict, rounds: int = 5, warmup: int = 1):
    """Median milliseconds of each fns[key]() on the current stream, the
    candidates timed in turn within every round so that slow drifts of the
    device clock affect them alike."""
    import torch

    for _ in range(warmup):
        for fn in fns.values():
            fn()
    times = {k: [] for k in fns}
    for _ in range(rounds):
        for k, fn in fns.items():
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record()
            fn()
            b.record()
            b.synchronize()
            times[k].append(a.elapsed_time(b))
    return {k: statistics.median(v) for k, v in times.items()}


def tune_lanes(sim, n_coarse: int, n_fine: int, dt: float, dt_out: float, sweep, t0: float = 0.0,
               reps: int = 7, use_cache: bool = True):
    """Best lanes_per_step for one simulation of this shape (0: phased launch
    on the resident waves, -k: phased on k waves per SIMD for k = 1..4, else
    lanes per coarse step); returns (lanes, table)."""
    import torch

    candidates = PHASED_CANDIDATES + lane_candidates(n_fine)
    key = _key(sim, "single", n_coarse, n_fine, tuple(candidates))
    if use_cache:
        with _lock:
            if key in _memo:
                return _memo[key], None
        disk = _load_disk()
        if key in disk:
            with _lock:
                _memo[key] = int(disk[key])
            return int(disk[key]), None
    dev = torch.device("cuda", torch.cuda.current_device())
    d = sim.dimension
    time_out = torch.empty(n_coarse, dtype=torch.float64, device=dev)
    u = torch.empty((n_coarse, d, d), dtype=torch.complex128, device=dev)
    sw = sweep if isinstance(sweep, torch.Tensor) else torch.as_tensor(sweep, dtype=torch.float64,
                                                                        device=dev)
    # three untimed rounds first: the device clock ramps up over the first
    # ~10 ms of load, which would otherwise favour the later candidates
    table = time_interleaved(
        {g: (lambda g=g: sim.launch_time_evolution(sw, t0, dt, dt_out, time_out, u, g))
         for g in candidates}, rounds=reps, warmup=3)
    best = min(table, key=table.get)
    with _lock:
        _memo[key] = best
    _store_disk(key, best)
    return best, table


def scan_chunk_candidates(total_rows: int, targets=(512, 1024, 2048, 4096, 8192)):
    """Chunk sizes (whole 1024-row tiles) giving about `targets` chunks."""
    from .batch import SCAN_TILE_ROWS

    tiles = max(1, -(-int(total_rows) // SCAN_TILE_ROWS))
    return sorted({SCAN_TILE_ROWS * max(1, -(-tiles // t)) for t in targets})


def tune_scan(plan, reps: int = 2, use_cache: bool = True, targets=(512, 1024, 2048, 4096, 8192),
              warmup: int = 1):
    """Best rows per chunk of a BatchPlan's state scan; applied to the plan
    (the costate scan of the adjoint uses the same chunking).  Times the
    scans over the plan's current operators."""
    candidates = scan_chunk_candidates(plan.total, targets) if plan.total else [0]
    key = _key(plan.sim, "scan", plan.n_sims, plan.total, plan.spin_out is not None,
               tuple(candidates))
    if use_cache:
        disk = _load_disk()
        if key in disk:
            plan.set_scan_chunk_rows(int(disk[key]))
            return plan.scan_chunk_rows, None
    # one workspace large enough for every candidate (the smallest chunks)
    plan.set_scan_chunk_rows(min(candidates))

    def run(c):
        plan.scan_chunk_rows = c
        plan.run_state_scan()
    table = time_interleaved({c: (lambda c=c: run(c)) for c in candidates}, rounds=reps,
                             warmup=warmup)
    best = min(table, key=table.get)
    plan.set_scan_chunk_rows(best)
    _store_disk(key, best)
    return best, table


def tune_batch(plan, reps: int = 2, use_cache: bool = True, lanes=None, items=(1, 4, 8, 16),
               warmup: int = 1, scan: bool = True):
    """Best (lanes_per_step, items_per_grab) for a BatchPlan, and with `scan`
    the state scan's chunk size (:func:`tune_scan`); applied to the plan.

    ``lanes`` defaults to :func:`lane_candidates` of the largest n_fine.
    """
    nf_max = int(plan.n_fine.max().item()) if plan.total else 1
    lanes = tuple(lanes) if lanes is not None else tuple(lane_candidates(nf_max))
    key = _key(plan.sim, "batch", plan.n_sims, plan.total, nf_max, lanes, tuple(items))
    best = None
    if use_cache:
        disk = _load_disk()
        if key in disk:
            best = tuple(map(int, disk[key]))
    table = None
    if best is None:
        def run(g, it):
            plan.lanes, plan.items_per_grab = g, it
            plan.run_time_evolution()
        table = time_interleaved({(g, it): (lambda g=g, it=it: run(g, it))
                                  for g in lanes for it in items}, rounds=reps, warmup=warmup)
        best = min(table, key=table.get)
        _store_disk(key, list(best))
    plan.lanes, plan.items_per_grab = best
    if scan and plan.total:
        plan.run_time_evolution()
        tune_scan(plan, reps=reps, use_cache=use_cache, warmup=warmup)
    return best, table
