"""Autotuner for the batched and plain steps.

Picks the grid cell edge, the bucket count (bucket_shift), the decide
workgroup size, the birds each decide workgroup scans per lane
(birds_per_thread), the lanes that share one bird's walk (decide_group) and
whether decide tests the bucket occupancy bitmap, runs as a persistent grid
or maps chunks XCD by XCD, whether it stages each chunk's positions in LDS,
whether cells are bucketed by the spatial hash,
whether the state stays in bucket order between steps and flies each chunk inside decide, and whether
that layout is rebuilt only every fourth step (lazy rebuilds), by timing
short runs of the real workload. Each candidate is timed over `steps` steps
after `warmup` untimed ones; the fastest few are timed again and the lowest
median kept. Results can be cached in a JSON file keyed by
the search version, the device name, the caller's workload label and the
workload's signature (flock sizes, steps, flags, search mode); a cached entry
is returned without timing anything.

Changing the cell edge changes the hash keys (the keys match the reference's
only at 4 m) but not the neighbour sets: the walk always covers the 2r box.
"""
from __future__ import annotations

import hashlib
import json
import os
import time

import numpy as np

from . import (FLAG_CANONICAL, FLAG_FLIGHT, FLAG_GRAPH, FLAG_LAZY, FLAG_OCC, FLAG_PERSISTENT, FLAG_RESIDENT, FLAG_SPATIAL, FLAG_UNFUSED,
               FLAG_WINDOW, FLAG_XCD, Batch, Flock, device_info)

BPT = (1, 2, 4, 8, 16)
BLOCKS = (64, 128, 256, 512)
GROUPS = (1, 2, 4, 8, 16, 64)
CELLS = (4.0, 3.0, 5.0, 6.0, 8.0, 10.0, 12.0)
SHIFTS = (0, 1, 2, 3, 4)
# on/off tunables and the flag each one sets
SWITCHES = {"occ": FLAG_OCC, "resident": FLAG_RESIDENT, "spatial": FLAG_SPATIAL, "persist": FLAG_PERSISTENT,
            "xcd": FLAG_XCD, "window": FLAG_WINDOW, "flight": FLAG_FLIGHT, "lazy": FLAG_LAZY}
# Part of every cache key: raised whenever the searched space or the kernels'
# costs change, so an entry found for other kernels or a smaller space is not
# reused.
SEARCH_VERSION = 12
FINALISTS = 6
CONFIRM = 3


def _signature(kind: str, sizes, steps: int, flags: int) -> str:
    h = hashlib.sha1(np.asarray(sizes, np.int64).tobytes()).hexdigest()[:16]
    return f"{kind}:{h}:{steps}:{flags}"


def _load(path):
    try:
        with open(path) as fp:
            return json.load(fp)
    except (OSError, ValueError):
        return {}


def _time_batch(starts, times, prm: dict, flags: int, warmup: int, steps: int) -> float:
    total = sum(int(w.shape[0]) for w in starts)
    buckets = sum(_buckets(int(w.shape[0])) for w in starts)
    with Batch(total + 256 * len(starts), buckets + 4096 * len(starts), len(starts), flags=flags, **prm) as b:
        for w, t in zip(starts, times):
            b.add(w, t)
        b.step(warmup)
        b.step(1)  # sync point after warm-up
        t0 = time.perf_counter()
        b.step(steps)
        return (time.perf_counter() - t0) / steps


def _time_plain(words, time0: float, prm: dict, flags: int, warmup: int, steps: int) -> float:
    with Flock.from_words(words, time0, flags=flags, **prm) as f:
        f.step(warmup)
        f.sync()
        t0 = time.perf_counter()
        f.step(steps)
        f.sync()
        return (time.perf_counter() - t0) / steps


def _buckets(n: int) -> int:
    want = 1
    while want < 2 * n:
        want <<= 1
    return min(max(want, 1 << 16), 1 << 22)


def tune(starts, kind: str = "batch", flags: int = FLAG_GRAPH, warmup: int = 3, steps: int = 8,
         cache_path: str | None = None, cells=CELLS, progress=None, thorough: bool = True, times=None,
         shifts=SHIFTS, spatial=(0, 1), label: str = "", refresh: bool = False, cached_only: bool = False):
    """Returns {"params": {...}, "flags": f, "ms_per_step": x, "trials": [...]}
    ("trials" only when searched; a cache entry holds the chosen point)
    for the batch of `starts` (kind "batch") or the single flock starts[0]
    ("plain"). `flags` is the caller's flags plus the switches the search
    turned on (FLAG_OCC, FLAG_RESIDENT, FLAG_SPATIAL, FLAG_PERSISTENT, FLAG_XCD); `spatial=(0,)` keeps
    the reference hash. With `thorough` the launch shape is searched
    exhaustively (group x birds_per_thread x block) before and after the
    cell and bucket searches; otherwise one coordinate sweep per tunable.
    `times` are the flocks' clocks (default 0), e.g. of states downloaded
    after some steps. `progress`, if given, is called with one line of text
    per trial. `label` names the workload in the cache key (flocks of the
    same sizes but another distribution want their own entry); `refresh`
    searches again and replaces a cached entry. A result read from the cache
    carries "from_cache": True. With `cached_only` nothing is searched: a
    workload the cache lacks returns None."""
    times = [0.0] * len(starts) if times is None else [float(t) for t in times]
    sizes = [int(w.shape[0]) for w in starts]
    key = (f"v{SEARCH_VERSION}|{device_info()['name']}|{label}|" + _signature(kind, sizes, steps, flags) +
           (":t" if thorough else ":q"))
    cache = _load(cache_path) if cache_path else {}
    if key in cache and not refresh:
        return {**cache[key], "from_cache": True}
    if cached_only:
        return None

    def timer(prm):
        prm = dict(prm)
        f = flags
        for name, bit in SWITCHES.items():
            f |= bit if prm.pop(name) else 0
        if f & FLAG_LAZY and (not (f & FLAG_RESIDENT and f & FLAG_FLIGHT) or f & (FLAG_WINDOW | FLAG_UNFUSED)):
            return float("inf")  # lazy rebuilds: resident layout, fused flight, no window
        if kind == "batch":
            return _time_batch(starts, times, prm, f, warmup, steps)
        return _time_plain(starts[0], times[0], prm, f, warmup, steps)

    best = {"cell": float(cells[0]), "decide_block": 256, "birds_per_thread": 8, "decide_group": 1, "bucket_shift": 0,
            "occ": 0, "resident": 0, "spatial": 0, "persist": 0, "xcd": 0, "window": 0, "flight": 0, "lazy": 0}
    trials = []
    seen = {}

    def run(prm):
        k = tuple(sorted(prm.items()))
        if k in seen:
            return seen[k]
        t = timer(prm)
        seen[k] = t
        trials.append({**prm, "ms": t * 1e3})
        if progress:
            progress(f"{kind} trial {len(trials)}: {prm} {t * 1e3:.4f} ms/step")
        return t

    best_t = run(best)

    def sweep(names_values):
        """Tries every combination of the listed tunables around the best."""
        nonlocal best, best_t
        combos = [{}]
        for name, values in names_values:
            combos = [{**c, name: v} for c in combos for v in values]
        for c in combos:
            cand = {**best, **c}
            t = run(cand)
            if t < best_t:
                best, best_t = cand, t

    shape = [("decide_group", GROUPS), ("birds_per_thread", BPT), ("decide_block", BLOCKS)]
    if thorough:
        sweep(shape)
    else:
        for nv in shape:
            sweep([nv])
    # the resident layout keeps no canonical order
    resident = (0, 1) if not (flags & FLAG_CANONICAL) else (0,)
    # The fused flight exists only in the resident layout, lazy rebuilds only
    # with the fused flight and without the window: the three are swept
    # together (a batch's fused flight alone is slower than the separate pass,
    # so a flight sweep on its own would never reach lazy rebuilds).
    for nvs in ((("resident", resident),), (("spatial", spatial),), (("cell", cells),), (("bucket_shift", shifts),),
                (("occ", (0, 1)),), (("persist", (0, 1)),), (("xcd", (0, 1)),),
                (("window", (0, 1)), ("flight", resident), ("lazy", resident)), (("cell", cells),),
                (("bucket_shift", shifts),)):
        sweep(list(nvs))
    # the first shape sweep ran on the default layout: the shape is searched
    # again on the chosen one (resident lazy batches prefer other blocks)
    if thorough:
        sweep(shape[:2])
    else:
        for nv in shape:
            sweep([nv])
    # Short trials are noisy, so a search can settle a few percent away from
    # the best point: the fastest FINALISTS distinct points are timed
    # CONFIRM times more on fresh flocks and the lowest median wins.
    finalists = sorted(seen.items(), key=lambda kv: kv[1])[:FINALISTS]
    confirmed = []
    for k, t in finalists:
        prm = dict(k)
        ts = sorted([t] + [timer(prm) for _ in range(CONFIRM)])
        med = ts[len(ts) // 2]
        confirmed.append((med, prm))
        trials.append({**prm, "ms": med * 1e3, "confirm": [x * 1e3 for x in ts]})
        if progress:
            progress(f"{kind} confirm: {prm} median {med * 1e3:.4f} ms/step")
    best_t, best = min(confirmed, key=lambda c: c[0])
    params = {k: v for k, v in best.items() if k not in SWITCHES}
    out_flags = flags
    for name, bit in SWITCHES.items():
        out_flags |= bit if best[name] else 0
    out = {"params": params, "flags": out_flags, "ms_per_step": best_t * 1e3,
           "trials": trials, "signature": key}
    if cache_path:
        # the cache keeps the chosen point; the trial log is returned only
        # entries of an older search space are dropped: their points were
        # found among fewer choices or for other kernels
        cache = {k: v for k, v in cache.items() if k.startswith(f"v{SEARCH_VERSION}|")}
        cache[key] = {"params": params, "flags": out_flags, "ms_per_step": best_t * 1e3, "signature": key,
                      "trial_count": len(trials), "confirmed": [c[0] * 1e3 for c in sorted(confirmed, key=lambda c: c[0])]}
        os.makedirs(os.path.dirname(os.path.abspath(cache_path)), exist_ok=True)
        with open(cache_path, "w") as fp:
            json.dump(cache, fp, indent=1)
    return out
