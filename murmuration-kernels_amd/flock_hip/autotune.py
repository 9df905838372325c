"""Autotuner for the batched and plain steps.

Picks the grid cell edge, the decide workgroup size, the birds each decide
lane scans (birds_per_thread) and the lanes that share one bird's walk
(decide_group) by timing short runs of the real workload: a coordinate
search (group, birds_per_thread, workgroup size, cell edge, bucket count,
then cell edge and group again),
each candidate timed over `steps` steps after `warmup` untimed ones. Results
are cached in a JSON file keyed by the workload's signature.

Changing the cell edge changes the hash keys (the keys match the reference's
only at 4 m) but not the neighbour sets: the walk always covers the 2r box.
"""
from __future__ import annotations

import hashlib
import json
import os
import time

import numpy as np

from . import FLAG_GRAPH, Batch, Flock

BPT = (2, 4, 8, 16)
BLOCKS = (64, 128, 256, 512)
GROUPS = (1, 2, 4, 8)
CELLS = (4.0, 3.0, 5.0, 6.0, 8.0, 10.0, 12.0)
SHIFTS = (0, 1, 2, 3, 4)


def _signature(kind: str, sizes, steps: int, flags: int) -> str:
    h = hashlib.sha1(np.asarray(sizes, np.int64).tobytes()).hexdigest()[:16]
    return f"{kind}:{h}:{steps}:{flags}"


def _load(path):
    try:
        with open(path) as fp:
            return json.load(fp)
    except (OSError, ValueError):
        return {}


def _time_batch(starts, prm: dict, flags: int, warmup: int, steps: int) -> float:
    total = sum(int(w.shape[0]) for w in starts)
    buckets = sum(_buckets(int(w.shape[0])) for w in starts)
    with Batch(total + 256 * len(starts), buckets + 4096 * len(starts), len(starts), flags=flags, **prm) as b:
        for w in starts:
            b.add(w)
        b.step(warmup)
        b.step(1)  # sync point after warm-up
        t0 = time.perf_counter()
        b.step(steps)
        return (time.perf_counter() - t0) / steps


def _time_plain(words, prm: dict, flags: int, warmup: int, steps: int) -> float:
    with Flock.from_words(words, flags=flags, **prm) as f:
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
         cache_path: str | None = None, cells=CELLS, progress=None) -> dict:
    """Returns {"params": {...}, "ms_per_step": x, "trials": [...]} for the
    batch of `starts` (kind "batch") or the single flock starts[0] ("plain").
    `progress`, if given, is called with one line of text per trial."""
    sizes = [int(w.shape[0]) for w in starts]
    key = _signature(kind, sizes, steps, flags)
    cache = _load(cache_path) if cache_path else {}
    if key in cache:
        return cache[key]
    timer = (lambda prm: _time_batch(starts, prm, flags, warmup, steps)) if kind == "batch" else \
        (lambda prm: _time_plain(starts[0], prm, flags, warmup, steps))
    best = {"cell": 4.0, "decide_block": 256, "birds_per_thread": 8, "decide_group": 1, "bucket_shift": 0}
    trials = []

    def run(prm):
        t = timer(prm)
        trials.append({**prm, "ms": t * 1e3})
        if progress:
            progress(f"{kind} trial {len(trials)}: {prm} {t * 1e3:.4f} ms/step")
        return t

    best_t = run(best)
    seen = {tuple(sorted(best.items()))}
    for name, values in (("decide_group", GROUPS), ("birds_per_thread", BPT), ("decide_block", BLOCKS),
                         ("cell", cells), ("bucket_shift", SHIFTS), ("cell", cells), ("decide_group", GROUPS)):
        for v in values:
            if v == best[name]:
                continue
            cand = {**best, name: v}
            if tuple(sorted(cand.items())) in seen:
                continue
            seen.add(tuple(sorted(cand.items())))
            t = run(cand)
            if t < best_t:
                best, best_t = cand, t
    out = {"params": best, "ms_per_step": best_t * 1e3, "trials": trials, "signature": key}
    if cache_path:
        cache[key] = out
        os.makedirs(os.path.dirname(os.path.abspath(cache_path)), exist_ok=True)
        with open(cache_path, "w") as fp:
            json.dump(cache, fp, indent=1)
    return out
