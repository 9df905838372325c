#!/usr/bin/env python3
"""Benchmark of the gfx950 flock step.

Default workload: the plain single-flock step at 1,000,000 birds (uniform
464.16 m cube about the roost = 0.64 birds per 4 m cell, isotropic headings at
10 +- 1 m/s, seed 1), 20 warm-up + 20 timed steps, state resident in HBM.
Prints one JSON line (the metric is bird-steps/s of the whole step).

    python bench.py [--gpus 1] [--steps 20] [--warmup 20] [--birds N] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "murmuration-kernels_amd"))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E peak


def step_bytes(n: int, buckets: int, due: int) -> dict:
    """Algorithmic HBM bytes per pass (FP32 SoA; n birds, B buckets, d due)."""
    return {
        "hash_count": 16 * n + 8 * buckets,
        "scan": 8 * buckets,
        "scatter": 56 * n + 4 * buckets,
        "decide": 28 * n + 4 * buckets + 32 * due,
        "integrate": 72 * n,
    }


def cpu_baseline(words: np.ndarray, n: int, sample: int) -> dict:
    """The oracle's brute-force step (single thread) on a bounded sample:
    `sample` due birds' decisions plus step_bird for every bird, extrapolated
    to one whole step (all due birds)."""
    from oracle import flockref as fr

    t = np.float32(0.0) + np.float32(0.005)
    due = np.nonzero(t >= words[:, 12])[0]
    pick = due[: min(sample, len(due))]
    t0 = time.perf_counter()
    for i in pick:
        fr.decide(words, int(i))
    t_dec = (time.perf_counter() - t0) / max(1, len(pick))
    t0 = time.perf_counter()
    for i in range(0, n, max(1, n // 200_000)):
        fr.step_bird(words[i])
    per_bird = (time.perf_counter() - t0) / len(range(0, n, max(1, n // 200_000)))
    step_s = t_dec * len(due) + per_bird * n
    return {
        "value": n / step_s,
        "unit": "bird-steps/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"one step of the C oracle (brute force, single thread): decide timed on {len(pick)} of the "
                   f"{len(due)} due birds, step_bird timed on a stride sample of the {n} birds, "
                   f"extrapolated to the whole step ({step_s:.1f} s)"),
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--birds", type=int, default=1_000_000)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--flags", type=int, default=4, help="fh flags for the timed run (default: graph)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=2500)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(flocks would shard by HBM bytes across GPUs, see the design notes)", file=sys.stderr)
        return 2

    import flock_hip as fh

    n = args.birds
    side = fh.cube_side_for(n)
    words = fh.gen_cube(n, side, args.seed)
    dev = fh.device_info()

    # Timed run: fused step replayed from a hipGraph; wall clock over K steps,
    # synchronised on both sides, inputs resident in HBM.
    with fh.Flock.from_words(words, flags=args.flags) as f:
        buckets = f.buckets
        f.step(args.warmup)
        f.sync()
        t0 = time.perf_counter()
        f.step(args.steps)
        f.sync()
        wall = time.perf_counter() - t0
    ms_per_step = wall * 1000.0 / args.steps

    # Per-pass breakdown with hipEvents on the library's stream (same start).
    with fh.Flock.from_words(words, flags=args.flags & ~fh.FLAG_GRAPH) as f:
        f.step(args.warmup)
        passes = f.step(args.steps, 0, timing=True)
    due_last = int(passes["due_last"])  # birds that decided in the last timed step

    nbytes = step_bytes(n, buckets, due_last)
    dec_ms = passes["decide"]
    achieved = nbytes["decide"] / (dec_ms * 1e-3) / 1e9
    step_gbs = sum(nbytes.values()) / (passes["step_total"] * 1e-3) / 1e9
    vgpr = fh.kernel_info("decide")["vgprs"]

    out = {
        "metric": "bird-steps/s served+batched vs plain; decide ms @1M birds; p50/p99 latency",
        "value": n / (ms_per_step / 1000.0),
        "unit": "bird-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"plain single-flock step, {n} birds, uniform {side:.2f} m cube (0.64 birds per 4 m cell), "
                        f"isotropic headings 10+-1 m/s, seed {args.seed}",
            "birds": n,
            "buckets": buckets,
            "cell_m": 4.0,
            "flags": args.flags,
            "device": dev["name"],
        },
        "decide_ms": dec_ms,
        "passes_ms": {k: passes[k] for k in ("grid_histogram", "grid_scan", "grid_scatter", "decide",
                                            "integrate", "step_total")},
        "decide_vgprs": vgpr,
        "due_birds": due_last,
        "step_roofline": {"achieved": step_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": step_gbs / HBM_PEAK_GBS, "bytes": sum(nbytes.values())},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": "decide",
                     "bytes_per_launch": nbytes["decide"]},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(words, n, args.cpu_sample)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
