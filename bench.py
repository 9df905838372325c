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
import glob
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "murmuration-kernels_amd"))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E peak
BATCH_TUNE_STEPS = 100  # timed steps of a batched-leg tuning trial (the leg's default length)
_T0 = time.perf_counter()


def log(msg: str) -> None:
    """One progress line on stderr (the JSON result goes to stdout)."""
    print(f"[bench {time.perf_counter() - _T0:7.1f} s] {msg}", file=sys.stderr, flush=True)


def step_bytes(n: int, buckets: int, due: int, resident: bool, rebuild_every: int = 1) -> dict:
    """Algorithmic HBM bytes per pass (n birds, B buckets, d due), keyed like
    the per-pass times: the standard step (histogram, scan, scatter of the
    sorted copies, decide, integrate) or the resident one (decide, integrate
    counting the movers, then scan and move as one 'grid_scatter' pass, which
    lazy rebuilds run once every `rebuild_every` steps: its bytes per step are
    spread over them, like its mean time)."""
    if resident:
        return {"decide": 28 * n + 4 * buckets + 32 * due, "integrate": 72 * n,
                "grid_scatter": (20 * buckets + 120 * n) / rebuild_every}
    return {
        "grid_histogram": 16 * n + 8 * buckets,
        "grid_scan": 8 * buckets,
        "grid_scatter": 56 * n + 4 * buckets,
        "decide": 28 * n + 4 * buckets + 32 * due,
        "integrate": 72 * n,
    }


def profiled_traffic(params: dict, flags: int):
    """HBM bytes per decide launch measured by the committed rocprofv3 counter
    runs (FETCH_SIZE and WRITE_SIZE, each scaled by a copy-kernel
    calibration): the 'traffic:' line of a summary under profiles/ taken at
    these tunables and flags, or None when no profile matches them."""
    want = {**{k: float(v) for k, v in params.items()}, "flags": float(flags)}
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        with open(path) as fp:
            for line in fp:
                if not line.startswith("traffic: kernel=decide "):
                    continue
                fields = dict(f.split("=", 1) for f in line.split()[1:])
                got = {k: float(v) for k, v in (kv.split("=") for kv in fields["params"].split(","))}
                if got == want:
                    return {"bytes": float(fields["bytes"]), "profile": os.path.basename(path)}
    return None


def full_prm_block(prm: dict) -> int:
    """decide workgroup size of a tunables dict (the library default: 256)."""
    return int(prm.get("decide_block", 256) or 256)


def tune_cache(args, label: str) -> dict:
    """Cache arguments of autotune.tune for one leg's workload."""
    return {"cache_path": args.tune_cache or None, "label": label, "refresh": args.retune}


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fp:
            for line in fp:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _cpu_baseline_costs(flocks, due_sample: int, bird_sample: int) -> dict:
    """Single-thread costs of the oracle's brute-force step (model.rs:182-200)
    on the given start states: seconds per candidate of one decision (a due
    bird tests every other bird), timed on up to `due_sample` due birds of
    each flock, and seconds per bird of step_bird, on a stride sample of up
    to `bird_sample` birds of each."""
    from oracle import flockref as fr

    t = np.float32(0.0) + np.float32(0.005)
    dec_s = cand = 0.0
    bird_s = birds = 0.0
    for w in flocks:
        n = int(w.shape[0])
        due = np.nonzero(t >= w[:, 12])[0][:due_sample]
        t0 = time.perf_counter()
        for i in due:
            fr.decide(w, int(i))
        dec_s += time.perf_counter() - t0
        cand += len(due) * n
        idx = range(0, n, max(1, n // bird_sample))
        t0 = time.perf_counter()
        for i in idx:
            fr.step_bird(w[i])
        bird_s += time.perf_counter() - t0
        birds += len(idx)
    return {"per_candidate_s": dec_s / max(1.0, cand), "per_bird_s": bird_s / max(1.0, birds),
            "timed_s": dec_s + bird_s}


def _cpu_baseline_step_s(costs: dict, n: int, due: float) -> float:
    """Modelled single-thread seconds of one oracle step of n birds with
    `due` of them deciding."""
    return due * n * costs["per_candidate_s"] + n * costs["per_bird_s"]


def cpu_baseline(words: np.ndarray, n: int, sample: int) -> dict:
    """The oracle's brute-force step (single thread) on a bounded sample:
    `sample` due birds' decisions plus step_bird for every bird, extrapolated
    to one whole step (all due birds). Beside it, labelled as a variant, the
    same decisions in parallel over the due birds on up to 16 host threads
    (threads x `sample` of them)."""
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
    threads = max(1, min(16, os.cpu_count() or 1))
    many = due[: min(sample * threads, len(due))]
    t0 = time.perf_counter()
    fr.decide_many(words, many, threads=threads)
    t_dec_mt = (time.perf_counter() - t0) / max(1, len(many))
    step_mt = t_dec_mt * len(due) + per_bird * n / threads
    return {
        "value": n / step_s,
        "unit": "bird-steps/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"one step of the C oracle (brute force, single thread): decide timed on {len(pick)} of the "
                   f"{len(due)} due birds, step_bird timed on a stride sample of the {n} birds, "
                   f"extrapolated to the whole step ({step_s:.1f} s)"),
        "multicore_variant": {
            "value": n / step_mt,
            "threads": threads,
            "sample": (f"variant: the same brute-force decide in parallel over {len(many)} due birds (OpenMP, "
                       f"{threads} threads), step_bird's single-thread time divided by the threads, "
                       f"extrapolated to the whole step ({step_mt:.1f} s)"),
        },
        "host": {"nproc": os.cpu_count(), "cpu_model": _cpu_model()},
    }


def algo_step_bytes(n: int, due: float) -> float:
    """Algorithmic HBM bytes of one step of one flock: 172 n + 24 B + 32 d
    (histogram 16n + 8B, scan 8B, scatter 56n + 4B, decide 28n + 4B + 32d,
    integrate 72n; B = buckets_for(n), the reference's grid)."""
    import flock_hip as fh

    return 172.0 * n + 24.0 * fh.buckets_for(int(n)) + 32.0 * due


def hbm_roofline(nbytes: float, seconds: float, what: str) -> dict:
    gbs = nbytes / seconds / 1e9
    return {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
            "bytes": nbytes, "basis": what}


def plain_point(args) -> tuple[dict, int]:
    """Tunables and flags of the plain 1M-bird step (the autotuner's cached
    point for the default plain workload)."""
    import flock_hip as fh
    from flock_hip import autotune, workloads

    n = 1_000_000
    tuned = autotune.tune([workloads.plain(n, 1)], "plain", flags=fh.FLAG_GRAPH, warmup=20, steps=20,
                          progress=log, **{**tune_cache(args, f"plain-cube-{n}-seed1"), "refresh": False})
    return tuned["params"], tuned["flags"]


def plain_tunables(args, words, cached_only: bool = False) -> tuple[dict, dict | None, int] | None:
    """Tunables and flags of the plain step on `words`: those of --params, the
    library defaults (--no-tune), or the autotuner's (its cached point when
    the cache holds this workload). Returns (params, autotune result or None,
    flags); None when `cached_only` and the cache lacks the workload."""
    prm, tuned, flags = {}, None, args.flags
    if args.params:
        for item in args.params.split(","):
            k, v = item.split("=")
            if k == "flags":
                flags = int(v)
            else:
                prm[k] = float(v) if k == "cell" else int(v)
    elif not args.no_tune:
        from flock_hip import autotune

        # trials time the same window as the timed run (radii grow each decision)
        log(f"plain leg: autotune at {words.shape[0]} birds")
        tuned = autotune.tune([words], "plain", flags=args.flags, warmup=args.warmup, steps=args.steps, progress=log,
                              cached_only=cached_only,
                              **tune_cache(args, f"plain-cube-{words.shape[0]}-seed{args.seed}"))
        if tuned is None:
            return None
        prm, flags = tuned["params"], tuned["flags"]
    return prm, tuned, flags


def traffic_probe(args) -> int:
    """The program that live_traffic runs under rocprofv3 --pmc: a 256 MiB
    device-to-device copy (the counters' calibration), then the plain step's
    decide as the timed run has it (same workload and tunables, pass by pass
    so each launch is its own dispatch) for warm-up + timed steps. Prints the
    tunables it ran as one JSON line."""
    import torch

    import flock_hip as fh

    x = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    y.copy_(x)
    torch.cuda.synchronize()
    del x, y
    words = fh.gen_cube(args.birds, fh.cube_side_for(args.birds), args.seed)
    point = plain_tunables(args, words, cached_only=True)
    if point is None:  # no search under the profiler: the parent falls back
        print("traffic probe: the autotuner cache lacks this workload", file=sys.stderr)
        return 3
    prm, _, flags = point
    with fh.Flock.from_words(words, flags=flags & ~fh.FLAG_GRAPH, **prm) as f:
        f.step(args.warmup)
        f.step(args.steps)
        f.sync()
    print(json.dumps({"params": prm, "flags": flags, "copy_bytes": 256 << 20}), flush=True)
    return 0


def live_traffic(args) -> dict | None:
    """HBM bytes per decide launch of the plain step, measured in this bench
    run: two rocprofv3 --pmc runs of `bench.py --traffic-probe` (FETCH_SIZE,
    then WRITE_SIZE, in KiB per dispatch), averaged over the timed steps'
    decide dispatches and each scaled by what the counter reads for the
    probe's 256 MiB copy (FETCH_SIZE reads half the bytes a copy fetches,
    WRITE_SIZE all it writes). Runs before this process touches the GPU (the
    probe is a child process). None when rocprofv3 is missing or a run fails."""
    import csv
    import shutil
    import signal
    import subprocess
    import tempfile

    exe = shutil.which("rocprofv3")
    if not exe:
        return None
    probe = [sys.executable, os.path.abspath(__file__), "--traffic-probe", "--birds", str(args.birds),
             "--seed", str(args.seed), "--warmup", str(args.warmup), "--steps", str(args.steps),
             "--flags", str(args.flags), "--tune-cache", args.tune_cache]
    if args.params:
        probe += ["--params", args.params]
    if args.no_tune:
        probe.append("--no-tune")
    out = {}
    with tempfile.TemporaryDirectory(dir="/tmp") as tmp:
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            d = os.path.join(tmp, counter)
            cmd = [exe, "--pmc", counter, "--output-format", "csv", "-d", d, "-o", "t", "--", *probe]
            log(f"plain leg: live traffic, rocprofv3 --pmc {counter}")
            p = subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                                 start_new_session=True, env={**os.environ, "TMPDIR": "/tmp"})
            try:
                so, se = p.communicate(timeout=240)
            except subprocess.TimeoutExpired:
                os.killpg(p.pid, signal.SIGKILL)
                p.communicate()
                log(f"plain leg: live traffic run ({counter}) timed out")
                return None
            files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
            ran = [ln for ln in so.splitlines() if ln.startswith("{")]
            if p.returncode != 0 or not files or not ran:
                log(f"plain leg: live traffic run ({counter}) failed: rc {p.returncode}, {se.strip()[-300:]}")
                return None
            ran = json.loads(ran[-1])
            decide, copy = {}, []
            for path in files:
                with open(path) as fp:
                    for r in csv.DictReader(fp):
                        if r["Counter_Name"] != counter:
                            continue
                        if "k_decide<" in r["Kernel_Name"]:
                            decide[int(r["Dispatch_Id"])] = float(r["Counter_Value"])
                        elif "copyBuffer" in r["Kernel_Name"] and float(r["Counter_Value"]) > 65536:
                            copy.append(float(r["Counter_Value"]))
            timed = [decide[k] for k in sorted(decide)][-args.steps:]
            if len(timed) < args.steps or not copy:
                log(f"plain leg: live traffic run ({counter}): {len(decide)} decide and {len(copy)} copy dispatches")
                return None
            out[counter] = {"kib": float(np.mean(timed)), "scale": (ran["copy_bytes"] / 1024) / float(np.mean(copy))}
            out["params"], out["flags"] = ran["params"], ran["flags"]
    f, w = out["FETCH_SIZE"], out["WRITE_SIZE"]
    return {"bytes": (f["kib"] * f["scale"] + w["kib"] * w["scale"]) * 1024.0, "fetch_kib": f["kib"],
            "write_kib": w["kib"], "fetch_scale": f["scale"], "write_scale": w["scale"], "params": out["params"],
            "flags": out["flags"], "dispatches": args.steps}


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
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-reference", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="plain leg at the default tunables")
    ap.add_argument("--tune-cache", default=os.path.join(ROOT, "murmuration-kernels_amd", "flock_hip",
                                                         "tuned_mi355x.json"),
                    help="autotuner results by device and workload; a workload found there is not searched again")
    ap.add_argument("--retune", action="store_true", help="search every leg's tunables again (updates the cache)")
    ap.add_argument("--params", default="",
                    help="plain leg at these tunables instead of autotuning, e.g. "
                         "'cell=4,decide_block=512,birds_per_thread=4,decide_group=4,bucket_shift=0,flags=4'")
    ap.add_argument("--no-clumps", action="store_true")
    ap.add_argument("--no-c1", action="store_true")
    ap.add_argument("--only", choices=("clumps", "batched", "served", "reference", "c1"),
                    help="run this leg alone (for its profiles, or to retune it with --retune)")
    ap.add_argument("--no-plain", action="store_true",
                    help="batched leg without its per-flock comparators (a kernel trace of the batch alone)")
    ap.add_argument("--clump-birds", type=int, default=10_000_000)
    ap.add_argument("--clump-steps", type=int, default=20)
    ap.add_argument("--clump-warmup", type=int, default=0)
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--flocks", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=100)
    ap.add_argument("--requests", type=int, default=2000)
    ap.add_argument("--slots", type=int, default=64)
    ap.add_argument("--load", type=float, default=0.6, help="offered load / measured server capacity")
    ap.add_argument("--plain-sample", type=int, default=32, help="served requests also stepped alone")
    ap.add_argument("--copy-cal", action="store_true",
                    help="only run a 256 MiB device-to-device copy (calibrates FETCH_SIZE/WRITE_SIZE counters)")
    ap.add_argument("--no-live-traffic", action="store_true",
                    help="take the decide launch's HBM traffic from a committed profile instead of measuring it")
    ap.add_argument("--traffic-probe", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.traffic_probe:
        return traffic_probe(args)
    if args.copy_cal:
        import torch

        x = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_per_launch": x.numel() * 4, "launches": 3}))
        return 0
    if args.only:
        legs = {"clumps": lambda: clumps_leg(args), "batched": lambda: batched_leg(args, plain_point(args)),
                "served": lambda: served_leg(args, plain_point(args)), "reference": lambda: reference_leg(args),
                "c1": lambda: c1_leg(args)}
        print(json.dumps({args.only: legs[args.only]()}), flush=True)
        return 0
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(flocks would shard by HBM bytes across GPUs, see the design notes)", file=sys.stderr)
        return 2

    # the traffic runs start before this process touches the GPU
    live = None if (args.no_live_traffic or args.retune) else live_traffic(args)

    import flock_hip as fh

    n = args.birds
    side = fh.cube_side_for(n)
    words = fh.gen_cube(n, side, args.seed)
    dev = fh.device_info()
    prm, tuned, flags = plain_tunables(args, words)

    def timed_run(params: dict, flags: int) -> tuple[float, int]:
        """Fused step replayed from a hipGraph; wall clock over K steps,
        synchronised on both sides, inputs resident in HBM."""
        with fh.Flock.from_words(words, flags=flags, **params) as f:
            f.step(args.warmup)
            f.sync()
            t0 = time.perf_counter()
            f.step(args.steps)
            f.sync()
            return (time.perf_counter() - t0) * 1000.0 / args.steps, f.buckets

    # three runs of the same protocol (scripts/cloud-session.sh:62-63 reports
    # three per size); the value is their median
    runs = [timed_run(prm, flags) for _ in range(3)]
    buckets = runs[0][1]
    run_ms = [r[0] for r in runs]
    ms_per_step = float(np.median(run_ms))
    default_ms = timed_run({}, args.flags)[0] if tuned else ms_per_step
    log(f"plain leg: {ms_per_step:.4f} ms/step tuned (runs {', '.join(f'{x:.4f}' for x in run_ms)}), "
        f"{default_ms:.4f} ms/step at the default tunables")

    # Per-pass breakdown with hipEvents on the library's stream (same start).
    with fh.Flock.from_words(words, flags=flags & ~fh.FLAG_GRAPH, **prm) as f:
        f.step(args.warmup)
        passes = f.step(args.steps, 0, timing=True)
    due_last = int(passes["due_last"])  # birds that decided in the last timed step

    lazy = bool(flags & fh.FLAG_LAZY)
    nbytes = step_bytes(n, buckets, due_last, bool(flags & fh.FLAG_RESIDENT), fh.LAZY_PERIOD if lazy else 1)
    # With the fused flight, the decide launch also does the integrate pass's
    # work (its event interval is then empty): the dominant kernel's bytes are
    # both passes'. decide alone is timed in a run without the fusion.
    fused = bool(flags & fh.FLAG_FLIGHT) and bool(flags & fh.FLAG_RESIDENT)
    dec_ms = passes["decide"]
    dec_bytes = nbytes["decide"] + (nbytes["integrate"] if fused else 0)
    achieved = dec_bytes / (dec_ms * 1e-3) / 1e9
    step_gbs = sum(nbytes.values()) / (passes["step_total"] * 1e-3) / 1e9
    if fused:
        pass_gbs = {"decide+flight": achieved,
                    "grid_scatter": nbytes["grid_scatter"] / (passes["grid_scatter"] * 1e-3) / 1e9}
        with fh.Flock.from_words(words, flags=flags & ~(fh.FLAG_GRAPH | fh.FLAG_FLIGHT | fh.FLAG_LAZY), **prm) as f:
            f.step(args.warmup)
            decide_alone_ms = f.step(args.steps, 0, timing=True)["decide"]
    else:
        pass_gbs = {k: b / (passes[k] * 1e-3) / 1e9 for k, b in nbytes.items() if passes[k] > 0}
        decide_alone_ms = dec_ms
    dec_inst = (f"decide:{full_prm_block(prm)}:{prm.get('decide_group', 1)}:"
                f"{('m' if lazy else 'f' if flags & fh.FLAG_FLIGHT else 'r') if flags & fh.FLAG_RESIDENT else 's'}")
    kinfo = fh.kernel_info(dec_inst)
    vgpr = kinfo["vgprs"]
    # waves per SIMD the register file allows (512 VGPRs per lane, granule 8)
    waves = min(8, 512 // (((vgpr + 7) // 8) * 8)) if vgpr else None
    full_prm = {"cell": 4.0, "decide_block": 256, "birds_per_thread": 8, "decide_group": 1, "bucket_shift": 0, **prm}
    if live and live["flags"] == flags and live["params"] == prm:
        traffic = {**live, "source": "measured in this run: rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE, of "
                                     "bench.py --traffic-probe (same workload and tunables, timed steps' decide "
                                     "dispatches), each scaled by the probe's 256 MiB copy"}
    else:
        traffic = profiled_traffic(full_prm, flags)
        if traffic:
            traffic["source"] = f"profiles/{traffic['profile']} (counter runs at these tunables)"

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
            "flags": flags,
            "params": full_prm,
            "autotuned": tuned is not None,
            "tunables_from_cache": bool(tuned and tuned.get("from_cache")),
            "device": dev["name"],
        },
        "runs_ms_per_step": run_ms,
        "runs_spread": (max(run_ms) - min(run_ms)) / ms_per_step,
        "ms_per_step_default_params": default_ms,
        "decide_ms": decide_alone_ms,
        "decide_kernel_ms": dec_ms,
        "passes_ms": {k: passes[k] for k in ("grid_histogram", "grid_scan", "grid_scatter", "decide",
                                            "integrate", "step_total")},
        "passes_bytes": nbytes,
        "passes_gbs": pass_gbs,
        "decide_kernel": {"instance": dec_inst, "vgprs": vgpr, "scratch_bytes": kinfo["scratch_bytes"],
                          "waves_per_simd_by_vgprs": waves},
        "due_birds": due_last,
        "step_roofline": {"achieved": step_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": step_gbs / HBM_PEAK_GBS, "bytes": sum(nbytes.values())},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic["bytes"] if traffic else None,
                     "kernel": "decide with fused flight" if fused else "decide", "bytes_per_launch": dec_bytes,
                     "traffic_source": traffic["source"] if traffic else None,
                     "traffic_detail": ({k: traffic[k] for k in ("fetch_kib", "write_kib", "fetch_scale",
                                                                  "write_scale", "dispatches")}
                                        if traffic and "fetch_kib" in traffic else None)},
    }
    if not args.no_cpu:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(words, n, args.cpu_sample)
    del words
    if not args.no_c1:
        log("C1 leg: 10k birds, 20 + 20 steps")
        out["c1"] = c1_leg(args)
    if not args.no_reference:
        log("reference-protocol leg")
        out["reference_protocol"] = reference_leg(args)
    if not args.no_batched:
        log("batched leg")
        out["batched"] = batched_leg(args, (prm, flags))
        log(f"batched leg: speedup {out['batched']['speedup']:.2f}x over plain per-flock stepping")
        if not args.no_served:
            log("served leg")
            out["served"] = served_leg(args, (prm, flags))
    if not args.no_clumps:
        log("clumps leg")
        out["clumps"] = clumps_leg(args)
    log("done")
    print(json.dumps(out), flush=True)
    return 0


def reference_leg(args) -> dict:
    """The reference's own protocol (docs/SPEC.md:48-58): Flock::scatter at
    density 0.5 birds/m^3 (32 birds per 4 m cell), +x heading, seed 1, 20
    warm-up + 20 timed steps, mean GPU ms per pass from events (histogram as
    its own pass); the rows of
    docs/results/rtx4000ada-2026-10-21/bench-table-4000ada.txt. Each size is
    run at the default launch shape and at the autotuned one (same grid)."""
    import flock_hip as fh
    from flock_hip import autotune

    rows = {}
    for n in (100_000, 1_000_000):
        w = fh.gen_reference(n, 1)
        # the reference grid (4 m cells, buckets_for(n) buckets); launch shape autotuned
        tuned = autotune.tune([w], "plain", flags=fh.FLAG_UNFUSED, warmup=20, steps=5, cells=(4.0,), shifts=(0,),
                              spatial=(0,), thorough=False, progress=log,
                              **tune_cache(args, f"reference-protocol-{n}-seed1"))
        for name, prm, flags in (("default", {}, fh.FLAG_UNFUSED), ("tuned", tuned["params"], tuned["flags"])):
            with fh.Flock.from_words(w, flags=flags, **prm) as f:
                f.step(20)
                ms = f.step(20, 0, timing=True)
            row = {k: ms[k] for k in ("grid_histogram", "grid_scan", "grid_scatter", "decide", "integrate",
                                      "step_total")}
            row["bird_steps_per_s"] = n / (ms["step_total"] * 1e-3)
            row["params"] = prm
            row["flags"] = flags
            rows[f"{n}_{name}"] = row
    rows["published_rtx4000ada_decide_ms"] = {"100000": 1.24, "1000000": 9.46}
    rows["published_rtx4000ada_step_total_ms"] = {"100000": 1.276, "1000000": 9.979}
    return rows


def c1_leg(args) -> dict:
    """The plain step at 10,000 birds (uniform 100 m cube = 0.64 birds per 4 m
    cell, isotropic headings 10 +- 1 m/s, seed 1): 20 warm-up + 20 timed
    steps replayed from hipGraphs at autotuned tunables, three runs. Its CPU
    baseline runs the oracle over the same 40 steps in full and checks the
    GPU state after every one of them against it."""
    import flock_hip as fh
    from flock_hip import autotune, workloads

    t0 = time.perf_counter()
    n = 10_000
    w = workloads.plain(n, args.seed)
    tuned = autotune.tune([w], "plain", flags=fh.FLAG_GRAPH, warmup=20, steps=20, thorough=False, progress=log,
                          **tune_cache(args, f"plain-cube-{n}-seed{args.seed}"))
    prm, flags = tuned["params"], tuned["flags"]
    runs = []
    for _ in range(3):
        with fh.Flock.from_words(w, flags=flags, **prm) as f:
            f.step(20)
            f.sync()
            tt = time.perf_counter()
            f.step(20)
            f.sync()
            runs.append((time.perf_counter() - tt) * 1e3 / 20)
    ms = float(np.median(runs))
    with fh.Flock.from_words(w, flags=flags & ~fh.FLAG_GRAPH, **prm) as f:
        f.step(20)
        passes = f.step(20, 0, timing=True)
    nbytes = algo_step_bytes(n, passes["due_last"])
    out = {
        "workload": f"{n} birds, uniform 100 m cube (0.64 birds per 4 m cell), isotropic headings 10+-1 m/s, "
                    f"seed {args.seed}; 20 warm-up + 20 timed steps, three runs",
        "bird_steps_per_s": n / (ms * 1e-3),
        "ms_per_step": ms,
        "runs_ms_per_step": runs,
        "passes_ms": {k: passes[k] for k in ("grid_histogram", "grid_scan", "grid_scatter", "decide", "integrate",
                                            "step_total")},
        "roofline": hbm_roofline(nbytes, ms * 1e-3, "172n + 24B + 32d per step (d measured)"),
        "autotuned": prm,
        "flags": flags,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = c1_cpu_baseline(w, prm, flags)
    out["leg_seconds"] = time.perf_counter() - t0
    return out


def c1_cpu_baseline(w: np.ndarray, prm: dict, flags: int) -> dict:
    """The oracle (brute force, single thread) over the C1 protocol's 40
    steps in full, timed; the GPU steps the same start one step at a time at
    the leg's tunables, and every step is checked against the oracle: the
    reference's pointwise 1e-3 with equal neighbour counts
    (docs/SPEC.md:66-72)."""
    import flock_hip as fh
    from oracle import flockref as fr

    n = int(w.shape[0])
    gpu = []
    with fh.Flock.from_words(w, flags=flags & ~fh.FLAG_GRAPH, **prm) as f:
        for _ in range(40):
            f.step(1)
            gpu.append(f.download())
    cpu, tc = w, 0.0
    spent = 0.0
    dpos, dvel, nbm, ok = [], [], [], True
    for g, tg in gpu:
        t0 = time.perf_counter()
        cpu, tc = fr.step(cpu, tc, 1)
        spent += time.perf_counter() - t0
        d = fr.diff(cpu, g)
        dpos.append(d["dpos"])
        dvel.append(d["dvel"])
        nbm.append(d["nb_mismatch"])
        ok = ok and d["ok"] and tg == tc
    return {
        "value": n * 40 / spent,
        "unit": "bird-steps/s",
        "cores": 1,
        "kind": "port",
        "sample": f"the whole protocol: 40 oracle steps of all {n} birds (brute force, single thread), {spent:.2f} s",
        "parity_every_step": {"ok": ok, "max_dpos": dpos, "max_dvel": dvel, "nb_mismatch": nbm,
                              "bar": "max |dpos|, |dvel| <= 1e-3 and equal neighbour counts after every step"},
        "host": {"nproc": os.cpu_count(), "cpu_model": _cpu_model()},
    }


def clumps_leg(args) -> dict:
    """10,000,000 birds in 8 Gaussian clumps (sigma 5 m) with centres uniform
    in a 1 km cube, headings aligned per clump (10 degree cap), 10 m/s; cell,
    bucket and decide tunables autotuned (cells down to 1 m for the dense
    cores); 20 timed steps."""
    import flock_hip as fh
    from flock_hip import autotune, workloads

    t0 = time.perf_counter()
    n = args.clump_birds
    w = workloads.clumps(n, 8, 5.0, 1000.0, seed=args.seed)
    tuned = autotune.tune([w], "plain", flags=fh.FLAG_GRAPH, warmup=1, steps=2,
                          cells=(4.0, 1.0, 2.0, 3.0), progress=log, thorough=False,
                          **tune_cache(args, f"clumps-8-sigma5-{n}-seed{args.seed}"))
    prm, flags = tuned["params"], tuned["flags"]
    with fh.Flock.from_words(w, flags=flags, **prm) as f:
        f.step(args.clump_warmup)
        f.sync()
        tt = time.perf_counter()
        f.step(args.clump_steps)
        f.sync()
        wall = time.perf_counter() - tt
    with fh.Flock.from_words(w, flags=flags & ~fh.FLAG_GRAPH, **prm) as f:
        f.step(args.clump_warmup)
        passes = f.step(args.clump_steps, 0, timing=True)
    ms = wall * 1e3 / args.clump_steps
    out = {
        "workload": f"{n} birds, 8 Gaussian clumps sigma 5 m, centres uniform in a 1 km cube, aligned headings, "
                    f"10 m/s, seed {args.seed}; {args.clump_warmup} warm-up + {args.clump_steps} timed steps",
        "bird_steps_per_s": n * args.clump_steps / wall,
        "ms_per_step": ms,
        "passes_ms": {k: passes[k] for k in ("grid_scan", "grid_scatter", "decide", "integrate", "step_total")},
        "due_birds_last_step": passes["due_last"],
        "roofline": hbm_roofline(algo_step_bytes(n, n / 10.0), ms * 1e-3,
                                 "172n + 24B + 32d per step, d = n/10 (the mean over a decision period)"),
        "autotuned": prm,
        "flags": flags,
        "tunables_from_cache": bool(tuned.get("from_cache")),
    }
    if not args.no_cpu:
        costs = _cpu_baseline_costs([w], due_sample=12, bird_sample=20_000)
        step_s = _cpu_baseline_step_s(costs, n, n / 10.0)
        out["cpu_baseline"] = {
            "value": n / step_s, "unit": "bird-steps/s", "cores": 1, "kind": "port",
            "sample": (f"the oracle's brute-force decide (single thread) timed on 12 due birds of the 10M-bird "
                       f"flock and step_bird on 20,000 of its birds ({costs['timed_s']:.1f} s), extrapolated "
                       f"to one step ({step_s:.0f} s)"),
        }
    out["leg_seconds"] = time.perf_counter() - t0
    return out


def _plain_per_flock(starts, prm: dict, flags: int, steps: int) -> float:
    """Seconds of `steps` steps of every flock stepped alone, one flock after
    another (each on its own handle, created, warmed for 3 steps, timed and
    destroyed before the next: one flock's device state and graphs live at a
    time)."""
    import flock_hip as fh

    total = 0.0
    for w in starts:
        with fh.Flock.from_words(w, flags=flags | fh.FLAG_GRAPH, **prm) as f:
            f.step(3)
            f.sync()
            t0 = time.perf_counter()
            f.step(steps)
            f.sync()
            total += time.perf_counter() - t0
    return total


def batched_leg(args, plain_pt) -> dict:
    """256 flocks of log-uniform sizes 1k-200k at 0.64 birds per cell: all of
    them in one batch (one launch per stage, hipGraph, autotuned) against the
    same flocks stepped one flock per launch sequence, each with its own
    graph, at the batch's tunables and at the plain 1M-bird step's own."""
    import flock_hip as fh
    from flock_hip import autotune, workloads

    t0 = time.perf_counter()
    starts = workloads.mix(args.flocks, seed=args.seed)
    sizes = [int(w.shape[0]) for w in starts]
    total = sum(sizes)
    # trials time the leg's own window (3 warm-up + 100 steps, whatever
    # --batch-steps says): the radii grow over the steps, and a short window
    # picks cells that lose over the leg (cell 6 against 10 m: 0.78 against 0.68 ms)
    tuned = autotune.tune(starts, "batch", flags=fh.FLAG_GRAPH, warmup=3, steps=BATCH_TUNE_STEPS,
                          progress=log, thorough=False, **tune_cache(args, f"mix-{args.flocks}-seed{args.seed}"))
    prm, flags = tuned["params"], tuned["flags"]
    buckets = sum(autotune._buckets(s) for s in sizes)
    with fh.Batch(total + 256 * len(starts), buckets + 4096 * len(starts), len(starts), flags=flags, **prm) as b:
        for w in starts:
            b.add(w)
        b.step(3)
        tb = time.perf_counter()
        b.step(args.batch_steps)
        batched_s = time.perf_counter() - tb
        due = b.step(1, 0, timing=True)["due_last"]
    if args.no_plain:
        plain_s = plain2_s = float("nan")
    else:
        log("batched leg: flocks alone at the batch's tunables")
        plain_s = _plain_per_flock(starts, prm, flags, args.batch_steps)
        log("batched leg: flocks alone at the plain step's tunables")
        plain2_s = _plain_per_flock(starts, plain_pt[0], plain_pt[1], args.batch_steps)
    bs = total * args.batch_steps
    step_bytes_all = sum(172.0 * n + 24.0 * fh.buckets_for(n) for n in sizes) + 32.0 * due
    out = {
        "workload": f"{len(starts)} flocks, sizes log-uniform 1e3..2e5 (seed {args.seed}), {total} birds, "
                    f"0.64 birds per 4 m cell, {args.batch_steps} steps",
        "birds": total,
        "steps": args.batch_steps,
        "batched_bird_steps_per_s": bs / batched_s,
        "batched_ms_per_step": batched_s * 1e3 / args.batch_steps,
        "plain_bird_steps_per_s": bs / min(plain_s, plain2_s),
        "plain_at_batch_tunables": {"bird_steps_per_s": bs / plain_s, "ms_per_step": plain_s * 1e3 / args.batch_steps},
        "plain_at_plain_tunables": {"bird_steps_per_s": bs / plain2_s,
                                    "ms_per_step": plain2_s * 1e3 / args.batch_steps,
                                    "params": plain_pt[0], "flags": plain_pt[1]},
        # against the faster of the two per-flock comparators
        "speedup": min(plain_s, plain2_s) / batched_s,
        "roofline": hbm_roofline(step_bytes_all, batched_s / args.batch_steps,
                                 "sum over the flocks of 172n + 24B + 32d per step (d measured)"),
        "autotuned": prm,
        "flags": flags,
        "tunables_from_cache": bool(tuned.get("from_cache")),
    }
    if not args.no_cpu:
        order = np.argsort(sizes)
        pick = [starts[int(order[i])] for i in (0, 64, 128, 192, len(order) - 1)]
        costs = _cpu_baseline_costs(pick, due_sample=200, bird_sample=20_000)
        frac = due / total
        step_s = sum(_cpu_baseline_step_s(costs, n, frac * n) for n in sizes)
        out["cpu_baseline"] = {
            "value": total / step_s, "unit": "bird-steps/s", "cores": 1, "kind": "port",
            "sample": (f"the oracle's brute-force decide (single thread) timed on up to 200 due birds of 5 of the "
                       f"flocks (smallest, quartiles, largest) and step_bird on up to 20,000 birds of each "
                       f"({costs['timed_s']:.1f} s), extrapolated to one step of all {len(sizes)} flocks "
                       f"({step_s:.0f} s)"),
        }
    out["leg_seconds"] = time.perf_counter() - t0
    return out


def _replay(fh, req, args, prm, flags) -> dict:
    res, st = fh.server_replay(req["birds"], req["steps"], req["seeds"], req["arrival_s"],
                               max_active=args.slots, hbm_budget_bytes=200e9, steps_per_iter=8,
                               flags=flags, **prm)
    ok = [r for r in res if r["status"] == 0]
    lat = np.array([r["done_s"] - r["arrival_s"] for r in ok])
    span = max(r["done_s"] for r in ok) - min(r["arrival_s"] for r in ok)
    done = np.array([r["birds"] * r["steps"] for r in ok], np.float64)
    nbytes = sum(float(r["steps"]) * algo_step_bytes(int(r["birds"]), r["birds"] / 10.0) for r in ok)
    return {
        "requests": len(req["birds"]),
        "completed": len(ok),
        "bird_steps_per_s": float(done.sum()) / span,
        "bird_steps_per_busy_s": st["bird_steps"] / st["busy_s"],
        "p50_latency_s": float(np.percentile(lat, 50)),
        "p99_latency_s": float(np.percentile(lat, 99)),
        "peak_active": st["peak_active"],
        "peak_hbm_bytes": st["peak_bytes"],
        "iterations": st["iterations"],
        # where the server's wall time went: batched steps, and the GPU-idle
        # admissions (add) and retirements (download + remove)
        "busy_s": st["busy_s"],
        "admit_s": st["admit_s"],
        "retire_s": st["retire_s"],
        "gap_s": st["gap_s"],
        "gap_share": st["gap_s"] / max(1e-9, st["gap_s"] + st["busy_s"]),
        "roofline": hbm_roofline(nbytes, span, "sum over the completed requests of steps x (172n + 24B + 3.2n), "
                                               "over the span from the first arrival to the last completion"),
    }


def _plain_requests(fh, req, count: int, prm: dict, flags: int) -> dict:
    """The first `count` requests of the stream, each stepped alone, one after
    another (one flock per launch sequence, hipGraph): upload, steps and
    download timed; start states generated outside the timed region."""
    total_s, bird_steps = 0.0, 0.0
    for b, st, seed in list(zip(req["birds"], req["steps"], req["seeds"]))[:count]:
        w = fh.gen_request(int(b), int(seed))
        t0 = time.perf_counter()
        with fh.Flock.from_words(w, flags=flags | fh.FLAG_GRAPH, **prm) as f:
            f.step(int(st))
            f.download()
        total_s += time.perf_counter() - t0
        bird_steps += float(b) * float(st)
    return {"requests": count, "bird_steps_per_s": bird_steps / total_s, "seconds": total_s, "params": prm,
            "flags": flags | fh.FLAG_GRAPH}


def served_leg(args, plain_pt) -> dict:
    """Flock-simulation requests (birds log-uniform 2e3..5e5, steps uniform
    50..500, own seeds) through the continuous-batching server: admission by
    HBM bytes against a 200 GB budget and a slot cap, state resident on the
    device, active flocks packed into the batched step.

    1. capacity: 500 requests arriving within 0.1 s (the server saturated);
    2. latency: the 2000-request Poisson stream at `load` x that capacity.
    Tunables are autotuned on request-sized flocks warmed for 250 steps (the
    mean request length), where the adaptive radii have grown."""
    import flock_hip as fh
    from flock_hip import autotune, workloads

    t0 = time.perf_counter()
    warm_req = workloads.stream(32, 1.0, seed=args.seed + 7)
    with fh.Batch(40 << 20, 64 << 20, 64, flags=fh.FLAG_GRAPH) as b:
        ids = [b.add(fh.gen_request(int(n), int(s))) for n, s in zip(warm_req["birds"], warm_req["seeds"])]
        b.step(250)
        warm = [b.download(i) for i in ids]
    log("served leg: autotune on 32 warmed request flocks")
    tuned = autotune.tune([w for w, _ in warm], "batch", flags=fh.FLAG_GRAPH, warmup=1, steps=4,
                          times=[t for _, t in warm],
                          cells=(4.0, 6.0, 8.0, 10.0, 12.0), progress=log,
                          **tune_cache(args, f"served-warm250-seed{args.seed + 7}"))
    prm, flags = tuned["params"], tuned["flags"] & ~fh.FLAG_GRAPH
    del warm
    sat = workloads.stream(500, 5000.0, seed=args.seed + 1)
    capacity = _replay(fh, sat, args, prm, flags)
    log(f"served leg: capacity {capacity['bird_steps_per_s']:.3e} bird-steps/s")
    mean_bird_steps = 90_200 * 275.0  # E[birds] x E[steps] of the request mix
    rate = args.load * capacity["bird_steps_per_s"] / mean_bird_steps
    req = workloads.stream(args.requests, rate, seed=args.seed)
    stream = _replay(fh, req, args, prm, flags)
    log(f"served leg: stream p50 {stream['p50_latency_s']:.3f} s, p99 {stream['p99_latency_s']:.3f} s")
    plain = _plain_requests(fh, req, args.plain_sample, prm, flags)
    plain2 = _plain_requests(fh, req, args.plain_sample, plain_pt[0], plain_pt[1])
    best_plain = max(plain["bird_steps_per_s"], plain2["bird_steps_per_s"])
    out = {
        "workload": f"{args.requests} requests, Poisson {rate:.1f}/s ({args.load:.0%} of the measured capacity), "
                    f"birds log-uniform 2e3..5e5, steps uniform 50..500, slot cap {args.slots}, 200 GB HBM budget "
                    f"({(fh.flock_bytes(512, flags=flags) - fh.flock_bytes(256, flags=flags)) // 256} B per bird "
                    f"at these flags, birds padded to 256, + bytes per bucket)",
        "rate_per_s": rate,
        **stream,
        "capacity": capacity,
        "plain": plain,
        "plain_at_plain_tunables": plain2,
        # against the faster of the two per-request comparators
        "capacity_vs_plain": capacity["bird_steps_per_s"] / best_plain,
        "autotuned": prm,
        "flags": flags,
        "tunables_from_cache": bool(tuned.get("from_cache")),
    }
    if not args.no_cpu:
        order = np.argsort(req["birds"][: args.plain_sample])
        pick = [fh.gen_request(int(req["birds"][int(i)]), int(req["seeds"][int(i)])) for i in
                (order[0], order[len(order) // 2], order[-1])]
        costs = _cpu_baseline_costs(pick, due_sample=200, bird_sample=20_000)
        birds = req["birds"].astype(np.float64)
        steps = req["steps"].astype(np.float64)
        cpu_s = float(np.sum(steps * (birds * birds / 10.0 * costs["per_candidate_s"] + birds * costs["per_bird_s"])))
        out["cpu_baseline"] = {
            "value": float(np.sum(birds * steps)) / cpu_s, "unit": "bird-steps/s", "cores": 1, "kind": "port",
            "sample": (f"the oracle's brute-force decide (single thread) timed on up to 200 due birds of 3 request "
                       f"start states (smallest, median, largest of the first {args.plain_sample}) and step_bird "
                       f"on up to 20,000 birds of each ({costs['timed_s']:.1f} s), extrapolated to all "
                       f"{args.requests} requests' steps ({cpu_s:.3g} s; one tenth of the birds decide per step)"),
        }
    out["leg_seconds"] = time.perf_counter() - t0
    return out


if __name__ == "__main__":
    sys.exit(main())
