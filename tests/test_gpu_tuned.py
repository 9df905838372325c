"""The autotuned points the benchmark times, each run on its own workload
against the oracle.

Every entry of the committed tuning cache (flock_hip/tuned_mi355x.json) that
bench.py reads is run here at exactly its tunables and flags:
  - plain 1M-bird cube: one step against the oracle's grid walk, then the
    drift over 20 steps;
  - the 256-flock mix: every flock against stepping it alone, a sample of
    flocks against the oracle (brute force for the small ones), and the
    20-step drift of that sample;
  - the served stream (requests of 2k-500k birds, 50-500 steps) through the
    server: every result against stepping it alone, the smallest against the
    oracle;
  - the clumped 10M-bird flock: one step checked on 50,000 sampled due birds
    against the oracle's grid walk over the whole flock, and the 20-step drift
    of a 100k-bird flock of the same clump shape.

Bars: the reference's pointwise 1e-3 (positions and velocities, absolute) with
equal neighbour counts for short horizons, outliers below 1e-2 up to 100
steps, and its statdiff beyond (docs/SPEC.md:66-72, scripts/parity.sh:23-34).
Decision times and clocks are always exact. A flock in a resident batch takes
its arrivals in atomic order, so it matches stepping alone to rounding of
reordered sums.
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)
DT = np.float32(0.005)
CACHE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "murmuration-kernels_amd",
                     "flock_hip", "tuned_mi355x.json")


def tuned(label):
    """(params, flags) of the cache entry for a workload label."""
    with open(CACHE) as fp:
        cache = json.load(fp)
    hits = [v for k, v in cache.items() if k.split("|")[2] == label]
    assert len(hits) == 1, label
    return dict(hits[0]["params"]), int(hits[0]["flags"])


def abs_err(a, b):
    return float(np.max(np.abs(a.astype(np.float64) - b.astype(np.float64)), initial=0.0))


def rel_err(a, b):
    a = a.astype(np.float64)
    b = b.astype(np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(1.0, np.abs(b)), initial=0.0))


def horizon_ok(fr, want, got, steps):
    """The reference's ladder bar for a run of `steps` steps: pointwise up to
    30 steps, 2 outliers below 1e-2 up to 100 steps at the ladder's size
    (2000 birds, up to 5000 here), statdiff beyond. Over long horizons a
    neighbour count can flip at the radius in a large flock, which the
    pointwise bars do not allow."""
    if steps <= 30:
        d = fr.diff(want, got)
    elif steps <= 100 and want.shape[0] <= 5000:
        d = fr.diff(want, got, allow_outliers=2, outlier_max=1e-2)
    else:
        d = fr.statdiff(want, got)
    return d["ok"], d


def drift_ok(fr, want, got):
    """20 steps of a large flock: a neighbour at the radius can flip in one
    bird in ten thousand and move that bird beyond 1e-3; the rest stay
    within the pointwise bar and the aggregates within statdiff."""
    d = fr.diff(want, got)
    far = int(np.sum(np.max(np.abs(got[:, 0:6].astype(np.float64) - want[:, 0:6]), axis=1) > 1e-3))
    cap = max(2, want.shape[0] // 10_000)
    d["beyond_1e-3"] = far
    return fr.statdiff(want, got)["ok"] and far <= cap and d["nb_mismatch"] <= cap, d


def near_alone(got, want, steps):
    """A resident-batch result against the same flock stepped alone."""
    assert np.array_equal(got[:, 12], want[:, 12])  # decision times
    if steps <= 30:
        assert abs_err(got[:, 0:6], want[:, 0:6]) <= 1e-3
        assert np.mean(got[:, 13].view(np.uint32) == want[:, 13].view(np.uint32)) >= 0.99


def test_plain_1m_tuned_point(fh, fr):
    from flock_hip import workloads

    prm, flags = tuned("plain-cube-1000000-seed1")
    w0 = workloads.plain(1_000_000, 1)
    print("plain 1M tuned point", prm, flags, flush=True)
    with fh.Flock.from_words(w0, flags=flags, **prm) as f:
        f.step(1)
        gpu, t = f.download()
        cpu, tc = fr.step_grid(w0, 0.0, 1, threads=THREADS)
        assert t == tc
        d = fr.diff(cpu, gpu)
        print("1 step:", d, "rel", rel_err(gpu[:, 0:6], cpu[:, 0:6]), flush=True)
        assert d["ok"], d
        assert rel_err(gpu[:, 0:6], cpu[:, 0:6]) <= 1e-4
        f.step(19)
        gpu, t = f.download()
    cpu, tc = fr.step_grid(cpu, tc, 19, threads=THREADS)
    assert t == tc
    assert np.array_equal(gpu[:, 12], cpu[:, 12])
    ok, d = drift_ok(fr, cpu, gpu)
    print(f"C2 20-step drift: max |dpos| {d['dpos']:.3e} |dvel| {d['dvel']:.3e}, "
          f"neighbour-count mismatches {d['nb_mismatch']} of 1000000, birds beyond 1e-3: {d['beyond_1e-3']}",
          flush=True)
    assert ok, d


def test_mix_256_tuned_point(fh, fr):
    from flock_hip import workloads

    prm, flags = tuned("mix-256-seed1")
    starts = workloads.mix(256, seed=1)
    sizes = [int(w.shape[0]) for w in starts]
    total = sum(sizes)
    print("mix 256 tuned point", prm, flags, total, "birds", flush=True)
    steps = 6
    bk = sum(fh.buckets_for(n) for n in sizes)
    with fh.Batch(total + 256 * len(starts), bk + 4096 * len(starts), len(starts), flags=flags, **prm) as b:
        ids = [b.add(w) for w in starts]
        b.step(steps)
        got = [b.download(i) for i in ids]
    # every flock against itself stepped alone (canonical order, default tunables)
    for k, ((g, t), w) in enumerate(zip(got, starts)):
        with fh.Flock.from_words(w, flags=fh.FLAG_CANONICAL) as f:
            f.step(steps)
            want, tw = f.download()
        assert t == tw, k
        near_alone(g, want, steps)
    # a sample against the oracle: the four smallest by brute force, four
    # more across the size range by the grid walk
    order = np.argsort(sizes)
    brute = [int(i) for i in order[:4]]
    grid = [int(order[i]) for i in (64, 128, 160, 192)]
    for k in brute + grid:
        if k in brute:
            cpu, tc = fr.step(starts[k], 0.0, steps, threads=THREADS)
        else:
            cpu, tc = fr.step_grid(starts[k], 0.0, steps, threads=THREADS)
        d = fr.diff(cpu, got[k][0])
        print(f"flock {k} ({sizes[k]} birds) {steps} steps vs oracle:", d, flush=True)
        assert got[k][1] == tc and d["ok"], (k, d)
    # 20-step drift of the sampled flocks
    sample = [starts[k] for k in brute + grid]
    sb = sum(fh.buckets_for(w.shape[0]) for w in sample)
    with fh.Batch(sum(w.shape[0] for w in sample) + 256 * 8, sb + 4096 * 8, 8, flags=flags, **prm) as b:
        ids = [b.add(w) for w in sample]
        b.step(20)
        got20 = [b.download(i) for i in ids]
    worst, flips = 0.0, 0
    for k, w, (g, t) in zip(brute + grid, sample, got20):
        cpu, tc = fr.step_grid(w, 0.0, 20, threads=THREADS)
        assert t == tc and np.array_equal(g[:, 12], cpu[:, 12])
        ok, d = drift_ok(fr, cpu, g)
        worst = max(worst, d["dpos"], d["dvel"])
        flips += d["nb_mismatch"]
        assert ok, (k, d)
    print(f"C3 sample 20-step drift: max |dpos|, |dvel| {worst:.3e}, neighbour-count mismatches {flips}, "
          f"over {len(sample)} flocks", flush=True)


def test_served_stream_tuned_point(fh, fr):
    from flock_hip import workloads

    prm, flags = tuned("served-warm250-seed8")
    flags &= ~fh.FLAG_GRAPH  # the served leg's flags (bench.py)
    req = workloads.stream(64, 100.0, seed=3)
    assert req["birds"].min() < 5000 and req["birds"].max() > 200_000
    print("served tuned point", prm, flags, int(req["birds"].sum()), "birds", flush=True)
    with fh.Server(max_active=64, hbm_budget_bytes=200e9, steps_per_iter=8, keep_states=1, flags=flags,
                   **prm) as srv:
        ids = [srv.submit(int(b), int(s), int(sd)) for b, s, sd in zip(req["birds"], req["steps"], req["seeds"])]
        srv.drain(600)
        res = {r["id"]: r for r in srv.poll()}
        states = {rid: srv.take_state(rid) for rid in ids}
    assert len(res) == 64 and all(r["status"] == 0 for r in res.values())
    worst = 0.0
    for rid, b, s, sd in zip(ids, req["birds"], req["steps"], req["seeds"]):
        got, t = states[rid]
        w0 = fh.gen_request(int(b), int(sd))
        with fh.Flock.from_words(w0, flags=fh.FLAG_CANONICAL) as f:
            f.step(int(s))
            want, tw = f.download()
        assert t == tw and res[rid]["final_time"] == t, rid
        near_alone(got, want, int(s))
        ok, d = horizon_ok(fr, want, got, int(s))
        assert ok, (rid, int(b), int(s), d)
        worst = max(worst, abs_err(got[:, 0:6], want[:, 0:6]))
    print(f"served vs alone: max |dpos|, |dvel| {worst:.3e} over 64 requests", flush=True)
    # the oracle for the four smallest requests (by bird-steps)
    cost = req["birds"].astype(np.float64) * req["steps"]
    for k in np.argsort(cost)[:4]:
        w0 = fh.gen_request(int(req["birds"][k]), int(req["seeds"][k]))
        cpu, tc = fr.step_grid(w0, 0.0, int(req["steps"][k]), threads=THREADS)
        got, t = states[ids[k]]
        assert t == tc
        ok, d = horizon_ok(fr, cpu, got, int(req["steps"][k]))
        print(f"request {k} ({int(req['birds'][k])} birds, {int(req['steps'][k])} steps) vs oracle:", d, flush=True)
        assert ok, d


def test_clumps_10m_tuned_point(fh, fr):
    from flock_hip import workloads

    prm, flags = tuned("clumps-8-sigma5-10000000-seed1")
    n = 10_000_000
    w = workloads.clumps(n, 8, 5.0, 1000.0, seed=1)
    print("clumps tuned point", prm, flags, flush=True)
    with fh.Flock.from_words(w, flags=flags, **prm) as f:
        f.step(1)
        gpu, tg = f.download()
    assert tg == DT
    due = np.nonzero(np.float32(0.0) + DT >= w[:, 12])[0]
    sample = np.sort(np.random.default_rng(1).choice(due, 50_000, replace=False))
    keys, counts, offsets, srt = fr.grid(w)
    d = fr.decide_grid_many(w, sample, offsets, srt, fr.buckets_for(n), threads=THREADS)
    assert np.array_equal(gpu[sample, 13].view(np.uint32), d["neighbours"])
    assert abs_err(gpu[sample, 8], d["radius"]) <= 1e-6  # radius_step's powf: device vs host libm
    assert rel_err(gpu[sample, 9:12], d["force"]) <= 1e-4
    assert rel_err(gpu[sample, 14], d["centrality"]) <= 1e-4
    for k in range(0, 2000):
        i = sample[k]
        r = w[i].copy()
        r[9:12] = d["force"][k]
        r[8] = d["radius"][k]
        r[12] += np.float32(0.05)
        want = fr.step_bird(r)
        assert abs_err(gpu[i, 0:6], want[0:6]) <= 1e-3 and rel_err(gpu[i, 0:6], want[0:6]) <= 1e-4, i


def test_clumps_sample_20_step_drift(fh, fr):
    # The clumped shape at a size the oracle steps in seconds: 100,000 birds in
    # 8 clumps of sigma 5 m, at the 10M-bird flock's tunables.
    from flock_hip import workloads

    prm, flags = tuned("clumps-8-sigma5-10000000-seed1")
    w0 = workloads.clumps(100_000, 8, 5.0, 1000.0, seed=4)
    with fh.Flock.from_words(w0, flags=flags, **prm) as f:
        f.step(1)
        g1, t1 = f.download()
        f.step(19)
        g20, t20 = f.download()
    c1, tc1 = fr.step_grid(w0, 0.0, 1, threads=THREADS)
    d1 = fr.diff(c1, g1)
    assert t1 == tc1 and d1["ok"], d1
    c20, tc20 = fr.step_grid(c1, tc1, 19, threads=THREADS)
    assert t20 == tc20 and np.array_equal(g20[:, 12], c20[:, 12])
    ok, d = drift_ok(fr, c20, g20)
    print(f"C5 sample (100k clumped birds) drift: 1 step |dpos| {d1['dpos']:.3e} |dvel| {d1['dvel']:.3e}; "
          f"20 steps |dpos| {d['dpos']:.3e} |dvel| {d['dvel']:.3e}, neighbour mismatches {d['nb_mismatch']}, "
          f"birds beyond 1e-3: {d['beyond_1e-3']}", flush=True)
    assert fr.statdiff(c20, g20)["ok"]


def test_c1_10k_tuned_point(fh, fr):
    # 10,000 birds, 100 m cube, seed 1, at its cached tunables: the oracle
    # after every one of 40 steps (the C1 protocol's warm-up and timed steps).
    from flock_hip import workloads

    prm, flags = tuned("plain-cube-10000-seed1")
    w0 = workloads.plain(10_000, 1)
    cpu, tc = w0, 0.0
    worst = 0.0
    with fh.Flock.from_words(w0, flags=flags, **prm) as f:
        for s in range(40):
            f.step(1)
            gpu, t = f.download()
            cpu, tc = fr.step(cpu, tc, 1, threads=THREADS)
            d = fr.diff(cpu, gpu)
            assert t == tc and d["ok"], (s, d)
            worst = max(worst, d["dpos"], d["dvel"])
            if s == 19:
                print(f"C1 20-step drift at the tuned point: max |dpos|, |dvel| {worst:.3e}", flush=True)


@pytest.mark.parametrize("n", [100_000, 1_000_000])
def test_reference_protocol_tuned_points(fh, fr, n):
    # Flock::scatter at density 0.5 (state.rs:79-101) at the cached tunables
    # of the reference-protocol leg: one step against the oracle's grid walk.
    prm, flags = tuned(f"reference-protocol-{n}-seed1")
    w0 = fh.gen_reference(n, 1)
    with fh.Flock.from_words(w0, flags=flags, **prm) as f:
        f.step(1)
        gpu, t = f.download()
    cpu, tc = fr.step_grid(w0, 0.0, 1, threads=THREADS)
    d = fr.diff(cpu, gpu)
    print(f"reference protocol {n}: {d}", flush=True)
    assert t == tc and d["ok"], d


def test_every_cached_point_is_checked():
    # each label of the tuning cache has a test above
    with open(CACHE) as fp:
        labels = {k.split("|")[2] for k in json.load(fp)}
    assert labels <= {"plain-cube-1000000-seed1", "mix-256-seed1", "served-warm250-seed8",
                      "clumps-8-sigma5-10000000-seed1", "plain-cube-10000-seed1", "reference-protocol-100000-seed1",
                      "reference-protocol-1000000-seed1"}, labels
