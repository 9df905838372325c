"""Resident layout (FH_FLAG_RESIDENT): fh_step keeps the state in bucket order
between steps and re-counts only the birds that change bucket. The grid it
maintains must be the reference grid of the current positions (keys, counts
and offsets bit-exact, per-bucket sets equal), and its steps must agree with
the oracle like the standard path's."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)
DT = np.float32(0.005)


def cube(fh, n, seed=1):
    return fh.gen_cube(n, fh.cube_side_for(n), seed)


def rel_err(a, b):
    a = a.astype(np.float64)
    b = b.astype(np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(1.0, np.abs(b)))) if a.size else 0.0


def per_bucket_sets_equal(g_sorted, offsets, o_sorted):
    for b in np.nonzero(np.diff(offsets.astype(np.int64)) > 1)[0]:
        s, e = offsets[b], offsets[b + 1]
        if set(g_sorted[s:e].tolist()) != set(o_sorted[s:e].tolist()):
            return False
    single = np.diff(offsets.astype(np.int64)) == 1
    return np.array_equal(g_sorted[offsets[:-1][single]], o_sorted[offsets[:-1][single]])


@pytest.mark.parametrize("n,kind,cell,extra", [(200_000, "cube", 4.0, 0), (50_000, "ref", 4.0, 0),
                                               (300_000, "clumps", 2.0, 0), (200_000, "cube", 4.0, 2048),
                                               (300_000, "clumps", 2.0, 2048)])
def test_resident_grid_is_the_reference_grid(fh, fr, n, kind, cell, extra):
    # After every run of resident steps the layout is the counting sort of the
    # current positions; downloading leaves the layout and the next step
    # enters it again.
    if kind == "cube":
        w0 = cube(fh, n, 3)
    elif kind == "ref":
        w0 = fh.gen_reference(n, 3)
    else:
        w0 = fh.gen_clumps(n, 4, 4.0, 200.0, 3)
    with fh.Flock.from_words(w0, cell=cell, flags=fh.FLAG_RESIDENT | fh.FLAG_GRAPH | extra) as f:
        for k in range(3):
            f.step(5 + 3 * k)
            g = f.read_grid()
            state, _ = f.download()
            keys, counts, offsets, srt = fr.grid(state, f.buckets, cell)
            print(kind, k, "movers-free check", int(counts.max()))
            assert np.array_equal(g.keys, keys)
            assert np.array_equal(g.counts, counts)
            assert np.array_equal(g.offsets, offsets)
            assert per_bucket_sets_equal(g.sorted, offsets, srt)


@pytest.mark.parametrize("flags,group", [(64, 1), (64 | 4, 4), (64 | 4 | 32, 2), (64 | 4, 16), (64 | 4, 64),
                                         (64 | 4 | 256, 4), (64 | 256, 1), (64 | 4 | 128 | 512, 4),
                                         (64 | 4 | 128 | 512 | 1024, 4), (64 | 1024, 1), (64 | 4 | 1024 | 256, 16),
                                         (64 | 4 | 2048, 2), (64 | 2048 | 128 | 1024, 4), (64 | 2048 | 256, 64),
                                         (64 | 4 | 2048 | 512, 1)])
def test_resident_step_matches_the_oracle(fh, fr, flags, group):
    w0 = cube(fh, 200_000, 5)
    with fh.Flock.from_words(w0, flags=flags, decide_group=group) as f:
        f.step(1)
        gpu, t = f.download()
    cpu, tc = fr.step_grid(w0, 0.0, 1, threads=THREADS)
    d = fr.diff(cpu, gpu)
    print("resident 1-step", flags, group, d)
    assert t == tc and d["ok"], d  # absolute 1e-3 and equal neighbour counts (docs/SPEC.md:66-72)
    assert rel_err(gpu[:, 0:6], cpu[:, 0:6]) <= 1e-4


def test_resident_c1_20_steps(fh, fr):
    # 10,000 birds, 100 m cube: 20 consecutive resident steps against the
    # oracle's 20 steps (the trajectories drift apart only by rounding).
    w0 = cube(fh, 10_000)
    with fh.Flock.from_words(w0, flags=fh.FLAG_RESIDENT | fh.FLAG_GRAPH) as f:
        f.step(20)
        gpu, t = f.download()
    cpu, tc = fr.step(w0, 0.0, 20, threads=THREADS)
    assert t == tc
    d = fr.diff(cpu, gpu)
    drift = rel_err(gpu[:, 0:6], cpu[:, 0:6])
    print("resident C1 20-step drift", drift, d)
    assert drift <= 1e-3


@pytest.mark.parametrize("group,extra", [(2, 32), (4, 0), (2, 4096)])
def test_resident_parity_ladder_2000(fh, fr, group, extra):
    # The ladder of scripts/parity.sh:23-34 (docs/SPEC.md:66-72) on the path of
    # the tuned 1M-bird step: resident layout, fused flight, spatial hash,
    # lane groups of 2 with the occupancy bits (and of 4, and of 2 with lazy
    # rebuilds). Same bars as the canonical ladder: pointwise 1e-3 up to 30
    # steps, at most 2 outliers under 1e-2 at 100, aggregates at 1000.
    w0 = fh.gen_reference(2000, 1)
    flags = fh.FLAG_RESIDENT | fh.FLAG_FLIGHT | fh.FLAG_SPATIAL | fh.FLAG_GRAPH | extra
    for k in (1, 10, 30, 100, 1000):
        with fh.Flock.from_words(w0, flags=flags, decide_group=group) as f:
            f.step(k)
            gpu, t = f.download()
        cpu, tc = fr.step(w0, 0.0, k, threads=THREADS)
        assert t == tc
        if k <= 30:
            d = fr.diff(cpu, gpu)
        elif k <= 100:
            d = fr.diff(cpu, gpu, allow_outliers=2, outlier_max=1e-2)
        else:
            d = fr.statdiff(cpu, gpu)
        print(f"resident ladder {k}: {d}", flush=True)
        assert d["ok"], (k, d)


def test_resident_and_standard_paths_agree(fh, fr):
    # The same 40 steps on the standard path and on the resident one, with
    # the pass API used in between (which leaves the resident layout).
    w0 = cube(fh, 100_000, 7)
    outs = {}
    for name, flags in (("standard", fh.FLAG_GRAPH), ("resident", fh.FLAG_GRAPH | fh.FLAG_RESIDENT)):
        with fh.Flock.from_words(w0, flags=flags) as f:
            f.step(20)
            f.hash_count()
            f.scan()
            f.scatter()
            f.step(20)
            outs[name] = f.download()
    (a, ta), (b, tb) = outs["standard"], outs["resident"]
    assert ta == tb
    assert rel_err(b[:, 0:6], a[:, 0:6]) <= 1e-3
    assert np.mean(b[:, 13].view(np.uint32) == a[:, 13].view(np.uint32)) > 0.999


def test_resident_due_count_and_timing(fh):
    w0 = cube(fh, 100_000, 9)
    with fh.Flock.from_words(w0, flags=fh.FLAG_RESIDENT) as f:
        ms = f.step(4, 1, timing=True)
        assert ms["timed_steps"] == 3
        assert ms["decide"] > 0 and ms["integrate"] > 0 and ms["grid_scatter"] > 0
        assert 5_000 < ms["due_last"] < 15_000  # about a tenth of the birds decide each step
        for _ in range(3):  # leaving and entering again restarts the counters
            f.download()
            ms = f.step(1, 0, timing=True)
            assert 5_000 < ms["due_last"] < 15_000


def test_resident_rejects_canonical(fh):
    with pytest.raises(fh.FlockHipError) as e:
        fh.Flock(1000, flags=fh.FLAG_RESIDENT | fh.FLAG_CANONICAL)
    assert e.value.status == fh.ERR_INVALID


@pytest.mark.parametrize("n", [1, 2, 257])
def test_resident_tiny_flocks(fh, fr, n):
    w0 = fh.gen_reference(n, 4)
    with fh.Flock.from_words(w0, flags=fh.FLAG_RESIDENT | fh.FLAG_GRAPH) as f:
        f.step(30)
        gpu, t = f.download()
    cpu, tc = fr.step(w0, 0.0, 30)
    assert t == tc
    assert rel_err(gpu[:, 0:6], cpu[:, 0:6]) <= 1e-3


def test_resident_dense_single_cell(fh, fr):
    # 5000 birds inside one 4 m cell (one bucket of 5000) plus a spread: every
    # stayer of the dense bucket counts the movers before it.
    rng = np.random.default_rng(8)
    w0 = fh.gen_reference(8000, 5)
    w0[:5000, 0:3] = (rng.random((5000, 3)) * 3.9 + np.array([0.05, 0.05, 100.05])).astype(np.float32)
    w0[:, 8] = np.float32(0.5)
    with fh.Flock.from_words(w0, flags=fh.FLAG_RESIDENT) as f:
        for k in range(4):
            f.step(5)
            g = f.read_grid()
            state, _ = f.download()
            keys, counts, offsets, srt = fr.grid(state, f.buckets)
            assert np.array_equal(g.keys, keys) and np.array_equal(g.offsets, offsets)
            assert per_bucket_sets_equal(g.sorted, offsets, srt)


def test_resident_persistent_after_leaving_at_odd_parity(fh):
    # Leaving the resident layout after an odd number of steps and entering
    # it again restarts at parity 0: its chunk counter must be zero again, or
    # the persistent decide would find every chunk taken.
    w0 = cube(fh, 100_000, 13)
    outs = []
    for flags in (fh.FLAG_RESIDENT | fh.FLAG_PERSISTENT | fh.FLAG_GRAPH, fh.FLAG_RESIDENT | fh.FLAG_GRAPH):
        with fh.Flock.from_words(w0, flags=flags) as f:
            f.step(3)
            f.download()
            f.step(4)
            outs.append(f.download()[0])
    # arrivals take their slots in atomic order, so sums may round apart; the
    # decision clocks (advanced only by decide) must agree exactly
    assert np.array_equal(outs[0][:, 12], outs[1][:, 12])
    assert rel_err(outs[0][:, 0:6], outs[1][:, 0:6]) <= 1e-5


@pytest.mark.parametrize("extra", [0, 1024 | 128, 256 | 512])
def test_fused_flight_equals_the_separate_pass(fh, extra):
    # The fused flight computes the same new state per bird as the integrate
    # pass; slots of arrivals follow atomic order in both, so positions agree
    # to rounding of reordered sums and decision clocks exactly.
    w0 = cube(fh, 300_000, 21)
    outs = []
    for fl in (fh.FLAG_FLIGHT, 0):
        with fh.Flock.from_words(w0, flags=fh.FLAG_RESIDENT | extra | fl) as f:  # per-pass events: no graph
            f.step(6)
            f.download()
            f.step(9)
            ms = f.step(3, 0, timing=True)
            outs.append((f.download(), ms))
    ((a, ta), ma), ((b, tb), mb) = outs
    assert ta == tb
    assert np.array_equal(a[:, 12], b[:, 12])
    assert rel_err(a[:, 0:6], b[:, 0:6]) <= 1e-5
    assert np.mean(a[:, 13].view(np.uint32) == b[:, 13].view(np.uint32)) > 0.999
    assert ma["decide"] > 0 and mb["integrate"] > 0
    assert ma["due_last"] == mb["due_last"]


def test_resident_upload_replaces_the_records(fh):
    # An upload while resident drops the resident records; the next steps
    # start from the uploaded state.
    a = cube(fh, 20_000, 11)
    b = cube(fh, 20_000, 12)
    with fh.Flock.from_words(a, flags=fh.FLAG_RESIDENT) as f:
        f.step(5)
        f.upload(b, 0.0)
        f.step(5)
        got, _ = f.download()
    with fh.Flock.from_words(b, flags=fh.FLAG_RESIDENT) as f:
        f.step(5)
        want, _ = f.download()
    assert rel_err(got[:, 0:6], want[:, 0:6]) <= 1e-5
