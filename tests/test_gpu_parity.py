"""GPU parity: the gfx950 step against the oracle, through the C ABI.

Bars (stated per test):
  - keys, counts, offsets: bit-exact on identical positions;
  - sorted order: per-bucket set equality (arrival order) or bit-exact
    (canonical order = the oracle's stable counting sort);
  - neighbour sets: equal to brute force except ties within 1e-5 of the radius;
  - one step: positions/velocities within 1e-4 relative, neighbour counts equal;
  - the reference ladder (scripts/parity.sh): 1/10/30 steps pointwise 1e-3,
    100 steps with at most 2 outliers below 1e-2, 1000 steps statdiff.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)
DT = np.float32(0.005)


def cube(fh, n, seed=1):
    return fh.gen_cube(n, fh.cube_side_for(n), seed)


def build_grid(f):
    f.hash_count()
    f.scan()
    f.scatter()
    return f.read_grid()


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


@pytest.mark.parametrize("n,kind", [(2000, "ref"), (10_000, "cube"), (100_000, "cube"),
                                    (1_000_000, "cube"), (1, "cube"), (5000, "ref")])
def test_grid_bit_exact(fh, fr, n, kind):
    w = fh.gen_reference(n, 1) if kind == "ref" else cube(fh, n)
    print(f"grid n={n} kind={kind}", flush=True)
    with fh.Flock.from_words(w) as f:
        f.hash_count()
        raw = f.read_grid(offsets=False, sorted=False)
        keys, counts, offsets, srt = fr.grid(w, f.buckets)
        assert np.array_equal(raw.keys, keys)
        assert np.array_equal(raw.counts, counts)
        f.scan()
        f.scatter()
        g = f.read_grid()
        assert np.array_equal(g.offsets, offsets)
        assert np.array_equal(g.counts, counts)
        assert fr.check_grid(g.keys, g.offsets, g.sorted, f.buckets) == (0, "")
        assert per_bucket_sets_equal(g.sorted, offsets, srt)


def test_keys_at_extreme_positions(fh, fr):
    # cell_of saturates like Rust `as i32` (NaN -> 0) on the device too: NaN,
    # infinities, values far beyond 2^31 cells and values at the i32 edges.
    w = fh.gen_reference(2000, 7)
    vals = np.array([np.nan, np.inf, -np.inf, 1e30, -1e30, 8.6e9, -8.6e9, 8.589934e9, -8.589935e9, -0.0,
                     3.9999998, 4.0, -1e-30, -4.0], np.float32)
    rng = np.random.default_rng(3)
    for i in range(700):
        w[i, rng.integers(0, 3)] = vals[i % len(vals)]
    with fh.Flock.from_words(w) as f:
        f.hash_count()
        raw = f.read_grid(offsets=False, sorted=False)
        keys, counts, _, _ = fr.grid(w, f.buckets)
    assert np.array_equal(raw.keys, keys)
    assert np.array_equal(raw.counts, counts)


@pytest.mark.parametrize("n", [3000, 200_000])
def test_canonical_order_is_the_stable_sort(fh, fr, n):
    w = fh.gen_reference(n, 2)
    with fh.Flock.from_words(w, flags=fh.FLAG_CANONICAL) as f:
        g = build_grid(f)
        _, _, offsets, srt = fr.grid(w, f.buckets)
        assert np.array_equal(g.offsets, offsets)
        assert np.array_equal(g.sorted, srt)


def test_canonical_order_in_big_buckets(fh, fr):
    # Cells of 1500, 5000 and 12000 birds take the LDS sort of big buckets
    # (bit-exact stable order); 17000 birds exceed its 16384 cap and keep the
    # arrival order (set equality). Bird indices of a cell are interleaved.
    rng = np.random.default_rng(4)
    n = 40_000
    w = fh.gen_reference(n, 5)
    sizes = [1500, 5000, 12000, 17000]
    idx = rng.permutation(n)
    at = 0
    for c, m in enumerate(sizes):
        pick = idx[at:at + m]
        w[pick, 0:3] = (rng.random((m, 3)) * 3.9 + np.array([0.05 + 40.0 * c, 0.05, 100.05])).astype(np.float32)
        at += m
    with fh.Flock.from_words(w, flags=fh.FLAG_CANONICAL) as f:
        g = build_grid(f)
        _, counts, offsets, srt = fr.grid(w, f.buckets)
    assert np.array_equal(g.offsets, offsets)
    big = np.nonzero(counts > 16384)[0]
    assert len(big) == 1 and np.any(counts[counts <= 16384] >= 12000)
    s, e = offsets[big[0]], offsets[big[0] + 1]
    assert set(g.sorted[s:e].tolist()) == set(srt[s:e].tolist())
    assert np.array_equal(np.delete(g.sorted, np.arange(s, e)), np.delete(srt, np.arange(s, e)))


def test_dense_cell_and_hash_collisions(fh, fr):
    # 3000 birds inside one 4 m cell plus a spread, hashed into only 64 buckets.
    rng = np.random.default_rng(1)
    w = fh.gen_reference(6000, 3)
    w[:3000, 0:3] = (rng.random((3000, 3)) * 3.9 + np.array([0.05, 0.05, 100.05])).astype(np.float32)
    w[:, 8] = np.float32(1.0)
    with fh.Flock.from_words(w, buckets=64) as f:
        g = build_grid(f)
        keys, counts, offsets, srt = fr.grid(w, 64)
        assert np.array_equal(g.keys, keys) and np.array_equal(g.offsets, offsets)
        assert per_bucket_sets_equal(g.sorted, offsets, srt)
        for i in [0, 1, 2999, 3000, 4500, 5999]:
            for walk in (1, 2):
                got = f.neighbours(i, walk, cap=8192)
                want = fr.neighbours_brute(w, i, walk)
                assert np.array_equal(got, want), (i, walk)


@pytest.mark.parametrize("group", [16, 64])
def test_split_walks_with_hash_collisions_match_brute_force(fh, fr, group):
    # The split walks (lanes share a bucket's candidates) where most buckets
    # hold birds of several cells: 64 buckets for a dense cell and a spread
    # over hundreds of cells.
    rng = np.random.default_rng(4)
    w = fh.gen_reference(6000, 5)
    w[:2500, 0:3] = (rng.random((2500, 3)) * 3.9 + np.array([0.05, 0.05, 100.05])).astype(np.float32)
    w[:, 8] = np.float32(1.5)
    cells = np.unique(np.floor(w[:, 0:3].astype(np.float64) / 4.0), axis=0)
    assert len(cells) > 3 * 64  # three cells per bucket on average
    t = np.float32(0.0) + DT
    with fh.Flock.from_words(w, buckets=64, decide_group=group, decide_block=256) as f:
        build_grid(f)
        f.decide(t)
        got, _ = f.download()
    due = np.nonzero(t >= w[:, 12])[0]
    assert len(due) > 500
    for i in due:
        d = fr.decide(w, int(i))
        assert got[i, 13].view(np.uint32) == d["neighbours"], i
        assert rel_err(got[i, 9:12], d["force"]) <= 1e-4, i
        assert abs(got[i, 14] - d["centrality"]) <= 1e-5, i


@pytest.mark.parametrize("group,buckets", [(2, 512), (4, 4096), (16, 512), (64, 4096)])
def test_spatial_walks_without_cell_test_match_brute_force(fh, fr, group, buckets):
    # Under the spatial hash a walk leaves out the cell test where the buckets
    # of its box, widened by one cell per side, are all distinct: at most
    # 2^(bits/3) cells per axis (8 with 512 buckets, 16 with 4096) inside one
    # block of 1024 cells. Every bucket here holds birds of many cells, so a
    # skip where the buckets are not distinct would change the sums. The flock
    # sits away from the origin (cells -1 and 0 lie in different blocks).
    rng = np.random.default_rng(5)
    w = fh.gen_reference(20_000, 7)
    w[:2500, 0:3] = (rng.random((2500, 3)) * 3.9 + np.array([200.05, 200.05, 100.05])).astype(np.float32)
    w[2500:, 0:3] = (rng.random((17_500, 3)) * 120.0 + np.array([150.0, 150.0, 40.0])).astype(np.float32)
    w[:, 8] = np.float32(1.5)
    cell, flags = np.float32(4.0), 128  # FH_FLAG_SPATIAL
    # the walks that skip the test, by the rule above (float32 cell coordinates)
    inv = np.float32(1.0) / cell
    rr = np.float32(2.0) * w[:, 8]
    lo = np.floor((w[:, 0:3] - rr[:, None]) * inv).astype(np.int64) - 1
    hi = np.floor((w[:, 0:3] + rr[:, None]) * inv).astype(np.int64) + 1
    span = 1 << (int(np.log2(buckets)) // 3)
    skip = np.all((hi - lo < span) & (lo >> 10 == hi >> 10), axis=1)
    cells = np.unique(np.floor(w[:, 0:3].astype(np.float64) / 4.0), axis=0)
    assert len(cells) > 3 * buckets  # three cells per bucket on average
    t = np.float32(0.0) + DT
    with fh.Flock.from_words(w, cell=float(cell), buckets=buckets, decide_group=group, decide_block=256,
                             flags=flags) as f:
        build_grid(f)
        f.decide(t)
        got, _ = f.download()
    due = np.nonzero(t >= w[:, 12])[0]
    assert len(due) > 1000 and skip[due].mean() > 0.5
    for i in due:
        d = fr.decide(w, int(i))
        assert got[i, 13].view(np.uint32) == d["neighbours"], i
        assert rel_err(got[i, 9:12], d["force"]) <= 1e-4, i
        assert abs(got[i, 14] - d["centrality"]) <= 1e-5, i


def ties_ok(w, i, walk, got, want):
    """Set difference allowed only for birds within 1e-5 of the radius."""
    r = float(w[i, 8]) * walk
    for j in set(got.tolist()) ^ set(want.tolist()):
        d = float(np.linalg.norm(w[j, :3].astype(np.float64) - w[i, :3].astype(np.float64)))
        if abs(d - r) >= 1e-5:
            return False
    return True


@pytest.mark.parametrize("n,radius", [(2000, None), (20_000, 2.5), (1_000_000, None)])
def test_neighbour_sets_match_brute_force(fh, fr, n, radius):
    w = fh.gen_reference(n, 1) if n == 2000 else cube(fh, n, 4)
    if radius:
        w[:, 8] = np.float32(radius)
    with fh.Flock.from_words(w) as f:
        build_grid(f)
        for i in np.linspace(0, n - 1, 40).astype(int):
            for walk in (1, 2):
                got = f.neighbours(int(i), walk)
                want = fr.neighbours_brute(w, int(i), walk)
                assert ties_ok(w, int(i), walk, got, want), (i, walk)


@pytest.mark.parametrize("flags", [0, 1, 32, 256])  # arrival order, canonical, occupancy skip, persistent
def test_decide_matches_brute_force(fh, fr, flags):
    w = fh.gen_reference(4000, 6)
    t = np.float32(0.0) + DT
    with fh.Flock.from_words(w, flags=flags) as f:
        build_grid(f)
        f.decide(t)
        got, _ = f.download()
    due = np.nonzero(t >= w[:, 12])[0]
    assert len(due) > 300
    for i in due:
        d = fr.decide(w, int(i))
        assert got[i, 13].view(np.uint32) == d["neighbours"], i
        assert rel_err(got[i, 9:12], d["force"]) <= 1e-4, i
        assert abs(got[i, 8] - d["radius"]) <= 1e-6 * d["radius"]
        assert abs(got[i, 14] - d["centrality"]) <= 1e-5
        assert got[i, 12] == w[i, 12] + np.float32(0.05)
    other = np.setdiff1d(np.arange(4000), due)
    assert np.array_equal(got[other].view(np.uint32), w[other].view(np.uint32))


def test_integrate_matches_step_bird(fh, fr):
    w = cube(fh, 5000, 8)
    w[:, 9:12] = (np.random.default_rng(2).standard_normal((5000, 3)) * 0.2).astype(np.float32)
    w[:, 6] = np.float32(0.3)
    with fh.Flock.from_words(w) as f:
        f.integrate()
        got, _ = f.download()
    want = np.stack([fr.step_bird(w[i]) for i in range(5000)])
    assert rel_err(got[:, 0:6], want[:, 0:6]) <= 1e-5
    assert np.max(np.abs(got[:, 6] - want[:, 6])) <= 1e-5
    assert np.max(np.abs(got[:, 7] - want[:, 7])) <= 1e-3


@pytest.mark.parametrize("flags", [0, 1, 2, 5])
def test_one_step_parity_reference_start(fh, fr, flags):
    w = fh.gen_reference(2000, 1)
    with fh.Flock.from_words(w, flags=flags) as f:
        f.step(1)
        gpu, t = f.download()
    cpu, tc = fr.step(w, 0.0, 1)
    d = fr.diff(cpu, gpu)
    print("1-step", flags, d)
    assert t == tc and d["ok"] and d["nb_mismatch"] == 0
    assert rel_err(gpu[:, 0:6], cpu[:, 0:6]) <= 1e-4


def test_parity_ladder_2000(fh, fr):
    # scripts/parity.sh:23-34 and docs/SPEC.md:66-72
    w0 = fh.gen_reference(2000, 1)
    for k in (1, 10, 30, 100, 1000):
        with fh.Flock.from_words(w0, flags=fh.FLAG_CANONICAL) as f:
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
        print(f"ladder {k}: {d}", flush=True)
        assert d["ok"], (k, d)


def test_c1_cube_10k_20_steps_parity_every_step(fh, fr):
    # 10,000 birds, 100 m cube, seed 1: oracle parity after every step.
    w0 = cube(fh, 10_000)
    cpu, tc = w0, 0.0
    drift = []
    with fh.Flock.from_words(w0, flags=fh.FLAG_CANONICAL) as f:
        for s in range(20):
            f.step(1)
            gpu, t = f.download()
            cpu, tc = fr.step(cpu, tc, 1, threads=THREADS)
            assert t == tc
            d = fr.diff(cpu, gpu)
            drift.append(rel_err(gpu[:, 0:6], cpu[:, 0:6]))
            assert d["ok"], (s, d)  # absolute 1e-3 and equal neighbour counts (docs/SPEC.md:66-72)
            if s == 0:
                assert drift[0] <= 1e-4
    print("C1 drift per step (max rel pos/vel):", ["%.2e" % x for x in drift])
    assert max(drift) <= 1e-3


def test_c2_cube_1m_one_step_parity(fh, fr):
    # 1,000,000 birds, 464.16 m cube: one step against the oracle's grid walk
    # (same neighbour semantics as brute force, pinned in test_oracle.py), and
    # brute-force neighbour sets for a sample of due birds.
    w0 = cube(fh, 1_000_000)
    with fh.Flock.from_words(w0, flags=fh.FLAG_CANONICAL) as f:
        f.step(1)
        gpu, t = f.download()
    cpu, tc = fr.step_grid(w0, 0.0, 1, threads=THREADS)
    assert t == tc
    d = fr.diff(cpu, gpu)
    print("C2 1-step:", d, "rel", rel_err(gpu[:, 0:6], cpu[:, 0:6]))
    assert d["ok"], d
    assert rel_err(gpu[:, 0:6], cpu[:, 0:6]) <= 1e-4
    due = np.nonzero(DT >= w0[:, 12])[0][:200]
    for i in due:
        want = fr.neighbours_brute(w0, int(i), 1)
        assert gpu[i, 13].view(np.uint32) == len(want)


@pytest.mark.parametrize("cell,group,block,bpt,shift", [(4.0, 4, 256, 8, 0), (6.0, 2, 128, 4, 1),
                                                        (3.0, 8, 64, 16, 0), (5.0, 1, 512, 2, 2),
                                                        (8.0, 4, 128, 8, 3), (12.0, 2, 256, 4, 4),
                                                        (4.0, 16, 256, 8, 0), (2.0, 64, 128, 4, 1)])
def test_c2_one_step_parity_at_tuned_params(fh, fr, cell, group, block, bpt, shift):
    # Tunables change the hash keys (cell) and the summation grouping (group),
    # never the neighbour sets: one step at 1M birds stays within 1e-4 of the
    # oracle's grid walk at the reference cell size.
    w0 = cube(fh, 1_000_000, 2)
    with fh.Flock.from_words(w0, cell=cell, decide_group=group, decide_block=block, birds_per_thread=bpt,
                             bucket_shift=shift, flags=fh.FLAG_CANONICAL) as f:
        f.step(1)
        gpu, t = f.download()
    cpu, tc = fr.step_grid(w0, 0.0, 1, threads=THREADS)
    d = fr.diff(cpu, gpu)
    print("tuned 1-step", cell, group, d)
    assert t == tc and d["ok"], d
    assert rel_err(gpu[:, 0:6], cpu[:, 0:6]) <= 1e-4


@pytest.mark.parametrize("group,block", [(2, 256), (4, 256), (8, 256), (16, 256), (64, 64), (64, 512)])
def test_grouped_decide_matches_brute_force(fh, fr, group, block):
    w = fh.gen_reference(4000, 6)
    t = np.float32(0.0) + DT
    with fh.Flock.from_words(w, decide_group=group, decide_block=block) as f:
        build_grid(f)
        f.decide(t)
        got, _ = f.download()
    for i in np.nonzero(t >= w[:, 12])[0]:
        d = fr.decide(w, int(i))
        assert got[i, 13].view(np.uint32) == d["neighbours"], i
        assert rel_err(got[i, 9:12], d["force"]) <= 1e-4, i
        assert abs(got[i, 14] - d["centrality"]) <= 1e-5


def test_step_is_deterministic_in_canonical_mode(fh):
    w0 = cube(fh, 50_000, 3)
    outs = []
    for _ in range(2):
        with fh.Flock.from_words(w0, flags=fh.FLAG_CANONICAL) as f:
            f.step(25)
            outs.append(f.download()[0])
    assert np.array_equal(outs[0].view(np.uint32), outs[1].view(np.uint32))


@pytest.mark.parametrize("n,cell,group,shift", [(200_000, 4.0, 1, 0), (200_000, 6.0, 4, 0), (100_000, 4.0, 8, 3)])
def test_occupancy_skip_is_bit_identical(fh, n, cell, group, shift):
    # Skipping empty buckets through the scan's occupancy bitmap must not
    # change a single bit (a wrong bit would drop a neighbour).
    w0 = cube(fh, n, 6)
    outs = []
    for flags in (fh.FLAG_CANONICAL | fh.FLAG_OCC, fh.FLAG_CANONICAL):
        with fh.Flock.from_words(w0, flags=flags, cell=cell, decide_group=group, bucket_shift=shift) as f:
            f.step(12)
            outs.append(f.download()[0])
    assert np.array_equal(outs[0].view(np.uint32), outs[1].view(np.uint32))


@pytest.mark.parametrize("n,group,block,bpt,graph,mode", [(200_000, 1, 256, 8, 0, 256), (1_000_000, 4, 512, 1, 4, 256),
                                                         (100_000, 16, 128, 2, 4, 256), (3000, 64, 64, 1, 0, 256),
                                                         (1_000_000, 4, 512, 4, 4, 512), (200_003, 1, 64, 2, 0, 512),
                                                         (5000, 16, 128, 1, 0, 512),
                                                         (1_000_000, 4, 512, 2, 4, 1024 | 512 | 128),
                                                         (200_000, 1, 256, 8, 0, 1024),
                                                         (100_000, 16, 128, 4, 0, 1024 | 256),
                                                         (300_000, 64, 512, 16, 4, 1024 | 128),
                                                         (2000, 2, 64, 1, 0, 1024)])
def test_chunk_mapping_is_bit_identical(fh, n, group, block, bpt, graph, mode):
    # Persistent grid (256): chunks are taken from a counter the scan zeroes
    # each step. XCD mapping (512): a permutation of workgroups to chunks.
    # LDS window (1024): candidate positions near the chunk read from LDS.
    # Either way every chunk must be decided exactly once, in every step.
    # Both runs use the same hash: with G >= 16 the lanes split each bucket's
    # candidates by their place in the bucket, so another hash (other
    # collisions) groups the partial sums differently.
    w0 = cube(fh, n, 8)
    outs = []
    for flags in (fh.FLAG_CANONICAL | mode | graph, fh.FLAG_CANONICAL | (mode & fh.FLAG_SPATIAL)):
        with fh.Flock.from_words(w0, flags=flags, decide_group=group, decide_block=block,
                                 birds_per_thread=bpt) as f:
            f.step(9)
            outs.append(f.download()[0])
    assert np.array_equal(outs[0].view(np.uint32), outs[1].view(np.uint32))


def test_persistent_decide_in_a_batch(fh):
    starts = [cube(fh, n, 20 + n) for n in (5000, 60_000, 1)]
    res = []
    for flags in (fh.FLAG_CANONICAL | fh.FLAG_PERSISTENT, fh.FLAG_CANONICAL):
        with fh.Batch(100_000, 1 << 21, 4, flags=flags, birds_per_thread=2) as b:
            ids = [b.add(w) for w in starts]
            b.step(7)
            res.append([b.download(i)[0] for i in ids])
    for a, b in zip(*res):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_graph_and_unfused_replays_match_eager(fh):
    w0 = cube(fh, 30_000, 5)
    res = {}
    for name, flags in [("eager", 1), ("graph", 1 | 4), ("unfused", 1 | 2)]:
        with fh.Flock.from_words(w0, flags=flags) as f:
            snaps = []
            for _ in range(3):  # several replays with host work in between
                f.step(7)
                snaps.append(f.download()[0])
            res[name] = snaps
    for name in ("graph", "unfused"):
        for a, b in zip(res["eager"], res[name]):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), name


@pytest.mark.parametrize("plan", [[40], [16, 16], [17, 33, 5], [3, 20]])
def test_graph_runs_match_eager(fh, plan):
    # fh_step replays runs of 16 steps as one graph (never across the
    # warm-up boundary) and single-step graphs for the rest: every split of
    # the same steps gives the eager state bit for bit (canonical order).
    w0 = cube(fh, 40_000, 6)
    out = {}
    for name, flags in [("eager", 1), ("graph", 1 | 4)]:
        with fh.Flock.from_words(w0, flags=flags) as f:
            for k, n in enumerate(plan):
                f.step(n, warmup=n // 3 if k == 0 else 0, timing=(k == 0))
            out[name] = f.download()
    (a, ta), (b, tb) = out["eager"], out["graph"]
    assert ta == tb
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("plan", [[40], [1, 32, 7], [17, 16]])
def test_resident_graph_runs_match_eager(fh, plan):
    # Resident runs start at buffer parity 0: after an odd number of steps a
    # single step comes first. Slots of arrivals follow atomic order, so sums
    # are reordered from run to run: the check is the decision clocks exactly
    # and positions to 1e-4 after 40 steps (rounding drift, about 1e-5).
    w0 = cube(fh, 40_000, 7)
    out = {}
    for name, flags in [("eager", 64), ("graph", 64 | 4), ("fused", 64 | 4 | 2048)]:
        with fh.Flock.from_words(w0, flags=flags) as f:
            for n in plan:
                f.step(n)
            out[name] = f.download()
    a, ta = out["eager"]
    for name in ("graph", "fused"):
        b, tb = out[name]
        assert ta == tb, name
        assert np.array_equal(a[:, 12], b[:, 12]), name
        assert rel_err(b[:, 0:6], a[:, 0:6]) <= 1e-4, name


def test_pass_order_is_enforced(fh):
    with fh.Flock.from_words(cube(fh, 100)) as f:
        with pytest.raises(fh.FlockHipError) as e:
            f.scatter()
        assert e.value.status == fh.ERR_STATE
        with pytest.raises(fh.FlockHipError) as e:
            f.upload(cube(fh, 99))
        assert e.value.status == fh.ERR_INVALID


def test_empty_flock_steps(fh):
    with fh.Flock(0) as f:
        f.step(3)
        w, t = f.download()
        assert w.shape == (0, 15) and abs(t - 0.015) < 1e-6


def test_decide_instances_fit_their_register_budget(fh):
    # Lane groups of 4 and 8 are compiled for at least 5 waves per SIMD (at
    # most 96 VGPRs of the 512 per lane), groups of 16 and 64 for 6 (80), and
    # groups of 2 for 4 (128). No instance spills, except that the split walks
    # with the mover lists (groups of 16 and 64, kind m) keep a few words in
    # scratch outside the candidate loops.
    for block in (64, 256, 512):
        for group in (2, 4, 8, 16, 64):
            for kind in "rsfm":
                k = fh.kernel_info(f"decide:{block}:{group}:{kind}")
                cap = 128 if group == 2 else 80 if group >= 16 else 96
                assert 0 < k["vgprs"] <= cap, (block, group, kind, k)
                limit = 16 if (kind == "m" and group >= 16) else 0
                assert k["scratch_bytes"] <= limit, (block, group, kind, k)
    with pytest.raises(fh.FlockHipError):
        fh.kernel_info("decide:512:3:r")


def test_timing_report(fh):
    with fh.Flock.from_words(cube(fh, 100_000)) as f:
        ms = f.step(6, 2, timing=True)
    assert ms["timed_steps"] == 4 and ms["decide"] > 0 and ms["step_total"] > 0
    print("timing", ms)
