"""Batched stepping in the resident layout (FH_FLAG_RESIDENT on a batch):
every flock's state stays in bucket order between steps, as for one flock.
Arrivals take their slots in atomic order, so results are compared with
stepping alone by the decision clocks (exactly) and positions, velocities
and neighbour counts (to rounding of reordered sums); the oracle agrees
with a batched flock like with a flock stepped alone.

With FH_FLAG_LAZY (and the fused flight) the layout is rebuilt every fourth
step; in between the flocks' clocks and budgets follow from the segment
table and the step's phase, and every membership, budget or download call
first rebuilds the layout of the current positions. The same comparisons
hold at every phase: the lazy flag sets below run through each test."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RES, GRAPH, CANON = 64, 4, 1
FLIGHT, LAZY = 2048, 4096
LZ = RES | FLIGHT | LAZY


def cube(fh, n, seed):
    return fh.gen_cube(n, fh.cube_side_for(n), seed)


def alone(fh, w, steps, flags, time=0.0):
    with fh.Flock.from_words(w, time, flags=flags) as f:
        f.step(steps)
        return f.download()


def rel_err(a, b):
    a = a.astype(np.float64)
    b = b.astype(np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(1.0, np.abs(b)))) if a.size else 0.0


def close(got, want, tol=1e-4):
    assert np.array_equal(got[:, 12], want[:, 12])  # decision times
    assert rel_err(got[:, 0:6], want[:, 0:6]) <= tol
    assert np.mean(got[:, 13].view(np.uint32) == want[:, 13].view(np.uint32)) >= 0.99


@pytest.mark.parametrize("flags", [RES, RES | GRAPH, RES | GRAPH | 2048, RES | 128 | 512 | GRAPH,
                                   RES | 256 | 1024, RES | 2048 | 128, LZ, LZ | GRAPH | 128, LZ | 256 | 512])
def test_resident_batch_equals_alone(fh, flags):
    sizes = [1000, 3000, 257, 20_000, 1, 64_000]
    starts = [cube(fh, n, 30 + i) for i, n in enumerate(sizes)]
    with fh.Batch(200_000, 1 << 22, 16, flags=flags, decide_group=4) as b:
        ids = [b.add(w) for w in starts]
        b.step(12)
        for fid, w in zip(ids, starts):
            got, t = b.download(fid)
            want, tw = alone(fh, w, 12, CANON)
            assert t == tw
            close(got, want)


@pytest.mark.parametrize("flags", [RES | GRAPH | 2048, LZ | GRAPH | 128])
def test_resident_batch_matches_the_oracle(fh, fr, flags):
    w0 = fh.gen_reference(2000, 1)
    other = cube(fh, 30_000, 2)
    with fh.Batch(100_000, 1 << 21, 4, flags=flags) as b:
        b.add(other)
        fid = b.add(w0)
        b.step(10)
        got, t = b.download(fid)
    cpu, tc = fr.step(w0, 0.0, 10)
    d = fr.diff(cpu, got)
    assert t == tc and d["ok"], d


@pytest.mark.parametrize("res_flags", [RES | GRAPH, LZ | GRAPH])
def test_resident_batch_join_and_leave(fh, res_flags):
    # Adding and removing splice the flock into or out of the resident layout
    # in place; a download flushes that flock's slots only.
    a, c, d = cube(fh, 5000, 1), cube(fh, 9000, 2), cube(fh, 700, 3)
    out = {}
    for name, flags in (("resident", res_flags), ("canonical", CANON)):
        with fh.Batch(50_000, 1 << 21, 8, flags=flags) as b:
            ia = b.add(a)
            b.step(5)
            ic = b.add(c, 0.0)
            b.step(7)
            b.download(ia)
            b.step(3)  # an odd number of steps: the next entry restarts at parity 0
            b.remove(ia)
            idd = b.add(d, 0.0)
            b.step(4)
            out[name] = (b.download(ic), b.download(idd))
    for (g, tg), (w, tw) in zip(out["resident"], out["canonical"]):
        assert tg == tw
        close(g, w)


@pytest.mark.parametrize("flags", [RES | GRAPH, RES | GRAPH | 2048 | 128, RES | 256 | 512, LZ | GRAPH | 128])
def test_resident_batch_splices_follow_the_canonical_batch(fh, flags):
    # Flocks join and leave a resident batch in place: at the end of the
    # bucket space, in the ranges of flocks that left, the first and a middle
    # flock leaving, two changes between the same steps, at both buffer
    # parities. Every flock still live must follow the canonical batch.
    ws = [cube(fh, n, 20 + k) for k, n in enumerate((3000, 40_000, 700, 12_000, 2500, 5000, 60_000))]
    out = {}
    for name, fl in (("resident", flags), ("canonical", CANON)):
        with fh.Batch(300_000, 1 << 22, 8, flags=fl) as b:
            ids = [b.add(ws[0]), b.add(ws[1]), b.add(ws[2])]
            b.step(3)
            ids.append(b.add(ws[3], 0.0))  # joins after every flock
            b.step(2)
            b.remove(ids[0])  # the first flock leaves
            b.step(1)
            ids.append(b.add(ws[4], 0.0))  # fits the first flock's ranges
            b.remove(ids[2])  # a middle flock leaves
            ids.append(b.add(ws[5], 0.0))
            b.step(5)
            mid = b.download(ids[3])  # a download leaves the layout as it is
            b.step(2)
            ids.append(b.add(ws[6], 0.0))
            b.step(3)
            out[name] = [mid] + [b.download(i) for i in (ids[1], ids[3], ids[4], ids[5], ids[6])]
    for (g, tg), (w, tw) in zip(out["resident"], out["canonical"]):
        assert tg == tw
        close(g, w)


@pytest.mark.parametrize("res_flags", [RES | 128 | 2048, LZ | 128, LZ | GRAPH])
def test_resident_batch_random_joins_and_leaves(fh, res_flags):
    # A seeded sequence of adds, removes, downloads and steps over up to 8
    # flocks of mixed sizes; the resident batch (splicing flocks in and out)
    # must keep every live flock equal to the canonical batch.
    rng = np.random.default_rng(5)
    sizes = [int(x) for x in rng.integers(300, 30_000, 14)]
    ws = [cube(fh, n, 40 + k) for k, n in enumerate(sizes)]
    ops = []
    live, nxt = [], 0
    for _ in range(24):
        r = rng.random()
        if (r < 0.4 or not live) and nxt < len(ws) and len(live) < 8:
            ops.append(("add", nxt))
            live.append(nxt)
            nxt += 1
        elif r < 0.6 and len(live) > 1:
            k = live.pop(int(rng.integers(len(live))))
            ops.append(("remove", k))
        elif r < 0.7 and live:
            ops.append(("download", live[int(rng.integers(len(live)))]))
        else:
            ops.append(("step", int(rng.integers(1, 4))))
    ops.append(("step", 3))
    out = {}
    for name, fl in (("resident", res_flags), ("canonical", CANON)):
        with fh.Batch(300_000, 1 << 22, 8, flags=fl) as b:
            ids, mids = {}, []
            for op, arg in ops:
                if op == "add":
                    ids[arg] = b.add(ws[arg], 0.0)
                elif op == "remove":
                    b.remove(ids.pop(arg))
                elif op == "download":
                    mids.append(b.download(ids[arg]))
                else:
                    b.step(arg)
            out[name] = mids + [b.download(ids[k]) for k in sorted(ids)]
    assert len(out["resident"]) == len(out["canonical"]) >= 3
    for (g, tg), (w, tw) in zip(out["resident"], out["canonical"]):
        assert tg == tw
        close(g, w)


@pytest.mark.parametrize("flags", [RES, LZ])
def test_resident_batch_clock_per_flock(fh, flags):
    w = cube(fh, 2000, 4)
    with fh.Batch(10_000, 1 << 18, 4, flags=flags) as b:
        i0 = b.add(w, 0.0)
        i1 = b.add(w, 1.25)
        b.step(3)
        _, t0 = b.download(i0)
        got1, t1 = b.download(i1)
    want1, tw1 = alone(fh, w, 3, CANON, time=1.25)
    assert t0 == np.float32(np.float32(np.float32(0.005) + np.float32(0.005)) + np.float32(0.005))
    assert t1 == tw1
    close(got1, want1)


@pytest.mark.parametrize("flags", [RES, LZ])
def test_resident_batch_timing_and_bytes(fh, flags):
    starts = [cube(fh, n, 40 + n) for n in (30_000, 50_000)]
    with fh.Batch(100_000, 1 << 21, 4, flags=flags) as b:
        for w in starts:
            b.add(w)
        ms = b.step(4, 1, timing=True)
    assert ms["timed_steps"] == 3
    assert ms["decide"] > 0 and ms["grid_scatter"] > 0
    assert ms["integrate"] > 0 or flags & FLIGHT  # the fused flight runs inside decide
    assert 4_000 < ms["due_last"] < 12_000  # about a tenth of the 80,000 birds decide each step
    assert fh.flock_bytes(1000, flags=RES) > fh.flock_bytes(1000)
    assert fh.flock_bytes(1000, flags=LZ) > fh.flock_bytes(1000, flags=RES | FLIGHT)


def test_resident_batch_rejects_canonical(fh):
    with pytest.raises(fh.FlockHipError) as e:
        fh.Batch(1000, 1 << 16, 2, flags=RES | CANON)
    assert e.value.status == fh.ERR_INVALID


@pytest.mark.parametrize("flags", [RES, RES | GRAPH, RES | GRAPH | 2048, LZ, LZ | GRAPH])
def test_step_budgets_freeze_each_flock_on_its_last_step(fh, flags):
    # Budgets of 3, 7 and unlimited steps, then 10 steps of the batch: each
    # flock matches that many steps alone, and a frozen flock continues once
    # its budget is raised. Flocks of 100k+ birds give the flight hundreds of
    # workgroups, so a budget committed while some of them still read it
    # would freeze part of a flock a step early.
    starts = [cube(fh, n, 50 + n) for n in (120_000, 150_000, 2500)]
    with fh.Batch(400_000, 1 << 22, 4, flags=flags) as b:
        ids = [b.add(w) for w in starts]
        b.limit(ids[0], 3)
        b.limit(ids[1], 7)
        b.step(6)
        b.step(4)
        got = [b.download(i) for i in ids]
        b.limit(ids[0], 2)  # thawed for two more steps
        b.step(5)
        again, t_again = b.download(ids[0])
    for (g, t), w, k in zip(got, starts, (3, 7, 10)):
        want, tw = alone(fh, w, k, CANON)
        assert t == tw, k
        close(g, want)
    want, tw = alone(fh, starts[0], 5, CANON)
    assert t_again == tw
    close(again, want)


def test_step_budgets_need_the_resident_layout(fh):
    with fh.Batch(10_000, 1 << 18, 2, flags=CANON) as b:
        fid = b.add(cube(fh, 100, 1))
        with pytest.raises(fh.FlockHipError) as e:
            b.limit(fid, 5)
        assert e.value.status == fh.ERR_INVALID


@pytest.mark.parametrize("flags", [RES, RES | GRAPH | 2048, LZ | GRAPH])
def test_empty_resident_batch_steps_and_takes_new_flocks(fh, flags):
    # Every flock leaves a resident batch, the empty batch steps, and a new
    # flock joins it: that flock must follow stepping alone.
    a, c = cube(fh, 3000, 61), cube(fh, 5000, 62)
    with fh.Batch(50_000, 1 << 21, 4, flags=flags) as b:
        ia = b.add(a)
        b.step(3)
        b.remove(ia)
        assert b.info()["live_flocks"] == 0
        b.step(2)
        ic = b.add(c, 0.0)
        b.step(5)
        got, t = b.download(ic)
    want, tw = alone(fh, c, 5, CANON)
    assert t == tw
    close(got, want)


@pytest.mark.parametrize("flags", [RES | GRAPH, RES | GRAPH | 2048 | 128, LZ | GRAPH | 128])
def test_resident_scan_beyond_the_direct_tile_count(fh, fr, flags):
    # A resident layout of more than 2048 scan tiles (over 2M buckets) sums
    # the tile totals through super-tile totals; its flocks must follow the
    # canonical batch, and a small one the oracle.
    small = fh.gen_reference(2000, 1)
    ws = [cube(fh, 300_000, 71), cube(fh, 400_000, 72), small]
    out = {}
    for name, fl in (("resident", flags), ("canonical", CANON)):
        with fh.Batch(800_000, 5 << 20, 4, flags=fl) as b:
            ids = [b.add(w) for w in ws]
            assert b.info()["bucket_range"] > 2048 * 1024
            b.step(6)
            out[name] = [b.download(i) for i in ids]
    for (g, tg), (w, tw) in zip(out["resident"], out["canonical"]):
        assert tg == tw
        close(g, w)
    cpu, tc = fr.step(small, 0.0, 6)
    d = fr.diff(cpu, out["resident"][2][0])
    assert out["resident"][2][1] == tc and d["ok"], d


@pytest.mark.parametrize("steps", [1, 2, 3, 4, 5, 7, 8, 9, 13])
def test_lazy_batch_at_every_phase(fh, steps):
    # Steps ending at every phase of the period (so downloads settle from
    # each), in graph periods and single steps; flocks of the batch overlap in
    # space, so a mover listed by another flock would show in the neighbour
    # counts. Budgets end inside periods.
    sizes = [40_000, 3000, 1, 257, 20_000, 90_000]
    starts = [cube(fh, n, 80 + i) for i, n in enumerate(sizes)]
    budgets = [None, 2, None, 5, 3, None]
    with fh.Batch(200_000, 1 << 22, 8, flags=LZ | GRAPH | 128, decide_group=4) as b:
        ids = [b.add(w) for w in starts]
        for i, k in zip(ids, budgets):
            if k is not None:
                b.limit(i, k)
        b.step(steps)
        got = [b.download(i) for i in ids]
    for (g, t), w, k in zip(got, starts, budgets):
        want, tw = alone(fh, w, steps if k is None else min(k, steps), CANON)
        assert t == tw
        close(g, want)


def test_lazy_batch_graph_periods_match_single_steps(fh):
    # The same lazy batch stepped as graph periods and as single steps,
    # split in several ways: equal clocks and decision times.
    starts = [cube(fh, n, 90 + n) for n in (60_000, 7000, 25_000)]
    out = {}
    for name, flags, splits in (("graph", LZ | GRAPH, (8, 3, 5, 4)), ("eager", LZ, (1, 6, 2, 11))):
        with fh.Batch(200_000, 1 << 22, 4, flags=flags) as b:
            ids = [b.add(w) for w in starts]
            for k in splits:
                b.step(k)
            out[name] = [b.download(i) for i in ids]
    for (g, tg), (w, tw) in zip(out["graph"], out["eager"]):
        assert tg == tw
        close(g, w)
