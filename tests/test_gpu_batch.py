"""Batched stepping of many flocks: each flock's result equals stepping it
alone (bit-identical in canonical order), the oracle agrees with batched
flocks, and flocks can join and leave between steps."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CANON = 1
GRAPH = 4


def cube(fh, n, seed):
    return fh.gen_cube(n, fh.cube_side_for(n), seed)


def alone(fh, w, steps, flags=CANON, time=0.0):
    with fh.Flock.from_words(w, time, flags=flags) as f:
        f.step(steps)
        return f.download()


def same_bits(a, b):
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("flags", [CANON, CANON | GRAPH])
def test_batch_equals_alone(fh, flags):
    sizes = [1000, 3000, 257, 20_000, 1, 64_000]
    starts = [cube(fh, n, 10 + i) for i, n in enumerate(sizes)]
    with fh.Batch(200_000, 1 << 22, 16, flags=flags) as b:
        ids = [b.add(w) for w in starts]
        b.step(12)
        for fid, w in zip(ids, starts):
            got, t = b.download(fid)
            want, tw = alone(fh, w, 12)
            assert t == tw
            assert same_bits(got, want), fid


def test_batch_matches_the_oracle(fh, fr):
    w0 = fh.gen_reference(2000, 1)
    other = cube(fh, 30_000, 2)
    with fh.Batch(100_000, 1 << 21, 4, flags=CANON) as b:
        b.add(other)
        fid = b.add(w0)
        b.step(10)
        got, t = b.download(fid)
    cpu, tc = fr.step(w0, 0.0, 10)
    d = fr.diff(cpu, got)
    assert t == tc and d["ok"], d


def test_flocks_join_and_leave_between_steps(fh):
    a, c, d = cube(fh, 5000, 1), cube(fh, 9000, 2), cube(fh, 700, 3)
    with fh.Batch(50_000, 1 << 21, 8, flags=CANON) as b:
        ia = b.add(a)
        b.step(5)
        ic = b.add(c, 0.0)           # joins at its own clock 0
        b.step(7)
        b.remove(ia)
        idd = b.add(d, 0.0)          # reuses the freed room
        b.step(4)
        got_c, tc = b.download(ic)
        got_d, td = b.download(idd)
    want_c, twc = alone(fh, c, 11)
    want_d, twd = alone(fh, d, 4)
    assert tc == twc and td == twd
    assert same_bits(got_c, want_c) and same_bits(got_d, want_d)


def test_clock_per_flock(fh):
    w = cube(fh, 2000, 4)
    with fh.Batch(10_000, 1 << 18, 4, flags=CANON) as b:
        i0 = b.add(w, 0.0)
        i1 = b.add(w, 1.25)
        b.step(3)
        _, t0 = b.download(i0)
        got1, t1 = b.download(i1)
    want1, tw1 = alone(fh, w, 3, time=1.25)
    assert t0 == np.float32(np.float32(np.float32(0.005) + np.float32(0.005)) + np.float32(0.005))
    assert t1 == tw1 and same_bits(got1, want1)


def test_budget_is_enforced(fh):
    with fh.Batch(4096, 1 << 17, 2) as b:
        b.add(cube(fh, 4000, 1))
        with pytest.raises(fh.FlockHipError) as e:
            b.add(cube(fh, 200, 2))
        assert e.value.status == fh.ERR_BUDGET
    with fh.Batch(100_000, 1 << 20, 1) as b:
        b.add(cube(fh, 100, 1))
        with pytest.raises(fh.FlockHipError) as e:
            b.add(cube(fh, 100, 2))
        assert e.value.status == fh.ERR_BUDGET


def test_info_and_bytes(fh):
    with fh.Batch(100_000, 1 << 21, 4) as b:
        b.add(cube(fh, 1000, 1))
        b.add(cube(fh, 70_000, 2))
        info = b.info()
    assert info["live_flocks"] == 2 and info["live_birds"] == 71_000
    assert info["bucket_range"] >= (1 << 16) + (1 << 18)
    assert fh.flock_bytes(1000) == 1024 * 108 + (1 << 16) * 8 + (1 << 16) // 8


@pytest.mark.parametrize("flags", [CANON, CANON | GRAPH])
def test_scan_beyond_the_merged_tile_count(fh, fr, flags):
    # More than 2048 scan tiles of 4096 buckets (a bucket range over 8.4M):
    # the scan takes its separate tile-prefix pass. Three flocks of 2^22
    # buckets each; every flock must equal stepping it alone, and the
    # reference start the oracle.
    ws = [fh.gen_reference(2000, 1), cube(fh, 7000, 81), cube(fh, 30_000, 82)]
    with fh.Batch(100_000, 13 << 20, 4, flags=flags, buckets=1 << 22) as b:
        ids = [b.add(w) for w in ws]
        assert b.info()["bucket_range"] > 2048 * 4096
        b.step(5)
        got = [b.download(i) for i in ids]
    for (g, t), w in zip(got, ws):
        with fh.Flock.from_words(w, flags=CANON, buckets=1 << 22) as f:
            f.step(5)
            want, tw = f.download()
        assert t == tw
        assert same_bits(g, want)
    cpu, tc = fr.step(ws[0], 0.0, 5)
    d = fr.diff(cpu, got[0][0])
    assert got[0][1] == tc and d["ok"], d
