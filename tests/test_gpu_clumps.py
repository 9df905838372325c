"""Clustered flock: Gaussian clumps with thousands of birds per 4 m cell.

Parity on sampled due birds against the oracle's grid walk over the whole
flock (the same neighbour semantics as brute force, pinned in
test_oracle.py), and against brute force for a few of them."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)
DT = np.float32(0.005)


def rel_err(a, b):
    a = a.astype(np.float64)
    b = b.astype(np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(1.0, np.abs(b)), initial=0.0))


@pytest.mark.parametrize("n,cell,flags,group", [(1_000_000, 4.0, 0, 4), (1_000_000, 2.0, 8, 4),
                                               (1_000_000, 1.0, 32, 16), (1_000_000, 2.0, 0, 64),
                                               (1_000_000, 2.0, 0, 2)])
def test_clumped_decide_matches_the_oracle(fh, fr, n, cell, flags, group):
    w = fh.gen_clumps(n, 8, 5.0, 1000.0, 3)
    t = np.float32(0.0) + DT
    with fh.Flock.from_words(w, cell=cell, flags=flags, decide_group=group) as f:
        f.hash_count()
        g = f.read_grid(offsets=False, sorted=False)
        # dense buckets really occur: the peak density of a 125k-bird clump of
        # sigma 5 m is 63.5 birds/m^3, i.e. ~500 per 2 m cell and ~4000 per 4 m cell
        assert int(g.counts.max()) > 300 * (cell / 2.0) ** 3
        f.scan()
        f.scatter()
        f.decide(t)
        got, _ = f.download()
    due = np.nonzero(t >= w[:, 12])[0]
    sample = np.sort(np.random.default_rng(0).choice(due, 2000, replace=False))
    keys, counts, offsets, srt = fr.grid(w)
    d = fr.decide_grid_many(w, sample, offsets, srt, fr.buckets_for(n), threads=THREADS)
    assert np.array_equal(got[sample, 13].view(np.uint32), d["neighbours"])
    assert rel_err(got[sample, 9:12], d["force"]) <= 1e-4
    assert rel_err(got[sample, 14], d["centrality"]) <= 1e-4
    for k, i in enumerate(sample[:6]):
        assert fr.decide(w, int(i))["neighbours"] == d["neighbours"][k]


@pytest.mark.parametrize("cell,block,bpt", [(1.0, 128, 2), (2.0, 256, 8), (4.0, 512, 16)])
def test_tiled_decide_is_bit_identical_to_the_lane_walk(fh, cell, block, bpt):
    # Dense clumps take the tiled path; it must give the per-lane walk's bits.
    w = fh.gen_clumps(300_000, 4, 4.0, 200.0, 5)
    outs = []
    for flags in (fh.FLAG_CANONICAL | fh.FLAG_OCC, fh.FLAG_CANONICAL | fh.FLAG_NO_TILE):
        with fh.Flock.from_words(w, cell=cell, decide_block=block, birds_per_thread=bpt, flags=flags) as f:
            f.step(3)
            outs.append(f.download()[0])
    assert np.array_equal(outs[0].view(np.uint32), outs[1].view(np.uint32))


def test_tiled_decide_matches_the_oracle(fh, fr):
    n = 400_000
    w = fh.gen_clumps(n, 8, 5.0, 300.0, 9)
    t = np.float32(0.0) + DT
    with fh.Flock.from_words(w, cell=1.0, decide_block=256, birds_per_thread=4) as f:
        f.hash_count()
        f.scan()
        f.scatter()
        f.decide(t)
        got, _ = f.download()
    due = np.nonzero(t >= w[:, 12])[0]
    sample = np.sort(np.random.default_rng(2).choice(due, 3000, replace=False))
    keys, counts, offsets, srt = fr.grid(w)
    d = fr.decide_grid_many(w, sample, offsets, srt, fr.buckets_for(n), threads=THREADS)
    assert np.array_equal(got[sample, 13].view(np.uint32), d["neighbours"])
    assert rel_err(got[sample, 9:12], d["force"]) <= 1e-4
    assert rel_err(got[sample, 14], d["centrality"]) <= 1e-4


def test_clumped_step_parity_10m_sample(fh, fr):
    # The configured workload, 10M birds: one step, checked on 50,000 due birds
    # sampled from the whole flock.
    n = 10_000_000
    w = fh.gen_clumps(n, 8, 5.0, 1000.0, 1)
    with fh.Flock.from_words(w, cell=2.0, decide_group=4) as f:
        f.step(1)
        gpu, tg = f.download()
    due = np.nonzero(np.float32(0.0) + DT >= w[:, 12])[0]
    sample = np.sort(np.random.default_rng(1).choice(due, 50_000, replace=False))
    keys, counts, offsets, srt = fr.grid(w)
    d = fr.decide_grid_many(w, sample, offsets, srt, fr.buckets_for(n), threads=THREADS)
    assert np.array_equal(gpu[sample, 13].view(np.uint32), d["neighbours"])
    assert rel_err(gpu[sample, 9:12], d["force"]) <= 1e-4
    # positions and velocities after the step follow step_bird of the decided state
    for k in range(0, 2000):
        i = sample[k]
        r = w[i].copy()
        r[9:12] = d["force"][k]
        r[8] = d["radius"][k]
        r[12] += np.float32(0.05)
        want = fr.step_bird(r)
        assert rel_err(gpu[i, 0:6], want[0:6]) <= 1e-4, i
