"""Pins the CPU oracle before it is trusted as the checker.

Every check here is independent of the oracle's own formulas: the reference's
unit tests (crates/flock-ref/src/model.rs:208-269, state.rs:177-197,
cutile/src/main.rs:368-393), known answers held in the reference's docs and
logs, closed forms worked out by hand, and an independent library (scipy's
k-d tree) for neighbour sets.
"""
import math
import os

import numpy as np
import pytest

DT, DU, V0, MASS = 0.005, 0.05, 10.0, 0.08


def rec(pos=(0, 0, 100), vel=(10, 0, 0), radius=1.5, **kw):
    r = np.zeros(15, np.float32)
    r[0:3], r[3:6], r[8] = pos, vel, radius
    for k, v in kw.items():
        r[{"roll": 6, "spin": 7, "next": 12}[k]] = v
    return r


# ---- the reference's own unit tests ------------------------------------------

def test_level_cruise_holds_height_and_speed(fr):
    # model.rs:208-220
    b = rec(pos=(0, 0, 100), vel=(V0, 0, 0), radius=0)
    for _ in range(200):
        b = fr.step_bird(b, DT)
    assert abs(b[2] - 100.0) < 1.0
    assert abs(np.linalg.norm(b[3:6]) - V0) < 0.5


def test_a_lone_bird_gets_only_roost_and_speed_forces(fr):
    # model.rs:222-233
    w = rec(pos=(0, 0, 100), vel=(V0, 0, 0), radius=1.5)[None, :]
    d = fr.decide(w, 0)
    assert d["neighbours"] == 0
    assert np.linalg.norm(d["force"]) < 1e-6


def test_two_close_birds_separate(fr):
    # model.rs:235-258
    w = np.stack([rec((0, 0, 100), (V0, 0, 0), 2.0), rec((0, 0.5, 100), (V0, 0, 0), 2.0)])
    d = fr.decide(w, 0)
    assert d["neighbours"] == 1
    assert d["force"][1] < 0


def test_settled_flock_neighbour_count_is_near_n_c(fr):
    # model.rs:260-269
    w, _ = fr.step(fr.scatter(2000, 1), 0.0, 200, threads=os.cpu_count() or 1)
    mean = w[:, 13].view(np.uint32).astype(np.float32).mean()
    assert abs(mean - 6.5) < 2.0, mean


def test_file_round_trips(fr, tmp_path):
    # state.rs:177-187
    w = fr.scatter(50, 7)
    w[3, 13] = np.uint32(9).view(np.float32)
    p = tmp_path / "a.flk"
    assert fr.write_flk(p, w, 1.25) == 0
    back, t, rc = fr.read_flk(p)
    assert rc == 0 and t == np.float32(1.25)
    assert np.array_equal(back.view(np.uint32), w.view(np.uint32))


def test_scatter_is_repeatable_and_at_density(fr):
    # state.rs:189-197
    a, b = fr.scatter(1000, 1), fr.scatter(1000, 1)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    side = np.float32(np.cbrt(np.float32(1000.0) / np.float32(0.5)))
    assert np.all(np.abs(a[:, 0]) <= side / 2)
    assert np.all(np.abs(np.linalg.norm(a[:, 3:6], axis=1) - V0) < 1e-3)


def test_check_grid_accepts_a_correct_grid(fr):
    # cutile/src/main.rs:368-375
    assert fr.check_grid([2, 0, 3, 3, 0], [0, 2, 2, 3, 5], [1, 4, 0, 2, 3], 4) == (0, "")


def test_check_grid_rejects_a_wrong_total(fr):
    # cutile/src/main.rs:377-384
    rc, msg = fr.check_grid([2, 0, 3, 3, 0], [0, 2, 2, 3, 4], [1, 4, 0, 2, 3], 4)
    assert rc != 0 and "grid_scan_add" in msg


def test_check_grid_rejects_a_bird_in_the_wrong_range(fr):
    # cutile/src/main.rs:386-392
    rc, msg = fr.check_grid([2, 0, 3, 3, 0], [0, 2, 2, 3, 5], [1, 4, 2, 0, 3], 4)
    assert rc != 0 and "grid_scatter" in msg


# ---- known answers held in the reference ---------------------------------------

def test_known_answer_191_birds_decide_on_step_one(fr):
    # docs/progress.md:346 and docs/ARTICLE.md:76: the 2000-bird seed-1 start
    # (scripts/cloud-session.sh:41) has 191 birds deciding on the first step.
    w = fr.scatter(2000, 1)
    assert fr.count_due(w, np.float32(0.0) + np.float32(DT)) == 191


def test_known_answer_1000_step_aggregates(fr):
    # docs/results/rtx4000ada-2026-10-21/wgpu-session/wgpu-parity-1000.log:1-4,
    # column A (the CPU oracle, scripts/cloud-wgpu.sh:28-32): 2000 birds, seed 1,
    # 1000 steps. The run is chaotic: ulp-level changes move mean neighbours by
    # ~0.02 (measured by perturbing every bird by one ulp), so the check is the
    # reference's own statdiff tolerance (main.rs:111).
    w, t = fr.step(fr.scatter(2000, 1), 0.0, 1000, threads=os.cpu_count() or 1)
    s = fr.stats(w)
    ref = np.array([9.9927, 6.6795, 1.5390, 0.9983])
    tol = np.array([0.02, 0.1, 0.01, 0.005])
    assert np.all(np.abs(s - ref) <= tol), (s, ref)
    assert abs(t - 5.0) < 1e-4


# ---- closed forms -----------------------------------------------------------

def test_constants_match_params_rs(fr):
    c = fr.constants()
    f = np.float32
    want = [f(0.005), f(0.05), f(10), f(0.08), f(3.3), f(9.81), f(0.08) * f(9.81),
            (f(0.08) * f(9.81)) / f(3.3), f(1), f(0.15), f(6), f(80) * f(math.pi) / f(180), f(5), f(20),
            f(6.5), f(100), f(0.1), f(0.2), f(1.37), f(1), f(0.5), f(0.65), f(-math.sqrt(0.5)), f(0.2),
            f(40), f(100), f(4), f(0.1) / f(3)]
    assert c.view(np.uint32).tolist() == np.array(want, np.float32).view(np.uint32).tolist()


def test_buckets_for_table(fr):
    # next_pow2(2n) clamped to [2^16, 2^22] (cuda/flock.cu:368-372)
    table = {0: 1 << 16, 1: 1 << 16, 32768: 1 << 16, 32769: 1 << 17, 10_000: 1 << 16,
             1_000_000: 1 << 21, 2_097_152: 1 << 22, 10_000_000: 1 << 22}
    for n, b in table.items():
        assert fr.buckets_for(n) == b, n


def test_cell_hash_closed_form(fr):
    m = (1 << 32) - 1
    assert fr.bucket_of(0, 0, 0, m) == 0
    assert fr.bucket_of(1, 0, 0, m) == 73856093
    assert fr.bucket_of(0, 1, 0, m) == 19349663
    assert fr.bucket_of(0, 0, 1, m) == 83492791
    assert fr.bucket_of(-1, 0, 0, m) == (1 << 32) - 73856093
    assert fr.bucket_of(1, 1, 1, m) == 73856093 ^ 19349663 ^ 83492791
    assert fr.bucket_of(2, 0, 0, 0xFFFF) == (2 * 73856093) & 0xFFFF


def test_cell_coord_floor_and_saturation(fr):
    assert fr.cell_coord(0.0) == 0 and fr.cell_coord(3.999) == 0 and fr.cell_coord(4.0) == 1
    assert fr.cell_coord(-0.001) == -1 and fr.cell_coord(-4.0) == -1 and fr.cell_coord(-4.001) == -2
    assert fr.cell_coord(float("nan")) == 0
    assert fr.cell_coord(1e30) == 2**31 - 1 and fr.cell_coord(-1e30) == -(2**31)


def test_sphere_test_slack(fr):
    # nearest point of cell (1,0,0) to p=(0,2,2) is (4,2,2): distance 4
    p = (0.0, 2.0, 2.0)
    assert fr.cell_intersects_sphere((1, 0, 0), p, 16.0)
    assert fr.cell_intersects_sphere((1, 0, 0), p, 16.0 - 0.5e-4)  # inside the +1e-4 slack
    assert not fr.cell_intersects_sphere((1, 0, 0), p, 16.0 - 2e-4)


def test_radius_step_closed_form(fr):
    # r (N_C / max(n,1))^(0.1/3) clamped to [R_H, R_MAX] (params.rs:61-65)
    for r, n in [(1.5, 0), (1.5, 1), (1.5, 6), (1.5, 7), (2.0, 13), (5.0, 100)]:
        want = r * (6.5 / max(n, 1)) ** (0.1 / 3)
        assert abs(fr.radius_step(r, n) - want) <= 2e-7 * want
    assert fr.radius_step(1.5, 0) == fr.radius_step(1.5, 1)
    assert fr.radius_step(0.1, 1000) == np.float32(0.2)
    assert fr.radius_step(99.9, 0) == np.float32(100.0)


def test_frame_closed_forms(fr):
    assert np.allclose(fr.forward([3, 4, 0]), [0.6, 0.8, 0], atol=1e-7)
    assert np.array_equal(fr.forward([0, 0, 0]), [1, 0, 0])
    assert np.array_equal(fr.side_of([1, 0, 0]), [0, -1, 0])
    assert np.array_equal(fr.side_of([0, 0, 1]), [1, 0, 0])  # fallback when f is vertical
    assert np.allclose(fr.body_up([1, 0, 0], 0.0), [0, 0, 1], atol=1e-7)
    assert np.allclose(fr.body_up([1, 0, 0], math.pi / 2), [0, -1, 0], atol=1e-6)


def test_step_bird_roll_rate_clamp(fr):
    # hard sideways steering: target bank = MAX_BANK, rate clamped to 6 rad/s
    b = rec(vel=(V0, 0, 0), radius=0)
    b[9:12] = (0.0, -100.0, 0.0)  # along side_of(+x) = -y
    out = fr.step_bird(b, DT)
    assert out[7] == np.float32(6.0)
    assert abs(out[6] - 0.03) < 1e-7


def test_step_bird_speed_limits(fr):
    fast = fr.step_bird(rec(vel=(30, 0, 0), radius=0), DT)
    assert abs(np.linalg.norm(fast[3:6]) - 20.0) < 1e-5
    slow = fr.step_bird(rec(vel=(1, 0, 0), radius=0), DT)
    assert abs(np.linalg.norm(slow[3:6]) - 5.0) < 1e-5


def test_decide_pair_is_antisymmetric(fr):
    w = np.stack([rec((0, 0, 100), (V0, 0, 0), 2.0), rec((0, 0.5, 100), (V0, 0, 0), 2.0)])
    a, b = fr.decide(w, 0), fr.decide(w, 1)
    assert a["neighbours"] == b["neighbours"] == 1
    assert a["force"][1] == -b["force"][1] and a["force"][0] == b["force"][0]


def test_neighbour_sets_match_kdtree(fr):
    spatial = pytest.importorskip("scipy.spatial")
    w = fr.scatter(3000, 3)
    w[:, 8] = np.float32(2.0)
    pts = w[:, :3].astype(np.float64)
    tree = spatial.cKDTree(pts)
    for i in range(0, 3000, 97):
        for walk, rad in ((1, 2.0), (2, 4.0)):
            got = set(fr.neighbours_brute(w, i, walk).tolist())
            inner = set(tree.query_ball_point(pts[i], rad - 1e-5)) - {i}
            outer = set(tree.query_ball_point(pts[i], rad + 1e-5)) - {i}
            assert inner <= got <= outer, (i, walk)


def test_grid_decide_matches_brute_force(fr):
    # The grid walk (cuda/flock.cu:190-290) and the brute force (model.rs:53-145)
    # must find the same neighbours; sums differ only in order.
    w = fr.scatter(3000, 5)
    keys, counts, offsets, srt = fr.grid(w)
    b = fr.buckets_for(3000)
    for i in range(0, 3000, 7):
        a, g = fr.decide(w, i), fr.decide_grid(w, i, offsets, srt, b)
        assert a["neighbours"] == g["neighbours"]
        assert np.allclose(a["force"], g["force"], rtol=1e-5, atol=1e-6)
        assert abs(a["centrality"] - g["centrality"]) <= 1e-5
        assert a["radius"] == g["radius"]


def test_grid_is_a_valid_stable_counting_sort(fr):
    w = fr.scatter(5000, 2)
    keys, counts, offsets, srt = fr.grid(w)
    b = fr.buckets_for(5000)
    assert fr.check_grid(keys, offsets, srt, b) == (0, "")
    assert np.array_equal(np.bincount(keys, minlength=b), counts)
    # stable: sorting by key with a stable sort over bird index
    assert np.array_equal(np.argsort(keys, kind="stable"), srt)


def test_threads_do_not_change_the_result(fr):
    w0 = fr.scatter(1500, 4)
    a, _ = fr.step(w0, 0.0, 30, threads=1)
    b, _ = fr.step(w0, 0.0, 30, threads=4)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_decide_many_equals_decide_per_bird(fr):
    # the multi-core baseline's entry point: the same brute-force decisions,
    # bit for bit, whatever the thread count
    w = fr.scatter(2500, 6)
    birds = np.arange(0, 2500, 9)
    for threads in (1, 4):
        m = fr.decide_many(w, birds, threads=threads)
        for k, i in enumerate(birds):
            d = fr.decide(w, int(i))
            assert np.array_equal(m["force"][k].view(np.uint32), d["force"].view(np.uint32))
            assert m["neighbours"][k] == d["neighbours"]
            assert m["radius"][k] == d["radius"] and m["centrality"][k] == d["centrality"]


def test_diff_and_statdiff_semantics(fr):
    a = fr.scatter(100, 1)
    b = a.copy()
    assert fr.diff(a, b)["ok"]
    b[5, 0] += np.float32(5e-3)
    d = fr.diff(a, b)
    assert not d["ok"] and d["outliers"] == 1
    assert fr.diff(a, b, allow_outliers=2, outlier_max=1e-2)["ok"]
    b[6, 13] = np.uint32(3).view(np.float32)
    assert fr.diff(a, b, allow_outliers=2)["nb_mismatch"] == 1
    assert fr.statdiff(a, a)["ok"]


def test_flk_errors(fr, tmp_path):
    p = tmp_path / "bad.flk"
    p.write_bytes(b"FLK2" + bytes(12))
    assert fr.read_flk(p)[2] == 2
    p.write_bytes(b"FLK1" + np.array([1, 14, 0], np.uint32).tobytes())
    assert fr.read_flk(p)[2] == 3
    p.write_bytes(b"FLK1" + np.array([2, 15, 0], np.uint32).tobytes() + bytes(60))
    assert fr.read_flk(p)[2] == 4
    assert fr.read_flk(tmp_path / "missing.flk")[2] == 1


@pytest.mark.parametrize("kind,n,cell", [("ref", 2000, 4.0), ("cube", 10_000, 4.0), ("cube", 10_000, 8.0)])
def test_grid_step_follows_the_brute_force_step(fr, kind, n, cell):
    # fr.step_grid (the grid walk of cuda/flock.cu:190-290 over the stable
    # counting sort) is the checker of every large configuration; it must
    # follow fr.step (brute force, model.rs:182-200) at the reference's own
    # bar: 1, 5 and 20 steps pointwise within 1e-3 with equal neighbour
    # counts and decision times (docs/SPEC.md:66-72). Sums differ only in
    # order, so one step is also within 1e-5.
    if kind == "ref":
        w0 = fr.scatter(n, 1)
    else:
        side = 4.0 * (n / 0.64) ** (1.0 / 3.0)
        rng = np.random.default_rng(3)
        w0 = np.zeros((n, 15), np.float32)
        w0[:, 0:3] = rng.uniform(-side / 2, side / 2, (n, 3))
        w0[:, 2] += 100.0
        d = rng.normal(size=(n, 3))
        w0[:, 3:6] = d / np.linalg.norm(d, axis=1, keepdims=True) * rng.uniform(9, 11, (n, 1))
        w0[:, 8] = 1.5
        w0[:, 12] = rng.uniform(0, 0.05, n)
    for steps in (1, 5, 20):
        a, ta = fr.step(w0, 0.0, steps, threads=4)
        g, tg = fr.step_grid(w0, 0.0, steps, cell=cell, threads=4)
        assert ta == tg
        assert np.array_equal(a[:, 12], g[:, 12])
        d = fr.diff(a, g)
        assert d["ok"], (steps, d)
        if steps == 1:
            assert d["dpos"] <= 1e-5 and d["dvel"] <= 1e-5, d
