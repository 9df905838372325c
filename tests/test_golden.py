"""The committed .flk fixtures of tests/golden/ (written by make_golden.py
with the C oracle): intact, readable by the library's own .flk reader, and
still what the oracle produces from the reference's 2000-bird seed-1 start,
whose first step has 191 due birds (docs/progress.md:345-346)."""
import hashlib
import json
import os

import numpy as np

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
META = json.load(open(os.path.join(HERE, "golden.json")))
DT = np.float32(0.005)


def path(name):
    return os.path.join(HERE, name)


def test_fixtures_are_intact():
    for name, m in META["files"].items():
        with open(path(name), "rb") as fp:
            data = fp.read()
        assert hashlib.sha256(data).hexdigest() == m["sha256"], name
        assert len(data) == 16 + 60 * m["birds"], name  # header + 15 f32 words per bird


def test_start_is_the_reference_scatter_with_191_due(fr):
    w, t, rc = fr.read_flk(path("s2000_seed1.flk"))
    assert rc == 0 and t == 0.0
    assert np.array_equal(w.view(np.uint32), fr.scatter(2000, 1).view(np.uint32))
    assert fr.count_due(w, np.float32(0.0) + DT) == META["known_answers"]["due_on_step_1"]["value"]


def test_oracle_still_produces_the_stepped_fixtures(fr):
    start, _, _ = fr.read_flk(path("s2000_seed1.flk"))
    for k in (1, 20):
        want, tw, rc = fr.read_flk(path(f"s2000_seed1_step{k}.flk"))
        assert rc == 0
        got, tg = fr.step(start, 0.0, k, threads=os.cpu_count() or 1)
        assert np.float32(tg) == np.float32(tw)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), k


def test_library_reader_agrees_with_the_oracle_reader(fh, fr):
    for name in META["files"]:
        a, ta = fh.read_flk(path(name))
        b, tb, _ = fr.read_flk(path(name))
        assert np.float32(ta) == np.float32(tb)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
