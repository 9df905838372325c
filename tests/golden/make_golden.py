#!/usr/bin/env python3
"""Writes the golden .flk fixtures of this directory with the C oracle.

The reference commits no fixtures, so these are self-generated. They are
pinned by checks that do not depend on the oracle's own formulas:
- the start is the reference's `Flock::scatter(2000, 1)` (crates/flock-ref/src/state.rs:58-101), the
  s2000.flk of scripts/cloud-session.sh:41, and 191 of its birds decide on
  the first step (docs/progress.md:345-346, docs/ARTICLE.md:76);
- the oracle that steps it is pinned in tests/test_oracle.py (the reference's
  unit tests, its 1000-step aggregates, closed forms, a k-d tree).

Files: s2000_seed1.flk (time 0), s2000_seed1_step1.flk and
s2000_seed1_step20.flk (the oracle's brute-force Flock::step, model.rs:182-200),
and golden.json (their SHA-256 and the known answers).

    python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from oracle import flockref as fr  # noqa: E402

DT = np.float32(0.005)


def sha(path: str) -> str:
    with open(path, "rb") as fp:
        return hashlib.sha256(fp.read()).hexdigest()


def main() -> None:
    threads = os.cpu_count() or 1
    start = fr.scatter(2000, 1)
    states = {"s2000_seed1.flk": (start, 0.0)}
    for k in (1, 20):
        w, t = fr.step(start, 0.0, k, threads=threads)
        states[f"s2000_seed1_step{k}.flk"] = (w, t)
    meta = {"files": {}, "known_answers": {
        "due_on_step_1": {"value": 191, "source": "docs/progress.md:345-346; docs/ARTICLE.md:76"},
    }}
    for name, (w, t) in states.items():
        path = os.path.join(HERE, name)
        assert fr.write_flk(path, w, t) == 0, name
        meta["files"][name] = {"birds": int(w.shape[0]), "time": t, "sha256": sha(path)}
    assert fr.count_due(start, np.float32(0.0) + DT) == 191
    with open(os.path.join(HERE, "golden.json"), "w") as fp:
        json.dump(meta, fp, indent=1)
        fp.write("\n")
    print(json.dumps(meta, indent=1))


if __name__ == "__main__":
    main()
