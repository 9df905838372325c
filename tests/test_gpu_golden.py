"""The HIP step against the committed fixtures of tests/golden/ alone (no
oracle call at run time): the reference's 2000-bird seed-1 start, stepped 1
and 20 times by the oracle."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
META = json.load(open(os.path.join(HERE, "golden.json")))


def load(fh, name):
    return fh.read_flk(os.path.join(HERE, name))


def abs_err(a, b):
    return float(np.max(np.abs(a.astype(np.float64) - b.astype(np.float64))))


def rel_err(a, b):
    a = a.astype(np.float64)
    b = b.astype(np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(1.0, np.abs(b))))


@pytest.mark.parametrize("flags", [0, 1, 4, 64 | 4, 128 | 4, 512 | 256, 1024 | 128 | 64 | 4, 2048 | 64 | 4])
def test_one_step_matches_the_fixture(fh, flags):
    # one step: neighbour counts exact, positions and velocities within 1e-4
    start, _ = load(fh, "s2000_seed1.flk")
    want, tw = load(fh, "s2000_seed1_step1.flk")
    with fh.Flock.from_words(start, flags=flags) as f:
        ms = f.step(1, 0, timing=True)
        got, t = f.download()
    assert np.float32(t) == np.float32(tw)
    assert ms["due_last"] == META["known_answers"]["due_on_step_1"]["value"]
    assert np.array_equal(got[:, 13].view(np.uint32), want[:, 13].view(np.uint32))
    assert np.array_equal(got[:, 12], want[:, 12])  # next decision times
    assert rel_err(got[:, 0:6], want[:, 0:6]) <= 1e-4
    # the reference's own GPU backends differ from its oracle by 4.8e-7 m
    # after one step at this start (logs/cuda-parity-1.log:1-5)
    assert abs_err(got[:, 0:6], want[:, 0:6]) <= 1e-5


@pytest.mark.parametrize("flags", [0, 64 | 4, 2048 | 64 | 4])
def test_twenty_steps_stay_near_the_fixture(fh, flags):
    start, _ = load(fh, "s2000_seed1.flk")
    want, tw = load(fh, "s2000_seed1_step20.flk")
    with fh.Flock.from_words(start, flags=flags) as f:
        f.step(20)
        got, t = f.download()
    assert np.float32(t) == np.float32(tw)
    drift = abs_err(got[:, 0:6], want[:, 0:6])
    print("20-step drift from the fixture (max abs pos/vel)", flags, drift)
    # the reference's 100-step parity at this start is 7.6e-6 m
    # (logs/cuda-parity-100.log:1-5); its pointwise bar is 1e-3 (docs/SPEC.md:70)
    assert drift <= 1e-4
    assert np.mean(got[:, 13].view(np.uint32) == want[:, 13].view(np.uint32)) >= 0.99
