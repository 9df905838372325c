"""C-ABI checks that need no GPU: the library loads, exports every symbol the
header declares, its constants, generators and .flk I/O agree with the oracle,
and device calls fail loudly (FH_ERR_NODEVICE) when no GPU is present."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", f) for f in sorted(os.listdir(os.path.join(ROOT, "include")))
           if f.endswith(".h")]


def declared_symbols():
    names = []
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(fh_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_pass_api():
    names = declared_symbols()
    for n in ["fh_create", "fh_hash_count", "fh_scan", "fh_scatter", "fh_decide", "fh_integrate",
              "fh_integrate_hash", "fh_step", "fh_upload_words", "fh_download_words", "fh_read_grid"]:
        assert n in names


def test_library_exports_every_declared_symbol(fh):
    lib = ctypes.CDLL(fh.LIB_PATH)
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_constants_equal_the_oracle(fh, fr):
    a, b = fh.constants(), fr.constants()
    assert a.view(np.uint32).tolist() == b.view(np.uint32).tolist()


def test_buckets_for_equals_the_oracle(fh, fr):
    for n in [0, 1, 2, 1000, 32768, 32769, 65536, 10**5, 10**6, 2**21, 10**7, 2**31]:
        assert fh.buckets_for(n) == fr.buckets_for(n), n


@pytest.mark.parametrize("n,seed", [(1, 1), (50, 7), (2000, 1), (100_000, 3)])
def test_reference_generator_is_bit_identical(fh, fr, n, seed):
    assert np.array_equal(fh.gen_reference(n, seed).view(np.uint32), fr.scatter(n, seed).view(np.uint32))


def test_flk_round_trip_across_both_sides(fh, fr, tmp_path):
    w = fh.gen_cube(777, 40.0, 9)
    w[:, 13] = np.arange(777, dtype=np.uint32).view(np.float32)
    fh.write_flk(tmp_path / "a.flk", w, 0.125)
    back, t, rc = fr.read_flk(tmp_path / "a.flk")
    assert rc == 0 and t == np.float32(0.125) and np.array_equal(back.view(np.uint32), w.view(np.uint32))
    fr.write_flk(tmp_path / "b.flk", w, 2.5)
    back2, t2 = fh.read_flk(tmp_path / "b.flk")
    assert t2 == np.float32(2.5) and np.array_equal(back2.view(np.uint32), w.view(np.uint32))


def test_flk_errors_are_reported(fh, tmp_path):
    p = tmp_path / "bad.flk"
    p.write_bytes(b"FLK2" + bytes(12))
    with pytest.raises(fh.FlockHipError) as e:
        fh.read_flk(p)
    assert e.value.status == fh.ERR_FLK
    p.write_bytes(b"FLK1" + np.array([1, 14, 0], np.uint32).tobytes())
    with pytest.raises(fh.FlockHipError) as e:
        fh.read_flk(p)
    assert e.value.status == fh.ERR_FLK and "words" in str(e.value)
    p.write_bytes(b"FLK1" + np.array([3, 15, 0], np.uint32).tobytes() + bytes(60))
    with pytest.raises(fh.FlockHipError) as e:
        fh.read_flk(p)
    assert e.value.status == fh.ERR_FLK and "truncated" in str(e.value)
    with pytest.raises(fh.FlockHipError) as e:
        fh.read_flk(tmp_path / "missing.flk")
    assert e.value.status == fh.ERR_IO


def test_cube_generator_properties(fh):
    n, side = 20_000, 100.0
    w = fh.gen_cube(n, side, 1)
    assert np.array_equal(w.view(np.uint32), fh.gen_cube(n, side, 1).view(np.uint32))
    assert np.all(np.abs(w[:, 0]) <= side / 2) and np.all(np.abs(w[:, 1]) <= side / 2)
    assert np.all(np.abs(w[:, 2] - 100.0) <= side / 2)
    sp = np.linalg.norm(w[:, 3:6].astype(np.float64), axis=1)
    assert np.all((sp >= 9.0 - 1e-4) & (sp <= 11.0 + 1e-4))
    d = w[:, 3:6] / sp[:, None]
    assert np.all(np.abs(d.mean(axis=0)) < 0.03)  # isotropic
    assert np.all(np.abs((d ** 2).mean(axis=0) - 1 / 3) < 0.02)
    assert np.all((w[:, 12] >= 0) & (w[:, 12] < 0.05)) and np.all(w[:, 8] == np.float32(1.5))


def test_clump_generator_properties(fh):
    n = 40_000
    w = fh.gen_clumps(n, 8, 5.0, 1000.0, 1)
    sp = np.linalg.norm(w[:, 3:6].astype(np.float64), axis=1)
    assert np.all(np.abs(sp - 10.0) < 1e-4)
    # 8 clumps: every bird lies within 8 sigma of some clump centre
    from scipy.cluster.vq import kmeans2
    cent, lab = kmeans2(w[:, :3].astype(np.float64), 8, seed=1, minit="++")
    spread = np.linalg.norm(w[:, :3] - cent[lab], axis=1)
    assert np.median(spread) < 5.0 * 2.0
    for c in range(8):
        v = w[lab == c, 3:6] / 10.0
        if len(v) < 10:
            continue
        h = v.mean(axis=0)
        h /= np.linalg.norm(h)
        ang = np.degrees(np.arccos(np.clip(v @ h, -1, 1)))
        assert np.max(ang) < 20.5  # each direction within 10 deg of the clump heading


def test_cube_side_for_matches_the_configs(fh):
    assert abs(fh.cube_side_for(10_000) - 100.0) < 1e-9
    assert abs(fh.cube_side_for(1_000_000) - 464.1588833612779) < 1e-6


def test_device_calls_fail_loudly_without_a_gpu(fh):
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    with pytest.raises(fh.FlockHipError) as e:
        fh.Flock(100)
    assert e.value.status == fh.ERR_NODEVICE


@pytest.mark.parametrize("kw", [dict(decide_group=32), dict(decide_group=3), dict(decide_block=100),
                                dict(birds_per_thread=65), dict(bucket_shift=9), dict(buckets=1000),
                                dict(flags=4096 | 64), dict(flags=4096 | 2048), dict(flags=4096 | 2048 | 64 | 1024),
                                dict(flags=4096 | 2048 | 64 | 2)])
def test_bad_params_are_rejected_before_the_device(fh, kw):
    # Parameter validation runs before any device call, so it reports
    # FH_ERR_INVALID on any machine.
    with pytest.raises(fh.FlockHipError) as e:
        fh.Flock(100, **kw)
    assert e.value.status == fh.ERR_INVALID, kw


def test_spatial_keys_known_answers(fh):
    # Cells (0,0,0), (1,0,0), (0,1,0), (0,0,1), (1,1,1), (2,0,0), (1023,0,0) at
    # cell 4 m: Morton codes 0, 1, 2, 4, 7, 8, 0x9249249 (block hash 0); cell
    # (1024,0,0) lies in block (1,0,0): Morton 0 + 73856093.
    cells = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 1], [2, 0, 0], [1023, 0, 0],
                      [1024, 0, 0]], np.float32)
    w = np.zeros((len(cells), fh.WORDS), np.float32)
    w[:, 0:3] = cells * 4.0 + 0.5
    keys = fh.spatial_keys(w, 1 << 30)
    assert keys.tolist() == [0, 1, 2, 4, 7, 8, 0x9249249, 73856093]
