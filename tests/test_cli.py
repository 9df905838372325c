"""The drop-in command line (`flock_hip`, in the form of the reference's step
binary, cuda/flock.cu:374-523) and the oracle's `flock-ref`, driven the way
scripts/parity.sh:23-34 drives the reference: generate a start state, step it
with both binaries, compare with `flock-ref diff` / `statdiff`. Usage and file
errors end with the reference's exit codes (2 for usage, flock.cu:386-388;
1 for a bad .flk, flock.cu:344-351) before any device call."""
import os
import struct
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "murmuration-kernels_amd", "_build", "flock_hip")
REF = os.path.join(ROOT, "oracle", "_build", "flock-ref")
THREADS = str(min(16, os.cpu_count() or 1))


def run(*args, timeout=300):
    for exe in (CLI, REF):
        if args[0] == exe and not os.path.exists(exe):
            raise FileNotFoundError(f"{exe} is not built (run __graft_entry__.build())")
    return subprocess.run(list(args), capture_output=True, text=True, timeout=timeout)


@pytest.mark.parametrize("args", [["--bogus"], [], ["--in"], ["--in", "x.flk", "--steps", "-1"]])
def test_usage_errors_exit_2(args):
    r = run(CLI, *args)
    assert r.returncode == 2, (args, r.stdout, r.stderr)


def test_bad_files_exit_1(tmp_path):
    missing = tmp_path / "missing.flk"
    bad_magic = tmp_path / "magic.flk"
    bad_magic.write_bytes(struct.pack("<4I", 0x12345678, 1, 15, 0) + bytes(60))
    short = tmp_path / "short.flk"
    short.write_bytes(struct.pack("<IIIf", 0x314B4C46, 10, 15, 0.0) + bytes(60 * 3))
    words = tmp_path / "words.flk"
    words.write_bytes(struct.pack("<IIIf", 0x314B4C46, 1, 14, 0.0) + bytes(56))
    for path in (missing, bad_magic, short, words):
        r = run(CLI, "--in", str(path), "--steps", "1")
        assert r.returncode == 1, (path.name, r.stdout, r.stderr)


@pytest.mark.gpu
def test_parity_ladder_through_the_command_lines(tmp_path):
    # parity.sh's ladder at 2000 birds: pointwise 1e-3 up to 30 steps, at most
    # 2 outliers under 1e-2 at 100, aggregates at 1000.
    start = tmp_path / "p2000.flk"
    r = run(REF, "gen", "--birds", "2000", "--seed", "1", "--out", str(start))
    assert r.returncode == 0, r.stderr
    for k in (1, 10, 30, 100, 1000):
        gpu, cpu = tmp_path / f"gpu-{k}.flk", tmp_path / f"cpu-{k}.flk"
        r = run(CLI, "--in", str(start), "--steps", str(k), "--warmup", "0", "--out", str(gpu))
        assert r.returncode == 0, (k, r.stderr)
        for row in ("grid_histogram", "grid_scan", "grid_scatter", "decide", "integrate", "step total",
                    "bird-steps/s"):
            assert row in r.stdout, (k, row, r.stdout)
        r = run(REF, "step", "--in", str(start), "--steps", str(k), "--out", str(cpu), "--threads", THREADS)
        assert r.returncode == 0, (k, r.stderr)
        if k <= 30:
            r = run(REF, "diff", str(cpu), str(gpu))
        elif k <= 100:
            r = run(REF, "diff", str(cpu), str(gpu), "--allow-outliers", "2", "--outlier-max", "1e-2")
        else:
            r = run(REF, "statdiff", str(cpu), str(gpu))
        print(k, r.stdout.strip().splitlines()[-1:] if r.stdout else r.stderr)
        assert r.returncode == 0, (k, r.stdout, r.stderr)
