"""The flight's divisions by model constants (fh::div_model_const, model.h):
a product by the rounded reciprocal and one fused remainder correction,
claimed bit-identical to IEEE division for every float with
2^-104 <= |x| <= 2^100 and each of MASS, G, TAU_ROLL and V0 (other values
divide). The host build of model.h is checked here on every 997th float of
that range and on the values that take the division; `div_check 1` runs the
whole range (about 3.4e9 floats per constant, a few minutes)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.fixture(scope="module")
def div_check(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc is not installed")
    exe = str(tmp_path_factory.mktemp("div") / "div_check")
    subprocess.run([HIPCC, "-O2", "-std=c++17", "-ffp-contract=off", "-o", exe, os.path.join(HERE, "div_check.cpp")],
                   check=True, capture_output=True)
    return exe


def run(exe, *args):
    p = subprocess.run([exe, *args], capture_output=True, text=True)
    rows = {name: (int(bad), int(n)) for name, bad, n in (line.split() for line in p.stdout.split("\n") if line)}
    return p.returncode, rows


def test_constant_division_is_the_division(div_check):
    rc, rows = run(div_check, "997")
    assert set(rows) == {"MASS", "G", "TAU_ROLL", "V0"}
    for name, (bad, n) in rows.items():
        assert n > 3_000_000 and bad == 0, (name, bad, n)
    assert rc == 0


def test_check_catches_a_missing_correction(div_check):
    # the product alone is off by an ulp for a share of the floats
    rc, rows = run(div_check, "997", "plain")
    assert rc == 1 and all(bad > 1000 for bad, _ in rows.values()), rows
