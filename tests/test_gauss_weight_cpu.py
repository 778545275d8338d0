"""The gather's Gaussian weight without IEEE divisions (phx_div_recip) equals
the literal expression bit for bit: exhaustively for the constant
denominator, on random radius/distance pairs for the per-pixel one, and
through exp = 1 where the quotient is too small for the fma steps
(tests/cpp/gauss_weight_check.cpp)."""
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
SRC = ROOT / "tests" / "cpp" / "gauss_weight_check.cpp"


def test_gauss_weight_matches_literal(tmp_path):
    exe = tmp_path / "gauss_weight_check"
    cc = subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-pthread", str(SRC), "-o", str(exe)],
                        capture_output=True, text=True)
    if cc.returncode != 0:
        pytest.fail(cc.stderr)
    res = subprocess.run([str(exe), "20000000"], capture_output=True, text=True, timeout=600)
    print(res.stdout, res.stderr)
    assert res.returncode == 0 and res.stdout.startswith("ok"), res.stdout + res.stderr
