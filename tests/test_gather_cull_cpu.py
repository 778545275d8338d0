"""The gather's cell and octant culling never skips a cell or a half-cell
octant holding a photon that the distance test accepts: random grids up to 4096³ with origins anywhere in
±1024, photons at cell edges, hitpoints at distance ≈ r
(tests/cpp/gather_cull_check.cpp).  The same trials with a fixed margin of
1e-4 cell must find violations, so the trials reach the cases the margin is
sized for."""
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
SRC = ROOT / "tests" / "cpp" / "gather_cull_check.cpp"


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    out = tmp_path_factory.mktemp("cull") / "gather_cull_check"
    cc = subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-pthread", str(SRC), "-o", str(out)],
                        capture_output=True, text=True)
    if cc.returncode != 0:
        pytest.fail(cc.stderr)
    return out


def test_cell_culling_is_conservative(exe):
    res = subprocess.run([str(exe), "20000000"], capture_output=True, text=True, timeout=600)
    print(res.stdout, res.stderr)
    assert res.returncode == 0 and res.stdout.startswith("ok"), res.stdout + res.stderr


def test_fixed_margin_is_caught(exe):
    res = subprocess.run([str(exe), "20000000", "fixed"], capture_output=True, text=True, timeout=600)
    print(res.stdout, res.stderr)
    assert res.returncode == 0 and res.stdout.startswith("fixed"), res.stdout + res.stderr
