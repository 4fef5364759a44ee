"""dvc_amd/csrc/dvc_math.h (the device's float transcendentals) against the
host C library, compiled for the host: the full BP domain of phi and the
domains of logf (positive floats), expf and tanhf on strides."""
import os
import subprocess
import tempfile
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


@pytest.fixture(scope="module")
def checker():
    d = tempfile.mkdtemp()
    exe = os.path.join(d, "math_check")
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fopenmp", "-o", exe,
                    str(ROOT / "tests" / "native" / "math_check.c"), "-lm"], check=True)
    return exe


@pytest.mark.parametrize("func,first,last,stride", [
    ("phi", "0", "7f800000", "1"),          # every non-negative finite float
    ("phi_nonneg", "0", "80000000", "1"),   # the branch-free form, same domain plus -0
    ("phi_fast", "0", "80000000", "1"),     # its shortened form (the BP kernel's; IEEE quotients on the host)
    ("logf", "1", "7f7fffff", "7"),
    ("expf", "0", "ffffffff", "13"),
    ("tanhf", "0", "ffffffff", "13"),
    ("expm1f", "0", "ffffffff", "17"),
])
def test_matches_host_libm(checker, func, first, last, stride):
    r = subprocess.run([checker, func, first, last, stride], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "mismatches=0" in r.stdout
