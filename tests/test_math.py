"""dvc_amd/csrc/dvc_math.h (the device's transcendentals) against the host C
library, compiled for the host: the full BP domain of phi, the domains of
logf (positive floats), expf and tanhf on strides, and double exp on
pseudo-random arguments (with reconstruction's own argument forms); and the
exp table against its definition."""
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
    ("exp", "400000000", "5eed", "0"),      # double exp: 400M pseudo-random arguments
])
def test_matches_host_libm(checker, func, first, last, stride):
    r = subprocess.run([checker, func, first, last, stride], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "mismatches=0" in r.stdout


def test_exp_table_matches_its_definition():
    """dvcm_exp_tab[2k], [2k+1] = bits of T, bits of H - (k << 45), where H is
    the double nearest 2^(k/128) and T the double nearest (2^(k/128) - H)/H,
    recomputed here with 80-digit decimal arithmetic."""
    import re
    import struct
    from decimal import Decimal, getcontext
    getcontext().prec = 80
    text = (ROOT / "dvc_amd" / "csrc" / "dvc_math.h").read_text()
    body = text[text.index("dvcm_exp_tab[256] = {"):]
    vals = [int(v, 16) for v in re.findall(r"0x([0-9a-f]{16})ull", body[:body.index("};")])]
    assert len(vals) == 256
    for k in range(128):
        v = Decimal(2) ** (Decimal(k) / Decimal(128))
        h = float(v)
        t = float((v - Decimal(h)) / Decimal(h))
        bits = lambda d: struct.unpack("<Q", struct.pack("<d", d))[0]
        assert vals[2 * k] == bits(t), k
        assert vals[2 * k + 1] == (bits(h) - (k << 45)) % (1 << 64), k
