"""Host-side logic without a GPU: the serving loop's decode order and stream
set, and bench.py's one-GPU contract."""
import subprocess
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))


def decode_gop_order(begin, end):
    """decodeGOP (DVC_Decoder.cpp:26-29, 471-472) restated as a list of
    (cur, past, fut) in the order the reference decodes them."""
    out = []

    def rec(b, e):
        mid = (b + e) // 2
        if mid > b:
            out.append((mid, b, e))
            rec(b, mid)
            rec(mid, e)
    rec(begin, end)
    return out


def test_gop_order_matches_decode_gop():
    from dvc_amd import serve
    for gop in (2, 4, 8):
        got = []
        serve.gop_order(0, gop, got)
        assert got == decode_gop_order(0, gop)
        assert sorted(c for c, _, _ in got) == list(range(1, gop))
        # every reference is a key frame or decoded earlier in the list
        done = {0, gop}
        for cur, past, fut in got:
            assert past in done and fut in done
            done.add(cur)


def test_serving_streams():
    from dvc_amd import serve
    streams = serve.make_streams(n_qcif=3, n_cif=1, frames=9, seed=7)
    assert [s.width for s in streams] == [176, 176, 176, 352]
    for s in streams:
        assert s.gop in (2, 4, 8) and 1 <= s.qindex <= 8
        assert set(s.wz) == {c for c, _, _ in s.order}
        assert all(f % s.gop == 0 or f == 8 for f in s.decoded)
        assert s.video.shape == (9, s.height, s.width) and s.video.dtype == np.uint8


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
