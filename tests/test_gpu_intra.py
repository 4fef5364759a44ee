"""Intra blocks and the deblocking filter on the GPU against the reference
(tests/golden/intra_streams.json, from tests/golden/make_golden.py): whole
sequences coded by the reference encoder's main loop with skip blocks, intra
blocks (H.264 intra 4x4 DC prediction, CAVLC) and deblocking as each case
sets them, kept as the WZ and intra bitstreams, and decoded by the
reference's own main path (readWZBitstream, readIntraBitstream, decodeGOP).
Here the library parses both bitstreams (dvc_wz_stream_*, with its JBIG and
CAVLC decoders) and decodes every GOP through the serving engine and through
the plain path; codes, entropy bits, bitrate, PSNR and reconstruction must
equal the reference's."""
import base64
import hashlib
import json
import struct
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"
CASES = json.loads((GOLD / "intra_streams.json").read_text())["cases"]
MODE1 = {c["name"]: c for c in json.loads((GOLD / "intra_streams_mode1.json").read_text())["cases"]}


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def _fbits(x):
    return struct.unpack("<I", struct.pack("<f", float(x)))[0]


def _video(c):
    """The case's synthetic video with its noise box (the originals, which are
    also the key frames the decoder receives)."""
    from dvc_amd import workload as W
    vid = W.synthetic_video(c["W"], c["H"], c["frames"], c["shift"], c["seed"]).astype(np.int32)
    if c["noise_box"] is not None:
        rng = np.random.default_rng(c["seed"] + 99)
        y0, y1, x0, x1, amp = c["noise_box"]
        for k in range(c["frames"]):
            vid[k, y0:y1, x0:x1] += rng.integers(-amp, amp + 1, (y1 - y0, x1 - x0))
    return np.clip(vid, 0, 255).astype(np.uint8)


def _check(g, rec, st, where):
    assert st["levels"] == g["levels"], where
    assert [_fbits(e) for e in st["entropy"]] == g["entropy_bits"], where
    assert st["bitrate"] == g["bitrate"], where
    assert st["psnr"] == g["psnr"], where
    assert _sha(rec) == g["recons_sha256"], where


def _decode_and_check(ctx, case, decoded_gold):
    from dvc_amd import _lib, serve
    vid = _video(case)
    info, frames = _lib.read_wz_stream(base64.b64decode(case["wz_stream"]), base64.b64decode(case["intra_stream"]))
    parsed = {f["idx"][1]: f["wz"] for f in frames}
    gold = {g["cur"]: g for g in decoded_gold}
    print(f"case {case['name']}: {len(frames)} WZ frames", flush=True)
    # the serving engine, GOP by GOP, frames submitted as their references are decoded
    eng = _lib.Engine(ctx, max_inflight=8, lanes=2)
    try:
        for i in range(0, case["frames"], case["gop"]):
            end = min(i + case["gop"], case["frames"] - 1)
            out, order = serve.decode_gop(eng, vid, parsed, i, end, sid=case["seed"] + i)
            for cur, past, fut in order:
                assert (past, fut) == (gold[cur]["past"], gold[cur]["fut"])
                rec, st = out[cur]
                _check(gold[cur], rec, st, f"engine frame {cur}")
    finally:
        eng.close()
    # the plain path, frame by frame in decoding order, decoded frames as references
    decoded = {k: vid[k] for k in range(case["frames"]) if k not in gold}
    for f in frames:
        past, cur, fut = f["idx"]
        fr = dict(past=decoded[past], fut=decoded[fut], orig=vid[cur], wz=f["wz"], idx=f["idx"])
        (rec, st), = ctx.decode_wz_frames([fr], batched=False)
        _check(gold[cur], rec, st, f"plain frame {cur}")
        decoded[cur] = rec


@pytest.mark.parametrize("case", CASES, ids=lambda c: c["name"])
def test_stream_matches_reference(ctx, case):
    _decode_and_check(ctx, case, case["decoded"])


@pytest.mark.parametrize("case", CASES, ids=lambda c: c["name"])
def test_stream_mode1_matches_reference(case):
    """The same streams with ParallelMode 1 BP semantics (the reference run
    with MALLOC_PERTURB_=85, as the other ParallelMode 1 fixtures)."""
    from dvc_amd import _lib
    with _lib.Context(max_iter=50, parallel_mode=1) as c1:
        _decode_and_check(c1, case, MODE1[case["name"]]["decoded"])
