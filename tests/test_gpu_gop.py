"""Hierarchical GOP 4 and GOP 8 decodes on the GPU against the reference's
decodeGOP (DVC_Decoder.cpp:26-474, recursion :471-472), where decoded WZ
frames become the references of the later frames of the GOP.  Fixtures:
tests/golden/frames_gop.json (ParallelMode 0: QCIF Q4 and CIF Q8, GOP 4 and
8) and frames_gop_mode1.json (ParallelMode 1, QCIF Q4 GOP 8), written by
tests/golden/make_golden.py from the reference's own sources.

Each GOP is decoded through the serving engine with each frame submitted as
soon as both of its references are decoded (dvc_amd.serve.decode_gop), and
once more frame by frame in the reference's order through the plain path
(dvc_decode_wz_frames, batched = 0)."""
import hashlib
import json
import struct
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = Path(__file__).resolve().parent / "golden"


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def _fbits(x):
    return struct.unpack("<I", struct.pack("<f", float(x)))[0]


def _case_inputs(case):
    from dvc_amd import workload as W
    vid = W.synthetic_video(case["W"], case["H"], case["gop"] + 1, case["shift"], case["seed"])
    wz = {k: W.encode_wz_frame(vid[k], case["qindex"]) for k in range(1, case["gop"])}
    return vid, wz


def _check_frame(gold, rec, st, where):
    assert st["levels"] == gold["levels"], where
    assert [_fbits(e) for e in st["entropy"]] == gold["entropy_bits"], where
    assert st["bitrate"] == gold["bitrate"], where
    assert st["psnr"] == gold["psnr"], where
    assert _sha(rec) == gold["recons_sha256"], where


def _cases(name):
    return json.loads((GOLD / name).read_text())["cases"]


@pytest.mark.parametrize("case", _cases("frames_gop.json"), ids=lambda c: f"{c['W']}x{c['H']}-q{c['qindex']}-gop{c['gop']}")
def test_gop_engine_matches_reference(ctx, case):
    from dvc_amd import _lib, serve
    vid, wz = _case_inputs(case)
    eng = _lib.Engine(ctx, max_inflight=8, lanes=2)
    try:
        out, order = serve.decode_gop(eng, vid, wz, 0, case["gop"], sid=case["seed"])
    finally:
        eng.close()
    assert [[c, p, f] for c, p, f in order] == [[g["cur"], g["past"], g["fut"]] for g in case["frames"]]
    for g in case["frames"]:
        rec, st = out[g["cur"]]
        _check_frame(g, rec, st, f"frame {g['cur']}")


def test_gop8_plain_path_matches_reference(ctx):
    """The plain drop-in (one frame at a time, sequential rate search) on the
    CIF Q8 GOP 8 case, frames in the reference's order."""
    case = [c for c in _cases("frames_gop.json") if c["W"] == 352 and c["gop"] == 8][0]
    vid, wz = _case_inputs(case)
    decoded = {0: vid[0], case["gop"]: vid[case["gop"]]}
    for g in case["frames"]:
        fr = dict(past=decoded[g["past"]], fut=decoded[g["fut"]], orig=vid[g["cur"]], wz=wz[g["cur"]],
                  idx=(g["past"], g["cur"], g["fut"]))
        (rec, st), = ctx.decode_wz_frames([fr], batched=False)
        _check_frame(g, rec, st, f"frame {g['cur']}")
        decoded[g["cur"]] = rec


@pytest.mark.parametrize("case", _cases("frames_gop_mode1.json"), ids=lambda c: f"{c['W']}x{c['H']}-q{c['qindex']}-gop{c['gop']}")
def test_gop_mode1_engine_matches_reference(case):
    """ParallelMode 1 semantics (beliefPropagation_OpenMP): every WZ frame
    starts with fresh carried decisions, as each decodeGOP call mallocs its
    own decodedBitplane (DVC_Decoder.cpp:124)."""
    from dvc_amd import _lib, serve
    vid, wz = _case_inputs(case)
    with _lib.Context(max_iter=50, parallel_mode=1) as c1:
        eng = _lib.Engine(c1, max_inflight=8, lanes=2)
        try:
            out, _ = serve.decode_gop(eng, vid, wz, 0, case["gop"], sid=case["seed"])
        finally:
            eng.close()
    for g in case["frames"]:
        rec, st = out[g["cur"]]
        _check_frame(g, rec, st, f"frame {g['cur']}")
