"""Skip blocks on the GPU against the reference (tests/golden/frames_skip.json,
wz_stream.json, from tests/golden/make_golden.py): the reference encoder with
skip blocks on (findSkipBlock, the map's JBIG data), readGOP's map decoding
and decodeGOP with the map (DVC_Decoder.cpp:131, 155-157, 203-207 and the
BlockModeFlag checks of noiseParameter, initSIR, putBitplaneToCoeff,
reconstruction).  The map comes from the library's own JBIG decoder; the
frames are decoded through the plain path and the batched engine, and a WZ
bitstream with skip frames in GOPs of 4 is parsed by the library's reader and
decoded GOP by GOP."""
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


def _check(gold, rec, st, where):
    assert st["levels"] == gold["levels"], where
    assert [_fbits(e) for e in st["entropy"]] == gold["entropy_bits"], where
    assert st["bitrate"] == gold["bitrate"], where
    assert st["psnr"] == gold["psnr"], where
    assert _sha(rec) == gold["recons_sha256"], where


def _skip_frame(case):
    from dvc_amd import _lib, workload as W
    vid = W.synthetic_video(case["W"], case["H"], case["frames"], case["shift"], case["seed"])
    m, used = _lib.jbig_decode_map(bytes.fromhex(case["skip_jbig"]) + bytes(32), case["W"] // 8, case["H"] // 8)
    mode = (m.reshape(-1) * W.SKIP_BLOCK).astype(np.uint8)
    assert _sha(mode) == case["block_mode_sha256"]
    wz = W.encode_wz_frame(vid[case["wz"]], case["qindex"], block_mode=mode, mode_bits=(used << 3) + 1 + 1)
    return vid, wz


@pytest.mark.parametrize("case", json.loads((GOLD / "frames_skip.json").read_text())["cases"],
                         ids=lambda c: f"{c['W']}x{c['H']}-q{c['qindex']}")
def test_skip_frame_matches_reference(ctx, case):
    vid, wz = _skip_frame(case)
    fr = dict(past=vid[0], fut=vid[2], orig=vid[case["wz"]], wz=wz)
    for batched in (False, True):
        (rec, st), = ctx.decode_wz_frames([fr], batched=batched)
        _check(case, rec, st, f"batched={batched}")
        # skip blocks are the past key frame's pixels
        m8 = wz.block_mode.reshape(case["H"] // 8, case["W"] // 8) == 2
        m = np.repeat(np.repeat(m8, 8, 0), 8, 1)
        assert np.array_equal(rec[m], vid[0][m])


def test_skip_frames_batched_with_plain_frames(ctx):
    """Frames with and without skip maps in one batch: each as decoded alone."""
    from dvc_amd import workload as W
    cases = json.loads((GOLD / "frames_skip.json").read_text())["cases"]
    frames, gold = [], []
    for c in cases:
        vid, wz = _skip_frame(c)
        frames.append(dict(past=vid[0], fut=vid[2], orig=vid[c["wz"]], wz=wz))
        gold.append(c)
        plain = W.encode_wz_frame(vid[c["wz"]], c["qindex"])
        frames.append(dict(past=vid[0], fut=vid[2], orig=vid[c["wz"]], wz=plain))
        gold.append(None)
    out = ctx.decode_wz_frames(frames, batched=True)
    solo = [ctx.decode_wz_frames([f], batched=False)[0] for f in frames]
    for k, ((rec, st), (rs, ss), g) in enumerate(zip(out, solo, gold)):
        assert np.array_equal(rec, rs) and st["levels"] == ss["levels"] and st["bitrate"] == ss["bitrate"], k
        if g is not None:
            _check(g, rec, st, f"frame {k}")


def test_wz_stream_gops_match_reference(ctx):
    """The WZ bitstream parsed by dvc_wz_stream_* and decoded GOP by GOP
    through the engine (each frame submitted once its references are
    decoded)."""
    from dvc_amd import _lib, serve, workload as W
    d = json.loads((GOLD / "wz_stream.json").read_text())
    c = d["case"]
    vid = W.synthetic_video(c["W"], c["H"], c["frames"], c["shift"], c["seed"])
    jbig = {int(k): bytes.fromhex(v) for k, v in d["skip_jbig"].items()}
    wz = {}
    for g in d["decoded"]:
        mode = W.skip_candidates(vid[g["cur"]], vid[g["past"]]) if g["cur"] in jbig else None
        wz[g["cur"]] = W.encode_wz_frame(vid[g["cur"]], c["qindex"], block_mode=mode)
    stream = W.write_wz_bitstream(c["W"], c["H"], c["qindex"], c["gop"], c["frames"], wz, skip_jbig=jbig)
    info, frames = _lib.read_wz_stream(stream)
    parsed = {f["idx"][1]: f["wz"] for f in frames}
    gold = {g["cur"]: g for g in d["decoded"]}
    eng = _lib.Engine(ctx, max_inflight=8, lanes=2)
    try:
        for i in range(0, c["frames"], c["gop"]):
            end = min(i + c["gop"], c["frames"] - 1)
            out, order = serve.decode_gop(eng, vid, parsed, i, end, sid=c["seed"] + i)
            for cur, past, fut in order:
                g = gold[cur]
                assert (past, fut) == (g["past"], g["fut"])
                rec, st = out[cur]
                _check(g, rec, st, f"frame {cur}")
    finally:
        eng.close()


def test_skip_map_rejects_intra_blocks(ctx):
    from dvc_amd import _lib, workload as W
    vid = W.synthetic_video(176, 144, 3, 0, 5)
    mode = np.zeros(22 * 18, np.uint8)
    mode[7] = 1  # an intra block
    wz = W.encode_wz_frame(vid[1], 4, block_mode=mode, mode_bits=2)
    with pytest.raises(_lib.DvcError):
        ctx.decode_wz_frames([dict(past=vid[0], fut=vid[2], orig=vid[1], wz=wz)], batched=False)
