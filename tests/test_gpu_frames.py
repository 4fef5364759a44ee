"""Whole WZ-frame decoding on the GPU (dvc_decode_wz_frames) against the
reference's own decodes (golden fixtures from decodeGOP) and the oracle:
reconstruction, accepted code per bitplane, entropy bits, bitrate and PSNR
must be identical, for the plain and the batched mode, and a frame decoded in
a mixed batch equals the same frame decoded alone."""
import hashlib
import json
import struct
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as O
from dvc_amd import workload as W

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def fbits(x):
    return struct.unpack("<I", struct.pack("<f", float(x)))[0]


def _case_inputs(c):
    vid = W.synthetic_video(c["W"], c["H"], c["frames"], c["shift"], c["seed"])
    wz = W.encode_wz_frame(vid[c["wz"]], c["qindex"])
    return dict(past=vid[0], fut=vid[2], orig=vid[c["wz"]], wz=wz)


def _check_against_golden(rec, st, c):
    print(c["W"], c["qindex"], "levels", st["levels"][:12], "golden", c["levels"][:12], flush=True)
    assert st["levels"] == c["levels"]
    assert [fbits(e) for e in st["entropy"]] == c["entropy_bits"]
    assert st["bitrate"] == c["bitrate"]
    assert sha(rec) == c["recons_sha256"]
    assert st["psnr"] == c["psnr"]


@pytest.mark.parametrize("case", range(4))
@pytest.mark.parametrize("batched", [False, True])
def test_frame_matches_reference_golden(ctx, case, batched):
    c = json.loads((GOLD / "frames.json").read_text())["cases"][case]
    (rec, st), = ctx.decode_wz_frames([_case_inputs(c)], batched=batched)
    _check_against_golden(rec, st, c)


@pytest.mark.parametrize("case", range(5))
@pytest.mark.parametrize("batched", [False, True])
def test_cif_qindex_matches_reference_golden(ctx, case, batched):
    """CIF at Qindex 1, 3, 5, 6 and 7 (tests/golden/frames_cif_q.json, the
    reference's decodeGOP): the quantiser tables, band plans and bitplane
    counts frames.json does not reach."""
    c = json.loads((GOLD / "frames_cif_q.json").read_text())["cases"][case]
    (rec, st), = ctx.decode_wz_frames([_case_inputs(c)], batched=batched)
    _check_against_golden(rec, st, c)


def test_mixed_batch_equals_solo(ctx):
    cases = json.loads((GOLD / "frames.json").read_text())["cases"]
    inputs = [_case_inputs(c) for c in cases]
    outs = ctx.decode_wz_frames(inputs, batched=True)
    for (rec, st), c in zip(outs, cases):
        _check_against_golden(rec, st, c)


def test_frames_against_oracle_fresh_seed(ctx):
    vid = W.synthetic_video(176, 144, 3, 3, seed=777)
    for q in (1, 6):
        wz = W.encode_wz_frame(vid[1], q)
        orec, oinfo = O.decode_wz_frame(vid[0], vid[2], vid[1], wz, q)
        (rec, st), = ctx.decode_wz_frames([dict(past=vid[0], fut=vid[2], orig=vid[1], wz=wz)], batched=True)
        assert np.array_equal(rec, orec)
        assert st["levels"] == oinfo["levels"] and st["bitrate"] == oinfo["bitrate"]
        assert st["iterations"] == oinfo["iters"]


@pytest.mark.parametrize("size", [(176, 144, 3), (352, 288, 6)])
def test_frames_against_oracle_max_iter_100(size):
    """The measured configuration caps BP at 100 iterations per code (the
    reference's compiled cap is 50): GPU plain and batched decodes equal the
    oracle run with the same cap, at Qindex 8."""
    from dvc_amd import _lib
    w, h, shift = size
    vid = W.synthetic_video(w, h, 3, shift, seed=1100)
    wz = W.encode_wz_frame(vid[1], 8)
    orec, oinfo = O.decode_wz_frame(vid[0], vid[2], vid[1], wz, 8, max_iter=100)
    c100 = _lib.Context(max_iter=100)
    try:
        for batched in (False, True):
            (rec, st), = c100.decode_wz_frames([dict(past=vid[0], fut=vid[2], orig=vid[1], wz=wz)], batched=batched)
            print(size, batched, "levels", st["levels"][:8], flush=True)
            assert np.array_equal(rec, orec)
            assert st["levels"] == oinfo["levels"] and st["bitrate"] == oinfo["bitrate"]
            assert st["iterations"] == oinfo["iters"]
            assert st["psnr"] == oinfo["psnr"]
    finally:
        c100.close()


@pytest.mark.parametrize("lanes", [1, 3])
def test_engine_staggered_arrivals_equal_solo(ctx, lanes):
    """Continuous batching: frames submitted while others are mid-decode
    (different sizes, quantisers and arrival steps) give the same results as
    when decoded alone, with one lane or with lanes stepping concurrently."""
    from dvc_amd import _lib
    cases = json.loads((GOLD / "frames.json").read_text())["cases"]
    inputs = [_case_inputs(c) for c in cases]
    eng = _lib.Engine(ctx, max_inflight=3, lanes=lanes)
    tickets = {}
    step = 0
    pending = list(range(len(inputs)))
    arrive = {0: 0, 1: 3, 2: 7, 3: 20}
    results = {}
    while pending or tickets:
        for i in list(pending):
            if arrive[i] <= step:
                tickets[eng.submit(inputs[i], past_key=(i + 1) << 32, fut_key=((i + 1) << 32) | 2)] = i
                pending.remove(i)
        eng.step()
        step += 1
        for t in list(tickets):
            if eng.done(t):
                results[tickets.pop(t)] = eng.result(t)
        assert step < 2000
    eng.close()
    for i, c in enumerate(cases):
        rec, st = results[i]
        _check_against_golden(rec, st, c)


@pytest.mark.parametrize("lane_steps", [0, 1 << 20])
def test_engine_lanes_mixed_batch(ctx, lane_steps):
    """Frames spread over several lanes (QCIF and CIF, Qindex 1..8, shared key
    frames) decode to the same output as each frame alone on the plain path,
    with the lanes in lockstep or stepping independently."""
    from dvc_amd import _lib
    frames = []
    for k, (w, h, sh) in enumerate([(176, 144, 3), (352, 288, 6)]):
        vid = W.synthetic_video(w, h, 5, sh, seed=900 + k)
        for q in (1, 4, 8):
            for cur in (1, 3):
                frames.append(dict(past=vid[cur - 1], fut=vid[cur + 1], orig=vid[cur],
                                   wz=W.encode_wz_frame(vid[cur], q), idx=(cur - 1, cur, cur + 1)))
    eng = _lib.Engine(ctx, max_inflight=len(frames), lanes=4, lane_steps=lane_steps)
    outs = eng.decode_all(frames)
    eng.close()
    for fr, (rec, st) in zip(frames, outs):
        (rec1, st1), = ctx.decode_wz_frames([fr], batched=False)
        assert np.array_equal(rec, rec1)
        assert st["levels"] == st1["levels"] and st["bitrate"] == st1["bitrate"] and st["psnr"] == st1["psnr"]


@pytest.mark.parametrize("lanes", [1, 2])
def test_engine_graph_replay_equals_eager(ctx, lanes):
    """Steps in which no frame starts or ends a band replay a captured graph
    (dozens of replays per frame at Qindex 8, with eager band starts and ends
    and synchronisations between them); every output equals the eager run."""
    from dvc_amd import _lib
    frames = []
    for k, (w, h, sh, q) in enumerate([(352, 288, 6, 8), (176, 144, 3, 8), (352, 288, 6, 4), (176, 144, 3, 6)]):
        vid = W.synthetic_video(w, h, 3, sh, seed=950 + k)
        frames.append(dict(past=vid[0], fut=vid[2], orig=vid[1], wz=W.encode_wz_frame(vid[1], q)))
    runs = []
    for eager in (False, True):
        eng = _lib.Engine(ctx, max_inflight=len(frames), lanes=lanes, eager_steps=eager)
        runs.append(eng.decode_all(frames))
        eng.close()
    for (rg, sg), (re_, se) in zip(*runs):
        assert np.array_equal(rg, re_)
        assert sg["levels"] == se["levels"] and sg["iterations"] == se["iterations"]
        assert sg["bitrate"] == se["bitrate"] and sg["psnr"] == se["psnr"]


def test_engine_reused_across_batches(ctx):
    """A long-lived engine (device blocks of retired frames and key frames
    pooled and reused, captured graphs kept) decodes later batches, of other
    sizes and quantisers, to the same output as solo decodes."""
    from dvc_amd import _lib
    batches = []
    for k, (w, h, sh, qs) in enumerate([(352, 288, 6, (8, 3)), (176, 144, 3, (8, 5, 2)), (352, 288, 6, (8, 1))]):
        vid = W.synthetic_video(w, h, 5, sh, seed=970 + k)
        batches.append([dict(past=vid[2 * i], fut=vid[2 * i + 2], orig=vid[2 * i + 1],
                             wz=W.encode_wz_frame(vid[2 * i + 1], q), idx=(2 * i, 2 * i + 1, 2 * i + 2))
                        for i, q in zip(range(2), qs)] + [dict(past=vid[1], fut=vid[3], orig=vid[2],
                                                                wz=W.encode_wz_frame(vid[2], qs[-1]), idx=(1, 2, 3))])
    eng = _lib.Engine(ctx, max_inflight=4, lanes=2, lane_steps=1 << 20)
    for gen, frames in enumerate(batches):
        tickets = [eng.submit(fr, past_key=(gen << 48) | fr["idx"][0], fut_key=(gen << 48) | fr["idx"][2])
                   for fr in frames]
        pending = set(tickets)
        while pending:
            eng.step()
            pending = {t for t in pending if not eng.done(t)}
        for fr, t in zip(frames, tickets):
            rec, st = eng.result(t)
            (rec1, st1), = ctx.decode_wz_frames([fr], batched=False)
            assert np.array_equal(rec, rec1)
            assert st["levels"] == st1["levels"] and st["iterations"] == st1["iterations"]
            assert st["bitrate"] == st1["bitrate"] and st["psnr"] == st1["psnr"]
    eng.close()


def test_engine_device_buffers_equal_host(ctx):
    """Requests whose key frames, syndromes and outputs live in HBM (device
    pointers, shared key frames under stream keys) decode to the same frames
    and statistics as host-buffer requests."""
    from dvc_amd import _lib
    vid = W.synthetic_video(352, 288, 5, 6, seed=990)
    frames = [dict(past=vid[k], fut=vid[k + 2], orig=vid[k + 1], wz=W.encode_wz_frame(vid[k + 1], q), idx=(k, k + 1, k + 2))
              for k, q in ((0, 8), (2, 5))]
    host = _lib.Engine(ctx, max_inflight=2).decode_all(frames)
    keys = {f: _lib.DeviceBuffer.from_array(ctx, vid[f]) for f in (0, 2, 4)}
    accs = [_lib.DeviceBuffer.from_array(ctx, np.asarray(fr["wz"].acc, np.uint8)) for fr in frames]
    recs = [_lib.DeviceBuffer(ctx, 288 * 352) for _ in frames]
    eng = _lib.Engine(ctx, max_inflight=2, device_buffers=True)
    tickets = [eng.submit_device(352, 288, fr["wz"], keys[fr["idx"][0]].ptr, keys[fr["idx"][2]].ptr,
                                 acc.ptr, rec.ptr, idx=fr["idx"], past_key=100 + fr["idx"][0],
                                 fut_key=100 + fr["idx"][2])
               for fr, acc, rec in zip(frames, accs, recs)]
    pending = set(tickets)
    while pending:
        eng.step()
        pending = {t for t in pending if not eng.done(t)}
    for t, rec, (hrec, hst) in zip(tickets, recs, host):
        _, st = eng.result(t)
        assert np.array_equal(rec.download((288, 352), np.uint8), hrec)
        assert st["levels"] == hst["levels"] and st["iterations"] == hst["iterations"] and st["bitrate"] == hst["bitrate"]
    eng.close()
