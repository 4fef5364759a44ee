"""Serving loop (dvc_amd.serve): Poisson arrivals from concurrent streams of
mixed size, GOP and quantiser, admitted into the continuous-batching engine
as their references become available (in GOP 4/8 decoded WZ frames are the
references of later ones, DVC_Decoder.cpp:471-472).

The serving configuration's shape: 32 streams, 24 QCIF and 8 CIF, GOP drawn
from {2, 4, 8} and Qindex from 1..8 per stream (seed 4001 draws every GOP
size at both frame sizes and every Qindex, Q8 among the CIF streams), 9
frames per stream.  Every served frame must equal its decode alone (an
engine with one frame in flight, given the same references), and a sample
with one frame per (frame size, GOP size) must equal the oracle's decode
(decodeGOP's frame body restated on the CPU) given the same references."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEED = 4001


def test_served_frames_equal_solo(ctx):
    from dvc_amd import serve
    streams = serve.make_streams(n_qcif=3, n_cif=1, frames=9, seed=4100)
    assert {s.gop for s in streams} - {2}, "the mix should hold a GOP above 2"
    res, outputs = serve.run(ctx, streams, rate=400.0, max_inflight=8, lanes=2,
                             verify=serve.verify_solo(ctx, sample=0))
    print(res, flush=True)
    assert res["completed"] == res["requests"] == sum(len(s.order) for s in streams)
    assert res["verified"]["mismatches"] == 0 and res["verified"]["checked"] == res["requests"]
    assert res["latency_ms"]["p99"] >= res["latency_ms"]["p50"] > 0


@pytest.fixture(scope="module")
def served(ctx):
    from dvc_amd import serve
    streams = serve.make_streams(n_qcif=24, n_cif=8, frames=9, seed=SEED)
    assert {(s.width, s.gop) for s in streams} == {(w, g) for w in (176, 352) for g in (2, 4, 8)}
    assert {s.qindex for s in streams} == set(range(1, 9))
    res, outputs = serve.run(ctx, streams, rate=300.0, max_inflight=64, lanes=2)
    print({k: v for k, v in res.items()}, flush=True)
    return streams, res, outputs


def test_config4_every_served_frame_equals_solo(ctx, served):
    from dvc_amd import serve
    streams, res, outputs = served
    total = sum(len(s.order) for s in streams)
    assert res["completed"] == res["requests"] == total
    assert res["latency_ms"]["p99"] >= res["latency_ms"]["p50"] > 0
    ver = serve.verify_solo(ctx, sample=0)(outputs, streams)
    print(ver, flush=True)
    assert ver["checked"] == total and ver["mismatches"] == 0


def test_config4_sample_equals_oracle(served):
    """One served frame per (frame size, GOP size) against the oracle; for GOP
    4 and 8 the frame picked is the last one of its stream's decode order,
    whose references include decoded WZ frames."""
    import oracle_lib as O
    from dvc_amd import serve
    streams, _, outputs = served
    picks = {}
    for s in streams:
        key = (s.width, s.gop)
        if key not in picks or s.qindex > picks[key][0].qindex:
            picks[key] = (s, s.order[-1][0])
    for n in (1584, 6336):
        O.ladder(n)

    def oracle(item):
        s, cur = item
        past, fut = serve.references(s, cur)
        return O.decode_wz_frame(s.decoded[past], s.decoded[fut], s.video[cur], s.wz[cur], s.qindex,
                                 idx=(past, cur, fut))

    items = [picks[k] for k in sorted(picks)]
    first = oracle(items[0])  # serial first call: the oracle's transform scales are set up once
    with ThreadPoolExecutor(len(items) - 1) as ex:
        rest = list(ex.map(oracle, items[1:]))
    for (s, cur), (orec, ost) in zip(items, [first] + rest):
        rec, st = outputs[(s.sid, cur)]
        where = f"stream {s.sid} {s.width}x{s.height} gop {s.gop} q{s.qindex} frame {cur}"
        if s.gop > 2:
            past, fut = serve.references(s, cur)
            assert any(r not in range(0, 10**6, s.gop) for r in (past, fut)), where
        assert np.array_equal(rec, orec), where
        assert st["levels"] == ost["levels"] and st["psnr"] == ost["psnr"] and st["bitrate"] == ost["bitrate"], where
