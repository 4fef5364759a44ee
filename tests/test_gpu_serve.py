"""Serving loop (dvc_amd.serve): Poisson arrivals from concurrent streams of
mixed size, GOP and quantiser, admitted into the continuous-batching engine
as their references become available (GOP 4/8 decode decoded WZ frames'
successors); every served frame equals its solo decode."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


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
