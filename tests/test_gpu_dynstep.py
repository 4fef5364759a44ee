"""The DYNAMICSTEP rate search on the GPU (dvc_params.rate_step = 1): codes
start + 0, 1, 3, 6, 10, 14, ... then the full rate, as the reference decoder
built with DYNAMICSTEP defined in LdpcaDecoder.cpp tries them.  Bitplanes and
whole WZ frames must equal that build's own results (golden fixtures from
oracle/_ref/libdvcref_dyn.so) and the oracle's rate_step 1, on the plain
sequential path and the speculative batched path."""
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


@pytest.fixture(scope="module")
def dctx():
    from dvc_amd import _lib
    c = _lib.Context(device=0, max_iter=50, rate_step=1)
    yield c
    c.close()


def _golden_planes():
    g = json.loads((GOLD / "ldpca_dynstep.json").read_text())
    sets = {}
    out = []
    for it in g["items"]:
        key = tuple(it["set"])
        if key not in sets:
            sets[key] = W.ldpca_bitplane_set(n_cif=key[0], n_qcif=key[1], seed=key[2])
        out.append((sets[key][it["index"]], it))
    return out


def _check_plane(dec, st, pl, it):
    odec, ost, oent = O.ldpca_decode(pl["LR"], pl["acc"], pl["crc"], rate_step=1)
    assert fbits(st["entropy"]) == it["entropy_bits"]
    assert st["level"] == it["level"] and st["rate"] == it["rate"]
    assert sha(dec) == it["decoded_sha256"]
    assert st["start_level"] == ost["start_level"] and st["iterations"] == ost["total_iters"]
    assert st["requests"] == ost["requests"] and st["used_inverse"] == ost["used_inverse"]
    assert np.array_equal(dec, odec)


def test_bitplanes_plain_match_dynstep_reference(dctx):
    for pl, it in _golden_planes():
        dec, st = dctx.ldpca_decode(pl["LR"], pl["acc"], pl["crc"])
        _check_plane(dec, st, pl, it)


def test_bitplanes_batched_match_dynstep_reference(dctx):
    items = _golden_planes()
    decs, sts = dctx.ldpca_decode_batch([p["LR"] for p, _ in items], [p["acc"] for p, _ in items],
                                        [p["crc"] for p, _ in items])
    for (pl, it), dec, st in zip(items, decs, sts):
        _check_plane(dec, st, pl, it)


@pytest.mark.parametrize("batched", [False, True])
def test_frames_match_dynstep_reference(dctx, batched):
    cases = json.loads((GOLD / "frames_dynstep.json").read_text())["cases"]
    frames = []
    for c in cases:
        vid = W.synthetic_video(c["W"], c["H"], c["frames"], c["shift"], c["seed"])
        frames.append(dict(past=vid[0], fut=vid[2], orig=vid[c["wz"]], wz=W.encode_wz_frame(vid[c["wz"]], c["qindex"])))
    outs = dctx.decode_wz_frames(frames, batched=batched)
    for (rec, st), c in zip(outs, cases):
        assert st["levels"] == c["levels"]
        assert [fbits(e) for e in st["entropy"]] == c["entropy_bits"]
        assert st["bitrate"] == c["bitrate"] and st["psnr"] == c["psnr"]
        assert sha(rec) == c["recons_sha256"]


def test_dynstep_needs_parallel_mode_0():
    from dvc_amd import _lib
    with pytest.raises(_lib.DvcError):
        _lib.Context(device=0, parallel_mode=1, rate_step=1)
