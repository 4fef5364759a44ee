"""ParallelMode 1 semantics on the GPU (beliefPropagation_OpenMP,
LdpcaDecoder.cpp:164-306: decisions LLR < 0, failure after four consecutive
iterations with unchanged decisions, compared first against the decisions
the previous call left): whole frames against the reference's own mode-1
decodes (golden fixtures made with one OpenMP thread and a 0xAA-filled
decision buffer), and bitplanes against the oracle's mode-1 rate search with
carried decisions that provoke the repeat case, in the plain and the
speculative batched forms."""
import hashlib
import json
import struct
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as O
from dvc_amd import _lib, workload as W

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def ctx1():
    c = _lib.Context(max_iter=50, parallel_mode=1)
    yield c
    c.close()


def _fbits(x):
    return struct.unpack("<I", struct.pack("<f", float(x)))[0]


@pytest.mark.parametrize("case", range(4))
@pytest.mark.parametrize("batched", [False, True])
def test_mode1_frames_match_reference_golden(ctx1, case, batched):
    c = json.loads((GOLD / "frames_mode1.json").read_text())["cases"][case]
    vid = W.synthetic_video(c["W"], c["H"], c["frames"], c["shift"], c["seed"])
    fr = dict(past=vid[0], fut=vid[2], orig=vid[c["wz"]], wz=W.encode_wz_frame(vid[c["wz"]], c["qindex"]))
    (rec, st), = ctx1.decode_wz_frames([fr], batched=batched)
    print(c["W"], c["qindex"], batched, st["levels"][:10], c["levels"][:10], flush=True)
    assert st["levels"] == c["levels"]
    assert [_fbits(e) for e in st["entropy"]] == c["entropy_bits"]
    assert st["bitrate"] == c["bitrate"]
    assert hashlib.sha256(np.ascontiguousarray(rec).tobytes()).hexdigest() == c["recons_sha256"]
    assert st["psnr"] == c["psnr"]


def _carries(planes, rng):
    """Carried decisions per bitplane: none matching (0xAA), zeros, and the
    hard decisions of the channel (close to a code's first decisions, which
    exercises the repeat case)."""
    out = []
    for i, p in enumerate(planes):
        kind = i % 3
        if kind == 0:
            out.append(np.full(p["N"], 0xAA, np.uint8))
        elif kind == 1:
            out.append(np.zeros(p["N"], np.uint8))
        else:
            out.append((p["LR"] < 1.0).astype(np.uint8))
    return out


def test_mode1_bitplanes_against_oracle(ctx1):
    planes = W.ldpca_bitplane_set(n_cif=18, n_qcif=30, seed=77)
    carries = _carries(planes, np.random.default_rng(0))
    dec_b, st_b = ctx1.ldpca_decode_batch([p["LR"] for p in planes], [p["acc"] for p in planes],
                                          [p["crc"] for p in planes], carries=carries)
    for p, car, db, sb in zip(planes, carries, dec_b, st_b):
        od, ost, _ = O.ldpca_decode(p["LR"], p["acc"], p["crc"], mode=1, max_iter=50, nthreads=1, decoded_init=car)
        dp, sp = ctx1.ldpca_decode(p["LR"], p["acc"], p["crc"], carry=car)
        for d, s in ((db, sb), (dp, sp)):
            assert np.array_equal(d, od)
            assert s["start_level"] == ost["start_level"]
            assert s["level"] == ost["level"] and s["rate"] == ost["rate"]
            assert s["iterations"] == ost["total_iters"]


@pytest.mark.parametrize("size", [(176, 144, 3), (352, 288, 6)])
def test_mode1_frames_against_oracle_max_iter_100(size):
    w, h, shift = size
    vid = W.synthetic_video(w, h, 3, shift, seed=1500)
    wz = W.encode_wz_frame(vid[1], 8)
    orec, oinfo = O.decode_wz_frame(vid[0], vid[2], vid[1], wz, 8, mode=1, nthreads=1, max_iter=100)
    c = _lib.Context(max_iter=100, parallel_mode=1)
    try:
        for batched in (False, True):
            (rec, st), = c.decode_wz_frames([dict(past=vid[0], fut=vid[2], orig=vid[1], wz=wz)], batched=batched)
            assert np.array_equal(rec, orec)
            assert st["levels"] == oinfo["levels"] and st["iterations"] == oinfo["iters"]
            assert st["bitrate"] == oinfo["bitrate"] and st["psnr"] == oinfo["psnr"]
    finally:
        c.close()


def test_mode1_repeat_case_resolution(ctx1):
    """Bitplanes whose decisions never move (saturated channel LLRs): every
    code after the first repeats the decisions carried into it, so under
    ParallelMode 1 it fails after 4 iterations instead of 5.  The speculative
    launch runs them as if they did not repeat and the resolution must turn
    each into the 4-iteration failure; with 200 flipped bits and LR 1e12 some
    codes move and one later converges.  Counts and bits equal the oracle's."""
    rng = np.random.default_rng(5)
    planes, carries = [], []
    for n, ones, lrv, zero_carry in [(1584, 12, 1e30, False), (1584, 12, 1e30, True), (6336, 30, 1e30, False),
                                     (1584, 200, 1e12, False), (6336, 300, 1e12, True)]:
        bits = np.zeros(n, np.uint8)
        bits[rng.choice(n, ones, replace=False)] = 1
        acc, crc = W.ldpca_encode(W.ladder(n), bits)
        planes.append(dict(N=n, bits=bits, LR=np.full(n, np.float32(lrv), np.float32), acc=acc, crc=crc))
        carries.append(np.zeros(n, np.uint8) if zero_carry else np.full(n, 0xAA, np.uint8))
    dec_b, st_b = ctx1.ldpca_decode_batch([p["LR"] for p in planes], [p["acc"] for p in planes],
                                          [p["crc"] for p in planes], carries=carries)
    for p, car, db, sb in zip(planes, carries, dec_b, st_b):
        od, ost, _ = O.ldpca_decode(p["LR"], p["acc"], p["crc"], mode=1, max_iter=50, nthreads=1, decoded_init=car)
        dp, sp = ctx1.ldpca_decode(p["LR"], p["acc"], p["crc"], carry=car)
        print(p["N"], ost, sb["iterations"], sp["iterations"], flush=True)
        assert np.array_equal(od, p["bits"])
        for d, s in ((db, sb), (dp, sp)):
            assert np.array_equal(d, od)
            assert (s["level"], s["iterations"], s["used_inverse"]) == (ost["level"], ost["total_iters"],
                                                                         ost["used_inverse"])
