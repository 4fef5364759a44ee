"""The irregular QCIF ladder (irReg1584.lad: 66 codes, column degrees 2..21,
the reference's alternative at DvcDec/src/main.cpp:84-85) on the GPU against
the reference run with that ladder (tests/golden/ldpca_irreg.json,
frames_irreg.json, frames_irreg_mode1.json) and the oracle: bitplanes by the
plain sequential search and the speculative batch (mixed with CIF bitplanes,
which keep the regular ladder), and whole WZ frames through the engine and
the plain path, ParallelMode 0 and 1."""
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


@pytest.fixture(scope="module")
def ictx():
    from dvc_amd import _lib
    c = _lib.Context(max_iter=50, qcif_ladder="irreg")
    yield c
    c.close()


def test_irregular_bitplanes_match_reference(ictx):
    import oracle_lib as O
    from dvc_amd import workload as W
    d = json.loads((GOLD / "ldpca_irreg.json").read_text())
    planes = W.ldpca_bitplane_set(n_cif=0, n_qcif=24, seed=35, kind="irreg")
    cif = W.ldpca_bitplane_set(n_cif=4, n_qcif=0, seed=36)
    nc, _, _ = ictx.ladder_info(1584)
    assert nc == 66
    plain = [ictx.ldpca_decode(p["LR"], p["acc"], p["crc"]) for p in planes]
    mix = planes + cif
    bdec, bst = ictx.ldpca_decode_batch([p["LR"] for p in mix], [p["acc"] for p in mix], [p["crc"] for p in mix])
    for it in d["items"]:
        i = it["index"]
        ost = O.ldpca_decode(planes[i]["LR"], planes[i]["acc"], planes[i]["crc"], kind="irreg")[1]
        for how, dec, st in (("plain", plain[i][0], plain[i][1]), ("batched", bdec[i], bst[i])):
            where = f"item {i} {how}"
            assert st["level"] == it["level"] and st["rate"] == it["rate"], where
            assert _fbits(st["entropy"]) == it["entropy_bits"], where
            assert _sha(dec) == it["decoded_sha256"], where
            assert st["iterations"] == ost["total_iters"], where
    for k, p in enumerate(cif):  # the CIF bitplanes of the batch keep the regular ladder
        ost = O.ldpca_decode(p["LR"], p["acc"], p["crc"])[1]
        assert np.array_equal(bdec[len(planes) + k], p["bits"]) and bst[len(planes) + k]["level"] == ost["level"]


@pytest.mark.parametrize("fixture,mode", [("frames_irreg.json", 0), ("frames_irreg_mode1.json", 1)])
def test_irregular_frames_match_reference(fixture, mode):
    from dvc_amd import _lib, workload as W
    cases = json.loads((GOLD / fixture).read_text())["cases"]
    frames = []
    for c in cases:
        vid = W.synthetic_video(c["W"], c["H"], c["frames"], c["shift"], c["seed"])
        frames.append(dict(past=vid[0], fut=vid[2], orig=vid[c["wz"]],
                           wz=W.encode_wz_frame(vid[c["wz"]], c["qindex"], kind="irreg")))
    with _lib.Context(max_iter=50, parallel_mode=mode, qcif_ladder="irreg") as c1:
        for batched in (True, False):
            for c, (rec, st) in zip(cases, c1.decode_wz_frames(frames, batched=batched)):
                where = f"q{c['qindex']} batched={batched}"
                assert st["levels"] == c["levels"], where
                assert [_fbits(e) for e in st["entropy"]] == c["entropy_bits"], where
                assert st["bitrate"] == c["bitrate"] and st["psnr"] == c["psnr"], where
                assert _sha(rec) == c["recons_sha256"], where
