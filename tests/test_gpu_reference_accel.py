"""The drop-in proven on the reference itself: the reference decoder's own
objects (oracle/_ref, compiled from DvcDec/src unmodified), linked with the
binding oracle/ref_accel_stubs.cpp so that its calls to LdpcaDecoder,
noiseDistribution, condBitProb, updateSI and updateSMF (and the frame set-up
at initSMF) go through the C ABI to the GPU library, runs decodeGOP on the
golden frame cases.  Every accepted code, entropy bit pattern, bitrate, PSNR
and reconstruction must equal tests/golden/frames.json, which the unrouted
reference wrote.  Skipped where the routed build was not made (it needs the
reference tree at build time)."""
import ctypes
import hashlib
import json
import struct
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
ACCEL_SO = ROOT / "oracle" / "_ref" / "libdvcref_accel.so"
GOLD = ROOT / "tests" / "golden"


def _fbits(x):
    return struct.unpack("<I", struct.pack("<f", float(x)))[0]


@pytest.mark.skipif(not ACCEL_SO.exists(), reason="oracle/_ref/libdvcref_accel.so not built (needs the reference tree)")
def test_reference_decoder_routed_to_gpu_matches_golden(tmp_path):
    import oracle_lib as O
    from dvc_amd import _lib  # noqa: F401  (loads libdvc_accel.so from the package directory first)
    from dvc_amd import workload as W
    lib = ctypes.CDLL(str(ACCEL_SO))
    lib.ref_get_psnr.restype = ctypes.c_double
    lib.ref_accel_error.restype = ctypes.c_char_p
    lib.ref_accel_routed.restype = ctypes.c_long
    assert lib.ref_accel_open(50) == 0
    try:
        cases = json.loads((GOLD / "frames.json").read_text())["cases"]
        for c in cases:
            print(f"routed reference decodeGOP: {c['W']}x{c['H']} q{c['qindex']} seed {c['seed']}", flush=True)
            vid = W.synthetic_video(c["W"], c["H"], c["frames"], c["shift"], c["seed"])
            wz = W.encode_wz_frame(vid[c["wz"]], c["qindex"])
            R = O.RefDecoder(c["W"], c["H"], c["qindex"], mode=0, nthreads=1, tmpdir=str(tmp_path), lib=lib)
            rec, info = R.decode_wz_frame(vid[0], vid[2], vid[c["wz"]], wz)
            assert lib.ref_accel_error() == b"", lib.ref_accel_error()
            assert info["levels"] == c["levels"]
            assert [_fbits(e) for e in info["entropy"]] == c["entropy_bits"]
            assert info["bitrate"] == c["bitrate"]
            assert info["psnr"] == c["psnr"]
            assert hashlib.sha256(np.ascontiguousarray(rec).tobytes()).hexdigest() == c["recons_sha256"]
        nbp = sum(len(c["levels"]) for c in cases)
        routed = [lib.ref_accel_routed(k) for k in range(6)]
        print("routed calls (ldpca, frame init, noise dist, condbit, updateSI, updateSMF):", routed, flush=True)
        assert routed[0] == nbp and routed[3] == nbp
        assert routed[1] == len(cases) and min(routed[2], routed[4], routed[5]) > 0
    finally:
        lib.ref_accel_close()
