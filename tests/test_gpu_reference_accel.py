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


@pytest.mark.skipif(not ACCEL_SO.exists(), reason="oracle/_ref/libdvcref_accel.so not built (needs the reference tree)")
def test_reference_decoder_routed_to_gpu_decodes_skip_and_intra_stream(tmp_path):
    """The reference's main path (readWZBitstream, readIntraBitstream, then
    decodeGOP per GOP) on the GOP-4 stream with skip blocks, intra blocks and
    deblocking of tests/golden/intra_streams.json, with its per-band calls
    routed to the GPU.  initSIR leaves the skip and intra blocks out of the
    SIR (sideInfoCreation.cpp:554-596), so the per-band entry points work on
    that subset; every frame's codes, entropy bits, bitrate, PSNR and
    reconstruction must equal the unrouted reference's.  One sequence in this
    process, as the reference decodes one per run: deblock_filter keeps the QP
    of its first call in a function-local static (deblock.cpp:45)."""
    import base64
    import oracle_lib as O
    from dvc_amd import _lib  # noqa: F401
    from test_gpu_intra import _video
    case = next(c for c in json.loads((GOLD / "intra_streams.json").read_text())["cases"]
                if c["name"] == "qcif_q4_gop4_skip_intra_deblock")
    lib = ctypes.CDLL(str(ACCEL_SO))
    lib.ref_get_psnr.restype = ctypes.c_double
    lib.ref_accel_error.restype = ctypes.c_char_p
    lib.ref_accel_routed.restype = ctypes.c_long
    assert lib.ref_accel_open(50) == 0
    try:
        w, h, q, gop, nf = case["W"], case["H"], case["qindex"], case["gop"], case["frames"]
        wz_path, intra_path = tmp_path / "seq.wz", tmp_path / "seq.intra"
        wz_path.write_bytes(base64.b64decode(case["wz_stream"]))
        intra_path.write_bytes(base64.b64decode(case["intra_stream"]))
        R = O.RefDecoder(w, h, q, mode=0, nthreads=1, nframes=nf, tmpdir=str(tmp_path), lib=lib)
        hdr = np.zeros(8, np.int32)
        assert lib.ref_load_stream(str(wz_path).encode(), str(intra_path).encode(), O.P(hdr)) == case["intra_mode_exist"]
        assert hdr.tolist() == case["header"]
        vid = _video(case)
        order = []
        for i in range(0, nf, gop):
            O._gop_order(i, min(i + gop, nf - 1), order)
        wzset = {c for c, _, _ in order}
        for k in range(nf):
            pix = vid[k] if k not in wzset else np.zeros((h, w), np.uint8)
            lib.ref_set_frame(k, O.P(np.ascontiguousarray(pix)), O.P(np.ascontiguousarray(vid[k])))
        gold = {g["cur"]: g for g in case["decoded"]}
        assert sum(g["skip_blocks"] for g in gold.values()) > 0 and sum(g["intra_blocks"] for g in gold.values()) > 0
        for i in range(0, nf, gop):
            end = min(i + gop, nf - 1)
            sub = []
            O._gop_order(i, end, sub)
            print(f"routed reference decodeGOP({i}, {end}) on {case['name']}", flush=True)
            cnt = lib.ref_decode_gop(i, end)
            assert lib.ref_accel_error() == b"", lib.ref_accel_error()
            rates = np.zeros(4096, np.int32)
            levels = np.zeros(4096, np.int32)
            ent = np.zeros(4096, np.float32)
            lib.ref_get_log(O.P(rates), O.P(levels), O.P(ent))
            at = 0
            for cur, _, _ in sub:
                g = gold[cur]
                nb = len(g["levels"])
                rec = np.zeros((h, w), np.uint8)
                lib.ref_get_frame(cur, O.P(rec))
                assert levels[at:at + nb].tolist() == g["levels"], cur
                assert [_fbits(e) for e in ent[at:at + nb]] == g["entropy_bits"], cur
                assert lib.ref_get_bitrate(cur) == g["bitrate"], cur
                assert lib.ref_get_psnr(cur) == g["psnr"], cur
                assert hashlib.sha256(np.ascontiguousarray(rec).tobytes()).hexdigest() == g["recons_sha256"], cur
                at += nb
            assert at == cnt
        routed = [lib.ref_accel_routed(k) for k in range(6)]
        print("routed calls (ldpca, frame init, noise dist, condbit, updateSI, updateSMF):", routed, flush=True)
        assert routed[1] == len(gold) and min(routed[2:]) > 0
    finally:
        lib.ref_accel_close()
