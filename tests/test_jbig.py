"""The library's JBIG (T.82) decoder and WZ bitstream reader -- host code, no
GPU -- against the reference's JBIG-KIT and readWZBitstream
(tests/golden/jbig.json, wz_stream.json): every map of every option set
decodes to JBIG-KIT's image with the same byte count, readGOP's map path
(header rebuilt, data followed by more bitstream) reports simple_JBIG_dec's
byte count, and a WZ bitstream with skip frames parses to the fields the
reference reader fills in."""
import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

GOLD = Path(__file__).resolve().parent / "golden"


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def _unpack(hexrows, w, h):
    b = np.frombuffer(bytes.fromhex(hexrows), np.uint8).reshape(h, -1)
    return np.unpackbits(b, axis=1)[:, :w]


@pytest.fixture(scope="module")
def lib():
    from dvc_amd import _lib
    return _lib


def test_jbig_bie_matches_jbigkit(lib):
    cases = json.loads((GOLD / "jbig.json").read_text())["cases"]
    assert len({c["options"] for c in cases}) >= 7
    for c in cases:
        img, used = lib.jbig_decode(bytes.fromhex(c["bie"]))
        where = f"map {c['map']} ({c['kind']} {c['W']}x{c['H']}) options {c['options']}"
        assert img.shape == (c["H"], c["W"]), where
        assert np.array_equal(img, _unpack(c["image"], c["W"], c["H"])), where
        assert used == c["consumed"], where


def test_jbig_block_mode_map_path(lib):
    """readGOP's way: the data without its header, followed by the rest of the
    bitstream; the byte count is what skipBlockBits is made of."""
    for c in json.loads((GOLD / "jbig.json").read_text())["cases"]:
        if c["options"] != "block_mode_map":
            continue
        data = bytes.fromhex(c["bie"])[20:] + bytes(range(64))
        m, used = lib.jbig_decode_map(data, c["W"], c["H"])
        assert np.array_equal(m, _unpack(c["image"], c["W"], c["H"])), c["map"]
        assert used == c["map_decode_consumed"] == c["consumed"] - 20, c["map"]


def test_jbig_rejects_truncated_data(lib):
    c = next(c for c in json.loads((GOLD / "jbig.json").read_text())["cases"] if c["kind"] == "blocks")
    bie = bytes.fromhex(c["bie"])
    with pytest.raises(RuntimeError):
        lib.jbig_decode(bie[:len(bie) - 3])  # the stripe's end marker is gone
    with pytest.raises(RuntimeError):
        lib.jbig_decode(bie[:12])


def test_wz_stream_matches_reference_reader(lib):
    from dvc_amd import workload as W
    d = json.loads((GOLD / "wz_stream.json").read_text())
    c = d["case"]
    vid = W.synthetic_video(c["W"], c["H"], c["frames"], c["shift"], c["seed"])
    jbig = {int(k): bytes.fromhex(v) for k, v in d["skip_jbig"].items()}
    wz = {}
    for f in d["frames"]:
        cur = f["cur"]
        past = next(x["past"] for x in d["decoded"] if x["cur"] == cur)
        mode = W.skip_candidates(vid[cur], vid[past]) if cur in jbig else None
        wz[cur] = W.encode_wz_frame(vid[cur], c["qindex"], block_mode=mode)
    stream = W.write_wz_bitstream(c["W"], c["H"], c["qindex"], c["gop"], c["frames"], wz, skip_jbig=jbig)
    assert hashlib.sha256(stream).hexdigest() == d["stream_sha256"]
    info, frames = lib.read_wz_stream(stream)
    h = d["header"]
    assert (info["width"], info["height"], info["qindex"], info["intra_qp"], info["gop"], info["frames"],
            info["deblocking"], info["nbitplanes"]) == tuple(h)
    assert info["bytes_read"] == len(stream) == d["stream_bytes"]
    expect = []

    def gop(begin, end):  # readGOP's recursion (wzbitstream.cpp:13-126)
        cur = (begin + end) // 2
        if cur > begin:
            expect.append((begin, cur, end))
            gop(begin, cur)
            gop(cur, end)
    for i in range(0, c["frames"], c["gop"]):
        gop(i, min(i + c["gop"], c["frames"] - 1))
    assert [tuple(f["idx"]) for f in frames] == expect
    ref = {f["cur"]: f for f in d["frames"]}
    n8 = (c["W"] // 8) * (c["H"] // 8)
    # the AC ranges readGOP reads (the others stay unset in the reference)
    read = []
    for j in range(1, 16):
        if W.QTABLE[c["qindex"]][W.ZIGZAG[j]] <= 0:
            break
        read.append(W.ZIGZAG[j])
    for f in frames:
        r = ref[f["idx"][1]]
        wzf = f["wz"]
        mode = wzf.block_mode if wzf.block_mode is not None else np.zeros(n8, np.uint8)
        assert [int(wzf.ac_range[i]) for i in read] == [r["ac_range"][i] for i in read]
        assert (f["skip_mode"], f["intra_mode"]) == (r["skip_mode"], r["intra_mode"])
        assert wzf.mode_bits == r["skip_bits"] + r["intra_bits"]
        assert _sha(mode) == r["block_mode_sha256"]
        assert _sha(wzf.crc) == r["crc_sha256"] and _sha(wzf.acc) == r["acc_sha256"]


def test_wz_stream_rejects_bad_input(lib):
    with pytest.raises(RuntimeError):
        lib.read_wz_stream(bytes([2, 4, 0, 4, 0, 9, 0]))  # resolution byte 2
    with pytest.raises(RuntimeError):
        lib.read_wz_stream(bytes([0, 4, 0, 4, 0, 9, 0, 1, 2]))  # truncated first frame


def test_intra_stream_mode_bits_match_reference(lib):
    """The CAVLC parse of every intra frame uses exactly the bits the
    reference's readCoeff4x4_CAVLC used: each frame's mode bits (skip map +
    intra map + CAVLC bits) equal skipBlockBits + intraBlockBits after the
    reference's decodeGOP (tests/golden/intra_streams.json)."""
    import base64
    d = json.loads((GOLD / "intra_streams.json").read_text())
    assert any(f["intra_blocks"] for c in d["cases"] for f in c["decoded"])
    for c in d["cases"]:
        info, frames = lib.read_wz_stream(base64.b64decode(c["wz_stream"]), base64.b64decode(c["intra_stream"]))
        assert (info["deblocking"], info["intra_qp"]) == (c["header"][6], c["header"][3])
        gold = {f["cur"]: f for f in c["decoded"]}
        for f in frames:
            g = gold[f["idx"][1]]
            wz = f["wz"]
            assert wz.mode_bits == g["skip_block_bits"] + g["intra_block_bits"], (c["name"], f["idx"])
            n_intra = 0 if wz.block_mode is None else int((wz.block_mode == 1).sum())
            assert n_intra == g["intra_blocks"] and (wz.intra_coeffs is not None) == (n_intra > 0)
