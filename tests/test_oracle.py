"""Pins the CPU oracle: closed forms, invariants and the golden fixtures that
tests/golden/make_golden.py wrote from the reference's own sources.  Where the
reference build (oracle/_ref) is present the oracle is also compared with it
directly on fresh seeded inputs."""
import hashlib
import json
import struct
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as O
from dvc_amd import workload as W

GOLD = Path(__file__).resolve().parent / "golden"
LIBM = O.ctypes.CDLL("libm.so.6")
LIBM.logf.restype = O.ctypes.c_float
LIBM.logf.argtypes = [O.ctypes.c_float]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def fbits(x):
    return struct.unpack("<I", struct.pack("<f", float(x)))[0]


def test_crc8_check_value():
    # CRC-8 (poly 0x07, init 0, no reflection, xorout 0): check value 0xF4
    assert O.orc().orc_crc8(O.P(np.frombuffer(b"123456789", np.uint8).copy()), 9) == 0xF4
    assert W.crc8(b"123456789") == 0xF4


def test_crc_linearity_basis():
    L = W.ladder(1584)
    rng = np.random.default_rng(0)
    for _ in range(5):
        bits = (rng.random(1584) < 0.3).astype(np.uint8)
        assert W.crc8(np.packbits(bits)) == W.ldpca_encode(L, bits)[1]


def test_ladder_matches_reference_fixture():
    g = json.loads((GOLD / "ladder.json").read_text())
    L = O.ladder(1584)
    ir = np.stack([L.ir(c) for c in range(L.num_codes)]).astype(np.int32)
    assert L.num_codes == g["num_codes"] and L.E == g["E"]
    assert L.synd_length().tolist() == g["synd_length"]
    assert sha(ir) == g["decoder_ir_sha256"]
    assert hashlib.sha256(L.inv().tobytes()).hexdigest() == g["inverse_sha256"]


def test_irregular_ladder_matches_reference_fixture():
    """The irregular QCIF ladder derived from the pack (irRegEncoder1584.lad
    and irReg1584.lad's transmission pattern) equals irReg1584.lad's decoder
    rows; its full-rate inverse (which the reference does not ship) inverts
    the encoder matrix (test_inverse_times_h_is_identity)."""
    g = json.loads((GOLD / "ladder.json").read_text())["irreg"]
    L = O.ladder(1584, "irreg")
    ir = np.stack([L.ir(c) for c in range(L.num_codes)]).astype(np.int32)
    assert L.num_codes == g["num_codes"] == 66 and L.E == g["E"]
    assert L.synd_length().tolist() == g["synd_length"]
    assert sha(ir) == g["decoder_ir_sha256"]


def test_irregular_ldpca_matches_reference_fixture():
    d = json.loads((GOLD / "ldpca_irreg.json").read_text())
    planes = W.ldpca_bitplane_set(n_cif=0, n_qcif=24, seed=35, kind="irreg")
    for it in d["items"]:
        pl = planes[it["index"]]
        dec, st, ent = O.ldpca_decode(pl["LR"], pl["acc"], pl["crc"], kind="irreg")
        assert st["level"] == it["level"] and fbits(ent) == it["entropy_bits"], it["index"]
        assert st["rate"] == it["rate"] and sha(dec) == it["decoded_sha256"], it["index"]


@pytest.mark.parametrize("fixture", ["frames_irreg.json", "frames_irreg_mode1.json"])
def test_irregular_frames_match_reference_fixture(fixture):
    for c in json.loads((GOLD / fixture).read_text())["cases"]:
        vid = W.synthetic_video(c["W"], c["H"], c["frames"], c["shift"], c["seed"])
        wz = W.encode_wz_frame(vid[c["wz"]], c["qindex"], kind="irreg")
        rec, info = O.decode_wz_frame(vid[0], vid[2], vid[c["wz"]], wz, c["qindex"], mode=c["mode"], nthreads=1,
                                      kind="irreg")
        assert info["levels"] == c["levels"]
        assert [fbits(e) for e in info["entropy"]] == c["entropy_bits"]
        assert info["bitrate"] == c["bitrate"] and info["psnr"] == c["psnr"]
        assert sha(rec) == c["recons_sha256"]


def test_cif_ladder_row_degrees_match_reference_comments():
    g = json.loads((GOLD / "ladder.json").read_text())
    L = O.ladder(6336)
    for code, hist in g["cif_row_degrees"].items():
        deg = np.bincount(L.ir(int(code)))
        got = {str(d): int(c) for d, c in zip(*np.unique(deg, return_counts=True))}
        for d, c in hist.items():
            assert got.get(d) == c, (code, d, got)


@pytest.mark.parametrize("n", [1584, 6336])
def test_inverse_times_h_is_identity(n):
    L = O.ladder(n)
    pk = W.ladder(n)
    inv = np.unpackbits(L.inv().reshape(n, -1), axis=1)[:, :n].astype(np.uint8)  # [j, i]
    rows = pk.ir.reshape(n, 3)                                                      # column k -> 3 rows
    prod = inv[:, rows[:, 0]] ^ inv[:, rows[:, 1]] ^ inv[:, rows[:, 2]]             # [j, k]
    assert np.array_equal(prod, np.eye(n, dtype=np.uint8))


def test_irregular_inverse_times_h_is_identity():
    """The irregular ladder's full-rate inverse (column degrees 2..21) over
    GF(2), as a dense product."""
    L = O.ladder(1584, "irreg")
    pk = W.ladder(1584, "irreg")
    n = 1584
    inv = np.unpackbits(L.inv().reshape(n, -1), axis=1)[:, :n].astype(np.float32)  # [j, i]
    H = np.zeros((n, n), np.float32)                                                # [i, k]
    cols = np.repeat(np.arange(n), np.diff(pk.jc))
    np.add.at(H, (pk.ir, cols), 1.0)
    prod = (inv @ H).astype(np.int64) & 1
    assert sorted(set(np.diff(pk.jc).tolist())) == [2, 3, 7, 8, 19, 21]
    assert np.array_equal(prod, np.eye(n, dtype=np.int64))


def test_library_inverse_cache_matches_oracle():
    for n in (1584, 6336):
        cached = (Path(W.DATA) / f"reg{n}.invh").read_bytes()
        assert cached == O.ladder(n).inv().tobytes()


def test_accumulated_syndrome_identity():
    """H x = s on the full-rate code and on a merged code: forming the
    decoder syndrome from the accumulated syndrome equals summing the decoder
    rows directly."""
    L = W.ladder(1584)
    OL = O.ladder(1584)
    rng = np.random.default_rng(1)
    bits = (rng.random(1584) < 0.2).astype(np.uint8)
    acc, _ = W.ldpca_encode(L, bits)
    accb = np.unpackbits(acc)
    col = np.repeat(np.arange(1584), 3)
    for code in (0, 7, 40, 64):
        ir = OL.ir(code)
        m = OL.synd_length()[code]
        direct = np.bincount(ir, weights=bits[col], minlength=m).astype(np.int64) & 1
        assert np.array_equal(W.form_syndrome(L, code, accb), direct[:m])


def test_encoder_matches_reference_fixture():
    g = json.loads((GOLD / "encoder.json").read_text())
    for c in g["cases"]:
        vid = W.synthetic_video(c["W"], c["H"], c["frame"] + 2, c["shift"], c["seed"])
        wz = W.encode_wz_frame(vid[c["frame"]], c["qindex"])
        assert wz.ac_range.tolist() == c["ac_range"]
        assert wz.crc.tolist() == c["crc"]
        assert sha(wz.acc) == c["acc_sha256"]
        # and the oracle's restatement of the encoder
        g2 = O.geom(c["W"], c["H"])
        n = (c["W"] // 4) * (c["H"] // 4)
        enc = O.ladder(n).enc
        acr = np.zeros(16, np.int32)
        crc = np.zeros(c["nbitplanes"], np.uint8)
        acc = np.zeros((c["nbitplanes"], n // 8), np.uint8)
        fr = np.ascontiguousarray(vid[c["frame"]])
        O.orc().orc_encode_wz_frame(O.ctypes.byref(g2), c["qindex"], O.ctypes.byref(enc), O.P(fr), O.P(acr),
                                    O.P(crc), O.P(acc))
        assert acr.tolist() == c["ac_range"] and crc.tolist() == c["crc"] and sha(acc) == c["acc_sha256"]


@pytest.mark.parametrize("fixture", ["ldpca.json", "ldpca_dynstep.json"])
def test_ldpca_matches_reference_fixture(fixture):
    """ldpca_dynstep.json comes from the reference decoder built with
    DYNAMICSTEP; the oracle's rate_step 1 must reproduce it."""
    g = json.loads((GOLD / fixture).read_text())
    sets = {}
    for it in g["items"]:
        key = tuple(it["set"])
        if key not in sets:
            sets[key] = W.ldpca_bitplane_set(n_cif=key[0], n_qcif=key[1], seed=key[2])
        pl = sets[key][it["index"]]
        dec, st, ent = O.ldpca_decode(pl["LR"], pl["acc"], pl["crc"], max_iter=g["max_iter"],
                                      rate_step=g.get("rate_step", 0))
        assert fbits(ent) == it["entropy_bits"]
        assert st["rate"] == it["rate"] and st["level"] == it["level"]
        assert sha(dec) == it["decoded_sha256"]
        assert np.array_equal(dec, pl["bits"]) == it["source_ok"]


def _frame_cases(name):
    return json.loads((GOLD / name).read_text())["cases"]


@pytest.mark.parametrize("case", range(4))
@pytest.mark.parametrize("fixture", ["frames.json", "frames_mode1.json", "frames_dynstep.json"])
def test_frame_decode_matches_reference_fixture(case, fixture):
    c = _frame_cases(fixture)[case]
    vid = W.synthetic_video(c["W"], c["H"], c["frames"], c["shift"], c["seed"])
    wz = W.encode_wz_frame(vid[c["wz"]], c["qindex"])
    rate_step = json.loads((GOLD / fixture).read_text()).get("rate_step", 0)
    rec, info = O.decode_wz_frame(vid[0], vid[2], vid[c["wz"]], wz, c["qindex"], mode=c["mode"], nthreads=1,
                                  rate_step=rate_step)
    assert info["levels"] == c["levels"]
    assert [fbits(e) for e in info["entropy"]] == c["entropy_bits"]
    assert info["bitrate"] == c["bitrate"]
    assert info["psnr"] == c["psnr"]
    assert sha(rec) == c["recons_sha256"]


@pytest.mark.parametrize("case", [0, 1])
def test_cif_low_qindex_frames_match_reference_fixture(case):
    """The oracle on the CIF Qindex 1 and 3 frames of frames_cif_q.json (the
    higher Qindex cases of that file run on the GPU only, to keep this suite
    short)."""
    c = _frame_cases("frames_cif_q.json")[case]
    vid = W.synthetic_video(c["W"], c["H"], c["frames"], c["shift"], c["seed"])
    wz = W.encode_wz_frame(vid[c["wz"]], c["qindex"])
    rec, info = O.decode_wz_frame(vid[0], vid[2], vid[c["wz"]], wz, c["qindex"], mode=0, nthreads=1)
    assert info["levels"] == c["levels"]
    assert [fbits(e) for e in info["entropy"]] == c["entropy_bits"]
    assert info["bitrate"] == c["bitrate"]
    assert info["psnr"] == c["psnr"]
    assert sha(rec) == c["recons_sha256"]


@pytest.mark.parametrize("fixture", ["ldpca_nan.json", "ldpca_nan_mode1.json"])
def test_ldpca_nan_bitplanes_match_reference_fixture(fixture):
    """The oracle's LdpcaDecoder on bitplanes with NaN likelihood ratios (x86's
    default NaN at the fixture's positions) against the reference's results:
    ParallelMode 0 fails every code and takes the full rate; ParallelMode 1
    decides NaN as 0 and its early jump sets the iteration count."""
    d = json.loads((GOLD / fixture).read_text())
    sets = {}
    for it in d["items"]:
        key = tuple(it["set"])
        if key not in sets:
            sets[key] = W.ldpca_bitplane_set(n_cif=key[0], n_qcif=key[1], seed=key[2])
        pl = sets[key][it["index"]]
        lr = np.array(pl["LR"], np.float32)
        lr.view(np.uint32)[it["nan_positions"]] = d["nan_bits"]
        dec, st, ent = O.ldpca_decode(lr, pl["acc"], pl["crc"], mode=d["mode"], nthreads=1,
                                      entropy=None if d["mode"] == 0 else O.compute_ce(lr))
        assert st["level"] == it["level"] and fbits(ent) == it["entropy_bits"]
        assert sha(dec) == it["decoded_sha256"]
        if it["bp_iterations"] is not None:
            # the reference's counter advances in the for-increment
            # (LdpcaDecoder.cpp:186), so it leaves out the last iteration of
            # every code that ends with a break; here every code ends with
            # the early jump's break
            assert st["total_iters"] == it["bp_iterations"] + (st["level"] - st["start_level"])
        else:
            assert st["total_iters"] == d["max_iter"] * (st["level"] - st["start_level"])


def _gop_cases():
    out = []
    for fx in ("frames_gop.json", "frames_gop_mode1.json"):
        out += [c for c in json.loads((GOLD / fx).read_text())["cases"] if c["W"] == 176]
    return out


@pytest.mark.parametrize("case", _gop_cases(), ids=lambda c: f"mode{c['mode']}-gop{c['gop']}")
def test_gop_decode_matches_reference_fixture(case):
    """decodeGOP(0, gop) (DVC_Decoder.cpp:26-474, 471-472) restated with the
    oracle's frame body: decoded WZ frames are the references of the later
    frames.  Frames whose references are decoded run concurrently (ctypes
    calls release the interpreter lock; the oracle keeps no shared state)."""
    from concurrent.futures import ThreadPoolExecutor
    gop, q = case["gop"], case["qindex"]
    vid = W.synthetic_video(case["W"], case["H"], gop + 1, case["shift"], case["seed"])
    decoded = {0: vid[0], gop: vid[gop]}
    gold = {g["cur"]: g for g in case["frames"]}
    todo = [(g["cur"], g["past"], g["fut"]) for g in case["frames"]]
    O.ladder(case["W"] * case["H"] // 16)

    def one(item):
        cur, past, fut = item
        wz = W.encode_wz_frame(vid[cur], q)
        return cur, O.decode_wz_frame(decoded[past], decoded[fut], vid[cur], wz, q, mode=case["mode"], nthreads=1,
                                      idx=(past, cur, fut))
    with ThreadPoolExecutor(4) as ex:
        while todo:
            ready = [t for t in todo if t[1] in decoded and t[2] in decoded]
            todo = [t for t in todo if t not in ready]
            for cur, (rec, info) in ex.map(one, ready):
                g = gold[cur]
                assert info["levels"] == g["levels"], cur
                assert [fbits(e) for e in info["entropy"]] == g["entropy_bits"], cur
                assert info["bitrate"] == g["bitrate"] and info["psnr"] == g["psnr"], cur
                assert sha(rec) == g["recons_sha256"], cur
                decoded[cur] = rec


def test_side_info_matches_reference_fixture():
    for c in _frame_cases("frames.json"):
        vid = W.synthetic_video(c["W"], c["H"], c["frames"], c["shift"], c["seed"])
        si, res = oracle_side_info(vid[0], vid[2])
        assert sha(si) == c["si_sha256"] and sha(res) == c["residue_sha256"]


def oracle_side_info(past, fut, idx=(0, 1, 2)):
    h, w = past.shape
    g = O.geom(w, h)
    lib = O.orc()
    plp, flp = np.zeros_like(past), np.zeros_like(fut)
    ph, fh = np.zeros((2 * h, 2 * w), np.uint8), np.zeros((2 * h, 2 * w), np.uint8)
    past = np.ascontiguousarray(past); fut = np.ascontiguousarray(fut)
    lib.orc_lowpass(O.P(past), w, h, O.P(plp)); lib.orc_lowpass(O.P(fut), w, h, O.P(flp))
    lib.orc_fir(O.P(past), w, h, O.P(ph)); lib.orc_fir(O.P(fut), w, h, O.P(fh))
    si = np.zeros((h, w), np.uint8)
    res = np.zeros((h, w), np.int32)
    mbmv = np.zeros(g.b1616n, O.MV_DTYPE)
    bmv = np.zeros(g.b88n, O.MV_DTYPE)
    lib.orc_create_side_info(O.ctypes.byref(g), O.P(plp), O.P(flp), O.P(ph), O.P(fh), idx[0], idx[1], idx[2],
                             O.P(si), O.P(res), O.P(mbmv), O.P(bmv), 0)
    return si, res


def test_fir_is_dyadic_integer_filter():
    """The 6-tap half-pel filter equals its exact integer form
    clamp((S+16)>>5) / clamp((S2+512)>>10) (coefficients are dyadic)."""
    rng = np.random.default_rng(3)
    img = rng.integers(0, 256, (16, 24), dtype=np.uint8)
    out = np.zeros((32, 48), np.uint8)
    O.orc().orc_fir(O.P(img), 24, 16, O.P(out))
    tap = np.array([1, -5, 20, 20, -5, 1])
    pad = np.pad(img.astype(np.int64), 3, mode="edge")
    for i in range(32):
        for j in range(48):
            y, x = i // 2, j // 2
            if i % 2 == 0 and j % 2 == 0:
                exp = img[y, x]
            elif i % 2 == 0:
                exp = np.clip((pad[y + 3, x + 1:x + 7] @ tap + 16) >> 5, 0, 255)
            elif j % 2 == 0:
                exp = np.clip((pad[y + 1:y + 7, x + 3] @ tap + 16) >> 5, 0, 255)
            else:
                s = tap @ pad[y + 1:y + 7, x + 1:x + 7] @ tap
                exp = np.clip((s + 512) >> 10, 0, 255)
            assert out[i, j] == exp, (i, j)


def test_forward_me_finds_global_translation():
    """Noise-free global translation: away from the borders forward ME returns
    that motion for every macroblock (cost is minimal, 0, at the true shift)."""
    rng = np.random.default_rng(4)
    base = rng.integers(0, 256, (200, 240)).astype(np.uint8)
    past = base[20:164, 20:196].copy()
    fut = base[23:167, 18:194].copy()   # future = past shifted by (dy=+3, dx=-2)
    w, h = 176, 144
    g = O.geom(w, h)
    si1616 = np.zeros(g.b1616n, O.SEARCH_INFO_DTYPE)
    O.orc().orc_init_1616(O.ctypes.byref(g), O.P(si1616))
    mb = np.zeros(g.b1616n, O.MV_DTYPE)
    holes = np.zeros((g.b1616n, 2), np.float64)
    O.orc().orc_forward_me(O.ctypes.byref(g), O.P(si1616), O.P(fut), O.P(past), 0, 1, 2, O.P(mb), O.P(holes), 0)
    # the recorded per-MB motion is the nearest-hole MV scaled by 1/2; check
    # the forward ME directly through the holes: hole = left/2 + past.x/2
    for i in range(g.b1616n):
        top, left = si1616[i]["top"], si1616[i]["left"]
        if 32 <= top <= h - 48 and 32 <= left <= w - 48:
            assert holes[i, 0] == left * 0.5 + (left - 2) * 0.5
            assert holes[i, 1] == top * 0.5 + (top + 3) * 0.5


@pytest.mark.skipif(not O.have_reference_build(), reason="reference build (oracle/_ref) not present")
def test_oracle_bp_modes_match_reference_directly():
    R = O.RefDecoder(176, 144, 4)
    planes = W.ldpca_bitplane_set(n_cif=0, n_qcif=6, seed=41)
    OL = O.ladder(1584)
    for pl in planes:
        llr = np.array([LIBM.logf(float(v)) for v in pl["LR"]], np.float32)  # the C library's logf
        acc_bits = np.unpackbits(pl["acc"])
        for code in (3, 10, 30):
            synd = W.form_syndrome(W.ladder(1584), code, acc_bits).astype(np.uint8)
            for mode in (0, 1):
                d_ref = np.full(1584, 0xAA, np.uint8)
                ok_ref = R.lib.ref_bp(mode, code, O.P(llr.copy()), O.P(synd), O.P(d_ref))
                ok, d, _ = O.bp(llr, synd, code, mode=mode)
                assert ok == ok_ref and np.array_equal(d, d_ref), (code, mode)
        assert OL.num_codes == 65
