"""Writes the golden fixtures under tests/golden/ from the reference's own CPU
sources (compiled into oracle/_ref by `make -C oracle ref`) and data files.
Run in a checkout that has /root/reference:

    python tests/golden/make_golden.py

Fixtures (all small JSON):
  ladder.json     hashes of the reference's RegDecoder1584.lad decoder rows and
                  Inverse_Matrix_H_Reg1584.bin, and the row-degree histograms
                  the reference lists for N=6336 codes (LDPCA_kernel.cu comments)
  encoder.json    reference encoder (DvcEnc forwardTransform/quantize/
                  extractBitplane/LdpcaEncoder) outputs for seeded frames
  ldpca.json      reference LdpcaDecoder (ParallelMode 0) results for seeded
                  bitplanes: start code, accepted code, rate, entropy bits,
                  decoded-bitplane hash
  frames.json     reference decodeGOP results for seeded WZ frames: accepted
                  code per bitplane, entropy bits, bitrate, PSNR, hashes of
                  the side information and of the reconstruction
  frames_mode1.json  the same with ParallelMode 1
  ldpca_dynstep.json, frames_dynstep.json  the ldpca and frames fixtures from
                  the decoder built with DYNAMICSTEP defined in
                  LdpcaDecoder.cpp (DVC_GOLDEN_ONLY=dynstep writes just these)
  ldpca_irreg.json, frames_irreg.json, frames_irreg_mode1.json  the irregular
                  QCIF ladder (irReg1584.lad): reference LdpcaDecoder on
                  seeded bitplanes and decodeGOP on QCIF frames, ParallelMode
                  0 and 1 (DVC_GOLDEN_ONLY=irreg writes just these)
  ldpca_nan.json, ldpca_nan_mode1.json  reference LdpcaDecoder, ParallelMode
                  0 and 1, on seeded bitplanes with NaN likelihood ratios
                  (DVC_GOLDEN_ONLY=nan writes just these)
  frames_gop.json, frames_gop_mode1.json  reference decodeGOP(0, gop) for
                  GOP 4 and 8 (decoded WZ frames are references of later
                  ones), ParallelMode 0 and 1 (DVC_GOLDEN_ONLY=gop writes
                  just these)
  jbig.json       the reference's JBIG-KIT (DvcDec/src/jbig.cpp) on seeded
                  bi-level maps: each map's BIE as JBIG-KIT's encoder writes
                  it under several option sets (the block-mode maps' own,
                  typical prediction, two-line template, several stripes,
                  SDRST, adaptive template), and what JBIG-KIT's decoder and
                  readGOP's simple_JBIG_dec read back (image, bytes used)
  frames_skip.json  skip blocks: the reference encoder with skip blocks on
                  (findSkipBlock + the map's JBIG data) and decodeGOP on the
                  map readGOP decodes, QCIF and CIF, ParallelMode 0
  wz_stream.json  a WZ bitstream with skip frames (GOP 4) as the reference's
                  readWZBitstream parses it, and decodeGOP over its GOPs
  (DVC_GOLDEN_ONLY=skip writes just these three)
  intra_streams.json  whole sequences through the reference encoder's
                  main loop (encodeGOP with skip blocks, intra blocks and
                  deblocking as each case sets them, writeWZbitstream and the
                  intra bitstream), kept as the two bitstreams, and the
                  reference decoder's main path on them (readWZBitstream,
                  readIntraBitstream, decodeGOP per GOP): per WZ frame its
                  block modes, codes, entropy bits, bitrate, PSNR and
                  reconstruction hash; intra_streams_mode1.json the same streams
                  decoded in ParallelMode 1 (DVC_GOLDEN_ONLY=intra writes these two)
Inputs are regenerated from their seeds by dvc_amd.workload, so a fixture
holds only seeds, sizes and expected outputs (and the small JBIG images,
which take JBIG-KIT's encoder to make).
"""
import ctypes
import hashlib
import json
import os
import struct
import subprocess
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

import oracle_lib as O  # noqa: E402
from dvc_amd import workload as W  # noqa: E402

REF = Path("/root/reference")


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def fbits(x):
    return struct.unpack("<I", struct.pack("<f", float(x)))[0]


def ladder_fixture():
    L = O.Ladder(lad_path=REF / "release/DvcDec/LDPCA_LadderFile/RegDecoder1584.lad",
                 inv_path=REF / "release/DvcDec/LDPCA_LadderFile/Inverse_Matrix_H_Reg1584.bin")
    ir = np.stack([L.ir(c) for c in range(L.num_codes)]).astype(np.int32)
    inv = (REF / "release/DvcDec/LDPCA_LadderFile/Inverse_Matrix_H_Reg1584.bin").read_bytes()
    return {
        "N": 1584, "num_codes": int(L.num_codes), "E": int(L.E),
        "synd_length": L.synd_length().tolist(),
        "decoder_ir_sha256": sha(ir),
        "inverse_sha256": hashlib.sha256(inv).hexdigest(),
        # Row-degree histograms of the N=6336 decoder codes, as listed in the
        # reference's kernel comments (DvcDec/src/LDPCA_kernel.cu:321,364,465,584):
        # {code: {degree: rows}}
        # the irregular QCIF ladder irReg1584.lad (main.cpp:84-85): decoder rows
        # of its 66 codes; the reference ships no inverse for it
        "irreg": irreg_ladder_facts(),
        "cif_row_degrees": {
            "1": {"48": 96, "50": 5, "51": 84, "52": 7, "98": 10, "99": 78, "100": 8},
            "3": {"23": 5, "24": 84, "25": 7, "27": 96, "48": 192, "50": 10, "51": 78, "52": 8},
            "30": {"5": 15, "6": 2850, "7": 15, "9": 192},
            "63": {"2": 15, "3": 6114, "4": 15, "6": 96},
        },
    }


def irreg_ladder_facts():
    L = O.Ladder(lad_path=REF / "release/DvcDec/LDPCA_LadderFile/irReg1584.lad")
    ir = np.stack([L.ir(c) for c in range(L.num_codes)]).astype(np.int32)
    return {"N": 1584, "num_codes": int(L.num_codes), "E": int(L.E), "synd_length": L.synd_length().tolist(),
            "decoder_ir_sha256": sha(ir)}


def encoder_fixture():
    enc = O.ref_enc()
    cases = []
    for (w, h, q, seed, frame) in [(176, 144, 4, 1000, 1), (176, 144, 8, 1001, 3), (352, 288, 8, 1002, 1),
                                   (352, 288, 1, 1003, 1)]:
        vid = W.synthetic_video(w, h, frame + 2, 3 if w == 176 else 6, seed)
        n = (w // 4) * (h // 4)
        lad = REF / f"release/DvcEnc/LDPCA_LadderFile/RegEncoder{n}.lad"
        nbp = enc.refenc_setup(w, h, q, str(lad).encode())
        acr = np.zeros(16, np.int32)
        crc = np.zeros(nbp, np.uint8)
        acc = np.zeros((nbp, n // 8), np.uint8)
        enc.refenc_encode(O.P(np.ascontiguousarray(vid[frame])), O.P(acr), O.P(crc), O.P(acc))
        cases.append({"W": w, "H": h, "qindex": q, "seed": seed, "frame": frame,
                      "shift": 3 if w == 176 else 6, "nbitplanes": int(nbp),
                      "ac_range": acr.tolist(), "crc": crc.tolist(), "acc_sha256": sha(acc)})
    return {"cases": cases}


def ldpca_fixture():
    R = O.ref_dec()
    out = []
    for (n_cif, n_qcif, seed) in [(0, 24, 31), (8, 0, 32)]:
        w, h = (352, 288) if n_cif else (176, 144)
        O.RefDecoder(w, h, 4)
        planes = W.ldpca_bitplane_set(n_cif=n_cif, n_qcif=n_qcif, seed=seed)
        for i, pl in enumerate(planes):
            lr = pl["LR"].copy()
            ent = R.ref_compute_ce(O.P(lr))
            acc_bits = np.unpackbits(pl["acc"]).astype(np.uint8)
            dec = np.full(pl["N"], 0xAA, np.uint8)
            rate = ctypes.c_int()
            R.ref_ldpca(O.P(lr), O.P(acc_bits), ctypes.c_uint8(pl["crc"]), ctypes.c_float(ent), O.P(dec),
                        ctypes.byref(rate))
            level = [(rate.value - 8) // (pl["N"] // 66) - 2]
            out.append({"set": [n_cif, n_qcif, seed], "index": i, "N": pl["N"], "rate": rate.value,
                        "level": level[0] if level else -1, "entropy_bits": fbits(ent),
                        "decoded_sha256": sha(dec), "source_ok": bool(np.array_equal(dec, pl["bits"]))})
    return {"max_iter": 50, "mode": 0, "items": out}


def ldpca_dynstep_fixture():
    """The reference LdpcaDecoder built with DYNAMICSTEP (oracle/_ref/
    libdvcref_dyn.so) on the bitplanes of ldpca_fixture."""
    R = O.ref_dec(dynstep=True)
    out = []
    for (n_cif, n_qcif, seed) in [(0, 24, 31), (8, 0, 32)]:
        w, h = (352, 288) if n_cif else (176, 144)
        O.RefDecoder(w, h, 4, dynstep=True)
        planes = W.ldpca_bitplane_set(n_cif=n_cif, n_qcif=n_qcif, seed=seed)
        for i, pl in enumerate(planes):
            lr = pl["LR"].copy()
            ent = R.ref_compute_ce(O.P(lr))
            acc_bits = np.unpackbits(pl["acc"]).astype(np.uint8)
            dec = np.full(pl["N"], 0xAA, np.uint8)
            rate = ctypes.c_int()
            R.ref_ldpca(O.P(lr), O.P(acc_bits), ctypes.c_uint8(pl["crc"]), ctypes.c_float(ent), O.P(dec),
                        ctypes.byref(rate))
            out.append({"set": [n_cif, n_qcif, seed], "index": i, "N": pl["N"], "rate": rate.value,
                        "level": (rate.value - 8) // (pl["N"] // 66) - 2, "entropy_bits": fbits(ent),
                        "decoded_sha256": sha(dec), "source_ok": bool(np.array_equal(dec, pl["bits"]))})
    return {"max_iter": 50, "mode": 0, "rate_step": 1, "items": out}


FRAME_CASES = [
    # (W, H, qindex, seed, shift, frames, wz frame index)  -- config-0 style QCIF and two CIF frames
    (176, 144, 4, 1000, 3, 3, 1),
    (176, 144, 8, 1004, 3, 3, 1),
    (352, 288, 2, 1005, 6, 3, 1),
    (352, 288, 8, 1006, 6, 3, 1),
]


# CIF at the Qindex values FRAME_CASES leaves out (bitplanes per frame: Q1 10, Q3 17, Q5 36, Q6 45, Q7 50)
FRAME_CASES_CIF_Q = [
    (352, 288, 1, 1021, 6, 3, 1),
    (352, 288, 3, 1023, 6, 3, 1),
    (352, 288, 5, 1025, 6, 3, 1),
    (352, 288, 6, 1026, 6, 3, 1),
    (352, 288, 7, 1027, 6, 3, 1),
]


def frames_fixture(mode=0, dynstep=False, cases=None):
    out = []
    for (w, h, q, seed, shift, nf, fi) in (FRAME_CASES if cases is None else cases):
        vid = W.synthetic_video(w, h, nf, shift, seed)
        wz = W.encode_wz_frame(vid[fi], q)
        R = O.RefDecoder(w, h, q, mode=mode, nthreads=1, dynstep=dynstep)
        # side information of the reference (createSideInfo)
        si = np.zeros((h, w), np.uint8)
        res = np.zeros((h, w), np.int32)
        R.lib.ref_set_frame(0, O.P(np.ascontiguousarray(vid[0])), O.P(np.ascontiguousarray(vid[fi])))
        R.lib.ref_set_frame(2, O.P(np.ascontiguousarray(vid[2])), O.P(np.ascontiguousarray(vid[fi])))
        R.lib.ref_create_side_info(0, 2, 1, O.P(si), O.P(res), None, None)
        R = O.RefDecoder(w, h, q, mode=mode, nthreads=1, dynstep=dynstep)
        rec, info = R.decode_wz_frame(vid[0], vid[2], vid[fi], wz)
        out.append({"W": w, "H": h, "qindex": q, "seed": seed, "shift": shift, "frames": nf, "wz": fi,
                    "mode": mode, "levels": info["levels"], "rates": info["rates"],
                    "entropy_bits": [fbits(e) for e in info["entropy"]], "bitrate": info["bitrate"],
                    "psnr": info["psnr"], "si_sha256": sha(si), "residue_sha256": sha(res),
                    "recons_sha256": sha(rec)})
        print(f"frame case {w}x{h} q{q} mode{mode}: psnr {info['psnr']:.4f} bitrate {info['bitrate']}", flush=True)
    return {"max_iter": 50, "cases": out} if not dynstep else {"max_iter": 50, "rate_step": 1, "cases": out}


def ldpca_irreg_fixture():
    """Reference LdpcaDecoder (ParallelMode 0) with the irregular QCIF ladder
    irReg1584.lad (main.cpp:84-85) on seeded bitplanes whose syndromes come
    from irRegEncoder1584.lad."""
    R = O.ref_dec()
    O.RefDecoder(176, 144, 4, kind="irreg")
    out = []
    for i, pl in enumerate(W.ldpca_bitplane_set(n_cif=0, n_qcif=24, seed=35, kind="irreg")):
        lr = pl["LR"].copy()
        ent = R.ref_compute_ce(O.P(lr))
        acc_bits = np.unpackbits(pl["acc"]).astype(np.uint8)
        dec = np.full(pl["N"], 0xAA, np.uint8)
        rate = ctypes.c_int()
        R.ref_ldpca(O.P(lr), O.P(acc_bits), ctypes.c_uint8(pl["crc"]), ctypes.c_float(ent), O.P(dec),
                    ctypes.byref(rate))
        out.append({"set": [0, 24, 35], "index": i, "N": pl["N"], "rate": rate.value,
                    "level": (rate.value - 8) // (pl["N"] // 66) - 1, "entropy_bits": fbits(ent),
                    "decoded_sha256": sha(dec), "source_ok": bool(np.array_equal(dec, pl["bits"]))})
    return {"max_iter": 50, "mode": 0, "ladder": "irreg", "items": out}


IRREG_FRAME_CASES = [(176, 144, 4, 1020, 3, 3, 1), (176, 144, 8, 1021, 3, 3, 1)]


def frames_irreg_fixture(mode=0):
    """Reference decodeGOP on QCIF WZ frames with the irregular ladder."""
    out = []
    for (w, h, q, seed, shift, nf, fi) in IRREG_FRAME_CASES:
        vid = W.synthetic_video(w, h, nf, shift, seed)
        wz = W.encode_wz_frame(vid[fi], q, kind="irreg")
        R = O.RefDecoder(w, h, q, mode=mode, nthreads=1, kind="irreg")
        rec, info = R.decode_wz_frame(vid[0], vid[2], vid[fi], wz)
        out.append({"W": w, "H": h, "qindex": q, "seed": seed, "shift": shift, "frames": nf, "wz": fi,
                    "mode": mode, "levels": info["levels"], "rates": info["rates"],
                    "entropy_bits": [fbits(e) for e in info["entropy"]], "bitrate": info["bitrate"],
                    "psnr": info["psnr"], "recons_sha256": sha(rec)})
        print(f"irregular frame {w}x{h} q{q} mode{mode}: psnr {info['psnr']:.4f} bitrate {info['bitrate']}",
              flush=True)
    return {"max_iter": 50, "ladder": "irreg", "cases": out}


NAN_BITS = 0xFFC00000  # x86's default NaN: what 0/0 yields on the reference's CPU path


def nan_planes():
    """Seeded bitplanes (workload.ldpca_bitplane_set) with 1 or 3 likelihood
    ratios replaced by NaN, as condBitProb yields them when both of a block's
    interval sums underflow to zero (p0/(p0+p1) = 0/0)."""
    rng = np.random.default_rng(77)
    out = []
    for (n_cif, n_qcif, seed) in [(0, 12, 61), (3, 0, 62)]:
        for i, pl in enumerate(W.ldpca_bitplane_set(n_cif=n_cif, n_qcif=n_qcif, seed=seed)):
            nn = 1 if i % 2 == 0 else 3
            pos = sorted(int(v) for v in rng.choice(pl["N"], nn, replace=False))
            out.append(((n_cif, n_qcif, seed), i, pl, pos))
    return out


def with_nans(lr, pos):
    lr = np.array(lr, np.float32)
    lr.view(np.uint32)[pos] = NAN_BITS
    return lr


def ldpca_nan_fixture(mode):
    """Reference LdpcaDecoder (ParallelMode `mode`, one thread) on bitplanes
    with NaN likelihood ratios: accepted code, rate, entropy bits, decoded
    hash, and (ParallelMode 1) the BP iteration count."""
    R = O.ref_dec()
    items = []
    for key, i, pl, pos in nan_planes():
        w, h = (352, 288) if pl["N"] == 6336 else (176, 144)
        O.RefDecoder(w, h, 4, mode=mode, nthreads=1)
        lr = with_nans(pl["LR"], pos)
        ent = R.ref_compute_ce(O.P(lr.copy())) if mode == 0 else R.ref_compute_ce_omp(O.P(lr.copy()))
        acc_bits = np.unpackbits(pl["acc"]).astype(np.uint8)
        dec = np.full(pl["N"], 0xAA, np.uint8)
        rate = ctypes.c_int()
        it0 = R.ref_bp_iterations()
        R.ref_ldpca(O.P(lr), O.P(acc_bits), ctypes.c_uint8(pl["crc"]), ctypes.c_float(ent), O.P(dec), ctypes.byref(rate))
        items.append({"set": list(key), "index": i, "N": pl["N"], "nan_positions": pos, "rate": rate.value,
                      "level": (rate.value - 8) // (pl["N"] // 66) - 2, "entropy_bits": fbits(ent),
                      "decoded_sha256": sha(dec), "source_ok": bool(np.array_equal(dec, pl["bits"])),
                      "bp_iterations": (R.ref_bp_iterations() - it0) if mode == 1 else None})
    return {"max_iter": 50, "mode": mode, "nan_bits": NAN_BITS, "items": items}


GOP_CASES = [
    # (W, H, qindex, seed, shift, gop): frames 0..gop of a seeded video, key
    # frames at 0 and gop, the WZ frames in between decoded hierarchically
    (176, 144, 4, 1010, 3, 4),
    (176, 144, 4, 1011, 3, 8),
    (352, 288, 8, 1012, 6, 4),
    (352, 288, 8, 1013, 6, 8),
]
GOP_CASES_MODE1 = [(176, 144, 4, 1011, 3, 8)]


def frames_gop_fixture(mode=0):
    """Reference decodeGOP(0, gop) for GOP 4 and 8: per WZ frame in the
    reference's recursion order (DVC_Decoder.cpp:471-472), its references,
    accepted code per bitplane, entropy bits, bitrate, PSNR and the hash of
    its reconstruction (which later frames of the GOP take as a reference)."""
    out = []
    for (w, h, q, seed, shift, gop) in (GOP_CASES if mode == 0 else GOP_CASES_MODE1):
        vid = W.synthetic_video(w, h, gop + 1, shift, seed)
        wz = {k: W.encode_wz_frame(vid[k], q) for k in range(1, gop)}
        R = O.RefDecoder(w, h, q, mode=mode, nthreads=1, nframes=gop + 1)
        res, order = R.decode_gop(vid, wz, 0, gop)
        frames = []
        for cur, past, fut in order:
            rec, info = res[cur]
            frames.append({"cur": cur, "past": past, "fut": fut, "levels": info["levels"],
                           "entropy_bits": [fbits(e) for e in info["entropy"]], "bitrate": info["bitrate"],
                           "psnr": info["psnr"], "recons_sha256": sha(rec)})
            print(f"gop case {w}x{h} q{q} gop{gop} mode{mode} frame {cur} ({past},{fut}): "
                  f"psnr {info['psnr']:.4f} bitrate {info['bitrate']}", flush=True)
        out.append({"W": w, "H": h, "qindex": q, "seed": seed, "shift": shift, "gop": gop, "mode": mode,
                    "frames": frames})
    return {"max_iter": 50, "cases": out}


# ------------------------------------------------------------ skip blocks, JBIG, WZ bitstream
JBIG_ILEAVE, JBIG_SMID, JBIG_TPBON, JBIG_TPDON, JBIG_DPON, JBIG_LRLTWO, JBIG_SDRST = 0x02, 0x01, 0x08, 0x10, 0x04, 0x40, 0x200


def jbig_encode(m, order, options, l0, mx):
    lib = O.ref_dec()
    h, w = m.shape
    out = np.zeros(65536, np.uint8)
    n = lib.ref_jbig_encode(O.P(np.ascontiguousarray(m, np.uint8)), w, h, order, options, l0, mx, O.P(out), len(out))
    assert n > 0
    return out[:n].tobytes()


def jbig_decode_ref(bie):
    lib = O.ref_dec()
    buf = np.frombuffer(bie, np.uint8)
    img = np.zeros(1 << 20, np.uint8)
    w, h, used = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    rc = lib.ref_jbig_decode_bie(O.P(buf), len(buf), O.P(img), len(img), ctypes.byref(w), ctypes.byref(h),
                                 ctypes.byref(used))
    assert rc == 0, rc
    return img[:w.value * h.value].reshape(h.value, w.value), used.value


def jbig_maps(seed):
    """Seeded bi-level maps: random at several densities, blocky regions (as
    skip maps are), and periodic rows that make JBIG-KIT move its adaptive
    template pixel."""
    rng = np.random.default_rng(seed)
    maps = []
    for (w, h) in [(22, 18), (44, 36), (37, 23), (64, 40)]:
        for d in (0.02, 0.3, 0.7, 0.98):
            maps.append(("random", (rng.random((h, w)) < d).astype(np.uint8)))
        blk = np.zeros((h, w), np.uint8)
        for _ in range(6):
            y0, x0 = rng.integers(0, h), rng.integers(0, w)
            blk[y0:y0 + rng.integers(2, 12), x0:x0 + rng.integers(2, 16)] = 1
        maps.append(("blocks", blk))
        per = int(rng.integers(3, 9))
        row = (rng.random(per) < 0.5).astype(np.uint8)
        pm = np.tile(row, (h, w // per + 1))[:, :w].copy()
        pm ^= (rng.random((h, w)) < 0.03).astype(np.uint8)
        maps.append((f"periodic{per}", pm))
    maps.append(("zeros", np.zeros((36, 44), np.uint8)))
    maps.append(("ones", np.ones((36, 44), np.uint8)))
    return maps


JBIG_OPTION_SETS = [
    # (name, order, options, l0 (0: the image height), mx)
    ("block_mode_map", JBIG_ILEAVE | JBIG_SMID, JBIG_TPDON | JBIG_DPON, 0, 8),
    ("tpbon", JBIG_ILEAVE | JBIG_SMID, JBIG_TPBON | JBIG_TPDON | JBIG_DPON, 0, 8),
    ("two_line", JBIG_ILEAVE | JBIG_SMID, JBIG_LRLTWO | JBIG_TPBON, 0, 8),
    ("stripes4", JBIG_ILEAVE | JBIG_SMID, JBIG_TPBON, 4, 8),
    ("sdrst_stripes5", JBIG_ILEAVE | JBIG_SMID, JBIG_TPBON | JBIG_SDRST, 5, 8),
    ("no_at", JBIG_ILEAVE | JBIG_SMID, JBIG_TPDON | JBIG_DPON, 0, 0),
    ("at_mx16", JBIG_ILEAVE | JBIG_SMID, JBIG_TPBON, 7, 16),
]


def pack_rows(m):
    return np.packbits(np.ascontiguousarray(m, np.uint8), axis=1).tobytes().hex()


def jbig_fixture():
    cases = []
    for k, (kind, m) in enumerate(jbig_maps(2024)):
        h, w = m.shape
        for (name, order, options, l0, mx) in JBIG_OPTION_SETS:
            bie = jbig_encode(m, order, options, l0 or h, mx)
            img, used = jbig_decode_ref(bie)
            assert np.array_equal(img, m)
            case = {"map": k, "kind": kind, "W": w, "H": h, "options": name, "bie": bie.hex(),
                    "image": pack_rows(img), "consumed": used}
            if name == "block_mode_map":
                # readGOP's own path: the data only, header rebuilt, bytes used
                mp = np.zeros(w * h, np.uint8)
                data = np.frombuffer(bie[20:] + bytes(64), np.uint8)  # the bitstream goes on after the map
                case["map_decode_consumed"] = O.ref_dec().ref_jbig_map_decode(O.P(data), len(data), w, h, O.P(mp))
                assert np.array_equal(mp.reshape(h, w), m)
            cases.append(case)
    print(f"jbig: {len(cases)} cases", flush=True)
    return {"cases": cases}


SKIP_CASES = [
    # (W, H, qindex, seed, shift, frames, wz frame index)
    (176, 144, 4, 1101, 0, 3, 1),
    (176, 144, 8, 1102, 0, 3, 1),
    (352, 288, 8, 1103, 0, 3, 1),
    (352, 288, 4, 1104, 1, 3, 1),
]


def ref_encode_skip(frame, past, q):
    """The reference encoder with skip blocks on: (WZFrame, skip_jbig or None)."""
    enc = O.ref_enc()
    h, w = frame.shape
    n = (w // 4) * (h // 4)
    lad = REF / f"release/DvcEnc/LDPCA_LadderFile/RegEncoder{n}.lad"
    nbp = enc.refenc_setup(w, h, q, str(lad).encode())
    acr = np.zeros(16, np.int32)
    crc = np.zeros(nbp, np.uint8)
    acc = np.zeros((nbp, n // 8), np.uint8)
    flags = np.zeros((h // 8) * (w // 8), np.uint8)
    jb = np.zeros(4096, np.uint8)
    jl = ctypes.c_int()
    mode = enc.refenc_encode_skip(O.P(np.ascontiguousarray(frame)), O.P(np.ascontiguousarray(past)), O.P(acr), O.P(crc),
                                  O.P(acc), O.P(flags), O.P(jb), len(jb), ctypes.byref(jl))
    block_mode = (flags * W.SKIP_BLOCK).astype(np.uint8)
    mine = W.encode_wz_frame(frame, q, block_mode=block_mode)
    assert np.array_equal(mine.ac_range, acr) and np.array_equal(mine.crc, crc) and np.array_equal(mine.acc, acc), \
        "numpy encoder differs from the reference encoder with skip blocks"
    assert mode == 0 or jl.value > 0
    return mine, (jb[:jl.value].tobytes() if mode == 1 else None), block_mode


def frames_skip_fixture():
    out = []
    for (w, h, q, seed, shift, nf, fi) in SKIP_CASES:
        vid = W.synthetic_video(w, h, nf, shift, seed)
        wz, jbig, block_mode = ref_encode_skip(vid[fi], vid[0], q)
        assert jbig is not None, "case without skip mode"
        R = O.RefDecoder(w, h, q, mode=0, nthreads=1)
        rec, info = R.decode_wz_frame(vid[0], vid[2], vid[fi], wz, skip_jbig=jbig)
        out.append({"W": w, "H": h, "qindex": q, "seed": seed, "shift": shift, "frames": nf, "wz": fi,
                    "skip_blocks": int((block_mode == W.SKIP_BLOCK).sum()), "skip_jbig": jbig.hex(),
                    "block_mode_sha256": sha(block_mode), "levels": info["levels"],
                    "entropy_bits": [fbits(e) for e in info["entropy"]], "bitrate": info["bitrate"],
                    "psnr": info["psnr"], "recons_sha256": sha(rec)})
        print(f"skip case {w}x{h} q{q}: {out[-1]['skip_blocks']} skip blocks, jbig {len(jbig)} B, "
              f"psnr {info['psnr']:.4f} bitrate {info['bitrate']}", flush=True)
    return {"max_iter": 50, "cases": out}


STREAM_CASE = dict(W=176, H=144, qindex=4, seed=1201, shift=0, gop=4, frames=9)


def wz_stream_fixture():
    """A WZ bitstream (GOPs of 4 over 9 frames, skip frames as the reference
    encoder decides) parsed by the reference's readWZBitstream, and decoded
    GOP by GOP by the reference's decodeGOP."""
    c = STREAM_CASE
    w, h, q, gop, nf = c["W"], c["H"], c["qindex"], c["gop"], c["frames"]
    vid = W.synthetic_video(w, h, nf, c["shift"], c["seed"])
    wz, jbig = {}, {}
    for i in range(0, nf, gop):
        order = []
        O._gop_order(i, min(i + gop, nf - 1), order)
        for cur, past, fut in order:
            f, jb, _ = ref_encode_skip(vid[cur], vid[past], q)
            wz[cur] = f
            if jb is not None:
                jbig[cur] = jb
    stream = W.write_wz_bitstream(w, h, q, gop, nf, wz, skip_jbig=jbig)
    path = Path("/tmp/dvc_wz_stream_fixture.bin")
    path.write_bytes(stream)
    lib = O.ref_dec()
    O.RefDecoder(w, h, q)  # geometry and ladder
    hdr = np.zeros(8, np.int32)
    lib.ref_read_wz_bitstream(str(path).encode(), O.P(hdr))
    n8, nbp, n = (w // 8) * (h // 8), int(hdr[7]), (w // 4) * (h // 4)
    frames = []
    for cur in sorted(wz):
        acr = np.zeros(16, np.int32); modes = np.zeros(4, np.int32); flags = np.zeros(n8, np.uint8)
        crc = np.zeros(nbp, np.uint8); acc = np.zeros((nbp, n // 8), np.uint8)
        lib.ref_wz_frame(cur, O.P(acr), O.P(modes), O.P(flags), O.P(crc), O.P(acc))
        frames.append({"cur": cur, "ac_range": acr.tolist(), "skip_mode": int(modes[0]), "intra_mode": int(modes[1]),
                       "skip_bits": int(modes[2]), "intra_bits": int(modes[3]), "block_mode_sha256": sha(flags),
                       "crc_sha256": sha(crc), "acc_sha256": sha(acc)})
    # decodeGOP over each GOP on what readGOP read (skip maps from the JBIG data)
    decoded = {}
    for i in range(0, nf, gop):
        end = min(i + gop, nf - 1)
        R = O.RefDecoder(w, h, q, mode=0, nthreads=1, nframes=nf)
        res, order = R.decode_gop(vid, wz, i, end, skip_jbig=jbig)
        for cur, past, fut in order:
            rec, info = res[cur]
            decoded[cur] = {"cur": cur, "past": past, "fut": fut, "levels": info["levels"],
                            "entropy_bits": [fbits(e) for e in info["entropy"]], "bitrate": info["bitrate"],
                            "psnr": info["psnr"], "recons_sha256": sha(rec)}
            print(f"stream frame {cur} ({past},{fut}) skip={cur in jbig}: psnr {info['psnr']:.4f} "
                  f"bitrate {info['bitrate']}", flush=True)
    return {"case": c, "header": hdr.tolist(), "stream_sha256": hashlib.sha256(stream).hexdigest(),
            "stream_bytes": len(stream), "skip_jbig": {str(k): v.hex() for k, v in jbig.items()},
            "frames": frames, "decoded": [decoded[k] for k in sorted(decoded)], "max_iter": 50}


INTRA_CASES = [
    # (name, W, H, qindex, gop, frames, seed, shift, skip, intra, deblock, intra_qp, noise box)
    # the noise box (y0, y1, x0, x1, amplitude) adds fresh uniform noise to a region of every frame, which
    # the side information cannot predict, so the encoder codes blocks there in intra mode
    ("qcif_q4_gop2_skip_intra_deblock", 176, 144, 4, 2, 3, 1311, 0, 1, 1, 1, 28, (40, 100, 60, 140, 80)),
    ("qcif_q4_gop4_skip_intra_deblock", 176, 144, 4, 4, 5, 1314, 0, 1, 1, 1, 30, (20, 80, 20, 100, 70)),
    ("qcif_q8_gop2_intra_deblock", 176, 144, 8, 2, 3, 1312, 1, 0, 1, 1, 24, (0, 72, 0, 88, 40)),
    ("qcif_q4_gop2_deblock", 176, 144, 4, 2, 5, 1316, 1, 0, 0, 1, 30, None),
    ("cif_q4_gop2_skip_intra_deblock", 352, 288, 4, 2, 3, 1315, 0, 1, 1, 1, 30, (80, 200, 120, 280, 80)),
]


def intra_video(w, h, nf, shift, seed, box):
    """The synthetic video with the case's noise box added (clipped to 8 bits)."""
    vid = W.synthetic_video(w, h, nf, shift, seed).astype(np.int32)
    if box is not None:
        rng = np.random.default_rng(seed + 99)
        y0, y1, x0, x1, amp = box
        for k in range(nf):
            vid[k, y0:y1, x0:x1] += rng.integers(-amp, amp + 1, (y1 - y0, x1 - x0))
    return np.ascontiguousarray(np.clip(vid, 0, 255).astype(np.uint8))


def intra_streams_fixture():
    """Each case in a process of its own, as the reference codes and decodes
    one sequence per run: deblock_filter fixes its QP at its first call
    (a function-local static, deblock.cpp:47) and the encoder carries its
    intra model from frame to frame."""
    cases = []
    for k in range(len(INTRA_CASES)):
        out = Path(f"/tmp/dvc_intra_case{k}.json")
        env = dict(os.environ, DVC_GOLDEN_ONLY="intra_case", DVC_GOLDEN_CASE=str(k))
        subprocess.run([sys.executable, __file__], env=env, check=True)
        cases.append(json.loads(out.read_text()))
    return {"max_iter": 50, "cases": cases}


def intra_stream_case(k):
    import base64
    cases = []
    for (name, w, h, q, gop, nf, seed, shift, skip, intra, deblock, qp, box) in INTRA_CASES[k:k + 1]:
        vid = intra_video(w, h, nf, shift, seed, box)
        n = (w // 4) * (h // 4)
        lad = REF / f"release/DvcEnc/LDPCA_LadderFile/RegEncoder{n}.lad"
        wz_path, intra_path = Path(f"/tmp/dvc_{name}.wz"), Path(f"/tmp/dvc_{name}.intra")
        if intra_path.exists():
            intra_path.unlink()
        nwz = O.ref_enc().refenc_run(O.P(vid), nf, w, h, q, gop, skip, intra, deblock, qp, str(lad).encode(),
                                     str(wz_path).encode(), str(intra_path).encode())
        assert nwz > 0
        wz_bytes = wz_path.read_bytes()
        intra_bytes = intra_path.read_bytes() if intra_path.exists() else b""
        hdr, has_intra, frames = ref_decode_streams(vid, w, h, q, gop, nf, n, wz_path, intra_path, 0, name)
        cases.append({"name": name, "W": w, "H": h, "qindex": q, "gop": gop, "frames": nf, "seed": seed,
                      "shift": shift, "noise_box": box, "header": hdr.tolist(), "intra_mode_exist": int(has_intra),
                      "wz_stream": base64.b64encode(wz_bytes).decode(),
                      "intra_stream": base64.b64encode(intra_bytes).decode(), "decoded": frames})
    return cases[0]


def intra_streams_mode1_fixture():
    """The streams of intra_streams.json decoded by the reference in ParallelMode 1."""
    import base64
    src = json.loads((HERE / "intra_streams.json").read_text())
    k = int(os.environ["DVC_GOLDEN_CASE"])
    cases = []
    for c in src["cases"][k:k + 1]:
        w, h, q, gop, nf = c["W"], c["H"], c["qindex"], c["gop"], c["frames"]
        vid = intra_video(w, h, nf, c["shift"], c["seed"], tuple(c["noise_box"]) if c["noise_box"] else None)
        wz_path, intra_path = Path(f"/tmp/dvc_{c['name']}_m1.wz"), Path(f"/tmp/dvc_{c['name']}_m1.intra")
        wz_path.write_bytes(base64.b64decode(c["wz_stream"]))
        intra_path.write_bytes(base64.b64decode(c["intra_stream"]))
        _, _, frames = ref_decode_streams(vid, w, h, q, gop, nf, (w // 4) * (h // 4), wz_path, intra_path, 1,
                                          c["name"] + " mode1")
        cases.append({"name": c["name"], "decoded": frames})
    return cases[0]


def ref_decode_streams(vid, w, h, q, gop, nf, n, wz_path, intra_path, mode, name):
    """The reference decoder's main path on a WZ and an intra bitstream
    (main.cpp:36-42, then decodeGOP per GOP): header, IntraModeExist and per
    WZ frame its block modes, mode bits, codes, entropy, bitrate, PSNR and
    reconstruction hash."""
    if True:
        R = O.RefDecoder(w, h, q, mode=mode, nthreads=1, nframes=nf)
        hdr = np.zeros(8, np.int32)
        has_intra = R.lib.ref_load_stream(str(wz_path).encode(), str(intra_path).encode(), O.P(hdr))
        order = []
        for i in range(0, nf, gop):
            O._gop_order(i, min(i + gop, nf - 1), order)
        wzset = {c for c, _, _ in order}
        for k in range(nf):
            pix = vid[k] if k not in wzset else np.zeros((h, w), np.uint8)
            R.lib.ref_set_frame(k, O.P(np.ascontiguousarray(pix)), O.P(np.ascontiguousarray(vid[k])))
        n8 = (w // 8) * (h // 8)
        modes = {}
        for c, _, _ in order:
            acr = np.zeros(16, np.int32); md = np.zeros(4, np.int32); flags = np.zeros(n8, np.uint8)
            crc = np.zeros(int(hdr[7]), np.uint8); acc = np.zeros((int(hdr[7]), n // 8), np.uint8)
            R.lib.ref_wz_frame(c, O.P(acr), O.P(md), O.P(flags), O.P(crc), O.P(acc))
            modes[c] = (flags.copy(), md.tolist())
        frames = []
        for i in range(0, nf, gop):
            end = min(i + gop, nf - 1)
            sub = []
            O._gop_order(i, end, sub)
            cnt = R.lib.ref_decode_gop(i, end)
            rates = np.zeros(4096, np.int32); levels = np.zeros(4096, np.int32); ent = np.zeros(4096, np.float32)
            R.lib.ref_get_log(O.P(rates), O.P(levels), O.P(ent))
            at = 0
            for cur, past, fut in sub:
                nb = int(hdr[7])
                rec = np.zeros((h, w), np.uint8)
                R.lib.ref_get_frame(cur, O.P(rec))
                flags, md = modes[cur]
                acr = np.zeros(16, np.int32); md2 = np.zeros(4, np.int32); fl2 = np.zeros(n8, np.uint8)
                crc = np.zeros(int(hdr[7]), np.uint8); acc = np.zeros((int(hdr[7]), n // 8), np.uint8)
                R.lib.ref_wz_frame(cur, O.P(acr), O.P(md2), O.P(fl2), O.P(crc), O.P(acc))  # after decodeGOP
                frames.append({"cur": cur, "past": past, "fut": fut, "skip_mode": md[0], "intra_mode": md[1],
                               "skip_block_bits": int(md2[2]), "intra_block_bits": int(md2[3]),
                               "skip_blocks": int((flags == 2).sum()), "intra_blocks": int((flags == 1).sum()),
                               "block_mode_sha256": sha(flags), "levels": levels[at:at + nb].tolist(),
                               "entropy_bits": [fbits(e) for e in ent[at:at + nb]],
                               "bitrate": R.lib.ref_get_bitrate(cur), "psnr": R.lib.ref_get_psnr(cur),
                               "recons_sha256": sha(rec)})
                at += nb
                print(f"{name} frame {cur} ({past},{fut}): skip {frames[-1]['skip_blocks']} intra "
                      f"{frames[-1]['intra_blocks']} psnr {frames[-1]['psnr']:.4f} bitrate {frames[-1]['bitrate']}",
                      flush=True)
            assert at == cnt, (at, cnt)
        return hdr, has_intra, frames


def main():
    assert O.have_reference_build(), "run `make -C oracle ref` first"
    if os.environ.get("DVC_GOLDEN_ONLY") == "intra":
        (HERE / "intra_streams.json").write_text(json.dumps(intra_streams_fixture(), indent=1))
        m1 = []
        for k in range(len(INTRA_CASES)):
            env = dict(os.environ, DVC_GOLDEN_ONLY="intra_mode1", DVC_GOLDEN_CASE=str(k), MALLOC_PERTURB_="85",
                       OMP_NUM_THREADS="1")
            subprocess.run([sys.executable, __file__], env=env, check=True)
            m1.append(json.loads(Path(f"/tmp/dvc_intra_m1_case{k}.json").read_text()))
        (HERE / "intra_streams_mode1.json").write_text(json.dumps({"max_iter": 50, "mode": 1, "cases": m1}, indent=1))
        return
    if os.environ.get("DVC_GOLDEN_ONLY") == "intra_case":
        k = int(os.environ["DVC_GOLDEN_CASE"])
        Path(f"/tmp/dvc_intra_case{k}.json").write_text(json.dumps(intra_stream_case(k)))
        return
    if os.environ.get("DVC_GOLDEN_ONLY") == "intra_mode1":
        k = int(os.environ["DVC_GOLDEN_CASE"])
        Path(f"/tmp/dvc_intra_m1_case{k}.json").write_text(json.dumps(intra_streams_mode1_fixture()))
        return
    if os.environ.get("DVC_GOLDEN_ONLY") == "cif_q":
        (HERE / "frames_cif_q.json").write_text(json.dumps(frames_fixture(cases=FRAME_CASES_CIF_Q), indent=1))
        return
    if os.environ.get("DVC_GOLDEN_ONLY") == "skip":
        (HERE / "jbig.json").write_text(json.dumps(jbig_fixture(), indent=1))
        (HERE / "frames_skip.json").write_text(json.dumps(frames_skip_fixture(), indent=1))
        (HERE / "wz_stream.json").write_text(json.dumps(wz_stream_fixture(), indent=1))
        return
    if os.environ.get("DVC_GOLDEN_ONLY") == "ladder":
        (HERE / "ladder.json").write_text(json.dumps(ladder_fixture(), indent=1))
        return
    if os.environ.get("DVC_GOLDEN_ONLY") == "irreg":
        (HERE / "ldpca_irreg.json").write_text(json.dumps(ldpca_irreg_fixture(), indent=1))
        (HERE / "frames_irreg.json").write_text(json.dumps(frames_irreg_fixture(0), indent=1))
        env = dict(os.environ, DVC_GOLDEN_ONLY="irreg_mode1", MALLOC_PERTURB_="85", OMP_NUM_THREADS="1")
        subprocess.run([sys.executable, __file__], env=env, check=True)
        return
    if os.environ.get("DVC_GOLDEN_ONLY") == "irreg_mode1":
        (HERE / "frames_irreg_mode1.json").write_text(json.dumps(frames_irreg_fixture(1), indent=1))
        return
    if os.environ.get("DVC_GOLDEN_ONLY") == "nan":
        (HERE / "ldpca_nan.json").write_text(json.dumps(ldpca_nan_fixture(0), indent=1))
        env = dict(os.environ, DVC_GOLDEN_ONLY="nan_mode1", MALLOC_PERTURB_="85", OMP_NUM_THREADS="1")
        subprocess.run([sys.executable, __file__], env=env, check=True)
        return
    if os.environ.get("DVC_GOLDEN_ONLY") == "nan_mode1":
        (HERE / "ldpca_nan_mode1.json").write_text(json.dumps(ldpca_nan_fixture(1), indent=1))
        return
    if os.environ.get("DVC_GOLDEN_ONLY") == "gop":
        (HERE / "frames_gop.json").write_text(json.dumps(frames_gop_fixture(mode=0), indent=1))
        env = dict(os.environ, DVC_GOLDEN_ONLY="gop_mode1", MALLOC_PERTURB_="85", OMP_NUM_THREADS="1")
        subprocess.run([sys.executable, __file__], env=env, check=True)
        return
    if os.environ.get("DVC_GOLDEN_ONLY") == "gop_mode1":
        (HERE / "frames_gop_mode1.json").write_text(json.dumps(frames_gop_fixture(mode=1), indent=1))
        return
    if os.environ.get("DVC_GOLDEN_ONLY") == "dynstep":
        # the DYNAMICSTEP build's fixtures alone (the others are unchanged by it)
        (HERE / "ldpca_dynstep.json").write_text(json.dumps(ldpca_dynstep_fixture(), indent=1))
        (HERE / "frames_dynstep.json").write_text(json.dumps(frames_fixture(mode=0, dynstep=True), indent=1))
        return
    (HERE / "ladder.json").write_text(json.dumps(ladder_fixture(), indent=1))
    (HERE / "encoder.json").write_text(json.dumps(encoder_fixture(), indent=1))
    (HERE / "ldpca.json").write_text(json.dumps(ldpca_fixture(), indent=1))
    if os.environ.get("DVC_GOLDEN_CHILD") == "mode1":
        (HERE / "frames_mode1.json").write_text(json.dumps(frames_fixture(mode=1), indent=1))
        return
    (HERE / "frames.json").write_text(json.dumps(frames_fixture(mode=0), indent=1))
    # ParallelMode 1 keeps the BP decision buffer across calls; its first use
    # reads fresh malloc memory, made deterministic (0xAA) with MALLOC_PERTURB_.
    env = dict(os.environ, DVC_GOLDEN_CHILD="mode1", MALLOC_PERTURB_="85", OMP_NUM_THREADS="1")
    subprocess.run([sys.executable, __file__], env=env, check=True)


if __name__ == "__main__":
    main()
